// kernels_fused_edge_a.hip -- run-time-tap edge instances (float storage),
// ring periods 25 23 9 1 (fused_edge_part.hpp).
#include "fused_edge_part.hpp"

NDL_EDGE_PART_BEGIN(edge_part_a)
NDL_EDGE_GEN(NDL_F32, float, 25)
NDL_EDGE_GEN(NDL_F32, float, 23)
NDL_EDGE_GEN(NDL_F32, float, 9)
NDL_EDGE_GEN(NDL_F32, float, 1)
NDL_EDGE_PART_END()
