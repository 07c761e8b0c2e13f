// kernels_fused_edge_b.hip -- run-time-tap edge instances (float storage),
// ring periods 21 19 11 5 3 (fused_edge_part.hpp).
#include "fused_edge_part.hpp"

NDL_EDGE_PART_BEGIN(edge_part_b)
NDL_EDGE_GEN(NDL_F32, float, 21)
NDL_EDGE_GEN(NDL_F32, float, 19)
NDL_EDGE_GEN(NDL_F32, float, 11)
NDL_EDGE_GEN(NDL_F32, float, 5)
NDL_EDGE_GEN(NDL_F32, float, 3)
NDL_EDGE_PART_END()
