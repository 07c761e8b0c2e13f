// kernels_fused_edge_c.hip -- run-time-tap edge instances (float storage),
// ring periods 17 15 13 7 (fused_edge_part.hpp).
#include "fused_edge_part.hpp"

NDL_EDGE_PART_BEGIN(edge_part_c)
NDL_EDGE_GEN(NDL_F32, float, 17)
NDL_EDGE_GEN(NDL_F32, float, 15)
NDL_EDGE_GEN(NDL_F32, float, 13)
NDL_EDGE_GEN(NDL_F32, float, 7)
NDL_EDGE_PART_END()
