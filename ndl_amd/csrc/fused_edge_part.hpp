// fused_edge_part.hpp -- one group of edge-weight-row instances of the fused
// kernel (fused_tile.hpp, EDGE = true).  The instances are spread over
// several translation units so they compile in parallel; each unit defines
// one edge_part_<n>() from the macros below, and kernels_fused_edge.hip asks
// the parts in order (compile-time-tap part first).
#pragma once
#include "fused_tile.hpp"

namespace ndl {

enum EdgeOp { EDGE_QUERY = 0, EDGE_LAUNCH = 1, EDGE_OCCUPANCY = 2 };

// hipErrorInvalidValue: this part holds no instance for (dt, kx, ky, kz)
using EdgePartFn = hipError_t (*)(const FusedArgs* a, int dt, int kx, int ky, int kz, size_t lds, hipStream_t s,
                                  int op, int* occ);
hipError_t edge_part_spec(const FusedArgs*, int, int, int, int, size_t, hipStream_t, int, int*);
hipError_t edge_part_a(const FusedArgs*, int, int, int, int, size_t, hipStream_t, int, int*);
hipError_t edge_part_b(const FusedArgs*, int, int, int, int, size_t, hipStream_t, int, int*);
hipError_t edge_part_c(const FusedArgs*, int, int, int, int, size_t, hipStream_t, int, int*);

namespace fused_detail {

constexpr int kEdgeTX = 64, kEdgeTY = 16, kEdgeNT = kEdgeTX * kEdgeTY / 2;

template <class TS, int KXS, int KYS, int P>
hipError_t edge_op(const FusedArgs* a, size_t lds, hipStream_t s, int op, int* occ)
{
    const auto kern = fused_tile_kernel<TS, kEdgeTX, kEdgeTY, 2, KXS, KYS, P, 1, true>;
    if (op == EDGE_LAUNCH) {
        hipLaunchKernelGGL(kern, dim3(a->ntiles), dim3(kEdgeNT), lds, s, *a);
        return hipGetLastError();
    }
    if (op == EDGE_OCCUPANCY)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, reinterpret_cast<const void*>(kern), kEdgeNT, lds);
    return hipSuccess;
}

} // namespace fused_detail
} // namespace ndl

#define NDL_EDGE_PART_BEGIN(NAME)                                                                               \
    namespace ndl {                                                                                             \
    hipError_t NAME(const FusedArgs* a, int dt, int kx, int ky, int kz, size_t lds, hipStream_t s, int op,       \
                    int* occ)                                                                                   \
    {                                                                                                           \
        using namespace fused_detail;                                                                           \
        (void)kx;                                                                                               \
        (void)ky;
// compile-time x / y tap counts
#define NDL_EDGE_SPEC(DT, T, KX, KY, KZ) \
    if (dt == DT && kx == KX && ky == KY && kz == KZ) return edge_op<T, KX, KY, KZ>(a, lds, s, op, occ);
// run-time x / y tap counts, ring period KZ
#define NDL_EDGE_GEN(DT, T, KZ) \
    if (dt == DT && kz == KZ) return edge_op<T, 0, 0, KZ>(a, lds, s, op, occ);
#define NDL_EDGE_PART_END()         \
        return hipErrorInvalidValue; \
    }                                \
    }
