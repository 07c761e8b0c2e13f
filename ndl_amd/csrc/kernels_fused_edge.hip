// kernels_fused_edge.hip -- the fused separable kernel (fused_tile.hpp) with
// edge weight rows: the input adjoint of clamp and asymmetric-reflect
// filters in one pass, plus the device builder of those rows.
//
// Row i of a transposed axis operator A^T is the flipped band except within
// K rows of either end, where clamp / reflect fold several taps onto one
// index; those rows get a weight row each (EdgeTables).  The tile loads its
// window through the border map as in the forward pass: a weight row is
// zero wherever its source index is outside the axis, so what the border map
// puts there does not matter.
#include "fused_edge_part.hpp"

// compile-time taps: the 9-tap kernels of the 384^3 adjoint workload and the
// clamp Gaussians (2-D sigma 2; video sigma 1 + 5-tap binomial)
NDL_EDGE_PART_BEGIN(edge_part_spec)
NDL_EDGE_SPEC(NDL_F32, float, 9, 9, 9)
NDL_EDGE_SPEC(NDL_F32, float, 13, 13, 1)
NDL_EDGE_SPEC(NDL_F32, float, 7, 7, 5)
NDL_EDGE_PART_END()

namespace ndl {
namespace {

const EdgePartFn kEdgeParts[] = {edge_part_spec, edge_part_a, edge_part_b, edge_part_c};

// every edge instance: float accumulation, 64 x 16 tile, two x outputs per
// thread, no interleave
hipError_t edge_dispatch(const FusedArgs* a, const FusedConfig& cfg, int dt, int kx, int ky, size_t lds,
                         hipStream_t s, int op, int* occ)
{
    if (cfg.tx != 64 || cfg.ty != 16 || cfg.cpt != 2 || cfg.il != 1 || !cfg.edge) return hipErrorInvalidValue;
    for (EdgePartFn f : kEdgeParts) {
        hipError_t e = f(a, dt, kx, ky, cfg.ring, lds, s, op, occ);
        if (e != hipErrorInvalidValue) return e;
    }
    return hipErrorInvalidValue;
}

// one thread per (row, window slot); blockIdx.y = axis
__global__ void edge_table_kernel(const EdgeTables e)
{
    const int ax = blockIdx.y;
    const int K = e.K[ax], c = e.c[ax], W = e.W[ax], L = e.L[ax];
    const int64_t M = e.M[ax];
    const int EL = e.EL[ax], ER = e.ER[ax];
    const int n = (EL + ER) * W;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x) {
        const int r = idx / W, s = idx - r * W;
        const int64_t i = r < EL ? r : M - ER + (r - EL);
        const int64_t j = i + s - L;
        double v = 0.0;
        if (j >= 0 && j < M)
            for (int t = 0; t < K; t++)
                if (border_map(e.border[ax], M, j + t - c) == i) v += e.wide[ax] ? e.wide[ax][t] : e.taps[ax][t];
        if (e.outf[ax]) e.outf[ax][idx] = (float)v;
        if (e.outd[ax]) e.outd[ax][idx] = v;
    }
}

} // namespace

bool fused_edge_supported(const FusedConfig& cfg, int dt, int kx, int ky)
{
    return edge_dispatch(nullptr, cfg, dt, kx, ky, 0, nullptr, EDGE_QUERY, nullptr) == hipSuccess;
}

hipError_t launch_fused_edge(const FusedArgs& a, const FusedConfig& cfg, int dt, size_t lds, hipStream_t s)
{
    return edge_dispatch(&a, cfg, dt, a.kx, a.ky, lds, s, EDGE_LAUNCH, nullptr);
}

int fused_edge_blocks_per_cu(const FusedConfig& cfg, int dt, int kx, int ky, size_t lds)
{
    int n = 0;
    return edge_dispatch(nullptr, cfg, dt, kx, ky, lds, nullptr, EDGE_OCCUPANCY, &n) == hipSuccess ? n : 0;
}

hipError_t launch_edge_tables(const EdgeTables& e, hipStream_t s)
{
    int rows = 0;
    for (int a = 0; a < e.naxes; a++) rows = std::max(rows, (e.EL[a] + e.ER[a]) * e.W[a]);
    if (e.naxes <= 0 || rows == 0) return hipSuccess;
    hipLaunchKernelGGL(edge_table_kernel, dim3((rows + 255) / 256, e.naxes), dim3(256), 0, s, e);
    return hipGetLastError();
}

} // namespace ndl
