// ndl_internal.hpp -- launch interfaces shared between the kernel files and
// the C ABI implementation (api.hip).  Not installed; not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "ndl_common.hpp"

namespace ndl {

// Taps carried inline in kernel arguments; wider axes (up to NDL_MAX_TAPS)
// pass a device array instead (the generic per-axis and morphology passes).
constexpr int kArgTaps = 64;

// ------------------------------------------------------- generic axis pass
// One output per thread, outputs enumerated in `order` (dims sorted by
// |src stride|, so a wavefront reads contiguous memory for every tap).
enum PassMode { PASS_GATHER = 0, PASS_TABLE = 1 };

struct AxisPassArgs {
    const void* in;
    void* out;
    int64_t m;       // extent along the convolved axis
    int64_t in_as;   // src stride along the axis
    int32_t nd;      // iteration dims (the axis included)
    int32_t axis_pos; // position of the axis in the iteration order
    int64_t ext[NDL_MAX_DIM];
    int64_t ist[NDL_MAX_DIM];
    int64_t ost[NDL_MAX_DIM];
    FastDiv fdiv[NDL_MAX_DIM];
    uint32_t total;  // outputs (< 2^32)
    int32_t K, center, border, rounding, mode;
    int32_t EL, ER;       // PASS_TABLE: edge rows with their own weights
    int32_t W, L;         // PASS_TABLE: row width and left reach (row entry s reads position j + s - L)
    const void* table;    // PASS_TABLE: (EL+ER) rows of W weights (Acc type), device
    double taps_d[kArgTaps];
    float taps_f[kArgTaps];
    const double* wide_d;   // K > kArgTaps: the taps on the device (else null)
    const float* wide_f;
};

// accum: NDL_ACCUM_REF (FP64, separately rounded) or NDL_ACCUM_F32 (FP32 FMA)
hipError_t launch_axis_pass(const AxisPassArgs& a, int in_dt, int out_dt, int accum, hipStream_t s);

// ------------------------------------------------------- short-axis pass
// One pass along an axis of at most 16 elements (the channel axis of a
// colour image): a thread owns one line (every other index fixed), holds its
// m elements in registers and writes all m outputs (kernels_short.hip).
constexpr int kShortMax = 16;
struct ShortAxisArgs {
    const void* in;
    void* out;
    int32_t m;                        // extent of the axis
    int64_t in_as, out_as;            // strides along the axis
    int32_t nd;                       // other dimensions, fastest (by input stride) first
    int64_t ext[NDL_MAX_DIM];
    int64_t ist[NDL_MAX_DIM];
    int64_t ost[NDL_MAX_DIM];
    FastDiv fdiv[NDL_MAX_DIM];
    uint32_t lines;
    int32_t rounding;
    int32_t k, center, border;        // taps, centre, border mode
    double wd[kArgTaps];
    float wf[kArgTaps];
};
// REF: TIn -> FP64, FP64 -> TOut or FP64 -> FP64; F32: f32/f16/bf16 to the
// same type.  hipErrorInvalidValue: no instance.
hipError_t launch_short_axis(const ShortAxisArgs& a, int in_dt, int out_dt, int accum, hipStream_t s);

// ------------------------------------------------------ dense N-D kernel
struct NdArgs {
    const void* in;
    void* out;
    int32_t nd;
    int64_t ext[NDL_MAX_DIM];
    int64_t ist[NDL_MAX_DIM];
    int64_t ost[NDL_MAX_DIM];
    FastDiv fdiv[NDL_MAX_DIM];
    uint32_t total;
    int32_t border;
    int32_t ntaps;
    const int32_t* tap_off; // ntaps*nd offsets (coordinate - center), device
    const double* tap_val;  // ntaps nonzero weights, device
};
hipError_t launch_convolve_nd(const NdArgs& a, int in_dt, int out_dt, hipStream_t s);

// --------------------------------------------------------- morphology
// One min (erode) / max (dilate) pass along one axis over the taps whose
// mask entry is nonzero (processing/morphology.h:57-101).  Floating-point
// passes skip NaN (a NaN is kept only when every value is NaN); the last
// pass then applies the reference's rule for the first included tap of the
// whole structuring element (the window corner): when it is NaN the result
// is that NaN.  With strict comparisons in ascending tap order and passes in
// ascending dimension order, ties resolve to the first tap in kernel
// coordinate order, as the reference's single loop does.
struct MorphPassArgs {
    const void* in;
    void* out;
    int64_t m, in_as;
    int32_t nd, axis_pos;
    int64_t ext[NDL_MAX_DIM];
    int64_t ist[NDL_MAX_DIM];
    int64_t ost[NDL_MAX_DIM];
    FastDiv fdiv[NDL_MAX_DIM];
    uint32_t total;
    int32_t K, center, border;
    int32_t dilate;               // 0: minimum, 1: maximum
    uint8_t mask[kArgTaps];       // included taps of this axis (K <= kArgTaps)
    const uint8_t* wide_mask;     // K > kArgTaps: the mask on the device
    const void* corner;           // original source for the NaN rule, or null
    int64_t cst[NDL_MAX_DIM];     // its strides (iteration order)
    int32_t coff[NDL_MAX_DIM];    // first included tap - centre per dimension (0: none)
    int32_t cbord[NDL_MAX_DIM];   // border per dimension
};
hipError_t launch_morph_pass(const MorphPassArgs& a, int in_dt, int out_dt, hipStream_t s);
// erode/dilate with an N-D mask in one pass, the reference's loop exactly
// (NdArgs::tap_off lists the included taps; tap_val is unused)
hipError_t launch_morph_nd(const NdArgs& a, int dilate, int in_dt, int out_dt, hipStream_t s);

// ------------------------------------------------------ scale stacks
// First pass of several gaussian_blur() levels of one image: along one axis,
// every level's sum from the same input line (FP64, separately rounded,
// taps ascending, zero taps skipped), one dense FP64 output per level.
constexpr int NDL_MAX_LEVELS = 16;
struct MultiAxisArgs {
    AxisPassArgs p;                 // input, axis, iteration order, border (taps unused)
    int32_t nlev;
    void* outs[NDL_MAX_LEVELS];     // dense FP64, strides p.ost
    int32_t K[NDL_MAX_LEVELS], center[NDL_MAX_LEVELS];
    const double* taps;             // device, nlev x NDL_MAX_TAPS
};
hipError_t launch_axis_pass_multi(const MultiAxisArgs& a, int in_dt, hipStream_t s);
// out = b - a elementwise, in FP64, stored as out's type (dense or strided)
struct DiffArgs {
    const void* a;
    const void* b;
    void* out;
    int32_t nd;
    int64_t ext[NDL_MAX_DIM];
    int64_t ast[NDL_MAX_DIM], bst[NDL_MAX_DIM], ost[NDL_MAX_DIM];
    FastDiv fdiv[NDL_MAX_DIM];
    uint32_t total;
};
hipError_t launch_difference(const DiffArgs& a, int in_dt, int out_dt, hipStream_t s);

// --------------------------------------------------------- downsample
// Block mean, processing/convolution.h:241-305: one output per thread,
// FP64 sum over its clipped block in block order (dimension 0 fastest),
// std::round into integral types.
struct DownArgs {
    const void* in;
    void* out;
    int32_t nd, channel_axis, factor;
    int64_t ext[NDL_MAX_DIM];  // source extents
    int64_t oext[NDL_MAX_DIM]; // output extents
    int64_t ist[NDL_MAX_DIM];
    int64_t ost[NDL_MAX_DIM];
    FastDiv fdiv[NDL_MAX_DIM]; // by the output extents, dimension 0 first
    uint32_t total;
    // blocks of at most kDownPos positions (kernels_downsample.hip): the
    // decimated dimensions with extent > 1 (dimension 0 first), and per
    // position of a full block, in block order, its offset and its step along
    // each of them (positions past a ragged edge are skipped)
    int32_t P, nq;
    int32_t qdim[4];
    int64_t poff[16];
    uint8_t pdig[16][4];
};
constexpr int kDownPos = 16;
hipError_t launch_downsample(const DownArgs& a, int dt, hipStream_t s);
// hipErrorInvalidValue when a.P is 0 or above kDownPos
hipError_t launch_downsample_fast(const DownArgs& a, int dt, hipStream_t s);

// The last axis pass of a blur fused with the block mean: per output, each
// source position of its block in block order (dimension 0 fastest) gets the
// pass along `axis` (taps ascending, zero taps skipped), is cast to the
// output type as the blur would store it (`rounding`), and is summed in FP64;
// the mean is stored as downsample() stores it.
struct BlurDownArgs {
    const void* in;            // input of the last pass, strides ist
    void* out;
    int32_t nd, channel_axis, factor;
    int32_t axis, K, center, border, rounding;
    int64_t ext[NDL_MAX_DIM];  // full-resolution extents
    int64_t oext[NDL_MAX_DIM]; // output extents
    int64_t ist[NDL_MAX_DIM];
    int64_t ost[NDL_MAX_DIM];
    // work item: one column of outputs along `axis` (the other output
    // coordinates, decoded over ldim[0..nl-1], fastest first) and a chunk of
    // tz outputs along it; lanes consecutive in the column index
    int32_t nl;
    int32_t ldim[NDL_MAX_DIM];
    FastDiv lfd[NDL_MAX_DIM];
    uint32_t lanes;
    int32_t tz, tiles_z;
    double taps_d[kArgTaps];
    float taps_f[kArgTaps];
    const double* wide_d;      // K > kArgTaps: the taps on the device (else null)
    const float* wide_f;
    // march kernel: positions per block across the other decimated
    // dimensions (P <= 64), output columns per wave (64 / P), output type
    int32_t P, ncol, out_dt;
};
// accum NDL_ACCUM_REF: TIn any, FP64; NDL_ACCUM_F32: f32-family in and out
hipError_t launch_blur_downsample(const BlurDownArgs& a, int in_dt, int out_dt, int accum, hipStream_t s);
// The same result when `axis` is the slowest dimension of every block (each
// decimated dimension other than it comes earlier), the input is the
// accumulator type (FP64 for NDL_ACCUM_REF, FP32 for NDL_ACCUM_F32), K is odd
// in 3..25 with no zero tap and P <= 64: lanes march along the axis with a
// register ring (kernels_line.hip).  hipErrorInvalidValue: no instance.
hipError_t launch_blur_down_march(const BlurDownArgs& a, int accum, hipStream_t s);

// ------------------------------------------------------ wide one-axis pass
// One pass along any axis with a run-time tap count (kernels_wide.hip): a
// thread computes 8 consecutive outputs of one line from a register window
// of 15 elements, taps in groups of 8.  Taps on the device; no zero tap.
struct WideArgs {
    const void* in;
    void* out;
    int32_t m;                  // axis extent
    int64_t in_as, out_as;      // axis strides
    int32_t K, center, border, rounding;
    int32_t nd;                 // other dimensions, fastest (by input stride) first
    int64_t ext[NDL_MAX_DIM];
    int64_t ist[NDL_MAX_DIM];
    int64_t ost[NDL_MAX_DIM];
    FastDiv fdiv[NDL_MAX_DIM];
    uint32_t lines, chunks, items; // items = lines x chunks (8 outputs each)
    FastDiv fsplit;             // by chunks (chunk_fast) or by lines
    int32_t chunk_fast;         // 1: consecutive items step along the axis (unit-stride axis)
    const double* wd;           // K taps (NDL_ACCUM_REF)
    const float* wf;            // K taps (NDL_ACCUM_F32)
};
// REF: TIn -> FP64, FP64 -> TOut; F32: f32/f16/bf16 to the same type;
// hipErrorInvalidValue: no instance
hipError_t launch_wide_axis(const WideArgs& a, int in_dt, int out_dt, int accum, hipStream_t s);

// ---------------------------------------------------------------- gradient
// gradient(), convolution.h:359-388: every component in one launch.  One
// thread per source element; component a is the reference's convolve() with
// {-0.5, 0, 0.5} along a (the zero tap skipped): total = 0 + x[B(i-1)] *
// -0.5, then total + x[B(i+1)] * 0.5, in FP64, cast with truncation.
struct GradArgs {
    const void* in;
    void* out;
    int32_t nd;                 // source dimensions
    int64_t ext[NDL_MAX_DIM];
    int64_t ist[NDL_MAX_DIM];
    int64_t ost[NDL_MAX_DIM];   // dst strides of the spatial dimensions (dst dims 1..nd)
    int64_t cst;                // dst stride of the component axis (dst dim 0)
    FastDiv fdiv[NDL_MAX_DIM];
    uint32_t total;
    int32_t border;
};
hipError_t launch_gradient(const GradArgs& a, int in_dt, int out_dt, hipStream_t s);

// ------------------------------------------- two reference passes at once
// Two consecutive passes of the NDL_ACCUM_REF chain (kernels_xyref.hip):
// along x, then along y (two distinct dimensions, any strides), FP64 with
// every product and sum rounded separately, taps ascending, zero taps
// skipped; the first pass's FP64 results stay in LDS.  A workgroup owns a
// 64 x 32 tile of outputs (x by y) of one index of the other dimensions.
constexpr int kXYTileX = 64, kXYTileY = 32, kXYMaxTaps = 25;
struct XYRefArgs {
    const void* in;
    void* out;
    int32_t X, Y;               // extents of the two axes
    int64_t isx, isy, osx, osy; // strides (elements)
    int32_t kx, cx, bx, ky, cy, by;
    int32_t rounding;           // into the output type (integral outputs)
    int32_t tiles_x, tiles_y;
    FastDiv ftx, fty;
    int32_t nb;                 // other dimensions
    FastDiv bext[NDL_MAX_DIM];
    int64_t ib[NDL_MAX_DIM], ob[NDL_MAX_DIM];
    uint32_t ntiles;            // tiles_x * tiles_y * ceil(nbatch / nsl)
    uint32_t nbatch, nsl;       // indices of the other dimensions; per workgroup
    double wx[kXYMaxTaps], wy[kXYMaxTaps];
    // fx * fy > 1: store downsample()'s block means of the y pass results
    // (rounded as the blur stores them) over fx x fy blocks (1, 2 or 4 each),
    // at output strides osx / osy of the decimated image
    int32_t fx, fy;
};
// LDS bytes of one workgroup for these tap counts and input type size
size_t xy_ref_lds_bytes(int kx, int ky, int in_size);
// Three reference passes in one launch: a short first axis c of at most
// kCXYMaxC elements (colour channels), then x, then y (XYRefArgs for x / y;
// the tile runs every index of c).  Taps of c: wc, cc, cb; one tap count K
// for all three axes.
constexpr int kCXYMaxC = 4;
struct CXYRefArgs {
    XYRefArgs xy;
    int32_t m;                  // extent of c
    int64_t ics, ocs;           // strides of c (elements)
    int32_t cc, cb;             // centre and border of the c taps
    double wc[kXYMaxTaps];
};
size_t cxy_ref_lds_bytes(int k, int in_size);
// out_dt: NDL_F64 or in_dt (uint8 / uint16 / int16 / float32 input);
// hipErrorInvalidValue: no instance or more LDS than a workgroup may take
hipError_t launch_cxy_ref(const CXYRefArgs& a, int in_dt, int out_dt, hipStream_t s);
// out_dt: NDL_F64 or in_dt; hipErrorInvalidValue: no instance
hipError_t launch_xy_ref(const XYRefArgs& a, int in_dt, int out_dt, hipStream_t s);

// ------------------------------------------- tap-gradient partial sums
// partial[blk*K + t] = sum over the block's share of positions p of
// g[p] * u[B(p_axis + t - c)] (FP64), later summed in block order.
constexpr int kTapSeg = 16;    // tap-gradient positions per work item along the axis
struct TapGradArgs {
    const void* g;      // grad_out, strided
    const void* u;      // dense intermediate (Acc type)
    // lines along the axis: the other dimensions, fastest (by g stride) first
    int32_t nl;                    // number of other dimensions
    int64_t ext[NDL_MAX_DIM];
    int64_t gst[NDL_MAX_DIM];
    int64_t ust[NDL_MAX_DIM];
    FastDiv fdiv[NDL_MAX_DIM];
    uint32_t lines;                // product of the other extents
    int32_t seg;                   // 1: item = segment * lines + line; 0: line * nseg + segment
    uint32_t nseg;                 // segments per line
    uint32_t items;                // lines * nseg
    FastDiv ldiv;                  // divides by lines (seg = 1) or by nseg (seg = 0)
    int64_t m, g_as, u_as;         // axis extent and strides
    int32_t K, center, border;
    double* partial;    // [nblocks][K]
    int32_t nblocks;
};
hipError_t launch_tapgrad_partial(const TapGradArgs& a, int g_dt, int u_is_f64, hipStream_t s);
hipError_t launch_tapgrad_final(const double* partial, int nblocks, int K, double* out, hipStream_t s);

// ------------------------------------------------------ fused tile kernel
// Up to three convolved axes in one pass over HBM: a tile of TX x TY columns
// (x = the unit-stride dimension) marching along z with a register ring.
struct FusedArgs {
    const void* src;
    void* dst;
    int32_t X, Y, Z;
    int64_t sy, sz, dy, dz;       // element strides (x stride is 1)
    int32_t nb;                   // batch dims
    int64_t bext[NDL_MAX_DIM];
    int64_t sb[NDL_MAX_DIM], db[NDL_MAX_DIM];
    FastDiv bdiv[NDL_MAX_DIM];
    int32_t kx, cx, bx, ky, cy, by, kz, cz, bz;
    int32_t tz;                   // z outputs per tile
    int32_t tiles_x, tiles_y, tiles_z;
    uint32_t ntiles;              // tiles_x*tiles_y*tiles_z*batch
    FastDiv fdx, fdy, fdz;
    int32_t xcd_remap;            // 1: give each XCD a contiguous range of tiles
    float wx[32], wy[32], wzlag[32];   // float accumulation taps
    // edge weight rows per role (x, y, z): eL rows at the start and eR at
    // the end of the axis, k floats each in window order; null = the band
    // everywhere.  Only the EDGE instances read them.
    const float* etab[3];
    int32_t eL[3], eR[3];
    // tap-gradient instances (FusedConfig::grad): dst is grad_out (read),
    // the ring holds x passed through the x/y taps, and each tile writes the
    // kz partial sums sum_p g[p] * U[B_z(p_z + t - cz)] to gpart[tile * kz + t]
    double* gpart;
};
struct FusedConfig {
    int tx, ty, ring, kmax;       // template selection
    int cpt;                      // x outputs per thread (threads = tx*ty/cpt)
    int il;                       // interleave: x taps il elements apart (channels)
    int edge;                     // 1: rows near the ends take weight rows (FusedArgs::etab)
    int grad;                     // 1: tap gradient of the z role (FusedArgs::gpart)
};
// Returns hipErrorInvalidValue when no instantiation matches cfg/dtype.
hipError_t launch_fused(const FusedArgs& a, const FusedConfig& cfg, int dt, bool int_acc,
                        hipStream_t s);
// Persistent work queue over many jobs (device arrays owned by a plan).
struct FusedQueue {
    const FusedArgs* jobs;        // device, njobs entries
    const uint32_t* tile_start;   // device, njobs+1 prefix sums of tile counts
    int njobs;
    uint32_t total;               // tiles over all jobs
    uint32_t* counter;            // device work counter, reset by the launch
    uint32_t grid;                // persistent workgroups
};
hipError_t launch_fused_queue(const FusedQueue& q, const FusedConfig& cfg, int dt, bool int_acc, size_t lds,
                              hipStream_t s);
int fused_ring_for(int kz);       // ring period for kz, or -1 when unsupported
// Whole-row fused kernel (kernels_slab.hip): float32 storage and FP32 sums,
// x (unit stride) extent even and at most 256, three axes with the
// compile-time tap counts of an instance, x kernel centred.  FusedArgs as
// for the tile kernel with tiles_x = 1, tiles_y = ceil(Y / 4) and tz the z
// chunk; one workgroup of 768 threads per tile.
bool slab_supported(int kx, int ky, int kz);
int slab_blocks_per_cu(int kx, int ky, int kz);
hipError_t launch_slab(const FusedArgs& a, hipStream_t s);
// Instances with edge weight rows (cfg.edge), kernels_fused_edge.hip.
hipError_t launch_fused_edge(const FusedArgs& a, const FusedConfig& cfg, int dt, size_t lds, hipStream_t s);
bool fused_edge_supported(const FusedConfig& cfg, int dt, int kx, int ky);
int fused_edge_blocks_per_cu(const FusedConfig& cfg, int dt, int kx, int ky, size_t lds);
// Tap-gradient instances (cfg.grad), kernels_fused_grad.hip.
hipError_t launch_fused_grad(const FusedArgs& a, const FusedConfig& cfg, int dt, size_t lds, hipStream_t s);
bool fused_grad_supported(const FusedConfig& cfg, int dt, int kx, int ky);
int fused_grad_blocks_per_cu(const FusedConfig& cfg, int dt, int kx, int ky, size_t lds);

// Edge weight rows of transposed axis operators, built on the device:
// for axis a of extent M with taps w (count K, centre c) and border B,
// row r covers output i = r (r < EL) or M - ER + (r - EL), and
// V[r][s] = sum_t w[t] [B(j + t - c) == i] with j = i + s - L, s < W
// (0 when j is outside [0, M)); sums in FP64, t ascending.  The window
// (left reach L, width W) holds every j that folds onto i: L = K-1-c and
// W = K for kernels centred within one tap, wider otherwise (a reflect fold
// of an off-centre kernel lands beyond the band's reach).
struct EdgeTables {
    int32_t naxes;
    int32_t K[3], c[3], border[3], EL[3], ER[3], W[3], L[3];
    int64_t M[3];
    double taps[3][kArgTaps];
    const double* wide[3];   // K > kArgTaps: that axis's taps on the device
    float* outf[3];   // float rows (or null)
    double* outd[3];  // double rows (or null)
};
hipError_t launch_edge_tables(const EdgeTables& e, hipStream_t s);
// One-axis pass along a dimension other than the unit-stride one
// (kernels_line.hip): every lane owns V consecutive x elements of one line
// (x block, batch index) and marches along z with a register ring.
// Workgroups: tiles_z x ceil(lanes / 256).
struct LineArgs {
    const void* src;
    void* dst;
    int32_t Z;               // convolved extent
    int64_t sz, dz;          // src / dst strides of z (elements)
    int32_t k, c, border;    // taps (odd count), centre, border mode
    int32_t tz, tiles_z;     // outputs per chunk along z; chunks
    uint32_t lanes;          // (X / V) x product of the batch extents
    FastDiv xv;              // X / V lanes per line
    int32_t nb;              // batch dimensions, fastest first
    FastDiv bext[NDL_MAX_DIM];
    int64_t sb[NDL_MAX_DIM];
    int64_t db[NDL_MAX_DIM];
    float w[kArgTaps];       // NDL_ACCUM_F32
    double wd[kArgTaps];     // NDL_ACCUM_REF
    int32_t rounding;        // into an integral output (NDL_ACCUM_REF)
    // last erode / dilate pass of a float image: the original source, whose
    // value at the structuring element's corner (per dimension: first-tap
    // offset, border) replaces the result when it is NaN; null otherwise
    const void* corner;
    int32_t X;                              // x extent
    int64_t cxs, czs;                       // corner strides of x and z
    int32_t cox, cbx, coz, cbz;             // corner offsets / borders of x and z
    int64_t cbs[NDL_MAX_DIM];               // ... of the batch dimensions
    int32_t cob[NDL_MAX_DIM], cbb[NDL_MAX_DIM];
};
// One-axis kernel selection (erode / dilate: float32 or uint8, the same type
// in and out, every tap included): storage types (NDL_F32 / NDL_F64 in and out),
// accumulation (NDL_ACCUM_F32: float in and out; NDL_ACCUM_REF: FP64, every
// product and sum rounded separately), tap count, x elements per lane v
// (1, 2, 4) and, for the line march, loads in flight per lane d.
struct LineKey {
    int in_dt, out_dt, accum, k, v, d;
    int op;  // 0 convolution, 1 erode (minimum), 2 dilate (maximum)
};
hipError_t launch_line_march(const LineArgs& a, const LineKey& key, hipStream_t s);
bool line_march_supported(const LineKey& key);
int line_march_blocks_per_cu(const LineKey& key);
// One-axis pass along the unit-stride dimension (kernels_line.hip): a wave
// takes a segment of 64 x V outputs of one row (a fixed index of every other
// dimension) and its halo through a wave-private LDS window.
struct RowArgs {
    const void* src;
    void* dst;
    int32_t X;               // row length
    int32_t k, c, border;    // taps (odd count), centre, border mode
    uint32_t total;          // rows x segments per row
    FastDiv segs;            // segments per row, ceil(X / (64 V))
    int32_t nb;              // row dimensions, fastest first
    FastDiv bext[NDL_MAX_DIM];
    int64_t sb[NDL_MAX_DIM];
    int64_t db[NDL_MAX_DIM];
    float w[kArgTaps];       // NDL_ACCUM_F32
    double wd[kArgTaps];     // NDL_ACCUM_REF
    int32_t rounding;        // into an integral output (NDL_ACCUM_REF)
};
// key.d is unused for rows
hipError_t launch_row_pass(const RowArgs& a, const LineKey& key, uint32_t blocks, hipStream_t s);
bool row_pass_supported(const LineKey& key);
int row_pass_blocks_per_cu(const LineKey& key);
// Two-axis passes over uint8 images with non-negative integer taps (at most
// 9 per axis, x reach <= 4), four pixels per lane (kernels_bytes.hip).
// x is the unit-stride dimension (X a multiple of 4 and at least 256,
// dword-aligned rows);
// byte strides.  Each wave owns a 256-pixel strip of `rows` output rows.
struct ByteArgs {
    const uint8_t* src;
    uint8_t* dst;
    int32_t X, Y;            // x and y extents
    int32_t sy, dy;          // y strides (bytes, positive)
    int32_t cx, bx, cy, by;  // centres and borders
    int32_t rows, strips, chunks;
    uint32_t waves;          // strips * chunks * batch
    FastDiv fstrips, fchunks;
    int32_t nb;              // batch dimensions
    FastDiv bext[NDL_MAX_DIM];
    int64_t sb[NDL_MAX_DIM], db[NDL_MAX_DIM];
    uint32_t wx[3];          // x taps, four 8-bit taps per word, window order, zero-padded
    uint32_t wxs[9];         // x taps one per word (interleaved colour)
    uint32_t wy[9];          // y taps
};
bool bytes_2d_supported(int kx, int ky);
// interleaved colour ({3, w, h}: channel fastest, x stride 3 bytes): X is
// the pixel count of a row (a multiple of 4, at least 256), x convolved
// with kx in {3, 5, 7, 9} taps, ky in {1, 3, 5, 7, 9}
bool bytes_rgb_supported(int kx, int ky);
constexpr int bytes_rgb_prefetch_rows(int ky) { return ky >= 5 ? ky : (ky == 3 ? 3 : 4); }
hipError_t launch_bytes_rgb(const ByteArgs& a, int kx, int ky, hipStream_t s);
// input rows a wave keeps in flight for ky y taps; a chunk whose rows + ky - 1
// input rows are a multiple of it runs with no per-row checks
constexpr int bytes_prefetch_rows(int ky) { return ky >= 5 ? 2 * ky : (ky == 3 ? 12 : 16); }
hipError_t launch_bytes_2d(const ByteArgs& a, int kx, int ky, hipStream_t s);
// resident workgroups per CU for this instantiation (0 if unknown)
int fused_blocks_per_cu(const FusedConfig& cfg, int dt, int kx, int ky, size_t lds);
// whether an instantiation exists (queue: the persistent work-queue kernel)
bool fused_supported(const FusedConfig& cfg, int dt, int kx, int ky, bool queue);
size_t fused_lds_bytes(const FusedArgs& a, const FusedConfig& cfg, bool acc8);

} // namespace ndl
