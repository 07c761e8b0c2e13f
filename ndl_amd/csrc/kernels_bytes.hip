// kernels_bytes.hip -- two-axis passes over 8-bit images with small
// non-negative integer taps (box filters, binomials), four pixels per lane.
//
// convolve() / the per-axis chain over uint8 images with integer taps sum
// exact integers in FP64 and narrow with static_cast<uint8_t>
// (convolution.h:64-70): the result is the low byte of the exact sum (the
// reference's x86-64 code truncates into int32 first; every sum here stays
// below 2^31, so that is the same byte).  This kernel computes the same sum
// in 32-bit integers, so it is bit-identical for every border and rounding
// mode (an integral sum rounds to itself).
//
// A wave owns a strip of 256 x positions (one dword of four pixels per lane)
// and marches down a chunk of y rows:
//   * per input row it loads its own dword and the two neighbouring dwords
//     (their pixels folded at the image edges: wrap reads the other end,
//     clamp broadcasts the edge pixel, reflect reverses the edge dword);
//   * each of the four outputs' x window (at most 9 pixels, reach <= 4) is
//     cut out of those 12 bytes with v_alignbyte and summed against the
//     packed 8-bit taps with v_dot4_u32_u8 (taps in window order);
//   * the x sums of the last KY rows stay in a register ring as packed
//     16-bit pairs; each output row is sum_t wy[t] * ring[t] with packed
//     16-bit integer multiply-adds, and the four low bytes are stored as
//     one dword.  Only the low byte of each sum is kept, and arithmetic
//     modulo 2^16 leaves the low byte of an integer sum unchanged.
// Taps ascending, zero taps contribute 0: the order of an exact integer sum
// does not change it.
#include "ndl_internal.hpp"

namespace ndl {
namespace {

template <int V> using ICb = std::integral_constant<int, V>;
template <class F, int... I> __device__ __forceinline__ void sfor_impl(F&& f, std::integer_sequence<int, I...>)
{
    (f(ICb<I>{}), ...);
}
template <int N, class F> __device__ __forceinline__ void sfor(F&& f)
{
    sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// 0..3 -> byte permutations of one dword (v_perm_b32 selectors)
constexpr uint32_t kKeep = 0x03020100u, kReverse = 0x00010203u, kFirst = 0x00000000u, kLast = 0x03030303u;

template <int KX, int KY>
__global__ __launch_bounds__(256) void bytes_2d_kernel(const ByteArgs a)
{
    static_assert(KX >= 1 && KX <= 9 && KY >= 1 && KY <= 9, "reach up to four pixels");
    constexpr int CX = KX / 2; // kernelCenter (detail.h:155-163); the host routes other centres elsewhere
    const uint32_t wave = blockIdx.x * 4u + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wave >= a.waves) return; // whole wave
    const int lane = threadIdx.x & 63;
    uint32_t q = a.fstrips.div(wave);
    const int strip = (int)(wave - q * (uint32_t)a.strips);
    uint32_t b = a.fchunks.div(q);
    const int chunk = (int)(q - b * (uint32_t)a.chunks);
    int64_t soff = 0, doff = 0;
    for (int d = 0; d < a.nb; d++) {
        const uint32_t qq = a.bext[d].div(b);
        const int64_t c = (int64_t)(b - qq * (uint32_t)a.bext[d].d);
        b = qq;
        soff += c * a.sb[d];
        doff += c * a.db[d];
    }
    const int x = strip * 256 + 4 * lane;
    const bool active = x < a.X;
    const int y0 = chunk * a.rows;
    const int ny = min(a.rows, a.Y - y0);
    const int nin = ny + KY - 1;

    const auto rs = __builtin_amdgcn_make_buffer_rsrc((void*)(a.src + soff), (short)0, 0x7fffffff, 0x00020000);
    const auto rd = __builtin_amdgcn_make_buffer_rsrc((void*)(a.dst + doff), (short)0, 0x7fffffff, 0x00020000);

    // neighbour dwords: byte offset within a row and the permutation that
    // folds them at the image edges (single fold: reach <= 4 <= X)
    int xl = x - 4, xr = x + 4;
    uint32_t sl = kKeep, sr = kKeep;
    if (xl < 0) {
        if (a.bx == NDL_WRAP) xl += a.X;
        else { xl = x; sl = a.bx == NDL_REFLECT ? kReverse : kFirst; }
    }
    if (xr >= a.X) {
        if (a.bx == NDL_WRAP) xr -= a.X;
        else { xr = x; sr = a.bx == NDL_REFLECT ? kReverse : kLast; }
    }
    if (!active) xl = xr = 0;
    const int xo = active ? x : 0;

    typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
    u16x2 ring[KY][2]; // x sums of pixels (0, 1) and (2, 3), modulo 2^16
#pragma unroll
    for (int t = 0; t < KY; t++) ring[t][0] = ring[t][1] = u16x2{0, 0};

    // window start of output k within the 12 bytes (left dword first): 4 + k - CX
    auto row_off = [&](int r) -> int {
        const int yy = border_map32(a.by, a.Y, y0 - a.cy + r);
        return __builtin_amdgcn_readfirstlane(yy * a.sy);
    };
    uint32_t bl[2], bc[2], br[2];
    auto load_row = [&](int r, int slot) __attribute__((always_inline)) {
        const int ro = row_off(r);
        bl[slot] = __builtin_amdgcn_raw_buffer_load_b32(rs, xl, ro, 0);
        bc[slot] = __builtin_amdgcn_raw_buffer_load_b32(rs, xo, ro, 0);
        br[slot] = __builtin_amdgcn_raw_buffer_load_b32(rs, xr, ro, 0);
    };
    load_row(0, 0);
    for (int r0 = 0; r0 < nin; r0 += 2 * KY) {
        sfor<2 * KY>([&](auto ic) __attribute__((always_inline)) {
            constexpr int u = decltype(ic)::value;
            constexpr int slot = u & 1;
            constexpr int U = u % KY; // ring slot of input row r
            const int r = r0 + u;
            if (r >= nin) return;
            if (r + 1 < nin) load_row(r + 1, slot ^ 1);
            const uint32_t L = __builtin_amdgcn_perm(bl[slot], bl[slot], sl);
            const uint32_t C = bc[slot];
            const uint32_t R = __builtin_amdgcn_perm(br[slot], br[slot], sr);
            uint32_t xsum[4];
            sfor<4>([&](auto kc) __attribute__((always_inline)) {
                constexpr int k = decltype(kc)::value;
                constexpr int o = 4 + k - CX; // 0..7
                uint32_t d0, d1, d2 = 0u;
                if constexpr (o < 4) {
                    d0 = __builtin_amdgcn_alignbyte(C, L, o);
                    d1 = __builtin_amdgcn_alignbyte(R, C, o);
                    if constexpr (KX > 8) d2 = __builtin_amdgcn_alignbyte(0u, R, o);
                } else {
                    d0 = __builtin_amdgcn_alignbyte(R, C, o - 4);
                    d1 = __builtin_amdgcn_alignbyte(0u, R, o - 4);
                }
                uint32_t s = __builtin_amdgcn_udot4(d0, a.wx[0], 0u, false);
                if constexpr (KX > 4) s = __builtin_amdgcn_udot4(d1, a.wx[1], s, false);
                if constexpr (KX > 8) s = __builtin_amdgcn_udot4(d2, a.wx[2], s, false);
                xsum[k] = s;
            });
            // low halves of (0, 1) and (2, 3)
            ring[U][0] = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(xsum[1], xsum[0], 0x05040100u));
            ring[U][1] = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(xsum[3], xsum[2], 0x05040100u));
            if (r >= KY - 1) {
                u16x2 acc0 = {0, 0}, acc1 = {0, 0};
#pragma unroll
                for (int t = 0; t < KY; t++) {
                    const unsigned short w = (unsigned short)a.wy[t];
                    const u16x2 w2 = {w, w};
                    acc0 = acc0 + w2 * ring[(U + 1 + t) % KY][0];
                    acc1 = acc1 + w2 * ring[(U + 1 + t) % KY][1];
                }
                // bytes 0 and 2 of each pair: the low bytes of the four sums
                const uint32_t packed = __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, acc1),
                                                              __builtin_bit_cast(uint32_t, acc0), 0x06040200u);
                const int yo = y0 + r - (KY - 1);
                if (active)
                    __builtin_amdgcn_raw_buffer_store_b32(packed, rd, x,
                                                          __builtin_amdgcn_readfirstlane(yo * a.dy), 0);
            }
        });
    }
}

template <class F> hipError_t with_bytes_instance(int kx, int ky, F&& f)
{
#define NDL_B(KX, KY) \
    if (kx == KX && ky == KY) return f(ICb<KX>{}, ICb<KY>{});
    NDL_B(3, 3) NDL_B(5, 5) NDL_B(7, 7) NDL_B(9, 9) NDL_B(3, 1) NDL_B(5, 1) NDL_B(7, 1) NDL_B(9, 1)
    NDL_B(1, 3) NDL_B(1, 5) NDL_B(1, 7) NDL_B(1, 9)
#undef NDL_B
    return hipErrorInvalidValue;
}

} // namespace

bool bytes_2d_supported(int kx, int ky)
{
    return with_bytes_instance(kx, ky, [](auto, auto) { return hipSuccess; }) == hipSuccess;
}

hipError_t launch_bytes_2d(const ByteArgs& a, int kx, int ky, hipStream_t s)
{
    if (a.waves == 0) return hipSuccess;
    return with_bytes_instance(kx, ky, [&](auto kxs, auto kys) {
        hipLaunchKernelGGL((bytes_2d_kernel<decltype(kxs)::value, decltype(kys)::value>), dim3((a.waves + 3) / 4),
                           dim3(256), 0, s, a);
        return hipGetLastError();
    });
}

} // namespace ndl
