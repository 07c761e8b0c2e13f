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
// and marches down a chunk of y rows, keeping the next rows' loads in flight:
//   * per input row each lane loads its own dword; the dwords left and right
//     of it come from the neighbouring lanes (DPP wave shifts), and one more
//     load brings the dwords beyond the strip's ends to lanes 0 and 63 (their
//     pixels folded at the image edges);
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

// YBOX: every y tap equals wy[0]; the y sum is then wy[0] times a running
// sum of the ring (add the new row, subtract the row it replaces).
template <int KX, int KY, bool YBOX>
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
    // the last strip of a row ends at X (X >= 256): when 256 does not divide
    // X it overlaps the strip before it, whose pixels it recomputes from the
    // same inputs and stores again with the same values.  Every lane is in
    // the image, so no load, store or lane exchange of the row loop is
    // conditional (a conditional store is left out of the compiler's load
    // wait counts, and each wait then also drains later stores).
    const int xs = min(strip * 256, a.X - 256);
    const int x = xs + 4 * lane;
    const int y0 = chunk * a.rows;
    const int ny = min(a.rows, a.Y - y0);
    const int nin = ny + KY - 1;

    const auto rs = __builtin_amdgcn_make_buffer_rsrc((void*)(a.src + soff), (short)0, 0x7fffffff, 0x00020000);
    const auto rd = __builtin_amdgcn_make_buffer_rsrc((void*)(a.dst + doff), (short)0, 0x7fffffff, 0x00020000);

    // The dwords left and right of a lane's own are its neighbour lanes' own
    // dwords (wave_shr:1 / wave_shl:1 lane shifts); lane 0 and lane 63 take
    // theirs from an edge load: the dword left of the strip and the dword
    // right of it, folded at the image edges (wrap reads the other end,
    // clamp broadcasts the edge pixel, reflect reverses the edge dword;
    // single fold: reach <= 4 <= X).  Lanes 1-31 read lane 0's edge address
    // and lanes 32-62 lane 63's, so the edge load touches two lines.
    int xl = xs - 4, xr = xs + 256;
    uint32_t sl = kKeep, sr = kKeep;
    if (xl < 0) {
        if (a.bx == NDL_WRAP) xl += a.X;
        else { xl = xs; sl = a.bx == NDL_REFLECT ? kReverse : kFirst; }
    }
    if (xr >= a.X) {
        if (a.bx == NDL_WRAP) xr -= a.X;
        else { xr = xs + 252; sr = a.bx == NDL_REFLECT ? kReverse : kLast; }
    }
    const int xe = lane < 32 ? xl : xr;
    const uint32_t se = lane == 0 ? sl : (lane == 63 ? sr : kKeep);

    typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
    u16x2 ring[KY][2]; // x sums of pixels (0, 1) and (2, 3), modulo 2^16
#pragma unroll
    for (int t = 0; t < KY; t++) ring[t][0] = ring[t][1] = u16x2{0, 0};
    u16x2 run0 = {0, 0}, run1 = {0, 0}; // YBOX: sum of the ring

    // input row r is image row y0 - cy + r; output row r - (KY - 1) is y0 + r - (KY - 1)
    const int ybase = y0 - a.cy;
    const int obase = (y0 - (KY - 1)) * a.dy;

    // One pass over the chunk's input rows; EDGE chunks border-map their
    // rows, interior chunks address them linearly.  Rows r + 1 .. r + PD are
    // in flight while row r is summed; the rows run in unrolled blocks of NU
    // (a multiple of the ring and of the prefetch depth) with no per-row
    // checks but the last block's.
    auto pass = [&](auto edge_tag) __attribute__((always_inline)) {
        constexpr bool EDGE = decltype(edge_tag)::value;
        constexpr int PD = bytes_prefetch_rows(KY);
        constexpr int NU = PD;
        static_assert(NU % KY == 0, "the ring slot of a row is fixed in a block");
        auto row_off = [&](int r) -> int {
            if constexpr (EDGE) return border_map32(a.by, a.Y, ybase + r) * a.sy;
            else return (ybase + r) * a.sy;
        };
        uint32_t be[PD], bc[PD];
        auto load_row = [&](int r, int slot) __attribute__((always_inline)) {
            const int ro = row_off(r);
            bc[slot] = __builtin_amdgcn_raw_buffer_load_b32(rs, x, ro, 0);
            be[slot] = __builtin_amdgcn_raw_buffer_load_b32(rs, xe, ro, 0);
            // keep the rows' loads in row order: the waits count loads in
            // issue order, and a row's loads hoisted into an earlier row's
            // group would make every wait of the loop a deep one
            __builtin_amdgcn_sched_barrier(0);
        };
        sfor<PD>([&](auto ip) __attribute__((always_inline)) {
            constexpr int p = decltype(ip)::value;
            load_row(min(p, nin - 1), p);
        });
        // FIRST: the block at r0 = 0 (its first KY - 1 rows store nothing);
        // GUARD: the last, partial block (rows past nin skipped, r0 checked)
        auto step = [&](auto ic, int r0, auto first_tag, auto guard_tag) __attribute__((always_inline)) {
            constexpr int u = decltype(ic)::value;
            constexpr bool FIRST = decltype(first_tag)::value;
            constexpr bool GUARD = decltype(guard_tag)::value;
            constexpr int slot = u % PD;
            constexpr int U = u % KY; // ring slot of input row r
            const int r = r0 + u;
            if constexpr (GUARD) {
                if (r >= nin) return;
            }
            // window start of output k within the 12 bytes (left dword first): 4 + k - CX
            const uint32_t C = bc[slot];
            const uint32_t E = __builtin_amdgcn_perm(be[slot], be[slot], se);
            const uint32_t L = __builtin_amdgcn_update_dpp(E, C, 0x138, 0xf, 0xf, false); // wave_shr:1, lane 0 keeps E
            const uint32_t R = __builtin_amdgcn_update_dpp(E, C, 0x130, 0xf, 0xf, false); // wave_shl:1, lane 63 keeps E
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
            // refill the slot once its dwords are consumed (a load into a
            // slot whose value is still live would move every slot each block)
            load_row(min(r + PD, nin - 1), slot); // past the end: a repeat of the last row
            // low halves of (0, 1) and (2, 3)
            const u16x2 n0 = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(xsum[1], xsum[0], 0x05040100u));
            const u16x2 n1 = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(xsum[3], xsum[2], 0x05040100u));
            if constexpr (YBOX) {
                run0 = run0 - ring[U][0] + n0;
                run1 = run1 - ring[U][1] + n1;
            }
            ring[U][0] = n0;
            ring[U][1] = n1;
            // the first KY - 1 input rows produce no output row
            bool out = true;
            if constexpr (FIRST && u < KY - 1) out = false;
            if constexpr (GUARD && u < KY - 1) out = r0 > 0;
            if (out) {
                u16x2 acc0, acc1;
                const unsigned short w0 = (unsigned short)a.wy[0];
                if constexpr (YBOX) {
                    const u16x2 w2 = {w0, w0};
                    acc0 = w2 * run0;
                    acc1 = w2 * run1;
                } else {
                    acc0 = acc1 = u16x2{0, 0};
#pragma unroll
                    for (int t = 0; t < KY; t++) {
                        const unsigned short w = (unsigned short)a.wy[t];
                        const u16x2 w2 = {w, w};
                        acc0 = acc0 + w2 * ring[(U + 1 + t) % KY][0];
                        acc1 = acc1 + w2 * ring[(U + 1 + t) % KY][1];
                    }
                }
                // bytes 0 and 2 of each pair: the low bytes of the four sums
                const uint32_t packed = __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, acc1),
                                                              __builtin_bit_cast(uint32_t, acc0), 0x06040200u);
                __builtin_amdgcn_raw_buffer_store_b32(packed, rd, x, obase + r * a.dy, 0);
            }
        };
        using F = std::false_type;
        using T = std::true_type;
        int r0 = 0;
        if (NU <= nin) {
            sfor<NU>([&](auto ic) __attribute__((always_inline)) { step(ic, 0, T{}, F{}); });
            r0 = NU;
        }
        for (; r0 + NU <= nin; r0 += NU)
            sfor<NU>([&](auto ic) __attribute__((always_inline)) { step(ic, r0, F{}, F{}); });
        if (r0 < nin) sfor<NU>([&](auto ic) __attribute__((always_inline)) { step(ic, r0, F{}, T{}); });
    };
    if (ybase >= 0 && ybase + nin <= a.Y) pass(std::false_type{});
    else pass(std::true_type{});
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
    bool ybox = ky > 1;
    for (int t = 1; t < ky; t++) ybox = ybox && a.wy[t] == a.wy[0];
    return with_bytes_instance(kx, ky, [&](auto kxs, auto kys) {
        constexpr int KX = decltype(kxs)::value, KY = decltype(kys)::value;
        if constexpr (KY > 1) {
            if (ybox) {
                hipLaunchKernelGGL((bytes_2d_kernel<KX, KY, true>), dim3((a.waves + 3) / 4), dim3(256), 0, s, a);
                return hipGetLastError();
            }
        }
        hipLaunchKernelGGL((bytes_2d_kernel<KX, KY, false>), dim3((a.waves + 3) / 4), dim3(256), 0, s, a);
        return hipGetLastError();
    });
}

} // namespace ndl
