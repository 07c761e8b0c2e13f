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
// wave index -> (strip, row chunk, batch offsets); strips fastest
__device__ __forceinline__ void decode_wave(const ByteArgs& a, uint32_t wave, int& strip, int& chunk, int64_t& soff,
                                            int64_t& doff)
{
    uint32_t q = a.fstrips.div(wave);
    strip = (int)(wave - q * (uint32_t)a.strips);
    uint32_t b = a.fchunks.div(q);
    chunk = (int)(q - b * (uint32_t)a.chunks);
    soff = 0;
    doff = 0;
    for (int d = 0; d < a.nb; d++) {
        const uint32_t qq = a.bext[d].div(b);
        const int64_t c = (int64_t)(b - qq * (uint32_t)a.bext[d].d);
        b = qq;
        soff += c * a.sb[d];
        doff += c * a.db[d];
    }
}

template <int KX, int KY, bool YBOX>
__global__ __launch_bounds__(256) void bytes_2d_kernel(const ByteArgs a)
{
    static_assert(KX >= 1 && KX <= 9 && KY >= 1 && KY <= 9, "reach up to four pixels");
    constexpr int CX = KX / 2; // kernelCenter (detail.h:155-163); the host routes other centres elsewhere
    const uint32_t wave = blockIdx.x * 4u + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wave >= a.waves) return; // whole wave
    const int lane = threadIdx.x & 63;
    int strip, chunk;
    int64_t soff, doff;
    decode_wave(a, wave, strip, chunk, soff, doff);
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

// ------------------------------------------------------ interleaved colour
// {3, w, h} uint8 (channel fastest, x stride 3 bytes), x and at most one
// more axis convolved, the channel untouched.  A lane owns four pixels
// (12 bytes, one dwordx3 load), a wave a 256-pixel strip.  The 12 bytes are
// split (v_perm) into per-channel pixel pairs (r0 r1) (r2 r3) (g0 g1) ...
// of packed 16-bit values, so the x taps act on whole pairs: the pairs of
// the pixels left and right of a lane's four come from the neighbouring
// lanes (DPP wave shifts), and lanes 0 and 63 take theirs from an edge load
// of the four pixels beyond the strip, folded per pixel at the image edges
// (a per-lane v_perm selector: keep, reverse for reflect, repeat the first
// or last pixel for clamp; wrap only moves the address).  A window that
// starts at an odd pixel is a pair straddling two aligned pairs
// (v_alignbyte by 2 bytes).  x and y sums are packed 16-bit multiply-adds,
// exact modulo 2^16, which keeps the low byte (as in the planar kernel);
// the four pixels' low bytes are packed back into 12 bytes for one dwordx3
// store.  Row loop, prefetch, chunks and the overlapping last strip as in
// the planar kernel.

// fold selectors for v_perm(pair (c d), pair (a b)) of a four-pixel group
// (a b c d): first output pair, second output pair
constexpr uint32_t kG01Keep = 0x03020100u, kG23Keep = 0x07060504u;       // (a b) (c d)
constexpr uint32_t kG01Rev = 0x05040706u, kG23Rev = 0x01000302u;         // (d c) (b a)
constexpr uint32_t kGFirst = 0x01000100u, kGLast = 0x07060706u;          // (a a) (a a) / (d d) (d d)

typedef unsigned int u32x3 __attribute__((ext_vector_type(3)));
typedef unsigned short u16x2r __attribute__((ext_vector_type(2)));

// 12 bytes r0 g0 b0 r1 | g1 b1 r2 g2 | b2 r3 g3 b3 -> pairs [channel][0: pixels 0,1; 1: pixels 2,3]
__device__ __forceinline__ void unpack_rgb(const u32x3 v, uint32_t (&p)[3][2])
{
    p[0][0] = __builtin_amdgcn_perm(v.x, v.x, 0x0c030c00u); // r0 r1
    p[0][1] = __builtin_amdgcn_perm(v.z, v.y, 0x0c050c02u); // r2 r3
    p[1][0] = __builtin_amdgcn_perm(v.y, v.x, 0x0c040c01u); // g0 g1
    p[1][1] = __builtin_amdgcn_perm(v.z, v.y, 0x0c060c03u); // g2 g3
    p[2][0] = __builtin_amdgcn_perm(v.y, v.x, 0x0c050c02u); // b0 b1
    p[2][1] = __builtin_amdgcn_perm(v.z, v.z, 0x0c030c00u); // b2 b3
}

// low bytes of the pairs (byte 0 and byte 2 of each) back into 12 bytes
__device__ __forceinline__ u32x3 pack_rgb(const uint32_t (&p)[3][2])
{
    const uint32_t R0 = p[0][0], R1 = p[0][1], G0 = p[1][0], G1 = p[1][1], B0 = p[2][0], B1 = p[2][1];
    u32x3 o;
    const uint32_t t0 = __builtin_amdgcn_perm(G0, R0, 0x0c020400u); // r0 g0 r1 .
    o.x = __builtin_amdgcn_perm(B0, t0, 0x02040100u);                // r0 g0 b0 r1
    const uint32_t u1 = __builtin_amdgcn_perm(B0, G0, 0x0c0c0602u); // g1 b1 . .
    const uint32_t u2 = __builtin_amdgcn_perm(G1, R1, 0x04000c0cu); // . . r2 g2
    o.y = u1 | u2;                                                   // g1 b1 r2 g2
    const uint32_t v1 = __builtin_amdgcn_perm(G1, R1, 0x0c06020cu); // . r3 g3 .
    o.z = __builtin_amdgcn_perm(B1, v1, 0x06020104u);                // b2 r3 g3 b3
    return o;
}

template <int KX, int KY, bool YBOX>
__global__ __launch_bounds__(256) void bytes_rgb_kernel(const ByteArgs a)
{
    static_assert(KX >= 3 && KX <= 9 && KY >= 1 && KY <= 9, "reach up to four pixels");
    constexpr int CX = KX / 2; // kernelCenter (detail.h:155-163); the host routes other centres elsewhere
    const uint32_t wave = blockIdx.x * 4u + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wave >= a.waves) return; // whole wave
    const int lane = threadIdx.x & 63;
    int strip, chunk;
    int64_t soff, doff;
    decode_wave(a, wave, strip, chunk, soff, doff);
    const int ps = min(strip * 256, a.X - 256); // the last strip ends at X (see the planar kernel)
    const int xb = 3 * (ps + 4 * lane);           // own 12 bytes
    const int y0 = chunk * a.rows;
    const int ny = min(a.rows, a.Y - y0);
    const int nin = ny + KY - 1;

    const auto rs = __builtin_amdgcn_make_buffer_rsrc((void*)(a.src + soff), (short)0, 0x7fffffff, 0x00020000);
    const auto rd = __builtin_amdgcn_make_buffer_rsrc((void*)(a.dst + doff), (short)0, 0x7fffffff, 0x00020000);

    // the four pixels left / right of the strip (lane 0 / lane 63), folded at
    // the image edges (single fold: reach <= 4 <= X)
    int gl = ps - 4, gr = ps + 256;
    uint32_t fl0 = kG01Keep, fl1 = kG23Keep, fr0 = kG01Keep, fr1 = kG23Keep;
    if (gl < 0) {
        if (a.bx == NDL_WRAP) gl += a.X;
        else if (a.bx == NDL_REFLECT) { gl = ps; fl0 = kG01Rev; fl1 = kG23Rev; }
        else { gl = ps; fl0 = fl1 = kGFirst; }
    }
    if (gr >= a.X) {
        if (a.bx == NDL_WRAP) gr -= a.X;
        else if (a.bx == NDL_REFLECT) { gr = a.X - 4; fr0 = kG01Rev; fr1 = kG23Rev; }
        else { gr = a.X - 4; fr0 = fr1 = kGLast; }
    }
    // lanes 1-31 read lane 0's edge group and lanes 32-62 lane 63's (two lines)
    // (separate selectors for the left and right shifts: each fold then
    // writes the shift's own register, which the DPP move updates in place)
    const int xe = 3 * (lane < 32 ? gl : gr);
    const uint32_t sl0 = lane == 0 ? fl0 : kG01Keep, sl1 = lane == 0 ? fl1 : kG23Keep;
    const uint32_t sr0 = lane == 63 ? fr0 : kG01Keep, sr1 = lane == 63 ? fr1 : kG23Keep;

    uint32_t ring[KY][3][2]; // x sums, packed 16-bit pairs modulo 2^16
#pragma unroll
    for (int t = 0; t < KY; t++)
#pragma unroll
        for (int c = 0; c < 3; c++) ring[t][c][0] = ring[t][c][1] = 0u;
    u16x2r run[3][2]; // YBOX: sum of the ring
#pragma unroll
    for (int c = 0; c < 3; c++) run[c][0] = run[c][1] = u16x2r{0, 0};

    const int ybase = y0 - a.cy;
    const int obase = (y0 - (KY - 1)) * a.dy;

    auto pass = [&](auto edge_tag) __attribute__((always_inline)) {
        constexpr bool EDGE = decltype(edge_tag)::value;
        constexpr int PD = bytes_rgb_prefetch_rows(KY);
        constexpr int NU = PD;
        static_assert(NU % KY == 0, "the ring slot of a row is fixed in a block");
        auto row_off = [&](int r) -> int {
            if constexpr (EDGE) return border_map32(a.by, a.Y, ybase + r) * a.sy;
            else return (ybase + r) * a.sy;
        };
        u32x3 bo[PD], be[PD];
        auto load_row = [&](int r, int slot) __attribute__((always_inline)) {
            const int ro = row_off(r);
            bo[slot] = __builtin_amdgcn_raw_buffer_load_b96(rs, xb, ro, 0);
            be[slot] = __builtin_amdgcn_raw_buffer_load_b96(rs, xe, ro, 0);
            __builtin_amdgcn_sched_barrier(0); // rows' loads stay in row order (see the planar kernel)
        };
        sfor<PD>([&](auto ip) __attribute__((always_inline)) {
            constexpr int p = decltype(ip)::value;
            load_row(min(p, nin - 1), p);
        });
        auto step = [&](auto ic, int r0, auto first_tag, auto guard_tag) __attribute__((always_inline)) {
            constexpr int u = decltype(ic)::value;
            constexpr bool FIRST = decltype(first_tag)::value;
            constexpr bool GUARD = decltype(guard_tag)::value;
            constexpr int slot = u % PD;
            constexpr int U = u % KY;
            const int r = r0 + u;
            if constexpr (GUARD) {
                if (r >= nin) return;
            }
            uint32_t O[3][2], E[3][2];
            unpack_rgb(bo[slot], O);
            unpack_rgb(be[slot], E);
            load_row(min(r + PD, nin - 1), slot); // both slots consumed; past the end: a repeat of the last row
            // window pairs by start pixel: aligned starts -4 -2 0 2 4 6
            uint32_t Wp[3][6];
#pragma unroll
            for (int c = 0; c < 3; c++) {
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    const uint32_t el = __builtin_amdgcn_perm(E[c][1], E[c][0], k ? sl1 : sl0);
                    const uint32_t er = __builtin_amdgcn_perm(E[c][1], E[c][0], k ? sr1 : sr0);
                    Wp[c][k] = __builtin_amdgcn_update_dpp(el, O[c][k], 0x138, 0xf, 0xf, false);     // wave_shr:1
                    Wp[c][4 + k] = __builtin_amdgcn_update_dpp(er, O[c][k], 0x130, 0xf, 0xf, false); // wave_shl:1
                    Wp[c][2 + k] = O[c][k];
                }
            }
            uint32_t xs_[3][2];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                sfor<2>([&](auto iq) __attribute__((always_inline)) {
                    constexpr int q = decltype(iq)::value;
                    u16x2r acc = {0, 0};
                    sfor<KX>([&](auto it) __attribute__((always_inline)) {
                        constexpr int t = decltype(it)::value;
                        constexpr int st = 2 * q + t - CX; // window start pixel, -4 .. 6
                        uint32_t v;
                        if constexpr ((st & 1) == 0) v = Wp[c][(st + 4) / 2];
                        else v = __builtin_amdgcn_alignbyte(Wp[c][(st + 5) / 2], Wp[c][(st + 3) / 2], 2);
                        const unsigned short w = (unsigned short)a.wxs[t];
                        acc = acc + u16x2r{w, w} * __builtin_bit_cast(u16x2r, v);
                    });
                    xs_[c][q] = __builtin_bit_cast(uint32_t, acc);
                });
            }
            uint32_t out[3][2];
#pragma unroll
            for (int c = 0; c < 3; c++) {
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const u16x2r n = __builtin_bit_cast(u16x2r, xs_[c][q]);
                    if constexpr (YBOX) run[c][q] = run[c][q] - __builtin_bit_cast(u16x2r, ring[U][c][q]) + n;
                    ring[U][c][q] = xs_[c][q];
                    u16x2r acc;
                    const unsigned short w0 = (unsigned short)a.wy[0];
                    if constexpr (YBOX) {
                        acc = u16x2r{w0, w0} * run[c][q];
                    } else {
                        acc = u16x2r{0, 0};
#pragma unroll
                        for (int t = 0; t < KY; t++) {
                            const unsigned short w = (unsigned short)a.wy[t];
                            acc = acc + u16x2r{w, w} * __builtin_bit_cast(u16x2r, ring[(U + 1 + t) % KY][c][q]);
                        }
                    }
                    out[c][q] = __builtin_bit_cast(uint32_t, acc);
                }
            }
            // the first KY - 1 input rows produce no output row
            bool emit = true;
            if constexpr (FIRST && u < KY - 1) emit = false;
            if constexpr (GUARD && u < KY - 1) emit = r0 > 0;
            if (emit) __builtin_amdgcn_raw_buffer_store_b96(pack_rgb(out), rd, xb, obase + r * a.dy, 0);
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

template <class F> hipError_t with_rgb_instance(int kx, int ky, F&& f)
{
#define NDL_B(KX, KY) \
    if (kx == KX && ky == KY) return f(ICb<KX>{}, ICb<KY>{});
    NDL_B(3, 3) NDL_B(5, 5) NDL_B(7, 7) NDL_B(9, 9) NDL_B(3, 1) NDL_B(5, 1) NDL_B(7, 1) NDL_B(9, 1)
#undef NDL_B
    return hipErrorInvalidValue;
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

bool bytes_rgb_supported(int kx, int ky)
{
    return with_rgb_instance(kx, ky, [](auto, auto) { return hipSuccess; }) == hipSuccess;
}

hipError_t launch_bytes_rgb(const ByteArgs& a, int kx, int ky, hipStream_t s)
{
    if (a.waves == 0) return hipSuccess;
    bool ybox = ky > 1;
    for (int t = 1; t < ky; t++) ybox = ybox && a.wy[t] == a.wy[0];
    return with_rgb_instance(kx, ky, [&](auto kxs, auto kys) {
        constexpr int KX = decltype(kxs)::value, KY = decltype(kys)::value;
        if constexpr (KY > 1) {
            if (ybox) {
                hipLaunchKernelGGL((bytes_rgb_kernel<KX, KY, true>), dim3((a.waves + 3) / 4), dim3(256), 0, s, a);
                return hipGetLastError();
            }
        }
        hipLaunchKernelGGL((bytes_rgb_kernel<KX, KY, false>), dim3((a.waves + 3) / 4), dim3(256), 0, s, a);
        return hipGetLastError();
    });
}

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
