// lscomp_block.h -- per-block arithmetic of the lsCOMP uint16 codec, written for
// a lane that owns one whole 8x8 block in 32 packed registers.
//
// Register layout: D[4*r + c] = v[r][2c] | v[r][2c+1] << 16 (row r of the
// block, column k = 2c or 2c+1).  Everything here is integer and bit-exact.
//
// The functions compile both for gfx950 (hipcc) and for the host (clang++),
// so the CPU test suite checks this exact code against the oracle before any
// GPU run; only v_perm_b32 and v_alignbit_b32 have host emulations.
//
// Semantics follow the reference kernel lsCOMP/src/cuLSZ_kernel.cu:
//   statistics 898-917, pooling/bin choice 919-935, rate/header 937-957,
//   plain packing 1169-1277, pooled packing 1102-1168, unpacking 1467-1627.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define LSC_HD __host__ __device__ __forceinline__
#else
#define LSC_HD inline __attribute__((always_inline))
#endif

namespace lsc {

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

LSC_HD u16x2 as_u16x2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
LSC_HD uint32_t as_u32(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// v_perm_b32: result byte i = byte sel.byte[i] of the 8-byte value {hi:lo}
// (selector 0..3 picks from lo, 4..7 from hi, 0x0c gives 0x00).  Only these
// selectors are used.
LSC_HD uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(hi, lo, sel);
#else
    uint64_t v = ((uint64_t)hi << 32) | lo;
    uint32_t r = 0;
    for (int i = 0; i < 4; i++) {
        const uint32_t s = (sel >> (8 * i)) & 0xffu;
        if (s == 0x0c) continue;   // constant zero byte
        r |= (uint32_t)((v >> (8 * (s & 7u))) & 0xffu) << (8 * i);
    }
    return r;
#endif
}

// Rotate right by s (mod 32): v_alignbit_b32(x, x, s).
LSC_HD uint32_t rotr(uint32_t x, uint32_t s)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(x, x, s);
#else
    s &= 31u;
    return s ? (x >> s) | (x << (32u - s)) : x;
#endif
}

LSC_HD uint32_t bit_width(uint32_t v) { return v ? 32u - (uint32_t)__builtin_clz(v) : 0u; }

// ---------------------------------------------------------------------------
// Division by a per-block bin: q = floor(v / d) for v < 2^16, any d >= 1,
// as q = trunc((float)v * f) with f = fl((1 + 2^-20) / d).
// Proof: f*v = (v/d)(1 + 2^-20)(1 + e), |e| <= 2^-23, so the product lies in
// [v/d, v/d + (v/d) 2^-19).  v/d <= q + 1 - 1/d and (v/d) 2^-19 < 1/d for
// v < 2^19, hence q <= f*v < q + 1.  Checked exhaustively on the CPU.
// Pooled values floor((a + b) / 2) / d are floor((a + b) / 2d), the same
// division of the pair sum (< 2^17) by 2d.
// ---------------------------------------------------------------------------
LSC_HD uint32_t quant(uint32_t v, float f) { return (uint32_t)((float)v * f); }

// ---------------------------------------------------------------------------
// Division by a per-block bin with a 24-bit multiply-high (v_mul_hi_u32_u24),
// used when every bin lies in [2, 16384]:
//   floor(s / d) = mulhi24(s * 2^a, m),  m = ceil(2^(32-a) / d),
//   mulhi24(x, y) = floor(x[23:0] * y[23:0] / 2^32).
// Proof: s*2^a*m / 2^32 = s/d + s*e / (d*2^(32-a)) with e = m*d - 2^(32-a) in
// [0, d); writing s = q*d + r, the value is q + (r + s*e/2^(32-a)) / d, and
// s*e < 2^(32-a) keeps r + s*e/2^(32-a) < r + 1 <= d, so its floor is q.
// For s < 2^S and d in (2^(j-1), 2^j], any a with 9-j <= a <= min(24-S,
// 32-S-j) gives s*2^a < 2^24, m < 2^24 and s*e < 2^(32-a); make_div24 picks
// the largest and checks the three bounds.  Plain values (S = 16) cover d in
// [2, 65536], pair sums (S = 17) d in [3, 32768], so bins in [2, 16384] cover
// every plain bin and every pooled divisor 2*bin.  Checked exhaustively on
// the CPU for every such divisor and every s.  The scaling s*2^a
// of the two halves of a packed pair is one v_dot2_u32_u16 with {2^a, 2^a}
// (the pair sum: pooled) or {2^a, 0} / {0, 2^a} (one element: plain).
// ---------------------------------------------------------------------------
LSC_HD uint32_t mulhi24(uint32_t a, uint32_t b)
{
    return (uint32_t)(((uint64_t)(a & 0xffffffu) * (uint64_t)(b & 0xffffffu)) >> 32);
}

// a.x * b.x + a.y * b.y + c over the 16-bit halves (v_dot2_u32_u16)
LSC_HD uint32_t dot2(uint32_t a, uint32_t b, uint32_t c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_udot2(as_u16x2(a), as_u16x2(b), c, false);
#else
    return (a & 0xffffu) * (b & 0xffffu) + (a >> 16) * (b >> 16) + c;
#endif
}

// x[k] for a per-block k in 0..3, as a select between values.  The empty asm
// keeps them values: the compiler would otherwise load through a selected
// address into the kernel's parameter block, which then has to live in
// scratch memory.
LSC_HD uint32_t pick4(const uint32_t x[4], uint32_t k)
{
    uint32_t a = x[0], b = x[1], c = x[2], d = x[3];
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(a), "+v"(b), "+v"(c), "+v"(d));
#endif
    return k == 0 ? a : k == 1 ? b : k == 2 ? c : d;
}

LSC_HD float pick4(const float x[4], uint32_t k)
{
    uint32_t u[4];
    for (int i = 0; i < 4; i++) u[i] = __builtin_bit_cast(uint32_t, x[i]);
    return __builtin_bit_cast(float, pick4(u, k));
}

// x[k] of a per-block k: pick4 for parameters passed in the kernel arguments
// (kMem = false), a plain indexed load for parameters that live in device
// memory (kMem = true: the batched form's item table), which needs no
// registers for the four candidates.  The kMem kernels also keep the float
// division: with the multiply-high path the batched compression kernel needs
// more than its 128 registers and spills.
template <bool kMem>
LSC_HD uint32_t sel4(const uint32_t x[4], uint32_t k)
{
    if constexpr (kMem) return x[k];
    else return pick4(x, k);
}
template <bool kMem>
LSC_HD float sel4(const float x[4], uint32_t k)
{
    if constexpr (kMem) return x[k];
    else return pick4(x, k);
}

// Fast u32 division by an invariant divisor (round-up multiplier method,
// Granlund & Montgomery 1994): q = (umulhi(n, m) + n) >> s.
struct FastDiv {
    uint32_t d, m, s, pad;
};
LSC_HD uint32_t fdiv(uint32_t n, FastDiv fd)
{
    uint32_t hi = (uint32_t)(((uint64_t)n * fd.m) >> 32);
    return (uint32_t)(((uint64_t)hi + n) >> fd.s);
}

// fd[k] for a per-block k (see pick4)
LSC_HD FastDiv pick4(const FastDiv fd[4], uint32_t k)
{
    const uint32_t d[4] = {fd[0].d, fd[1].d, fd[2].d, fd[3].d}, m[4] = {fd[0].m, fd[1].m, fd[2].m, fd[3].m};
    const uint32_t sh[4] = {fd[0].s, fd[1].s, fd[2].s, fd[3].s};
    FastDiv r{};
    r.d = pick4(d, k);
    r.m = pick4(m, k);
    r.s = pick4(sh, k);
    return r;
}

// ---------------------------------------------------------------------------
// Block statistics and decision.
// ---------------------------------------------------------------------------
struct Decide {
    uint32_t b[4];     // quantization bins
    float f[4];        // fl((1 + 2^-20) / b[k])
    float f2[4];       // fl((1 + 2^-20) / (2 b[k])): pooled pair sums
    uint32_t zth;      // pool iff zero_count >= zth (65: never)
    uint32_t flags;    // DEC_*
    FastDiv fd[4];     // exact 32-bit division by b[k] (uint32 volumes)
    // DEC_Q24: floor(v / b[k]) = mulhi24(dot2(pair, qs[k] or qs[k] << 16), qm[k])
    // for one element of a packed pair, floor((a + b) / 2b[k]) =
    // mulhi24(dot2(pair, ps[k]), pm[k]) for a pooled pair (see mulhi24)
    uint32_t qm[4], qs[4];   // plain: multiplier, 2^a (low half of the scale word)
    uint32_t pm[4], ps[4];   // pooled: multiplier, {2^a, 2^a}
};
enum : uint32_t {
    DEC_LOSSLESS = 1u,  // all bins 1 and no pooling: q == v
    DEC_POOL = 2u,      // pooling possible (zth <= 64)
    DEC_BINSEL = 4u,    // bin choice depends on the data (b0 != b3, b0 < 2^16)
    DEC_Q24 = 8u,       // every bin in [2, 16384]: 24-bit multiply-high division
};

// Compile-time specialization of the encoder for a known kind of settings
// (the caller checked p.flags): only that kind's code is compiled, which
// keeps the general case's scalar registers out of the hot loop.
enum : int {
    KIND_ANY = 0,
    KIND_LOSSLESS = 1,   // DEC_LOSSLESS
    KIND_Q24 = 2,        // DEC_Q24 (lossy, no bin of 1)
};
template <int kKind>
LSC_HD uint32_t kind_flags(const Decide &p)
{
    if (kKind == KIND_LOSSLESS) return DEC_LOSSLESS;
    if (kKind == KIND_Q24) return (p.flags & (DEC_POOL | DEC_BINSEL)) | DEC_Q24;
    return p.flags;
}

// Header byte of a block: (pool << 7) | (bin_choice << 5) | rate.  kMem: p
// lives in device memory (see sel4).  kKind: KIND_*.
template <bool kMem = false, int kKind = KIND_ANY>
LSC_HD uint32_t block_header(const uint32_t D[32], const Decide &p)
{
    const uint32_t flags = kind_flags<kKind>(p);
    // max of all 64 values (packed 16-bit max)
    u16x2 mx = as_u16x2(D[0]);
#pragma unroll
    for (int i = 1; i < 32; i++) mx = __builtin_elementwise_max(mx, as_u16x2(D[i]));
    uint32_t m1 = mx.x > mx.y ? mx.x : mx.y;
    if (flags & DEC_LOSSLESS) return (3u << 5) | bit_width(m1);

    uint32_t pool = 0, s2 = 0;
    if (flags & DEC_POOL) {
        // nonzero count: min(v,1) per half, two packed pairs summed per 32-bit
        // three-input add (no carry between the halves: each stays <= 32);
        // max pair sum
        uint32_t nzw = 0;
        uint32_t one_w = 0x00010001u;
#if defined(__HIP_DEVICE_COMPILE__)
        // an opaque 1: with a visible constant the packed min becomes a
        // compare, a select and a repack per half
        asm("" : "+v"(one_w));
#endif
        const u16x2 one = as_u16x2(one_w);
#pragma unroll
        for (int i = 0; i < 32; i += 2) {
            const u16x2 v = as_u16x2(D[i]), w = as_u16x2(D[i + 1]);
            nzw += as_u32(__builtin_elementwise_min(v, one)) + as_u32(__builtin_elementwise_min(w, one));
            const uint32_t s = (uint32_t)v.x + (uint32_t)v.y, t = (uint32_t)w.x + (uint32_t)w.y;
            const uint32_t st = s > t ? s : t;
            s2 = st > s2 ? st : s2;
        }
        const uint32_t zc = 64u - ((nzw & 0xffffu) + (nzw >> 16));
        pool = zc >= p.zth;
    }

    uint32_t bc = 3;
    if (flags & DEC_BINSEL) {
        // bin choice: bc = #{k in 1..3 : m0 >= b[k]}, m0 = min{v : v >= b0}
        // (no value in [b0,b1) <=> m0 >= b1; bins are sorted).  When m1 < b1
        // the maximum decides it: no value >= b0 (bc = 3), or m0 <= m1 < b1
        // (bc = 0); only blocks with a value >= b1 need m0 (on the GPU a wave
        // none of whose blocks has one skips the loop).  m0 is the minimum of
        // (v - b0) mod 2^16, which is < 2^16 - b0 iff some v >= b0, plus b0.
        bc = m1 < p.b[0] ? 3u : 0u;
        if (m1 >= p.b[1]) {
            const u16x2 b0 = {(unsigned short)p.b[0], (unsigned short)p.b[0]};
            u16x2 mt = as_u16x2(D[0]) - b0;
#pragma unroll
            for (int i = 1; i < 32; i++) mt = __builtin_elementwise_min(mt, as_u16x2(D[i]) - b0);
            const uint32_t t = mt.x < mt.y ? mt.x : mt.y;
            const uint32_t m0 = (t <= 65535u - p.b[0]) ? t + p.b[0] : 0xffffffffu;
            bc = (uint32_t)(m0 >= p.b[1]) + (uint32_t)(m0 >= p.b[2]) + (uint32_t)(m0 >= p.b[3]);
        }
    }
    // max floor(floor((a+b)/2) / bin) = floor(max(a+b) / 2bin)
    uint32_t maxq;
    if (!kMem && (flags & DEC_Q24)) {
        const uint32_t sc = pool ? sel4<kMem>(p.ps, bc) & 0xffffu : sel4<kMem>(p.qs, bc);
        maxq = mulhi24((pool ? s2 : m1) * sc, pool ? sel4<kMem>(p.pm, bc) : sel4<kMem>(p.qm, bc));
    } else {
        const float f = pool ? sel4<kMem>(p.f2, bc) : sel4<kMem>(p.f, bc);
        maxq = quant(pool ? s2 : m1, f);
    }
    return (pool << 7) | (bc << 5) | bit_width(maxq);
}

LSC_HD uint32_t header_payload_bytes(uint32_t hdr)
{
    return (hdr & 0x1fu) * ((hdr & 0x80u) ? 4u : 8u);
}

// ---------------------------------------------------------------------------
// Plain blocks.  Column words: CL[k] (rows 0..3) and CL[8+k] (rows 4..7) hold
// the low bytes of q[row][k], one row per byte; CH the high bytes.  Plane i
// (8 bytes, byte r = row r, bit 7-k = bit i of q[r][k]) is then
//   OR_k rotr(C[k], i + k - 7) & (0x01010101 << (7 - k))   per 4-row half.
// ---------------------------------------------------------------------------
// Column words of the low bytes (hi = false) or the high bytes (hi = true).
LSC_HD void columns_from_rows(const uint32_t D[32], uint32_t C[16], bool hi)
{
    const uint32_t s01 = hi ? 0x07030501u : 0x06020400u;
#pragma unroll
    for (int h = 0; h < 2; h++)           // row half: rows 4h..4h+3
#pragma unroll
        for (int c = 0; c < 4; c++) {     // column pair (2c, 2c+1)
            const uint32_t d0 = D[4 * (4 * h + 0) + c], d1 = D[4 * (4 * h + 1) + c];
            const uint32_t d2 = D[4 * (4 * h + 2) + c], d3 = D[4 * (4 * h + 3) + c];
            // (row0 k, row1 k, row0 k+1, row1 k+1) of the selected byte
            const uint32_t x01 = perm(d1, d0, s01), x23 = perm(d3, d2, s01);
            C[8 * h + 2 * c] = perm(x23, x01, 0x05040100u);
            C[8 * h + 2 * c + 1] = perm(x23, x01, 0x07060302u);
        }
}

LSC_HD void rows_from_columns(const uint32_t CL[16], const uint32_t CH[16], uint32_t D[32])
{
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const uint32_t le = CL[8 * h + 2 * c], lo = CL[8 * h + 2 * c + 1];
            const uint32_t he = CH[8 * h + 2 * c], ho = CH[8 * h + 2 * c + 1];
            uint32_t e01 = perm(he, le, 0x05010400u), e23 = perm(he, le, 0x07030602u);
            uint32_t o01 = perm(ho, lo, 0x05010400u), o23 = perm(ho, lo, 0x07030602u);
            D[4 * (4 * h + 0) + c] = perm(o01, e01, 0x05040100u);
            D[4 * (4 * h + 1) + c] = perm(o01, e01, 0x07060302u);
            D[4 * (4 * h + 2) + c] = perm(o23, e23, 0x05040100u);
            D[4 * (4 * h + 3) + c] = perm(o23, e23, 0x07060302u);
        }
}

// rows_from_columns when every high byte is zero (rate <= 8)
LSC_HD void rows_from_low_columns(const uint32_t CL[16], uint32_t D[32])
{
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const uint32_t le = CL[8 * h + 2 * c], lo = CL[8 * h + 2 * c + 1];
            // selector 0x0c yields a zero byte
            uint32_t e01 = perm(0u, le, 0x0c010c00u), e23 = perm(0u, le, 0x0c030c02u);
            uint32_t o01 = perm(0u, lo, 0x0c010c00u), o23 = perm(0u, lo, 0x0c030c02u);
            D[4 * (4 * h + 0) + c] = perm(o01, e01, 0x05040100u);
            D[4 * (4 * h + 1) + c] = perm(o01, e01, 0x07060302u);
            D[4 * (4 * h + 2) + c] = perm(o23, e23, 0x05040100u);
            D[4 * (4 * h + 3) + c] = perm(o23, e23, 0x07060302u);
        }
}

// bit plane i (0..7) of 8 column words -> 4 bytes
LSC_HD uint32_t plane_from_columns(const uint32_t *C, uint32_t i)
{
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
        acc |= rotr(C[k], i + (uint32_t)k - 7u) & (0x01010101u << (7 - k));
    return acc;
}

// OR bit plane i (0..7) back into 8 column words
LSC_HD void columns_add_plane(uint32_t *C, uint32_t P, uint32_t i)
{
#pragma unroll
    for (int k = 0; k < 8; k++)
        C[k] |= rotr(P, 7u - (uint32_t)k - i) & (0x01010101u << i);
}

// All 8 planes of 8 column words at once: an 8x8 bit transpose in every byte
// lane.  Rows R[r] = C[7 - r] (bit i of a byte: bit i of the value of column
// 7 - r); row i of the transpose is plane_from_columns(C, i).  Three stages
// swap the off-diagonal blocks of 4, 2 and 1 bits between rows r and r + s,
// a shift and a bit select for each row of a pair: 48 instructions for the 8
// planes of 8 column words, where one plane alone costs 19.
LSC_HD void swap_blocks(uint32_t &a, uint32_t &b, uint32_t s, uint32_t lo)
{
    const uint32_t na = (a & lo) | ((b << s) & ~lo);
    const uint32_t nb = (b & ~lo) | ((a >> s) & lo);
    a = na;
    b = nb;
}

LSC_HD void transpose8(uint32_t R[8])
{
#pragma unroll
    for (int r = 0; r < 4; r++) swap_blocks(R[r], R[r + 4], 4, 0x0f0f0f0fu);
#pragma unroll
    for (int r = 0; r < 8; r += (r & 1) ? 3 : 1) swap_blocks(R[r], R[r + 2], 2, 0x33333333u);   // r = 0, 1, 4, 5
#pragma unroll
    for (int r = 0; r < 8; r += 2) swap_blocks(R[r], R[r + 1], 1, 0x55555555u);
}

// C[i] := plane_from_columns(C, i) for i = 0..7, in place.
LSC_HD void planes_of_columns8(uint32_t C[8])
{
    uint32_t R[8];
#pragma unroll
    for (int r = 0; r < 8; r++) R[r] = C[7 - r];
    transpose8(R);
#pragma unroll
    for (int i = 0; i < 8; i++) C[i] = R[i];
}

// The inverse, in place: P[0..7] planes (zero where absent) -> column words,
// C[k] byte m bit i = byte m bit 7-k of P[i] (columns_add_plane of every plane).
LSC_HD void columns_of_planes8(uint32_t P[8])
{
    transpose8(P);   // row c: bit i = bit c of plane i; column k is row 7 - k
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t t = P[k];
        P[k] = P[7 - k];
        P[7 - k] = t;
    }
}

// Whether to build a group of planes with the transpose (all 8 at once) or
// one plane at a time.  On the GPU the choice is made for the whole wave (a
// lane's loop over its planes runs as often as the wave's largest count):
// any active lane with `from` planes or more.  Timed on the CSSI volume:
// lossy compression 0.727 -> 0.691 ms with the transpose from 3 planes on
// (its costly waves are the few with rates of 10-14 around the bright spots);
// decompression gains 0.2 % from 5 planes on.  Lossless compression lost 5 %
// with the transpose at any threshold from 3 to 6, so it keeps the plane
// loop, and so does uint32 compression (its 206-register kernel lost 9-12 %
// with the transpose).  The host build (tests) takes the block's own count,
// or forces either form.
constexpr uint32_t kPlanesEncode = 3, kPlanesDecode = 5;
#if !defined(__HIP_DEVICE_COMPILE__)
inline int g_planes_form = 0;   // host only: 0 by count, 1 plane by plane, 2 transpose
#endif
LSC_HD bool planes_at_once(uint32_t nplanes, uint32_t from)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ballot_w64(nplanes >= from) != 0;
#else
    return g_planes_form == 2 || (g_planes_form == 0 && nplanes >= from);
#endif
}

// Quantize 64 values in place.
LSC_HD void quantize_plain(uint32_t D[32], float f)
{
#pragma unroll
    for (int i = 0; i < 32; i++) {
        uint32_t lo = quant(D[i] & 0xffffu, f), hi = quant(D[i] >> 16, f);
        D[i] = lo | (hi << 16);
    }
}

// The same with the 24-bit multiply-high: scale word sc = 2^a of the bin
// (low half), multiplier m.  Five instructions per pair of values.
LSC_HD void quantize_plain24(uint32_t D[32], uint32_t sc, uint32_t m)
{
    const uint32_t sc_hi = sc << 16;
#pragma unroll
    for (int i = 0; i < 32; i++)
        D[i] = perm(mulhi24(dot2(D[i], sc_hi, 0u), m), mulhi24(dot2(D[i], sc, 0u), m), 0x05040100u);
}

// Pooled: Q[t] = floor((v[2t] + v[2t+1]) / 2) / bin = floor((v[2t] + v[2t+1]) / 2bin),
// t = 4r + c (pair c of row r); f2 = the pooled factor of the bin.
LSC_HD void quantize_pooled(const uint32_t D[32], float f2, uint32_t Q[32])
{
#pragma unroll
    for (int t = 0; t < 32; t++) Q[t] = quant((D[t] & 0xffffu) + (D[t] >> 16), f2);
}

// The same with the 24-bit multiply-high: the pair sum scaled by 2^a in one
// dot2 (scale word {2^a, 2^a}), then the multiply-high by m.
LSC_HD void quantize_pooled24(const uint32_t D[32], uint32_t sc, uint32_t m, uint32_t Q[32])
{
#pragma unroll
    for (int t = 0; t < 32; t++) Q[t] = mulhi24(dot2(D[t], sc, 0u), m);
}

// Pooled column words: C[u] byte m = low (hi = false) or high byte of Q[8m+u].
LSC_HD void pooled_columns(const uint32_t Q[32], uint32_t C[8], bool hi)
{
    const uint32_t sel = hi ? 0x07060302u : 0x05040100u;
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const uint32_t a = Q[u], b = Q[8 + u], c = Q[16 + u], d = Q[24 + u];
        const uint32_t ab = perm(b, a, 0x05010400u);   // (a.lo, b.lo, a.hi, b.hi)
        const uint32_t cd = perm(d, c, 0x05010400u);   // (c.lo, d.lo, c.hi, d.hi)
        C[u] = perm(cd, ab, sel);                      // (a, b, c, d) low or high bytes
    }
}

// Encode the payload of one block into words out[0..]; returns the number of
// 32-bit words (2*rate plain, rate pooled).  hdr from block_header().
// Quantizes D in place for plain blocks.  kMem: p lives in device memory.
// kKind: as for block_header.
template <bool kMem = false, int kKind = KIND_ANY>
LSC_HD uint32_t encode_block(uint32_t D[32], uint32_t hdr, const Decide &p, uint32_t *out)
{
    const uint32_t rate = hdr & 0x1fu;
    if (rate == 0) return 0;
    const uint32_t flags = kind_flags<kKind>(p);
    const uint32_t bc = (hdr >> 5) & 3u;
    const bool lossless = kKind == KIND_LOSSLESS;
    const bool identity = lossless || sel4<kMem>(p.b, bc) == 1u;
    if (!lossless && (hdr & 0x80u)) {
        uint32_t Q[32], C[8];
        if (!kMem && (flags & DEC_Q24)) quantize_pooled24(D, sel4<kMem>(p.ps, bc), sel4<kMem>(p.pm, bc), Q);
        else quantize_pooled(D, sel4<kMem>(p.f2, bc), Q);
        pooled_columns(Q, C, false);
        const uint32_t nlo = rate < 8 ? rate : 8u;
        if (planes_at_once(nlo, kPlanesEncode)) {
            planes_of_columns8(C);
#pragma unroll
            for (uint32_t i = 0; i < 8; i++)
                if (i < nlo) out[i] = C[i];
        } else {
            for (uint32_t i = 0; i < nlo; i++) out[i] = plane_from_columns(C, i);
        }
        if (rate > 8) {
            pooled_columns(Q, C, true);
            const uint32_t nhi = rate - 8;
            if (planes_at_once(nhi, kPlanesEncode)) {
                planes_of_columns8(C);
#pragma unroll
                for (uint32_t i = 0; i < 8; i++)
                    if (i < nhi) out[8 + i] = C[i];
            } else {
                for (uint32_t i = 8; i < rate; i++) out[i] = plane_from_columns(C, i - 8);
            }
        }
        return rate;
    }
    if (lossless) {
    } else if (!kMem && (flags & DEC_Q24)) {
        quantize_plain24(D, sel4<kMem>(p.qs, bc), sel4<kMem>(p.qm, bc));   // no bin is 1
    } else if (!identity) {
        quantize_plain(D, sel4<kMem>(p.f, bc));
    }
    uint32_t C[16];
    columns_from_rows(D, C, false);
    const uint32_t nlo = rate < 8 ? rate : 8u;
    const bool tp = !lossless && !(flags & DEC_LOSSLESS);   // lossless settings keep the plane loop
    if (tp && planes_at_once(nlo, kPlanesEncode)) {
        planes_of_columns8(C);
        planes_of_columns8(C + 8);
#pragma unroll
        for (uint32_t i = 0; i < 8; i++)
            if (i < nlo) {
                out[2 * i] = C[i];
                out[2 * i + 1] = C[8 + i];
            }
    } else {
        for (uint32_t i = 0; i < nlo; i++) {
            out[2 * i] = plane_from_columns(C, i);
            out[2 * i + 1] = plane_from_columns(C + 8, i);
        }
    }
    if (rate > 8) {
        columns_from_rows(D, C, true);
        const uint32_t nhi = rate - 8;
        if (tp && planes_at_once(nhi, kPlanesEncode)) {
            planes_of_columns8(C);
            planes_of_columns8(C + 8);
#pragma unroll
            for (uint32_t i = 0; i < 8; i++)
                if (i < nhi) {
                    out[16 + 2 * i] = C[i];
                    out[17 + 2 * i] = C[8 + i];
                }
        } else {
            for (uint32_t i = 8; i < rate; i++) {
                out[2 * i] = plane_from_columns(C, i - 8);
                out[2 * i + 1] = plane_from_columns(C + 8, i - 8);
            }
        }
    }
    return 2 * rate;
}

// Decode one block from its payload words (in[0..]) into D.  rate <= 16.
// De-quantization is modulo 2^16 like the reference's uint16 store.
LSC_HD void decode_block(const uint32_t *in, uint32_t hdr, const uint32_t bins[4], uint32_t D[32])
{
    const uint32_t rate = hdr & 0x1fu;
    const uint32_t bin = pick4(bins, (hdr >> 5) & 3u);
    if (hdr & 0x80u) {
        uint32_t CP[8] = {0, 0, 0, 0, 0, 0, 0, 0}, CPH[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const uint32_t nlo = rate < 8 ? rate : 8u, nhi = rate > 8 ? rate - 8 : 0u;
        if (planes_at_once(nlo, kPlanesDecode)) {
#pragma unroll
            for (uint32_t i = 0; i < 8; i++) CP[i] = i < nlo ? in[i] : 0u;
            columns_of_planes8(CP);
        } else {
            for (uint32_t i = 0; i < nlo; i++) columns_add_plane(CP, in[i], i);
        }
        if (nhi) {
            if (planes_at_once(nhi, kPlanesDecode)) {
#pragma unroll
                for (uint32_t i = 0; i < 8; i++) CPH[i] = i < nhi ? in[8 + i] : 0u;
                columns_of_planes8(CPH);
            } else {
                for (uint32_t i = 8; i < rate; i++) columns_add_plane(CPH, in[i], i - 8);
            }
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            // Q[8m+u] = CP[u].byte m | CPH[u].byte m << 8
            uint32_t q01 = perm(CPH[u], CP[u], 0x05010400u);   // Q[u], Q[8+u]
            uint32_t q23 = perm(CPH[u], CP[u], 0x07030602u);   // Q[16+u], Q[24+u]
            uint32_t qs[4] = {q01 & 0xffffu, q01 >> 16, q23 & 0xffffu, q23 >> 16};
#pragma unroll
            for (int m = 0; m < 4; m++) {
                uint32_t v = (qs[m] * bin) & 0xffffu;
                D[8 * m + u] = v | (v << 16);
            }
        }
        return;
    }
    uint32_t CL[16], CH[16];
#pragma unroll
    for (int j = 0; j < 16; j++) { CL[j] = 0; CH[j] = 0; }
    const uint32_t nlo = rate < 8 ? rate : 8u, nhi = rate > 8 ? rate - 8 : 0u;
    if (planes_at_once(nlo, kPlanesDecode)) {
#pragma unroll
        for (uint32_t i = 0; i < 8; i++) {
            CL[i] = i < nlo ? in[2 * i] : 0u;
            CL[8 + i] = i < nlo ? in[2 * i + 1] : 0u;
        }
        columns_of_planes8(CL);
        columns_of_planes8(CL + 8);
    } else {
        for (uint32_t i = 0; i < nlo; i++) {
            columns_add_plane(CL, in[2 * i], i);
            columns_add_plane(CL + 8, in[2 * i + 1], i);
        }
    }
    if (nhi) {
        if (planes_at_once(nhi, kPlanesDecode)) {
#pragma unroll
            for (uint32_t i = 0; i < 8; i++) {
                CH[i] = i < nhi ? in[16 + 2 * i] : 0u;
                CH[8 + i] = i < nhi ? in[17 + 2 * i] : 0u;
            }
            columns_of_planes8(CH);
            columns_of_planes8(CH + 8);
        } else {
            for (uint32_t i = 8; i < rate; i++) {
                columns_add_plane(CH, in[2 * i], i - 8);
                columns_add_plane(CH + 8, in[2 * i + 1], i - 8);
            }
        }
    }
    if (rate > 8) rows_from_columns(CL, CH, D);
    else rows_from_low_columns(CL, D);
    if (bin != 1u) {
        const u16x2 bb = {(unsigned short)bin, (unsigned short)bin};
#pragma unroll
        for (int i = 0; i < 32; i++) D[i] = as_u32(as_u16x2(D[i]) * bb);
    }
}

// ---------------------------------------------------------------------------
// uint32 blocks (cuLSZ_kernel.cu:4-824): 64 values V[8r + k], same decision
// and bit layout as uint16, with 32-bit arithmetic: the pair average wraps
// modulo 2^32 like the reference's uint sum, divisions are exact 32-bit
// (fdiv), de-quantization wraps modulo 2^32, and rates reach 31.  The planes
// are built from the values' low and high 16-bit halves packed in pairs, so
// the uint16 column-word code does the bit transposition.
// ---------------------------------------------------------------------------

// Header byte of a uint32 block.  *rate32 is set when the quantized maximum
// needs 32 bits, which the 5-bit rate field cannot hold (the reference then
// corrupts the bin choice and leaves an unwritten 256-byte hole, so such
// volumes are rejected); the returned byte is then meaningless.
LSC_HD uint32_t block_header_u32(const uint32_t V[64], const Decide &p, bool *rate32)
{
    uint32_t m1 = 0;
#pragma unroll
    for (int i = 0; i < 64; i++) m1 = V[i] > m1 ? V[i] : m1;
    uint32_t pool = 0, m2 = 0;
    if (p.flags & DEC_POOL) {
        uint32_t nz = 0;
#pragma unroll
        for (int i = 0; i < 64; i++) nz += V[i] != 0u;
#pragma unroll
        for (int t = 0; t < 32; t++) {
            const uint32_t a = (V[2 * t] + V[2 * t + 1]) >> 1;   // wraps like the reference
            m2 = a > m2 ? a : m2;
        }
        pool = (64u - nz) >= p.zth;
    }
    // bin choice: bc = #{k in 1..3 : m0 >= b[k]}, m0 = min{v : v >= b0}
    // (none: 3), from the minimum of (v - b0) mod 2^32
    uint32_t bc = 3;
    if (p.b[0] != p.b[3]) {
        uint32_t mt = 0xffffffffu;
#pragma unroll
        for (int i = 0; i < 64; i++) {
            const uint32_t t = V[i] - p.b[0];
            mt = t < mt ? t : mt;
        }
        if (mt <= 0xffffffffu - p.b[0]) {
            const uint32_t m0 = mt + p.b[0];
            bc = (uint32_t)(m0 >= p.b[1]) + (uint32_t)(m0 >= p.b[2]) + (uint32_t)(m0 >= p.b[3]);
        }
    }
    const uint32_t mx = pool ? m2 : m1;
    const uint32_t maxq = pick4(p.b, bc) == 1u ? mx : fdiv(mx, pick4(p.fd, bc));
    const uint32_t rate = bit_width(maxq);
    *rate32 = rate > 31u;
    return (pool << 7) | (bc << 5) | (rate & 0x1fu);
}

// Encode the payload of a uint32 block into words out[0..] (2*rate words
// plain, rate pooled); returns the word count.  Overwrites V.
LSC_HD uint32_t encode_block_u32(uint32_t V[64], uint32_t hdr, const Decide &p, uint32_t *out)
{
    const uint32_t rate = hdr & 0x1fu;
    if (rate == 0) return 0;
    const uint32_t bc = (hdr >> 5) & 3u;
    const bool identity = pick4(p.b, bc) == 1u;
    const FastDiv fd = pick4(p.fd, bc);
    if (hdr & 0x80u) {
        uint32_t Q[32], C[8];
#pragma unroll
        for (int t = 0; t < 32; t++) {
            const uint32_t a = (V[2 * t] + V[2 * t + 1]) >> 1;
            Q[t] = identity ? a : fdiv(a, fd);
        }
        for (uint32_t g = 0; 8 * g < rate; g++) {   // planes 8g .. 8g+7: byte g of each value
            if (g == 2) {
#pragma unroll
                for (int t = 0; t < 32; t++) Q[t] >>= 16;
            }
            pooled_columns(Q, C, (g & 1u) != 0);
            const uint32_t end = rate < 8 * g + 8 ? rate : 8 * g + 8;
            for (uint32_t i = 8 * g; i < end; i++) out[i] = plane_from_columns(C, i - 8 * g);
        }
        return rate;
    }
    if (!identity) {
#pragma unroll
        for (int e = 0; e < 64; e++) V[e] = fdiv(V[e], fd);
    }
    // D[i] = pair i (elements 2i, 2i+1) of the low (first pass) or high halves
    uint32_t D[32], C[16];
#pragma unroll
    for (int i = 0; i < 32; i++) D[i] = perm(V[2 * i + 1], V[2 * i], 0x05040100u);
    for (uint32_t g = 0; 8 * g < rate; g++) {
        if (g == 2) {
#pragma unroll
            for (int i = 0; i < 32; i++) D[i] = perm(V[2 * i + 1], V[2 * i], 0x07060302u);
        }
        columns_from_rows(D, C, (g & 1u) != 0);
        const uint32_t end = rate < 8 * g + 8 ? rate : 8 * g + 8;
        for (uint32_t i = 8 * g; i < end; i++) {
            out[2 * i] = plane_from_columns(C, i - 8 * g);
            out[2 * i + 1] = plane_from_columns(C + 8, i - 8 * g);
        }
    }
    return 2 * rate;
}

// Decode a uint32 block from its payload words.  rate <= 31.
LSC_HD void decode_block_u32(const uint32_t *in, uint32_t hdr, const uint32_t bins[4], uint32_t V[64])
{
    const uint32_t rate = hdr & 0x1fu;
    const uint32_t bin = pick4(bins, (hdr >> 5) & 3u);
#pragma unroll
    for (int e = 0; e < 64; e++) V[e] = 0;
    if (rate == 0) return;
    if (hdr & 0x80u) {
        for (uint32_t g = 0; 8 * g < rate; g++) {
            uint32_t CP[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            const uint32_t n = rate - 8 * g < 8u ? rate - 8 * g : 8u;   // planes of this group
            if (planes_at_once(n, kPlanesDecode)) {
#pragma unroll
                for (uint32_t i = 0; i < 8; i++) CP[i] = i < n ? in[8 * g + i] : 0u;
                columns_of_planes8(CP);
            } else {
                for (uint32_t i = 0; i < n; i++) columns_add_plane(CP, in[8 * g + i], i);
            }
#pragma unroll
            for (int u = 0; u < 8; u++)
#pragma unroll
                for (int m = 0; m < 4; m++)   // CP[u] byte m = byte g of Q[8m + u]
                    V[2 * (8 * m + u)] |= ((CP[u] >> (8 * m)) & 0xffu) << (8 * g);
        }
#pragma unroll
        for (int t = 0; t < 32; t++) {
            const uint32_t v = V[2 * t] * bin;
            V[2 * t] = v;
            V[2 * t + 1] = v;
        }
        return;
    }
    uint32_t D[32];
    for (uint32_t h = 0; 16 * h < rate; h++) {   // low then high 16-bit halves
        uint32_t CL[16], CH[16];
#pragma unroll
        for (int j = 0; j < 16; j++) { CL[j] = 0; CH[j] = 0; }
        // planes 16h .. 16h+7 (low bytes) and 16h+8 .. 16h+15 (high bytes)
        const uint32_t nl = rate - 16 * h < 8u ? rate - 16 * h : 8u;
        const uint32_t nh = rate > 16 * h + 8 ? (rate - 16 * h - 8 < 8u ? rate - 16 * h - 8 : 8u) : 0u;
        const uint32_t *inl = in + 32 * h, *inh = in + 32 * h + 16;
        if (planes_at_once(nl, kPlanesDecode)) {
#pragma unroll
            for (uint32_t i = 0; i < 8; i++) {
                CL[i] = i < nl ? inl[2 * i] : 0u;
                CL[8 + i] = i < nl ? inl[2 * i + 1] : 0u;
            }
            columns_of_planes8(CL);
            columns_of_planes8(CL + 8);
        } else {
            for (uint32_t i = 0; i < nl; i++) {
                columns_add_plane(CL, inl[2 * i], i);
                columns_add_plane(CL + 8, inl[2 * i + 1], i);
            }
        }
        if (nh) {
            if (planes_at_once(nh, kPlanesDecode)) {
#pragma unroll
                for (uint32_t i = 0; i < 8; i++) {
                    CH[i] = i < nh ? inh[2 * i] : 0u;
                    CH[8 + i] = i < nh ? inh[2 * i + 1] : 0u;
                }
                columns_of_planes8(CH);
                columns_of_planes8(CH + 8);
            } else {
                for (uint32_t i = 0; i < nh; i++) {
                    columns_add_plane(CH, inh[2 * i], i);
                    columns_add_plane(CH + 8, inh[2 * i + 1], i);
                }
            }
        }
        rows_from_columns(CL, CH, D);
#pragma unroll
        for (int i = 0; i < 32; i++) {   // pair i: elements 2i (low half of D) and 2i+1
            V[2 * i] |= (D[i] & 0xffffu) << (16 * h);
            V[2 * i + 1] |= (D[i] >> 16) << (16 * h);
        }
    }
    if (bin != 1u) {
#pragma unroll
        for (int e = 0; e < 64; e++) V[e] *= bin;
    }
}

}  // namespace lsc
