// Host build of the kernels' per-block code (lscomp_block.h, lscomp_params.h)
// for the CPU test suite: the same source the gfx950 kernels compile, driven
// one block at a time so it can be compared with the oracle without a GPU.
#include <stdint.h>
#include "lscomp_block.h"
#include "lscomp_params.h"

using namespace lsc;

static void pack_rows(const uint16_t v[64], uint32_t D[32])
{
    for (int r = 0; r < 8; r++)
        for (int c = 0; c < 4; c++)
            D[4 * r + c] = (uint32_t)v[8 * r + 2 * c] | ((uint32_t)v[8 * r + 2 * c + 1] << 16);
}

extern "C" {

// Form of the bit-plane code: 0 by each block's plane count (as a wave of
// equal blocks on the GPU), 1 one plane at a time, 2 all 8 planes of a group
// at once (what a block gets on the GPU when another lane of its wave needs
// 3 planes or more).
void bh_set_planes_form(int form) { g_planes_form = form; }

// Header byte and payload bytes of one block; payload written to out (<= 128 B).
uint32_t bh_encode(const uint16_t v[64], const uint32_t bins[4], float p, uint8_t *out, uint32_t *hdr)
{
    Decide d = make_decide(bins, p);
    uint32_t D[32];
    pack_rows(v, D);
    uint32_t h = block_header(D, d);
    uint32_t words[32];
    uint32_t n = encode_block(D, h, d, words);
    for (uint32_t i = 0; i < n; i++)
        for (int b = 0; b < 4; b++) out[4 * i + b] = (uint8_t)(words[i] >> (8 * b));
    *hdr = h;
    return 4 * n;
}

void bh_decode(uint32_t hdr, const uint8_t *payload, const uint32_t bins[4], uint16_t v[64])
{
    uint32_t words[32] = {0};
    uint32_t n = header_payload_bytes(hdr) / 4;
    for (uint32_t i = 0; i < n && i < 32; i++)
        words[i] = (uint32_t)payload[4 * i] | ((uint32_t)payload[4 * i + 1] << 8) |
                   ((uint32_t)payload[4 * i + 2] << 16) | ((uint32_t)payload[4 * i + 3] << 24);
    uint32_t D[32];
    if ((hdr & 0x1f) == 0) {
        for (int i = 0; i < 32; i++) D[i] = 0;
    } else {
        decode_block(words, hdr, bins, D);
    }
    for (int r = 0; r < 8; r++)
        for (int c = 0; c < 4; c++) {
            v[8 * r + 2 * c] = (uint16_t)(D[4 * r + c] & 0xffff);
            v[8 * r + 2 * c + 1] = (uint16_t)(D[4 * r + c] >> 16);
        }
}

// uint32 block: header byte (*rate32 set when the rate would be 32) and payload.
uint32_t bh_encode_u32(const uint32_t v[64], const uint32_t bins[4], float p, uint8_t *out, uint32_t *hdr,
                       uint32_t *rate32)
{
    Decide d = make_decide(bins, p);
    uint32_t V[64];
    for (int i = 0; i < 64; i++) V[i] = v[i];
    bool r32 = false;
    uint32_t h = block_header_u32(V, d, &r32);
    *rate32 = r32 ? 1u : 0u;
    *hdr = h;
    if (r32) return 0;
    uint32_t words[62];
    uint32_t n = encode_block_u32(V, h, d, words);
    for (uint32_t i = 0; i < n; i++)
        for (int b = 0; b < 4; b++) out[4 * i + b] = (uint8_t)(words[i] >> (8 * b));
    return 4 * n;
}

void bh_decode_u32(uint32_t hdr, const uint8_t *payload, const uint32_t bins[4], uint32_t v[64])
{
    uint32_t words[62] = {0};
    uint32_t n = header_payload_bytes(hdr) / 4;
    for (uint32_t i = 0; i < n && i < 62; i++)
        words[i] = (uint32_t)payload[4 * i] | ((uint32_t)payload[4 * i + 1] << 8) |
                   ((uint32_t)payload[4 * i + 2] << 16) | ((uint32_t)payload[4 * i + 3] << 24);
    decode_block_u32(words, hdr, bins, v);
}

uint32_t bh_fdiv(uint32_t n, uint32_t d) { return fdiv(n, make_fastdiv(d)); }

// Counts (v, d) pairs with v in [0, 65536) and d in [d_lo, d_hi] where the
// kernel's division q = trunc((float)v * f_d) differs from v / d.
uint64_t bh_quant_mismatches(uint32_t d_lo, uint32_t d_hi)
{
    uint64_t bad = 0;
    for (uint64_t d = d_lo; d <= d_hi; d++) {
        const uint32_t b4[4] = {(uint32_t)d, (uint32_t)d, (uint32_t)d, (uint32_t)d};
        const float f = make_decide(b4, 1.0f).f[0];
        for (uint32_t v = 0; v < 65536; v++) bad += quant(v, f) != v / (uint32_t)d;
    }
    return bad;
}

// Counts the divisors d in [d_lo, d_hi] for which the pooled division
// q = trunc((float)s * f2_d) differs from s / 2d for some pair sum s < 2^17.
// trunc(s * f2) is non-decreasing in s, so it equals s / 2d everywhere iff it
// does at both ends of every step [2dk, 2d(k+1) - 1].
uint64_t bh_quant2_mismatches(uint32_t d_lo, uint32_t d_hi)
{
    uint64_t bad = 0;
    for (uint64_t d = d_lo; d <= d_hi; d++) {
        const uint32_t b4[4] = {(uint32_t)d, (uint32_t)d, (uint32_t)d, (uint32_t)d};
        const float f2 = make_decide(b4, 1.0f).f2[0];
        const uint64_t step = 2 * d;
        for (uint64_t lo = 0; lo < (1u << 17); lo += step) {
            const uint64_t hi = lo + step - 1 < (1u << 17) - 1 ? lo + step - 1 : (1u << 17) - 1;
            bad += quant((uint32_t)lo, f2) != lo / step;
            bad += quant((uint32_t)hi, f2) != hi / step;
        }
    }
    return bad;
}

// The 24-bit multiply-high division as the kernels run it (quantize_plain24 /
// quantize_pooled24 on packed pairs).  Counts mismatches against exact
// division over every plain value v < 2^16 (divisor d) and every pair sum
// s < 2^17 (divisor 2d), for each bin d in [d_lo, d_hi]; a bin for which
// make_decide does not choose the multiply-high path counts 2^40.
uint64_t bh_q24_mismatches(uint32_t d_lo, uint32_t d_hi)
{
    uint64_t bad = 0;
    for (uint64_t d = d_lo; d <= d_hi; d++) {
        const uint32_t b4[4] = {(uint32_t)d, (uint32_t)d, (uint32_t)d, (uint32_t)d};
        const Decide dc = make_decide(b4, 1.0f);
        if (!(dc.flags & DEC_Q24)) {
            bad += 1ull << 40;
            continue;
        }
        uint32_t D[32], Q[32];
        for (uint32_t v0 = 0; v0 < 65536; v0 += 32) {   // plain: v in the low half, 65535 - v in the high half
            for (int i = 0; i < 32; i++) D[i] = (v0 + i) | ((65535u - (v0 + i)) << 16);
            quantize_plain24(D, dc.qs[0], dc.qm[0]);
            for (int i = 0; i < 32; i++) {
                bad += (D[i] & 0xffffu) != (v0 + i) / d;
                bad += (D[i] >> 16) != (65535u - (v0 + i)) / d;
            }
        }
        for (uint32_t s0 = 0; s0 < (1u << 17); s0 += 32) {   // pooled: every pair sum s <= 2 * 65535
            for (int i = 0; i < 32; i++) {                    // split into two values (the last lanes repeat s0)
                const uint32_t s = s0 + i <= 131070u ? s0 + i : s0, a = s < 65535u ? s : 65535u;
                D[i] = a | ((s - a) << 16);
            }
            quantize_pooled24(D, dc.ps[0], dc.pm[0], Q);
            for (int i = 0; i < 32; i++) bad += Q[i] != ((s0 + i <= 131070u ? s0 + i : s0) / (2 * d));
        }
    }
    return bad;
}

uint32_t bh_decide_flags(const uint32_t bins[4], float p) { return make_decide(bins, p).flags; }

uint32_t bh_zth(float p)
{
    const uint32_t b4[4] = {1, 1, 1, 1};
    return make_decide(b4, p).zth;
}

}  // extern "C"
