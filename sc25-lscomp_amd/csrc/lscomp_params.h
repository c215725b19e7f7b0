// lscomp_params.h -- host-side derivation of kernel parameters (shared by the
// library and the CPU test harness, so both use the same code).
#pragma once
#include <stdint.h>
#include "lscomp_block.h"

namespace lsc {

// Round-up multiplier for fdiv(): s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1.
inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f{};
    uint32_t s = 0;
    while (s < 32 && (1ull << s) < d) s++;
    f.d = d;
    f.s = s;
    f.m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << s) - d)) / d + 1);
    return f;
}

// Parameters of floor(s / d) = mulhi24(s * 2^a, m) for every s < 2^sbits
// (lscomp_block.h): the largest shift a with s * 2^a < 2^24, m =
// ceil(2^(32-a) / d) < 2^24 and s * (m*d - 2^(32-a)) < 2^(32-a) for the
// largest s.  Returns false when no shift satisfies the three bounds.
inline bool make_div24(uint64_t d, uint32_t sbits, uint32_t &m_out, uint32_t &a_out)
{
    if (d < 2) return false;
    const uint64_t smax = (1ull << sbits) - 1;
    for (int a = 24 - (int)sbits; a >= 0; a--) {
        const uint64_t two = 1ull << (32 - a);
        const uint64_t m = (two + d - 1) / d;
        const uint64_t e = m * d - two;
        if ((smax << a) < (1ull << 24) && m < (1ull << 24) && smax * e < two) {
            m_out = (uint32_t)m;
            a_out = (uint32_t)a;
            return true;
        }
    }
    return false;
}

// Decision parameters for bins b[0..3] (sorted, > 0) and pooling threshold p.
inline Decide make_decide(const uint32_t bins[4], float p)
{
    Decide d{};
    bool q24 = true;
    for (int k = 0; k < 4; k++) {
        d.b[k] = bins[k];
        d.f[k] = (float)((1.0 + 0x1p-20) / (double)bins[k]);
        d.f2[k] = (float)((1.0 + 0x1p-20) / (2.0 * (double)bins[k]));
        d.fd[k] = make_fastdiv(bins[k]);
        // plain values < 2^16 divided by the bin, pooled pair sums < 2^17 by twice the bin
        uint32_t m = 0, a = 0;
        if (bins[k] <= 16384u && make_div24(bins[k], 16, m, a)) {
            d.qm[k] = m;
            d.qs[k] = 1u << a;
        } else {
            q24 = false;
        }
        if (bins[k] <= 16384u && make_div24(2ull * bins[k], 17, m, a)) {
            d.pm[k] = m;
            d.ps[k] = (1u << a) | (1u << (16 + a));
        } else {
            q24 = false;
        }
    }
    // pool iff (float)zero_count / 64 > p, as in cuLSZ_kernel.cu:920-921
    d.zth = 65;
    for (uint32_t z = 0; z <= 64; z++) {
        volatile float sp = (float)z / 64;
        if (sp > p) { d.zth = z; break; }
    }
    uint32_t fl = 0;
    if (d.zth <= 64) fl |= DEC_POOL;
    if (bins[0] != bins[3] && bins[0] <= 65535u) fl |= DEC_BINSEL;
    if (bins[0] == 1 && bins[1] == 1 && bins[2] == 1 && bins[3] == 1 && !(fl & DEC_POOL))
        fl |= DEC_LOSSLESS;
    if (q24) fl |= DEC_Q24;
    d.flags = fl;
    return d;
}

}  // namespace lsc
