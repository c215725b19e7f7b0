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

// Decision parameters for bins b[0..3] (sorted, > 0) and pooling threshold p.
inline Decide make_decide(const uint32_t bins[4], float p)
{
    Decide d{};
    for (int k = 0; k < 4; k++) {
        d.b[k] = bins[k];
        d.f[k] = (float)((1.0 + 0x1p-20) / (double)bins[k]);
        d.f2[k] = (float)((1.0 + 0x1p-20) / (2.0 * (double)bins[k]));
        d.fd[k] = make_fastdiv(bins[k]);
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
    d.flags = fl;
    return d;
}

}  // namespace lsc
