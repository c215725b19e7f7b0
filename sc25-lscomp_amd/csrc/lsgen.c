/*
 * lsgen.c -- seeded synthetic photon-counting detector volumes (host, OpenMP).
 *
 * Every value is a pure function of (seed, frame, pixel) through a
 * counter-based hash, so a volume is identical for any thread count and the
 * first k frames of a volume equal the k-frame volume of the same seed.
 * Peak shapes use libm exp(), so volumes are reproducible on one machine;
 * parity checks always feed the same host buffer to the oracle and the GPU.
 *
 * Kinds:
 *   LSGEN_SPARSE  Poisson(0.05) counts plus 0.1% hot pixels uniform in [200,4000].
 *   LSGEN_CSSI    Poisson(0.3) background, 50 Gaussian Bragg-like peaks per frame
 *                 (sigma 1..3 px, amplitude up to 60000, clipped to 65535), and
 *                 two 8-px-wide full-height gap stripes set to 0.
 *   LSGEN_XPCS    speckle: per-pixel rate Exp(mean 0.02) x Gamma(k=4, theta=1/4)
 *                 fixed across frames, Poisson counts per frame.
 */
#include <stdint.h>
#include <math.h>
#include <string.h>

#define LSGEN_SPARSE 1
#define LSGEN_CSSI 2
#define LSGEN_XPCS 3

static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t hash3(uint64_t seed, uint64_t a, uint64_t b)
{
    return mix64(mix64(mix64(seed) ^ a) ^ b);
}

static inline double unif(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

/* Poisson by sequential inversion; lambda is small here (<= a few). */
static inline uint32_t poisson(double lambda, double u)
{
    double p = exp(-lambda), f = p;
    uint32_t k = 0;
    while (u > f && k < 1000) {
        k++;
        p *= lambda / k;
        f += p;
    }
    return k;
}

/* Poisson by inversion against a precomputed CDF (fixed lambda). */
static inline uint32_t poisson_tab(const double *cdf, int n, double u)
{
    uint32_t k = 0;
    while ((int)k < n - 1 && u > cdf[k]) k++;
    return k;
}

static void gen_background(int kind, uint64_t seed, uint64_t frame, uint32_t y, uint32_t z,
                           uint16_t *out)
{
    uint64_t npx = (uint64_t)y * z;
    double cdf[24];
    double lam = (kind == LSGEN_SPARSE) ? 0.05 : 0.3, p = exp(-lam), f = p;
    for (int k = 0; k < 24; k++) {
        cdf[k] = f;
        p *= lam / (k + 1);
        f += p;
    }
    for (uint64_t i = 0; i < npx; i++) {
        uint64_t h = hash3(seed, frame, i);
        uint32_t v;
        if (kind == LSGEN_SPARSE) {
            uint64_t h2 = hash3(seed ^ 0x5151, frame, i);
            if (unif(h2) < 0.001)
                v = 200 + (uint32_t)(unif(mix64(h2)) * 3801.0);
            else
                v = poisson_tab(cdf, 24, unif(h));
        } else if (kind == LSGEN_CSSI) {
            v = poisson_tab(cdf, 24, unif(h));
        } else {
            /* speckle rate: fixed per pixel (frame index not hashed in) */
            uint64_t hp = hash3(seed ^ 0xA5A5, 0, i);
            double e = -log(1.0 - unif(hp)) * 0.02;
            double g = 0.0;
            for (int j = 0; j < 4; j++) g += -log(1.0 - unif(mix64(hp + 1 + j)));
            g *= 0.25;
            v = poisson(e * g, unif(h));
        }
        out[i] = (uint16_t)(v > 65535 ? 65535 : v);
    }
}

static void add_peaks(uint64_t seed, uint64_t frame, uint32_t y, uint32_t z, uint16_t *out)
{
    for (int p = 0; p < 50; p++) {
        uint64_t h = hash3(seed ^ 0xBEEF, frame, (uint64_t)p);
        double cy = unif(h) * y;
        double cz = unif(mix64(h ^ 1)) * z;
        double sigma = 1.0 + 2.0 * unif(mix64(h ^ 2));
        double amp = 100.0 + 59900.0 * unif(mix64(h ^ 3));
        int rad = (int)ceil(4.0 * sigma);
        int y0 = (int)floor(cy) - rad, y1 = (int)floor(cy) + rad;
        int z0 = (int)floor(cz) - rad, z1 = (int)floor(cz) + rad;
        for (int yy = y0; yy <= y1; yy++) {
            if (yy < 0 || yy >= (int)y) continue;
            for (int zz = z0; zz <= z1; zz++) {
                if (zz < 0 || zz >= (int)z) continue;
                double dy = yy + 0.5 - cy, dz = zz + 0.5 - cz;
                double a = amp * exp(-(dy * dy + dz * dz) / (2.0 * sigma * sigma));
                uint64_t idx = (uint64_t)yy * z + zz;
                double s = (double)out[idx] + floor(a + 0.5);
                out[idx] = (uint16_t)(s > 65535.0 ? 65535.0 : s);
            }
        }
    }
}

static void add_gaps(uint32_t y, uint32_t z, uint16_t *out)
{
    /* two 8-px-wide full-height stripes; [512,520) and [1032,1040) on a
     * 1558-wide frame, otherwise at one and two thirds of the width */
    uint32_t s0, s1;
    if (z > 1040) { s0 = 512; s1 = 1032; }
    else { s0 = (z / 3) & ~7u; s1 = (2 * z / 3) & ~7u; }
    if (z < 24) return;
    for (uint32_t yy = 0; yy < y; yy++)
        for (uint32_t k = 0; k < 8; k++) {
            if (s0 + k < z) out[(uint64_t)yy * z + s0 + k] = 0;
            if (s1 + k < z) out[(uint64_t)yy * z + s1 + k] = 0;
        }
}

/* One frame (y rows of z pixels) with global frame index `frame`. */
void lsgen_frame(int kind, uint64_t seed, uint64_t frame, uint32_t y, uint32_t z, uint16_t *out)
{
    gen_background(kind, seed, frame, y, z, out);
    if (kind == LSGEN_CSSI) {
        add_peaks(seed, frame, y, z, out);
        add_gaps(y, z, out);
    }
}

/* x frames, frame f of the volume has global frame index first_frame + f. */
void lsgen_volume(int kind, uint64_t seed, uint64_t first_frame, uint32_t x, uint32_t y,
                  uint32_t z, uint16_t *out)
{
    uint64_t fpx = (uint64_t)y * z;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t f = 0; f < (int64_t)x; f++)
        lsgen_frame(kind, seed, first_frame + (uint64_t)f, y, z, out + (uint64_t)f * fpx);
}
