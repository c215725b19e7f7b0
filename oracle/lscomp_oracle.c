/*
 * lscomp_oracle.c -- CPU restatement of the lsCOMP uint16 codec.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library (liblscomp.so) never
 * links or calls it, and has no CPU fallback.
 *
 * Every function follows the reference's per-block logic; the file:line
 * citations point into the reference tree (hyfshishen/SC25-lsCOMP):
 *
 *   blocking, block ids, grid        lsCOMP/src/cuLSZ_entry.cu:93-131
 *   gather with zero padding         lsCOMP/src/cuLSZ_kernel.cu:864-896
 *   block statistics                 lsCOMP/src/cuLSZ_kernel.cu:898-917
 *   pooling + bin choice             lsCOMP/src/cuLSZ_kernel.cu:919-935
 *   rate + 1-byte header             lsCOMP/src/cuLSZ_kernel.cu:937-957
 *   payload placement (scan)         lsCOMP/src/cuLSZ_kernel.cu:962-1097
 *   quantize + bit-plane packing     lsCOMP/src/cuLSZ_kernel.cu:1099-1277
 *   decode: headers, unpack, store   lsCOMP/src/cuLSZ_kernel.cu:1322-1640
 *   SSIM of two volumes              lsCOMP/src/cuLSZ_utility.cu:757-886
 *
 * Stream format (block order b = 0..B-1, B = x*ceil(y/8)*ceil(z/8)):
 *   [header byte of every block][zero pad to R = 4*ceil(B/4)][payload(0)]...
 * Payload of a block = rate planes of 8 bytes (plain) or 4 bytes (pooled).
 *
 * Divergences from the reference, each where the reference output is
 * undefined or wrong (defined here, and identically in the HIP path):
 *   - pad bytes [B, R) are written as 0 (never written by the reference);
 *   - a rate-0 block decodes to zeros (the reference skips the store,
 *     cuLSZ_kernel.cu:1460-1461, leaving prior memory contents);
 *   - the reported size is R + sum(payload) (the reference reports
 *     sum(payload) + B, cuLSZ_entry.cu:125, and B alone for grids of one CTA);
 *   - element indices are 64-bit (the reference uses 32-bit, :857, :886).
 *
 * uint32 volumes (the reference's cuLSZ_*_uint32_bsize64, lsCOMP/src/
 * cuLSZ_kernel.cu:4-824, lsCOMP/src/cuLSZ_entry.cu:18-91) use the same
 * per-block functions: the pair average wraps modulo 2^32 like the
 * reference's uint sum (cuLSZ_kernel.cu:91, :286), de-quantization wraps
 * modulo 2^32, and rates reach 31.  A block whose quantized maximum needs 32
 * bits would get rate 32, which sets bit 5 of the header (cuLSZ_kernel.cu:
 * 120-122, 129-131) and corrupts its bin choice; the reference then writes no
 * payload but advances the offset by 256 bytes (:259-261 vs :121, :130), so
 * its stream cannot be decoded.  Here such a volume is rejected
 * (ORC_ERR_RANGE), and the worst case per block is 31 planes (248 bytes).
 *
 * The OpenMP variants (oracle_*_omp) produce the same bytes with a two-pass
 * schedule (headers + sizes, exclusive scan, pack) and are what bench.py times
 * as the CPU baseline.
 *
 * ORACLE_MUTANT=n (compile-time) plants one plausible mistake; the test suite
 * builds each mutant and checks that its pins reject it.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

#define ORC_OK 0
#define ORC_ERR_ARG 1
#define ORC_ERR_CAPACITY 2
#define ORC_ERR_CORRUPT 3
#define ORC_ERR_RANGE 4

/* ---- blocking (cuLSZ_entry.cu:97-99, cuLSZ_kernel.cu:843) ---------------- */

uint64_t oracle_num_blocks(uint32_t x, uint32_t y, uint32_t z)
{
    uint64_t by = ((uint64_t)y + 7) / 8, bz = ((uint64_t)z + 7) / 8;
    return (uint64_t)x * by * bz;
}

uint64_t oracle_header_bytes(uint64_t nblocks)
{
#if ORACLE_MUTANT == 5
    return nblocks;                 /* mutant: forgets the 4-byte pad */
#else
    return (nblocks + 3) / 4 * 4;   /* rate_ofs, cuLSZ_kernel.cu:843 */
#endif
}

uint64_t oracle_max_compressed_bytes(uint32_t x, uint32_t y, uint32_t z)
{
    uint64_t b = oracle_num_blocks(x, y, z);
    return oracle_header_bytes(b) + 128 * b;
}

/* ---- gather one 8x8 block, zero padded (cuLSZ_kernel.cu:864-896) --------- */

void oracle_gather_block(const uint16_t *vol, uint32_t x, uint32_t y, uint32_t z,
                         uint64_t b, uint32_t v[64])
{
    uint64_t nby = ((uint64_t)y + 7) / 8, nbz = ((uint64_t)z + 7) / 8;
    uint64_t per_frame = nby * nbz;
    uint64_t bx = b / per_frame;
    uint64_t by = (b % per_frame) / nbz;
    uint64_t bz = (b % per_frame) % nbz;
    (void)x;
    for (int r = 0; r < 8; r++) {
        uint64_t yy = by * 8 + r;
        for (int k = 0; k < 8; k++) {
            uint64_t zz = bz * 8 + k;
            if (yy < y && zz < z)
                v[r * 8 + k] = vol[bx * (uint64_t)y * z + yy * z + zz];
            else
                v[r * 8 + k] = 0;
        }
    }
}

static void scatter_block(uint16_t *vol, uint32_t x, uint32_t y, uint32_t z,
                          uint64_t b, const uint32_t v[64])
{
    uint64_t nby = ((uint64_t)y + 7) / 8, nbz = ((uint64_t)z + 7) / 8;
    uint64_t per_frame = nby * nbz;
    uint64_t bx = b / per_frame;
    uint64_t by = (b % per_frame) / nbz;
    uint64_t bz = (b % per_frame) % nbz;
    (void)x;
    for (int r = 0; r < 8; r++) {
        uint64_t yy = by * 8 + r;
        for (int k = 0; k < 8; k++) {
            uint64_t zz = bz * 8 + k;
            if (yy < y && zz < z)   /* cuLSZ_kernel.cu:1638 */
                vol[bx * (uint64_t)y * z + yy * z + zz] = (uint16_t)v[r * 8 + k];
        }
    }
}

/* uint32 volumes: same blocking, 4-byte elements (cuLSZ_kernel.cu:41-72). */
void oracle_gather_block_u32(const uint32_t *vol, uint32_t x, uint32_t y, uint32_t z,
                             uint64_t b, uint32_t v[64])
{
    uint64_t nby = ((uint64_t)y + 7) / 8, nbz = ((uint64_t)z + 7) / 8;
    uint64_t per_frame = nby * nbz;
    uint64_t bx = b / per_frame;
    uint64_t by = (b % per_frame) / nbz;
    uint64_t bz = (b % per_frame) % nbz;
    (void)x;
    for (int r = 0; r < 8; r++) {
        uint64_t yy = by * 8 + r;
        for (int k = 0; k < 8; k++) {
            uint64_t zz = bz * 8 + k;
            v[r * 8 + k] = (yy < y && zz < z) ? vol[bx * (uint64_t)y * z + yy * z + zz] : 0;
        }
    }
}

static void scatter_block_u32(uint32_t *vol, uint32_t y, uint32_t z, uint64_t b, const uint32_t v[64])
{
    uint64_t nby = ((uint64_t)y + 7) / 8, nbz = ((uint64_t)z + 7) / 8;
    uint64_t per_frame = nby * nbz;
    uint64_t bx = b / per_frame;
    uint64_t by = (b % per_frame) / nbz;
    uint64_t bz = (b % per_frame) % nbz;
    for (int r = 0; r < 8; r++) {
        uint64_t yy = by * 8 + r;
        for (int k = 0; k < 8; k++) {
            uint64_t zz = bz * 8 + k;
            if (yy < y && zz < z)   /* cuLSZ_kernel.cu:815 */
                vol[bx * (uint64_t)y * z + yy * z + zz] = v[r * 8 + k];
        }
    }
}

/* ---- bit width, 32 - __clz(v) with __clz(0) == 32 ------------------------ */

static uint32_t bit_width(uint32_t v)
{
    uint32_t n = 0;
    while (v) { n++; v >>= 1; }
#if ORACLE_MUTANT == 7
    if (n) n++;                     /* mutant: rate one too high */
#endif
    return n;
}

/* ---- per-block decision (cuLSZ_kernel.cu:898-957) ------------------------ */

uint8_t oracle_block_header(const uint32_t v[64], const uint32_t bins[4],
                            float pooling_th, uint32_t *payload_bytes)
{
    uint32_t zero_count = 0;
    uint32_t below[4] = {0, 0, 0, 0};
    uint32_t max_plain = 0, max_pooled = 0;
    for (int i = 0; i < 64; i++) {
        uint32_t val = v[i];
        zero_count += (val == 0);
        for (int j = 0; j < 4; j++)
#if ORACLE_MUTANT == 4
            below[j] += (val <= bins[j]);   /* mutant: <= instead of < */
#else
            below[j] += (val < bins[j]);
#endif
        if (val > max_plain) max_plain = val;
        if (i % 2) {
#if ORACLE_MUTANT == 6
            uint32_t avg = (v[i - 1] + v[i] + 1) / 2;   /* mutant: rounds up */
#else
            uint32_t avg = (v[i - 1] + v[i]) / 2;
#endif
            if (avg > max_pooled) max_pooled = avg;
        }
    }

    /* (float)zero_count / 64 > poolingTH, evaluated in float. */
    float sparsity = (float)zero_count / 64;
#if ORACLE_MUTANT == 3
    int pool = (sparsity >= pooling_th);    /* mutant: >= instead of > */
#else
    int pool = (sparsity > pooling_th);
#endif

    uint32_t bc = 0;                        /* progressive bin choice */
    if (below[1] == below[0]) {
        bc = 1;
        if (below[2] == below[1]) {
            bc = 2;
            if (below[3] == below[2]) bc = 3;
        }
    }

    uint32_t maxq = (pool ? max_pooled : max_plain) / bins[bc];
    uint32_t rate = bit_width(maxq);
    *payload_bytes = rate * (pool ? 4u : 8u);
    return (uint8_t)(((uint32_t)pool << 7) | (bc << 5) | rate);
}

static uint32_t header_payload_bytes(uint8_t hdr)
{
    uint32_t rate = hdr & 0x1f;
    return (hdr >> 7) ? rate * 4u : rate * 8u;
}

/* ---- quantize + bit-plane pack (cuLSZ_kernel.cu:1099-1277) --------------- */
/* Plane i (i = 0 is the least significant bit) is 8 bytes for a plain block:
 * byte r holds bit i of q[r][0..7] with q[r][0] in the most significant bit.
 * A pooled block averages horizontal pairs to 32 values t = 0..31 (pair
 * (r, 2c),(r, 2c+1) -> t = 4r + c) and its planes are 4 bytes: byte m holds
 * bit i of qp[8m .. 8m+7], qp[8m] in the most significant bit.               */

uint32_t oracle_encode_block(const uint32_t v[64], uint8_t hdr, const uint32_t bins[4],
                             uint8_t *out)
{
    uint32_t rate = hdr & 0x1f;
    uint32_t bin = bins[(hdr >> 5) & 3];
    uint32_t n = 0;
    if (rate == 0) return 0;
    if (hdr >> 7) {
        uint32_t qp[32];
        for (int t = 0; t < 32; t++) qp[t] = ((v[2 * t] + v[2 * t + 1]) / 2) / bin;
        for (uint32_t i = 0; i < rate; i++) {
            for (int m = 0; m < 4; m++) {
                uint32_t byte = 0;
                for (int u = 0; u < 8; u++)
                    byte |= ((qp[8 * m + u] >> i) & 1u) << (7 - u);
                out[n++] = (uint8_t)byte;
            }
        }
    } else {
        uint32_t q[64];
        for (int e = 0; e < 64; e++) q[e] = v[e] / bin;
        for (uint32_t i = 0; i < rate; i++) {
            for (int r = 0; r < 8; r++) {
                uint32_t byte = 0;
                for (int k = 0; k < 8; k++) {
#if ORACLE_MUTANT == 1
                    byte |= ((q[r * 8 + k] >> i) & 1u) << k;         /* mutant: LSB-first */
#elif ORACLE_MUTANT == 2
                    byte |= ((q[k * 8 + r] >> i) & 1u) << (7 - k);   /* mutant: transposed */
#else
                    byte |= ((q[r * 8 + k] >> i) & 1u) << (7 - k);
#endif
                }
                out[n++] = (uint8_t)byte;
            }
        }
    }
    return n;
}

/* ---- unpack + de-quantize (cuLSZ_kernel.cu:1460-1627) -------------------- */

void oracle_decode_block(uint8_t hdr, const uint8_t *payload, const uint32_t bins[4],
                         uint32_t v[64])
{
    uint32_t rate = hdr & 0x1f;
    uint32_t bin = bins[(hdr >> 5) & 3];
    for (int e = 0; e < 64; e++) v[e] = 0;
    if (rate == 0) return;
    if (hdr >> 7) {
        uint32_t qp[32] = {0};
        for (uint32_t i = 0; i < rate; i++)
            for (int m = 0; m < 4; m++) {
                uint32_t byte = payload[i * 4 + m];
                for (int u = 0; u < 8; u++)
                    qp[8 * m + u] |= ((byte >> (7 - u)) & 1u) << i;
            }
        for (int t = 0; t < 32; t++) {
            v[2 * t] = qp[t] * bin;
#if ORACLE_MUTANT == 8
            v[2 * t + 1] = 0;               /* mutant: second pair element lost */
#else
            v[2 * t + 1] = v[2 * t];
#endif
        }
    } else {
        for (uint32_t i = 0; i < rate; i++)
            for (int r = 0; r < 8; r++) {
                uint32_t byte = payload[i * 8 + r];
                for (int k = 0; k < 8; k++)
                    v[r * 8 + k] |= ((byte >> (7 - k)) & 1u) << i;
            }
        for (int e = 0; e < 64; e++) v[e] *= bin;
    }
}

/* ---- parameter domain (lsCOMP_xpcs.cpp:51-85) ---------------------------- */

int oracle_check_params(uint32_t x, uint32_t y, uint32_t z, const uint32_t bins[4],
                        float pooling_th)
{
    if (x == 0 || y == 0 || z == 0) return ORC_ERR_ARG;
    if (bins[0] == 0 || bins[1] == 0 || bins[2] == 0 || bins[3] == 0) return ORC_ERR_ARG;
    if (!(bins[0] <= bins[1] && bins[1] <= bins[2] && bins[2] <= bins[3])) return ORC_ERR_ARG;
    if (!(pooling_th >= 0.0f && pooling_th <= 1.0f)) return ORC_ERR_ARG;
    return ORC_OK;
}

/* ---- whole-volume compress: block-order concatenation -------------------- */

int oracle_compress_u16(const uint16_t *in, uint32_t x, uint32_t y, uint32_t z,
                        const uint32_t bins[4], float pooling_th,
                        uint8_t *out, uint64_t capacity, uint64_t *out_bytes)
{
    int st = oracle_check_params(x, y, z, bins, pooling_th);
    if (st) return st;
    uint64_t nb = oracle_num_blocks(x, y, z);
    uint64_t hb = oracle_header_bytes(nb);
    if (capacity < hb) { *out_bytes = hb; return ORC_ERR_CAPACITY; }
    uint64_t pos = hb;
    for (uint64_t i = nb; i < hb; i++) out[i] = 0;
    uint32_t v[64];
    uint8_t tmp[256];   /* 128 suffices for rate <= 16; slack for the mutants */
    int over = 0;
    for (uint64_t b = 0; b < nb; b++) {
        uint32_t pb;
        oracle_gather_block(in, x, y, z, b, v);
        uint8_t h = oracle_block_header(v, bins, pooling_th, &pb);
        out[b] = h;
        uint32_t n = oracle_encode_block(v, h, bins, tmp);
        if (!over && pos + n <= capacity) memcpy(out + pos, tmp, n);
        else over = 1;
        pos += n;
    }
    *out_bytes = pos;
    return over ? ORC_ERR_CAPACITY : ORC_OK;
}

int oracle_decompress_u16(const uint8_t *in, uint64_t in_bytes,
                          uint32_t x, uint32_t y, uint32_t z, const uint32_t bins[4],
                          uint16_t *out)
{
    if (oracle_check_params(x, y, z, bins, 0.0f)) return ORC_ERR_ARG;
    uint64_t nb = oracle_num_blocks(x, y, z);
    uint64_t hb = oracle_header_bytes(nb);
    if (in_bytes < hb) return ORC_ERR_CORRUPT;
    uint64_t pos = hb;
    uint32_t v[64];
    int bad = 0;
    for (uint64_t b = 0; b < nb; b++) {
        uint8_t h = in[b];
        uint32_t n = header_payload_bytes(h);
        if ((h & 0x1f) > 16 || pos + n > in_bytes) {
            bad = 1;
            memset(v, 0, sizeof v);
        } else {
            oracle_decode_block(h, in + pos, bins, v);
        }
        scatter_block(out, x, y, z, b, v);
        pos += n;
    }
    return bad ? ORC_ERR_CORRUPT : ORC_OK;
}

/* ---- uint32 volumes (cuLSZ_kernel.cu:4-824, cuLSZ_entry.cu:18-91) -------- */

uint64_t oracle_max_compressed_bytes_u32(uint32_t x, uint32_t y, uint32_t z)
{
    uint64_t b = oracle_num_blocks(x, y, z);
    return oracle_header_bytes(b) + 248 * b;
}

int oracle_compress_u32(const uint32_t *in, uint32_t x, uint32_t y, uint32_t z,
                        const uint32_t bins[4], float pooling_th,
                        uint8_t *out, uint64_t capacity, uint64_t *out_bytes)
{
    int st = oracle_check_params(x, y, z, bins, pooling_th);
    if (st) return st;
    uint64_t nb = oracle_num_blocks(x, y, z);
    uint64_t hb = oracle_header_bytes(nb);
    if (capacity < hb) { *out_bytes = hb; return ORC_ERR_CAPACITY; }
    uint64_t pos = hb;
    for (uint64_t i = nb; i < hb; i++) out[i] = 0;
    uint32_t v[64];
    uint8_t tmp[256];
    int over = 0;
    for (uint64_t b = 0; b < nb; b++) {
        uint32_t pb;
        oracle_gather_block_u32(in, x, y, z, b, v);
        uint8_t h = oracle_block_header(v, bins, pooling_th, &pb);
        if (pb / ((h >> 7) ? 4u : 8u) > 31) { *out_bytes = 0; return ORC_ERR_RANGE; }   /* rate 32 */
        out[b] = h;
        uint32_t n = oracle_encode_block(v, h, bins, tmp);
        if (!over && pos + n <= capacity) memcpy(out + pos, tmp, n);
        else over = 1;
        pos += n;
    }
    *out_bytes = pos;
    return over ? ORC_ERR_CAPACITY : ORC_OK;
}

int oracle_decompress_u32(const uint8_t *in, uint64_t in_bytes,
                          uint32_t x, uint32_t y, uint32_t z, const uint32_t bins[4],
                          uint32_t *out)
{
    if (oracle_check_params(x, y, z, bins, 0.0f)) return ORC_ERR_ARG;
    uint64_t nb = oracle_num_blocks(x, y, z);
    uint64_t hb = oracle_header_bytes(nb);
    if (in_bytes < hb) return ORC_ERR_CORRUPT;
    uint64_t pos = hb;
    uint32_t v[64];
    int bad = 0;
    for (uint64_t b = 0; b < nb; b++) {
        uint8_t h = in[b];
        uint32_t n = header_payload_bytes(h);
        if (pos + n > in_bytes) {
            bad = 1;
            memset(v, 0, sizeof v);
        } else {
            oracle_decode_block(h, in + pos, bins, v);
        }
        scatter_block_u32(out, y, z, b, v);
        pos += n;
    }
    return bad ? ORC_ERR_CORRUPT : ORC_OK;
}

/* ---- OpenMP two-pass variant (same bytes) -------------------------------- */

int oracle_compress_u16_omp(const uint16_t *in, uint32_t x, uint32_t y, uint32_t z,
                            const uint32_t bins[4], float pooling_th,
                            uint8_t *out, uint64_t capacity, uint64_t *out_bytes,
                            uint64_t *scratch /* >= nthreads+1 entries, or NULL */)
{
    int st = oracle_check_params(x, y, z, bins, pooling_th);
    if (st) return st;
    uint64_t nb = oracle_num_blocks(x, y, z);
    uint64_t hb = oracle_header_bytes(nb);
    if (capacity < hb) { *out_bytes = hb; return ORC_ERR_CAPACITY; }
    for (uint64_t i = nb; i < hb; i++) out[i] = 0;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    uint64_t part_local[257];
    uint64_t *part = scratch ? scratch : part_local;
    if (!scratch && nt > 256) nt = 256;
    int over = 0;
#pragma omp parallel num_threads(nt)
    {
        int t = 0, n = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        n = omp_get_num_threads();
#endif
        uint64_t lo = nb * (uint64_t)t / n, hi = nb * (uint64_t)(t + 1) / n;
        uint32_t v[64];
        uint64_t sum = 0;
        for (uint64_t b = lo; b < hi; b++) {      /* pass 1: headers + sizes */
            uint32_t pb;
            oracle_gather_block(in, x, y, z, b, v);
            out[b] = oracle_block_header(v, bins, pooling_th, &pb);
            sum += pb;
        }
        part[t + 1] = sum;
#pragma omp barrier
#pragma omp single
        {
            part[0] = hb;
            for (int i = 1; i <= n; i++) part[i] += part[i - 1];
            if (part[n] > capacity) over = 1;
        }
        if (!over) {
            uint64_t pos = part[t];
            for (uint64_t b = lo; b < hi; b++) {  /* pass 2: pack */
                oracle_gather_block(in, x, y, z, b, v);
                pos += oracle_encode_block(v, out[b], bins, out + pos);
            }
        }
#pragma omp single
        *out_bytes = part[n];
    }
    return over ? ORC_ERR_CAPACITY : ORC_OK;
}

int oracle_decompress_u16_omp(const uint8_t *in, uint64_t in_bytes,
                              uint32_t x, uint32_t y, uint32_t z, const uint32_t bins[4],
                              uint16_t *out, uint64_t *scratch)
{
    if (oracle_check_params(x, y, z, bins, 0.0f)) return ORC_ERR_ARG;
    uint64_t nb = oracle_num_blocks(x, y, z);
    uint64_t hb = oracle_header_bytes(nb);
    if (in_bytes < hb) return ORC_ERR_CORRUPT;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    uint64_t part_local[257];
    uint64_t *part = scratch ? scratch : part_local;
    if (!scratch && nt > 256) nt = 256;
    int bad = 0;
#pragma omp parallel num_threads(nt) reduction(| : bad)
    {
        int t = 0, n = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        n = omp_get_num_threads();
#endif
        uint64_t lo = nb * (uint64_t)t / n, hi = nb * (uint64_t)(t + 1) / n;
        uint64_t sum = 0;
        for (uint64_t b = lo; b < hi; b++) sum += header_payload_bytes(in[b]);
        part[t + 1] = sum;
#pragma omp barrier
#pragma omp single
        {
            part[0] = hb;
            for (int i = 1; i <= n; i++) part[i] += part[i - 1];
        }
        uint64_t pos = part[t];
        uint32_t v[64];
        for (uint64_t b = lo; b < hi; b++) {
            uint8_t h = in[b];
            uint32_t nbytes = header_payload_bytes(h);
            if ((h & 0x1f) > 16 || pos + nbytes > in_bytes) {
                bad = 1;
                memset(v, 0, sizeof v);
            } else {
                oracle_decode_block(h, in + pos, bins, v);
            }
            scatter_block(out, x, y, z, b, v);
            pos += nbytes;
        }
    }
    return bad ? ORC_ERR_CORRUPT : ORC_OK;
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- SSIM (cuLSZ_utility.cu:757-886) -------------------------------------
 * Mean structural similarity of two volumes over 7x7x7 windows moved by 2
 * along every axis (size0 fastest), as the reference's computeSSIM: the
 * uint16 values are the reference's float samples; means and the two-pass
 * variances and covariance in double, c1 and c2 from the original's range in
 * the window (0.01^2 and 0.03^2 when it is 0), c3 = c2 / 2.  No window fits
 * when a dimension is below 7: 0/0, as the reference. */
static double ssim_window(const uint16_t *data, const uint16_t *other, size_t size1, size_t size0,
                          size_t o0, size_t o1, size_t o2)
{
    const int w = 7;
    int np = 0;
    const size_t first = o0 + size0 * (o1 + size1 * o2);
    float xMin = (float)data[first], xMax = xMin, yMin = (float)other[first], yMax = yMin;
    double xSum = 0, ySum = 0;
    for (size_t i2 = o2; i2 < o2 + w; i2++)
        for (size_t i1 = o1; i1 < o1 + w; i1++)
            for (size_t i0 = o0; i0 < o0 + w; i0++) {
                const size_t k = i0 + size0 * (i1 + size1 * i2);
                const float a = (float)data[k], b = (float)other[k];
                np++;
                if (xMin > a) xMin = a;
                if (xMax < a) xMax = a;
                if (yMin > b) yMin = b;
                if (yMax < b) yMax = b;
                xSum += a;
                ySum += b;
            }
    const double xMean = xSum / np, yMean = ySum / np;
    double var_x = 0, var_y = 0, var_xy = 0;
    for (size_t i2 = o2; i2 < o2 + w; i2++)
        for (size_t i1 = o1; i1 < o1 + w; i1++)
            for (size_t i0 = o0; i0 < o0 + w; i0++) {
                const size_t k = i0 + size0 * (i1 + size1 * i2);
                const float a = (float)data[k], b = (float)other[k];
                var_x += (a - xMean) * (a - xMean);
                var_y += (b - yMean) * (b - yMean);
                var_xy += (a - xMean) * (b - yMean);
            }
    var_x /= np;
    var_y /= np;
    var_xy /= np;
    const double xSigma = sqrt(var_x), ySigma = sqrt(var_y), xyCov = var_xy;
#if ORACLE_MUTANT == 10
    const float range = yMax - yMin;   /* planted: the decoded volume's range */
#else
    const float range = xMax - xMin;
#endif
    (void)yMin;
    (void)yMax;
    double c1, c2;
    if (range == 0) {
        c1 = 0.01 * 0.01;
        c2 = 0.03 * 0.03;
    } else {
        c1 = 0.01 * 0.01 * range * range;
        c2 = 0.03 * 0.03 * range * range;
    }
    const double c3 = c2 / 2;
    const double luminance = (2 * xMean * yMean + c1) / (xMean * xMean + yMean * yMean + c1);
    const double contrast = (2 * xSigma * ySigma + c2) / (xSigma * xSigma + ySigma * ySigma + c2);
    const double structure = (xyCov + c3) / (xSigma * ySigma + c3);
    return luminance * contrast * structure;
}

double oracle_ssim_u16(const uint16_t *ori, const uint16_t *dec, uint32_t size2, uint32_t size1, uint32_t size0)
{
#if ORACLE_MUTANT == 9
    const size_t shift = 1;   /* planted: windows moved by 1 */
#else
    const size_t shift = 2;
#endif
    const size_t w = 7;
    double sum = 0;
    uint64_t nw = 0;
    for (size_t o2 = 0; o2 + w <= size2; o2 += shift)
        for (size_t o1 = 0; o1 + w <= size1; o1 += shift)
            for (size_t o0 = 0; o0 + w <= size0; o0 += shift) {
                nw++;
                sum += ssim_window(ori, dec, size1, size0, o0, o1, o2);
            }
    return sum / (double)nw;
}
