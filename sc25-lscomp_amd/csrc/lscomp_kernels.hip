// lscomp_kernels.hip -- gfx950 kernels of the lsCOMP uint16 codec.
//
// One workgroup = one tile of TILE (128/256/512) consecutive 8x8 blocks of an
// item, one block per lane; persistent workgroups loop over tiles.
// Compression (replaces cuLSZ_compression_kernel_uint16_bsize64,
// lsCOMP/src/cuLSZ_kernel.cu:827-1283):
//   1. dynamic tile ids (forward progress does not depend on dispatch order);
//   2. each lane loads its block once (8 16-byte row loads, coalesced across
//      lanes: lanes are consecutive blocks along z), computes the header byte;
//   3. workgroup scan of payload sizes, tile aggregate published at once;
//   4. a tile stays pending, packed in LDS, for two iterations while the
//      workgroup loads and packs the next ones; its look-back runs then;
//   5. decoupled look-back with a window of `win` (a launch parameter, 64 by
//      default) predecessors per probe round, one 64-bit word per tile
//      (epoch | flag | value) read and written with relaxed agent-scope
//      atomics, so no fences are needed on the multi-XCD part;
//   6. headers and payload leave LDS as contiguous coalesced 16-byte stores.
// Decompression (replaces cuLSZ_decompression_kernel_uint16_bsize64,
// cuLSZ_kernel.cu:1286-1646) takes tile offsets from a first phase over the
// header bytes (or a pre-pass kernel), stages each tile's payload in LDS with
// coalesced loads (for dense streams one tile ahead), and each lane unpacks
// its block and writes its rows.  Rate-0 blocks decode to zeros.  uint32 volumes run the same
// kernels with the uint32 codec.
#include "lscomp_internal.h"

namespace lsc {

constexpr uint32_t kSpinLimit = 1u << 16;

// 16 bytes at 4-byte alignment.  Volume rows and stream bytes are in general
// only dword aligned (volumes and stream buffers at 4-byte alignment are
// accepted, and stream offsets are multiples of 16 from the buffer's start);
// gfx950's global_load_dwordx4 / global_store_dwordx4 take that.  This type
// states the alignment these accesses rely on (a uint4 pointer would promise
// 16).  LDS accesses stay uint4: their addresses are 16-byte aligned.
typedef uint32_t u32x4_a4 __attribute__((__vector_size__(16), __aligned__(4)));

__device__ __forceinline__ uint4 ld16_a4(const void *p)
{
    const u32x4_a4 v = *(const u32x4_a4 *)p;
    return make_uint4(v[0], v[1], v[2], v[3]);
}

__device__ __forceinline__ void st16_a4(void *p, uint4 x)
{
    const u32x4_a4 v = {x.x, x.y, x.z, x.w};
    *(u32x4_a4 *)p = v;
}

__device__ __forceinline__ uint64_t lb_word(uint32_t epoch, uint32_t flag, uint64_t val)
{
    return ((uint64_t)epoch << 38) | ((uint64_t)flag << 36) | (val & kValMask);
}

__device__ __forceinline__ uint64_t ld_relaxed(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_relaxed(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Inclusive scan over the wave's 64 lanes with DPP moves: shifts by 1, 2, 4 and
// 8 lanes within each row of 16, then row 0's (row 2's) last lane added to row
// 1 (row 3) and lane 31 added to rows 2 and 3.  Lanes a move does not reach
// add 0.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane)
{
    (void)lane;
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);   // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);   // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);   // row_bcast:31
    return v;
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, uint32_t lane)
{
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)lane) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)lane);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// ---------------------------------------------------------------------------
// Block gather / scatter with zero padding (cuLSZ_kernel.cu:877-896, 1629-1640).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_block(const Item &it, uint32_t bx, uint32_t by, uint32_t bz,
                                           uint32_t D[32])
{
    const uint32_t y0 = by * 8, z0 = bz * 8;
    const uint16_t *p = it.vol_in + ((uint64_t)bx * it.y + y0) * it.z + z0;
    const uint32_t ny = it.y - y0 < 8u ? it.y - y0 : 8u;   // rows inside the volume
    const uint32_t nz = it.z - z0 < 8u ? it.z - z0 : 8u;   // columns inside the volume
    if (it.align != ALIGN2) {
        // 16-byte row loads; global_load_dwordx4 needs only dword alignment, so
        // rows of 4-byte-aligned volumes (z even) load whole.  A block at the
        // end of a row (nz < 8, even) loads 16 bytes as well while they lie
        // inside the volume (the excess is the next row's start) and clears the
        // pairs past z; only the volume's last row falls back to dword loads.
        const uint16_t *vol_end = it.vol_in + (uint64_t)it.x * it.y * it.z;
        const uint32_t keep = nz >> 1;   // dwords of a row inside the volume
        uint4 t[8];
        if (ny == 8u && p + 7ull * it.z + 8 <= vol_end) {
            // every row's 16 bytes lie inside the volume (almost every block):
            // eight loads with no per-row tests (per-row tests cost about 15
            // scalar and vector instructions a row; lossy CSSI compression
            // 0.752 -> 0.738 ms without them, 0.729 with the whole-row copy below)
#pragma unroll
            for (int r = 0; r < 8; r++) t[r] = ld16_a4(p + (uint64_t)r * it.z);
        } else {
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const uint16_t *q = p + (uint64_t)r * it.z;
                t[r] = make_uint4(0, 0, 0, 0);
                if ((uint32_t)r < ny) {
                    if (nz == 8u || q + 8 <= vol_end) {
                        t[r] = ld16_a4(q);
                    } else {
                        const uint32_t *q32 = (const uint32_t *)q;
                        t[r].x = q32[0];   // keep >= 1 (nz is even and nonzero)
                        if (keep > 1) t[r].y = q32[1];
                        if (keep > 2) t[r].z = q32[2];
                    }
                }
            }
        }
        if (keep == 4u) {
#pragma unroll
            for (int r = 0; r < 8; r++) {
                D[4 * r] = t[r].x;
                D[4 * r + 1] = t[r].y;
                D[4 * r + 2] = t[r].z;
                D[4 * r + 3] = t[r].w;
            }
        } else {   // a block at a row end: clear the pairs past z
#pragma unroll
            for (int r = 0; r < 8; r++) {
                D[4 * r] = t[r].x;
                D[4 * r + 1] = keep > 1 ? t[r].y : 0u;
                D[4 * r + 2] = keep > 2 ? t[r].z : 0u;
                D[4 * r + 3] = keep > 3 ? t[r].w : 0u;
            }
        }
    } else {
        // Rows only 2-byte aligned (z odd, or a volume starting at 2 mod 4).  A
        // row at a 4-byte boundary loads 16 bytes; a row at 2 mod 4 loads the 20
        // bytes from the dword boundary one element before it and shifts them
        // down by one element (v_alignbit_b32).  A block at a row end (nz < 8)
        // does the same while the bytes lie inside the volume and clears the
        // elements past z; rows whose bytes would leave the volume (its first
        // and last rows) load by elements.
        const uint16_t *vol_end = it.vol_in + (uint64_t)it.x * it.y * it.z;
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const bool rv = (uint32_t)r < ny;
            const uint16_t *q = p + (uint64_t)r * it.z;
            bool got = false;
            uint32_t w[4];
            if (rv && ((uintptr_t)q & 2u) == 0 && q + 8 <= vol_end) {
                const uint4 t = ld16_a4(q);
                w[0] = t.x;
                w[1] = t.y;
                w[2] = t.z;
                w[3] = t.w;
                got = true;
            } else if (rv && q > it.vol_in && q + 9 <= vol_end) {
                const uint32_t *a = (const uint32_t *)(q - 1);   // dword boundary
                const uint4 t = ld16_a4(a);
                const uint32_t e = a[4];
                w[0] = __builtin_amdgcn_alignbit(t.y, t.x, 16);
                w[1] = __builtin_amdgcn_alignbit(t.z, t.y, 16);
                w[2] = __builtin_amdgcn_alignbit(t.w, t.z, 16);
                w[3] = __builtin_amdgcn_alignbit(e, t.w, 16);
                got = true;
            }
            if (got) {
#pragma unroll
                for (int c = 0; c < 4; c++)   // elements 2c, 2c+1; clear those at or past nz
                    D[4 * r + c] = 2u * c + 1 < nz ? w[c] : 2u * c < nz ? (w[c] & 0xffffu) : 0u;
            } else {
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const uint32_t lo = (rv && 2u * c < nz) ? q[2 * c] : 0u;
                    const uint32_t hi = (rv && 2u * c + 1 < nz) ? q[2 * c + 1] : 0u;
                    D[4 * r + c] = lo | (hi << 16);
                }
            }
        }
    }
}

// kSplit16: a whole row at 4, 8 or 12 mod 16 is stored as two pieces that end
// and start at its 16-byte boundary (12 + 4, 8 + 8 or 4 + 12 bytes), a row at a
// boundary as one 16-byte store: CSSI-shaped lossless decode (pipelined
// decoder) 0.791 -> 0.765 ms.  The sparse decoder, whose row stores run at
// raised wave priority, keeps the 12 + 4-byte form (0.707 against 0.719 ms).
template <bool kSplit16>
__device__ __forceinline__ void store_block(const Item &it, uint32_t bx, uint32_t by, uint32_t bz,
                                            const uint32_t D[32])
{
    const uint32_t y0 = by * 8, z0 = bz * 8;
    uint16_t *p = it.vol_out + ((uint64_t)bx * it.y + y0) * it.z + z0;
    const uint32_t ny = it.y - y0 < 8u ? it.y - y0 : 8u;
    const uint32_t nz = it.z - z0 < 8u ? it.z - z0 : 8u;
    if (it.align != ALIGN2) {
        // z is even here, so nz is 2, 4, 6 or 8.  A row of 6 or 8 elements
        // stores its first 12 bytes, one of 2 or 4 its first dword; rows of 8
        // and of 4 then store their last dword with one shared instruction.
        // A wave that holds the block at the end of a row (nz 6 when z mod 8
        // is 6) issues the same two stores as the others.  Timed on the CSSI
        // volume, the 12 + 4-byte form was 1.5-2.5 % faster than one 16-byte
        // store per row at 4-byte alignment (lossless decode 0.829 -> 0.819
        // ms, lossy 0.748 -> 0.729), and 4 % faster than 8 16-byte stores
        // with no per-row branch for whole blocks.
        typedef uint32_t u32x3_a4 __attribute__((ext_vector_type(3), __aligned__(4)));
        if (it.align == ALIGN16) {
            // rows 16-byte aligned (z a multiple of 8, so nz is 8): one
            // aligned 16-byte store per row.  XPCS-shaped 512 x 1024 x 1024
            // lossless decode 0.236 -> 0.214 ms against the 12 + 4-byte form.
#pragma unroll
            for (int r = 0; r < 8; r++) {
                if ((uint32_t)r >= ny) break;
                *(uint4 *)(p + (uint64_t)r * it.z) = make_uint4(D[4 * r], D[4 * r + 1], D[4 * r + 2], D[4 * r + 3]);
            }
            return;
        }
        const bool wide = nz >= 6u, tail = nz == 8u || nz == 4u;
        const uint32_t tail_at = nz == 8u ? 3u : 1u;
        typedef uint32_t u32x2_a4 __attribute__((ext_vector_type(2), __aligned__(4)));
#pragma unroll
        for (int r = 0; r < 8; r++) {
            if ((uint32_t)r >= ny) break;
            uint32_t *q32 = (uint32_t *)(p + (uint64_t)r * it.z);   // dword aligned
            const uint32_t d0 = D[4 * r], d1 = D[4 * r + 1], d2 = D[4 * r + 2], d3 = D[4 * r + 3];
            if (kSplit16 && nz == 8u) {
                // lanes of one row of one frame share the row's offset mod 16
                const uint32_t k = (uint32_t)((uintptr_t)q32 >> 2) & 3u;
                if (k == 0u) {
                    st16_a4(q32, make_uint4(d0, d1, d2, d3));
                } else if (k == 1u) {
                    const u32x3_a4 t = {d0, d1, d2};
                    *(u32x3_a4 *)q32 = t;
                    q32[3] = d3;
                } else if (k == 2u) {
                    const u32x2_a4 a = {d0, d1}, b = {d2, d3};
                    *(u32x2_a4 *)q32 = a;
                    *(u32x2_a4 *)(q32 + 2) = b;
                } else {
                    const u32x3_a4 t = {d1, d2, d3};
                    q32[0] = d0;
                    *(u32x3_a4 *)(q32 + 1) = t;
                }
                continue;
            }
            if (wide) {
                const u32x3_a4 t = {d0, d1, d2};
                *(u32x3_a4 *)q32 = t;
            } else {
                q32[0] = d0;
            }
            if (tail) q32[tail_at] = nz == 8u ? d3 : d1;
        }
    } else {
        // Rows only 2-byte aligned: a row at a 4-byte boundary stores 16 bytes, a
        // row at 2 mod 4 its first element, the next pairs as dwords and its last
        // element; a block at a row end stores its nz elements the same way
        // (nothing outside the block is written).
#pragma unroll
        for (int r = 0; r < 8; r++) {
            if ((uint32_t)r >= ny) break;
            uint16_t *q = p + (uint64_t)r * it.z;
            const uint32_t d0 = D[4 * r], d1 = D[4 * r + 1], d2 = D[4 * r + 2], d3 = D[4 * r + 3];
            if (nz == 8u && ((uintptr_t)q & 2u) == 0) {
                st16_a4(q, make_uint4(d0, d1, d2, d3));
            } else if (nz == 8u) {
                q[0] = (uint16_t)(d0 & 0xffffu);
                uint32_t *m = (uint32_t *)(q + 1);   // dword boundary
                m[0] = __builtin_amdgcn_alignbit(d1, d0, 16);
                m[1] = __builtin_amdgcn_alignbit(d2, d1, 16);
                m[2] = __builtin_amdgcn_alignbit(d3, d2, 16);
                q[7] = (uint16_t)(d3 >> 16);
            } else if (((uintptr_t)q & 2u) == 0) {   // row end at a 4-byte boundary: whole pairs, then the odd element
                uint32_t *q32 = (uint32_t *)q;
                if (nz > 1u) q32[0] = d0;
                if (nz > 3u) q32[1] = d1;
                if (nz > 5u) q32[2] = d2;
                if (nz & 1u) {
                    const uint32_t last = nz == 1u ? d0 : nz == 3u ? d1 : nz == 5u ? d2 : d3;
                    q[nz - 1] = (uint16_t)(last & 0xffffu);
                }
            } else {   // row end at 2 mod 4: the first element, pairs (1,2) (3,4) as dwords, the last if unpaired
                q[0] = (uint16_t)(d0 & 0xffffu);
                uint32_t *m = (uint32_t *)(q + 1);   // dword boundary
                if (nz > 2u) m[0] = __builtin_amdgcn_alignbit(d1, d0, 16);
                if (nz > 4u) m[1] = __builtin_amdgcn_alignbit(d2, d1, 16);
                if (nz > 6u) m[2] = __builtin_amdgcn_alignbit(d3, d2, 16);
                if ((nz & 1u) == 0) {   // element nz - 1 (odd index) has no partner
                    const uint32_t last = nz == 2u ? d0 : nz == 4u ? d1 : nz == 6u ? d2 : d3;
                    q[nz - 1] = (uint16_t)(last >> 16);
                }
            }
        }
    }
}

constexpr uint32_t kWordsU32 = 62;   // payload words of a uint32 block at most (31 planes x 8 bytes)

__device__ __forceinline__ void load_block_u32(const Item &it, uint32_t bx, uint32_t by, uint32_t bz,
                                               uint32_t V[64])
{
    const uint32_t y0 = by * 8, z0 = bz * 8;
    const uint32_t *p = (const uint32_t *)it.vol_in + ((uint64_t)bx * it.y + y0) * it.z + z0;
    const uint32_t ny = it.y - y0 < 8u ? it.y - y0 : 8u;
    const uint32_t nz = it.z - z0 < 8u ? it.z - z0 : 8u;
    // two 16-byte loads per row (dword alignment suffices); a block at a row
    // end (nz < 8) loads them too while they lie inside the volume and clears
    // the elements past z, so it takes no slower path; only the volume's last
    // row loads by elements
    const uint32_t *vol_end = (const uint32_t *)it.vol_in + (uint64_t)it.x * it.y * it.z;
    uint4 t[16];
    if (ny == 8u && p + 7ull * it.z + 8 <= vol_end) {   // every row's 32 bytes inside the volume
#pragma unroll
        for (int r = 0; r < 8; r++) {
            t[2 * r] = ld16_a4(p + (uint64_t)r * it.z);
            t[2 * r + 1] = ld16_a4(p + (uint64_t)r * it.z + 4);
        }
    } else {
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const uint32_t *q = p + (uint64_t)r * it.z;
            t[2 * r] = make_uint4(0, 0, 0, 0);
            t[2 * r + 1] = make_uint4(0, 0, 0, 0);
            if ((uint32_t)r < ny) {
                if (nz == 8u || q + 8 <= vol_end) {
                    t[2 * r] = ld16_a4(q);
                    t[2 * r + 1] = ld16_a4(q + 4);
                } else {
                    uint32_t e[8];
#pragma unroll
                    for (int k = 0; k < 8; k++) e[k] = (uint32_t)k < nz ? q[k] : 0u;
                    t[2 * r] = make_uint4(e[0], e[1], e[2], e[3]);
                    t[2 * r + 1] = make_uint4(e[4], e[5], e[6], e[7]);
                }
            }
        }
    }
    if (nz == 8u) {
#pragma unroll
        for (int h = 0; h < 16; h++) {
            V[4 * h] = t[h].x;
            V[4 * h + 1] = t[h].y;
            V[4 * h + 2] = t[h].z;
            V[4 * h + 3] = t[h].w;
        }
    } else {
#pragma unroll
        for (int h = 0; h < 16; h++) {
            const uint32_t k0 = 4u * (h & 1);   // column of the first element
            V[4 * h] = k0 < nz ? t[h].x : 0u;
            V[4 * h + 1] = k0 + 1 < nz ? t[h].y : 0u;
            V[4 * h + 2] = k0 + 2 < nz ? t[h].z : 0u;
            V[4 * h + 3] = k0 + 3 < nz ? t[h].w : 0u;
        }
    }
}

__device__ __forceinline__ void store_block_u32(const Item &it, uint32_t bx, uint32_t by, uint32_t bz,
                                                const uint32_t V[64])
{
    const uint32_t y0 = by * 8, z0 = bz * 8;
    uint32_t *p = (uint32_t *)it.vol_out + ((uint64_t)bx * it.y + y0) * it.z + z0;
    const uint32_t ny = it.y - y0 < 8u ? it.y - y0 : 8u;
    const uint32_t nz = it.z - z0 < 8u ? it.z - z0 : 8u;
    if (ny == 8u && nz == 8u) {
#pragma unroll
        for (int r = 0; r < 8; r++) {
            st16_a4(p + (uint64_t)r * it.z, make_uint4(V[8 * r], V[8 * r + 1], V[8 * r + 2], V[8 * r + 3]));
            st16_a4(p + (uint64_t)r * it.z + 4, make_uint4(V[8 * r + 4], V[8 * r + 5], V[8 * r + 6], V[8 * r + 7]));
        }
    } else {   // by elements (a variant storing row ends as 16-, 8- and 4-byte pieces decoded 2.3x slower)
#pragma unroll
        for (int r = 0; r < 8; r++)
#pragma unroll
            for (int k = 0; k < 8; k++)
                if ((uint32_t)r < ny && (uint32_t)k < nz) p[(uint64_t)r * it.z + k] = V[8 * r + k];
    }
}

// What the tile kernels need of a value type: registers per block, payload
// words per block at most, and the per-block load, header, encode, decode and
// store.  kMem: the item descriptors live in device memory (table mode), so
// per-block parameters are read with indexed loads (sel4).  kKind: KIND_* of
// every item (block_header).
template <bool kMem, int kKind = KIND_ANY>
struct CodecU16T {
    static constexpr int kRegs = 32;              // 64 values as packed pairs
    static constexpr uint32_t kMaxWords = 32;     // 16 planes x 8 bytes
    __device__ __forceinline__ static void load(const Item &it, uint32_t bx, uint32_t by, uint32_t bz, uint32_t R[kRegs])
    {
        load_block(it, bx, by, bz, R);
    }
    __device__ __forceinline__ static uint32_t header(const uint32_t R[kRegs], const Item &it, Ctrl *)
    {
        return block_header<kMem, kKind>(R, it.dec);
    }
    __device__ __forceinline__ static void encode(uint32_t R[kRegs], uint32_t hdr, const Item &it, uint32_t *out)
    {
        encode_block<kMem, kKind>(R, hdr, it.dec, out);
    }
    __device__ __forceinline__ static bool bad_header(uint32_t hdr) { return (hdr & 0x1fu) > 16u; }
    __device__ __forceinline__ static void decode(const uint32_t *in, uint32_t hdr, const Item &it, uint32_t R[kRegs])
    {
        decode_block(in, hdr, it.dec.b, R);
    }
    __device__ __forceinline__ static void store(const Item &it, uint32_t bx, uint32_t by, uint32_t bz, const uint32_t R[kRegs])
    {
        store_block<true>(it, bx, by, bz, R);
    }
};

using CodecU16 = CodecU16T<false>;

struct CodecU32 {
    static constexpr int kRegs = 64;
    static constexpr uint32_t kMaxWords = kWordsU32;
    __device__ __forceinline__ static void load(const Item &it, uint32_t bx, uint32_t by, uint32_t bz, uint32_t R[kRegs])
    {
        load_block_u32(it, bx, by, bz, R);
    }
    // A block whose quantized maximum needs 32 bits is not representable: the
    // launch is flagged and the block's rate cleared, so the stream stays
    // consistent.
    __device__ __forceinline__ static uint32_t header(const uint32_t R[kRegs], const Item &it, Ctrl *ctrl)
    {
        bool rate32 = false;
        uint32_t hdr = block_header_u32(R, it.dec, &rate32);
        if (rate32) {
            __hip_atomic_fetch_or(&ctrl->error, ERRF_RANGE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            hdr &= 0xe0u;
        }
        return hdr;
    }
    __device__ __forceinline__ static void encode(uint32_t R[kRegs], uint32_t hdr, const Item &it, uint32_t *out)
    {
        encode_block_u32(R, hdr, it.dec, out);
    }
    __device__ __forceinline__ static bool bad_header(uint32_t) { return false; }   // every 5-bit rate decodes
    __device__ __forceinline__ static void decode(const uint32_t *in, uint32_t hdr, const Item &it, uint32_t R[kRegs])
    {
        decode_block_u32(in, hdr, it.dec.b, R);
    }
    __device__ __forceinline__ static void store(const Item &it, uint32_t bx, uint32_t by, uint32_t bz, const uint32_t R[kRegs])
    {
        store_block_u32(it, bx, by, bz, R);
    }
};

// Workgroup exclusive scan of per-lane word counts; returns the lane's offset,
// sets *agg to the tile total.  Contains barriers.
template <int TILE>
__device__ __forceinline__ uint32_t tile_scan(uint32_t words, uint32_t tid, uint32_t *s_wsum,
                                              uint32_t *agg)
{
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    const uint32_t incl = wave_incl_scan(words, lane);
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < TILE / 64; w++) {
        const uint32_t s = s_wsum[w];
        before += (w < (int)wave) ? s : 0u;
        total += s;
    }
    *agg = total;
    return before + incl - words;
}

// Copy `words` 4-byte words from LDS to global memory at dst (4-byte
// aligned): dword stores up to 16-byte alignment, then 16-byte stores, then
// the dword tail.
template <int TILE>
__device__ __forceinline__ void store_payload(const uint32_t *pay, uint32_t words, uint8_t *dst, uint32_t tid)
{
    uint32_t head = (uint32_t)((16 - ((uintptr_t)dst & 15)) & 15) >> 2;
    if (head > words) head = words;
    uint32_t *d32 = (uint32_t *)dst;
    if (tid < head) d32[tid] = pay[tid];
    const uint32_t nchunk = (words - head) >> 2;
    uint4 *d128 = (uint4 *)(d32 + head);
    for (uint32_t c = tid; c < nchunk; c += TILE) {
        const uint32_t w = head + 4 * c;
        d128[c] = make_uint4(pay[w], pay[w + 1], pay[w + 2], pay[w + 3]);
    }
    const uint32_t done = head + 4 * nchunk;
    if (tid < words - done) d32[done + tid] = pay[done + tid];
}

// 16-byte chunk c of a stream counted from the 16-byte aligned byte a0; a chunk
// that crosses s_len is read by dwords (nothing at or past s_len is read,
// missing words are 0).
__device__ __forceinline__ uint4 load_chunk(const uint8_t *s_in, uint64_t s_len, uint64_t a0, uint32_t c)
{
    const uint64_t at = a0 + 16ull * c;
    if (at + 16 <= s_len) return ld16_a4(s_in + at);
    const uint32_t *s32 = (const uint32_t *)(s_in + at);
    uint32_t t[4] = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (at + 4 * j + 4 <= s_len) t[j] = s32[j];
    return make_uint4(t[0], t[1], t[2], t[3]);
}

// Stage `words` 4-byte words of a stream starting at byte g0 (4-byte aligned)
// into LDS: 16-byte loads of the aligned chunks covering them, written to
// pay[] from index 0 = the word at (g0 & ~15), so the first payload word sits at
// pay[(g0 & 15) / 4].  Nothing at or past s_len is read (dword tail).
// Chunks before first_chunk are left alone.
template <int TILE>
__device__ __forceinline__ void load_payload(const uint8_t *s_in, uint64_t s_len, uint64_t g0, uint32_t words,
                                             uint32_t *pay, uint32_t tid, uint32_t first_chunk = 0)
{
    const uint64_t a0 = g0 & ~15ull;
    const uint32_t shift = (uint32_t)(g0 - a0) >> 2;
    const uint32_t nchunk = (shift + words + 3) >> 2;
    const auto pass = [&](uint32_t base) {
        uint4 v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {      // up to 4 loads in flight per lane
            const uint32_t c = base + (uint32_t)k * TILE + tid;
            v[k] = c < nchunk ? load_chunk(s_in, s_len, a0, c) : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t c = base + (uint32_t)k * TILE + tid;
            if (c < nchunk) *(uint4 *)(pay + 4 * c) = v[k];
        }
    };
    // the first pass outside the loop: a loop header before it made the
    // loads wait for every store in flight
    if (first_chunk < nchunk) pass(first_chunk);
    for (uint32_t base = first_chunk + 4u * TILE; base < nchunk; base += 4u * TILE) pass(base);
}

// Look-back window `win` (a launch parameter, a multiple of 64 up to the
// tile): one probe round reads the `win` predecessor words nearest to the
// tile, one per thread of the first win / 64 waves (whole waves: a wave's
// unread lanes would count as empty predecessors).  Tiles are resolved two
// iterations after they publish their aggregates, so an inclusive prefix is
// nearly always close by; every tile pays for the window's status reads
// (agent-scope loads of lines all workgroups are writing).  Timed on the CSSI
// volume, compression lossless / lossy: at tile 512 (uint16) 512 words 0.861 /
// 0.858 ms, 256 0.82 / 0.81, 128 0.78-0.79 / 0.81, 64 0.79 / 0.81; at tile
// 256 (uint32) 256 words 1.68 / 1.65, 128 1.62 / 1.61, 64 1.57-1.58 / 1.59.
// The default is a quarter of the tile, at least 64 (default_window).
constexpr int lb_per(int) { return 1; }   // probe words per thread

// LDS of the workgroup-wide look-back, per slice of 64 window positions (one
// wave's k-th words): nearest inclusive and empty positions and the sum of the
// values up to the nearest inclusive one (all values without one).  Two
// buffers, used alternately, so one barrier per round suffices.
template <int NSLICE>
struct LbShared {
    uint32_t pos_incl[2][NSLICE], pos_empty[2][NSLICE];
    uint64_t sum[2][NSLICE];
    uint32_t err[2];
};

template <int TILE, uint32_t W = 32>
struct TileShared {
    static constexpr int kWaves = TILE / 64;
    alignas(16) uint32_t pay[TILE * W + 8];   // payload staging (W words per block max) + shift slack
    uint32_t hdr[4][TILE / 4];     // header bytes, one buffer per tile in flight (compression)
    uint32_t wsum[2][kWaves];      // wave totals of a scan, by iteration parity
    LbShared<kWaves * lb_per(TILE)> lb;    // look-back rounds
    LbShared<kWaves * lb_per(TILE)> lbf;   // compression: first round, by iteration parity
    uint32_t tile, next, epoch, ntot;
    uint32_t next2[2];             // compression: next tile id, by iteration parity
    uint32_t cls;                  // the workgroup's tile-id class
    uint32_t tagg[128];            // decode offset phase: payload words of a chunk's tiles
    uint32_t chunk, chunk_agg;     // decode offset phase: chunk id, chunk total
};

// One probe round: this thread's words of the predecessors at distances
// 1 + k*TILE + tid (k < lb_per(TILE)) behind `base + 1`, for the threads inside
// the window (the others read nothing); tiles before `first` (the item's first
// tile) read as an inclusive prefix of 0.
template <int TILE>
__device__ __forceinline__ void lb_probe(const uint64_t *status, int64_t base, uint32_t first, uint32_t epoch,
                                         uint32_t win, uint64_t w[lb_per(TILE)])
{
#pragma unroll
    for (int k = 0; k < lb_per(TILE); k++) {
        const int64_t idx = base - (int64_t)k * TILE - (int64_t)threadIdx.x;
        const bool inwin = (uint32_t)k * TILE + threadIdx.x < win;
        w[k] = !inwin ? 0ull : idx >= (int64_t)first ? ld_relaxed(status + idx) : lb_word(epoch, LB_INCL, 0);
    }
}

enum : uint32_t { LB_RETRY = 0, LB_NEXT = 1, LB_DONE = 2 };

// First half of the evaluation of one probe round (words w[] of the window
// that ends at the tile being resolved minus base): this wave's slices (64
// window positions each; waves past the window report nothing) go to LDS
// buffer `par`: nearest inclusive and empty
// positions and the sum of the values up to the nearest inclusive one (all
// values without one).  An aggregate is at most 2^21 words (a status word
// covers at most 128 tiles of at most 512 blocks of at most 32 words, or 256
// blocks of at most 62), so the 64 aggregate lanes of a slice are summed in 32
// bits and the inclusive lane's value is added in 64.  No barrier.
template <int TILE>
__device__ __forceinline__ void lb_slices(const uint64_t w[lb_per(TILE)], uint32_t epoch, uint32_t kWin,
                                          LbShared<TILE / 64 * lb_per(TILE)> &sh, uint32_t par)
{
    constexpr int P = lb_per(TILE), NW = TILE / 64;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
#pragma unroll
    for (int k = 0; k < P; k++) {
        if ((uint32_t)k * TILE + (tid & ~63u) >= kWin) {   // a wave outside the window reports nothing
            if (lane == 0) {
                sh.pos_incl[par][k * NW + wave] = kWin;
                sh.pos_empty[par][k * NW + wave] = kWin;
                sh.sum[par][k * NW + wave] = 0;
            }
            continue;
        }
        const uint32_t flag = ((uint32_t)(w[k] >> 38) == epoch) ? (uint32_t)(w[k] >> 36) & 3u : 0u;
        const uint64_t im = __ballot(flag == LB_INCL), em = __ballot(flag == LB_EMPTY);
        const uint32_t cut = im ? (uint32_t)__builtin_ctzll(im) : 64u;
        const uint32_t lo = wave_incl_scan(lane < cut ? (uint32_t)w[k] : 0u, lane);
        uint64_t part = (uint32_t)__builtin_amdgcn_readlane((int)lo, 63);
        if (im) part += readlane_u64(w[k], cut) & kValMask;
        if (lane == 0) {
            const uint32_t at = (uint32_t)k * TILE + (tid & ~63u);
            sh.pos_incl[par][k * NW + wave] = im ? at + cut : kWin;
            sh.pos_empty[par][k * NW + wave] = em ? at + (uint32_t)__builtin_ctzll(em) : kWin;
            sh.sum[par][k * NW + wave] = part;
        }
    }
}

// Second half, after a barrier that follows every wave's lb_slices: adds the
// values up to the nearest inclusive predecessor to excl.  Returns LB_DONE
// (excl is the exclusive prefix), LB_NEXT (no inclusive predecessor in the
// window: continue one window further back) or LB_RETRY (an empty predecessor
// lies before the nearest inclusive one: read the same window again).
template <int TILE>
__device__ __forceinline__ uint32_t lb_reduce(const LbShared<TILE / 64 * lb_per(TILE)> &sh, uint32_t par,
                                              uint32_t kWin, uint64_t &excl)
{
    constexpr int NS = TILE / 64 * lb_per(TILE);
    uint32_t j = kWin, e = kWin;
    for (int v = 0; v < NS; v++) {
        const uint32_t x = sh.pos_incl[par][v], y = sh.pos_empty[par][v];
        j = x < j ? x : j;
        e = y < e ? y : e;
    }
    // an empty predecessor nearer than the nearest inclusive one (or anywhere,
    // when the window holds no inclusive one) must be read again
    if (e < j) return LB_RETRY;
    // slices are in window order (slice v covers positions 64v .. 64v+63); the
    // slices before j's hold only aggregates, j's slice sums up to j
    uint64_t total = 0;
    for (int v = 0; v < NS; v++)
        if (64u * (uint32_t)v <= j) total += sh.sum[par][v];
    excl += total;
    return j < kWin ? LB_DONE : LB_NEXT;
}

// One probe round evaluated by the workgroup (lb_slices, barrier, lb_reduce).
// *err receives the launch error word when `poll` is set (else 0).  `par`
// alternates between consecutive calls (LDS buffer).  Called by every thread.
template <int TILE>
__device__ __forceinline__ uint32_t lb_eval(const uint64_t w[lb_per(TILE)], uint32_t epoch, uint32_t win,
                                            LbShared<TILE / 64 * lb_per(TILE)> &sh, Ctrl *ctrl, bool poll,
                                            uint32_t par, uint64_t &excl, uint32_t *err)
{
    lb_slices<TILE>(w, epoch, win, sh, par);
    if (threadIdx.x == 0)
        sh.err[par] = poll ? __hip_atomic_load(&ctrl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
    __syncthreads();
    *err = sh.err[par];
    return lb_reduce<TILE>(sh, par, win, excl);
}

// Look-back from the window ending at tile `base` (inclusive) towards the
// item's first tile, with `excl` already accumulated and `spins` probes spent:
// returns the exclusive prefix (in words).  Bounded wait: after kSpinLimit
// re-reads, or once any tile of the launch has given up, it stops waiting and
// flags the launch (ERRF_TIMEOUT).  Called by every thread; barriers inside.
template <int TILE>
__device__ __forceinline__ uint64_t lookback_from(uint64_t *status, int64_t base, uint32_t first, uint32_t epoch,
                                  uint32_t win, LbShared<TILE / 64 * lb_per(TILE)> &sh, Ctrl *ctrl, uint64_t excl,
                                  uint32_t spins, uint32_t par)
{
    for (;;) {
        uint64_t w[lb_per(TILE)];
        lb_probe<TILE>(status, base, first, epoch, win, w);
        // the launch-wide error word is polled only while waiting, every 64
        // probes: it must not become a hot spot for every look-back
        uint32_t err;
        const uint32_t r = lb_eval<TILE>(w, epoch, win, sh, ctrl, (spins & 63u) == 63u, par, excl, &err);
        par ^= 1u;
        if (r == LB_DONE) return excl;
        if (r == LB_NEXT) {
            base -= (int64_t)win;
            continue;
        }
        if ((err & ERRF_TIMEOUT) || ++spins > kSpinLimit) {
            if (threadIdx.x == 0)
                __hip_atomic_fetch_or(&ctrl->error, ERRF_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return 0;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Exclusive prefix (in words) of tile `tile` over the tiles [first, tile) of
// its item by decoupled look-back.
template <int TILE>
__device__ __forceinline__ uint64_t lookback_tile(uint64_t *status, uint32_t tile, uint32_t first,
                                                  uint32_t epoch, uint32_t win,
                                                  LbShared<TILE / 64 * lb_per(TILE)> &sh, Ctrl *ctrl)
{
    return lookback_from<TILE>(status, (int64_t)tile - 1, first, epoch, win, sh, ctrl, 0, 0, 0);
}

__device__ __forceinline__ void block_coords(const Item &it, uint32_t b, uint32_t &bx,
                                             uint32_t &by, uint32_t &bz)
{
    bx = fdiv(b, it.div_frame);
    const uint32_t rem = b - bx * it.per_frame;
    by = fdiv(rem, it.div_nbz);
    bz = rem - by * it.nbz;
}

__device__ __forceinline__ void finish_launch(Ctrl *ctrl, uint32_t epoch)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t old = __hip_atomic_fetch_add(&ctrl->done_ctr, 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        if (old == gridDim.x - 1) {
            __hip_atomic_store(&ctrl->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctrl->chunk.v, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for (uint32_t c = 0; c < kMaxCounters; c++)
                __hip_atomic_store(&ctrl->ctr[c].v, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctrl->done_ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctrl->epoch, (epoch + 1u) & kEpochMask, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Dynamic tile ids (thread 0 of a workgroup only).  Tile k*nc + c is the k-th
// id of counter c; a workgroup's class c is its ticket (its place in the start
// order) mod nc, taken once, so every class has a running member as soon as nc
// workgroups run, and a tile id is handed only to a running workgroup of its
// class.  Forward progress: the lowest tile without an inclusive prefix is
// either held by a running workgroup, whose look-backs all wait for lower
// tiles only, or not yet taken, and then no member of its class waits (their
// pending tiles are lower), so one of them takes it.  Only where no tile
// waits for another (decode with precomputed offsets) are ids taken ahead: a
// reserved tile must never hold up a look-back.
//
// take() returns the counter's raw value k and id(k) the tile: a value
// computed from the atomic's result at once would make the wave wait for its
// round trip, and gfx9 counts loads, stores and atomics in one in-order
// counter, so that wait would also cover every access issued before it.  The
// raw value goes to LDS where the old tile was written anyway (the wait falls
// where the loads it covers have returned) and is converted after the barrier.
//
// Compression with several counters hands out the first gridDim.x tiles by
// the start ticket itself (the workgroup with ticket j starts with tile j,
// whose class j mod nc is its own), so one atomic gives both; the class
// counters then start after those tiles (`base`: the number of tiles of the
// class below gridDim.x).  The argument above holds unchanged: the first
// tiles are held by running workgroups in start order.
struct TileQ {
    uint32_t cls, nc, base;   // base: ids of this class handed out by the start tickets
    // the first of n consecutive raw values of this class's counter
    __device__ __forceinline__ uint32_t take(Ctrl *ctrl, uint32_t n = 1u) const
    {
        return __hip_atomic_fetch_add(&ctrl->ctr[cls].v, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ uint32_t id(uint32_t k) const { return (k + base) * nc + cls; }
};

// Thread 0: the workgroup's class (one ticket atomic when there are several
// counters).  It is broadcast through LDS so that every thread holds the
// queue in scalar registers.
__device__ __forceinline__ uint32_t tile_class(Ctrl *ctrl, Sched sc)
{
    if (sc.nc <= 1) return 0;
    return __hip_atomic_fetch_add(&ctrl->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) % sc.nc;
}

__device__ __forceinline__ TileQ uniform_queue(uint32_t cls, Sched sc, uint32_t base = 0u)
{
    return TileQ{(uint32_t)__builtin_amdgcn_readfirstlane(cls), (uint32_t)__builtin_amdgcn_readfirstlane(sc.nc),
                 (uint32_t)__builtin_amdgcn_readfirstlane(base)};
}

// The single item, read again from the kernel argument segment where it is
// used.  Kernels that take their Item by value otherwise load all of its
// fields at entry and keep them live in scalar registers for the whole
// persistent loop; with the loop's own state that spilled 60-105 scalar
// registers into vector lanes (hundreds of v_readlane per tile).  Read
// through a pointer the compiler cannot hoist (the empty asm), the fields are
// scalar loads from the argument segment, near the use (lossy CSSI
// compression 0.779 -> 0.751 ms, uint32 1.59 -> 1.41 ms).  Only for kernels
// whose first parameter is `const Item`: the first explicit argument sits at
// offset 0 of the segment.
typedef const __attribute__((address_space(4))) Item const_item;
__device__ __forceinline__ const Item &arg_item()
{
    const const_item *p = (const const_item *)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return *(const Item *)p;
}

// Item of a global tile: the single item, or items[tile_item[tile]].
template <bool kTable>
__device__ __forceinline__ const Item &tile_item(const Item &single, const Item *items, const uint32_t *tile_item,
                                                 uint32_t tile)
{
    (void)single;
    if (!kTable) return arg_item();
    return items[__builtin_amdgcn_readfirstlane(tile_item[tile])];
}

// ---------------------------------------------------------------------------
// Compression.
// ---------------------------------------------------------------------------
template <bool kTable>
__device__ __forceinline__ const Item &item_at(const Item &single, const Item *items, uint32_t idx)
{
    (void)single;
    if (!kTable) return arg_item();
    return items[idx];
}

// A tile between its packing and its stores: payload staged in LDS at
// pay[start, start + agg), headers in hdr[hb], aggregate published.
struct PendTile {
    uint32_t tile, tl, item, agg, start, hb;
};

// Start of a run of `len` words in the staging ring [0, ring) behind the
// pending runs (a = oldest, b = newest when nlive == 2; a run of length 0
// occupies nothing): after the newest run, or at 0 when that passes the end of
// the ring and the run fits before the oldest.  ~0u when neither fits.
__device__ __forceinline__ uint32_t ring_place(uint32_t len, uint32_t nlive, const PendTile &a, const PendTile &b,
                                               uint32_t ring)
{
    const bool ua = nlive >= 1 && a.agg != 0, ub = nlive == 2 && b.agg != 0;
    if (!ua && !ub) return 0;
    if (ua != ub) {   // one run
        const uint32_t start = ua ? a.start : b.start, end = ua ? a.start + a.agg : b.start + b.agg;
        if (end + len <= ring) return end;
        return len <= start ? 0u : ~0u;
    }
    const uint32_t tail = b.start + b.agg;
    if (b.start > a.start) {   // in address order
        if (tail + len <= ring) return tail;
        return len <= a.start ? 0u : ~0u;
    }
    return tail + len <= a.start ? tail : ~0u;   // the newest run wrapped to the front
}

// Stores a tile whose exclusive prefix (in words) is excl: its TILE header
// bytes clipped to the padded header region (pad bytes are 0; 16-byte stores,
// the last partial chunk by dwords) and its payload; the item's last tile also
// writes the stream size.
template <int TILE>
__device__ __forceinline__ void store_tile(const Item &it, uint32_t tl, uint32_t agg, const uint32_t *hdr,
                                           const uint32_t *pay, uint64_t excl, uint32_t tid)
{
    const uint64_t g0 = (uint64_t)it.hdr_bytes + 4ull * excl;   // byte offset of the tile payload
    if (tid < (uint32_t)TILE / 16) {
        const uint64_t o = (uint64_t)tl * TILE + 16ull * tid;
        if (o + 16 <= it.hdr_bytes) {
            st16_a4(it.s_out + o, make_uint4(hdr[4 * tid], hdr[4 * tid + 1], hdr[4 * tid + 2], hdr[4 * tid + 3]));
        } else {
#pragma unroll
            for (uint32_t k = 0; k < 4; k++)
                if (o + 4 * k < it.hdr_bytes) *(uint32_t *)(it.s_out + o + 4 * k) = hdr[4 * tid + k];
        }
    }
    if (g0 + 4ull * agg <= it.s_len) store_payload<TILE>(pay, agg, it.s_out + g0, tid);
    if (tid == 0 && tl == it.ntiles - 1 && it.result) *it.result = g0 + 4ull * agg;
}

// Three-stage pipeline per workgroup.  A tile is loaded, scanned and packed
// into LDS in one iteration (its aggregate published at once) and stays
// pending for two more; the iteration that takes on tile n finishes the oldest
// pending tile p:
//   1. issue the loads of n and the look-back probe of p, then take the next
//      tile id (kept in a register of thread 0 until the next scan, so the
//      atomic's round trip overlaps the loads)
//   2. header of n; n's scan and the first half of p's look-back evaluation
//      share one barrier; publish n's aggregate
//   3. pack n into the LDS ring behind the pending tiles when it fits (the
//      common case)
//   4. finish p's look-back (re-reading while predecessors are still empty),
//      publish its inclusive prefix, store its headers and payload; if n did
//      not fit, finish further pending tiles until it does and pack it.
// By the time p is resolved its predecessors have had a whole iteration to
// publish their aggregates, so the look-back rarely waits, and a workgroup
// that does wait does not hold up the tiles behind its reserved id for long.
// The scan barrier is the only one an iteration needs: what is written before
// it (wave totals, the first look-back round, the next id) has one LDS buffer
// per iteration parity; header bytes have four buffers (a tile's buffer is
// rewritten four tiles later); a ring region is rewritten only by a later
// iteration's pack, after its scan barrier.  (Timed alternatives: issuing the
// next loads early, storing p before packing n, and a closing barrier after
// each iteration or after large tiles were all slower or no faster.)
// Forward progress: a look-back waits only for lower tile ids; the lowest tile
// without an inclusive prefix belongs to a workgroup that, whatever stage the
// tile is in, either resolves it (all its predecessors are inclusive) or
// finishes lower pending tiles of its own first, so it always advances.
template <class C, int TILE, bool kTable>
__device__ __forceinline__ void compress_tiles(const Item single, const Item *__restrict__ items,
                                               const uint32_t *__restrict__ tile_items, uint32_t total_tiles,
                                               const uint32_t *d_total_tiles, Ctrl *ctrl, uint64_t *status,
                                               Sched sc)
{
    constexpr uint32_t kRing = (uint32_t)TILE * C::kMaxWords;   // staging words: one worst-case tile
    __shared__ TileShared<TILE, C::kMaxWords> sh;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t win = sc.win;
    if (tid == 0) {
        sh.epoch = __hip_atomic_load(&ctrl->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sh.ntot = d_total_tiles ? *d_total_tiles : total_tiles;
        // the first tile: with one counter its first value, with several the
        // start ticket (TileQ); raw, turned into the class after the barrier
        sh.next = sc.nc <= 1 ? TileQ{0u, 1u, 0u}.take(ctrl)
                             : __hip_atomic_fetch_add(&ctrl->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const uint32_t epoch = sh.epoch, ntot = sh.ntot;
    const uint32_t n0 = sh.next;
    const uint32_t cls = sc.nc <= 1 ? 0u : n0 % sc.nc;
    const TileQ q = uniform_queue(cls, sc, sc.nc <= 1 ? 0u : (gridDim.x - cls + sc.nc - 1u) / sc.nc);

    uint32_t np = 0, hb = 0;   // pending tiles (p0 oldest), header buffer of the next tile
    PendTile p0 = {0, 0, 0, 0, 0, 0}, p1 = {0, 0, 0, 0, 0, 0};
    uint32_t n = n0, par = 0;   // par: iteration parity (LDS buffers written before the scan barrier)
    for (;; par ^= 1u) {
        const bool have_n = n < ntot;
        if (!have_n && np == 0) break;
        const bool resolve = np == 2 || (!have_n && np != 0);

        // 1. loads of n, probe of p0, next id
        uint32_t D[C::kRegs];
        uint32_t n_item = 0, n_tl = 0;
        bool nvalid = false;
        if (have_n) {
            n_item = kTable ? __builtin_amdgcn_readfirstlane(tile_items[n]) : 0u;
            const Item &nit = item_at<kTable>(single, items, n_item);
            n_tl = n - nit.first_tile;
            const uint32_t b = n_tl * (uint32_t)TILE + tid;
            nvalid = b < nit.nblocks;
            if (nvalid) {
                uint32_t bx, by, bz;
                block_coords(nit, b, bx, by, bz);
                C::load(nit, bx, by, bz, D);
            }
        }
        const bool p_lb = resolve && p0.tl != 0;   // p0 needs a look-back
        uint64_t w0[lb_per(TILE)];
        if (p_lb)
            lb_probe<TILE>(status, (int64_t)p0.tile - 1, item_at<kTable>(single, items, p0.item).first_tile, epoch,
                           win, w0);
        // Issued last: returns are counted in order, so waiting for the loads or
        // the probe must not include the atomic's longer round trip (the raw
        // value is only written to LDS before the scan barrier).
        uint32_t next_k = ~0u;
        if (tid == 0 && have_n) next_k = q.take(ctrl);

        // 2. header of n; n's scan and the first half of p0's evaluation share a barrier
        uint32_t n_hdr = 0, n_words = 0;
        if (have_n) {
            if (nvalid) {
                n_hdr = C::header(D, item_at<kTable>(single, items, n_item), ctrl);
                n_words = header_payload_bytes(n_hdr) >> 2;
            }
            ((uint8_t *)sh.hdr[hb])[tid] = (uint8_t)n_hdr;
        }
        const uint32_t incl = wave_incl_scan(n_words, lane);
        if (lane == 63) sh.wsum[par][wave] = incl;
        if (p_lb) lb_slices<TILE>(w0, epoch, win, sh.lbf, par);
        if (tid == 0) sh.next2[par] = next_k;
        __syncthreads();
        const uint32_t next_raw = sh.next2[par];
        const uint32_t n_next = next_raw == ~0u ? n : q.id(next_raw);
        uint32_t before = 0, n_agg = 0;
#pragma unroll
        for (int v = 0; v < TILE / 64; v++) {
            const uint32_t x = sh.wsum[par][v];
            before += (v < (int)wave) ? x : 0u;
            n_agg += x;
        }
        const uint32_t n_off = before + incl - n_words;
        if (have_n && tid == 0) st_relaxed(status + n, lb_word(epoch, n_tl == 0 ? LB_INCL : LB_AGG, n_agg));

        // 3. pack n behind the pending tiles if it fits in the staging ring
        uint32_t n_start = have_n ? ring_place(n_agg, np, p0, p1, kRing) : ~0u;
        if (n_start != ~0u && n_words)
            C::encode(D, n_hdr, item_at<kTable>(single, items, n_item), &sh.pay[n_start + n_off]);

        // 4. finish p0
        if (resolve) {
            const Item &pit = item_at<kTable>(single, items, p0.item);
            uint64_t excl = 0;
            if (p_lb) {
                const uint32_t r = lb_reduce<TILE>(sh.lbf, par, win, excl);
                if (r != LB_DONE) {
                    const int64_t base = (int64_t)p0.tile - 1 - (r == LB_NEXT ? (int64_t)win : 0);
                    excl = lookback_from<TILE>(status, base, pit.first_tile, epoch, win, sh.lb, ctrl, excl,
                                               r == LB_RETRY ? 1u : 0u, 0u);
                }
                if (tid == 0) st_relaxed(status + p0.tile, lb_word(epoch, LB_INCL, excl + p0.agg));
            }
            store_tile<TILE>(pit, p0.tl, p0.agg, sh.hdr[p0.hb], sh.pay + p0.start, excl, tid);
            p0 = p1;
            np--;
        }
        if (have_n && n_start == ~0u) {
            // n did not fit: finish pending tiles (oldest first) until it does
            __syncthreads();   // p0's payload has left LDS
            for (;;) {
                n_start = ring_place(n_agg, np, p0, p1, kRing);
                if (n_start != ~0u) break;
                const Item &pit = item_at<kTable>(single, items, p0.item);
                uint64_t excl = 0;
                if (p0.tl != 0) {
                    excl = lookback_tile<TILE>(status, p0.tile, pit.first_tile, epoch, win, sh.lb, ctrl);
                    if (tid == 0) st_relaxed(status + p0.tile, lb_word(epoch, LB_INCL, excl + p0.agg));
                }
                store_tile<TILE>(pit, p0.tl, p0.agg, sh.hdr[p0.hb], sh.pay + p0.start, excl, tid);
                p0 = p1;
                np--;
                __syncthreads();
            }
            if (n_words) C::encode(D, n_hdr, item_at<kTable>(single, items, n_item), &sh.pay[n_start + n_off]);
            __syncthreads();
        }
        if (have_n) {
            const PendTile t = {n, n_tl, n_item, n_agg, n_start, hb};
            if (np == 0) p0 = t; else p1 = t;
            np++;
            hb = (hb + 1u) & 3u;
        }
        n = n_next;
    }
    finish_launch(ctrl, epoch);
}

// kKind: KIND_* of every item.  The general kernel's code for every kind of
// settings costs scalar register spills: lossless CSSI compression was 2-3 %
// slower in it than in its own kernel.
template <int TILE, bool kTable, int kKind = KIND_ANY>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(4))) void compress_kernel(
    const Item single, const Item *__restrict__ items, const uint32_t *__restrict__ tile_items, uint32_t total_tiles,
    const uint32_t *d_total_tiles, Ctrl *ctrl, uint64_t *status, Sched sc)
{
    compress_tiles<CodecU16T<kTable, kKind>, TILE, kTable>(single, items, tile_items, total_tiles, d_total_tiles,
                                                               ctrl, status, sc);
}

// ---------------------------------------------------------------------------
// Decompression.
// ---------------------------------------------------------------------------

// payload words of the 4 headers packed in one dword
__device__ __forceinline__ uint32_t words4(uint32_t x)
{
    const uint32_t rate = x & 0x1f1f1f1fu;
    const uint32_t plain = ((~x >> 7) & 0x01010101u) * 0x1fu;   // 0x1f in bytes of plain blocks
    const uint32_t w = rate + (rate & plain);                     // rate * (plain ? 2 : 1), <= 62 per byte
    return (w * 0x01010101u) >> 24;                               // byte sum <= 248
}

// Payload words of the 16 header bytes at [off, off + 16) of an item; bytes at
// or past the last block count nothing (the last piece reads dwords below R).
__device__ __forceinline__ uint32_t words16(const Item &it, uint64_t off)
{
    if (off >= it.nblocks) return 0;
    uint4 x;
    if (off + 16 <= it.nblocks) {
        x = ld16_a4(it.s_in + off);
    } else {
        const uint32_t *s32 = (const uint32_t *)(it.s_in + off);
        uint32_t t[4] = {0, 0, 0, 0};
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const uint64_t at = off + 4u * (uint32_t)d;
            if (at < it.nblocks) {
                const uint64_t keep = it.nblocks - at;
                t[d] = keep < 4 ? s32[d] & (0xffffffffu >> (32 - 8 * (uint32_t)keep)) : s32[d];
            }
        }
        x = make_uint4(t[0], t[1], t[2], t[3]);
    }
    return words4(x.x) + words4(x.y) + words4(x.z) + words4(x.w);
}

// Payload offset (in words after the header section) of tile u of a single
// item: the word the offset phase or the pre-pass wrote, tagged with this
// launch's epoch and the inclusive flag (a cleared word, epoch 0 and flag
// empty, never passes).  Waits, bounded, while the word is not yet of this
// launch (its chunk is still being resolved); on a timeout the launch is
// flagged and an offset past any stream is returned, so the tile decodes to
// zeros.
// The word is loaded early with a plain load (off_word) and checked only
// where the offset is used (off_value), so the check never waits for a load
// in flight; a stale cached copy carries another epoch and is read again.
// The early load is a scalar load (through the constant address space): its
// wait counts scalar loads and LDS operations only.  A vector load of the
// uniform word is moved into scalar registers where it is issued, or copied
// between loop iterations, and on gfx9 either waits for every vector memory
// operation in flight, the previous tile's stores included (uint32 decode
// 1.49 -> 1.98 ms).  The scalar cache does not see other workgroups' writes
// made during the launch: such a stale word fails the epoch check and is read
// again with a vector load in off_value.
typedef const __attribute__((address_space(4))) uint64_t const_u64;
__device__ __forceinline__ uint64_t off_word(const uint64_t *tile_off, uint32_t u)
{
    return ((const_u64 *)tile_off)[u];
}

// The offset (uniform: every lane checked the same word) in scalar registers.
__device__ __forceinline__ uint64_t off_value(const uint64_t *tile_off, uint32_t u, uint64_t w, uint32_t epoch,
                                              Ctrl *ctrl)
{
    const uint32_t want = (epoch << 2) | LB_INCL;   // bits 36.. of the word
    uint64_t v;
    for (uint32_t spins = 1;; spins++) {
        if ((uint32_t)(w >> 36) == want) {
            v = w & kValMask;
            break;
        }
        if (spins > kSpinLimit ||
            ((spins & 63u) == 0 &&
             (__hip_atomic_load(&ctrl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ERRF_TIMEOUT))) {
            __hip_atomic_fetch_or(&ctrl->error, ERRF_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            v = ~0ull >> 3;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        w = ld_relaxed(tile_off + u);
    }
    return ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
           (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
}

__device__ __forceinline__ uint64_t tile_offset(const uint64_t *tile_off, uint32_t u, uint32_t epoch, Ctrl *ctrl)
{
    return off_value(tile_off, u, off_word(tile_off, u), epoch, ctrl);
}

// First phase of a decode launch (a single volume, or a batch: below): the
// payload offset of every tile, from the header bytes alone (1/128 of the
// volume).  Chunks of 128
// tiles come from their own counter, so every chunk is taken before any
// workgroup starts decoding.  TILE/128 threads sum 8 16-byte header pieces
// each of one tile (each load instruction of a wave reads 64 contiguous header
// bytes of each of 64 * 128 / TILE tiles), the tile sums of a chunk are
// scanned (one wave, two tiles per lane), chunks are chained by the look-back
// (one status word per chunk), and tile_off[t] = lb_word(epoch, LB_INCL,
// offset of tile t).  A 600 x 1813 x 1558 volume at tile 512 has 406 chunks,
// fewer than the 512 workgroups of a full grid.  Called by every thread;
// barriers inside.  Returns the workgroup's first chunk ticket (uniform): its
// place in the start order, which the decode uses as its class ticket instead
// of a second launch-wide atomic.  When the grid covers every chunk in the
// first round no workgroup takes a second ticket: each of these start-up
// atomics goes to one address from every workgroup at once.
//
// Batches (kTable): the tiles of every item in global tile order, each tile's
// header bytes read from its own item; tile_off[t] is then the prefix over all
// tiles before t, and a tile's offset within its item is tile_off[t] minus
// tile_off of the item's first tile.
template <int TILE, bool kTable = false, class SH>
__device__ __forceinline__ uint32_t offset_phase(const Item &single, const Item *items, const uint32_t *tile_items,
                                                 uint32_t ntot, Ctrl *ctrl, uint64_t *status, uint64_t *tile_off,
                                                 uint32_t epoch, uint32_t win, SH &sh)
{
    constexpr uint32_t kChunk = 128;          // tiles per chunk
    constexpr uint32_t kPer = TILE / kChunk;  // threads per tile; 8 pieces each cover its TILE header bytes
    static_assert(kPer >= 1 && kPer * 8 * 16 == TILE, "offset phase: tile of 128, 256 or 512 blocks");
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t nchunks = (ntot + kChunk - 1) / kChunk;
    if (tid == 0) sh.chunk = __hip_atomic_fetch_add(&ctrl->chunk.v, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t ticket = sh.chunk;
    for (;;) {
        const uint32_t chunk = sh.chunk;
        if (chunk >= nchunks) break;
        const uint32_t j = tid / kPer, q = tid % kPer;
        const uint32_t t = chunk * kChunk + j;   // this thread's tile
        uint32_t sum = 0;
        if (t < ntot) {
            // per lane: the 64 / kPer tiles of a wave may belong to different items
            const Item &it = kTable ? items[tile_items[t]] : single;
            const uint64_t tbase = (uint64_t)(t - it.first_tile) * TILE;   // the tile's first header byte
#pragma unroll
            for (uint32_t k = 0; k < 8; k++) sum += words16(it, tbase + 16ull * (q + k * kPer));
        }
#pragma unroll
        for (uint32_t d = 1; d < kPer; d <<= 1) sum += __shfl_xor(sum, (int)d, 64);
        if (q == 0) sh.tagg[j] = sum;
        __syncthreads();
        uint32_t incl = 0, incl2 = 0;   // inclusive sums of tiles lane and 64 + lane
        if (tid < 64) {
            incl = wave_incl_scan(sh.tagg[lane], lane);
            incl2 = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63) + wave_incl_scan(sh.tagg[64 + lane], lane);
            if (lane == 63) {
                sh.chunk_agg = incl2;
                st_relaxed(status + chunk, lb_word(epoch, chunk == 0 ? LB_INCL : LB_AGG, incl2));
            }
        }
        __syncthreads();
        uint64_t excl = 0;
        if (chunk != 0) {
            excl = lookback_tile<TILE>(status, chunk, 0, epoch, win, sh.lb, ctrl);
            if (tid == 0) st_relaxed(status + chunk, lb_word(epoch, LB_INCL, excl + sh.chunk_agg));
        }
        if (tid < 64) {
            const uint32_t t = chunk * kChunk + lane;
            if (t < ntot) st_relaxed(tile_off + t, lb_word(epoch, LB_INCL, excl + incl - sh.tagg[lane]));
            if (t + 64 < ntot)
                st_relaxed(tile_off + t + 64, lb_word(epoch, LB_INCL, excl + incl2 - sh.tagg[64 + lane]));
        }
        __syncthreads();
        if (nchunks <= gridDim.x) break;   // tickets 0 .. gridDim.x - 1 went out in the first round
        if (tid == 0)
            sh.chunk = __hip_atomic_fetch_add(&ctrl->chunk.v, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
    }
    return ticket;
}

// Thread 0: the class of a workgroup whose start ticket is known.
__device__ __forceinline__ uint32_t class_of(uint32_t ticket, Sched sc)
{
    return sc.nc <= 1 ? 0u : ticket % sc.nc;
}

// kOff: the tile offsets are precomputed in tile_off (by the launch's offset
// phase or the pre-pass kernel); otherwise each tile finds its offset with a
// look-back (batches, and lscomp_set_decode_prepass(0)).
template <int TILE, bool kTable, bool kOff = false>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(4))) void decompress_kernel(const Item single, const Item *__restrict__ items,
                                                          const uint32_t *__restrict__ tile_items,
                                                          uint32_t total_tiles, const uint32_t *d_total_tiles,
                                                          Ctrl *ctrl, uint64_t *status,
                                                          uint64_t *__restrict__ tile_off, Sched sc, uint32_t phase)
{
    __shared__ TileShared<TILE> sh;
    const uint32_t tid = threadIdx.x;
    if (tid == 0) {
        sh.epoch = __hip_atomic_load(&ctrl->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sh.ntot = d_total_tiles ? *d_total_tiles : total_tiles;
    }
    __syncthreads();
    const uint32_t epoch = sh.epoch, ntot = sh.ntot;
    uint32_t ticket = 0;
    if (phase)
        ticket = offset_phase<TILE, kTable>(single, items, tile_items, ntot, ctrl, status, tile_off, epoch, sc.win, sh);
    if (tid == 0) {
        sh.cls = phase ? class_of(ticket, sc) : tile_class(ctrl, sc);
        const TileQ q0{sh.cls, sc.nc};
        // with offsets the first two ids come from one atomic (raw, converted
        // after the barrier); with the look-back ids are taken one at a time
        sh.tile = kOff ? q0.take(ctrl, 2u) : q0.id(q0.take(ctrl));
    }
    __syncthreads();
    const TileQ q = uniform_queue(sh.cls, sc);
    if (kOff) {
        const uint32_t k0 = sh.tile;
        __syncthreads();   // every thread has read the raw value
        if (tid == 0) {
            sh.tile = q.id(k0);
            sh.next = k0 + 1u;   // raw (TileQ)
        }
        __syncthreads();
    }

    // With precomputed tile offsets (no look-back) the header byte and the
    // offset of the next tile are loaded while the current tile decodes, and
    // tile ids are taken two ahead (the id taken in an iteration is used two
    // iterations later, so the atomic's round trip never stalls a wave); that
    // is safe because no tile waits for another.  With the look-back, ids are
    // taken one at a time.
    // The next tile's header byte crosses the iteration in LDS (this thread's
    // word of sh.hdr, unused by the decode): a register carried around the
    // loop is copied at the loop's end, and that copy waited for every store
    // of the iteration.
    uint32_t *const hdr_slot = &sh.hdr[0][0] + tid;
    uint32_t hdr = 0;
    uint64_t w_next = 0;   // offset word of the next tile, checked when used
    uint64_t wb_next = 0;  // batches: offset word of the next tile's item's first tile (its base)
    if (kOff && sh.tile < ntot) {
        const Item &it = tile_item<kTable>(single, items, tile_items, sh.tile);
        const uint32_t b = (sh.tile - it.first_tile) * (uint32_t)TILE + tid;
        hdr = b < it.nblocks ? (uint32_t)it.s_in[b] : 0u;
        w_next = off_word(tile_off, sh.tile);
        if (kTable) wb_next = off_word(tile_off, it.first_tile);
    }
    if (kOff) *hdr_slot = hdr;

    for (;;) {
        const uint32_t tile = sh.tile;
        if (tile >= ntot) break;
        const uint32_t ntile = kOff ? q.id(sh.next) : ntot;
        uint32_t after_next = 0;   // thread 0: the raw id two tiles ahead
        if (kOff && tid == 0) after_next = q.take(ctrl);
        const Item &it = tile_item<kTable>(single, items, tile_items, tile);
        const uint32_t tl = tile - it.first_tile;
        const uint32_t b = tl * (uint32_t)TILE + tid;
        const bool valid = b < it.nblocks;
        hdr = kOff ? *hdr_slot : valid ? (uint32_t)it.s_in[b] : 0u;

        // with precomputed tile offsets no look-back is needed (in a batch the
        // offsets run over every item's tiles: less the item's first tile's)
        uint64_t excl = kOff ? off_value(tile_off, tile, w_next, epoch, ctrl) : 0ull;
        if (kOff && kTable) excl -= off_value(tile_off, it.first_tile, wb_next, epoch, ctrl);
        const uint32_t words = header_payload_bytes(hdr) >> 2;
        bool bad = (hdr & 0x1fu) > 16u;

        uint32_t agg;
        const uint32_t off = tile_scan<TILE>(words, tid, sh.wsum[0], &agg);
        if (!kOff) {
            if (tid == 0) st_relaxed(status + tile, lb_word(epoch, tl == 0 ? LB_INCL : LB_AGG, agg));
            excl = 0;
            if (tl != 0) {
                excl = lookback_tile<TILE>(status, tile, it.first_tile, epoch, sc.win, sh.lb, ctrl);
                if (tid == 0) st_relaxed(status + tile, lb_word(epoch, LB_INCL, excl + agg));
            }
        }
        const uint64_t g0 = (uint64_t)it.hdr_bytes + 4ull * excl;
        const bool fits = agg <= (uint32_t)TILE * 32u && g0 + 4ull * agg <= it.s_len;
        // the tile's payload lands in LDS at word `shift` (g0 mod 16, in words)
        const uint32_t shift = (uint32_t)(g0 & 15) >> 2;
        // next tile's header byte and offset, issued before this tile's payload
        // loads and waited for with them, before this tile's stores: a wait for
        // a load issued after the stores would wait for the stores too (one
        // in-order counter for loads and stores)
        const uint32_t hdr_cur = hdr;
        uint32_t hdr_next = 0;
        if (ntile < ntot) {
            const Item &nit = tile_item<kTable>(single, items, tile_items, ntile);
            const uint32_t nb = (ntile - nit.first_tile) * (uint32_t)TILE + tid;
            hdr_next = nb < nit.nblocks ? (uint32_t)nit.s_in[nb] : 0u;
            if (kOff) w_next = off_word(tile_off, ntile);
            if (kOff && kTable) wb_next = off_word(tile_off, nit.first_tile);
        }
        if (fits) load_payload<TILE>(it.s_in, it.s_len, g0, agg, sh.pay, tid);
        if (kOff) *hdr_slot = hdr_next;   // waits for the byte here, before this tile's stores
        if (kOff && tid == 0) sh.next = after_next;   // every thread read sh.next before the scan's barrier
        __syncthreads();

        if (valid) {
            uint32_t D[32];
            bad = bad || !fits;
            if (!bad && (hdr_cur & 0x1fu)) {
                decode_block(&sh.pay[shift + off], hdr_cur, it.dec.b, D);
            } else {
#pragma unroll
                for (int i = 0; i < 32; i++) D[i] = 0;
            }
            uint32_t bx, by, bz;
            block_coords(it, b, bx, by, bz);
            // the row stores go out at raised wave priority: lossy CSSI decode
            // 0.721 -> 0.713 ms (the same in the pipelined decoder took its
            // lossless decode 0.810 -> 0.854 ms, and compression's stores
            // gained nothing)
            __builtin_amdgcn_s_setprio(2);
            store_block<false>(it, bx, by, bz, D);
            __builtin_amdgcn_s_setprio(0);
            if (bad) __hip_atomic_fetch_or(&ctrl->error, ERRF_CORRUPT, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        if (tid == 0) {
            if (kOff) {
                sh.tile = ntile;
            } else {
                sh.tile = q.id(q.take(ctrl));   // no tile is taken ahead here
            }
        }
        __syncthreads();
    }
    finish_launch(ctrl, epoch);
}

// Decompression of a single item of at most kSmallTiles tiles: one workgroup
// per tile, which sums the payload sizes of every block before its tile from
// their header bytes itself (at most 255 * TILE bytes, 16-byte pieces strided
// over the workgroup, four loads in flight per thread).  No workgroup waits
// for another and no counter or epoch is used, so the launch is a chain of
// three memory round trips (header bytes, payload, row stores) instead of the
// offset phase's chunk ticket, header pass, look-back and tile ids ahead of
// the decode.  The header bytes before a tile are read once per later tile,
// but the whole header section is at most 128 KB and stays in L2.
template <class C, int TILE>
__device__ __forceinline__ void decompress_small_tiles(const Item &it, Ctrl *ctrl)
{
    __shared__ TileShared<TILE, C::kMaxWords> sh;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t t = blockIdx.x;
    const uint32_t b = t * (uint32_t)TILE + tid;
    const bool valid = b < it.nblocks;
    const uint32_t hdr = valid ? (uint32_t)it.s_in[b] : 0u;
    // payload words before this tile: at most 255 tiles of 512 blocks of 32
    // words (or 256 blocks of 62), so 32 bits hold every partial sum
    uint32_t before = 0;
    const uint32_t npieces = t * ((uint32_t)TILE / 16u);
    for (uint32_t i = tid; i < npieces; i += 4u * TILE) {
        uint32_t w[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t j = i + k * (uint32_t)TILE;
            w[k] = j < npieces ? words16(it, 16ull * j) : 0u;
        }
        before += (w[0] + w[1]) + (w[2] + w[3]);
    }
    const uint32_t wtot = (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(before, lane), 63);
    if (lane == 0) sh.wsum[1][wave] = wtot;
    const uint32_t words = header_payload_bytes(hdr) >> 2;
    uint32_t agg;
    const uint32_t off = tile_scan<TILE>(words, tid, sh.wsum[0], &agg);   // its barrier also covers wsum[1]
    uint32_t excl = 0;
#pragma unroll
    for (int w = 0; w < TILE / 64; w++) excl += sh.wsum[1][w];
    const uint64_t g0 = (uint64_t)it.hdr_bytes + 4ull * excl;
    const bool fits = agg <= (uint32_t)TILE * C::kMaxWords && g0 + 4ull * agg <= it.s_len;
    if (fits) load_payload<TILE>(it.s_in, it.s_len, g0, agg, sh.pay, tid);
    __syncthreads();
    if (valid) {
        const bool bad = C::bad_header(hdr) || !fits;
        uint32_t D[C::kRegs];
        if (!bad && (hdr & 0x1fu)) {
            C::decode(&sh.pay[((uint32_t)(g0 & 15) >> 2) + off], hdr, it, D);
        } else {
#pragma unroll
            for (int i = 0; i < C::kRegs; i++) D[i] = 0;
        }
        uint32_t bx, by, bz;
        block_coords(it, b, bx, by, bz);
        C::store(it, bx, by, bz, D);
        if (bad) __hip_atomic_fetch_or(&ctrl->error, ERRF_CORRUPT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int TILE>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(4))) void decompress_small_kernel(
    const Item it, Ctrl *ctrl)
{
    decompress_small_tiles<CodecU16, TILE>(it, ctrl);
}

// Decompression of a single item with precomputed tile offsets, pipelined so
// that a tile's payload loads overlap the decoding of the tile before it.  The
// iteration that decodes tile t (payload in the LDS ring at t_start):
//   1. scans the header bytes of tile t1 (loaded one iteration ahead): word
//      offsets and total; the id of tile t2 arrives with the same barrier
//   2. issues t2's header byte and offset and t1's payload loads (up to 4
//      chunks of 16 bytes per thread) into registers
//   3. decodes t's blocks from LDS into registers
//   4. writes t1's payload into the ring beside t's, then stores t's blocks
//      (gfx9 counts loads and stores in one in-order counter, so waiting for
//      the payload after the stores also waited for the stores: 0.818 ->
//      0.813 ms lossless CSSI decode in this order)
//   5. barrier.  A payload that does not fit beside t's, or needs more than 4
//      chunks per thread, is staged after the barrier instead (then one more).
// Tile ids are taken three ahead: no tile waits for another here.  Chosen for
// payload-dense streams (see use_pipe_decode); on sparse ones, where a tile's
// payload is a few hundred bytes, the decode with the payload loaded in its own
// iteration was 2-5 % faster in timings.
template <class C, int TILE>
__device__ __forceinline__ void decompress_pipe_tiles(const Item it_arg, Ctrl *ctrl, uint64_t *status,
                                                      uint64_t *__restrict__ tile_off, Sched sc, uint32_t phase)
{
    const Item &it = it_arg;
    constexpr uint32_t kMax = (uint32_t)TILE * C::kMaxWords;   // payload words of a tile at most
    constexpr uint32_t kRing = kMax + 8u;                     // staging words (sh.pay)
    __shared__ TileShared<TILE, C::kMaxWords> sh;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t ntot = it.ntiles;
    if (tid == 0) sh.epoch = __hip_atomic_load(&ctrl->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t epoch = sh.epoch;
    uint32_t ticket = 0;
    if (phase)
        ticket = offset_phase<TILE>(it, nullptr, nullptr, it.ntiles, ctrl, status, tile_off, epoch, sc.win, sh);
    if (tid == 0) {
        sh.cls = phase ? class_of(ticket, sc) : tile_class(ctrl, sc);
        const TileQ q0{sh.cls, sc.nc};
        sh.tile = q0.take(ctrl, 3u);   // raw values k0 .. k0 + 2: this tile, the next, the one after
    }
    __syncthreads();
    const TileQ q = uniform_queue(sh.cls, sc);
    const uint32_t k0 = sh.tile;
    uint32_t pend = k0 + 2u;   // thread 0: the raw id of the tile two ahead of the next one (TileQ)
    uint32_t t = q.id(k0), t1 = q.id(k0 + 1u);

    // header byte of this thread's block in tile u, and u's payload byte offset
    const auto hdr_of = [&](const Item &it, uint32_t u) -> uint32_t {
        const uint32_t b = u * (uint32_t)TILE + tid;
        return (u < ntot && b < it.nblocks) ? (uint32_t)it.s_in[b] : 0u;
    };
    const auto g0_of = [&](const Item &it, uint32_t u) -> uint64_t {
        return u < ntot ? (uint64_t)it.hdr_bytes + 4ull * tile_offset(tile_off, u, epoch, ctrl) : 0ull;
    };
    // from a word loaded earlier (off_word), checked now
    const auto g0_from = [&](const Item &it, uint32_t u, uint64_t w) -> uint64_t {
        return u < ntot ? (uint64_t)it.hdr_bytes + 4ull * off_value(tile_off, u, w, epoch, ctrl) : 0ull;
    };

    // tile t: scan, then its payload staged at the start of the ring; t1's
    // offset word is checked where its payload loads need it
    uint32_t hdr = hdr_of(it, t), hdr1 = hdr_of(it, t1);
    uint64_t g0 = g0_of(it, t), w1 = t1 < ntot ? off_word(tile_off, t1) : 0ull;
    uint32_t agg = 0, t_start = 0;
    uint32_t off = tile_scan<TILE>(header_payload_bytes(hdr) >> 2, tid, sh.wsum[0], &agg);
    bool fits = agg <= kMax && g0 + 4ull * agg <= it.s_len;
    if (t < ntot && fits) load_payload<TILE>(it.s_in, it.s_len, g0, agg, sh.pay, tid);
    __syncthreads();

    while (t < ntot) {
        const Item &it = arg_item();   // read near its uses (arg_item); uint32 decode 1.52 -> 1.47 ms
        // 1. scan of t1's headers; t2's id
        const uint32_t words1 = header_payload_bytes(hdr1) >> 2;
        const uint32_t incl = wave_incl_scan(words1, lane);
        if (lane == 63) sh.wsum[0][wave] = incl;
        if (tid == 0) sh.next = pend;
        __syncthreads();
        const uint32_t t2 = q.id(__builtin_amdgcn_readfirstlane(sh.next));
        uint32_t before = 0, agg1 = 0;
#pragma unroll
        for (int v = 0; v < TILE / 64; v++) {
            const uint32_t x = sh.wsum[0][v];
            before += (v < (int)wave) ? x : 0u;
            agg1 += x;
        }
        const uint32_t off1 = before + incl - words1;
        if (tid == 0) pend = q.take(ctrl);

        // 2. t2's header byte and offset word, t1's payload loads (in this
        //    order: the wait for the payload in step 4 then covers all of them)
        const uint32_t hdr2 = hdr_of(it, t2);
        const uint64_t w2 = t2 < ntot ? off_word(tile_off, t2) : 0ull;
        const uint64_t g1 = g0_from(it, t1, w1);
        const bool fits1 = t1 < ntot && agg1 <= kMax && g1 + 4ull * agg1 <= it.s_len;
        const uint64_t a1 = g1 & ~15ull;
        const uint32_t nchunk1 = fits1 ? (((uint32_t)(g1 - a1) >> 2) + agg1 + 3) >> 2 : 0u;
        const uint32_t len0 = fits ? 4u * ((((uint32_t)(g0 & 15) >> 2) + agg + 3) >> 2) : 0u;
        // ring place of t1 beside t: after it, or at 0 when t starts far enough in
        uint32_t t1_start = ~0u;
        if (nchunk1 <= 4u * TILE) {
            if (t_start + len0 + 4u * nchunk1 <= kRing) t1_start = t_start + len0;
            else if (4u * nchunk1 <= t_start) t1_start = 0;
        }
        // the payload loads go out at raised wave priority (s_setprio): issued
        // ahead of other waves' decode work they overlap more of it (lossless
        // CSSI decode 0.857 -> 0.840 ms; compression's loads gained nothing)
        uint4 V[4];
        __builtin_amdgcn_s_setprio(3);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t c = (uint32_t)k * TILE + tid;
            V[k] = c < nchunk1 ? load_chunk(it.s_in, it.s_len, a1, c) : make_uint4(0, 0, 0, 0);
        }
        __builtin_amdgcn_s_setprio(0);

        // 3. decode t
        const uint32_t b = t * (uint32_t)TILE + tid;
        const bool valid = b < it.nblocks;
        const bool bad = C::bad_header(hdr) || !fits;
        uint32_t D[C::kRegs];
        if (valid && !bad && (hdr & 0x1fu)) {
            C::decode(&sh.pay[t_start + ((uint32_t)(g0 & 15) >> 2) + off], hdr, it, D);
        } else {
#pragma unroll
            for (int i = 0; i < C::kRegs; i++) D[i] = 0;
        }

        // 4. t1's payload into the ring, then t's stores
        if (t1_start != ~0u) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t c = (uint32_t)k * TILE + tid;
                if (c < nchunk1) *(uint4 *)(sh.pay + t1_start + 4 * c) = V[k];
            }
        }
        if (valid) {
            uint32_t bx, by, bz;
            block_coords(it, b, bx, by, bz);
            C::store(it, bx, by, bz, D);
            if (bad) __hip_atomic_fetch_or(&ctrl->error, ERRF_CORRUPT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();   // t's payload has left LDS; t1's is visible
        if (t1_start == ~0u) {   // 5. t1 did not fit beside t (or is large): stage it now
            t1_start = 0;
            if (fits1) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t c = (uint32_t)k * TILE + tid;
                    if (c < nchunk1) *(uint4 *)(sh.pay + 4 * c) = V[k];
                }
                load_payload<TILE>(it.s_in, it.s_len, g1, agg1, sh.pay, tid, 4u * TILE);
            }
            __syncthreads();
        }
        t = t1; hdr = hdr1; g0 = g1; agg = agg1; off = off1; fits = fits1; t_start = t1_start;
        t1 = t2; hdr1 = hdr2; w1 = w2;
    }
    finish_launch(ctrl, epoch);
}

template <int TILE>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(4))) void decompress_pipe_kernel(
    const Item it, Ctrl *ctrl, uint64_t *status, uint64_t *__restrict__ tile_off, Sched sc, uint32_t phase)
{
    decompress_pipe_tiles<CodecU16, TILE>(it, ctrl, status, tile_off, sc, phase);
}

// ---------------------------------------------------------------------------
// Decode pre-pass: payload offset of every tile of a single item, from the
// header bytes alone.  A workgroup of 256 threads takes a chunk of 64 tiles
// (4 threads per tile, TILE/4 header bytes each, read as 16-byte pieces so
// that a wave's load covers 64 contiguous bytes of each of 16 tiles), reduces
// sizes per tile, scans them, and chains chunks with the same look-back.
// tile_off[t] is the word offset of tile t's payload after the header
// section, tagged with the epoch of the decode launch that follows (this
// launch's epoch + 1: its last workgroup advances the epoch by one).
// ---------------------------------------------------------------------------
struct OffShared {
    uint32_t tagg[64];
    LbShared<4 * lb_per(256)> lb;
    uint32_t tile, next, epoch, ntot;
    uint32_t excl_lo, excl_hi;
};

template <int TILE>
__global__ __launch_bounds__(256) void offsets_kernel(const Item it, Ctrl *ctrl, uint64_t *status,
                                                     uint64_t *__restrict__ tile_off)
{
    __shared__ OffShared sh;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t nchunks = (it.ntiles + 63) / 64;
    const TileQ tq = {0u, 1u};   // one counter: chunks are few
    if (tid == 0) {
        sh.epoch = __hip_atomic_load(&ctrl->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sh.tile = tq.id(tq.take(ctrl));
    }
    __syncthreads();
    const uint32_t epoch = sh.epoch;
    constexpr uint32_t kBytes = TILE / 4;   // header bytes per thread

    for (;;) {
        const uint32_t chunk = sh.tile;
        if (chunk >= nchunks) break;
        // thread (j, q) sums the 16-byte pieces q, q+4, q+8, ... of tile j: each
        // load instruction of a wave reads 64 contiguous bytes of 16 tiles
        const uint32_t j = tid >> 2, q = tid & 3u;
        const uint32_t tile = chunk * 64 + j;
        const uint64_t tbase = (uint64_t)tile * TILE;   // the tile's first header byte
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < kBytes / 16; k++) {
            const uint64_t off = tbase + 16ull * (4 * k + q);
            if (off >= it.nblocks) break;
            uint4 x;
            if (off + 16 <= it.nblocks) {
                x = ld16_a4(it.s_in + off);
            } else {   // the section's last piece: dwords below R, bytes past the last block cleared
                const uint32_t *s32 = (const uint32_t *)(it.s_in + off);
                uint32_t t[4] = {0, 0, 0, 0};
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    const uint64_t at = off + 4u * (uint32_t)d;
                    if (at < it.nblocks) {
                        const uint64_t keep = it.nblocks - at;
                        t[d] = keep < 4 ? s32[d] & (0xffffffffu >> (32 - 8 * (uint32_t)keep)) : s32[d];
                    }
                }
                x = make_uint4(t[0], t[1], t[2], t[3]);
            }
            sum += words4(x.x) + words4(x.y) + words4(x.z) + words4(x.w);
        }
        sum += __shfl_xor(sum, 1, 64);
        sum += __shfl_xor(sum, 2, 64);
        if (q == 0) sh.tagg[j] = sum;
        __syncthreads();
        uint32_t incl = 0, agg = 0;
        if (tid < 64) {
            incl = wave_incl_scan(sh.tagg[lane], lane);
            agg = __shfl(incl, 63, 64);
        }
        // chunk aggregate published by thread 0, look-back over chunks
        if (tid == 0) {
            sh.excl_lo = agg;
            st_relaxed(status + chunk, lb_word(epoch, chunk == 0 ? LB_INCL : LB_AGG, agg));
        }
        __syncthreads();
        const uint32_t cagg = sh.excl_lo;
        uint64_t excl = 0;
        if (chunk != 0) {
            excl = lookback_tile<256>(status, chunk, 0, epoch, default_window(256), sh.lb, ctrl);
            if (tid == 0) st_relaxed(status + chunk, lb_word(epoch, LB_INCL, excl + cagg));
        }
        if (tid < 64) {
            const uint32_t t = chunk * 64 + lane;
            if (t < it.ntiles)
                st_relaxed(tile_off + t, lb_word((epoch + 1u) & kEpochMask, LB_INCL, excl + incl - sh.tagg[lane]));
        }
        __syncthreads();
        if (tid == 0) sh.tile = tq.id(tq.take(ctrl));
        __syncthreads();
    }
    finish_launch(ctrl, epoch);
}

__global__ void fill_u64_kernel(uint64_t *p, uint64_t n, uint64_t value)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// Pre-pass kernel for the tile offsets of a single-item decode (OFF_KERNEL).
template <int TILE>
static void launch_offsets(const Item &it, const LaunchCfg &cfg, Ctrl *ctrl, uint64_t *status, hipStream_t stream)
{
    const uint32_t nchunks = (it.ntiles + 63) / 64;
    const uint32_t g = nchunks < cfg.grid ? nchunks : cfg.grid;
    hipLaunchKernelGGL((offsets_kernel<TILE>), dim3(g), dim3(256), 0, stream, it, ctrl, status, cfg.tile_off);
}

// The pipelined decode pays off once payloads are dense: measured on one
// MI355X (600x1813x1558 CSSI, 512x1024x1024 XPCS), it was 8-10 % faster at
// 6-16 payload bytes per block and 1-5 % slower at 0-1 byte per block (where
// a tile's payload load is short and the stores dominate).
static bool use_pipe_decode(const Item &it)
{
    return dense_stream(it);
}

template <int TILE>
static hipError_t launch_t(bool compress, const Item &single, const Item *items, const uint32_t *tile_items, LaunchCfg cfg,
                           Ctrl *ctrl, uint64_t *status, hipStream_t stream)
{
    const dim3 grid(cfg.grid), block(TILE);
    const Sched sc = cfg.sched;
    if (compress) {
        if (items)
            hipLaunchKernelGGL((compress_kernel<TILE, true>), grid, block, 0, stream, single, items, tile_items,
                               cfg.total_tiles, cfg.d_total_tiles, ctrl, status, sc);
        else if (single.dec.flags & DEC_LOSSLESS)
            hipLaunchKernelGGL((compress_kernel<TILE, false, KIND_LOSSLESS>), grid, block, 0, stream, single, items,
                               tile_items, cfg.total_tiles, cfg.d_total_tiles, ctrl, status, sc);
        else if (single.dec.flags & DEC_Q24)
            hipLaunchKernelGGL((compress_kernel<TILE, false, KIND_Q24>), grid, block, 0, stream, single, items,
                               tile_items, cfg.total_tiles, cfg.d_total_tiles, ctrl, status, sc);
        else
            hipLaunchKernelGGL((compress_kernel<TILE, false>), grid, block, 0, stream, single, items, tile_items,
                               cfg.total_tiles, cfg.d_total_tiles, ctrl, status, sc);
    } else {
        if (items && cfg.tile_off)
            hipLaunchKernelGGL((decompress_kernel<TILE, true, true>), grid, block, 0, stream, single, items,
                               tile_items, cfg.total_tiles, cfg.d_total_tiles, ctrl, status, cfg.tile_off, sc, 1u);
        else if (items)
            hipLaunchKernelGGL((decompress_kernel<TILE, true>), grid, block, 0, stream, single, items, tile_items,
                               cfg.total_tiles, cfg.d_total_tiles, ctrl, status, (uint64_t *)nullptr, sc, 0u);
        else if (single.ntiles <= kSmallTiles)
            hipLaunchKernelGGL((decompress_small_kernel<TILE>), dim3(single.ntiles), block, 0, stream, single, ctrl);
        else
        {
            const uint32_t phase = cfg.tile_off && cfg.off_mode == OFF_PHASE ? 1u : 0u;
            if (cfg.tile_off && !phase) launch_offsets<TILE>(single, cfg, ctrl, status, stream);
            if (cfg.tile_off && use_pipe_decode(single))
                hipLaunchKernelGGL((decompress_pipe_kernel<TILE>), grid, block, 0, stream, single, ctrl, status,
                                   cfg.tile_off, sc, phase);
            else if (cfg.tile_off)
                hipLaunchKernelGGL((decompress_kernel<TILE, false, true>), grid, block, 0, stream, single, items,
                                   tile_items, cfg.total_tiles, cfg.d_total_tiles, ctrl, status, cfg.tile_off, sc,
                                   phase);
            else
                hipLaunchKernelGGL((decompress_kernel<TILE, false>), grid, block, 0, stream, single, items,
                                   tile_items, cfg.total_tiles, cfg.d_total_tiles, ctrl, status, cfg.tile_off, sc,
                                   phase);
        }
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Served form: the device-resident request loop (SrvArgs, lscomp_internal.h).
// The workgroup that takes ticket 0 is the dispatcher; every other one runs
// the compression pipeline above on published tiles, from counter class
// (ticket - 1) mod nc.
// ---------------------------------------------------------------------------
template <int TILE>
struct SrvShared {
    TileShared<TILE> t;                               // the pipeline's LDS (the dispatcher uses t.pay as scratch)
    alignas(16) uint32_t itm[4][sizeof(Item) / 4];    // item of each tile in flight, by header buffer
    uint32_t nst[2], nr[2], ntl[2];                   // state of the next tile, by iteration parity
    uint64_t wr, now;
    uint32_t stop, total;
};

// Tile-map word of a published tile: epoch (26 bits) | item + 1 (16 bits) |
// tile within the item (22 bits).  A cleared word (item field 0) never passes.
__device__ __forceinline__ uint64_t map_word(uint32_t epoch, uint32_t r, uint32_t tl)
{
    return ((uint64_t)epoch << 38) | ((uint64_t)(r + 1) << 22) | tl;
}

// The dispatcher: takes new requests from the host ring in order, copies each
// item into the item table (setting its first tile and its size slot), clears
// its tile count, makes all of that visible (barrier, agent-scope release),
// then publishes every new tile's map word.  Closes the instance after
// idle_ticks without a request, on the host's stop word, or when the item
// table or the tile map is full (the requests left over stay in the ring for
// the next instance).  Called by every thread of its workgroup.
template <int TILE>
__device__ __forceinline__ void serve_dispatch(const SrvArgs &a, uint32_t epoch, SrvShared<TILE> &sh)
{
    constexpr uint32_t kW = sizeof(Item) / 4;
    const uint32_t tid = threadIdx.x;
    uint32_t *ex = sh.t.pay;   // first tile of each new request, relative to the instance's tile count
    if (tid == 0) {
        sh.wr = __hip_atomic_load(&a.host->rd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        sh.now = __builtin_amdgcn_s_memrealtime();
    }
    __syncthreads();
    uint64_t rd = sh.wr, last = sh.now;   // next request to take; time of the last one taken
    uint32_t nreq = 0, ntiles = 0;        // items and tiles published by this instance
    __syncthreads();
    for (;;) {
        if (tid == 0) {
            sh.wr = __hip_atomic_load(&a.host->wr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            sh.stop = __hip_atomic_load(&a.host->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            sh.now = __builtin_amdgcn_s_memrealtime();
        }
        __syncthreads();
        const uint64_t wr = sh.wr, now = sh.now;
        const uint32_t stop = sh.stop;
        uint64_t avail = wr > rd ? wr - rd : 0;
        if (avail > (uint64_t)TILE) avail = TILE;
        if (avail > (uint64_t)(kSrvMaxReq - nreq)) avail = kSrvMaxReq - nreq;
        const uint32_t take = (uint32_t)avail;
        if (take == 0) {
            if (stop || now - last > a.idle_ticks || nreq == kSrvMaxReq) break;
            __syncthreads();   // everyone has read sh.* before thread 0 rewrites it
            __builtin_amdgcn_s_sleep(8);
            continue;
        }
        // tile counts of the new requests, one per thread; a request whose
        // tiles would pass the end of the tile map waits for the next instance
        uint32_t nt = 0;
        if (tid < take)
            nt = __hip_atomic_load(&a.ring[(rd + tid) % kSrvRing].item.ntiles, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
        uint32_t agg;
        const uint32_t off = tile_scan<TILE>(nt, tid, sh.t.wsum[0], &agg);
        const bool fit = tid < take && (uint64_t)ntiles + off + nt <= kSrvMaxTiles;
        const uint32_t nfit = (uint32_t)__syncthreads_count(fit);   // a prefix of the new requests
        if (nfit == 0) break;
        if (fit) ex[tid] = off;
        if (tid + 1 == nfit) sh.total = off + nt;
        // copy the items from the ring
        for (uint32_t k = tid; k < nfit * kW; k += TILE) {
            const uint32_t j = k / kW, w = k - j * kW;
            const uint32_t *src = (const uint32_t *)&a.ring[(rd + j) % kSrvRing].item;
            ((uint32_t *)&a.items[nreq + j].it)[w] =
                __hip_atomic_load(src + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        __syncthreads();
        const uint32_t total = sh.total;
        if (tid < nfit) {
            Item &it = a.items[nreq + tid].it;
            it.first_tile = ntiles + ex[tid];
            it.result = a.sizes + nreq + tid;
            a.done[nreq + tid] = 0;
        }
        // items and tile counts reach memory before any map word
        __syncthreads();
        if (tid == 0) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        for (uint32_t t = tid; t < total; t += TILE) {
            uint32_t lo = 0, hi = nfit - 1;   // the request j with ex[j] <= t < ex[j + 1]
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (ex[mid] <= t) lo = mid;
                else hi = mid - 1;
            }
            st_relaxed(a.map + ntiles + t, map_word(epoch, nreq + lo, t - ex[lo]));
        }
        nreq += nfit;
        ntiles += total;
        rd += nfit;
        last = now;
        if (tid == 0) __hip_atomic_store(&a.host->rd, rd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __syncthreads();   // ex[] and sh.* are rewritten next round
    }
    // no tile at or past ntiles will be published by this instance
    if (tid == 0) st_relaxed(&a.dev->close, ((uint64_t)ntiles << 1) | 1u);
}

// A consumer's tile of item r is stored (every thread's stores issued): make
// them visible at system scope, count the tile, and if it was the item's last
// one write the completion record (bytes, status, then seq + 1).  Called by
// every thread; one barrier.
__device__ __forceinline__ void serve_tile_done(const SrvArgs &a, Ctrl *ctrl, uint32_t r, const Item &it)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        const uint32_t old = __hip_atomic_fetch_add(a.done + r, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1 == it.ntiles) {
            // every tile is stored; the last one (in block order) wrote the size
            const uint64_t bytes = ld_relaxed(a.sizes + r);
            uint32_t st = bytes > it.s_len ? SRV_CAPACITY : 0u;
            if (__hip_atomic_load(&ctrl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ERRF_TIMEOUT)
                st |= SRV_TIMEOUT;
            SrvComp *c = a.comp + it.seq % kSrvRing;
            __hip_atomic_store(&c->bytes, bytes, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&c->status, st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&c->seq, it.seq + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// A consumer: the compression pipeline of compress_tiles (same stages, same
// LDS ring) on published tiles.  Before each iteration thread 0 reads the map
// word of the reserved tile n with an agent-scope acquire (the item and the
// frame are read after it; the acquire also drops stale cache lines of a
// frame buffer the host reused).  When n is not yet published the iteration
// finishes pending tiles instead, so no tile waits for one that a waiting
// workgroup holds; with nothing pending the workgroup sleeps, and it leaves
// once the dispatcher has closed the instance below n.
template <int TILE>
__device__ __forceinline__ void serve_consume(const SrvArgs &a, Ctrl *ctrl, uint64_t *status, Sched sc,
                                              uint32_t cls, uint32_t epoch, SrvShared<TILE> &ss)
{
    using C = CodecU16T<true>;
    constexpr uint32_t kRing = (uint32_t)TILE * C::kMaxWords;
    constexpr uint32_t kW = sizeof(Item) / 4;
    TileShared<TILE> &sh = ss.t;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t win = sc.win;
    const TileQ q = {cls, sc.nc};
    if (tid == 0) sh.next = q.take(ctrl);   // raw (TileQ)
    __syncthreads();
    const auto item_in = [&](uint32_t buf) -> const Item & { return *(const Item *)ss.itm[buf]; };

    uint32_t np = 0, hb = 0;
    PendTile p0 = {0, 0, 0, 0, 0, 0}, p1 = {0, 0, 0, 0, 0, 0};
    uint32_t n = q.id(sh.next), par = 0;
    uint32_t late = 0;   // thread 0: rounds waited for n after the instance closed above it
    for (;; par ^= 1u) {
        // 0. is n published? (1: yes, 2: never in this instance, 0: not yet)
        if (tid == 0) {
            uint32_t st = 0;
            if (n < kSrvMaxTiles) {
                const uint64_t w = __hip_atomic_load(a.map + n, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t r1 = (uint32_t)(w >> 22) & 0xffffu;
                if ((uint32_t)(w >> 38) == epoch && r1 != 0) {
                    st = 1;
                    ss.nr[par] = r1 - 1;
                    ss.ntl[par] = (uint32_t)w & 0x3fffffu;
                }
            }
            if (st == 0) {
                const uint64_t c = ld_relaxed(&a.dev->close);
                if ((c & 1u) && (uint64_t)n >= (c >> 1)) st = 2;
                // closed above n, yet n's word never shows: give up (bounded wait)
                if ((c & 1u) && st == 0 && ++late > kSpinLimit) {
                    __hip_atomic_fetch_or(&ctrl->error, ERRF_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    st = 2;
                }
            }
            ss.nst[par] = st;
        }
        __syncthreads();
        const uint32_t st = ss.nst[par];
        const bool have_n = st == 1;
        if (!have_n && np == 0) {
            if (st == 2) break;
            __builtin_amdgcn_s_sleep(4);
            continue;
        }
        const bool resolve = np == 2 || (!have_n && np != 0);
        const uint32_t n_item = have_n ? ss.nr[par] : 0u, n_tl = have_n ? ss.ntl[par] : 0u;
        if (have_n) {   // the item of n into its buffer (last used four tiles ago)
            if (tid < kW) ss.itm[hb][tid] = ((const uint32_t *)&a.items[n_item].it)[tid];
            __syncthreads();
        }

        // 1. loads of n, probe of p0, next id
        uint32_t D[C::kRegs];
        bool nvalid = false;
        if (have_n) {
            const Item &nit = item_in(hb);
            const uint32_t b = n_tl * (uint32_t)TILE + tid;
            nvalid = b < nit.nblocks;
            if (nvalid) {
                uint32_t bx, by, bz;
                block_coords(nit, b, bx, by, bz);
                C::load(nit, bx, by, bz, D);
            }
        }
        const bool p_lb = resolve && p0.tl != 0;
        uint64_t w0[lb_per(TILE)];
        if (p_lb) lb_probe<TILE>(status, (int64_t)p0.tile - 1, item_in(p0.hb).first_tile, epoch, win, w0);
        uint32_t next_k = ~0u;   // raw (TileQ), written to LDS before the scan barrier
        if (tid == 0 && have_n) next_k = q.take(ctrl);

        // 2. header of n; scan and the first half of p0's evaluation share a barrier
        uint32_t n_hdr = 0, n_words = 0;
        if (have_n) {
            if (nvalid) {
                n_hdr = C::header(D, item_in(hb), ctrl);
                n_words = header_payload_bytes(n_hdr) >> 2;
            }
            ((uint8_t *)sh.hdr[hb])[tid] = (uint8_t)n_hdr;
        }
        const uint32_t incl = wave_incl_scan(n_words, lane);
        if (lane == 63) sh.wsum[par][wave] = incl;
        if (p_lb) lb_slices<TILE>(w0, epoch, win, sh.lbf, par);
        if (tid == 0) sh.next2[par] = next_k;
        __syncthreads();
        const uint32_t next_raw = sh.next2[par];
        const uint32_t n_next = next_raw == ~0u ? n : q.id(next_raw);
        uint32_t before = 0, n_agg = 0;
#pragma unroll
        for (int v = 0; v < TILE / 64; v++) {
            const uint32_t x = sh.wsum[par][v];
            before += (v < (int)wave) ? x : 0u;
            n_agg += x;
        }
        const uint32_t n_off = before + incl - n_words;
        if (have_n && tid == 0) st_relaxed(status + n, lb_word(epoch, n_tl == 0 ? LB_INCL : LB_AGG, n_agg));

        // 3. pack n behind the pending tiles if it fits in the staging ring
        uint32_t n_start = have_n ? ring_place(n_agg, np, p0, p1, kRing) : ~0u;
        if (n_start != ~0u && n_words) C::encode(D, n_hdr, item_in(hb), &sh.pay[n_start + n_off]);

        // 4. finish p0
        if (resolve) {
            const Item &pit = item_in(p0.hb);
            uint64_t excl = 0;
            if (p_lb) {
                const uint32_t r = lb_reduce<TILE>(sh.lbf, par, win, excl);
                if (r != LB_DONE) {
                    const int64_t base = (int64_t)p0.tile - 1 - (r == LB_NEXT ? (int64_t)win : 0);
                    excl = lookback_from<TILE>(status, base, pit.first_tile, epoch, win, sh.lb, ctrl, excl,
                                               r == LB_RETRY ? 1u : 0u, 0u);
                }
                if (tid == 0) st_relaxed(status + p0.tile, lb_word(epoch, LB_INCL, excl + p0.agg));
            }
            store_tile<TILE>(pit, p0.tl, p0.agg, sh.hdr[p0.hb], sh.pay + p0.start, excl, tid);
            serve_tile_done(a, ctrl, p0.item, pit);
            p0 = p1;
            np--;
        }
        if (have_n && n_start == ~0u) {
            __syncthreads();
            for (;;) {
                n_start = ring_place(n_agg, np, p0, p1, kRing);
                if (n_start != ~0u) break;
                const Item &pit = item_in(p0.hb);
                uint64_t excl = 0;
                if (p0.tl != 0) {
                    excl = lookback_tile<TILE>(status, p0.tile, pit.first_tile, epoch, win, sh.lb, ctrl);
                    if (tid == 0) st_relaxed(status + p0.tile, lb_word(epoch, LB_INCL, excl + p0.agg));
                }
                store_tile<TILE>(pit, p0.tl, p0.agg, sh.hdr[p0.hb], sh.pay + p0.start, excl, tid);
                serve_tile_done(a, ctrl, p0.item, pit);
                p0 = p1;
                np--;
                __syncthreads();
            }
            if (n_words) C::encode(D, n_hdr, item_in(hb), &sh.pay[n_start + n_off]);
            __syncthreads();
        }
        if (have_n) {
            const PendTile t = {n, n_tl, n_item, n_agg, n_start, hb};
            if (np == 0) p0 = t; else p1 = t;
            np++;
            hb = (hb + 1u) & 3u;
            n = n_next;
        }
    }
}

template <int TILE>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(4))) void serve_kernel(
    const SrvArgs a, Ctrl *ctrl, uint64_t *status, Sched sc)
{
    __shared__ SrvShared<TILE> sh;
    if (threadIdx.x == 0) {
        sh.t.epoch = __hip_atomic_load(&ctrl->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sh.t.cls = __hip_atomic_fetch_add(&ctrl->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const uint32_t epoch = sh.t.epoch;
    const uint32_t ticket = __builtin_amdgcn_readfirstlane(sh.t.cls);
    __syncthreads();
    if (ticket == 0) serve_dispatch<TILE>(a, epoch, sh);
    else serve_consume<TILE>(a, ctrl, status, sc, (ticket - 1) % sc.nc, epoch, sh);
    finish_launch(ctrl, epoch);
}

hipError_t launch_serve(const SrvArgs &a, uint32_t grid, Sched sc, Ctrl *ctrl, uint64_t *status, hipStream_t stream)
{
    if (grid < 2 || sc.nc < 1 || sc.nc > grid - 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL((serve_kernel<(int)kSrvTile>), dim3(grid), dim3(kSrvTile), 0, stream, a, ctrl, status, sc);
    return hipGetLastError();
}

int workgroups_per_cu_serve()
{
    int per_cu = 0;
    const hipError_t e =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, serve_kernel<(int)kSrvTile>, (int)kSrvTile, 0);
    return (e == hipSuccess && per_cu > 0) ? per_cu : 1;
}

// ---------------------------------------------------------------------------
// uint32 volumes (replace cuLSZ_{compression,decompression}_kernel_uint32_bsize64,
// lsCOMP/src/cuLSZ_kernel.cu:4-824): the uint16 compression and pipelined
// decompression with the uint32 codec (CodecU32: one block of 64 uint32 values
// per lane in 64 VGPRs, up to 62 payload words).  Tile 256: 62 KB of payload
// staging per workgroup, 2 workgroups per CU (2 waves per SIMD, so up to 256
// VGPRs).  The pipelined decode beat staging each payload in its own
// iteration on sparse uint32 streams as well (lossy CSSI: 1.49 against
// 1.68 ms), so it is the only uint32 decode.  A block whose quantized maximum
// needs 32 bits flags the launch (ERRF_RANGE).
// ---------------------------------------------------------------------------
template <int TILE>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(2))) void compress_u32_kernel(
    const Item it, Ctrl *ctrl, uint64_t *status, Sched sc)
{
    compress_tiles<CodecU32, TILE, false>(it, nullptr, nullptr, it.ntiles, nullptr, ctrl, status, sc);
}

template <int TILE>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(2))) void decompress_u32_pipe_kernel(
    const Item it, Ctrl *ctrl, uint64_t *status, uint64_t *__restrict__ tile_off, Sched sc, uint32_t phase)
{
    decompress_pipe_tiles<CodecU32, TILE>(it, ctrl, status, tile_off, sc, phase);
}

template <int TILE>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(2))) void decompress_u32_small_kernel(
    const Item it, Ctrl *ctrl)
{
    decompress_small_tiles<CodecU32, TILE>(it, ctrl);
}

hipError_t launch_u32(bool compress, const Item &it, LaunchCfg cfg, Ctrl *ctrl, uint64_t *status,
                      hipStream_t stream)
{
    constexpr int T = (int)kTileU32, TD = (int)kTileU32Dec;
    if (cfg.tile != (uint32_t)(compress ? T : TD)) return hipErrorInvalidValue;
    const dim3 grid(cfg.grid);
    if (compress) {
        hipLaunchKernelGGL((compress_u32_kernel<T>), grid, dim3(T), 0, stream, it, ctrl, status, cfg.sched);
    } else if (it.ntiles <= kSmallTiles) {
        hipLaunchKernelGGL((decompress_u32_small_kernel<TD>), dim3(it.ntiles), dim3(TD), 0, stream, it, ctrl);
    } else {
        if (!cfg.tile_off) return hipErrorInvalidValue;
        const uint32_t phase = cfg.off_mode == OFF_PHASE ? 1u : 0u;
        if (!phase) launch_offsets<TD>(it, cfg, ctrl, status, stream);
        hipLaunchKernelGGL((decompress_u32_pipe_kernel<TD>), grid, dim3(TD), 0, stream, it, ctrl, status,
                           cfg.tile_off, cfg.sched, phase);
    }
    return hipGetLastError();
}

static hipError_t launch(bool compress, const Item &single, const Item *items, const uint32_t *tile_items, LaunchCfg cfg,
                         Ctrl *ctrl, uint64_t *status, hipStream_t stream)
{
    switch (cfg.tile) {
    case 128: return launch_t<128>(compress, single, items, tile_items, cfg, ctrl, status, stream);
    case 256: return launch_t<256>(compress, single, items, tile_items, cfg, ctrl, status, stream);
    case 512: return launch_t<512>(compress, single, items, tile_items, cfg, ctrl, status, stream);
    case kPipeTile:   // the pipelined decode of one dense volume with the offset phase only
        if (compress || items || !cfg.tile_off || cfg.off_mode != OFF_PHASE || !dense_stream(single))
            return hipErrorInvalidValue;
        hipLaunchKernelGGL((decompress_pipe_kernel<kPipeTile>), dim3(cfg.grid), dim3(kPipeTile), 0, stream, single,
                           ctrl, status, cfg.tile_off, cfg.sched, 1u);
        return hipGetLastError();
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_compress(const Item &single, const Item *items, const uint32_t *tile_items, LaunchCfg cfg, Ctrl *ctrl,
                           uint64_t *status, hipStream_t stream)
{
    return launch(true, single, items, tile_items, cfg, ctrl, status, stream);
}

hipError_t launch_decompress(const Item &single, const Item *items, const uint32_t *tile_items, LaunchCfg cfg, Ctrl *ctrl,
                             uint64_t *status, hipStream_t stream)
{
    return launch(false, single, items, tile_items, cfg, ctrl, status, stream);
}

hipError_t launch_fill_u64(uint64_t *p, uint64_t n, uint64_t value, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    uint64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((uint32_t)g), dim3(256), 0, stream, p, n, value);
    return hipGetLastError();
}

template <int TILE>
static int per_cu_t(bool compress, bool table)
{
    int per_cu = 0;
    hipError_t e;
    if (compress)
        e = table ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, compress_kernel<TILE, true>, TILE, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, compress_kernel<TILE, false>, TILE, 0);
    else
        e = table ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, decompress_kernel<TILE, true>, TILE, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, decompress_pipe_kernel<TILE>, TILE, 0);
    return (e == hipSuccess && per_cu > 0) ? per_cu : 1;
}

int workgroups_per_cu(bool compress, bool table, int tile)
{
    switch (tile) {
    case 128: return per_cu_t<128>(compress, table);
    case 256: return per_cu_t<256>(compress, table);
    case 512: return per_cu_t<512>(compress, table);
    case kPipeTile: {
        int per_cu = 0;
        const hipError_t e =
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, decompress_pipe_kernel<kPipeTile>, kPipeTile, 0);
        return (!compress && !table && e == hipSuccess && per_cu > 0) ? per_cu : 1;
    }
    default: return 1;
    }
}

int workgroups_per_cu_u32(bool compress)
{
    constexpr int T = (int)kTileU32, TD = (int)kTileU32Dec;
    int per_cu = 0;
    const hipError_t e = compress
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, compress_u32_kernel<T>, T, 0)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, decompress_u32_pipe_kernel<TD>, TD, 0);
    return (e == hipSuccess && per_cu > 0) ? per_cu : 1;
}

int num_cus()
{
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return ncu;
}

}  // namespace lsc
