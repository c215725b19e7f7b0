// lscomp_kernels.hip -- gfx950 kernels of the lsCOMP uint16 codec.
//
// One workgroup = one tile of kTile consecutive 8x8 blocks of an item, one
// block per lane.  Compression (replaces cuLSZ_compression_kernel_uint16_bsize64,
// lsCOMP/src/cuLSZ_kernel.cu:827-1283):
//   1. fetch a dynamic tile id (forward progress does not depend on dispatch order);
//   2. each lane loads its block once (8 row loads, coalesced across lanes:
//      lanes are consecutive blocks along z), computes the header byte;
//   3. workgroup scan of payload sizes, tile aggregate published at once;
//   4. payload packed into LDS while predecessors resolve;
//   5. wave-wide decoupled look-back over 64 predecessors per probe, one
//      64-bit word per tile (epoch | flag | value) read and written with
//      relaxed agent-scope atomics, so no fences are needed on the multi-XCD part;
//   6. headers and payload leave LDS as contiguous coalesced dword stores.
// Decompression (replaces cuLSZ_decompression_kernel_uint16_bsize64,
// cuLSZ_kernel.cu:1286-1646) runs the same scan over header sizes, stages the
// tile's payload in LDS with coalesced loads, and each lane unpacks its block
// and writes its rows.  Rate-0 blocks decode to zeros.
#include "lscomp_internal.h"

namespace lsc {

constexpr uint32_t kSpinLimit = 1u << 18;

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

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Exclusive prefix (in words) of tile `tile` over the tiles [first, tile) of
// its item.  Called by all 64 lanes of one wave.
__device__ uint64_t lookback(uint64_t *status, uint32_t tile, uint32_t first, uint32_t epoch,
                             uint32_t lane, Ctrl *ctrl)
{
    uint64_t excl = 0;
    int64_t base = (int64_t)tile - 1;
    uint32_t spins = 0;
    for (;;) {
        const int64_t idx = base - (int64_t)lane;
        const bool inr = idx >= (int64_t)first;
        uint64_t w, incl_m, need;
        for (;;) {
            w = inr ? ld_relaxed(status + idx) : lb_word(epoch, LB_INCL, 0);
            const uint32_t flag = ((uint32_t)(w >> 38) == epoch) ? (uint32_t)(w >> 36) & 3u : 0u;
            incl_m = __ballot(flag == LB_INCL);
            const uint64_t empty_m = __ballot(flag == LB_EMPTY);
            need = incl_m ? (((incl_m & (~incl_m + 1ull)) << 1) - 1ull) : ~0ull;
            if (!(empty_m & need)) break;
            // Bounded wait: after kSpinLimit polls, or once any tile of the
            // launch has given up, stop waiting (output is then flagged bad).
            ++spins;
            if ((spins & 255u) == 0) {
                const uint32_t e = __hip_atomic_load(&ctrl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((e & ERRF_TIMEOUT) || spins > kSpinLimit) {
                    if (lane == 0) __hip_atomic_fetch_or(&ctrl->error, ERRF_TIMEOUT, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
                    return 0;
                }
            }
            __builtin_amdgcn_s_sleep(1);
        }
        excl += wave_sum_u64(((need >> lane) & 1ull) ? (w & kValMask) : 0ull);
        if (incl_m) return excl;
        base -= 64;
    }
}

// Locate the item of a global tile (binary search over first_tile).
__device__ __forceinline__ uint32_t find_item(const Item *items, uint32_t n, uint32_t tile)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (items[mid].first_tile <= tile) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Block gather / scatter with zero padding (cuLSZ_kernel.cu:877-896, 1629-1640).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_block(const Item &it, uint32_t bx, uint32_t by, uint32_t bz,
                                           uint32_t D[32])
{
    const uint32_t y0 = by * 8, z0 = bz * 8;
    const uint16_t *p = it.vol_in + ((uint64_t)bx * it.y + y0) * it.z + z0;
    const bool full = (y0 + 8 <= it.y) && (z0 + 8 <= it.z);
    if (full && it.align == ALIGN16) {
        uint4 t[8];
#pragma unroll
        for (int r = 0; r < 8; r++) t[r] = *(const uint4 *)(p + (uint64_t)r * it.z);
#pragma unroll
        for (int r = 0; r < 8; r++) {
            D[4 * r] = t[r].x; D[4 * r + 1] = t[r].y; D[4 * r + 2] = t[r].z; D[4 * r + 3] = t[r].w;
        }
    } else if (full && it.align == ALIGN4) {
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const uint32_t *q = (const uint32_t *)(p + (uint64_t)r * it.z);
#pragma unroll
            for (int c = 0; c < 4; c++) D[4 * r + c] = q[c];
        }
    } else {
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const bool rv = y0 + r < it.y;
            const uint16_t *q = p + (uint64_t)r * it.z;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const uint32_t lo = (rv && z0 + 2 * c < it.z) ? q[2 * c] : 0u;
                const uint32_t hi = (rv && z0 + 2 * c + 1 < it.z) ? q[2 * c + 1] : 0u;
                D[4 * r + c] = lo | (hi << 16);
            }
        }
    }
}

__device__ __forceinline__ void store_block(const Item &it, uint32_t bx, uint32_t by, uint32_t bz,
                                            const uint32_t D[32])
{
    const uint32_t y0 = by * 8, z0 = bz * 8;
    uint16_t *p = it.vol_out + ((uint64_t)bx * it.y + y0) * it.z + z0;
    const bool full = (y0 + 8 <= it.y) && (z0 + 8 <= it.z);
    if (full && it.align == ALIGN16) {
#pragma unroll
        for (int r = 0; r < 8; r++)
            *(uint4 *)(p + (uint64_t)r * it.z) = make_uint4(D[4 * r], D[4 * r + 1], D[4 * r + 2], D[4 * r + 3]);
    } else if (full && it.align == ALIGN4) {
#pragma unroll
        for (int r = 0; r < 8; r++) {
            uint32_t *q = (uint32_t *)(p + (uint64_t)r * it.z);
#pragma unroll
            for (int c = 0; c < 4; c++) q[c] = D[4 * r + c];
        }
    } else {
#pragma unroll
        for (int r = 0; r < 8; r++) {
            if (y0 + r >= it.y) break;
            uint16_t *q = p + (uint64_t)r * it.z;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                if (z0 + 2 * c < it.z) q[2 * c] = (uint16_t)(D[4 * r + c] & 0xffffu);
                if (z0 + 2 * c + 1 < it.z) q[2 * c + 1] = (uint16_t)(D[4 * r + c] >> 16);
            }
        }
    }
}

// Workgroup exclusive scan of per-lane word counts; returns the lane's offset,
// sets *agg to the tile total.  Contains barriers.
__device__ __forceinline__ uint32_t tile_scan(uint32_t words, uint32_t tid, uint32_t *s_wsum,
                                              uint32_t *agg)
{
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    const uint32_t incl = wave_incl_scan(words, lane);
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
        const uint32_t s = s_wsum[w];
        before += (w < (int)wave) ? s : 0u;
        total += s;
    }
    *agg = total;
    return before + incl - words;
}

struct TileShared {
    uint32_t pay[kTile * 32];      // payload staging, 4-byte words (128 B per block max)
    uint32_t hdr[kTile / 4];       // header bytes of the tile
    uint32_t wsum[kWaves];
    uint32_t tile, epoch, item;
    uint32_t excl_lo, excl_hi;
};

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
            __hip_atomic_store(&ctrl->tile_ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctrl->done_ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctrl->epoch, (epoch + 1u) & kEpochMask, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__device__ __forceinline__ void next_tile(TileShared &sh, Ctrl *ctrl)
{
    if (threadIdx.x == 0)
        sh.tile = __hip_atomic_fetch_add(&ctrl->tile_ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <bool kTable>
__device__ __forceinline__ const Item &tile_item(TileShared &sh, const Item &single, const Item *items,
                                                 uint32_t nitems, uint32_t tile)
{
    if (!kTable) return single;
    if (threadIdx.x == 0) sh.item = find_item(items, nitems, tile);
    __syncthreads();
    return items[__builtin_amdgcn_readfirstlane(sh.item)];
}

// ---------------------------------------------------------------------------
// Compression.
// ---------------------------------------------------------------------------
template <bool kTable>
__global__ __launch_bounds__(kTile) void compress_kernel(const Item single, const Item *__restrict__ items,
                                                        uint32_t nitems, uint32_t total_tiles,
                                                        Ctrl *ctrl, uint64_t *status)
{
    __shared__ TileShared sh;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid == 0) sh.epoch = __hip_atomic_load(&ctrl->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    next_tile(sh, ctrl);
    __syncthreads();
    const uint32_t epoch = sh.epoch;

    for (;;) {
        const uint32_t tile = sh.tile;
        if (tile >= total_tiles) break;
        const Item &it = tile_item<kTable>(sh, single, items, nitems, tile);
        const uint32_t tl = tile - it.first_tile;
        const uint32_t b = tl * (uint32_t)kTile + tid;
        const bool valid = b < it.nblocks;

        uint32_t D[32];
        uint32_t hdr = 0, words = 0;
        if (valid) {
            uint32_t bx, by, bz;
            block_coords(it, b, bx, by, bz);
            load_block(it, bx, by, bz, D);
            hdr = block_header(D, it.dec);
            words = header_payload_bytes(hdr) >> 2;
        }
        ((uint8_t *)sh.hdr)[tid] = (uint8_t)hdr;

        uint32_t agg;
        const uint32_t off = tile_scan(words, tid, sh.wsum, &agg);
        if (tid == 0) st_relaxed(status + tile, lb_word(epoch, tl == 0 ? LB_INCL : LB_AGG, agg));

        if (words) encode_block(D, hdr, it.dec, &sh.pay[off]);

        if (wave == 0) {
            uint64_t excl = 0;
            if (tl != 0) {
                excl = lookback(status, tile, it.first_tile, epoch, lane, ctrl);
                if (lane == 0) st_relaxed(status + tile, lb_word(epoch, LB_INCL, excl + agg));
            }
            if (lane == 0) { sh.excl_lo = (uint32_t)excl; sh.excl_hi = (uint32_t)(excl >> 32); }
        }
        __syncthreads();

        const uint64_t excl = ((uint64_t)sh.excl_hi << 32) | sh.excl_lo;
        const uint64_t g0 = (uint64_t)it.hdr_bytes + 4ull * excl;   // byte offset of the tile payload
        // header bytes [tl*kTile, tl*kTile + kTile) clipped to R (pad bytes are 0)
        for (uint32_t w = tid; w < (uint32_t)kTile / 4; w += kTile) {
            const uint64_t o = (uint64_t)tl * kTile + 4ull * w;
            if (o < it.hdr_bytes) *(uint32_t *)(it.s_out + o) = sh.hdr[w];
        }
        if (g0 + 4ull * agg <= it.s_len) {
            uint32_t *dst = (uint32_t *)(it.s_out + g0);
            for (uint32_t w = tid; w < agg; w += kTile) dst[w] = sh.pay[w];
        }
        if (tid == 0 && tl == it.ntiles - 1 && it.result) *it.result = g0 + 4ull * agg;

        __syncthreads();
        next_tile(sh, ctrl);
        __syncthreads();
    }
    finish_launch(ctrl, epoch);
}

// ---------------------------------------------------------------------------
// Decompression.
// ---------------------------------------------------------------------------
template <bool kTable>
__global__ __launch_bounds__(kTile) void decompress_kernel(const Item single, const Item *__restrict__ items,
                                                          uint32_t nitems, uint32_t total_tiles,
                                                          Ctrl *ctrl, uint64_t *status)
{
    __shared__ TileShared sh;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid == 0) sh.epoch = __hip_atomic_load(&ctrl->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    next_tile(sh, ctrl);
    __syncthreads();
    const uint32_t epoch = sh.epoch;

    for (;;) {
        const uint32_t tile = sh.tile;
        if (tile >= total_tiles) break;
        const Item &it = tile_item<kTable>(sh, single, items, nitems, tile);
        const uint32_t tl = tile - it.first_tile;
        const uint32_t b = tl * (uint32_t)kTile + tid;
        const bool valid = b < it.nblocks;

        const uint32_t hdr = valid ? (uint32_t)it.s_in[b] : 0u;
        const uint32_t words = header_payload_bytes(hdr) >> 2;
        bool bad = (hdr & 0x1fu) > 16u;

        uint32_t agg;
        const uint32_t off = tile_scan(words, tid, sh.wsum, &agg);
        if (tid == 0) st_relaxed(status + tile, lb_word(epoch, tl == 0 ? LB_INCL : LB_AGG, agg));
        if (wave == 0) {
            uint64_t excl = 0;
            if (tl != 0) {
                excl = lookback(status, tile, it.first_tile, epoch, lane, ctrl);
                if (lane == 0) st_relaxed(status + tile, lb_word(epoch, LB_INCL, excl + agg));
            }
            if (lane == 0) { sh.excl_lo = (uint32_t)excl; sh.excl_hi = (uint32_t)(excl >> 32); }
        }
        __syncthreads();

        const uint64_t excl = ((uint64_t)sh.excl_hi << 32) | sh.excl_lo;
        const uint64_t g0 = (uint64_t)it.hdr_bytes + 4ull * excl;
        const bool fits = agg <= (uint32_t)kTile * 32u && g0 + 4ull * agg <= it.s_len;
        if (fits) {
            const uint32_t *src = (const uint32_t *)(it.s_in + g0);
            for (uint32_t w = tid; w < agg; w += kTile) sh.pay[w] = src[w];
        }
        __syncthreads();

        if (valid) {
            uint32_t D[32];
            bad = bad || !fits;
            if (!bad && (hdr & 0x1fu)) {
                decode_block(&sh.pay[off], hdr, it.dec.b, D);
            } else {
#pragma unroll
                for (int i = 0; i < 32; i++) D[i] = 0;
            }
            uint32_t bx, by, bz;
            block_coords(it, b, bx, by, bz);
            store_block(it, bx, by, bz, D);
            if (bad) __hip_atomic_fetch_or(&ctrl->error, ERRF_CORRUPT, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        next_tile(sh, ctrl);
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

hipError_t launch_compress(const Item &single, const Item *items, uint32_t nitems, LaunchCfg cfg,
                           Ctrl *ctrl, uint64_t *status, hipStream_t stream)
{
    if (items)
        hipLaunchKernelGGL(compress_kernel<true>, dim3(cfg.grid), dim3(kTile), 0, stream, single, items,
                           nitems, cfg.total_tiles, ctrl, status);
    else
        hipLaunchKernelGGL(compress_kernel<false>, dim3(cfg.grid), dim3(kTile), 0, stream, single, items,
                           nitems, cfg.total_tiles, ctrl, status);
    return hipGetLastError();
}

hipError_t launch_decompress(const Item &single, const Item *items, uint32_t nitems, LaunchCfg cfg,
                             Ctrl *ctrl, uint64_t *status, hipStream_t stream)
{
    if (items)
        hipLaunchKernelGGL(decompress_kernel<true>, dim3(cfg.grid), dim3(kTile), 0, stream, single, items,
                           nitems, cfg.total_tiles, ctrl, status);
    else
        hipLaunchKernelGGL(decompress_kernel<false>, dim3(cfg.grid), dim3(kTile), 0, stream, single, items,
                           nitems, cfg.total_tiles, ctrl, status);
    return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t *p, uint64_t n, uint64_t value, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    uint64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((uint32_t)g), dim3(256), 0, stream, p, n, value);
    return hipGetLastError();
}

int max_resident_workgroups(bool compress)
{
    int dev = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    hipError_t e = compress
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, compress_kernel<false>, kTile, 0)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, decompress_kernel<false>, kTile, 0);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    return ncu * per_cu;
}

}  // namespace lsc
