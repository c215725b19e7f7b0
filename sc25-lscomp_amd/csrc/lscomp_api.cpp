// lscomp_api.cpp -- host side of the C ABI (include/lscomp.h).
//
// Replaces the reference's host entries cuLSZ_{compression,decompression}_
// uint16_bsize64 (lsCOMP/src/cuLSZ_entry.cu:93-166).  Differences in how, not
// what: validation happens in the library (the reference validates only in
// its CLI, lsCOMP_xpcs.cpp:51-85); the look-back scratch is a persistent
// per-(device, stream) area instead of three cudaMalloc/cudaMemset/cudaFree
// per call (cuLSZ_entry.cu:109-130); errors are returned, not exit()ed.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <map>
#include <mutex>
#include <utility>

#include "../../include/lscomp.h"
#include "lscomp_internal.h"
#include "lscomp_params.h"
#include "lscomp_host.h"

using namespace lsc;

namespace lsc {
// ---------------------------------------------------------------------------
// Parameters.
// ---------------------------------------------------------------------------
bool bins_ok(lscomp_bins4 b)
{
    if (b.x == 0 || b.y == 0 || b.z == 0 || b.w == 0) return false;
    return b.x <= b.y && b.y <= b.z && b.z <= b.w;
}

uint32_t align_class(const void *vol, uint32_t z)
{
    uintptr_t a = (uintptr_t)vol;
    if (a % 16 == 0 && z % 8 == 0) return ALIGN16;
    if (a % 4 == 0 && z % 2 == 0) return ALIGN4;
    return ALIGN2;
}

uint64_t num_blocks(lscomp_dim3 d)
{
    return (uint64_t)d.x * (((uint64_t)d.y + 7) / 8) * (((uint64_t)d.z + 7) / 8);
}

// Launch shape per direction (0 compress, 1 decompress), packed 16 bits per
// field (tile, workgroups per CU, window, counters) into one word.  Set by
// lscomp_set_launch_shape[_ex] / the autotuner.
namespace {
uint64_t pack_shape(const Shape &s)
{
    return (uint64_t)(s.tile & 0xffffu) | (uint64_t)(s.wgpcu & 0xffffu) << 16 | (uint64_t)(s.win & 0xffffu) << 32 |
           (uint64_t)(s.nc & 0xffffu) << 48;
}
Shape unpack_shape(uint64_t v)
{
    return Shape{(uint32_t)(v & 0xffffu), (uint32_t)(v >> 16 & 0xffffu), (uint32_t)(v >> 32 & 0xffffu),
                 (uint32_t)(v >> 48 & 0xffffu)};
}
Shape default_shape(int d)
{
    return Shape{kDefaultTile[d], 0u, default_window(kDefaultTile[d]), kDefaultCounters[d]};
}
std::atomic<uint64_t> g_shape[2] = {{pack_shape(default_shape(0))}, {pack_shape(default_shape(1))}};
}  // namespace

Shape load_shape(int direction) { return unpack_shape(g_shape[direction].load()); }
void store_shape(int direction, const Shape &s) { g_shape[direction].store(pack_shape(s)); }

int normalize_shape(int direction, Shape &s)
{
    if (s.tile == 0) s.tile = kDefaultTile[direction];
    if (s.tile != 128 && s.tile != 256 && s.tile != 512) return LSCOMP_ERR_INVALID_ARG;
    if (s.wgpcu > 64) return LSCOMP_ERR_INVALID_ARG;
    if (s.win == 0) s.win = default_window(s.tile);
    if (s.win % 64 != 0 || s.win > s.tile) return LSCOMP_ERR_INVALID_ARG;
    if (s.nc == 0) s.nc = kDefaultCounters[direction];
    if (s.nc > kMaxCounters) return LSCOMP_ERR_INVALID_ARG;
    return LSCOMP_OK;
}
// How a single-volume decompression finds its tile offsets: 1 = a first phase
// of the decode launch (default), 2 = a separate pre-pass kernel, 0 = a
// decoupled look-back inside the decode (LSCOMP_DECODE_PREPASS in the
// environment sets the initial mode).
std::atomic<int> g_decode_mode{[] {
    const char *e = std::getenv("LSCOMP_DECODE_PREPASS");
    return (e && (e[0] == '0' || e[0] == '2')) ? e[0] - '0' : 1;
}()};

int hip_fail(hipError_t e, const char *what)
{
    if (e == hipSuccess) return LSCOMP_OK;
    std::fprintf(stderr, "lscomp: %s failed: %s\n", what, hipGetErrorString(e));
    return LSCOMP_ERR_HIP;
}

// Fills the geometry part of an Item; returns a status.
int make_item(lscomp_dim3 dims, lscomp_bins4 bins, float p, Item &it, uint32_t tile)
{
    if (dims.x == 0 || dims.y == 0 || dims.z == 0) return LSCOMP_ERR_INVALID_ARG;
    if (!bins_ok(bins)) return LSCOMP_ERR_INVALID_ARG;
    if (!(p >= 0.0f && p <= 1.0f)) return LSCOMP_ERR_INVALID_ARG;
    const uint64_t nb = num_blocks(dims);
    if (nb > kMaxBlocksPerItem) return LSCOMP_ERR_TOO_LARGE;
    std::memset(&it, 0, sizeof it);
    it.elem = 2;
    it.x = dims.x;
    it.y = dims.y;
    it.z = dims.z;
    // in 64 bits: y or z near 2^32 must not wrap (nb <= 2^31 - 1 bounds both results)
    it.nbz = (uint32_t)(((uint64_t)dims.z + 7) / 8);
    it.per_frame = (uint32_t)((((uint64_t)dims.y + 7) / 8) * it.nbz);
    it.nblocks = (uint32_t)nb;
    it.hdr_bytes = (uint32_t)((nb + 3) / 4 * 4);
    it.first_tile = 0;
    it.ntiles = (uint32_t)((nb + tile - 1) / tile);
    it.div_frame = make_fastdiv(it.per_frame);
    it.div_nbz = make_fastdiv(it.nbz);
    const uint32_t b4[4] = {bins.x, bins.y, bins.z, bins.w};
    it.dec = make_decide(b4, p);
    return LSCOMP_OK;
}

LaunchCfg make_cfg(uint64_t total_tiles, bool compress, bool table, const Shape &shape)
{
    static std::atomic<int> occ[2][2][4];   // [compress][table][tile 128/256/512/1024]
    static std::atomic<int> ncu{0};
    const uint32_t tile = shape.tile;
    const int ti = tile == 128 ? 0 : tile == 256 ? 1 : tile == 512 ? 2 : 3;
    int o = occ[compress][table][ti].load();
    if (o == 0) {
        o = workgroups_per_cu(compress, table, (int)tile);
        occ[compress][table][ti].store(o);
    }
    int n = ncu.load();
    if (n == 0) {
        n = num_cus();
        ncu.store(n);
    }
    uint32_t per_cu = (uint32_t)o;
    if (shape.wgpcu && shape.wgpcu < per_cu) per_cu = shape.wgpcu;
    LaunchCfg c{};
    c.total_tiles = (uint32_t)total_tiles;
    c.tile = tile;
    uint64_t g = (uint64_t)(n > 0 ? n : 1) * per_cu;
    if (g > total_tiles) g = total_tiles;
    c.grid = (uint32_t)(g ? g : 1);
    // every class of tile ids needs a workgroup: at most one counter per workgroup
    c.sched.nc = shape.nc < c.grid ? shape.nc : c.grid;
    if (c.sched.nc == 0) c.sched.nc = 1;
    c.sched.win = shape.win ? shape.win : default_window(tile);
    return c;
}

}  // namespace lsc

namespace {

// ---------------------------------------------------------------------------
// Scratch areas: one per (device, stream).
// ---------------------------------------------------------------------------
struct Scratch {
    int device = -1;
    hipStream_t stream = nullptr;
    Ctrl *ctrl = nullptr;          // device, zero-initialised
    uint64_t *status = nullptr;    // device, cap_tiles look-back words
    uint64_t cap_tiles = 0;
    uint64_t *tile_off = nullptr;  // device, cap_tiles + 1 decode tile offsets
    uint64_t launches = 0;         // epochs used since the status words were cleared
    Ctrl *host = nullptr;          // pinned readback of ctrl
    std::mutex mu;
};

std::mutex g_mu;
std::map<std::pair<int, hipStream_t>, Scratch *> g_scratch;

Scratch *get_scratch(hipStream_t stream)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    auto key = std::make_pair(dev, stream);
    auto itr = g_scratch.find(key);
    if (itr != g_scratch.end()) return itr->second;
    Scratch *s = new Scratch;
    s->device = dev;
    s->stream = stream;
    if (hipMalloc((void **)&s->ctrl, sizeof(Ctrl)) != hipSuccess ||
        hipHostMalloc((void **)&s->host, sizeof(Ctrl), hipHostMallocDefault) != hipSuccess) {
        if (s->ctrl) (void)hipFree(s->ctrl);
        delete s;
        return nullptr;
    }
    if (launch_fill_u64((uint64_t *)s->ctrl, sizeof(Ctrl) / 8, 0, stream) != hipSuccess) {
        (void)hipFree(s->ctrl);
        (void)hipHostFree(s->host);
        delete s;
        return nullptr;
    }
    g_scratch[key] = s;
    return s;
}

// Caller holds s->mu.  Makes room for `tiles` look-back words.
int ensure_tiles(Scratch *s, uint64_t tiles)
{
    if (tiles <= s->cap_tiles) return LSCOMP_OK;
    uint64_t cap = s->cap_tiles ? s->cap_tiles : 1024;
    while (cap < tiles) cap *= 2;
    if (s->status) {
        int st = hip_fail(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        if (st) return st;
        (void)hipFree(s->status);
        (void)hipFree(s->tile_off);
        s->status = nullptr;
        s->tile_off = nullptr;
        s->cap_tiles = 0;
    }
    int st = hip_fail(hipMalloc((void **)&s->status, cap * sizeof(uint64_t)), "hipMalloc");
    if (!st) st = hip_fail(hipMalloc((void **)&s->tile_off, (cap + 1) * sizeof(uint64_t)), "hipMalloc");
    if (st) return st;
    st = hip_fail(launch_fill_u64(s->status, cap, 0, s->stream), "fill");
    if (!st) st = hip_fail(launch_fill_u64(s->tile_off, cap + 1, 0, s->stream), "fill");
    if (st) return st;
    s->cap_tiles = cap;
    s->launches = 0;
    return LSCOMP_OK;
}

// Caller holds s->mu.  Called once per launch that uses the status words.
int count_launch(Scratch *s)
{
    if (++s->launches >= kEpochClearPeriod) {
        int st = hip_fail(launch_fill_u64(s->status, s->cap_tiles, 0, s->stream), "fill");
        if (!st) st = hip_fail(launch_fill_u64(s->tile_off, s->cap_tiles + 1, 0, s->stream), "fill");
        if (st) return st;
        s->launches = 1;
    }
    return LSCOMP_OK;
}

int read_ctrl(Scratch *s)
{
    int st = hip_fail(hipMemcpyAsync(s->host, s->ctrl, kCtrlHead, hipMemcpyDeviceToHost, s->stream),
                      "hipMemcpyAsync");
    if (st) return st;
    return hip_fail(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
}

// Returns the status for sticky error bits and clears them.
int take_errors(Scratch *s)
{
    const uint32_t e = s->host->error;
    if (!e) return LSCOMP_OK;
    int st = hip_fail(launch_fill_u64((uint64_t *)&s->ctrl->error, 1, 0, s->stream), "fill");
    if (st) return st;
    st = hip_fail(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    if (st) return st;
    if (e & ERRF_TIMEOUT) return LSCOMP_ERR_INTERNAL;
    if (e & ERRF_RANGE) return LSCOMP_ERR_RANGE;
    return LSCOMP_ERR_CORRUPT;
}

}  // namespace

namespace lsc {

// Validates a compression call and fills its Item (no HIP calls).
int prepare_compress(const uint16_t *d_in, uint8_t *d_out, size_t cap, lscomp_dim3 dims, lscomp_bins4 bins,
                     float p, uint32_t tile, Item &it)
{
    int st = make_item(dims, bins, p, it, tile);
    if (st) return st;
    if (!d_in || !d_out || ((uintptr_t)d_in & 1) || ((uintptr_t)d_out & 3)) return LSCOMP_ERR_INVALID_ARG;
    if (cap < it.hdr_bytes) return LSCOMP_ERR_CAPACITY;
    it.vol_in = d_in;
    it.s_out = d_out;
    it.s_len = cap;
    it.align = align_class(d_in, dims.z);
    return LSCOMP_OK;
}

int prepare_decompress(uint16_t *d_out, const uint8_t *d_in, size_t in_bytes, lscomp_dim3 dims,
                       lscomp_bins4 bins, uint32_t tile, Item &it)
{
    int st = make_item(dims, bins, 0.0f, it, tile);
    if (st) return st;
    if (!d_in || !d_out || ((uintptr_t)d_out & 1) || ((uintptr_t)d_in & 3)) return LSCOMP_ERR_INVALID_ARG;
    if (in_bytes < it.hdr_bytes) return LSCOMP_ERR_CORRUPT;
    it.vol_out = d_out;
    it.s_in = d_in;
    it.s_len = in_bytes;
    it.align = align_class(d_out, dims.z);
    return LSCOMP_OK;
}

}  // namespace lsc

namespace {

// uint32 volumes: validation and item (tiles of kTileU32 blocks; no HIP calls).
int prepare_compress_u32(const uint32_t *d_in, uint8_t *d_out, size_t cap, lscomp_dim3 dims, lscomp_bins4 bins,
                         float p, Item &it)
{
    int st = make_item(dims, bins, p, it, kTileU32);
    if (st) return st;
    if (!d_in || !d_out || ((uintptr_t)d_in & 3) || ((uintptr_t)d_out & 3)) return LSCOMP_ERR_INVALID_ARG;
    if (cap < it.hdr_bytes) return LSCOMP_ERR_CAPACITY;
    it.vol_in = (const uint16_t *)d_in;
    it.s_out = d_out;
    it.s_len = cap;
    it.elem = 4;
    return LSCOMP_OK;
}

int prepare_decompress_u32(uint32_t *d_out, const uint8_t *d_in, size_t in_bytes, lscomp_dim3 dims,
                           lscomp_bins4 bins, Item &it)
{
    int st = make_item(dims, bins, 0.0f, it, kTileU32Dec);
    if (st) return st;
    if (!d_in || !d_out || ((uintptr_t)d_out & 3) || ((uintptr_t)d_in & 3)) return LSCOMP_ERR_INVALID_ARG;
    if (in_bytes < it.hdr_bytes) return LSCOMP_ERR_CORRUPT;
    it.vol_out = (uint16_t *)d_out;
    it.s_in = d_in;
    it.s_len = in_bytes;
    it.elem = 4;
    return LSCOMP_OK;
}

// Caller holds s->mu.  A uint32 item; decompression always uses the offsets pre-pass.
int launch_item_u32(Scratch *s, Item &it, bool compress)
{
    int st = ensure_tiles(s, it.ntiles);
    if (st) return st;
    st = count_launch(s);
    if (st) return st;
    static std::atomic<int> occ[2];
    int o = occ[compress].load();
    if (o == 0) {
        o = workgroups_per_cu_u32(compress);
        occ[compress].store(o);
    }
    LaunchCfg cfg{};
    cfg.tile = tile_u32(compress);
    cfg.total_tiles = it.ntiles;
    uint64_t g = (uint64_t)(num_cus() > 0 ? num_cus() : 1) * (uint64_t)o;
    if (g > it.ntiles) g = it.ntiles;
    cfg.grid = (uint32_t)(g ? g : 1);
    // the direction's counters; the window of 64 words timed best at this tile
    const uint32_t nc = load_shape(compress ? 0 : 1).nc;
    cfg.sched.nc = nc < cfg.grid ? (nc ? nc : 1u) : cfg.grid;
    cfg.sched.win = default_window(cfg.tile);
    if (compress) {
        if (!it.result) it.result = &s->ctrl->result;
    } else {
        cfg.tile_off = s->tile_off;
        cfg.off_mode = g_decode_mode.load() == 2 ? OFF_KERNEL : OFF_PHASE;
        if (cfg.off_mode == OFF_KERNEL) {
            st = count_launch(s);   // the pre-pass is a launch of its own
            if (st) return st;
        }
    }
    return hip_fail(launch_u32(compress, it, cfg, s->ctrl, s->status, s->stream),
                    compress ? "compress launch" : "decompress launch");
}

// Caller holds s->mu.  it.ntiles was computed for shape.tile (the same shape
// the call loaded once).
int launch_item(Scratch *s, Item &it, bool compress, const Shape &shape)
{
    int st = ensure_tiles(s, it.ntiles);
    if (st) return st;
    st = count_launch(s);
    if (st) return st;
    const LaunchCfg cfg = make_cfg(it.ntiles, compress, false, shape);
    if (compress) {
        if (!it.result) it.result = &s->ctrl->result;
        return hip_fail(launch_compress(it, nullptr, nullptr, cfg, s->ctrl, s->status, s->stream), "compress launch");
    }
    LaunchCfg dcfg = cfg;
    const int mode = g_decode_mode.load();
    if (mode == 1 && shape.tile == 512 && it.ntiles > kSmallTiles && dense_stream(it)) {
        // the pipelined decoder on tiles of kPipeTile blocks, one 16-wave workgroup
        // per CU: CSSI-shaped lossless decode 0.766 -> 0.686 ms against tiles of 512
        it.ntiles = (uint32_t)((it.nblocks + kPipeTile - 1) / kPipeTile);
        Shape big = shape;
        big.tile = kPipeTile;
        dcfg = make_cfg(it.ntiles, false, false, big);
    }
    if (mode != 0) {
        dcfg.tile_off = s->tile_off;
        dcfg.off_mode = mode == 2 ? OFF_KERNEL : OFF_PHASE;
        if (mode == 2) {
            st = count_launch(s);   // the pre-pass is a launch of its own
            if (st) return st;
        }
    }
    return hip_fail(launch_decompress(it, nullptr, nullptr, dcfg, s->ctrl, s->status, s->stream), "decompress launch");
}

}  // namespace

extern "C" {

size_t lscomp_num_blocks_u16(lscomp_dim3 dims) { return (size_t)num_blocks(dims); }

size_t lscomp_header_bytes_u16(lscomp_dim3 dims) { return (size_t)((num_blocks(dims) + 3) / 4 * 4); }

size_t lscomp_max_compressed_bytes_u16(lscomp_dim3 dims)
{
    const uint64_t nb = num_blocks(dims);
    return (size_t)((nb + 3) / 4 * 4 + 128 * nb);
}

int lscomp_compress_u16_async(const uint16_t *d_in, uint8_t *d_out, size_t out_capacity, uint64_t *d_out_bytes,
                              lscomp_dim3 dims, lscomp_bins4 bins, float pooling_th, hipStream_t stream)
{
    if (!d_out_bytes || ((uintptr_t)d_out_bytes & 7)) return LSCOMP_ERR_INVALID_ARG;
    const Shape shape = load_shape(0);
    Item it;
    int st = prepare_compress(d_in, d_out, out_capacity, dims, bins, pooling_th, shape.tile, it);
    if (st) return st;
    it.result = d_out_bytes;
    Scratch *s = get_scratch(stream);
    if (!s) return LSCOMP_ERR_HIP;
    std::lock_guard<std::mutex> lk(s->mu);
    return launch_item(s, it, true, shape);
}

int lscomp_compress_u16(const uint16_t *d_in, uint8_t *d_out, size_t out_capacity, size_t *out_bytes,
                        lscomp_dim3 dims, lscomp_bins4 bins, float pooling_th, hipStream_t stream)
{
    if (!out_bytes) return LSCOMP_ERR_INVALID_ARG;
    const Shape shape = load_shape(0);
    Item it;
    int st = prepare_compress(d_in, d_out, out_capacity, dims, bins, pooling_th, shape.tile, it);
    if (st == LSCOMP_ERR_CAPACITY) *out_bytes = lscomp_header_bytes_u16(dims);
    if (st) return st;
    Scratch *s = get_scratch(stream);
    if (!s) return LSCOMP_ERR_HIP;
    std::lock_guard<std::mutex> lk(s->mu);
    st = launch_item(s, it, true, shape);
    if (st) return st;
    st = read_ctrl(s);
    if (st) return st;
    *out_bytes = (size_t)s->host->result;
    st = take_errors(s);
    if (st) return st;
    return *out_bytes > out_capacity ? LSCOMP_ERR_CAPACITY : LSCOMP_OK;
}

int lscomp_decompress_u16_async(uint16_t *d_out, const uint8_t *d_in, size_t in_bytes, lscomp_dim3 dims,
                                lscomp_bins4 bins, hipStream_t stream)
{
    const Shape shape = load_shape(1);
    Item it;
    int st = prepare_decompress(d_out, d_in, in_bytes, dims, bins, shape.tile, it);
    if (st) return st;
    Scratch *s = get_scratch(stream);
    if (!s) return LSCOMP_ERR_HIP;
    std::lock_guard<std::mutex> lk(s->mu);
    return launch_item(s, it, false, shape);
}

int lscomp_decompress_u16(uint16_t *d_out, const uint8_t *d_in, size_t in_bytes, lscomp_dim3 dims,
                          lscomp_bins4 bins, hipStream_t stream)
{
    const Shape shape = load_shape(1);
    Item it;
    int st = prepare_decompress(d_out, d_in, in_bytes, dims, bins, shape.tile, it);
    if (st) return st;
    Scratch *s = get_scratch(stream);
    if (!s) return LSCOMP_ERR_HIP;
    std::lock_guard<std::mutex> lk(s->mu);
    st = launch_item(s, it, false, shape);
    if (st) return st;
    st = read_ctrl(s);
    if (st) return st;
    return take_errors(s);
}

size_t lscomp_max_compressed_bytes_u32(lscomp_dim3 dims)
{
    const uint64_t nb = num_blocks(dims);
    return (size_t)((nb + 3) / 4 * 4 + 248 * nb);
}

int lscomp_compress_u32_async(const uint32_t *d_in, uint8_t *d_out, size_t out_capacity, uint64_t *d_out_bytes,
                              lscomp_dim3 dims, lscomp_bins4 bins, float pooling_th, hipStream_t stream)
{
    if (!d_out_bytes || ((uintptr_t)d_out_bytes & 7)) return LSCOMP_ERR_INVALID_ARG;
    Item it;
    int st = prepare_compress_u32(d_in, d_out, out_capacity, dims, bins, pooling_th, it);
    if (st) return st;
    it.result = d_out_bytes;
    Scratch *s = get_scratch(stream);
    if (!s) return LSCOMP_ERR_HIP;
    std::lock_guard<std::mutex> lk(s->mu);
    return launch_item_u32(s, it, true);
}

int lscomp_compress_u32(const uint32_t *d_in, uint8_t *d_out, size_t out_capacity, size_t *out_bytes,
                        lscomp_dim3 dims, lscomp_bins4 bins, float pooling_th, hipStream_t stream)
{
    if (!out_bytes) return LSCOMP_ERR_INVALID_ARG;
    Item it;
    int st = prepare_compress_u32(d_in, d_out, out_capacity, dims, bins, pooling_th, it);
    if (st == LSCOMP_ERR_CAPACITY) *out_bytes = lscomp_header_bytes_u16(dims);
    if (st) return st;
    Scratch *s = get_scratch(stream);
    if (!s) return LSCOMP_ERR_HIP;
    std::lock_guard<std::mutex> lk(s->mu);
    st = launch_item_u32(s, it, true);
    if (st) return st;
    st = read_ctrl(s);
    if (st) return st;
    *out_bytes = (size_t)s->host->result;
    st = take_errors(s);
    if (st) return st;
    return *out_bytes > out_capacity ? LSCOMP_ERR_CAPACITY : LSCOMP_OK;
}

int lscomp_decompress_u32_async(uint32_t *d_out, const uint8_t *d_in, size_t in_bytes, lscomp_dim3 dims,
                                lscomp_bins4 bins, hipStream_t stream)
{
    Item it;
    int st = prepare_decompress_u32(d_out, d_in, in_bytes, dims, bins, it);
    if (st) return st;
    Scratch *s = get_scratch(stream);
    if (!s) return LSCOMP_ERR_HIP;
    std::lock_guard<std::mutex> lk(s->mu);
    return launch_item_u32(s, it, false);
}

int lscomp_decompress_u32(uint32_t *d_out, const uint8_t *d_in, size_t in_bytes, lscomp_dim3 dims,
                          lscomp_bins4 bins, hipStream_t stream)
{
    Item it;
    int st = prepare_decompress_u32(d_out, d_in, in_bytes, dims, bins, it);
    if (st) return st;
    Scratch *s = get_scratch(stream);
    if (!s) return LSCOMP_ERR_HIP;
    std::lock_guard<std::mutex> lk(s->mu);
    st = launch_item_u32(s, it, false);
    if (st) return st;
    st = read_ctrl(s);
    if (st) return st;
    return take_errors(s);
}

int lscomp_stream_status(hipStream_t stream)
{
    Scratch *s = get_scratch(stream);
    if (!s) return LSCOMP_ERR_HIP;
    std::lock_guard<std::mutex> lk(s->mu);
    int st = read_ctrl(s);
    if (st) return st;
    return take_errors(s);
}

int lscomp_reserve_u16(lscomp_dim3 dims, hipStream_t stream)
{
    Item it;
    int st = make_item(dims, lscomp_bins4{1, 1, 1, 1}, 1.0f, it, 128);   // smallest tile: most words
    if (st) return st;
    Scratch *s = get_scratch(stream);
    if (!s) return LSCOMP_ERR_HIP;
    std::lock_guard<std::mutex> lk(s->mu);
    return ensure_tiles(s, it.ntiles);
}

void lscomp_release_all(void)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto itr = g_scratch.begin(); itr != g_scratch.end();) {
        Scratch *s = itr->second;
        if (s->device != dev) { ++itr; continue; }
        (void)hipStreamSynchronize(s->stream);
        if (s->status) (void)hipFree(s->status);
        if (s->tile_off) (void)hipFree(s->tile_off);
        if (s->ctrl) (void)hipFree(s->ctrl);
        if (s->host) (void)hipHostFree(s->host);
        delete s;
        itr = g_scratch.erase(itr);
    }
}

int lscomp_set_launch_shape(int direction, uint32_t tile, uint32_t wg_per_cu)
{
    if (direction != 0 && direction != 1) return LSCOMP_ERR_INVALID_ARG;
    Shape s{tile, wg_per_cu, 0u, 0u};
    const int st = normalize_shape(direction, s);
    if (st) return st;
    store_shape(direction, s);
    return LSCOMP_OK;
}

int lscomp_set_launch_shape_ex(int direction, const lscomp_launch_shape *shape)
{
    if ((direction != 0 && direction != 1) || !shape) return LSCOMP_ERR_INVALID_ARG;
    Shape s{shape->tile, shape->wg_per_cu, shape->lb_window, shape->counters};
    const int st = normalize_shape(direction, s);
    if (st) return st;
    store_shape(direction, s);
    return LSCOMP_OK;
}

int lscomp_get_launch_shape_ex(int direction, lscomp_launch_shape *shape)
{
    if ((direction != 0 && direction != 1) || !shape) return LSCOMP_ERR_INVALID_ARG;
    const Shape s = load_shape(direction);
    shape->tile = s.tile;
    shape->wg_per_cu = s.wgpcu;
    shape->lb_window = s.win;
    shape->counters = s.nc;
    return LSCOMP_OK;
}

int lscomp_set_decode_prepass(int mode)
{
    if (mode < 0 || mode > 2) return -LSCOMP_ERR_INVALID_ARG;
    return g_decode_mode.exchange(mode);
}

int lscomp_get_launch_shape(int direction, uint32_t *tile, uint32_t *wg_per_cu)
{
    if ((direction != 0 && direction != 1) || !tile || !wg_per_cu) return LSCOMP_ERR_INVALID_ARG;
    const Shape s = load_shape(direction);
    *tile = s.tile;
    *wg_per_cu = s.wgpcu;
    return LSCOMP_OK;
}

const char *lscomp_status_string(int status)
{
    switch (status) {
    case LSCOMP_OK: return "ok";
    case LSCOMP_ERR_INVALID_ARG: return "invalid argument";
    case LSCOMP_ERR_CAPACITY: return "output buffer too small";
    case LSCOMP_ERR_CORRUPT: return "corrupt or truncated stream";
    case LSCOMP_ERR_HIP: return "HIP runtime error";
    case LSCOMP_ERR_TOO_LARGE: return "volume has more than 2^31-1 blocks";
    case LSCOMP_ERR_INTERNAL: return "internal error (look-back timeout)";
    case LSCOMP_ERR_RANGE: return "uint32 value range needs rate 32, which the format cannot hold";
    default: return "unknown status";
    }
}

int lscomp_version(void) { return 100; }

}  // extern "C"
