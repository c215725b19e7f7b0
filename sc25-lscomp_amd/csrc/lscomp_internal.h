// lscomp_internal.h -- types shared by the gfx950 kernels and the host API.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "lscomp_block.h"

namespace lsc {

// Blocks per tile == threads per workgroup: 128, 256 or 512 (a launch-time
// choice; the stream does not depend on it).  Default per direction, from
// timings of a 600x1813x1558 volume: 512 for both (one tile-counter atomic per
// tile caps the tile rate near 80 per microsecond, so larger tiles pay off).
constexpr uint32_t kDefaultTile[2] = {512, 512};
constexpr uint32_t kMaxBlocksPerItem = 0x7fffffffu;   // prefix values fit 36-bit words

// Alignment class of a volume: how a row of one block (8 x u16) may be loaded.
enum : uint32_t { ALIGN2 = 0, ALIGN4 = 1, ALIGN16 = 2 };

// Tile ids of a launch come from `nc` counters: tile k*nc + c is the k-th id
// of counter c, and a workgroup draws from counter c = (its ticket, i.e. its
// place in the start order) mod nc.  One same-address device-scope atomic
// serializes at about 12.5 ns, which caps a single counter near 80 tiles per
// microsecond; counters on separate 256-byte lines spread that load.  Classes
// follow the start order, so the first nc running workgroups cover every
// class and a tile id is only ever handed to a running workgroup.
constexpr uint32_t kMaxCounters = 16;
struct CtrLine {
    uint32_t v;
    uint32_t pad[63];
};

// Launch-state block of a scratch area.  All fields start at zero; the last
// workgroup of every launch resets the counters and advances the epoch, so a
// launch needs no host-side clearing (and replays inside a hipGraph).
// Each field that many workgroups touch sits on its own line: a tile counter
// takes one atomic per tile and must not share its line with anything else
// that is read or written concurrently.  The host reads back only the first
// 16 bytes (error, result).
struct Ctrl {
    uint32_t error;      // sticky ERRF_* bits, cleared by the host
    uint32_t pad2;
    uint64_t result;     // stream bytes of a single-item compression
    uint64_t pad3[14];
    uint32_t done_ctr;   // workgroups finished
    uint32_t epoch;      // tags look-back words of the current launch (read once per workgroup)
    uint32_t pad1[30];
    uint32_t ticket;     // class of each workgroup (start order)
    uint32_t pad0[63];
    CtrLine chunk;       // decode: chunks of the tile-offset phase
    CtrLine ctr[kMaxCounters];   // tile ids, one counter per class
};
static_assert(sizeof(Ctrl) == 768 + 256 * kMaxCounters, "Ctrl layout");
constexpr size_t kCtrlHead = 16;   // bytes of Ctrl the host reads back (error, result)

// Scheduling parameters of a launch (uniform over the grid).
struct Sched {
    uint32_t nc;    // tile-id counters, 1 .. kMaxCounters
    uint32_t win;   // look-back window: predecessor words per probe round, a multiple of 64, <= tile
};
// ERRF_RANGE: a uint32 block's quantized maximum needs 32 bits (rate 32).
enum : uint32_t { ERRF_CORRUPT = 1u, ERRF_TIMEOUT = 2u, ERRF_RANGE = 4u };

// Look-back word: epoch (26 bits) | flag (2 bits) | value in 4-byte words (36 bits).
constexpr uint32_t kEpochMask = (1u << 26) - 1u;
constexpr uint64_t kValMask = (1ull << 36) - 1ull;
enum : uint32_t { LB_EMPTY = 0, LB_AGG = 1, LB_INCL = 2 };

// One independent stream: a volume of x frames of y rows of z pixels.
struct Item {
    const uint16_t *vol_in;   // compress: source volume
    uint16_t *vol_out;        // decompress: destination volume
    const uint8_t *s_in;      // decompress: source stream
    uint8_t *s_out;           // compress: destination stream
    uint64_t s_len;           // compress: capacity; decompress: stream bytes
    uint64_t *result;         // compress: required stream bytes (or null)
    uint32_t x, y, z, nbz;    // dims; nbz = ceil(z/8)
    uint32_t per_frame;       // ceil(y/8) * ceil(z/8)
    uint32_t nblocks;         // x * per_frame
    uint32_t hdr_bytes;       // R = 4 * ceil(nblocks/4)
    uint32_t first_tile;      // first global tile of this item
    uint32_t ntiles;          // ceil(nblocks / tile)
    uint32_t align;           // ALIGN*
    uint32_t elem;            // element bytes: 2 (uint16) or 4 (uint32; vol_in/vol_out hold uint32)
    uint32_t pad1;
    FastDiv div_frame, div_nbz;
    Decide dec;
    uint64_t seq;             // served form: the request's sequence number
};

// A single-volume decode's stream holds 2 or more payload bytes per block on
// average: the pipelined decoder overlaps its payload loads with decoding
// (below that the decoder that stages each tile's payload in its own
// iteration was faster).
inline __host__ __device__ bool dense_stream(const Item &it)
{
    return it.s_len >= (uint64_t)it.hdr_bytes + 2ull * it.nblocks;
}

// Launch shape of one direction: the stream never depends on it.
struct Shape {
    uint32_t tile;      // blocks per tile (= threads per workgroup): 128, 256 or 512
    uint32_t wgpcu;     // workgroups per CU (0: occupancy limit)
    uint32_t win;       // look-back window (0: default for the tile)
    uint32_t nc;        // tile-id counters (0: default)
};
// Default tile-id counters per direction (compression, decompression) and
// look-back window.  Timed on the 600x1813x1558 CSSI volume (tile 512, one
// MI355X, the same box for every variant): compression lossless 0.807 /
// 0.781 / 0.780 ms with 1 / 2 / 4 counters (lossy 0.829 / 0.825 / 0.827),
// decompression lossless 0.919 / 0.898 / 0.900 ms (lossy 0.750 / 0.742 /
// 0.744); a window of 64 words against 128: lossless compression 0.780
// against 0.790 ms, lossy 0.827 against 0.823 ms.
constexpr uint32_t kDefaultCounters[2] = {4, 2};
constexpr uint32_t default_window(uint32_t) { return 64u; }

struct LaunchCfg {
    uint32_t grid;            // workgroups (persistent: each loops over tiles)
    uint32_t total_tiles;
    uint32_t tile;            // blocks per tile: 128, 256 or 512
    Sched sched;              // counters and look-back window
    const uint32_t *d_total_tiles;   // if set, the tile count is read from device memory
    uint64_t *tile_off;              // decompress, single item: if set, tile offsets (ntiles + 1
                                     // epoch-tagged words) are computed first and the decode needs no look-back
    uint32_t off_mode;               // with tile_off: OFF_PHASE or OFF_KERNEL
};
// How a single-volume decode gets its tile offsets: a first phase of the decode
// launch (OFF_PHASE) or a separate pre-pass kernel (OFF_KERNEL).
enum : uint32_t { OFF_PHASE = 1, OFF_KERNEL = 2 };

// Kernel launchers (lscomp_kernels.hip).  `items` is a device array of item
// descriptors and `tile_items` (device, total_tiles entries) the item of each
// global tile; both null for the single item passed by value.
hipError_t launch_compress(const Item &single, const Item *items, const uint32_t *tile_items,
                           LaunchCfg cfg, Ctrl *ctrl, uint64_t *status, hipStream_t stream);
hipError_t launch_decompress(const Item &single, const Item *items, const uint32_t *tile_items,
                             LaunchCfg cfg, Ctrl *ctrl, uint64_t *status, hipStream_t stream);
hipError_t launch_fill_u64(uint64_t *p, uint64_t n, uint64_t value, hipStream_t stream);
// uint32 volumes: one item, tiles of kTileU32 blocks to compress and of
// kTileU32Dec to decompress (the stream does not depend on the tile; one
// 8-wave workgroup per CU decoding tiles of 512 was 7-11 % faster than two
// 4-wave ones on tiles of 256, and 6-11 % slower compressing); decompression
// always computes tile offsets first (cfg.tile_off must be set).
constexpr uint32_t kTileU32 = 256;
// Decode geometry of a single uint16 volume.  At most kSmallTiles tiles: one
// workgroup per tile that sums the header bytes before it.  A payload-dense
// stream (dense_stream) at the default and largest launch tile (512) with the
// offset phase decodes in the pipelined decoder on tiles of kPipeTile blocks,
// one 16-wave workgroup per CU.
constexpr uint32_t kSmallTiles = 256;
constexpr uint32_t kPipeTile = 1024;
constexpr uint32_t kTileU32Dec = 512;
constexpr uint32_t tile_u32(bool compress) { return compress ? kTileU32 : kTileU32Dec; }
hipError_t launch_u32(bool compress, const Item &it, LaunchCfg cfg, Ctrl *ctrl, uint64_t *status,
                      hipStream_t stream);
// ---------------------------------------------------------------------------
// Served form: a device-resident request loop (serve_kernel, lscomp_server.cpp).
// The host writes prepared items into a ring in host-pinned coherent memory
// and publishes a write index; one workgroup of the persistent kernel (the
// dispatcher) copies new items into a device item table, gives each its tile
// range and publishes the tiles through tagged tile-map words; every other
// workgroup compresses published tiles and the last tile of a request writes
// its completion record into host-pinned memory.  A kernel instance closes
// after an idle period, on the host's stop word or when its item table or
// tile map is full; the host relaunches it when requests wait.
// ---------------------------------------------------------------------------
constexpr uint32_t kSrvRing = 4096;            // ring slots (requests queued or running)
constexpr uint32_t kSrvMaxReq = 1u << 15;      // requests per kernel instance (device item table)
constexpr uint32_t kSrvMaxTiles = 1u << 22;    // tiles per kernel instance (tile map, look-back words)
constexpr uint32_t kSrvTile = 512;             // blocks per tile of the served kernel

struct alignas(128) ItemSlot {
    Item it;
};
// One ring slot, written by the host before it publishes `wr`.  The item is
// complete except first_tile and result, which the dispatcher sets.
struct alignas(128) SrvReq {
    Item item;
};
// Completion record of a ring slot, written by the device: bytes and status,
// then seq = the request's sequence number + 1 (system-scope release).
struct alignas(32) SrvComp {
    uint64_t seq;
    uint64_t bytes;
    uint32_t status;      // 0, or SRV_CAPACITY / SRV_TIMEOUT
    uint32_t pad[3];
};
enum : uint32_t { SRV_CAPACITY = 1u, SRV_TIMEOUT = 2u };
// Host-pinned control words, one line each.
struct SrvHost {
    uint64_t wr;          // host: requests published (sequence numbers below wr are in the ring)
    uint64_t pad0[15];
    uint64_t rd;          // device: requests the dispatchers have taken (all instances)
    uint64_t pad1[15];
    uint32_t stop;        // host: the running instance closes
    uint32_t pad2[31];
};
// Device state of one kernel instance, cleared by the host before each launch.
struct SrvDev {
    uint64_t close;       // (tile limit << 1) | 1: no tile at or past the limit will be published
    uint64_t pad0[15];
};
struct SrvArgs {
    const SrvReq *ring;   // host-pinned
    SrvComp *comp;        // host-pinned
    SrvHost *host;        // host-pinned
    SrvDev *dev;
    ItemSlot *items;      // kSrvMaxReq
    uint64_t *map;        // kSrvMaxTiles: epoch | item + 1 | tile within the item
    uint32_t *done;       // kSrvMaxReq: tiles stored per item
    uint64_t *sizes;      // kSrvMaxReq: stream bytes per item
    uint64_t idle_ticks;  // s_memrealtime ticks (100 MHz) without a new request before an instance closes
};
hipError_t launch_serve(const SrvArgs &a, uint32_t grid, Sched sc, Ctrl *ctrl, uint64_t *status, hipStream_t stream);
int workgroups_per_cu_serve();

int workgroups_per_cu(bool compress, bool table, int tile);   // occupancy limit
int workgroups_per_cu_u32(bool compress);
int num_cus();

}  // namespace lsc
