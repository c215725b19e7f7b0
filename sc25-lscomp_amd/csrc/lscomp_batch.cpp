// lscomp_batch.cpp -- batched form: many items, one persistent launch.
//
// The kernels' table mode reads an Item descriptor per item and the item of
// each global tile from device tables; the tile count lives in device memory
// too, so a hipGraph captured once replays any later batch of the plan.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/lscomp.h"
#include "lscomp_host.h"
#include "lscomp_internal.h"

using namespace lsc;

struct lscomp_plan_s {
    int device = 0;
    int direction = 0;            // 0 compress, 1 decompress
    Shape shape{};                // launch shape (lscomp_plan_set_shape; a captured graph keeps it)
    uint32_t max_items = 0;
    uint64_t max_blocks = 0;
    uint64_t max_tiles = 0;       // tile-table capacity, for the smallest tile (128 blocks)
    std::vector<lscomp_item_u16> src;   // the current batch as the caller gave it
    uint32_t nitems = 0, ntiles = 0;
    // device
    Item *d_items = nullptr;
    uint32_t *d_tile_items = nullptr;
    uint32_t *d_total = nullptr;  // tile count of the current batch
    uint64_t *d_sizes = nullptr;  // per-item results
    Ctrl *ctrl = nullptr;
    uint64_t *status = nullptr;
    uint64_t *tile_off = nullptr; // decompression: payload offset of every tile over the batch (offset phase)
    // host staging (pinned) for the table uploads, two buffers alternating
    uint8_t *h_stage[2] = {nullptr, nullptr};
    hipEvent_t stage_done[2] = {nullptr, nullptr};
    int stage_next = 0;
    size_t stage_bytes = 0;
    Ctrl *h_ctrl = nullptr;
    uint64_t launches = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::mutex mu;
};

namespace {

size_t table_bytes(uint32_t max_items, uint64_t max_tiles)
{
    return sizeof(Item) * max_items + sizeof(uint32_t) * (max_tiles + 1);
}

void free_plan(lscomp_plan p)
{
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (int i = 0; i < 2; i++) {
        if (p->stage_done[i]) (void)hipEventDestroy(p->stage_done[i]);
        if (p->h_stage[i]) (void)hipHostFree(p->h_stage[i]);
    }
    if (p->h_ctrl) (void)hipHostFree(p->h_ctrl);
    if (p->d_items) (void)hipFree(p->d_items);
    if (p->d_tile_items) (void)hipFree(p->d_tile_items);
    if (p->d_total) (void)hipFree(p->d_total);
    if (p->d_sizes) (void)hipFree(p->d_sizes);
    if (p->ctrl) (void)hipFree(p->ctrl);
    if (p->status) (void)hipFree(p->status);
    if (p->tile_off) (void)hipFree(p->tile_off);
    delete p;
}

int count_launch(lscomp_plan p, hipStream_t stream)
{
    if (++p->launches >= kEpochClearPeriod) {
        int st = hip_fail(launch_fill_u64(p->status, p->max_tiles, 0, stream), "fill");
        if (!st && p->tile_off) st = hip_fail(launch_fill_u64(p->tile_off, p->max_tiles, 0, stream), "fill");
        if (st) return st;
        p->launches = 1;
    }
    return LSCOMP_OK;
}

LaunchCfg plan_cfg(lscomp_plan p)
{
    // Persistent grid sized for the plan's capacity so that a captured graph
    // stays valid for every later batch; the tile count comes from d_total.
    LaunchCfg c = make_cfg(p->max_tiles, p->direction == 0, true, p->shape);
    c.total_tiles = 0;
    c.d_total_tiles = p->d_total;
    // decompression: every tile's offset from the header bytes first, so no
    // tile waits for another and tile ids are taken ahead
    c.tile_off = p->direction == 0 ? nullptr : p->tile_off;
    c.off_mode = OFF_PHASE;
    return c;
}

// Caller holds p->mu.  Builds the item descriptors and the tile table of
// p->src for the plan's tile size in the next staging buffer and uploads them
// (stream-ordered).
int upload_tables(lscomp_plan p, hipStream_t stream)
{
    const uint32_t n = (uint32_t)p->src.size();
    // build the tables in the next staging buffer (wait until its last upload ran)
    const int sb = p->stage_next;
    int st = hip_fail(hipEventSynchronize(p->stage_done[sb]), "hipEventSynchronize");
    if (st) return st;
    Item *h_items = (Item *)p->h_stage[sb];
    uint32_t *h_tiles = (uint32_t *)(p->h_stage[sb] + sizeof(Item) * p->max_items);
    uint64_t tiles = 0, blocks = 0;
    const bool comp = p->direction == 0;
    for (uint32_t i = 0; i < n; i++) {
        const lscomp_item_u16 &src = p->src[i];
        Item &it = h_items[i];
        st = make_item(src.dims, src.bins, comp ? src.pooling_th : 0.0f, it, p->shape.tile);
        if (st) return st;
        if (!src.d_volume || !src.d_stream || ((uintptr_t)src.d_volume & 1) || ((uintptr_t)src.d_stream & 3))
            return LSCOMP_ERR_INVALID_ARG;
        if (src.stream_bytes < it.hdr_bytes) return comp ? LSCOMP_ERR_CAPACITY : LSCOMP_ERR_CORRUPT;
        blocks += it.nblocks;
        if (tiles + it.ntiles > p->max_tiles || blocks > p->max_blocks) return LSCOMP_ERR_TOO_LARGE;
        it.first_tile = (uint32_t)tiles;
        it.align = align_class(src.d_volume, src.dims.z);
        it.s_len = src.stream_bytes;
        it.result = p->d_sizes + i;
        if (comp) {
            it.vol_in = src.d_volume;
            it.s_out = src.d_stream;
        } else {
            it.vol_out = src.d_volume;
            it.s_in = src.d_stream;
        }
        for (uint32_t t = 0; t < it.ntiles; t++) h_tiles[tiles + t] = i;
        tiles += it.ntiles;
    }
    h_tiles[p->max_tiles] = (uint32_t)tiles;   // staged tile count (copied to d_total)
    st = hip_fail(hipMemcpyAsync(p->d_items, h_items, sizeof(Item) * n, hipMemcpyHostToDevice, stream), "upload");
    if (st) return st;
    st = hip_fail(hipMemcpyAsync(p->d_tile_items, h_tiles, sizeof(uint32_t) * tiles, hipMemcpyHostToDevice, stream),
                  "upload");
    if (st) return st;
    st = hip_fail(hipMemcpyAsync(p->d_total, h_tiles + p->max_tiles, sizeof(uint32_t), hipMemcpyHostToDevice, stream),
                  "upload");
    if (st) return st;
    st = hip_fail(hipEventRecord(p->stage_done[sb], stream), "hipEventRecord");
    if (st) return st;
    p->stage_next = sb ^ 1;
    p->nitems = n;
    p->ntiles = (uint32_t)tiles;
    return LSCOMP_OK;
}

// Caller holds p->mu.  A new shape: a captured graph is dropped (it holds the
// old grid and kernel), and the tables are rebuilt when the tile size changed.
int apply_shape(lscomp_plan p, const Shape &sh, hipStream_t stream)
{
    const bool retile = sh.tile != p->shape.tile;
    const Shape old = p->shape;
    p->shape = sh;
    if (p->exec) {
        (void)hipGraphExecDestroy(p->exec);
        (void)hipGraphDestroy(p->graph);
        p->exec = nullptr;
        p->graph = nullptr;
    }
    if (retile && !p->src.empty()) {
        const int st = upload_tables(p, stream);
        if (st) {
            p->shape = old;
            return st;
        }
    }
    return LSCOMP_OK;
}

hipError_t plan_kernel(lscomp_plan p, hipStream_t stream)
{
    Item unused;
    std::memset(&unused, 0, sizeof unused);
    const LaunchCfg c = plan_cfg(p);
    return p->direction == 0 ? launch_compress(unused, p->d_items, p->d_tile_items, c, p->ctrl, p->status, stream)
                             : launch_decompress(unused, p->d_items, p->d_tile_items, c, p->ctrl, p->status, stream);
}

}  // namespace

extern "C" {

int lscomp_plan_create(int direction, uint32_t max_items, uint64_t max_blocks, lscomp_plan *plan)
{
    if (!plan || (direction != 0 && direction != 1) || max_items == 0 || max_blocks == 0)
        return LSCOMP_ERR_INVALID_ARG;
    lscomp_plan p = new lscomp_plan_s;
    p->direction = direction;
    p->shape = load_shape(direction);
    p->max_items = max_items;
    p->max_blocks = max_blocks;
    // tables for the smallest tile, so the shape can change later; every item
    // wastes at most one partial tile
    p->max_tiles = (max_blocks + 127) / 128 + max_items;
    if (p->max_tiles > 0xffffffffull) { delete p; return LSCOMP_ERR_TOO_LARGE; }
    if (hip_fail(hipGetDevice(&p->device), "hipGetDevice")) { delete p; return LSCOMP_ERR_HIP; }
    p->stage_bytes = table_bytes(max_items, p->max_tiles);
    bool ok = !hip_fail(hipMalloc((void **)&p->d_items, sizeof(Item) * max_items), "hipMalloc") &&
              !hip_fail(hipMalloc((void **)&p->d_tile_items, sizeof(uint32_t) * (p->max_tiles + 1)), "hipMalloc") &&
              !hip_fail(hipMalloc((void **)&p->d_total, sizeof(uint32_t) * 2), "hipMalloc") &&
              !hip_fail(hipMalloc((void **)&p->d_sizes, sizeof(uint64_t) * max_items), "hipMalloc") &&
              !hip_fail(hipMalloc((void **)&p->ctrl, sizeof(Ctrl)), "hipMalloc") &&
              !hip_fail(hipMalloc((void **)&p->status, sizeof(uint64_t) * p->max_tiles), "hipMalloc") &&
              (direction == 0 ||
               !hip_fail(hipMalloc((void **)&p->tile_off, sizeof(uint64_t) * p->max_tiles), "hipMalloc")) &&
              !hip_fail(hipHostMalloc((void **)&p->h_ctrl, sizeof(Ctrl), hipHostMallocDefault), "hipHostMalloc");
    for (int i = 0; ok && i < 2; i++)
        ok = !hip_fail(hipHostMalloc((void **)&p->h_stage[i], p->stage_bytes, hipHostMallocDefault), "hipHostMalloc") &&
             !hip_fail(hipEventCreateWithFlags(&p->stage_done[i], hipEventDisableTiming), "hipEventCreate");
    if (!ok) { free_plan(p); return LSCOMP_ERR_HIP; }
    // zero scratch and tables with a fill kernel, then wait (creation is not stream-ordered)
    ok = !hip_fail(launch_fill_u64((uint64_t *)p->ctrl, sizeof(Ctrl) / 8, 0, nullptr), "fill") &&
         !hip_fail(launch_fill_u64(p->status, p->max_tiles, 0, nullptr), "fill") &&
         (!p->tile_off || !hip_fail(launch_fill_u64(p->tile_off, p->max_tiles, 0, nullptr), "fill")) &&
         !hip_fail(launch_fill_u64((uint64_t *)p->d_total, 1, 0, nullptr), "fill") &&
         !hip_fail(launch_fill_u64(p->d_sizes, max_items, 0, nullptr), "fill") &&
         !hip_fail(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    if (!ok) { free_plan(p); return LSCOMP_ERR_HIP; }
    *plan = p;
    return LSCOMP_OK;
}

int lscomp_plan_set_items_u16(lscomp_plan p, const lscomp_item_u16 *items, uint32_t n, hipStream_t stream)
{
    if (!p || !items || n == 0 || n > p->max_items) return LSCOMP_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(p->mu);
    std::vector<lscomp_item_u16> prev;
    prev.swap(p->src);
    p->src.assign(items, items + n);
    const int st = upload_tables(p, stream);
    if (st) p->src.swap(prev);
    return st;
}

int lscomp_plan_set_shape(lscomp_plan p, const lscomp_launch_shape *shape, hipStream_t stream)
{
    if (!p || !shape) return LSCOMP_ERR_INVALID_ARG;
    Shape sh{shape->tile, shape->wg_per_cu, shape->lb_window, shape->counters};
    int st = normalize_shape(p->direction, sh);
    if (st) return st;
    std::lock_guard<std::mutex> lk(p->mu);
    return apply_shape(p, sh, stream);
}

int lscomp_plan_get_shape(lscomp_plan p, lscomp_launch_shape *shape)
{
    if (!p || !shape) return LSCOMP_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(p->mu);
    shape->tile = p->shape.tile;
    shape->wg_per_cu = p->shape.wgpcu;
    shape->lb_window = p->shape.win;
    shape->counters = p->shape.nc;
    return LSCOMP_OK;
}

int lscomp_plan_launch(lscomp_plan p, hipStream_t stream)
{
    if (!p) return LSCOMP_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(p->mu);
    if (p->nitems == 0) return LSCOMP_ERR_INVALID_ARG;
    int st = count_launch(p, stream);
    if (st) return st;
    return hip_fail(plan_kernel(p, stream), "plan launch");
}

int lscomp_plan_launch_graph(lscomp_plan p, hipStream_t stream)
{
    if (!p) return LSCOMP_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(p->mu);
    if (p->nitems == 0) return LSCOMP_ERR_INVALID_ARG;
    if (!p->exec) {
        hipStream_t cap;
        int st = hip_fail(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
        if (st) return st;
        st = hip_fail(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        if (!st) {
            const hipError_t le = plan_kernel(p, cap);
            const hipError_t ce = hipStreamEndCapture(cap, &p->graph);
            st = hip_fail(le != hipSuccess ? le : ce, "graph capture");
        }
        if (!st) st = hip_fail(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        (void)hipStreamDestroy(cap);
        if (st) {
            if (p->graph) (void)hipGraphDestroy(p->graph);
            p->graph = nullptr;
            p->exec = nullptr;
            return st;
        }
    }
    int st = count_launch(p, stream);
    if (st) return st;
    return hip_fail(hipGraphLaunch(p->exec, stream), "hipGraphLaunch");
}

int lscomp_plan_autotune(lscomp_plan p, int reps, hipStream_t stream, lscomp_tune_entry *table, uint32_t *n_table)
{
    if (!p || reps < 1) return LSCOMP_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(p->mu);
    if (p->nitems == 0) return LSCOMP_ERR_INVALID_ARG;
    const Shape saved = p->shape;
    std::vector<lscomp_tune_entry> log;
    auto measure = [&](const Shape &s) -> float {
        if (apply_shape(p, s, stream) != LSCOMP_OK) return -1.0f;
        int rs = LSCOMP_OK;
        const float ms = time_launches(stream, reps, [&] {
            const int c = count_launch(p, stream);
            return c ? c : hip_fail(plan_kernel(p, stream), "plan launch");
        }, &rs);
        return rs == LSCOMP_OK ? ms : -1.0f;
    };
    Shape best{};
    int st = tune_shape(p->direction, true, saved, measure, best, log);
    if (!st) st = apply_shape(p, best, stream);
    if (st) {
        (void)apply_shape(p, saved, stream);
        return st;
    }
    if (table && n_table) {
        uint32_t k = 0;
        for (; k < *n_table && k < log.size(); k++) table[k] = log[k];
        *n_table = k;
    }
    return LSCOMP_OK;
}

const uint64_t *lscomp_plan_device_sizes(lscomp_plan p) { return p ? p->d_sizes : nullptr; }

}  // extern "C"

namespace lsc {
// Stream-ordered: copies the plan's error word into *h_err (pinned host
// memory) and clears it, so a batch's completions can carry the launch's
// status (served form).
int plan_take_error_async(lscomp_plan p, uint32_t *h_err, hipStream_t stream)
{
    if (!p || !h_err) return LSCOMP_ERR_INVALID_ARG;
    int st = hip_fail(hipMemcpyAsync(h_err, &p->ctrl->error, sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
                      "hipMemcpyAsync");
    if (!st) st = hip_fail(launch_fill_u64((uint64_t *)&p->ctrl->error, 1, 0, stream), "fill");
    return st;
}
}  // namespace lsc

extern "C" {

int lscomp_plan_sizes(lscomp_plan p, uint64_t *host_sizes, hipStream_t stream)
{
    if (!p || !host_sizes) return LSCOMP_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(p->mu);
    int st = hip_fail(hipMemcpyAsync(host_sizes, p->d_sizes, sizeof(uint64_t) * p->nitems, hipMemcpyDeviceToHost,
                                     stream), "hipMemcpyAsync");
    if (st) return st;
    return hip_fail(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

int lscomp_plan_status(lscomp_plan p, hipStream_t stream)
{
    if (!p) return LSCOMP_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(p->mu);
    int st = hip_fail(hipMemcpyAsync(p->h_ctrl, p->ctrl, kCtrlHead, hipMemcpyDeviceToHost, stream), "readback");
    if (st) return st;
    st = hip_fail(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (st) return st;
    const uint32_t e = p->h_ctrl->error;
    if (!e) return LSCOMP_OK;
    st = hip_fail(launch_fill_u64((uint64_t *)&p->ctrl->error, 1, 0, stream), "fill");
    if (!st) st = hip_fail(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (st) return st;
    return (e & ERRF_TIMEOUT) ? LSCOMP_ERR_INTERNAL : LSCOMP_ERR_CORRUPT;
}

int lscomp_plan_destroy(lscomp_plan p)
{
    if (!p) return LSCOMP_ERR_INVALID_ARG;
    (void)hipDeviceSynchronize();
    free_plan(p);
    return LSCOMP_OK;
}

}  // extern "C"
