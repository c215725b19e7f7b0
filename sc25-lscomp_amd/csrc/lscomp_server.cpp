// lscomp_server.cpp -- served form: continuous batching of frame compression
// requests under an HBM budget, with two engines behind one API.
//
// Device engine (lscomp_server_create_device): a device-resident request
//   loop.  serve_kernel (lscomp_kernels.hip) runs persistently: its dispatcher
//   workgroup takes requests from a ring in host-pinned memory and publishes
//   their tiles, the other workgroups compress them, and each request's last
//   tile writes a completion record into host-pinned memory.  The host thread
//   here only admits requests (HBM arena range + ring slot), reaps completion
//   records, and launches a new kernel instance when the previous one has
//   closed (idle, or its tables full) while requests wait: no launch per batch.
// Host engine (lscomp_server_create): a scheduler thread packs the requests
//   admitted since its last launch into one table-mode launch of one of
//   several plans (slots), each on its own stream, so batches overlap.
//
// Admission (both engines): in arrival order, while the request's worst-case
// stream fits into the HBM arena (first-fit free list with coalescing).  An
// output range stays reserved until lscomp_server_release.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/lscomp.h"
#include "lscomp_host.h"

namespace {

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// First-fit arena over [0, size) with coalescing free ranges.
class Arena {
public:
    explicit Arena(uint64_t size) { free_[0] = size; }
    bool alloc(uint64_t n, uint64_t &off)
    {
        n = (n + 255) & ~255ull;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                off = it->first;
                const uint64_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[off + n] = rest;
                used_[off] = n;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off)
    {
        auto u = used_.find(off);
        if (u == used_.end()) return;
        uint64_t n = u->second;
        used_.erase(u);
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + n == next->first) {
            n += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += n;
                return;
            }
        }
        free_[off] = n;
    }

private:
    std::map<uint64_t, uint64_t> free_, used_;
};

struct Req {
    uint64_t id = 0;
    lscomp_request_u16 r{};
    lsc::Item item{};          // device engine: prepared at submit
    uint64_t cap = 0, off = 0;
    double t_submit = 0, t_admit = 0, t_launch = 0;
};

// Host engine: one plan per slot.
struct Slot {
    lscomp_plan plan = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    uint64_t *h_sizes = nullptr;   // pinned
    uint32_t *h_err = nullptr;     // pinned: the launch's error word
    std::vector<Req> reqs;
    bool busy = false;
};

// Device engine state.
struct DevLoop {
    lsc::SrvReq *ring = nullptr;       // host-pinned (coherent)
    lsc::SrvComp *comp = nullptr;      // host-pinned (coherent)
    lsc::SrvHost *hc = nullptr;        // host-pinned (coherent)
    lsc::SrvArgs args{};               // device-side pointers
    lsc::Ctrl *ctrl = nullptr;
    uint64_t *status = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    uint32_t grid = 0;
    lsc::Sched sc{};
    bool running = false;
    uint64_t wr = 0;                   // next sequence number to publish
    // Published requests not yet complete: sequence number q lives in slot
    // q % kSrvRing while live[] is set; lo is the oldest live one (or wr).  A
    // reaping pass scans [lo, wr) in order over two flat arrays (a node-based
    // map scanned at a full ring of 4096 cost most of the host thread's time).
    std::vector<Req> req;
    std::vector<uint8_t> live;
    uint64_t lo = 0, nlive = 0;
    uint64_t instances = 0;
};

// Map and status words are cleared every kEpochClearPeriod instances: their
// epoch tags cannot repeat in between.

uint64_t load_u64(const uint64_t *p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
uint32_t load_u32(const uint32_t *p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
void store_u64(uint64_t *p, uint64_t v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }
void store_u32(uint32_t *p, uint32_t v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }

int status_of(uint32_t errf)
{
    if (errf & lsc::ERRF_TIMEOUT) return LSCOMP_ERR_INTERNAL;
    if (errf & lsc::ERRF_CORRUPT) return LSCOMP_ERR_CORRUPT;
    return LSCOMP_OK;
}

}  // namespace

struct lscomp_server_s {
    bool device = false;               // engine
    uint64_t max_blocks = 0;
    uint64_t budget = 0;
    uint8_t *d_arena = nullptr;
    Arena *arena = nullptr;
    // host engine
    uint32_t max_items = 0;
    std::vector<Slot> slots;
    std::deque<Req> admitted;
    // device engine
    DevLoop dl;
    // common
    std::deque<Req> pending;
    std::deque<lscomp_completion> done;
    std::unordered_map<uint64_t, uint64_t> held;   // finished id -> arena offset
    uint64_t next_id = 1, outstanding = 0;
    uint64_t launches = 0, completed = 0;
    int first_error = LSCOMP_OK;                   // first non-OK completion since the last drain
    std::mutex mu;
    std::condition_variable cv, cv_done;
    bool stop = false;
    std::thread worker;
    int device_id = 0;
};

namespace {

// Caller holds s->mu.
void complete(lscomp_server s, const Req &q, uint64_t bytes, int st, double t)
{
    lscomp_completion c{};
    c.id = q.id;
    c.tag = q.r.tag;
    c.stream_bytes = bytes;
    c.status = st != LSCOMP_OK ? st : (bytes > q.cap ? LSCOMP_ERR_CAPACITY : LSCOMP_OK);
    c.d_stream = c.status == LSCOMP_OK ? s->d_arena + q.off : nullptr;
    c.t_submit = q.t_submit;
    c.t_admit = q.t_admit;
    c.t_launch = q.t_launch;
    c.t_done = t;
    if (c.status != LSCOMP_OK && s->first_error == LSCOMP_OK) s->first_error = c.status;
    s->held[q.id] = q.off;
    s->done.push_back(c);
    s->outstanding--;
    s->completed++;
}

// Caller holds s->mu.  Arrival order: the head waits for room.
bool admit_one(lscomp_server s, Req &q)
{
    uint64_t off;
    if (!s->arena->alloc(q.cap, off)) return false;
    q.off = off;
    q.t_admit = now_s();
    return true;
}

// ---- host engine -----------------------------------------------------------

void host_retire(lscomp_server s)
{
    for (Slot &sl : s->slots) {
        if (!sl.busy || hipEventQuery(sl.done) != hipSuccess) continue;
        const int st = status_of(*sl.h_err);
        const double t = now_s();
        for (size_t i = 0; i < sl.reqs.size(); i++) complete(s, sl.reqs[i], sl.h_sizes[i], st, t);
        sl.reqs.clear();
        sl.busy = false;
        s->cv_done.notify_all();
    }
}

void host_admit(lscomp_server s)
{
    while (!s->pending.empty()) {
        Req &q = s->pending.front();
        if (!admit_one(s, q)) break;
        s->admitted.push_back(q);
        s->pending.pop_front();
    }
}

void host_launch(lscomp_server s)
{
    for (Slot &sl : s->slots) {
        if (s->admitted.empty()) return;
        if (sl.busy) continue;
        std::vector<lscomp_item_u16> items;
        uint64_t blocks = 0;
        while (!s->admitted.empty() && items.size() < s->max_items) {
            Req &q = s->admitted.front();
            const uint64_t nb = lscomp_num_blocks_u16(q.r.dims);
            if (!items.empty() && blocks + nb > s->max_blocks) break;
            lscomp_item_u16 it{};
            it.dims = q.r.dims;
            it.bins = q.r.bins;
            it.pooling_th = q.r.pooling_th;
            it.d_volume = const_cast<uint16_t *>(q.r.d_frame);
            it.d_stream = s->d_arena + q.off;
            it.stream_bytes = q.cap;
            items.push_back(it);
            blocks += nb;
            sl.reqs.push_back(q);
            s->admitted.pop_front();
        }
        int st = lscomp_plan_set_items_u16(sl.plan, items.data(), (uint32_t)items.size(), sl.stream);
        if (st == LSCOMP_OK) st = lscomp_plan_launch(sl.plan, sl.stream);
        if (st == LSCOMP_OK)
            st = lsc::hip_fail(hipMemcpyAsync(sl.h_sizes, lscomp_plan_device_sizes(sl.plan), 8 * items.size(),
                                              hipMemcpyDeviceToHost, sl.stream), "hipMemcpyAsync");
        if (st == LSCOMP_OK) st = lsc::plan_take_error_async(sl.plan, sl.h_err, sl.stream);
        if (st == LSCOMP_OK) st = lsc::hip_fail(hipEventRecord(sl.done, sl.stream), "hipEventRecord");
        const double t = now_s();
        for (Req &q : sl.reqs) q.t_launch = t;
        if (st != LSCOMP_OK) {   // fail this batch's requests without running them
            for (Req &q : sl.reqs) complete(s, q, 0, st, t);
            sl.reqs.clear();
            s->cv_done.notify_all();
            continue;
        }
        s->launches++;
        sl.busy = true;
    }
}

bool host_busy(lscomp_server s)
{
    for (const Slot &sl : s->slots)
        if (sl.busy) return true;
    return false;
}

// ---- device engine ---------------------------------------------------------

void dev_reap(lscomp_server s)
{
    DevLoop &d = s->dl;
    bool any = false;
    for (uint64_t seq = d.lo; seq < d.wr && d.nlive; seq++) {
        const size_t k = seq % lsc::kSrvRing;
        if (!d.live[k]) continue;
        lsc::SrvComp &c = d.comp[k];
        if (load_u64(&c.seq) != seq + 1) continue;
        const uint64_t bytes = load_u64(&c.bytes);
        const uint32_t f = load_u32(&c.status);
        const int st = (f & lsc::SRV_TIMEOUT) ? LSCOMP_ERR_INTERNAL : LSCOMP_OK;   // capacity: from the size
        complete(s, d.req[k], bytes, st, now_s());
        d.live[k] = 0;
        d.nlive--;
        any = true;
    }
    while (d.lo < d.wr && !d.live[d.lo % lsc::kSrvRing]) d.lo++;
    if (any) s->cv_done.notify_all();
}

void dev_admit(lscomp_server s)
{
    DevLoop &d = s->dl;
    while (!s->pending.empty()) {
        // the ring slot of sequence number wr is free once wr - kSrvRing has completed
        if (d.nlive && d.wr - d.lo >= lsc::kSrvRing) break;
        Req &q = s->pending.front();
        if (!admit_one(s, q)) break;
        q.item.vol_in = q.r.d_frame;
        q.item.s_out = s->d_arena + q.off;
        q.item.s_len = q.cap;
        q.item.seq = d.wr;
        q.t_launch = now_s();   // handed to the running loop
        std::memcpy(&d.ring[d.wr % lsc::kSrvRing].item, &q.item, sizeof(lsc::Item));
        if (!d.nlive) d.lo = d.wr;
        d.req[d.wr % lsc::kSrvRing] = q;
        d.live[d.wr % lsc::kSrvRing] = 1;
        d.nlive++;
        d.wr++;
        store_u64(&d.hc->wr, d.wr);   // release: the slot is written before it is published
        s->pending.pop_front();
    }
}

int dev_launch(lscomp_server s)
{
    DevLoop &d = s->dl;
    if (d.running && hipEventQuery(d.ev) == hipSuccess) d.running = false;
    if (d.running || load_u64(&d.hc->rd) >= d.wr) return LSCOMP_OK;
    // a new instance: clear its close word (and, rarely, the tagged words)
    int st = LSCOMP_OK;
    if (++d.instances % lsc::kEpochClearPeriod == 0) {
        st = lsc::hip_fail(lsc::launch_fill_u64(d.args.map, lsc::kSrvMaxTiles, 0, d.stream), "fill");
        if (!st) st = lsc::hip_fail(lsc::launch_fill_u64(d.status, lsc::kSrvMaxTiles, 0, d.stream), "fill");
    }
    if (!st) st = lsc::hip_fail(lsc::launch_fill_u64(&d.args.dev->close, 1, 0, d.stream), "fill");
    if (!st) st = lsc::hip_fail(lsc::launch_serve(d.args, d.grid, d.sc, d.ctrl, d.status, d.stream), "serve launch");
    if (!st) st = lsc::hip_fail(hipEventRecord(d.ev, d.stream), "hipEventRecord");
    if (st) return st;
    d.running = true;
    s->launches++;
    return LSCOMP_OK;
}

// Fail every request the loop cannot run (the launch failed).
void dev_fail_all(lscomp_server s, int st)
{
    const double t = now_s();
    DevLoop &d = s->dl;
    for (uint64_t seq = d.lo; seq < d.wr; seq++) {
        const size_t k = seq % lsc::kSrvRing;
        if (!d.live[k]) continue;
        complete(s, d.req[k], 0, st, t);
        d.live[k] = 0;
    }
    d.nlive = 0;
    d.lo = d.wr;
    s->cv_done.notify_all();
}

void worker_main(lscomp_server s)
{
    (void)hipSetDevice(s->device_id);
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop) {
        bool busy;
        if (s->device) {
            dev_reap(s);
            dev_admit(s);
            const int st = dev_launch(s);
            if (st != LSCOMP_OK) dev_fail_all(s, st);
            busy = s->dl.nlive != 0 || s->dl.running;
        } else {
            host_retire(s);
            host_admit(s);
            host_launch(s);
            busy = host_busy(s);
        }
        if (busy) {
            // poll completion records / events without holding the lock
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        } else if (s->pending.empty() && (s->device || s->admitted.empty())) {
            s->cv.wait(lk, [&] { return s->stop || !s->pending.empty(); });
        } else {
            // requests wait for arena room: a release wakes this thread
            s->cv.wait_for(lk, std::chrono::microseconds(200));
        }
    }
}

int dev_create(lscomp_server s, uint32_t idle_us)
{
    using namespace lsc;
    DevLoop &d = s->dl;
    d.req.resize(lsc::kSrvRing);
    d.live.assign(lsc::kSrvRing, 0);
    const unsigned pf = hipHostMallocCoherent | hipHostMallocMapped;
    if (hipHostMalloc((void **)&d.ring, sizeof(SrvReq) * kSrvRing, pf) != hipSuccess ||
        hipHostMalloc((void **)&d.comp, sizeof(SrvComp) * kSrvRing, pf) != hipSuccess ||
        hipHostMalloc((void **)&d.hc, sizeof(SrvHost), pf) != hipSuccess)
        return LSCOMP_ERR_HIP;
    std::memset(d.ring, 0, sizeof(SrvReq) * kSrvRing);
    std::memset(d.comp, 0, sizeof(SrvComp) * kSrvRing);
    std::memset(d.hc, 0, sizeof(SrvHost));
    SrvArgs &a = d.args;
    if (hipHostGetDevicePointer((void **)&a.ring, d.ring, 0) != hipSuccess ||
        hipHostGetDevicePointer((void **)&a.comp, d.comp, 0) != hipSuccess ||
        hipHostGetDevicePointer((void **)&a.host, d.hc, 0) != hipSuccess)
        return LSCOMP_ERR_HIP;
    int st = hip_fail(hipMalloc((void **)&a.dev, sizeof(SrvDev)), "hipMalloc");
    if (!st) st = hip_fail(hipMalloc((void **)&a.items, sizeof(ItemSlot) * kSrvMaxReq), "hipMalloc");
    if (!st) st = hip_fail(hipMalloc((void **)&a.map, sizeof(uint64_t) * kSrvMaxTiles), "hipMalloc");
    if (!st) st = hip_fail(hipMalloc((void **)&a.done, sizeof(uint32_t) * kSrvMaxReq), "hipMalloc");
    if (!st) st = hip_fail(hipMalloc((void **)&a.sizes, sizeof(uint64_t) * kSrvMaxReq), "hipMalloc");
    if (!st) st = hip_fail(hipMalloc((void **)&d.status, sizeof(uint64_t) * kSrvMaxTiles), "hipMalloc");
    if (!st) st = hip_fail(hipMalloc((void **)&d.ctrl, sizeof(Ctrl)), "hipMalloc");
    if (!st) st = hip_fail(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking), "hipStreamCreate");
    if (!st) st = hip_fail(hipEventCreateWithFlags(&d.ev, hipEventDisableTiming), "hipEventCreate");
    if (st) return st;
    st = hip_fail(launch_fill_u64((uint64_t *)a.dev, sizeof(SrvDev) / 8, 0, d.stream), "fill");
    if (!st) st = hip_fail(launch_fill_u64(a.map, kSrvMaxTiles, 0, d.stream), "fill");
    if (!st) st = hip_fail(launch_fill_u64(d.status, kSrvMaxTiles, 0, d.stream), "fill");
    if (!st) st = hip_fail(launch_fill_u64((uint64_t *)d.ctrl, sizeof(Ctrl) / 8, 0, d.stream), "fill");
    if (!st) st = hip_fail(hipStreamSynchronize(d.stream), "hipStreamSynchronize");
    if (st) return st;
    a.idle_ticks = (uint64_t)idle_us * 100u;   // s_memrealtime runs at 100 MHz
    const int ncu = num_cus();
    d.grid = (uint32_t)((ncu > 0 ? ncu : 1) * workgroups_per_cu_serve());
    if (d.grid < 2) d.grid = 2;
    const Shape shp = load_shape(0);
    d.sc.nc = shp.nc < d.grid - 1 ? (shp.nc ? shp.nc : 1u) : d.grid - 1;
    d.sc.win = shp.win ? (shp.win <= kSrvTile ? shp.win : kSrvTile) : default_window(kSrvTile);
    return LSCOMP_OK;
}

void dev_destroy(lscomp_server s)
{
    DevLoop &d = s->dl;
    if (d.hc) store_u32(&d.hc->stop, 1u);   // a running instance closes
    if (d.stream) (void)hipStreamSynchronize(d.stream);
    if (d.ev) (void)hipEventDestroy(d.ev);
    if (d.stream) (void)hipStreamDestroy(d.stream);
    lsc::SrvArgs &a = d.args;
    if (a.dev) (void)hipFree(a.dev);
    if (a.items) (void)hipFree(a.items);
    if (a.map) (void)hipFree(a.map);
    if (a.done) (void)hipFree(a.done);
    if (a.sizes) (void)hipFree(a.sizes);
    if (d.status) (void)hipFree(d.status);
    if (d.ctrl) (void)hipFree(d.ctrl);
    if (d.ring) (void)hipHostFree(d.ring);
    if (d.comp) (void)hipHostFree(d.comp);
    if (d.hc) (void)hipHostFree(d.hc);
}

int common_create(lscomp_server s, uint64_t max_blocks, uint64_t budget_bytes)
{
    s->max_blocks = max_blocks;
    s->budget = budget_bytes;
    if (hipGetDevice(&s->device_id) != hipSuccess || hipMalloc((void **)&s->d_arena, budget_bytes) != hipSuccess)
        return LSCOMP_ERR_HIP;
    s->arena = new Arena(budget_bytes);
    return LSCOMP_OK;
}

}  // namespace

extern "C" {

int lscomp_server_create(uint32_t slots, uint32_t max_items, uint64_t max_blocks, uint64_t budget_bytes,
                         lscomp_server *out)
{
    if (!out || slots == 0 || slots > 16 || max_items == 0 || max_blocks == 0 || budget_bytes < 256)
        return LSCOMP_ERR_INVALID_ARG;
    lscomp_server s = new lscomp_server_s;
    s->max_items = max_items;
    int st = common_create(s, max_blocks, budget_bytes);
    if (st) {
        lscomp_server_destroy(s);
        return st;
    }
    s->slots.resize(slots);
    for (Slot &sl : s->slots) {
        st = lscomp_plan_create(0, max_items, max_blocks, &sl.plan);
        if (st == LSCOMP_OK &&
            (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
             hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
             hipHostMalloc((void **)&sl.h_sizes, 8 * max_items, hipHostMallocDefault) != hipSuccess ||
             hipHostMalloc((void **)&sl.h_err, sizeof(uint32_t), hipHostMallocDefault) != hipSuccess))
            st = LSCOMP_ERR_HIP;
        if (st != LSCOMP_OK) {
            lscomp_server_destroy(s);
            return st;
        }
        *sl.h_err = 0;
    }
    s->worker = std::thread(worker_main, s);
    *out = s;
    return LSCOMP_OK;
}

int lscomp_server_create_device(uint64_t max_blocks, uint64_t budget_bytes, uint32_t idle_us, lscomp_server *out)
{
    if (!out || max_blocks == 0 || max_blocks > lsc::kMaxBlocksPerItem || budget_bytes < 256 || idle_us == 0)
        return LSCOMP_ERR_INVALID_ARG;
    lscomp_server s = new lscomp_server_s;
    s->device = true;
    int st = common_create(s, max_blocks, budget_bytes);
    if (!st) st = dev_create(s, idle_us);
    if (st) {
        lscomp_server_destroy(s);
        return st;
    }
    s->worker = std::thread(worker_main, s);
    *out = s;
    return LSCOMP_OK;
}

int lscomp_server_submit_u16(lscomp_server s, const lscomp_request_u16 *req, uint64_t *id)
{
    if (!s || !req || !req->d_frame || ((uintptr_t)req->d_frame & 1)) return LSCOMP_ERR_INVALID_ARG;
    Req q;
    int st = lsc::make_item(req->dims, req->bins, req->pooling_th, q.item, s->device ? lsc::kSrvTile : 128);
    if (st) return st;
    const uint64_t nb = lscomp_num_blocks_u16(req->dims);
    q.r = *req;
    q.cap = lscomp_max_compressed_bytes_u16(req->dims);
    if (nb > s->max_blocks) return LSCOMP_ERR_TOO_LARGE;
    // a request whose worst case cannot fit the arena would block every later one
    if (((q.cap + 255) & ~255ull) > s->budget) return LSCOMP_ERR_CAPACITY;
    q.item.align = lsc::align_class(req->d_frame, req->dims.z);
    q.t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        q.id = s->next_id++;
        s->pending.push_back(q);
        s->outstanding++;
    }
    s->cv.notify_one();
    if (id) *id = q.id;
    return LSCOMP_OK;
}

int lscomp_server_poll(lscomp_server s, lscomp_completion *out, uint32_t max, uint32_t *n)
{
    if (!s || !n || (max && !out)) return LSCOMP_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    uint32_t k = 0;
    while (k < max && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    *n = k;
    return LSCOMP_OK;
}

int lscomp_server_release(lscomp_server s, uint64_t id)
{
    if (!s) return LSCOMP_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->held.find(id);
        if (it == s->held.end()) return LSCOMP_ERR_INVALID_ARG;
        s->arena->release(it->second);
        s->held.erase(it);
    }
    s->cv.notify_one();
    return LSCOMP_OK;
}

int lscomp_server_drain(lscomp_server s)
{
    if (!s) return LSCOMP_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
    int st = s->first_error;
    s->first_error = LSCOMP_OK;
    if (!s->device) {
        for (Slot &sl : s->slots) {
            const int e = lscomp_plan_status(sl.plan, sl.stream);
            if (e != LSCOMP_OK && st == LSCOMP_OK) st = e;
        }
    }
    return st;
}

int lscomp_server_stats(lscomp_server s, lscomp_server_stats_t *out)
{
    if (!s || !out) return LSCOMP_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    out->submitted = s->next_id - 1;
    out->completed = s->completed;
    out->launches = s->launches;
    out->device_loop = s->device ? 1u : 0u;
    out->reserved = 0;
    return LSCOMP_OK;
}

int lscomp_server_replay_u16(lscomp_server s, const lscomp_request_u16 *reqs, const double *arrival_s, uint32_t n,
                             lscomp_completion *out, uint32_t flags)
{
    if (!s || !reqs || !arrival_s || !out || (flags & ~(uint32_t)LSCOMP_REPLAY_HOLD)) return LSCOMP_ERR_INVALID_ARG;
    std::unordered_map<uint64_t, uint32_t> index;
    std::vector<lscomp_completion> buf(64);
    uint32_t got = 0;
    auto collect = [&]() {
        uint32_t k = 0;
        lscomp_server_poll(s, buf.data(), (uint32_t)buf.size(), &k);
        for (uint32_t i = 0; i < k; i++) {
            auto it = index.find(buf[i].id);
            if (it == index.end()) continue;
            out[it->second] = buf[i];
            if (!(flags & LSCOMP_REPLAY_HOLD)) {
                lscomp_server_release(s, buf[i].id);
                out[it->second].d_stream = nullptr;   // the range may be reused at once
            }
            got++;
        }
    };
    const double t0 = now_s();
    for (uint32_t i = 0; i < n; i++) {
        while (now_s() - t0 < arrival_s[i]) collect();
        uint64_t id = 0;
        int st = lscomp_server_submit_u16(s, &reqs[i], &id);
        if (st) return st;
        index[id] = i;
    }
    while (got < n) collect();
    return LSCOMP_OK;
}

int lscomp_server_destroy(lscomp_server s)
{
    if (!s) return LSCOMP_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    if (s->device) dev_destroy(s);
    for (Slot &sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.plan) lscomp_plan_destroy(sl.plan);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.h_sizes) (void)hipHostFree(sl.h_sizes);
        if (sl.h_err) (void)hipHostFree(sl.h_err);
    }
    if (s->d_arena) (void)hipFree(s->d_arena);
    delete s->arena;
    delete s;
    return LSCOMP_OK;
}

}  // extern "C"
