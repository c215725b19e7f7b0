// lscomp_server.cpp -- served form: continuous batching over several plans.
//
// Scheduler thread loop:
//   1. retire finished batches (hipEventQuery on each busy slot), publish
//      their completions with the per-item sizes read back by the slot;
//   2. admit pending requests in arrival order while their worst-case stream
//      fits into the HBM arena (first-fit free list with coalescing);
//   3. if a slot is free, put every admitted request (up to the slot's plan
//      capacity) into one batch: upload its tables, one persistent launch,
//      size readback into pinned memory, event record -- all on the slot's
//      stream, so several batches run on the device at once.
#include <hip/hip_runtime.h>

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
    uint64_t id;
    lscomp_request_u16 r;
    uint64_t cap = 0, off = 0;
    double t_submit = 0, t_admit = 0, t_launch = 0;
};

struct Slot {
    lscomp_plan plan = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    uint64_t *h_sizes = nullptr;   // pinned
    std::vector<Req> reqs;
    bool busy = false;
};

}  // namespace

struct lscomp_server_s {
    uint32_t max_items = 0;
    uint64_t max_blocks = 0;
    uint8_t *d_arena = nullptr;
    Arena *arena = nullptr;
    std::vector<Slot> slots;
    std::deque<Req> pending, admitted;
    std::deque<lscomp_completion> done;
    std::unordered_map<uint64_t, uint64_t> held;   // finished id -> arena offset
    uint64_t next_id = 1, outstanding = 0;
    std::mutex mu;
    std::condition_variable cv, cv_done;
    bool stop = false;
    int error = LSCOMP_OK;
    std::thread worker;
    int device = 0;
};

namespace {

// Caller holds s->mu.
void retire(lscomp_server s)
{
    for (Slot &sl : s->slots) {
        if (!sl.busy || hipEventQuery(sl.done) != hipSuccess) continue;
        // compression reports per-item overflow through the sizes; launch-level
        // errors (look-back timeout) are collected by lscomp_server_drain
        const int st = LSCOMP_OK;
        const double t = now_s();
        for (size_t i = 0; i < sl.reqs.size(); i++) {
            const Req &q = sl.reqs[i];
            lscomp_completion c{};
            c.id = q.id;
            c.tag = q.r.tag;
            c.stream_bytes = sl.h_sizes[i];
            c.status = st != LSCOMP_OK ? st : (c.stream_bytes > q.cap ? LSCOMP_ERR_CAPACITY : LSCOMP_OK);
            c.d_stream = s->d_arena + q.off;
            c.t_submit = q.t_submit;
            c.t_admit = q.t_admit;
            c.t_launch = q.t_launch;
            c.t_done = t;
            s->held[q.id] = q.off;
            s->done.push_back(c);
            s->outstanding--;
        }
        sl.reqs.clear();
        sl.busy = false;
        s->cv_done.notify_all();
    }
}

// Caller holds s->mu.
void admit(lscomp_server s)
{
    while (!s->pending.empty()) {
        Req &q = s->pending.front();
        uint64_t off;
        if (!s->arena->alloc(q.cap, off)) break;   // arrival order: wait for room
        q.off = off;
        q.t_admit = now_s();
        s->admitted.push_back(q);
        s->pending.pop_front();
    }
}

// Caller holds s->mu.
void launch(lscomp_server s)
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
        if (st == LSCOMP_OK) st = lsc::hip_fail(hipEventRecord(sl.done, sl.stream), "hipEventRecord");
        const double t = now_s();
        for (Req &q : sl.reqs) q.t_launch = t;
        if (st != LSCOMP_OK) {
            // fail this batch's requests without running them
            for (Req &q : sl.reqs) {
                lscomp_completion c{};
                c.id = q.id;
                c.tag = q.r.tag;
                c.status = st;
                c.t_submit = q.t_submit;
                c.t_admit = q.t_admit;
                c.t_launch = c.t_done = t;
                s->held[q.id] = q.off;
                s->done.push_back(c);
                s->outstanding--;
            }
            sl.reqs.clear();
            s->cv_done.notify_all();
            continue;
        }
        sl.busy = true;
    }
}

void worker_main(lscomp_server s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop) {
        retire(s);
        admit(s);
        launch(s);
        bool busy = false;
        for (const Slot &sl : s->slots) busy |= sl.busy;
        if (busy) {
            // poll the device without holding the lock
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        } else if (s->pending.empty() && s->admitted.empty()) {
            s->cv.wait(lk, [&] { return s->stop || !s->pending.empty(); });
        } else {
            // admitted work but no free slot cannot happen when idle; pending
            // work blocked on the budget waits for a release
            s->cv.wait_for(lk, std::chrono::microseconds(200));
        }
    }
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
    s->max_blocks = max_blocks;
    if (hipGetDevice(&s->device) != hipSuccess || hipMalloc((void **)&s->d_arena, budget_bytes) != hipSuccess) {
        delete s;
        return LSCOMP_ERR_HIP;
    }
    s->arena = new Arena(budget_bytes);
    s->slots.resize(slots);
    for (Slot &sl : s->slots) {
        int st = lscomp_plan_create(0, max_items, max_blocks, &sl.plan);
        if (st == LSCOMP_OK &&
            (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
             hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
             hipHostMalloc((void **)&sl.h_sizes, 8 * max_items, hipHostMallocDefault) != hipSuccess))
            st = LSCOMP_ERR_HIP;
        if (st != LSCOMP_OK) {
            s->stop = true;
            lscomp_server_destroy(s);
            return st;
        }
    }
    s->worker = std::thread(worker_main, s);
    *out = s;
    return LSCOMP_OK;
}

int lscomp_server_submit_u16(lscomp_server s, const lscomp_request_u16 *req, uint64_t *id)
{
    if (!s || !req || !req->d_frame || ((uintptr_t)req->d_frame & 1)) return LSCOMP_ERR_INVALID_ARG;
    lsc::Item probe;
    int st = lsc::make_item(req->dims, req->bins, req->pooling_th, probe, 128);
    if (st) return st;
    const uint64_t nb = lscomp_num_blocks_u16(req->dims);
    Req q;
    q.r = *req;
    q.cap = lscomp_max_compressed_bytes_u16(req->dims);
    q.t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (nb > s->max_blocks) return LSCOMP_ERR_TOO_LARGE;
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
    int st = LSCOMP_OK;
    for (Slot &sl : s->slots) {
        const int e = lscomp_plan_status(sl.plan, sl.stream);
        if (e != LSCOMP_OK) st = e;
    }
    return st;
}

int lscomp_server_replay_u16(lscomp_server s, const lscomp_request_u16 *reqs, const double *arrival_s, uint32_t n,
                             lscomp_completion *out)
{
    if (!s || !reqs || !arrival_s || !out) return LSCOMP_ERR_INVALID_ARG;
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
            lscomp_server_release(s, buf[i].id);
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
    for (Slot &sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.plan) lscomp_plan_destroy(sl.plan);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.h_sizes) (void)hipHostFree(sl.h_sizes);
    }
    if (s->d_arena) (void)hipFree(s->d_arena);
    delete s->arena;
    delete s;
    return LSCOMP_OK;
}

}  // extern "C"
