// lscomp_tune.cpp -- autotuner for the launch shape.
//
// The stream never depends on the shape, so tuning only measures.  Four axes:
// tile (128 / 256 / 512 blocks), workgroups per CU (each value that gives a
// distinct grid: requests at or above the occupancy limit are the same launch,
// so they are one candidate), look-back window (64 .. tile words) and tile-id
// counters (1 .. 16).  Coordinate descent from the current shape: each axis
// in turn, every value of it with the others held, keeping a value only when
// it is at least 1 % faster.  Single volumes tune both directions on the
// caller's volume (lscomp_autotune_u16, cached per size class and mode);
// plans tune on their own batch (lscomp_plan_autotune, lscomp_batch.cpp).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/lscomp.h"
#include "lscomp_host.h"

namespace lsc {

float time_launches(hipStream_t stream, int reps, const std::function<int()> &fn, int *status)
{
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return -1.0f;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return -1.0f;
    }
    *status = fn();   // warm-up
    (void)hipEventRecord(a, stream);
    for (int i = 0; i < reps && *status == LSCOMP_OK; i++) *status = fn();
    (void)hipEventRecord(b, stream);
    (void)hipEventSynchronize(b);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return *status == LSCOMP_OK ? ms / (float)reps : -1.0f;
}

namespace {

std::vector<Shape> axis_values(int axis, const Shape &base, int direction, bool table)
{
    std::vector<Shape> out;
    auto with = [&](Shape s) {
        if (s.win > s.tile) s.win = s.tile;
        out.push_back(s);
    };
    switch (axis) {
    case 0:
        for (uint32_t t : {128u, 256u, 512u}) {
            Shape s = base;
            s.tile = t;
            with(s);
        }
        break;
    case 1: {
        const int occ = workgroups_per_cu(direction == 0, table, (int)base.tile);
        Shape s = base;
        s.wgpcu = 0;   // the occupancy limit
        with(s);
        for (int w = 1; w < occ && w <= 4; w++) {
            s.wgpcu = (uint32_t)w;
            with(s);
        }
        break;
    }
    case 2:
        for (uint32_t w : {64u, 128u, 256u, 512u}) {
            if (w > base.tile) break;
            Shape s = base;
            s.win = w;
            with(s);
        }
        break;
    default:
        for (uint32_t c : {1u, 2u, 4u, 8u, 16u}) {
            Shape s = base;
            s.nc = c;
            with(s);
        }
        break;
    }
    return out;
}

bool same_shape(const Shape &a, const Shape &b)
{
    return a.tile == b.tile && a.wgpcu == b.wgpcu && a.win == b.win && a.nc == b.nc;
}

}  // namespace

int tune_shape(int direction, bool table, Shape start, const std::function<float(const Shape &)> &measure,
               Shape &best, std::vector<lscomp_tune_entry> &log)
{
    int st = normalize_shape(direction, start);
    if (st) return st;
    std::vector<std::pair<Shape, float>> seen;
    auto run = [&](const Shape &s) -> float {
        for (const auto &e : seen)
            if (same_shape(e.first, s)) return e.second;
        const float ms = measure(s);
        seen.push_back({s, ms});
        if (ms > 0.0f) log.push_back({(uint32_t)direction, s.tile, s.wgpcu, s.win, s.nc, ms});
        return ms;
    };
    best = start;
    float best_ms = run(best);
    if (best_ms <= 0.0f) return LSCOMP_ERR_INTERNAL;
    for (int axis = 0; axis < 4; axis++) {
        for (const Shape &s : axis_values(axis, best, direction, table)) {
            const float ms = run(s);
            if (ms > 0.0f && ms < 0.99f * best_ms) {
                best = s;
                best_ms = ms;
            }
        }
    }
    return LSCOMP_OK;
}

}  // namespace lsc

namespace {

struct Best {
    lsc::Shape s[2];
};

std::mutex g_mu;
std::map<std::tuple<int, int, int>, Best> g_cache;   // (device, log2 blocks, mode) -> shapes

int size_class(uint64_t blocks)
{
    int c = 0;
    while ((1ull << (c + 1)) <= blocks) c++;
    return c;
}

int mode_of(lscomp_bins4 b, float p)
{
    if (b.x == 1 && b.y == 1 && b.z == 1 && b.w == 1 && p >= 1.0f) return 0;   // lossless
    return 1;
}

}  // namespace

extern "C" int lscomp_autotune_u16(const uint16_t *d_in, uint8_t *d_stream, size_t stream_capacity,
                                   uint16_t *d_decoded, lscomp_dim3 dims, lscomp_bins4 bins, float pooling_th,
                                   int reps, hipStream_t stream, lscomp_tune_entry *table, uint32_t *n_table)
{
    using namespace lsc;
    if (!d_in || !d_stream || !d_decoded || reps < 1) return LSCOMP_ERR_INVALID_ARG;
    if (stream_capacity < lscomp_max_compressed_bytes_u16(dims)) return LSCOMP_ERR_CAPACITY;
    Item probe;
    int st = make_item(dims, bins, pooling_th, probe, 128);
    if (st) return st;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return LSCOMP_ERR_HIP;
    const auto key = std::make_tuple(dev, size_class(lscomp_num_blocks_u16(dims)), mode_of(bins, pooling_th));
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_cache.find(key);
        if (it != g_cache.end()) {
            for (int d = 0; d < 2; d++) store_shape(d, it->second.s[d]);
            if (n_table) *n_table = 0;
            return LSCOMP_OK;
        }
    }
    st = lscomp_reserve_u16(dims, stream);   // allocation outside the timed region
    if (st) return st;
    uint64_t *d_size = nullptr;
    if (hipMalloc((void **)&d_size, sizeof(uint64_t)) != hipSuccess) return LSCOMP_ERR_HIP;
    const Shape saved[2] = {load_shape(0), load_shape(1)};
    std::vector<lscomp_tune_entry> log;
    Best best{};
    // compression on the caller's volume
    auto measure_c = [&](const Shape &s) -> float {
        store_shape(0, s);
        int rs = LSCOMP_OK;
        const float ms = time_launches(stream, reps, [&] {
            return lscomp_compress_u16_async(d_in, d_stream, stream_capacity, d_size, dims, bins, pooling_th, stream);
        }, &rs);
        return rs == LSCOMP_OK ? ms : -1.0f;
    };
    st = tune_shape(0, false, saved[0], measure_c, best.s[0], log);
    uint64_t n = 0;
    if (!st) {
        store_shape(0, best.s[0]);
        st = lscomp_compress_u16_async(d_in, d_stream, stream_capacity, d_size, dims, bins, pooling_th, stream);
        if (!st && (hipMemcpyAsync(&n, d_size, sizeof n, hipMemcpyDeviceToHost, stream) != hipSuccess ||
                    hipStreamSynchronize(stream) != hipSuccess))
            st = LSCOMP_ERR_HIP;
    }
    // decompression of that stream
    auto measure_d = [&](const Shape &s) -> float {
        store_shape(1, s);
        int rs = LSCOMP_OK;
        const float ms = time_launches(stream, reps, [&] {
            return lscomp_decompress_u16_async(d_decoded, d_stream, n, dims, bins, stream);
        }, &rs);
        return rs == LSCOMP_OK ? ms : -1.0f;
    };
    if (!st) st = tune_shape(1, false, saved[1], measure_d, best.s[1], log);
    (void)hipFree(d_size);
    if (st == LSCOMP_OK) st = lscomp_stream_status(stream);
    if (st != LSCOMP_OK) {
        for (int d = 0; d < 2; d++) store_shape(d, saved[d]);
        return st;
    }
    for (int d = 0; d < 2; d++) store_shape(d, best.s[d]);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_cache[key] = best;
    }
    if (table && n_table) {
        uint32_t k = 0;
        for (; k < *n_table && k < log.size(); k++) table[k] = log[k];
        *n_table = k;
    }
    return LSCOMP_OK;
}
