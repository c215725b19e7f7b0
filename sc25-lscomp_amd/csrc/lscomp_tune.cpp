// lscomp_tune.cpp -- autotuner for the launch shape (tile size, workgroups per CU).
//
// The stream never depends on the shape, so tuning only measures: every
// candidate runs `reps` compressions and decompressions of the caller's own
// volume on the caller's stream, timed with hipEvents; the fastest shape per
// direction is applied and remembered for the (size class, mode) key.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/lscomp.h"
#include "lscomp_host.h"

namespace {

struct Best {
    uint32_t tile[2], wgpcu[2];
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

float time_ms(hipStream_t stream, int reps, int (*fn)(void *), void *arg, int *status)
{
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return -1.0f;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return -1.0f;
    }
    *status = fn(arg);   // warm-up
    (void)hipEventRecord(a, stream);
    for (int i = 0; i < reps && *status == LSCOMP_OK; i++) *status = fn(arg);
    (void)hipEventRecord(b, stream);
    (void)hipEventSynchronize(b);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / (float)reps;
}

struct Args {
    const uint16_t *d_in;
    uint8_t *d_stream;
    size_t cap;
    uint16_t *d_dec;
    uint64_t *d_size;
    size_t in_bytes;
    lscomp_dim3 dims;
    lscomp_bins4 bins;
    float p;
    hipStream_t stream;
};

int run_compress(void *v)
{
    Args *a = (Args *)v;
    return lscomp_compress_u16_async(a->d_in, a->d_stream, a->cap, a->d_size, a->dims, a->bins, a->p, a->stream);
}

int run_decompress(void *v)
{
    Args *a = (Args *)v;
    return lscomp_decompress_u16_async(a->d_dec, a->d_stream, a->in_bytes, a->dims, a->bins, a->stream);
}

}  // namespace

extern "C" int lscomp_autotune_u16(const uint16_t *d_in, uint8_t *d_stream, size_t stream_capacity,
                                   uint16_t *d_decoded, lscomp_dim3 dims, lscomp_bins4 bins, float pooling_th,
                                   int reps, hipStream_t stream, lscomp_tune_entry *table, uint32_t *n_table)
{
    if (!d_in || !d_stream || !d_decoded || reps < 1) return LSCOMP_ERR_INVALID_ARG;
    if (stream_capacity < lscomp_max_compressed_bytes_u16(dims)) return LSCOMP_ERR_CAPACITY;
    lsc::Item probe;
    int st = lsc::make_item(dims, bins, pooling_th, probe, 128);
    if (st) return st;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return LSCOMP_ERR_HIP;
    const auto key = std::make_tuple(dev, size_class(lscomp_num_blocks_u16(dims)), mode_of(bins, pooling_th));
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_cache.find(key);
        if (it != g_cache.end()) {
            for (int d = 0; d < 2; d++) lscomp_set_launch_shape(d, it->second.tile[d], it->second.wgpcu[d]);
            if (n_table) *n_table = 0;
            return LSCOMP_OK;
        }
    }
    st = lscomp_reserve_u16(dims, stream);   // allocation outside the timed region
    if (st) return st;
    uint64_t *d_size = nullptr;
    if (hipMalloc((void **)&d_size, sizeof(uint64_t)) != hipSuccess) return LSCOMP_ERR_HIP;
    Args a{d_in, d_stream, stream_capacity, d_decoded, d_size, 0, dims, bins, pooling_th, stream};

    uint32_t saved_tile[2], saved_w[2];
    for (int d = 0; d < 2; d++) lscomp_get_launch_shape(d, &saved_tile[d], &saved_w[d]);
    const uint32_t tiles[3] = {128, 256, 512};
    const uint32_t wgs[5] = {1, 2, 3, 4, 0};
    std::vector<lscomp_tune_entry> res;
    Best best{};
    float best_ms[2] = {1e30f, 1e30f};
    for (uint32_t t : tiles) {
        for (uint32_t w : wgs) {
            lscomp_set_launch_shape(0, t, w);
            lscomp_set_launch_shape(1, t, w);
            lscomp_tune_entry e{t, w, 0.0f, 0.0f};
            e.compress_ms = time_ms(stream, reps, run_compress, &a, &st);
            if (st) break;
            uint64_t n = 0;
            if (hipMemcpyAsync(&n, d_size, sizeof n, hipMemcpyDeviceToHost, stream) != hipSuccess ||
                hipStreamSynchronize(stream) != hipSuccess) {
                st = LSCOMP_ERR_HIP;
                break;
            }
            a.in_bytes = n;
            e.decompress_ms = time_ms(stream, reps, run_decompress, &a, &st);
            if (st) break;
            res.push_back(e);
            for (int d = 0; d < 2; d++) {
                const float ms = d == 0 ? e.compress_ms : e.decompress_ms;
                if (ms > 0.0f && ms < best_ms[d]) {
                    best_ms[d] = ms;
                    best.tile[d] = t;
                    best.wgpcu[d] = w;
                }
            }
        }
        if (st) break;
    }
    (void)hipFree(d_size);
    if (st == LSCOMP_OK) st = lscomp_stream_status(stream);
    if (st != LSCOMP_OK || res.empty()) {
        for (int d = 0; d < 2; d++) lscomp_set_launch_shape(d, saved_tile[d], saved_w[d]);
        return st ? st : LSCOMP_ERR_INTERNAL;
    }
    for (int d = 0; d < 2; d++) lscomp_set_launch_shape(d, best.tile[d], best.wgpcu[d]);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_cache[key] = best;
    }
    if (table && n_table) {
        uint32_t k = 0;
        for (; k < *n_table && k < res.size(); k++) table[k] = res[k];
        *n_table = k;
    }
    return LSCOMP_OK;
}
