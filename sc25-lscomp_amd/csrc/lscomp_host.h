// lscomp_host.h -- host helpers shared by the plain (lscomp_api.cpp) and the
// batched (lscomp_batch.cpp) entry points.
#pragma once
#include <atomic>
#include <functional>
#include <vector>
#include <stdint.h>

#include "../../include/lscomp.h"
#include "lscomp_internal.h"

namespace lsc {

// Launch shape of a direction (0 compress, 1 decompress).  It is kept packed
// in one atomic word, so a call reads a consistent shape with a single load
// and uses that one copy for its tile count, scratch size and launch.
Shape load_shape(int direction);
void store_shape(int direction, const Shape &s);
// Validates a shape and fills the direction's defaults (tile 0 -> 512,
// window 0 -> default_window, counters 0 -> kDefaultCounters); returns a status.
int normalize_shape(int direction, Shape &s);

int hip_fail(hipError_t e, const char *what);
bool bins_ok(lscomp_bins4 b);
uint32_t align_class(const void *vol, uint32_t z);
uint64_t num_blocks(lscomp_dim3 d);
// Validates dims/bins/pooling and fills the geometry and decision parts of an Item.
int make_item(lscomp_dim3 dims, lscomp_bins4 bins, float p, Item &it, uint32_t tile);
// Stream-ordered copy of a plan's error word (ERRF_*) into pinned host memory,
// then cleared on the device (lscomp_batch.cpp).
int plan_take_error_async(lscomp_plan p, uint32_t *h_err, hipStream_t stream);
// Autotuner (lscomp_tune.cpp): the average ms of `reps` calls of fn on stream
// after one warm-up (negative if a call fails; its status in *status), and a
// coordinate descent over the shape axes from `start` with `measure` (ms of a
// shape, negative on failure), logging every measured shape.
float time_launches(hipStream_t stream, int reps, const std::function<int()> &fn, int *status);
int tune_shape(int direction, bool table, Shape start, const std::function<float(const Shape &)> &measure,
               Shape &best, std::vector<lscomp_tune_entry> &log);
// Grid and scheduling of a launch of `total_tiles` tiles with a normalized shape.
LaunchCfg make_cfg(uint64_t total_tiles, bool compress, bool table, const Shape &shape);
// Validate a single-volume call and fill its Item (tiles of `tile` blocks; no
// HIP calls; returns a status).
int prepare_compress(const uint16_t *d_in, uint8_t *d_out, size_t cap, lscomp_dim3 dims, lscomp_bins4 bins,
                     float p, uint32_t tile, Item &it);
int prepare_decompress(uint16_t *d_out, const uint8_t *d_in, size_t in_bytes, lscomp_dim3 dims,
                       lscomp_bins4 bins, uint32_t tile, Item &it);
// Look-back words carry a 26-bit epoch: cleared at least every 2^25 launches
// that use them, no stale word of an equal epoch survives.
constexpr uint64_t kEpochClearPeriod = 1ull << 25;

}  // namespace lsc
