// lscomp_host.h -- host helpers shared by the plain (lscomp_api.cpp) and the
// batched (lscomp_batch.cpp) entry points.
#pragma once
#include <atomic>
#include <stdint.h>

#include "../../include/lscomp.h"
#include "lscomp_internal.h"

namespace lsc {

extern std::atomic<uint32_t> g_tile[2];    // blocks per tile, per direction
extern std::atomic<uint32_t> g_wgpcu[2];   // workgroups per CU (0: occupancy limit)

int hip_fail(hipError_t e, const char *what);
bool bins_ok(lscomp_bins4 b);
uint32_t align_class(const void *vol, uint32_t z);
uint64_t num_blocks(lscomp_dim3 d);
// Validates dims/bins/pooling and fills the geometry and decision parts of an Item.
int make_item(lscomp_dim3 dims, lscomp_bins4 bins, float p, Item &it, uint32_t tile);
LaunchCfg make_cfg(uint64_t total_tiles, bool compress, bool table, uint32_t tile, uint32_t wg_per_cu);

}  // namespace lsc
