// lscomp_noalloc.hip -- single-volume compression that allocates nothing.
//
// For callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation): the look-back control
// block and status words live in `scratch`, a file-scope device array of
// kScratchBytes per GPU, held until the process ends.  Host code never passes
// the array's name to a kernel: the name is a host-side stand-in.  Its device
// address is looked up once per device (hipGetSymbolAddress, with that device
// current) and only that address is used.
//
// Every call on one GPU shares that GPU's array and finds it as the last call
// left it: the status words carry the launch epoch (kept in the control block,
// advanced by every launch), so stale words never match, and the host clears
// them every kEpochClearPeriod calls with a fill kernel (no memset node; the
// count is of calls, so a captured call is good for that many replays).  Calls
// that overlap on one GPU (two streams, or two host threads with no ordering
// between their streams) race on the array; the library serializes only the
// host side.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/lscomp.h"
#include "lscomp_internal.h"
#include "lscomp_host.h"

using namespace lsc;

namespace {

constexpr size_t kScratchBytes = LSCOMP_NOALLOC_SCRATCH_BYTES;
constexpr size_t kCtrlBytes = 8192;   // the control block, rounded up to keep status words aligned
static_assert(sizeof(Ctrl) <= kCtrlBytes, "control block");
static_assert(kScratchBytes % 256 == 0, "scratch size");
constexpr uint64_t kMaxTiles = (kScratchBytes - kCtrlBytes) / sizeof(uint64_t);

}  // namespace

static __device__ __align__(256) uint64_t scratch[kScratchBytes / sizeof(uint64_t)];

namespace {

constexpr int kMaxDevices = 64;

struct DeviceScratch {
    std::mutex mu;
    uint8_t *ptr = nullptr;    // device address of this device's `scratch`
    uint64_t launches = 0;     // launches since the status words were cleared
};
DeviceScratch g_dev[kMaxDevices];

// Caller holds d.mu, with device `dev` current.  Looks up the array's address
// on this device and clears it (control block and status words start at zero).
int setup(DeviceScratch &d, hipStream_t stream)
{
    void *p = nullptr;
    int st = hip_fail(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
    if (st) return st;
    if (!p) return LSCOMP_ERR_HIP;
    st = hip_fail(launch_fill_u64((uint64_t *)p, kScratchBytes / sizeof(uint64_t), 0, stream), "fill");
    if (st) return st;
    d.ptr = (uint8_t *)p;
    d.launches = 0;
    return LSCOMP_OK;
}

}  // namespace

extern "C" int lscomp_compress_u16_noalloc_async(const uint16_t *d_in, uint8_t *d_out, size_t out_capacity,
                                                 uint64_t *d_out_bytes, lscomp_dim3 dims, lscomp_bins4 bins,
                                                 float pooling_th, hipStream_t stream)
{
    if (!d_out_bytes || ((uintptr_t)d_out_bytes & 7)) return LSCOMP_ERR_INVALID_ARG;
    const Shape shape = load_shape(0);
    Item it;
    int st = prepare_compress(d_in, d_out, out_capacity, dims, bins, pooling_th, shape.tile, it);
    if (st) return st;
    if (it.ntiles > kMaxTiles) return LSCOMP_ERR_TOO_LARGE;
    it.result = d_out_bytes;
    int dev = 0;
    st = hip_fail(hipGetDevice(&dev), "hipGetDevice");
    if (st) return st;
    if (dev < 0 || dev >= kMaxDevices) return LSCOMP_ERR_INVALID_ARG;
    DeviceScratch &d = g_dev[dev];
    std::lock_guard<std::mutex> lk(d.mu);
    if (!d.ptr) {
        st = setup(d, stream);
        if (st) return st;
    }
    Ctrl *ctrl = (Ctrl *)d.ptr;
    uint64_t *status = (uint64_t *)(d.ptr + kCtrlBytes);
    if (++d.launches >= kEpochClearPeriod) {
        st = hip_fail(launch_fill_u64(status, kMaxTiles, 0, stream), "fill");
        if (st) return st;
        d.launches = 1;
    }
    const LaunchCfg cfg = make_cfg(it.ntiles, true, false, shape);
    return hip_fail(launch_compress(it, nullptr, nullptr, cfg, ctrl, status, stream), "compress launch");
}
