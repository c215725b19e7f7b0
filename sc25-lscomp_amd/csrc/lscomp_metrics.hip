// lscomp_metrics.hip -- error metrics between an original and a decoded volume.
//
// Definitions follow the reference's computePSNR (lsCOMP/src/cuLSZ_utility.cu:887-946):
// mse = sum(err^2)/n, range = max(orig) - min(orig), psnr = 20 log10(range)
// - 10 log10(mse), normErr = sqrt(sum err^2), normErr_norm = normErr /
// sqrt(sum dec^2), nrmse = sqrt(mse)/range; plus the maximum absolute error
// and the maximum pointwise relative error over nonzero originals.  Sums are
// exact 64-bit integers (uint16 data), reduced per wave and then atomically.
// One pass, HBM-bound: 4 bytes read per element pair.  Also the plain copy
// used to measure copy bandwidth and to calibrate the HBM byte counters.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "../../include/lscomp.h"
#include "lscomp_host.h"

namespace lsc {

struct MetricAcc {
    unsigned long long sum_sq_err, sum_sq_dec;
    uint32_t max_err, min_orig, max_orig, max_rel_bits;   // max_rel as float bits (non-negative)
};

__global__ void metrics_init_kernel(MetricAcc *acc)
{
    acc->sum_sq_err = 0;
    acc->sum_sq_dec = 0;
    acc->max_err = 0;
    acc->min_orig = 0xffffu;
    acc->max_orig = 0;
    acc->max_rel_bits = 0;
}

__device__ __forceinline__ void acc_pair(uint32_t a, uint32_t d, uint64_t &se, uint64_t &sd, uint32_t &me,
                                         uint32_t &mn, uint32_t &mx, float &mr)
{
    const uint32_t e = a > d ? a - d : d - a;
    se += (uint64_t)e * e;
    sd += (uint64_t)d * d;
    me = e > me ? e : me;
    mn = a < mn ? a : mn;
    mx = a > mx ? a : mx;
    if (a) {
        const float r = (float)e / (float)a;
        mr = r > mr ? r : mr;
    }
}

__global__ __launch_bounds__(256) void metrics_kernel(const uint16_t *__restrict__ orig,
                                                      const uint16_t *__restrict__ dec, uint64_t n,
                                                      MetricAcc *acc)
{
    uint64_t se = 0, sd = 0;
    uint32_t me = 0, mn = 0xffffu, mx = 0;
    float mr = 0.0f;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n8 = n / 8;
    const bool vec = (((uintptr_t)orig | (uintptr_t)dec) & 15) == 0;
    uint64_t tail = 0;
    if (vec) {
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n8; i += stride) {
            const uint4 a = ((const uint4 *)orig)[i], d = ((const uint4 *)dec)[i];
            const uint32_t av[4] = {a.x, a.y, a.z, a.w}, dv[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                acc_pair(av[k] & 0xffffu, dv[k] & 0xffffu, se, sd, me, mn, mx, mr);
                acc_pair(av[k] >> 16, dv[k] >> 16, se, sd, me, mn, mx, mr);
            }
        }
        tail = n8 * 8;
    }
    for (uint64_t i = tail + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        acc_pair(orig[i], dec[i], se, sd, me, mn, mx, mr);
    // wave reduction, then one atomic per wave
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        se += __shfl_xor(se, o, 64);
        sd += __shfl_xor(sd, o, 64);
        const uint32_t me2 = __shfl_xor(me, o, 64), mn2 = __shfl_xor(mn, o, 64), mx2 = __shfl_xor(mx, o, 64);
        const float mr2 = __shfl_xor(mr, o, 64);
        me = me2 > me ? me2 : me;
        mn = mn2 < mn ? mn2 : mn;
        mx = mx2 > mx ? mx2 : mx;
        mr = mr2 > mr ? mr2 : mr;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&acc->sum_sq_err, (unsigned long long)se);
        atomicAdd(&acc->sum_sq_dec, (unsigned long long)sd);
        atomicMax(&acc->max_err, me);
        atomicMin(&acc->min_orig, mn);
        atomicMax(&acc->max_orig, mx);
        atomicMax(&acc->max_rel_bits, __float_as_uint(mr));   // non-negative floats order like their bits
    }
}

// Plain copy: each workgroup copies one contiguous 16 KiB chunk, four 16-byte
// loads per lane issued before the four stores.
__global__ __launch_bounds__(256) void copy_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src,
                                                   uint64_t n16)
{
    const uint64_t base = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
    uint4 v[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (base + 256 * k < n16) v[k] = src[base + 256 * k];
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (base + 256 * k < n16) dst[base + 256 * k] = v[k];
}

__global__ void copy_tail_kernel(uint8_t *dst, const uint8_t *src, uint32_t n)
{
    if (threadIdx.x < n) dst[threadIdx.x] = src[threadIdx.x];
}

}  // namespace lsc

extern "C" int lscomp_copy_bytes(void *d_dst, const void *d_src, size_t bytes, hipStream_t stream)
{
    using namespace lsc;
    if (bytes == 0) return LSCOMP_OK;
    if (!d_dst || !d_src || (((uintptr_t)d_dst | (uintptr_t)d_src) & 15)) return LSCOMP_ERR_INVALID_ARG;
    const uint8_t *s = (const uint8_t *)d_src;
    uint8_t *d = (uint8_t *)d_dst;
    if (d < s + bytes && s < d + bytes) return LSCOMP_ERR_INVALID_ARG;   // overlap
    const uint64_t n16 = bytes / 16;
    if (n16) {
        const uint64_t blocks = (n16 + 1023) / 1024;   // 16 KiB per workgroup
        if (blocks > 0xffffffffull) return LSCOMP_ERR_INVALID_ARG;
        hipLaunchKernelGGL(copy_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, (uint4 *)d,
                           (const uint4 *)s, n16);
    }
    const uint32_t tail = (uint32_t)(bytes - 16 * n16);
    if (tail)
        hipLaunchKernelGGL(copy_tail_kernel, dim3(1), dim3(16), 0, stream, d + 16 * n16, s + 16 * n16, tail);
    return hip_fail(hipGetLastError(), "copy launch");
}

extern "C" int lscomp_error_stats_u16(const uint16_t *d_orig, const uint16_t *d_dec, size_t n,
                                      lscomp_error_stats *out, hipStream_t stream)
{
    using namespace lsc;
    if (!d_orig || !d_dec || !out || n == 0 || (((uintptr_t)d_orig | (uintptr_t)d_dec) & 1))
        return LSCOMP_ERR_INVALID_ARG;
    MetricAcc *acc = nullptr, host{};
    int st = hip_fail(hipMalloc((void **)&acc, sizeof(MetricAcc)), "hipMalloc");
    if (st) return st;
    hipLaunchKernelGGL(metrics_init_kernel, dim3(1), dim3(1), 0, stream, acc);
    uint64_t blocks = (n / 8 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(metrics_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, d_orig, d_dec, (uint64_t)n, acc);
    st = hip_fail(hipGetLastError(), "metrics launch");
    if (!st) st = hip_fail(hipMemcpyAsync(&host, acc, sizeof host, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
    if (!st) st = hip_fail(hipStreamSynchronize(stream), "hipStreamSynchronize");
    (void)hipFree(acc);
    if (st) return st;
    std::memset(out, 0, sizeof *out);
    out->n = n;
    out->max_abs_err = host.max_err;
    out->min_orig = host.min_orig;
    out->max_orig = host.max_orig;
    out->sum_sq_err = host.sum_sq_err;
    out->sum_sq_dec = host.sum_sq_dec;
    float mr;
    std::memcpy(&mr, &host.max_rel_bits, sizeof mr);
    out->max_rel_err = mr;
    out->mse = (double)host.sum_sq_err / (double)n;
    out->range = (double)host.max_orig - (double)host.min_orig;
    out->psnr = 20.0 * std::log10(out->range) - 10.0 * std::log10(out->mse);   // inf when lossless
    out->norm_err = std::sqrt((double)host.sum_sq_err);
    out->norm_err_norm = out->norm_err / std::sqrt((double)host.sum_sq_dec);
    out->nrmse = std::sqrt(out->mse) / out->range;
    return LSCOMP_OK;
}
