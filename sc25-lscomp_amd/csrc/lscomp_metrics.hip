// lscomp_metrics.hip -- error metrics between an original and a decoded volume.
//
// Definitions follow the reference's computePSNR (lsCOMP/src/cuLSZ_utility.cu:887-946):
// mse = sum(err^2)/n, range = max(orig) - min(orig), psnr = 20 log10(range)
// - 10 log10(mse), normErr = sqrt(sum err^2), normErr_norm = normErr /
// sqrt(sum dec^2), nrmse = sqrt(mse)/range; plus the maximum absolute error
// and the maximum pointwise relative error over nonzero originals.  Sums are
// exact 64-bit integers (uint16 data), reduced per wave and then atomically.
// One pass, HBM-bound: 4 bytes read per element pair.  Also the mean SSIM
// of the reference's computeSSIM (cuLSZ_utility.cu:757-886), and the plain
// copy used to measure copy bandwidth and to calibrate the HBM byte counters.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

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

// ---------------------------------------------------------------------------
// SSIM (cuLSZ_utility.cu:757-886): the mean over 7x7x7 windows moved by 2
// along every axis of luminance * contrast * structure, with c1 = (0.01 R)^2,
// c2 = (0.03 R)^2 (R: the original's range in the window; 0.01^2 and 0.03^2
// when R = 0) and c3 = c2 / 2.
//
// Separable window sums instead of the reference's two passes over every
// window's 343 points.  A workgroup owns 64 x 4 windows along (size0, size1)
// and up to 64 windows along size2, and walks that slab plane by plane: each
// plane's 13 x 133 patch of both volumes goes to LDS; row sums over 7 columns
// (at column stride 2) for 13 rows x 64 windows; each thread sums 7 rows of
// its window into the plane's moments P; along size2 a window is the sum of
// three plane pairs and one plane (planes 2k .. 2k+6), kept in registers.
// Moments are exact: integer sums of values, double sums of squares and
// products (integers below 2^53), so the variances and the covariance are
// n*S_xx - S_x^2 (exact 64-bit integers) over n^2, correctly rounded, where the
// reference rounds a two-pass double sum.  Each workgroup writes one partial
// sum, reduced in a fixed order; the host adds the partials in order, so the
// result does not depend on scheduling.
// ---------------------------------------------------------------------------
constexpr uint32_t kSsimT0 = 64, kSsimT1 = 4, kSsimChunk = 64;   // windows per workgroup along size0, size1, size2
constexpr uint32_t kSsimPC = 2 * kSsimT0 + 5, kSsimPR = 2 * kSsimT1 + 5;   // patch columns (133) and rows (13)
constexpr uint32_t kSsimPitch = 136;                                       // u16 per patch row (even, >= 134)

struct SsimMom {
    uint32_t sx, sy, mn, mx;   // sums of both volumes, min and max of the original
    double sxx, syy, sxy;      // exact integer sums
};

__device__ __forceinline__ SsimMom mom_add(const SsimMom &a, const SsimMom &b)
{
    return {a.sx + b.sx, a.sy + b.sy, a.mn < b.mn ? a.mn : b.mn, a.mx > b.mx ? a.mx : b.mx,
            a.sxx + b.sxx, a.syy + b.syy, a.sxy + b.sxy};
}

__device__ double ssim_of(const SsimMom &m)
{
#pragma clang fp contract(off)
    const double np = 343.0;
    const double x_mean = (double)m.sx / np, y_mean = (double)m.sy / np;
    const int64_t sx = m.sx, sy = m.sy;
    const double n2 = np * np;
    const double var_x = (double)(343 * (int64_t)m.sxx - sx * sx) / n2;
    const double var_y = (double)(343 * (int64_t)m.syy - sy * sy) / n2;
    const double cov = (double)(343 * (int64_t)m.sxy - sx * sy) / n2;
    const double x_sigma = sqrt(var_x), y_sigma = sqrt(var_y);
    const float range = (float)(m.mx - m.mn);
    double c1, c2;
    if (range == 0.0f) {
        c1 = 0.01 * 0.01;
        c2 = 0.03 * 0.03;
    } else {
        c1 = 0.01 * 0.01 * range * range;
        c2 = 0.03 * 0.03 * range * range;
    }
    const double c3 = c2 / 2;
    const double luminance = (2 * x_mean * y_mean + c1) / (x_mean * x_mean + y_mean * y_mean + c1);
    const double contrast = (2 * x_sigma * y_sigma + c2) / (x_sigma * x_sigma + y_sigma * y_sigma + c2);
    const double structure = (cov + c3) / (x_sigma * y_sigma + c3);
    return luminance * contrast * structure;
}

// One workgroup per (size0 tile, size1 tile, size2 chunk) of windows; grid
// nb0 * nb1 * nchunks workgroups of 256 threads; partial[blockIdx.x] = the
// sum of the SSIM of its windows.
__global__ __launch_bounds__(256) void ssim_kernel(const uint16_t *__restrict__ ori, const uint16_t *__restrict__ dec,
                                                   uint32_t X, uint32_t Y, uint32_t Z, uint32_t n0, uint32_t n1,
                                                   uint32_t n2, uint32_t nb0, uint32_t nb1, double *partial)
{
    __shared__ __align__(16) uint16_t xs[kSsimPR][kSsimPitch], ys[kSsimPR][kSsimPitch];
    __shared__ uint32_t r_sx[kSsimPR][kSsimT0], r_sy[kSsimPR][kSsimT0], r_mn[kSsimPR][kSsimT0],
        r_mx[kSsimPR][kSsimT0];
    __shared__ double r_sxx[kSsimPR][kSsimT0], r_syy[kSsimPR][kSsimT0], r_sxy[kSsimPR][kSsimT0];
    __shared__ double wsum[4];
    const uint32_t tid = threadIdx.x, t0 = tid & 63u, t1 = tid >> 6;
    uint32_t blk = blockIdx.x;
    const uint32_t b0 = blk % nb0;
    blk /= nb0;
    const uint32_t b1 = blk % nb1, c2 = blk / nb1;
    const uint32_t o0b = b0 * kSsimT0, o1b = b1 * kSsimT1, o2b = c2 * kSsimChunk;
    const uint32_t o2e = o2b + kSsimChunk < n2 ? o2b + kSsimChunk : n2;
    const uint64_t col0 = 2ull * o0b, row0 = 2ull * o1b;
    const bool mine = o0b + t0 < n0 && o1b + t1 < n1;
    // planes [2 o2b, 2 (o2e - 1) + 7): the last is at most X - 1 since o2e <= n2
    const uint32_t pb = 2u * o2b, pe = 2u * (o2e - 1u) + 7u;
    (void)X;

    SsimMom q0{}, q1{}, q2{}, pev{};   // the last three plane pairs, the last even plane
    double acc = 0.0;
    for (uint32_t p = pb; p < pe; p++) {
        // 1. the plane's patch (zeros outside the volume: only windows that
        //    are not counted read them)
        for (uint32_t e = tid; e < kSsimPR * kSsimPC; e += 256u) {
            const uint32_t r = e / kSsimPC, c = e - r * kSsimPC;
            const uint64_t gy = row0 + r, gz = col0 + c;
            uint16_t a = 0, b = 0;
            if (gy < Y && gz < Z) {
                const uint64_t i = ((uint64_t)p * Y + gy) * Z + gz;
                a = ori[i];
                b = dec[i];
            }
            xs[r][c] = a;
            ys[r][c] = b;
        }
        __syncthreads();
        // 2. row sums: 7 columns from column 2w, read as four aligned pairs
        for (uint32_t e = tid; e < kSsimPR * kSsimT0; e += 256u) {
            const uint32_t r = e >> 6, w = e & 63u;
            const uint32_t *xr = (const uint32_t *)&xs[r][0] + w, *yr = (const uint32_t *)&ys[r][0] + w;
            uint32_t sx = 0, sy = 0, mn = 0xffffu, mx = 0;
            double sxx = 0, syy = 0, sxy = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t xw = xr[k], yw = yr[k];
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    if (k == 3 && h == 1) break;   // 7 columns
                    const uint32_t a = h ? xw >> 16 : xw & 0xffffu, b = h ? yw >> 16 : yw & 0xffffu;
                    sx += a;
                    sy += b;
                    mn = a < mn ? a : mn;
                    mx = a > mx ? a : mx;
                    const double da = (double)a, db = (double)b;
                    sxx = fma(da, da, sxx);
                    syy = fma(db, db, syy);
                    sxy = fma(da, db, sxy);
                }
            }
            r_sx[r][w] = sx;
            r_sy[r][w] = sy;
            r_mn[r][w] = mn;
            r_mx[r][w] = mx;
            r_sxx[r][w] = sxx;
            r_syy[r][w] = syy;
            r_sxy[r][w] = sxy;
        }
        __syncthreads();
        // 3. this thread's window in the plane: rows 2 t1 .. 2 t1 + 6
        SsimMom P{0, 0, 0xffffu, 0, 0.0, 0.0, 0.0};
#pragma unroll
        for (uint32_t k = 0; k < 7; k++) {
            const uint32_t r = 2u * t1 + k;
            P = mom_add(P, SsimMom{r_sx[r][t0], r_sy[r][t0], r_mn[r][t0], r_mx[r][t0], r_sxx[r][t0], r_syy[r][t0],
                                   r_sxy[r][t0]});
        }
        // 4. along size2: window k covers pairs k, k+1, k+2 and plane 2k + 6
        const uint32_t j = p - pb;
        if ((j & 1u) == 0) {
            if (j >= 6 && mine) acc += ssim_of(mom_add(mom_add(q0, q1), mom_add(q2, P)));
            pev = P;
        } else {
            q0 = q1;
            q1 = q2;
            q2 = mom_add(pev, P);
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((tid & 63u) == 0) wsum[t1] = acc;
    __syncthreads();
    if (tid == 0) partial[blockIdx.x] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
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

extern "C" int lscomp_ssim_u16(const uint16_t *d_orig, const uint16_t *d_dec, lscomp_dim3 dims, double *ssim,
                               hipStream_t stream)
{
    using namespace lsc;
    if (!d_orig || !d_dec || !ssim || (((uintptr_t)d_orig | (uintptr_t)d_dec) & 1)) return LSCOMP_ERR_INVALID_ARG;
    if (dims.x < 7 || dims.y < 7 || dims.z < 7) return LSCOMP_ERR_INVALID_ARG;   // no window fits
    const uint32_t n0 = (dims.z - 7) / 2 + 1, n1 = (dims.y - 7) / 2 + 1, n2 = (dims.x - 7) / 2 + 1;
    const uint32_t nb0 = (n0 + kSsimT0 - 1) / kSsimT0, nb1 = (n1 + kSsimT1 - 1) / kSsimT1;
    const uint32_t nc = (n2 + kSsimChunk - 1) / kSsimChunk;
    const uint64_t blocks = (uint64_t)nb0 * nb1 * nc;
    if (blocks > 0x7fffffffull) return LSCOMP_ERR_TOO_LARGE;
    double *partial = nullptr;
    int st = hip_fail(hipMalloc((void **)&partial, blocks * sizeof(double)), "hipMalloc");
    if (st) return st;
    std::vector<double> host(blocks);
    hipLaunchKernelGGL(ssim_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, d_orig, d_dec, dims.x, dims.y,
                       dims.z, n0, n1, n2, nb0, nb1, partial);
    st = hip_fail(hipGetLastError(), "ssim launch");
    if (!st)
        st = hip_fail(hipMemcpyAsync(host.data(), partial, blocks * sizeof(double), hipMemcpyDeviceToHost, stream),
                      "hipMemcpyAsync");
    if (!st) st = hip_fail(hipStreamSynchronize(stream), "hipStreamSynchronize");
    (void)hipFree(partial);
    if (st) return st;
    double sum = 0.0;
    for (double v : host) sum += v;
    *ssim = sum / ((double)n0 * (double)n1 * (double)n2);
    return LSCOMP_OK;
}
