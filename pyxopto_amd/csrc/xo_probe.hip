/*
 * xo_probe.hip — the measured ceilings the kernels are compared against:
 * fp32 fused multiply-add throughput (the VALU issue rate of the transport
 * kernels), a streaming device-to-device copy (HBM bandwidth) and
 * random-address u64 atomic adds over a large array (the HBM/L2 atomic
 * rate the fluence and detector deposits run into). No counterpart in the
 * reference. Also the check of the range-restricted correctly rounded
 * square root and division of xo_device.h against the general ones.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>

#include "../../include/xomc.h"
#include "xo_ctx.h"
#include "xo_device.h"
#include "xo_seedhash.h"

namespace {

__global__ void __launch_bounds__(256) copy_kernel(const uint4 *__restrict__ src,
                                                   uint4 *__restrict__ dst, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

/* Thread t adds 1 at (hash(seed, t, k) mod nelem) for k < per_thread, for
 * the first total atomics overall. */
__global__ void __launch_bounds__(256) atomic_kernel(unsigned long long *__restrict__ arr,
                                                     uint64_t nelem, uint64_t total,
                                                     uint32_t per_thread, uint64_t seed) {
    const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t k = 0; k < per_thread; ++k) {
        const uint64_t j = (uint64_t)k * nthreads + t;
        if (j >= total) break;
        const uint64_t h = xo_mix64(seed + j * 0x9E3779B97F4A7C15ull);
        atomicAdd(&arr[h % nelem], 1ull);
    }
}

/* The atomic_kernel with the targets drawn from the first `hot` elements,
 * each wave adding into the copy of the array of the XCD it runs on (the
 * wave's HW_REG_XCC_ID, bits 3:0; eight copies of nelem) with
 * workgroup-scope atomics, which the XCD's L2 performs for all of its
 * CUs. */
__global__ void __launch_bounds__(256) atomic_xcd_kernel(unsigned long long *__restrict__ arr,
                                                         uint64_t nelem, uint64_t hot,
                                                         uint64_t total, uint32_t per_thread,
                                                         uint64_t seed) {
    const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t xcc = __builtin_amdgcn_s_getreg((3 << 11) | 20) & 7u;
    unsigned long long *mine = arr + (uint64_t)xcc * nelem;
    for (uint32_t k = 0; k < per_thread; ++k) {
        const uint64_t j = (uint64_t)k * nthreads + t;
        if (j >= total) break;
        const uint64_t h = xo_mix64(seed + j * 0x9E3779B97F4A7C15ull);
        __hip_atomic_fetch_add(&mine[h % hot], 1ull, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

/* Eight independent fused multiply-add chains per lane, `iters` rounds:
 * 8 iters lane-ops per thread, no memory traffic but the final store. */
__global__ void __launch_bounds__(256) valu_kernel(float *out, uint32_t iters, float a) {
    float v[8];
    for (int k = 0; k < 8; ++k) v[k] = (float)(threadIdx.x + k) * 1e-3f;
    for (uint32_t i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = __builtin_fmaf(v[k], a, 0.5f);
    float s = 0.f;
    for (int k = 0; k < 8; ++k) s += v[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

/* Every float x = +0 or 2^-96 <= x < 2^96: sqrt_in_range(x) against the
 * correctly rounded sqrtf. counts[0] checked, counts[1] mismatches. */
__global__ void __launch_bounds__(256) cr_sqrt_kernel(unsigned long long *counts) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long checked = 0, bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 0x7F800000ull;
         i += stride) {
        const float x = __uint_as_float((uint32_t)i);
        if (!(i == 0 || (x >= 0x1p-96f && x < 0x1p96f))) continue;
        ++checked;
        if (__float_as_uint(xo::sqrt_in_range(x)) != __float_as_uint(__builtin_sqrtf(x))) ++bad;
    }
    atomicAdd(&counts[0], checked);
    if (bad) atomicAdd(&counts[1], bad);
}

/* A float of random sign and mantissa with exponent in [-60, 59]. */
__device__ __forceinline__ float cr_operand(uint64_t h) {
    const uint32_t e = (uint32_t)((h >> 23) % 120u) + 67u;
    return __uint_as_float(((uint32_t)(h >> 40) & 1u) << 31 | e << 23 | ((uint32_t)h & 0x7FFFFFu));
}

/* n random operand pairs: div_in_range(a, b) against a / b; counts[2]
 * mismatches. */
__global__ void __launch_bounds__(256) cr_div_kernel(uint64_t n, uint64_t seed,
                                                     unsigned long long *counts) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float a = cr_operand(xo_mix64(seed + 2 * i + 1));
        const float b = cr_operand(xo_mix64(seed + 2 * i + 2));
        if (__float_as_uint(xo::div_in_range(a, b)) != __float_as_uint(a / b)) ++bad;
    }
    if (bad) atomicAdd(&counts[2], bad);
}

/* out[i] = f(in[i]) for the function `which` of a math policy (the
 * arithmetic of the transport kernels). */
template <class M>
__global__ void __launch_bounds__(256) math_kernel(int which, const float *in, float *out,
                                                   uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = in[i];
    float r = 0.0f, c = 0.0f;
    switch (which) {
    case XO_MATH_LOG: r = M::log(v); break;
    case XO_MATH_SIN_TURNS: M::sincos_turns(v, &r, &c); break;
    case XO_MATH_COS_TURNS: M::sincos_turns(v, &c, &r); break;
    case XO_MATH_SQRT: r = M::sqrt(v); break;
    case XO_MATH_RCP: r = M::div(1.0f, v); break;
    case XO_MATH_RSQRT: r = M::rsqrt_r(v); break;
    case XO_MATH_CBRT: r = M::cbrt(v); break;
    case XO_MATH_EXP: r = M::exp(v); break;
    default: break;
    }
    out[i] = r;
}

} // namespace

extern "C" {

int xo_probe_valu(xo_ctx *ctx, uint32_t iters, double *ms, double *lane_ops) {
    xo::clear_error();
    if (!ctx || !ms || !lane_ops || iters == 0) {
        xo::set_error("xo_probe_valu: invalid argument");
        return -1;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    XO_HIP(hipSetDevice(ctx->device));
    const uint32_t blocks = ctx->cus * 32;
    float *out = nullptr;
    XO_HIP(hipMalloc(&out, sizeof(float) * blocks * 256));
    hipLaunchKernelGGL(valu_kernel, dim3(blocks), dim3(256), 0, ctx->stream, out, iters / 8 + 1,
                       0.999f);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(valu_kernel, dim3(blocks), dim3(256), 0, ctx->stream, out, iters,
                           0.999f);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    float kms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1);
    (void)hipFree(out);
    if (e != hipSuccess) {
        xo::set_error("xo_probe_valu: %s", hipGetErrorString(e));
        return -1;
    }
    *ms = kms;
    *lane_ops = 8.0 * (double)iters * blocks * 256.0;
    return 0;
}

int xo_probe_copy(xo_ctx *ctx, uint64_t bytes, uint32_t reps, double *ms_per_copy) {
    xo::clear_error();
    if (!ctx || !ms_per_copy || bytes < 16 || reps == 0) {
        xo::set_error("xo_probe_copy: invalid argument");
        return -1;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    XO_HIP(hipSetDevice(ctx->device));
    const uint64_t n = bytes / 16;
    void *src = nullptr, *dst = nullptr;
    XO_HIP(hipMalloc(&src, n * 16));
    if (hipMalloc(&dst, n * 16) != hipSuccess) {
        (void)hipFree(src);
        xo::set_error("xo_probe_copy: out of device memory");
        return -1;
    }
    int rc = 0;
    const uint32_t blocks = ctx->cus * 8;
    hipError_t e = hipMemsetAsync(src, 1, n * 16, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           (const uint4 *)src, (uint4 *)dst, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
    for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) {
        hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           (const uint4 *)src, (uint4 *)dst, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    if (e != hipSuccess) {
        xo::set_error("xo_probe_copy: %s", hipGetErrorString(e));
        rc = -1;
    }
    *ms_per_copy = (double)ms / reps;
    (void)hipFree(src);
    (void)hipFree(dst);
    return rc;
}

/* natomics random u64 adds into a zeroed array of array_bytes (eight of
 * them, one per XCD, when per_xcd), the targets among the first hot_bytes:
 * kernel ms and the sum of everything allocated. */
static int atomic_probe(xo_ctx *ctx, const char *fn, uint64_t array_bytes, uint64_t hot_bytes,
                        uint64_t natomics, uint64_t seed, int per_xcd, double *ms,
                        uint64_t *checksum) {
    xo::clear_error();
    if (!ctx || !ms || !checksum || array_bytes < 8 || hot_bytes < 8 ||
        hot_bytes > array_bytes || natomics == 0) {
        xo::set_error("%s: invalid argument", fn);
        return -1;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    XO_HIP(hipSetDevice(ctx->device));
    const uint64_t nelem = array_bytes / 8, hot = hot_bytes / 8;
    const uint64_t ncopies = per_xcd ? 8u : 1u;
    const uint64_t nall = nelem * ncopies;
    unsigned long long *arr = nullptr;
    XO_HIP(hipMalloc(&arr, nall * 8));
    const uint32_t blocks = ctx->cus * 8;
    const uint64_t nthreads = (uint64_t)blocks * 256;
    const uint64_t per = (natomics + nthreads - 1) / nthreads;
    int rc = 0;
    hipError_t e = hipMemsetAsync(arr, 0, nall * 8, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
    if (e == hipSuccess) {
        if (per_xcd)
            hipLaunchKernelGGL(atomic_xcd_kernel, dim3(blocks), dim3(256), 0, ctx->stream, arr,
                               nelem, hot, natomics, (uint32_t)per, seed);
        else
            hipLaunchKernelGGL(atomic_kernel, dim3(blocks), dim3(256), 0, ctx->stream, arr,
                               hot, natomics, (uint32_t)per, seed);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    float kms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1);
    /* every atomic landed: everything sums to natomics */
    uint64_t sum = 0;
    if (e == hipSuccess) {
        const uint64_t chunk = 1u << 20;
        uint64_t *host = new uint64_t[chunk];
        for (uint64_t off = 0; off < nall && e == hipSuccess; off += chunk) {
            uint64_t m = nall - off < chunk ? nall - off : chunk;
            e = hipMemcpy(host, arr + off, m * 8, hipMemcpyDeviceToHost);
            for (uint64_t i = 0; i < m && e == hipSuccess; ++i) sum += host[i];
        }
        delete[] host;
    }
    if (e != hipSuccess) {
        xo::set_error("%s: %s", fn, hipGetErrorString(e));
        rc = -1;
    }
    *ms = kms;
    *checksum = sum;
    (void)hipFree(arr);
    return rc;
}

int xo_probe_atomic_u64(xo_ctx *ctx, uint64_t array_bytes, uint64_t natomics, uint64_t seed,
                        double *ms, uint64_t *checksum) {
    return atomic_probe(ctx, "xo_probe_atomic_u64", array_bytes, array_bytes, natomics, seed,
                        0, ms, checksum);
}

int xo_probe_atomic_u64_xcd(xo_ctx *ctx, uint64_t array_bytes, uint64_t hot_bytes,
                            uint64_t natomics, uint64_t seed, int per_xcd, double *ms,
                            uint64_t *checksum) {
    return atomic_probe(ctx, "xo_probe_atomic_u64_xcd", array_bytes, hot_bytes, natomics, seed,
                        per_xcd, ms, checksum);
}

int xo_probe_cr_math(xo_ctx *ctx, uint64_t ndiv, uint64_t seed, uint64_t *sqrt_checked,
                     uint64_t *sqrt_mismatches, uint64_t *div_mismatches) {
    xo::clear_error();
    if (!ctx || !sqrt_checked || !sqrt_mismatches || !div_mismatches) {
        xo::set_error("xo_probe_cr_math: null argument");
        return -1;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    XO_HIP(hipSetDevice(ctx->device));
    unsigned long long *counts = nullptr;
    XO_HIP(hipMalloc(&counts, 3 * sizeof(unsigned long long)));
    unsigned long long host[3] = {0, 0, 0};
    hipError_t e = hipMemsetAsync(counts, 0, sizeof(host), ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(cr_sqrt_kernel, dim3(ctx->cus * 32), dim3(256), 0, ctx->stream, counts);
        e = hipGetLastError();
    }
    if (e == hipSuccess && ndiv) {
        hipLaunchKernelGGL(cr_div_kernel, dim3(ctx->cus * 32), dim3(256), 0, ctx->stream, ndiv,
                           seed, counts);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(host, counts, sizeof(host), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(counts);
    if (e != hipSuccess) {
        xo::set_error("xo_probe_cr_math: %s", hipGetErrorString(e));
        return -1;
    }
    *sqrt_checked = host[0];
    *sqrt_mismatches = host[1];
    *div_mismatches = host[2];
    return 0;
}

int xo_probe_math(xo_ctx *ctx, int fast, int which, uint32_t n, const float *in, float *out) {
    xo::clear_error();
    if (!ctx || !in || !out || which < 0 || which >= XO_MATH_COUNT) {
        xo::set_error("xo_probe_math: invalid argument");
        return -1;
    }
    if (n == 0) return 0;
    std::lock_guard<std::mutex> lock(ctx->mu);
    XO_HIP(hipSetDevice(ctx->device));
    float *buf = nullptr;
    XO_HIP(hipMalloc(&buf, 2 * (size_t)n * sizeof(float)));
    hipError_t e = hipMemcpyAsync(buf, in, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        const dim3 grid((n + 255u) / 256u);
        if (fast)
            hipLaunchKernelGGL(math_kernel<xo::MathFast>, grid, dim3(256), 0, ctx->stream, which,
                               buf, buf + n, n);
        else
            hipLaunchKernelGGL(math_kernel<xo::MathExact>, grid, dim3(256), 0, ctx->stream,
                               which, buf, buf + n, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, buf + n, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) {
        xo::set_error("xo_probe_math: %s", hipGetErrorString(e));
        return -1;
    }
    return 0;
}

} // extern "C"
