/*
 * tp_peak.hip -- microbenchmarks of the device's achievable peaks
 * (include/tp_peak.h): HBM copy and read sweeps, FP64 and FP32 FMA chains,
 * and a bf16 MFMA loop.  Each kernel is register-only apart from its
 * streaming buffers and one store per lane, so that the compiler can neither
 * drop nor fold its work.
 */
#include <hip/hip_runtime.h>
#include "tp_peak.h"

namespace {

constexpr int BLOCK = 256;
constexpr int CHAINS = 16;       /* independent FMA chains per lane */
constexpr int MCHAINS = 4;       /* independent MFMA accumulators per wave */
constexpr int UNROLL = 4;        /* 16-byte loads in flight per lane (HBM sweeps) */

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f4 __attribute__((ext_vector_type(4)));

/* each lane keeps UNROLL 16-byte loads in flight; the copy's stores bypass
 * the caches (nontemporal), as a streaming copy's should */
__global__ __launch_bounds__(BLOCK) void k_copy16(const f4 *__restrict__ src,
                                                  f4 *__restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * BLOCK;
    size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    for (; i + (UNROLL - 1) * stride < n; i += UNROLL * stride) {
        f4 v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) __builtin_nontemporal_store(v[u], &dst[i + u * stride]);
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

__global__ __launch_bounds__(BLOCK) void k_read16(const f4 *__restrict__ src,
                                                  float *__restrict__ out, size_t n)
{
    const size_t stride = (size_t)gridDim.x * BLOCK;
    const size_t t = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    f4 s[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; u++) s[u] = (f4){0.f, 0.f, 0.f, 0.f};
    size_t i = t;
    for (; i + (UNROLL - 1) * stride < n; i += UNROLL * stride) {
#pragma unroll
        for (int u = 0; u < UNROLL; u++) s[u] += src[i + u * stride];
    }
    for (; i < n; i += stride) s[0] += src[i];
    f4 a = s[0];
#pragma unroll
    for (int u = 1; u < UNROLL; u++) a += s[u];
    out[t] = a.x + a.y + a.z + a.w;
}

/* x <- x a + b converges to b / (1 - a) for 0 < a < 1, so the chains stay
 * finite however long they run */
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_fma(T *__restrict__ out, int iters, T a, T b)
{
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    T x[CHAINS];
#pragma unroll
    for (int u = 0; u < CHAINS; u++) x[u] = (T)(threadIdx.x & 63) * (T)1e-3 + (T)u;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int u = 0; u < CHAINS; u++) x[u] = __builtin_fma(x[u], a, b);
    }
    T s = 0;
#pragma unroll
    for (int u = 0; u < CHAINS; u++) s += x[u];
    out[t] = s;
}

template <>
__global__ __launch_bounds__(BLOCK) void k_fma<float>(float *__restrict__ out, int iters,
                                                       float a, float b)
{
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    float x[CHAINS];
#pragma unroll
    for (int u = 0; u < CHAINS; u++) x[u] = (float)(threadIdx.x & 63) * 1e-3f + (float)u;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int u = 0; u < CHAINS; u++) x[u] = __builtin_fmaf(x[u], a, b);
    }
    float s = 0.f;
#pragma unroll
    for (int u = 0; u < CHAINS; u++) s += x[u];
    out[t] = s;
}

__global__ __launch_bounds__(BLOCK) void k_mfma_bf16(float *__restrict__ out, int iters)
{
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    bf16x8 a, b;
#pragma unroll
    for (int u = 0; u < 8; u++) {
        a[u] = (__bf16)((float)((threadIdx.x + u) & 7) * 0.125f);
        b[u] = (__bf16)((float)((threadIdx.x * 3 + u) & 7) * -0.0625f);
    }
    f32x16 c[MCHAINS];
#pragma unroll
    for (int m = 0; m < MCHAINS; m++)
#pragma unroll
        for (int u = 0; u < 16; u++) c[m][u] = 0.f;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int m = 0; m < MCHAINS; m++)
            c[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c[m], 0, 0, 0);
    }
    float s = 0.f;
#pragma unroll
    for (int m = 0; m < MCHAINS; m++)
#pragma unroll
        for (int u = 0; u < 16; u++) s += c[m][u];
    out[t] = s;
}

}  // namespace

extern "C" int tp_peak_measure(int device, int what, size_t bytes, int reps, double *value,
                               double *ms)
{
    if (!value || reps < 1 || what < TP_PEAK_HBM_COPY || what > TP_PEAK_BF16_MFMA)
        return TP_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return TP_E_NODEV;
    if (device < 0 || device >= ndev) return TP_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return TP_E_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return TP_E_HIP;
    const int ncu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 1;

    const bool hbm = what == TP_PEAK_HBM_COPY || what == TP_PEAK_HBM_READ;
    if (bytes == 0) bytes = (size_t)4 << 30;
    bytes &= ~(size_t)4095;
    if (hbm && bytes < ((size_t)1 << 20)) return TP_E_ARG;
    /* HBM sweeps: grid-stride, at 4, 16 and 64 blocks per CU (the best
     * counts); arithmetic loops: 8 waves of 64 lanes per SIMD (4 SIMDs per
     * CU) in one round */
    const int hbm_shapes[3] = {4, 16, 64};
    const int nshapes = hbm ? 3 : 1;
    const size_t max_threads = (size_t)ncu * (hbm ? 64 : 8) * BLOCK;
    const int iters = what == TP_PEAK_FP64_FMA ? 32768
                    : what == TP_PEAK_FP32_FMA ? 65536 : 16384;

    void *src = nullptr, *dst = nullptr, *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = TP_OK;
    float best = -1.f;
    if (hipMalloc(&out, max_threads * sizeof(double)) != hipSuccess) { out = nullptr; rc = TP_E_NOMEM; goto done; }
    if (hbm) {
        if (hipMalloc(&src, bytes) != hipSuccess) { src = nullptr; rc = TP_E_NOMEM; goto done; }
        if (hipMemset(src, 0, bytes) != hipSuccess) { rc = TP_E_HIP; goto done; }
        if (what == TP_PEAK_HBM_COPY &&
            hipMalloc(&dst, bytes) != hipSuccess) { dst = nullptr; rc = TP_E_NOMEM; goto done; }
    }
    if (hipStreamCreate(&st) != hipSuccess) { st = nullptr; rc = TP_E_HIP; goto done; }
    if (hipEventCreate(&e0) != hipSuccess) { e0 = nullptr; rc = TP_E_HIP; goto done; }
    if (hipEventCreate(&e1) != hipSuccess) { e1 = nullptr; rc = TP_E_HIP; goto done; }

    for (int sh = 0; sh < nshapes; sh++)
    for (int r = -1; r < reps; r++) {            /* r = -1: untimed first launch */
        const int blocks = ncu * (hbm ? hbm_shapes[sh] : 8);
        if (hipEventRecord(e0, st) != hipSuccess) { rc = TP_E_HIP; goto done; }
        const size_t n16 = bytes / 16;
        switch (what) {
        case TP_PEAK_HBM_COPY:
            k_copy16<<<blocks, BLOCK, 0, st>>>((const f4 *)src, (f4 *)dst, n16);
            break;
        case TP_PEAK_HBM_READ:
            k_read16<<<blocks, BLOCK, 0, st>>>((const f4 *)src, (float *)out, n16);
            break;
        case TP_PEAK_FP64_FMA:
            k_fma<double><<<blocks, BLOCK, 0, st>>>((double *)out, iters, 0.9999, 1e-4);
            break;
        case TP_PEAK_FP32_FMA:
            k_fma<float><<<blocks, BLOCK, 0, st>>>((float *)out, iters, 0.9999f, 1e-4f);
            break;
        default:
            k_mfma_bf16<<<blocks, BLOCK, 0, st>>>((float *)out, iters);
        }
        if (hipGetLastError() != hipSuccess) { rc = TP_E_HIP; goto done; }
        if (hipEventRecord(e1, st) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess) { rc = TP_E_HIP; goto done; }
        float t = 0.f;
        if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) { rc = TP_E_HIP; goto done; }
        if (r >= 0 && (best < 0.f || t < best)) best = t;
    }
    {
        const double sec = best * 1e-3;
        double work;
        switch (what) {
        case TP_PEAK_HBM_COPY: work = 2.0 * (double)bytes / 1e9; break;    /* GB */
        case TP_PEAK_HBM_READ: work = (double)bytes / 1e9; break;
        case TP_PEAK_BF16_MFMA:                                             /* TFLOP */
            work = (double)max_threads / 64 * iters * MCHAINS * 32768.0 / 1e12;
            break;
        default: work = (double)max_threads * iters * CHAINS * 2.0 / 1e12;
        }
        *value = work / sec;
        if (ms) *ms = best;
    }
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (out) (void)hipFree(out);
    return rc;
}
