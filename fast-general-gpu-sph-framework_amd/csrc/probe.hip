// probe.hip - the measurement-only library libgsph_probe.so (include/
// gsph_probe.h): calibration kernels for the roofline figures of the SPH
// step.  Not part of the product library; bench.py and one GPU test load it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "gsph.h"
#include "gsph_probe.h"

namespace {

constexpr int kNT = 256;
thread_local std::string g_error;

int fail(int code, const std::string &m) {
    g_error = m;
    return code;
}

// Plain float4 copy: calibration reference for HBM byte counters.
__global__ __launch_bounds__(kNT) void k_copy(const float4 *src, float4 *dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (size_t)gridDim.x * kNT)
        dst[i] = src[i];
}

// Eight independent FP32 FMA chains per lane, 64 FMAs per iteration: the
// peak VALU issue rate (calibration reference for the VALU counters).  The
// multiplier and addend are per-lane values, so every FMA reads vector
// registers only: with a scalar operand (a kernel argument) the FMAs issue at
// about half the rate that all-VGPR FP32 instructions reach on this GPU.
__global__ __launch_bounds__(kNT) void k_fma_probe(float *out, int iters, float b, float c) {
    float a[8];
    const float bl = b + 1e-9f * (float)threadIdx.x;
    const float cl = c * (float)(1 + (threadIdx.x & 7));
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = (float)(threadIdx.x + u);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = fmaf(a[u], bl, cl);
        }
    }
    float t = 0.f;
#pragma unroll
    for (int u = 0; u < 8; ++u) t += a[u];
    out[(size_t)blockIdx.x * kNT + threadIdx.x] = t;
}

void launch_copy(const float4 *src, float4 *dst, size_t n, hipStream_t s) {
    const int blocks = (int)std::min<size_t>((n + kNT - 1) / kNT, 8192);
    if (blocks > 0) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(kNT), 0, s, src, dst, n);
}

}  // namespace

extern "C" {

const char *gsph_probe_last_error(void) { return g_error.c_str(); }

int gsph_probe_copy(double bytes, int reps, float *ms) {
    if (!(bytes >= 16) || bytes > 64e9 || reps <= 0 || reps > 1000)
        return fail(GSPH_ERR_INVALID_ARG, "bad probe size");
    const size_t n = (size_t)(bytes / 16);
    float4 *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = GSPH_OK;
    if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess ||
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(GSPH_ERR_HIP, "probe allocation failed");
    } else {
        (void)hipMemsetAsync(a, 0, n * 16, s);
        launch_copy(a, b, n, s);
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; ++r) launch_copy(a, b, n, s);
        (void)hipEventRecord(e1, s);
        if (hipStreamSynchronize(s) != hipSuccess) rc = fail(GSPH_ERR_HIP, "probe failed");
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        if (ms) *ms = t / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int gsph_probe_valu(int iters, float *ms, double *lane_ops) {
    if (iters <= 0 || iters > (1 << 20)) return fail(GSPH_ERR_INVALID_ARG, "bad probe length");
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return fail(GSPH_ERR_NO_DEVICE, "no device");
    const int blocks = 8 * cus;   // 8 blocks of 4 waves per CU: 8 waves per SIMD
    float *out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = GSPH_OK;
    if (hipMalloc(&out, sizeof(float) * (size_t)blocks * kNT) != hipSuccess ||
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(GSPH_ERR_HIP, "probe allocation failed");
    } else {
        hipLaunchKernelGGL(k_fma_probe, dim3(blocks), dim3(kNT), 0, s, out, 16, 0.999f, 0.001f);
        (void)hipEventRecord(e0, s);
        hipLaunchKernelGGL(k_fma_probe, dim3(blocks), dim3(kNT), 0, s, out, iters, 0.999f, 0.001f);
        (void)hipEventRecord(e1, s);
        if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess)
            rc = fail(GSPH_ERR_HIP, "probe failed");
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        if (ms) *ms = t;
        if (lane_ops) *lane_ops = (double)blocks * kNT * (double)iters * 64.0;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (out) (void)hipFree(out);
    return rc;
}

}  // extern "C"
