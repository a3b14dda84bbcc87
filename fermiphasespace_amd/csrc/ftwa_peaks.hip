// ftwa_peaks.hip -- the device's own roofline peaks: an HBM copy, FP32 and FP64 FMA loops and an FP32
// MFMA loop, each timed with HIP events on a stream of its own.  The roofline fractions of the
// integrators are quoted against the published MI355X figures; these measured figures stand beside
// them (what a kernel on this device and toolchain can reach).  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>

#include "../../include/fps_ftwa.h"
#include "ftwa_internal.hpp"

namespace fpsamd {

namespace {

// Streaming copy: each workgroup moves chunks of 4 x 256 16-byte units with nontemporal loads and stores
// (which keep the streams out of the caches: 6.0 TB/s on 1 GiB here, against 3.6-3.9 TB/s for the same
// copy through ordinary loads and stores); n counts 16-byte units
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) peak_copy_kernel(const v4f* __restrict__ src, v4f* __restrict__ dst, size_t n)
{
    constexpr int U = 4;
    const size_t chunk = static_cast<size_t>(256) * U;
    const size_t stride = static_cast<size_t>(gridDim.x) * chunk;
    for (size_t base = static_cast<size_t>(blockIdx.x) * chunk; base < n; base += stride) {
        v4f v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t e = base + u * 256 + threadIdx.x;
            if (e < n) v[u] = __builtin_nontemporal_load(src + e);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t e = base + u * 256 + threadIdx.x;
            if (e < n) __builtin_nontemporal_store(v[u], dst + e);
        }
    }
}

// 8 independent multiply-add chains per lane (2 flop each per iteration)
template <typename T>
__global__ void __launch_bounds__(256) peak_fma_kernel(T* __restrict__ out, int iters, T b, T c)
{
    const T x = static_cast<T>(threadIdx.x) * T(1e-3);
    T a0 = x, a1 = x + T(1), a2 = x + T(2), a3 = x + T(3), a4 = x + T(4), a5 = x + T(5), a6 = x + T(6),
      a7 = x + T(7);
    for (int it = 0; it < iters; ++it) {
        a0 = a0 * b + c;
        a1 = a1 * b + c;
        a2 = a2 * b + c;
        a3 = a3 * b + c;
        a4 = a4 * b + c;
        a5 = a5 * b + c;
        a6 = a6 * b + c;
        a7 = a7 * b + c;
    }
    out[static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

// 4 independent 32x32x2 FP32 MFMA accumulators per wave: 4 x 4,096 flop per iteration
typedef float v16f __attribute__((ext_vector_type(16)));
__global__ void __launch_bounds__(256) peak_mfma_kernel(float* __restrict__ out, int iters)
{
    const float a = 1e-3f * static_cast<float>(threadIdx.x & 63), b = 1e-3f * static_cast<float>(threadIdx.x >> 6);
    v16f c0 = {}, c1 = {}, c2 = {}, c3 = {};
    for (int it = 0; it < iters; ++it) {
        c0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_32x32x2f32(b, a, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, a, c2, 0, 0, 0);
        c3 = __builtin_amdgcn_mfma_f32_32x32x2f32(b, b, c3, 0, 0, 0);
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += c0[k] + c1[k] + c2[k] + c3[k];
    out[static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x] = s;
}

struct Timer {
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Timer()
    {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    }
    hipError_t init()
    {
        hipError_t e = hipStreamCreate(&st);
        if (e == hipSuccess) e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        return e;
    }
    // best of `reps` timed launches after one untimed launch, in ms
    template <typename F>
    hipError_t best_ms(F launch, int reps, double& ms)
    {
        launch();
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        ms = 1e30;
        for (int r = 0; r < reps; ++r) {
            if ((e = hipEventRecord(e0, st)) != hipSuccess) return e;
            launch();
            if ((e = hipGetLastError()) != hipSuccess) return e;
            if ((e = hipEventRecord(e1, st)) != hipSuccess) return e;
            if ((e = hipEventSynchronize(e1)) != hipSuccess) return e;
            float t = 0.f;
            if ((e = hipEventElapsedTime(&t, e0, e1)) != hipSuccess) return e;
            ms = std::min(ms, static_cast<double>(t));
        }
        return hipSuccess;
    }
};

struct DevMem {
    void* p = nullptr;
    ~DevMem()
    {
        if (p) (void)hipFree(p);
    }
};

hipError_t measure(fpsftwa_peaks& pk)
{
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    Timer tm;
    if ((e = tm.init()) != hipSuccess) return e;

    // HBM: 1 GiB copied, read + write bytes per launch (far above the last-level caches)
    const size_t bytes = size_t(1) << 30, n4 = bytes / sizeof(v4f);
    DevMem src, dst, out;
    if ((e = hipMalloc(&src.p, bytes)) != hipSuccess) return e;
    if ((e = hipMalloc(&dst.p, bytes)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(src.p, 0, bytes, tm.st)) != hipSuccess) return e;
    const unsigned copy_blocks = static_cast<unsigned>(std::max(1, cus) * 16);
    double ms = 0;
    e = tm.best_ms([&] {
        hipLaunchKernelGGL(peak_copy_kernel, dim3(copy_blocks), dim3(256), 0, tm.st,
                           static_cast<const v4f*>(src.p), static_cast<v4f*>(dst.p), n4);
    }, 5, ms);
    if (e != hipSuccess) return e;
    pk.copy_bytes = 2.0 * static_cast<double>(bytes);
    pk.hbm_copy_gbps = pk.copy_bytes / (ms * 1e-3) / 1e9;

    // FMA loops: 8 workgroups of 4 waves per CU
    const unsigned blocks = static_cast<unsigned>(std::max(1, cus) * 8);
    if ((e = hipMalloc(&out.p, static_cast<size_t>(blocks) * 256 * sizeof(double))) != hipSuccess) return e;
    const int iters = 16384;
    const double lanes = static_cast<double>(blocks) * 256;
    e = tm.best_ms([&] {
        hipLaunchKernelGGL(peak_fma_kernel<float>, dim3(blocks), dim3(256), 0, tm.st, static_cast<float*>(out.p),
                           iters, 0.999f, 1e-4f);
    }, 5, ms);
    if (e != hipSuccess) return e;
    pk.fp32_fma_tflops = lanes * iters * 16.0 / (ms * 1e-3) / 1e12;
    e = tm.best_ms([&] {
        hipLaunchKernelGGL(peak_fma_kernel<double>, dim3(blocks), dim3(256), 0, tm.st, static_cast<double*>(out.p),
                           iters, 0.999, 1e-4);
    }, 5, ms);
    if (e != hipSuccess) return e;
    pk.fp64_fma_tflops = lanes * iters * 16.0 / (ms * 1e-3) / 1e12;

    // MFMA: 4 x 4,096 flop per wave and iteration
    const int mi = 4096;
    e = tm.best_ms([&] {
        hipLaunchKernelGGL(peak_mfma_kernel, dim3(blocks), dim3(256), 0, tm.st, static_cast<float*>(out.p), mi);
    }, 5, ms);
    if (e != hipSuccess) return e;
    pk.fp32_mfma_tflops = static_cast<double>(blocks) * 4 * mi * 4.0 * 4096.0 / (ms * 1e-3) / 1e12;
    return hipStreamSynchronize(tm.st);
}

}  // namespace

}  // namespace fpsamd

extern "C" int fpsftwa_measure_peaks(fpsftwa_peaks* out)
{
    if (!out) {
        fpsamd::set_last_error("null argument");
        return FPSFTWA_EINVAL;
    }
    *out = fpsftwa_peaks{};
    const hipError_t e = fpsamd::measure(*out);
    if (e != hipSuccess) {
        fpsamd::set_last_error(std::string("fpsftwa_measure_peaks: ") + hipGetErrorString(e));
        return FPSFTWA_EDEVICE;
    }
    return FPSFTWA_OK;
}
