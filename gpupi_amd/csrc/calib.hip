// Calibration kernels for the rooflines the benchmark reports (no
// counterpart in the reference):
//   - the integer multiply-add peak: v_mad_u64_u32 (32x32 -> 64 bits plus a
//     64-bit addend, the schoolbook and group-stage inner step) issued back
//     to back on every lane of every CU;
//   - a plain copy of a known number of bytes, against which the HBM traffic
//     counters (FETCH_SIZE, WRITE_SIZE) of the NTT kernels are scaled.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "gpupi_pi.h"

namespace {

constexpr int kChains = 8;

// Each lane runs kChains independent chains acc = a * lo32(acc) + acc, one
// v_mad_u64_u32 per step; the result is stored so nothing is removed.
__global__ void __launch_bounds__(256) mac_peak_kernel(uint64_t *out, uint32_t iters, uint32_t seed) {
    uint32_t a[kChains];
    uint64_t acc[kChains];
#pragma unroll
    for (int i = 0; i < kChains; ++i) {
        a[i] = (seed + i) * 0x9e3779b9u + threadIdx.x;
        acc[i] = ((uint64_t)blockIdx.x << 32) + threadIdx.x + i;
    }
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < kChains; ++i) acc[i] = (uint64_t)a[i] * (uint32_t)acc[i] + acc[i];
    }
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < kChains; ++i) s ^= acc[i];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 16-byte loads and stores, grid-stride.
__global__ void __launch_bounds__(256) copy_kernel(uint4 *dst, const uint4 *src, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

extern "C" {

hipError_t gpupi_calib_mac_peak(double *macs_per_second) {
    if (macs_per_second == nullptr) return hipErrorInvalidValue;
    int dev = 0, cus = 0;
    GPUPI_RETURN_IF(hipGetDevice(&dev));
    GPUPI_RETURN_IF(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t blocks = (uint32_t)cus * 32u, iters = 8192;
    uint64_t *out = nullptr;
    GPUPI_RETURN_IF(hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint64_t)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    float best = 0.0f;
    for (int r = 0; r < 4 && err == hipSuccess; ++r) {
        err = hipEventRecord(e0, nullptr);
        if (err == hipSuccess) hipLaunchKernelGGL(mac_peak_kernel, dim3(blocks), dim3(256), 0, nullptr, out, iters, 7u + r);
        if (err == hipSuccess) err = hipGetLastError();
        if (err == hipSuccess) err = hipEventRecord(e1, nullptr);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        if (err == hipSuccess && r > 0 && (best == 0.0f || ms < best)) best = ms;  // run 0 warms up
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (err != hipSuccess) return err;
    *macs_per_second = (double)blocks * 256.0 * iters * kChains / (best * 1e-3);
    return hipSuccess;
}

hipError_t gpupi_calib_copy(void *dst, const void *src, size_t bytes, hipStream_t stream) {
    if ((bytes & 15u) != 0 || (bytes > 0 && (dst == nullptr || src == nullptr))) return hipErrorInvalidValue;
    if (bytes == 0) return hipSuccess;
    const size_t n16 = bytes / 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(copy_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, static_cast<uint4 *>(dst),
                       static_cast<const uint4 *>(src), n16);
    return hipGetLastError();
}

}  // extern "C"
