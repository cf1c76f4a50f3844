// Streaming kernels that measure what HBM delivers to a plain access pattern, and that calibrate
// the FETCH_SIZE / WRITE_SIZE hardware counters against bytes actually moved (a diagnostic with no
// reference counterpart).
//
// Copy: every thread moves 16-byte vectors in a grid-stride loop, four in flight per iteration.
// Read: the same loads summed into one float per thread, stored once so the loads stay live.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "scene_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr int kUnroll = 4;
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kBlock) void k_stream_copy(const v4f* __restrict__ src, v4f* __restrict__ dst,
                                                        uint64_t n4) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    for (; i + (kUnroll - 1) * stride < n4; i += kUnroll * stride) {
        v4f v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) v[u] = __builtin_nontemporal_load(&src[i + u * stride]);
#pragma unroll
        for (int u = 0; u < kUnroll; u++) __builtin_nontemporal_store(v[u], &dst[i + u * stride]);
    }
    for (; i < n4; i += stride) dst[i] = src[i];
}

__global__ __launch_bounds__(kBlock) void k_stream_read(const v4f* __restrict__ src, float* __restrict__ sums,
                                                        uint64_t n4) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    float acc = 0.0f;
    for (; i + (kUnroll - 1) * stride < n4; i += kUnroll * stride) {
        v4f v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) v[u] = __builtin_nontemporal_load(&src[i + u * stride]);
#pragma unroll
        for (int u = 0; u < kUnroll; u++) acc += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    }
    for (; i < n4; i += stride) acc += (src[i].x + src[i].y) + (src[i].z + src[i].w);
    sums[(uint64_t)blockIdx.x * kBlock + threadIdx.x] = acc;
}

}  // namespace

extern "C" int vgpu_diag_stream(int32_t mode, const void* src, void* dst, uint64_t bytes, int32_t blocks_per_cu,
                                void* stream) {
    if (!src || !dst) return vgpu::fail(VGPU_ERR_ARG, "NULL buffer");
    if (bytes % 16 != 0 || ((uintptr_t)src | (uintptr_t)dst) % 16 != 0)
        return vgpu::fail(VGPU_ERR_ARG, "buffers and size must be 16-byte aligned");
    if (mode != VGPU_DIAG_COPY && mode != VGPU_DIAG_READ) return vgpu::fail(VGPU_ERR_ARG, "unknown mode %d", mode);
    if (blocks_per_cu < 1 || blocks_per_cu > 64) return vgpu::fail(VGPU_ERR_ARG, "blocks_per_cu must be in [1,64]");
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned grid = (unsigned)(cus * blocks_per_cu);
    const uint64_t n4 = bytes / 16;
    if (mode == VGPU_DIAG_COPY)
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(kBlock), 0, (hipStream_t)stream, (const v4f*)src,
                           (v4f*)dst, n4);
    else
        hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(kBlock), 0, (hipStream_t)stream, (const v4f*)src,
                           (float*)dst, n4);
    HIP_TRY(hipGetLastError());
    return VGPU_OK;
}
