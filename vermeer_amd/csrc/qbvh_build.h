// Internal interface of the host QBVH builders (qbvh_build.cpp).
#pragma once

#include <cstdint>
#include <hip/hip_runtime.h>
#include "../../include/vermeer_gpu.h"

namespace vgpu {

// Both return 0, -1 (bad argument / max_nodes too small: *nnodes_out still
// holds the required count) or -3 (degenerate input).
int build_qbvh(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leafMax, vgpu_qbvh_node* nodes_out,
               int32_t max_nodes, int32_t* nnodes_out, float* bounds_out, int32_t nthreads);
int build_qbvh_motion(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leafMax,
                      vgpu_mqbvh_node* nodes_out, int32_t max_nodes, int32_t* nnodes_out, int32_t nthreads);

// The GPU builder (bvh_build.hip): the same output as build_qbvh, built on the current device with
// work enqueued on `st`.  Returns 0, -1, -2 (HIP error), -3 (degenerate or non-finite input) or -4
// (a subtree deeper than the builder's stack; build_qbvh handles any depth).
int build_qbvh_gpu(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leafMax, vgpu_qbvh_node* nodes_out,
                   int32_t max_nodes, int32_t* nnodes_out, float* bounds_out, hipStream_t st);

// Host -> device copy through the GPU builder's pinned staging slots (a plain hipMemcpyAsync below
// 4 MiB), on the current device and ordered on `st`; staged copies are complete on return.
hipError_t stage_upload(void* dev, const void* host, size_t bytes, hipStream_t st);

}  // namespace vgpu
