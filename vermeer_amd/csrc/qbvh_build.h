// Internal interface of the host QBVH builders (qbvh_build.cpp).
#pragma once

#include <cstdint>
#include "../../include/vermeer_gpu.h"

namespace vgpu {

// Both return 0, -1 (bad argument / max_nodes too small: *nnodes_out still
// holds the required count) or -3 (degenerate input).
int build_qbvh(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leafMax, vgpu_qbvh_node* nodes_out,
               int32_t max_nodes, int32_t* nnodes_out, float* bounds_out, int32_t nthreads);
int build_qbvh_motion(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leafMax,
                      vgpu_mqbvh_node* nodes_out, int32_t max_nodes, int32_t* nnodes_out, int32_t nthreads);

}  // namespace vgpu
