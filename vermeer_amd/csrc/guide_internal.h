// Guide object shared by guide.cpp (its C ABI) and the path tracer.
#pragma once

#include "guide_kernels.h"

struct vgpu_guide {
    int device = 0;
    int num_cus = 0;
    uint32_t max_batch = 0;
    float* d_params = nullptr;    // fp32 master
    float* d_m = nullptr;         // Adam first moment
    float* d_v = nullptr;         // Adam second moment
    uint4* d_frag_fwd = nullptr;
    uint4* d_frag_bwd = nullptr;
    vgpu::guide::TrainBuffers tb{};
    void* d_stage = nullptr;      // host-pointer mode staging
    size_t stage_bytes = 0;
    float bounds[6] = {0, 0, 0, 0, 0, 0};

    vgpu::guide::Net net() const {
        vgpu::guide::Net n{};
        n.params = d_params;
        n.frag_fwd = d_frag_fwd;
        n.frag_bwd = d_frag_bwd;
        for (int k = 0; k < 6; k++) n.bounds[k] = bounds[k];
        return n;
    }
    ~vgpu_guide();
};
