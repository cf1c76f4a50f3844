// Path-guiding network object and its C-ABI entry points (vgpu_guide_*).
//
// No reference counterpart: the network is a new component of the GPU path
// tracer (the reference samples only its BSDFs).  Kernels: guide.hip.

#include <new>
#include <random>

#include "guide_internal.h"
#include "scene_internal.h"

using vgpu::fail;
using vgpu_detail::DeviceGuard;
namespace G = vgpu::guide;

vgpu_guide::~vgpu_guide() {
    int prev = 0;
    const bool ok = hipGetDevice(&prev) == hipSuccess;
    (void)hipSetDevice(device);
    void* ptrs[] = {d_params, d_m, d_v, d_frag_fwd, d_frag_bwd, tb.partial, tb.loss, tb.csum, tb.step, tb.grad,
                    tb.slot_fwd, tb.slot_bwd, d_stage};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    if (ok) (void)hipSetDevice(prev);
}

namespace {

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// He-uniform weights (bound sqrt(6 / fan_in)), zero biases; the output layer is scaled by
// 0.1 so that the initial mixture is broad (kappa ~ 1) with near-equal weights.
void init_params(uint64_t seed, std::vector<float>& P) {
    P.assign(G::NPARAM, 0.0f);
    uint64_t s = seed;
    auto uni = [&]() { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0; };
    const int ow[5] = {G::OW1, G::OW2, G::OW3, G::OW4, G::OW5};
    const int nin[5] = {G::NENC, G::NH, G::NH, G::NH, G::NH};
    const int nout[5] = {G::NH, G::NH, G::NH, G::NH, G::NOUT};
    for (int l = 0; l < 5; l++) {
        const double b = std::sqrt(6.0 / nin[l]) * (l == 4 ? 0.1 : 1.0);
        for (int i = 0; i < nout[l] * nin[l]; i++) P[ow[l] + i] = (float)(b * uni());
    }
}

int upload_params(vgpu_guide* g, const float* P) {
    HIP_TRY(hipMemcpy(g->d_params, P, sizeof(float) * G::NPARAM, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(g->d_m, 0, sizeof(float) * G::NPARAM));
    HIP_TRY(hipMemset(g->d_v, 0, sizeof(float) * G::NPARAM));
    HIP_TRY(hipMemset(g->tb.step, 0, 2 * sizeof(int32_t)));   // step count and its copy
    HIP_TRY(G::launch_pack(g->d_params, g->d_frag_fwd, g->d_frag_bwd, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    return VGPU_OK;
}

int ensure_stage(vgpu_guide* g, size_t bytes) {
    if (bytes <= g->stage_bytes) return VGPU_OK;
    if (g->d_stage) { HIP_TRY(hipFree(g->d_stage)); g->d_stage = nullptr; g->stage_bytes = 0; }
    HIP_TRY(hipMalloc(&g->d_stage, bytes));
    g->stage_bytes = bytes;
    return VGPU_OK;
}

}  // namespace

extern "C" {

int vgpu_guide_create(vgpu_guide** out, int device, const float* bounds, uint64_t seed, uint32_t max_batch) {
    if (!out || !bounds) return fail(VGPU_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (max_batch == 0 || max_batch % G::KSPLIT != 0 || max_batch > (1u << 22))
        return fail(VGPU_ERR_ARG, "max_batch %u must be a positive multiple of %d up to 2^22", max_batch, G::KSPLIT);
    for (int k = 0; k < 3; k++)
        if (!(bounds[3 + k] > bounds[k]) || !std::isfinite(bounds[k]) || !std::isfinite(bounds[3 + k]))
            return fail(VGPU_ERR_ARG, "bounds must be finite with hi > lo on every axis");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(VGPU_ERR_ARG, "device %d not available", device);
    vgpu_guide* g = new (std::nothrow) vgpu_guide();
    if (!g) return fail(VGPU_ERR_NOMEM, "out of host memory");
    g->device = device;
    g->max_batch = max_batch;
    std::memcpy(g->bounds, bounds, sizeof(g->bounds));
    DeviceGuard dg(device);
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    g->num_cus = prop.multiProcessorCount;
    const size_t P4 = sizeof(float) * G::NPARAM;
    const uint32_t nwg = max_batch / G::TRAIN_WG;
    if (e == hipSuccess) e = hipMalloc(&g->d_params, P4);
    if (e == hipSuccess) e = hipMalloc(&g->d_m, P4);
    if (e == hipSuccess) e = hipMalloc(&g->d_v, P4);
    if (e == hipSuccess) e = hipMalloc(&g->d_frag_fwd, (size_t)G::NF_FWD * 64 * 16);
    if (e == hipSuccess) e = hipMalloc(&g->d_frag_bwd, (size_t)G::NF_BWD * 64 * 16);
    if (e == hipSuccess) e = hipMalloc(&g->tb.partial, (size_t)(max_batch / G::TRAIN_WG) * P4);
    if (e == hipSuccess) e = hipMalloc(&g->tb.loss, sizeof(float) * max_batch);
    if (e == hipSuccess) e = hipMalloc(&g->tb.csum, sizeof(float) * (1 + nwg));
    if (e == hipSuccess) e = hipMalloc(&g->tb.step, 256);
    if (e == hipSuccess) e = hipMalloc(&g->tb.grad, P4);
    if (e == hipSuccess) e = hipMalloc(&g->tb.slot_fwd, sizeof(int32_t) * G::NPARAM);
    if (e == hipSuccess) e = hipMalloc(&g->tb.slot_bwd, sizeof(int32_t) * G::NPARAM);
    if (e == hipSuccess) e = hipMemset(g->tb.slot_fwd, 0xff, sizeof(int32_t) * G::NPARAM);
    if (e == hipSuccess) e = hipMemset(g->tb.slot_bwd, 0xff, sizeof(int32_t) * G::NPARAM);
    if (e == hipSuccess) e = G::launch_slot_map(g->tb.slot_fwd, g->tb.slot_bwd, nullptr);
    if (e != hipSuccess) { delete g; return fail(VGPU_ERR_HIP, "guide buffers: %s", hipGetErrorString(e)); }
    g->tb.max_batch = max_batch;
    std::vector<float> P;
    init_params(seed, P);
    const int rc = upload_params(g, P.data());
    if (rc) { delete g; return rc; }
    *out = g;
    return VGPU_OK;
}

int vgpu_guide_destroy(vgpu_guide* g) {
    delete g;
    return VGPU_OK;
}

int vgpu_guide_set_params(vgpu_guide* g, const float* params) {
    if (!g || !params) return fail(VGPU_ERR_ARG, "NULL argument");
    DeviceGuard dg(g->device);
    return upload_params(g, params);
}

int vgpu_guide_get_params(vgpu_guide* g, float* params) {
    if (!g || !params) return fail(VGPU_ERR_ARG, "NULL argument");
    DeviceGuard dg(g->device);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(params, g->d_params, sizeof(float) * G::NPARAM, hipMemcpyDeviceToHost));
    return VGPU_OK;
}

int vgpu_guide_get_grad(vgpu_guide* g, float* grad) {
    if (!g || !grad) return fail(VGPU_ERR_ARG, "NULL argument");
    DeviceGuard dg(g->device);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(grad, g->tb.grad, sizeof(float) * G::NPARAM, hipMemcpyDeviceToHost));
    return VGPU_OK;
}

int vgpu_guide_infer(vgpu_guide* g, uint32_t flags, const float* queries, uint64_t n, float* out, void* stream) {
    if (!g || (n && (!queries || !out))) return fail(VGPU_ERR_ARG, "NULL argument");
    if (n == 0) return VGPU_OK;
    DeviceGuard dg(g->device);
    hipStream_t st = (hipStream_t)stream;
    const bool host = (flags & VGPU_TRACE_DEVICE_PTRS) == 0;
    const float* dq = queries;
    float* dout = out;
    if (host) {
        const int rc = ensure_stage(g, n * 32 + n * 160);
        if (rc) return rc;
        dq = (const float*)g->d_stage;
        dout = (float*)((char*)g->d_stage + n * 32);
        HIP_TRY(hipMemcpyAsync((void*)dq, queries, n * 32, hipMemcpyHostToDevice, st));
    }
    HIP_TRY(G::launch_infer(g->net(), (const float4*)dq, n, nullptr, dout, (flags & VGPU_GUIDE_RAW) != 0, g->num_cus, st));
    if (host) {
        HIP_TRY(hipMemcpyAsync(out, dout, n * 160, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return VGPU_OK;
}

int vgpu_guide_train_step(vgpu_guide* g, uint32_t flags, const float* samples, uint32_t n, float lr, void* stream,
                          float* loss_out) {
    if (!g || !samples) return fail(VGPU_ERR_ARG, "NULL argument");
    if (n == 0 || n % G::KSPLIT != 0 || n > g->max_batch)
        return fail(VGPU_ERR_ARG, "batch %u must be a positive multiple of %d up to max_batch %u", n, G::KSPLIT,
                    g->max_batch);
    if (!(lr >= 0.0f) || !std::isfinite(lr)) return fail(VGPU_ERR_ARG, "learning rate must be finite and >= 0");
    DeviceGuard dg(g->device);
    hipStream_t st = (hipStream_t)stream;
    const bool host = (flags & VGPU_TRACE_DEVICE_PTRS) == 0;
    const G::TrainSample* ds = (const G::TrainSample*)samples;
    if (host) {
        const int rc = ensure_stage(g, (size_t)n * sizeof(G::TrainSample));
        if (rc) return rc;
        ds = (const G::TrainSample*)g->d_stage;
        HIP_TRY(hipMemcpyAsync((void*)ds, samples, (size_t)n * sizeof(G::TrainSample), hipMemcpyHostToDevice, st));
    }
    HIP_TRY(G::launch_train(g->net(), g->d_params, g->d_m, g->d_v, G::SampleSource{ds, nullptr, 0, 0}, n, lr,
                            (flags & VGPU_GUIDE_NO_UPDATE) == 0, g->tb, st));
    if (host) {
        if (loss_out) {
            // per-sample c_i (-log q_i); the batch loss divides their sum by sum c_i
            std::vector<float> L(n);
            float cs = 0.0f;
            HIP_TRY(hipMemcpyAsync(L.data(), g->tb.loss, sizeof(float) * n, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(&cs, g->tb.csum, sizeof(float), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            double s = 0.0;
            for (float v : L) s += v;
            *loss_out = cs > 0.0f ? (float)(s / cs) : 0.0f;
        }
        HIP_TRY(hipStreamSynchronize(st));
    }
    return VGPU_OK;
}

}  // extern "C"
