// Path-mode renderer object and its C-ABI entry points (vgpu_pt_*).
//
// One pass = camera rays, then per bounce: closest hit, hit shading, (guide
// inference), scatter with NEE, shadow any-hit, shadow resolve; then the FP64
// accumulation and, in guided mode, the training steps.  Queue lengths stay on
// the device: the host enqueues a fixed sequence of launches per pass and
// never waits inside it.

#include <new>

#include "guide_internal.h"
#include "pathtrace_kernels.h"
#include "scene_internal.h"

using vgpu::fail;
using namespace vgpu_detail;

struct vgpu_pt {
    vgpu_scene* scene = nullptr;
    vgpu_guide* guide = nullptr;
    vgpu::PTView v{};
    vgpu::PTBuffers b{};
    uint32_t train_steps = 0, batch = 0;
    float lr = 1e-3f;
    vgpu::PTMat* d_mats = nullptr;
    vgpu::PTLight* d_lights = nullptr;
    double* d_stage = nullptr;
    std::vector<void*> owned;
    vgpu_detail::GraphCache graph;   // VGPU_GRAPH passes

    ~vgpu_pt() {
        int prev = 0;
        const bool ok = hipGetDevice(&prev) == hipSuccess;
        if (scene) (void)hipSetDevice(scene->device);
        for (void* p : owned) if (p) (void)hipFree(p);
        if (ok) (void)hipSetDevice(prev);
    }
    hipError_t alloc(void** p, size_t bytes) {
        const hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 256));
        if (e == hipSuccess) owned.push_back(*p);
        return e;
    }
};

namespace {

float lum(const float* c) { return (0.2126f * c[0] + 0.7152f * c[1]) + 0.0722f * c[2]; }

}  // namespace

extern "C" {

int vgpu_pt_create(vgpu_pt** out, vgpu_scene* s, const vgpu_pt_material* materials, uint32_t nmaterials,
                   const vgpu_camera* cam, uint32_t width, uint32_t height, uint32_t max_depth, uint64_t seed) {
    if (!out || !s || !materials || !cam) return fail(VGPU_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (!s->committed) return fail(VGPU_ERR_STATE, "pt_create before scene commit");
    if (s->any_motion) return fail(VGPU_ERR_ARG, "path mode supports static scenes only");
    if (width == 0 || height == 0 || (uint64_t)width * height > (1ull << 24))
        return fail(VGPU_ERR_ARG, "image size %ux%u: 1 to 2^24 pixels", width, height);
    if (max_depth < 1 || max_depth > 32) return fail(VGPU_ERR_ARG, "max_depth %u outside [1,32]", max_depth);
    if (s->max_material >= (int32_t)nmaterials)
        return fail(VGPU_ERR_ARG, "a mesh refers to material %d of %u", s->max_material, nmaterials);
    for (uint32_t i = 0; i < nmaterials; i++)
        for (int k = 0; k < 3; k++) {
            const float a = materials[i].albedo[k], e = materials[i].emission[k];
            if (!(a >= 0.0f && a <= 1.0f)) return fail(VGPU_ERR_ARG, "material %u: albedo must be in [0,1]", i);
            if (!(e >= 0.0f) || !std::isfinite(e)) return fail(VGPU_ERR_ARG, "material %u: emission must be >= 0", i);
        }
    for (const MeshHost& m : s->meshes)
        if (m.sphere) return fail(VGPU_ERR_ARG, "path mode takes triangle meshes only (the scene holds a sphere)");
    // emissive triangles become the NEE light list (selection ~ area x luminance)
    std::vector<vgpu::PTLight> L;
    std::vector<double> w;
    for (size_t gi = 0; gi < s->meshes.size(); gi++) {
        const MeshHost& m = s->meshes[gi];
        for (uint32_t t = 0; t < m.ntris; t++) {
            const int si = m.shaderidx.empty() ? 0 : m.shaderidx[t];
            if (si >= (int)m.shaders.size()) continue;
            const vgpu_pt_material& mt = materials[m.shaders[si]];
            const float le = lum(mt.emission);
            if (!(le > 0.0f)) continue;
            const float* a = m.vert(0, m.idx[3 * (size_t)t]);
            const float* b = m.vert(0, m.idx[3 * (size_t)t + 1]);
            const float* c = m.vert(0, m.idx[3 * (size_t)t + 2]);
            const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            const float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            const float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            const float nl = std::sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
            if (!(nl > 0.0f)) continue;
            vgpu::PTLight l{};
            const float area = 0.5f * nl;
            l.v0 = make_float4(a[0], a[1], a[2], area);
            l.e1 = make_float4(e1[0], e1[1], e1[2], 0.0f);
            l.e2 = make_float4(e2[0], e2[1], e2[2], mt.one_sided ? 1.0f : 0.0f);
            l.n = make_float4(n[0] / nl, n[1] / nl, n[2] / nl, 0.0f);
            l.Le = make_float4(mt.emission[0], mt.emission[1], mt.emission[2], 0.0f);
            L.push_back(l);
            w.push_back((double)area * (double)le);
        }
    }
    double wsum = 0.0;
    for (double x : w) wsum += x;
    double acc = 0.0;
    for (size_t i = 0; i < L.size(); i++) {
        acc += w[i];
        L[i].e1.w = i + 1 == L.size() ? 1.0f : (float)(acc / wsum);
        L[i].Le.w = (float)((double)lum(&L[i].Le.x) / wsum);
    }
    std::vector<vgpu::PTMat> M(nmaterials);
    for (uint32_t i = 0; i < nmaterials; i++) {
        const vgpu_pt_material& mt = materials[i];
        M[i].a = make_float4(mt.albedo[0], mt.albedo[1], mt.albedo[2], mt.one_sided ? 1.0f : 0.0f);
        const float le = lum(mt.emission);
        M[i].e = make_float4(mt.emission[0], mt.emission[1], mt.emission[2],
                             (le > 0.0f && wsum > 0.0) ? (float)((double)le / wsum) : 0.0f);
    }
    vgpu_pt* pt = new (std::nothrow) vgpu_pt();
    if (!pt) return fail(VGPU_ERR_NOMEM, "out of host memory");
    pt->scene = s;
    vgpu::PTView& v = pt->v;
    v.sv = s->view;
    v.gshade = s->d_gshade;
    v.shaderidx = s->d_shaderidx;
    v.shaders = s->d_shaders;
    v.nlights = (int32_t)L.size();
    v.width = (int32_t)width;
    v.height = (int32_t)height;
    v.max_depth = (int32_t)max_depth;
    if (cam->radius != 0.0f) { delete pt; return fail(VGPU_ERR_ARG, "path mode takes a pinhole camera (radius 0)"); }
    std::memcpy(v.M, cam->M, sizeof(v.M));
    v.tan_theta_focal = cam->tan_theta_focal;
    v.aspect = cam->aspect;
    v.focal = cam->focal;
    v.seed_lo = (uint32_t)seed;
    v.seed_hi = (uint32_t)(seed >> 32);
    v.alpha = 0.0f;
    v.rec_stride = 0;
    std::lock_guard<std::recursive_mutex> lk(s->mu);
    DeviceGuard dg(s->device);
    const size_t np = (size_t)width * height;
    vgpu::PTBuffers& b = pt->b;
    hipError_t e = hipSuccess;
    for (int q = 0; q < 2 && e == hipSuccess; q++) {
        e = pt->alloc((void**)&b.qray[q], np * 32);
        if (e == hipSuccess) e = pt->alloc((void**)&b.qstate[q], np * 16);
        if (e == hipSuccess) e = pt->alloc((void**)&b.qmeta[q], np * 16);
    }
    if (e == hipSuccess) e = pt->alloc((void**)&b.count, 64);
    if (e == hipSuccess) e = pt->alloc((void**)&b.rays, 64);
    if (e == hipSuccess) e = pt->alloc((void**)&b.hits, np * 32);
    if (e == hipSuccess) e = pt->alloc((void**)&b.shade, np * 48);
    if (e == hipSuccess) e = pt->alloc((void**)&b.query, np * 32);
    if (e == hipSuccess) e = pt->alloc((void**)&b.sray, np * 32);
    if (e == hipSuccess) e = pt->alloc((void**)&b.scontrib, np * 32);
    if (e == hipSuccess) e = pt->alloc((void**)&b.shits, np * 16);
    if (e == hipSuccess) e = pt->alloc((void**)&b.lpass, np * 16);
    if (e == hipSuccess) e = pt->alloc((void**)&pt->d_mats, sizeof(vgpu::PTMat) * nmaterials);
    if (e == hipSuccess) e = pt->alloc((void**)&pt->d_lights, sizeof(vgpu::PTLight) * L.size());
    if (e == hipSuccess) e = hipMemcpy(pt->d_mats, M.data(), sizeof(vgpu::PTMat) * nmaterials, hipMemcpyHostToDevice);
    if (e == hipSuccess && !L.empty())
        e = hipMemcpy(pt->d_lights, L.data(), sizeof(vgpu::PTLight) * L.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(b.count, 0, 64);
    if (e == hipSuccess) e = hipMemset(b.rays, 0, 64);
    if (e != hipSuccess) { delete pt; return fail(VGPU_ERR_HIP, "path buffers: %s", hipGetErrorString(e)); }
    v.mats = pt->d_mats;
    v.lights = pt->d_lights;
    b.lobes = nullptr;
    b.rec = nullptr;
    b.plen = nullptr;
    vgpu::TraceLaunch cfg{s->cfg.block, s->cfg.lds_stack, s->cfg.waves_per_cu};
    const int rc = ensure_spill(s, std::max(vgpu::trace_grid_threads(cfg, np, s->num_cus),
                                            vgpu::trace_grid_threads(vgpu::kExtLaunch, np, s->num_cus)));
    if (rc) { delete pt; return rc; }
    *out = pt;
    return VGPU_OK;
}

int vgpu_pt_destroy(vgpu_pt* pt) {
    delete pt;
    return VGPU_OK;
}

int vgpu_pt_set_guide(vgpu_pt* pt, vgpu_guide* g, float alpha, uint32_t record_stride, uint32_t train_steps,
                      uint32_t batch, float lr) {
    if (!pt) return fail(VGPU_ERR_ARG, "NULL argument");
    if (!g) {
        pt->guide = nullptr;
        pt->v.alpha = 0.0f;
        pt->v.rec_stride = 0;
        pt->train_steps = 0;
        return VGPU_OK;
    }
    if (!(alpha > 0.0f && alpha < 1.0f)) return fail(VGPU_ERR_ARG, "alpha must be in (0,1): the BSDF keeps the estimator unbiased");
    if (g->device != pt->scene->device) return fail(VGPU_ERR_ARG, "guide and scene live on different devices");
    if (train_steps > 0 && (record_stride == 0 || batch == 0 || batch % vgpu::guide::KSPLIT != 0 || batch > g->max_batch))
        return fail(VGPU_ERR_ARG, "training needs record_stride > 0 and a batch that is a multiple of %d up to %u",
                    vgpu::guide::KSPLIT, g->max_batch);
    if (!(lr >= 0.0f) || !std::isfinite(lr)) return fail(VGPU_ERR_ARG, "learning rate");
    DeviceGuard dg(pt->scene->device);
    const size_t np = (size_t)pt->v.width * pt->v.height;
    vgpu::PTBuffers& b = pt->b;
    hipError_t e = hipSuccess;
    if (!b.lobes) e = pt->alloc((void**)&b.lobes, np * vgpu::guide::NOUT * sizeof(float));
    if (e == hipSuccess && record_stride > 0 && !b.rec) {
        e = pt->alloc((void**)&b.rec, np * pt->v.max_depth * vgpu::PT_REC_F4 * 16);
        if (e == hipSuccess) e = pt->alloc((void**)&b.plen, np * 4);
    }
    if (e == hipSuccess && train_steps > 0) {
        const uint64_t paths = np / record_stride + 4096;
        const uint64_t cap = std::min<uint64_t>(paths * (uint64_t)pt->v.max_depth, 1ull << 26);
        if (cap > b.tcap) {
            e = pt->alloc((void**)&b.tsamples, cap * sizeof(vgpu::guide::TrainSample));
            if (e == hipSuccess) b.tcap = (uint32_t)cap;
        }
    }
    if (e != hipSuccess) return fail(VGPU_ERR_HIP, "guide buffers: %s", hipGetErrorString(e));
    pt->guide = g;
    pt->v.alpha = alpha;
    pt->v.rec_stride = record_stride;
    pt->train_steps = train_steps;
    pt->batch = batch;
    pt->lr = lr;
    return VGPU_OK;
}

int vgpu_pt_pass(vgpu_pt* pt, uint32_t flags, uint32_t pass, double* accum, void* stream) {
    if (!pt || !accum) return fail(VGPU_ERR_ARG, "NULL argument");
    vgpu_scene* s = pt->scene;
    std::lock_guard<std::recursive_mutex> lk(s->mu);
    DeviceGuard dg(s->device);
    hipStream_t st = (hipStream_t)stream;
    const size_t np = (size_t)pt->v.width * pt->v.height;
    const bool host = (flags & VGPU_TRACE_DEVICE_PTRS) == 0;
    double* dacc = accum;
    if (host) {
        if (!pt->d_stage) {
            hipError_t e = pt->alloc((void**)&pt->d_stage, np * 3 * sizeof(double));
            if (e != hipSuccess) return fail(VGPU_ERR_HIP, "staging: %s", hipGetErrorString(e));
        }
        dacc = pt->d_stage;
        HIP_TRY(hipMemcpyAsync(dacc, accum, np * 3 * sizeof(double), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemsetAsync(s->d_err, 0, sizeof(int32_t), st));
    }
    const bool graph = (flags & VGPU_GRAPH) != 0;
    if (graph && host) return fail(VGPU_ERR_ARG, "VGPU_GRAPH needs VGPU_TRACE_DEVICE_PTRS");
    if (graph && s->timing) return fail(VGPU_ERR_STATE, "VGPU_GRAPH with kernel timing enabled");
    if (graph && s->cfg.sort > 0) { const int rc = ensure_sort(s, np); if (rc) return rc; }
    // The pass's launches, on the caller's stream or on the capture stream of the pass's graph.
    auto enqueue = [&](hipStream_t st) -> int {
        vgpu::PTView v = pt->v;
        vgpu::PTBuffers& b = pt->b;
        const bool guided = pt->guide != nullptr && v.alpha > 0.0f;
        const bool train = guided && pt->train_steps > 0 && v.rec_stride > 0 && (flags & VGPU_PT_NO_TRAIN) == 0;
        if (!train) v.rec_stride = 0;
        vgpu::PTBuffers bb = b;
        if (!train) bb.plen = nullptr;
        const char* stage = "camera";
        hipError_t e = vgpu::pt_launch_camera(v, pass, bb, st);
        for (int d = 0; d < v.max_depth && e == hipSuccess; d++) {
            const int cur = d & 1;
            stage = "closest hit";
            hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
            if (s->timing) {   // the same [sort start, trace start, trace end] records as enqueue_trace
                for (auto& x : ev) x = timing_event(s);
                if (!ev[0] || !ev[1] || !ev[2]) return fail(VGPU_ERR_HIP, "timing events");
                HIP_TRY(hipEventRecord(ev[0], st));
                HIP_TRY(hipEventRecord(ev[1], st));
            }
            e = vgpu::launch_trace_ext(s->view, false, bb.qray[cur], np, bb.hits, s->d_spT, s->d_spN, s->spill_threads,
                                       s->d_err, s->num_cus, st, bb.count + cur);
            if (ev[2] && e == hipSuccess) HIP_TRY(hipEventRecord(ev[2], st));
            if (e == hipSuccess) { stage = "hit"; e = vgpu::pt_launch_hit(v, pass, cur, bb, st); }
            if (e != hipSuccess || d + 1 >= v.max_depth) continue;
            if (guided) {
                stage = "guide inference";
                e = vgpu::guide::launch_infer(pt->guide->net(), bb.query, np, bb.count + cur, bb.lobes, false,
                                              pt->guide->num_cus, st);
            }
            if (e == hipSuccess) { stage = "scatter"; e = vgpu::pt_launch_scatter(v, pass, cur, bb, st); }
            if (e == hipSuccess) {
                const int rc = enqueue_trace(s, VGPU_TRACE_SHADOW | VGPU_TRACE_DEVICE_PTRS, bb.sray, np, bb.shits, nullptr,
                                             nullptr, st, s->cfg, bb.count + 2);
                if (rc) return rc;
                stage = "shadow";
                e = vgpu::pt_launch_shadow(v, bb, st);
            }
        }
        if (e == hipSuccess) { stage = "accumulate"; e = vgpu::pt_launch_accumulate(v, bb, dacc, st); }
        if (e == hipSuccess && train) {
            stage = "gather";
            e = vgpu::pt_launch_gather(v, pass, bb, st);
            vgpu_guide* g = pt->guide;
            for (uint32_t k = 0; k < pt->train_steps && e == hipSuccess; k++) {
                // each step draws its batch from the pass's training samples inside the training kernels
                stage = "training step";
                const vgpu::guide::SampleSource src{bb.tsamples, bb.count + 3, bb.tcap, pass * 7919u + k * 104729u};
                e = vgpu::guide::launch_train(g->net(), g->d_params, g->d_m, g->d_v, src, pt->batch, pt->lr, true, g->tb, st);
            }
        }
        if (e != hipSuccess) return fail(VGPU_ERR_HIP, "path pass (%s): %s", stage, hipGetErrorString(e));
        return VGPU_OK;
    };
    const int rc = graph ? pt->graph.run(st, enqueue) : enqueue(st);
    if (rc) return rc;
    if (host) {
        HIP_TRY(hipMemcpyAsync(accum, dacc, np * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
        int32_t err = 0;
        HIP_TRY(hipMemcpyAsync(&err, s->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (err) return fail(VGPU_ERR_STACK, "a ray needed more than %d traversal stack entries", VGPU_MAX_STACK);
    }
    return VGPU_OK;
}

int vgpu_pt_stats(vgpu_pt* pt, uint64_t* rays, uint32_t* train_samples) {
    if (!pt) return fail(VGPU_ERR_ARG, "NULL argument");
    DeviceGuard dg(pt->scene->device);
    HIP_TRY(hipDeviceSynchronize());
    if (rays) HIP_TRY(hipMemcpy(rays, pt->b.rays, sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (train_samples) HIP_TRY(hipMemcpy(train_samples, pt->b.count + 3, sizeof(uint32_t), hipMemcpyDeviceToHost));
    return VGPU_OK;
}

}  // extern "C"
