// Scene object and C-ABI entry points (include/vermeer_gpu.h).
//
// Host side of the drop-in boundary: it mirrors what the reference does at
// PreRender time — PolyMesh.initAccel per mesh (builtin/geom/polymesh/
// buildqbvh.go:14-212) and Scene.initAccel for the top level
// (builtin/scene/scene.go:135-268) — then flattens the result into the
// device layout of trace_kernels.h and uploads it once.  Traversal calls are
// batched: one kernel launch per call, never one call per ray.
//
// Compiled with -ffp-contract=off: the box/centroid arithmetic feeding the
// builder must round exactly as the reference's.

#include <new>
#include <thread>

#include "scene_internal.h"

namespace {

// Runs f(begin, end) over [0, n), split over up to 8 threads when n is large (the flatten's
// per-corner gathers of millions of triangles).  A thread that cannot be started runs its range on
// the calling thread.
template <class F>
void parallel_ranges(size_t n, const F& f) {
    constexpr int kMax = 8;
    const int T = n >= ((size_t)1 << 18) ? kMax : 1;
    const size_t per = (n + T - 1) / T;
    std::thread th[kMax];
    for (int t = 1; t < T; t++) {
        const size_t a = std::min(n, per * t), b = std::min(n, a + per);
        if (a >= b) continue;
        try {
            th[t] = std::thread([&f, a, b]() { f(a, b); });
        } catch (...) {
            f(a, b);
        }
    }
    f(0, std::min(n, per));
    for (int t = 1; t < T; t++)
        if (th[t].joinable()) th[t].join();
}

}  // namespace

namespace vgpu {

namespace {
thread_local std::string g_err;
}

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

void clear_error() { g_err.clear(); }

}  // namespace vgpu

using vgpu::fail;
using namespace vgpu_detail;

namespace vgpu_detail {

// Meshes with at least this many triangles build their BVH on the scene's GPU (vgpu::build_qbvh_gpu);
// smaller ones on the host, where the launch and transfer overheads would dominate.  Both give the
// reference's tree.  The host builder also takes any mesh whose tree is deeper than the GPU
// builder's stack (-4).
constexpr uint32_t kGpuBuildMinTris = 1u << 15;

int build_nodes(float* boxes, float* cent, int32_t* ids, uint32_t n, int leaf_max, std::vector<vgpu_qbvh_node>& nodes,
                float* bounds, int device) {
    int32_t nn = 0;
    nodes.assign(std::max<uint32_t>(n, 1), vgpu_qbvh_node{});
    int rc = -4;
    if (device >= 0 && n >= kGpuBuildMinTris) {
        DeviceGuard dg(device);
        rc = vgpu::build_qbvh_gpu(boxes, cent, ids, (int32_t)n, leaf_max, nodes.data(), (int32_t)nodes.size(), &nn,
                                  bounds, nullptr);
        if (rc == -2) return fail(VGPU_ERR_HIP, "HIP error in the GPU BVH builder");
    }
    if (rc == -4)
        rc = vgpu::build_qbvh(boxes, cent, ids, (int32_t)n, leaf_max, nodes.data(), (int32_t)nodes.size(), &nn,
                              bounds, 0);
    if (rc == -3) return fail(VGPU_ERR_BUILD, "BVH build failed: degenerate mesh (coincident or non-finite centroids)");
    if (rc != 0) return fail(VGPU_ERR_BUILD, "BVH build failed (%d)", rc);
    nodes.resize(nn);
    return VGPU_OK;
}

// Static mesh accel, as PolyMesh.initAccel (buildqbvh.go:60-129).
int build_static_mesh(MeshHost& m, int leaf_max, int device) {
    const uint32_t n = m.ntris;
    std::vector<float> boxes((size_t)n * 6), cent((size_t)n * 3);
    std::vector<int32_t> ids(n);
    parallel_ranges(n, [&](size_t a, size_t e) {
        for (size_t i = a; i < e; i++) {
            const float* V0 = m.vert(0, m.idx[i * 3 + 0]);
            const float* V1 = m.vert(0, m.idx[i * 3 + 1]);
            const float* V2 = m.vert(0, m.idx[i * 3 + 2]);
            Box6 b; b.reset(); b.grow_point(V0); b.grow_point(V1); b.grow_point(V2);
            std::memcpy(&boxes[i * 6], b.v, sizeof(b.v));
            for (int k = 0; k < 3; k++) cent[i * 3 + k] = (V0[k] + V1[k] + V2[k]) / 3.0f;
            ids[i] = (int32_t)i;
        }
    });
    std::vector<vgpu_qbvh_node> nodes;
    const int rc = build_nodes(boxes.data(), cent.data(), ids.data(), n, leaf_max, nodes, m.bounds, device);
    if (rc) return rc;
    m.nodes.swap(nodes);
    m.perm = ids;
    m.idxp_sorted.resize((size_t)n * 3);
    parallel_ranges(n, [&](size_t a, size_t e) {
        for (size_t i = a; i < e; i++)
            for (int c = 0; c < 3; c++) m.idxp_sorted[i * 3 + c] = m.idx[(size_t)ids[i] * 3 + c];
    });
    if (!m.shaderidx.empty()) {   // shaderidx follows the leaf order (buildqbvh.go:98-104)
        m.shaderidx_sorted.resize(n);
        for (uint32_t i = 0; i < n; i++) m.shaderidx_sorted[i] = m.shaderidx[ids[i]];
    }
    m.built = true;
    return VGPU_OK;
}

// Per-key node boxes of a motion topology (buildqbvh.go:171-212, scene.go:228-268).
void motion_boxes_rec(const std::vector<vgpu_mqbvh_node>& topo, int32_t node, float* boxes_key,
                      const std::function<void(int, int, Box6&)>& leaf, Box6& nodebox) {
    nodebox.reset();
    for (int c = 0; c < 4; c++) {
        const int32_t ch = topo[node].children[c];
        if (ch < 0) {
            if (ch == -1) continue;
            Box6 b; b.reset();
            leaf((int)((ch & 0x7ffffff) >> 4), (int)((ch & 0xf) + 1), b);
            for (int k = 0; k < 3; k++) { boxes_key[node * 24 + c + k * 4] = b.v[k]; boxes_key[node * 24 + c + 12 + k * 4] = b.v[3 + k]; }
            nodebox.grow_box(b.v);
        } else {
            Box6 b;
            motion_boxes_rec(topo, ch, boxes_key, leaf, b);
            for (int k = 0; k < 3; k++) { boxes_key[node * 24 + c + k * 4] = b.v[k]; boxes_key[node * 24 + c + 12 + k * 4] = b.v[3 + k]; }
            nodebox.grow_box(b.v);
        }
    }
}

// Motion mesh accel (buildqbvh.go:19-57 + initMotionBoxes :148-169).  BuildAccelMotion's topology is
// BuildAccel's with the boxes left out (the same splits, qbvh/motionbuild.go:12-106), so a large mesh
// takes it from the GPU builder.
int build_motion_mesh(MeshHost& m, int device) {
    const uint32_t n = m.ntris;
    std::vector<float> boxes((size_t)n * 6), cent((size_t)n * 3);
    std::vector<int32_t> ids(n);
    const float k = 0.5f * (float)(m.nkeys - 1);
    const float time = k - std::floor(k);
    const int key = (int)std::floor(k), key2 = (int)std::ceil(k);
    for (uint32_t i = 0; i < n; i++) {
        float V[3][3];
        for (int j = 0; j < 3; j++) lerp3(m.vert(key, m.idx[i * 3 + j]), m.vert(key2, m.idx[i * 3 + j]), time, V[j]);
        Box6 b; b.reset(); b.grow_point(V[0]); b.grow_point(V[1]); b.grow_point(V[2]);
        std::memcpy(&boxes[(size_t)i * 6], b.v, sizeof(b.v));
        for (int c = 0; c < 3; c++) cent[(size_t)i * 3 + c] = (V[0][c] + V[1][c] + V[2][c]) / 3.0f;
        ids[i] = (int32_t)i;
    }
    std::vector<vgpu_mqbvh_node> nodes;
    if (device >= 0 && n >= kGpuBuildMinTris) {
        std::vector<vgpu_qbvh_node> full;
        float bounds[6];
        const int rc = build_nodes(boxes.data(), cent.data(), ids.data(), n, 16, full, bounds, device);
        if (rc) return rc;
        nodes.resize(full.size());
        for (size_t i = 0; i < full.size(); i++) {
            vgpu_mqbvh_node& d = nodes[i];
            d.axis0 = (int32_t)full[i].axis0;
            d.axis1 = (int32_t)full[i].axis1;
            d.axis2 = (int32_t)full[i].axis2;
            for (int c = 0; c < 4; c++) d.children[c] = full[i].children[c];
            d.pad = 0;
        }
    } else {
        int32_t nn = 0;
        nodes.resize(std::max<uint32_t>(n, 1));
        int rc = vgpu::build_qbvh_motion(boxes.data(), cent.data(), ids.data(), (int32_t)n, 16, nodes.data(),
                                         (int32_t)nodes.size(), &nn, 0);
        if (rc == -3) return fail(VGPU_ERR_BUILD, "motion BVH build failed: degenerate mesh");
        if (rc != 0) return fail(VGPU_ERR_BUILD, "motion BVH build failed (%d)", rc);
        nodes.resize(nn);
    }
    m.mnodes.swap(nodes);
    m.perm = ids;
    const size_t nn = m.mnodes.size();
    m.mboxes.assign((size_t)m.nkeys * nn * 24, 0.0f);
    m.motion_bounds.assign((size_t)m.nkeys * 6, 0.0f);
    Box6 full; full.reset();
    for (uint32_t kk = 0; kk < m.nkeys; kk++) {
        auto leaf = [&](int lb, int lc, Box6& b) {
            for (int i = lb; i < lb + lc; i++) {
                const int f = m.perm[i];
                for (int j = 0; j < 3; j++) b.grow_point(m.vert(kk, m.idx[(size_t)f * 3 + j]));
            }
        };
        Box6 rb;
        motion_boxes_rec(m.mnodes, 0, &m.mboxes[(size_t)kk * nn * 24], leaf, rb);
        std::memcpy(&m.motion_bounds[kk * 6], rb.v, sizeof(rb.v));
        full.grow_box(rb.v);
    }
    std::memcpy(m.bounds, full.v, sizeof(full.v));
    m.built = true;
    return VGPU_OK;
}

void to_dev(const vgpu_qbvh_node& s, vgpu::DevNode& d) {
    std::memcpy(d.lox, s.boxes, 24 * sizeof(float));   // rows lox, loy, loz, hix, hiy, hiz
    for (int c = 0; c < 4; c++) d.children[c] = s.children[c];
    d.axes[0] = (int32_t)s.axis0; d.axes[1] = (int32_t)s.axis1; d.axes[2] = (int32_t)s.axis2; d.axes[3] = 0;
}

void append_motion_nodes(std::vector<float4>& out, const std::vector<vgpu_mqbvh_node>& topo, const std::vector<float>& boxes,
                         int nkeys) {
    const size_t nn = topo.size();
    for (size_t i = 0; i < nn; i++) {
        float4 t0, t1;
        std::memcpy(&t0, topo[i].children, 16);
        int32_t ax[4] = {topo[i].axis0, topo[i].axis1, topo[i].axis2, 0};
        std::memcpy(&t1, ax, 16);
        out.push_back(t0); out.push_back(t1);
        for (int k = 0; k < nkeys; k++) {
            const float* b = &boxes[((size_t)k * nn + i) * 24];
            for (int r = 0; r < 6; r++) out.push_back(make_float4(b[r * 4 + 0], b[r * 4 + 1], b[r * 4 + 2], b[r * 4 + 3]));
        }
    }
}

template <class T>
int upload(T** dptr, const std::vector<T>& v) {
    if (v.empty()) return VGPU_OK;
    HIP_TRY(hipMalloc((void**)dptr, v.size() * sizeof(T)));
    HIP_TRY(vgpu::stage_upload(*dptr, v.data(), v.size() * sizeof(T), nullptr));   // pinned slots when large
    HIP_TRY(hipStreamSynchronize(nullptr));
    return VGPU_OK;
}

int ensure_spill(vgpu_scene* s, uint64_t threads) {
    if (s->d_spT && s->spill_threads >= threads) return VGPU_OK;
    if (s->d_spT) { HIP_TRY(hipFree(s->d_spT)); s->d_spT = nullptr; }
    if (s->d_spN) { HIP_TRY(hipFree(s->d_spN)); s->d_spN = nullptr; }
    // size for the smallest LDS depth so later configuration changes never need more
    const uint64_t per = (uint64_t)(VGPU_MAX_STACK - 8);
    HIP_TRY(hipMalloc((void**)&s->d_spT, per * threads * sizeof(float)));
    HIP_TRY(hipMalloc((void**)&s->d_spN, per * threads * sizeof(int32_t)));
    s->spill_threads = threads;
    return VGPU_OK;
}

int check_config(const vgpu_trace_config& c) {
    if (c.block != 64 && c.block != 128 && c.block != 256) return fail(VGPU_ERR_ARG, "block must be 64, 128 or 256");
    if (c.lds_stack != 8 && c.lds_stack != 12 && c.lds_stack != 16 && c.lds_stack != 24 && c.lds_stack != 32)
        return fail(VGPU_ERR_ARG, "lds_stack must be 8, 12, 16, 24 or 32");
    if (c.block * c.lds_stack * 8 > 65536) return fail(VGPU_ERR_ARG, "block*lds_stack exceeds 64 KiB of LDS");
    if (c.waves_per_cu < 1 || c.waves_per_cu > 32) return fail(VGPU_ERR_ARG, "waves_per_cu must be in [1,32]");
    if (c.sort < 0 || c.sort > 6) return fail(VGPU_ERR_ARG, "sort must be 0 (off) or 1..6 grid bits per axis");
    if (c.refill < 0 || c.refill > 64) return fail(VGPU_ERR_ARG, "refill must be 0 (default) or 1..64 idle lanes");
    return VGPU_OK;
}

int ensure_sort(vgpu_scene* s, uint64_t n) {
    vgpu::RaySortBuffers& sb = s->sort;
    if (!sb.hist) {
        HIP_TRY(hipMalloc((void**)&sb.hist, sizeof(uint32_t) * vgpu::ray_sort_buckets(6)));
        HIP_TRY(hipMalloc((void**)&sb.bsum, sizeof(uint32_t) * vgpu::ray_sort_scan_blocks(6)));
    }
    if (sb.capacity >= n) return VGPU_OK;
    HIP_TRY(hipDeviceSynchronize());
    if (sb.keys) { HIP_TRY(hipFree(sb.keys)); sb.keys = nullptr; }
    if (sb.perm) { HIP_TRY(hipFree(sb.perm)); sb.perm = nullptr; }
    if (sb.rays) { HIP_TRY(hipFree(sb.rays)); sb.rays = nullptr; }
    sb.capacity = 0;
    HIP_TRY(hipMalloc((void**)&sb.keys, sizeof(uint32_t) * n));
    HIP_TRY(hipMalloc((void**)&sb.perm, sizeof(uint32_t) * n));
    HIP_TRY(hipMalloc((void**)&sb.rays, 32 * n));
    sb.capacity = n;
    return VGPU_OK;
}

GraphCache::~GraphCache() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (cap) (void)hipStreamDestroy(cap);
}

int GraphCache::run(hipStream_t st, const std::function<int(hipStream_t)>& body) {
    if (!cap) HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    HIP_TRY(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    const int rc = body(cap);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(cap, &graph);
    if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (ec != hipSuccess) return fail(VGPU_ERR_HIP, "graph capture: %s", hipGetErrorString(ec));
    hipError_t e = hipSuccess;
    if (exec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res;
        if (hipGraphExecUpdate(exec, graph, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(exec);
            exec = nullptr;
        }
    }
    if (!exec) {
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e == hipSuccess) instantiations++;
        else exec = nullptr;
    }
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(VGPU_ERR_HIP, "graph instantiation: %s", hipGetErrorString(e));
    HIP_TRY(hipGraphLaunch(exec, st));
    return VGPU_OK;
}

// Next free timing event (the pool grows; vgpu_scene_timing recycles it).
hipEvent_t timing_event(vgpu_scene* s) {
    if (s->ev_used == s->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        s->ev_pool.push_back(e);
    }
    return s->ev_pool[s->ev_used++];
}

int enqueue_trace(vgpu_scene* s, uint32_t flags, const void* d_rays, uint64_t n, void* d_hits, int32_t* d_geoms,
                  int32_t* d_counters, hipStream_t st, const vgpu_trace_config& cfg, const uint32_t* n_dev) {
    if (d_counters && s->view.any_xform)
        return fail(VGPU_ERR_ARG, "per-ray counters are not kept for scenes with instances, procs or mesh transforms");
    vgpu::TraceLaunch L{cfg.block, cfg.lds_stack, cfg.waves_per_cu, cfg.refill > 0 ? cfg.refill : vgpu::kRefillDefault};
    const uint64_t threads = vgpu::trace_grid_threads(L, n, s->num_cus);
    int rc = ensure_spill(s, threads);
    if (rc) return rc;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    if (s->timing && n > 0) {
        for (auto& e : ev) e = timing_event(s);
        if (!ev[0] || !ev[1] || !ev[2]) return fail(VGPU_ERR_HIP, "timing events");
        HIP_TRY(hipEventRecord(ev[0], st));
    }
    const uint32_t* perm = nullptr;
    if (cfg.sort > 0) {
        if (n > 0xffffffffull) return fail(VGPU_ERR_ARG, "ray sorting handles at most 2^32-1 rays per call");
        if ((rc = ensure_sort(s, n))) return rc;
        hipError_t e = vgpu::launch_ray_sort(s->box, cfg.sort, d_rays, n, n_dev, s->sort, s->num_cus, st);
        if (e != hipSuccess) return fail(VGPU_ERR_HIP, "ray sort launch: %s", hipGetErrorString(e));
        d_rays = s->sort.rays;
        perm = s->sort.perm;
    }
    if (ev[1]) HIP_TRY(hipEventRecord(ev[1], st));
    hipError_t e = vgpu::launch_trace(s->view, (flags & VGPU_TRACE_SHADOW) != 0, s->any_motion, L, d_rays, n, d_hits,
                                      d_geoms, d_counters, s->d_spT, s->d_spN, s->spill_threads, s->d_err, s->num_cus, st,
                                      n_dev, perm);
    if (e != hipSuccess) return fail(VGPU_ERR_HIP, "trace launch: %s", hipGetErrorString(e));
    if (ev[2]) HIP_TRY(hipEventRecord(ev[2], st));
    return VGPU_OK;
}

}  // namespace

// =============================================================== C ABI

extern "C" {

const char* vgpu_last_error(void) { return vgpu::g_err.c_str(); }

int vgpu_device_count(int* count) {
    if (!count) return fail(VGPU_ERR_ARG, "count is NULL");
    HIP_TRY(hipGetDeviceCount(count));
    return VGPU_OK;
}

int vgpu_qbvh_build(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leaf_max, vgpu_qbvh_node* nodes_out,
                    int32_t max_nodes, int32_t* nnodes_out, float* bounds_out, int32_t nthreads) {
    int rc = vgpu::build_qbvh(boxes, centroids, idx, n, leaf_max, nodes_out, max_nodes, nnodes_out, bounds_out, nthreads);
    if (rc == -3) return fail(VGPU_ERR_BUILD, "degenerate input: the reference recursion does not terminate");
    if (rc != 0) return fail(VGPU_ERR_ARG, "bad arguments or max_nodes too small (need %d)", nnodes_out ? *nnodes_out : -1);
    return VGPU_OK;
}

int vgpu_qbvh_build_gpu(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leaf_max,
                        vgpu_qbvh_node* nodes_out, int32_t max_nodes, int32_t* nnodes_out, float* bounds_out,
                        void* stream) {
    int rc = vgpu::build_qbvh_gpu(boxes, centroids, idx, n, leaf_max, nodes_out, max_nodes, nnodes_out, bounds_out,
                                  (hipStream_t)stream);
    if (rc == -3) return fail(VGPU_ERR_BUILD, "degenerate or non-finite input: the reference build fails or does not terminate");
    if (rc == -4) return fail(VGPU_ERR_STACK, "tree deeper than the GPU builder's stack (vgpu_qbvh_build handles it)");
    if (rc == -2) return fail(VGPU_ERR_HIP, "HIP error in the GPU builder");
    if (rc != 0) return fail(VGPU_ERR_ARG, "bad arguments or max_nodes too small (need %d)", nnodes_out ? *nnodes_out : -1);
    return VGPU_OK;
}

int vgpu_qbvh_build_motion(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leaf_max,
                           vgpu_mqbvh_node* nodes_out, int32_t max_nodes, int32_t* nnodes_out, int32_t nthreads) {
    int rc = vgpu::build_qbvh_motion(boxes, centroids, idx, n, leaf_max, nodes_out, max_nodes, nnodes_out, nthreads);
    if (rc == -3) return fail(VGPU_ERR_BUILD, "degenerate input: the reference recursion does not terminate");
    if (rc != 0) return fail(VGPU_ERR_ARG, "bad arguments or max_nodes too small (need %d)", nnodes_out ? *nnodes_out : -1);
    return VGPU_OK;
}

int vgpu_scene_create(vgpu_scene** out, int device) {
    if (!out) return fail(VGPU_ERR_ARG, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(VGPU_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
    vgpu_scene* s = new (std::nothrow) vgpu_scene();
    if (!s) return fail(VGPU_ERR_NOMEM, "out of host memory");
    s->device = device;
    DeviceGuard dg(device);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0) s->num_cus = cus;
    if (hipMalloc((void**)&s->d_err, sizeof(int32_t)) != hipSuccess || hipMemset(s->d_err, 0, sizeof(int32_t)) != hipSuccess) {
        delete s;
        return fail(VGPU_ERR_HIP, "cannot allocate the error flag");
    }
    // Once per device: the address of vgpu_trace_noalloc's scratch array on this device.
    const hipError_t ne = vgpu::noalloc_init(device);
    if (ne != hipSuccess) {
        delete s;
        return fail(VGPU_ERR_HIP, "allocation-free traversal scratch on device %d: %s", device, hipGetErrorString(ne));
    }
    *out = s;
    return VGPU_OK;
}

int vgpu_scene_destroy(vgpu_scene* scene) {
    delete scene;
    return VGPU_OK;
}

int vgpu_scene_add_mesh(vgpu_scene* s, const float* verts, uint32_t nverts, uint32_t nkeys, const uint32_t* idx,
                        uint32_t ntris, float ray_bias, uint32_t flags, int32_t* geom_out) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    if (s->committed) return fail(VGPU_ERR_STATE, "scene already committed");
    if (!verts || !idx || nverts == 0 || ntris == 0 || nkeys == 0) return fail(VGPU_ERR_ARG, "empty mesh");
    if (ntris >= (1u << 23)) return fail(VGPU_ERR_ARG, "mesh has %u triangles; the leaf encoding holds < 2^23 (qbvh/qbvh.go:64)", ntris);
    for (uint64_t i = 0; i < (uint64_t)ntris * 3; i++)
        if (idx[i] >= nverts) return fail(VGPU_ERR_ARG, "index %u out of range at %llu", idx[i], (unsigned long long)i);
    // A NaN or infinite coordinate gives a centroid no SAH bin holds: the reference's BuildAccel
    // indexes out of range or recurses without end on it.
    for (uint64_t i = 0; i < (uint64_t)nverts * nkeys * 3; i++)
        if (!std::isfinite(verts[i]))
            return fail(VGPU_ERR_ARG, "vertex coordinate %llu is not finite", (unsigned long long)i);
    MeshHost m;
    m.nverts = nverts; m.nkeys = nkeys; m.ntris = ntris; m.ray_bias = ray_bias; m.flags = flags;
    m.verts.assign(verts, verts + (size_t)nverts * nkeys * 3);
    m.idx.assign(idx, idx + (size_t)ntris * 3);
    s->meshes.push_back(std::move(m));
    if (geom_out) *geom_out = (int32_t)s->meshes.size() - 1;
    return VGPU_OK;
}

int vgpu_scene_add_shaded_mesh(vgpu_scene* s, const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntris,
                               const float* normals, const uint8_t* shaderidx, const int32_t* shaders, uint32_t nshaders,
                               int32_t* geom_out) {
    if (!shaders || nshaders == 0) return fail(VGPU_ERR_ARG, "a shaded mesh needs a shader table");
    for (uint32_t i = 0; i < nshaders; i++)
        if (shaders[i] < 0) return fail(VGPU_ERR_ARG, "shader %u has negative material id", i);
    if (shaderidx)
        for (uint32_t i = 0; i < ntris; i++)
            if (shaderidx[i] >= nshaders) return fail(VGPU_ERR_ARG, "triangle %u uses shader %u of %u", i, shaderidx[i], nshaders);
    int rc = vgpu_scene_add_mesh(s, verts, nverts, 1, idx, ntris, 0.0f, 0, geom_out);
    if (rc) return rc;
    MeshHost& m = s->meshes.back();
    if (normals) m.normals.assign(normals, normals + (size_t)nverts * 3);
    if (shaderidx) m.shaderidx.assign(shaderidx, shaderidx + ntris);
    m.shaders.assign(shaders, shaders + nshaders);
    return VGPU_OK;
}

int vgpu_scene_add_shaded_motion_mesh(vgpu_scene* s, const float* verts, uint32_t nverts, uint32_t nkeys,
                                      const uint32_t* idx, uint32_t ntris, const float* normals, uint32_t normal_keys,
                                      const uint8_t* shaderidx, const int32_t* shaders, uint32_t nshaders, uint32_t flags,
                                      int32_t* geom_out) {
    if (nkeys < 2) return fail(VGPU_ERR_ARG, "a moving mesh has at least 2 keys (vgpu_scene_add_shaded_mesh for static)");
    if ((flags & ~VGPU_MESH_MOTION_REFERENCE_INDEXING) != 0) return fail(VGPU_ERR_ARG, "unknown mesh flags 0x%x", flags);
    if (normals && normal_keys == 0) return fail(VGPU_ERR_ARG, "normals with 0 keys");
    if (!shaders || nshaders == 0) return fail(VGPU_ERR_ARG, "a shaded mesh needs a shader table");
    for (uint32_t i = 0; i < nshaders; i++)
        if (shaders[i] < 0) return fail(VGPU_ERR_ARG, "shader %u has negative material id", i);
    if (shaderidx && ntris)
        for (uint32_t i = 0; i < ntris; i++)
            if (shaderidx[i] >= nshaders) return fail(VGPU_ERR_ARG, "triangle %u uses shader %u of %u", i, shaderidx[i], nshaders);
    int rc = vgpu_scene_add_mesh(s, verts, nverts, nkeys, idx, ntris, 0.0f, flags, geom_out);
    if (rc) return rc;
    MeshHost& m = s->meshes.back();
    // trace.go:694: normals with another key count than the vertices are not used (N = Ng)
    if (normals && normal_keys == nkeys) m.mnormals.assign(normals, normals + (size_t)nkeys * ntris * 9);
    if (shaderidx) m.shaderidx.assign(shaderidx, shaderidx + ntris);
    m.shaders.assign(shaders, shaders + nshaders);
    return VGPU_OK;
}

int vgpu_scene_add_sphere(vgpu_scene* s, const float* centre, float radius, int32_t material, int32_t* geom_out) {
    if (!s || !centre) return fail(VGPU_ERR_ARG, "NULL argument");
    if (s->committed) return fail(VGPU_ERR_STATE, "scene already committed");
    if (!(radius > 0.0f) || !std::isfinite(radius)) return fail(VGPU_ERR_ARG, "sphere radius must be > 0");
    for (int k = 0; k < 3; k++) if (!std::isfinite(centre[k])) return fail(VGPU_ERR_ARG, "sphere centre must be finite");
    if (material < 0) return fail(VGPU_ERR_ARG, "sphere material must be >= 0");
    MeshHost m;
    m.sphere = true;
    m.built = true;
    for (int k = 0; k < 3; k++) m.sphere_c[k] = centre[k];
    m.sphere_r = radius;
    // Sphere.Bounds (builtin/geom/sphere/sphere.go:68-77)
    for (int k = 0; k < 3; k++) {
        m.bounds[k] = centre[k] - radius;
        m.bounds[3 + k] = centre[k] + radius;
    }
    m.shaders.assign(1, material);
    s->meshes.push_back(std::move(m));
    if (geom_out) *geom_out = (int32_t)s->meshes.size() - 1;
    return VGPU_OK;
}

namespace {
// Checks nkeys column-major matrices of a transform (finite; one key: invertible, as every ray inverts it).
int check_transform(const float* mats, uint32_t nkeys, bool allow_none) {
    if (nkeys == 0) return allow_none ? VGPU_OK : fail(VGPU_ERR_ARG, "a transform needs at least one key");
    if (!mats) return fail(VGPU_ERR_ARG, "transform matrices are NULL");
    if (nkeys > 64) return fail(VGPU_ERR_ARG, "at most 64 transform keys");
    for (size_t i = 0; i < (size_t)nkeys * 16; i++)
        if (!std::isfinite(mats[i])) return fail(VGPU_ERR_ARG, "transform matrices must be finite");
    float inv[16];
    if (nkeys == 1 && !vgpu::xf_inverse(mats, inv)) return fail(VGPU_ERR_ARG, "transform matrix is singular");
    return VGPU_OK;
}
int check_box(const float* bmin, const float* bmax) {
    if (!bmin || !bmax) return fail(VGPU_ERR_ARG, "bounds are NULL");
    for (int k = 0; k < 3; k++)
        if (!std::isfinite(bmin[k]) || !std::isfinite(bmax[k])) return fail(VGPU_ERR_ARG, "bounds must be finite");
    return VGPU_OK;
}
void set_user_box(MeshHost& m, const float* bmin, const float* bmax) {
    // BoundingBox.GrowVec3 of BMin[0] then BMax[0] from an empty box (instance.go:126-133)
    Box6 b;
    b.reset();
    b.grow_point(bmin);
    b.grow_point(bmax);
    std::memcpy(m.ubox, b.v, sizeof(m.ubox));
}
}  // namespace

int vgpu_scene_set_transform(vgpu_scene* s, int32_t geom, const float* mats, uint32_t nkeys) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    if (s->committed) return fail(VGPU_ERR_STATE, "scene already committed");
    if (geom < 0 || geom >= (int32_t)s->meshes.size()) return fail(VGPU_ERR_ARG, "geom %d out of range", geom);
    MeshHost& m = s->meshes[geom];
    if (m.kind != vgpu::kXfNone && m.kind != vgpu::kXfMesh) return fail(VGPU_ERR_ARG, "geom %d is an instance or proc", geom);
    if (m.sphere) return fail(VGPU_ERR_ARG, "geom %d is an analytic sphere (no Transform)", geom);
    int rc = check_transform(mats, nkeys, true);
    if (rc) return rc;
    m.xmats.assign(mats, mats + (size_t)nkeys * 16);
    m.kind = nkeys ? vgpu::kXfMesh : vgpu::kXfNone;
    return VGPU_OK;
}

int vgpu_scene_set_hidden(vgpu_scene* s, int32_t geom, int32_t hidden) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    if (s->committed) return fail(VGPU_ERR_STATE, "scene already committed");
    if (geom < 0 || geom >= (int32_t)s->meshes.size()) return fail(VGPU_ERR_ARG, "geom %d out of range", geom);
    s->meshes[geom].hidden = hidden != 0;
    return VGPU_OK;
}

namespace {
// GeomInstance (kind kXfInstance, one child) or Proc (kXfProc): nkeys transform keys (a proc with none
// holds the identity, proc.go:154-157), bounds the box of bmin, bmax.
int add_wrapper(vgpu_scene* s, int kind, const int32_t* children, uint32_t nchild, const float* mats, uint32_t nkeys,
                const float* bmin, const float* bmax, int32_t* geom_out) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    if (s->committed) return fail(VGPU_ERR_STATE, "scene already committed");
    if (nchild == 0 || !children) return fail(VGPU_ERR_ARG, "no geoms to hold");
    for (uint32_t i = 0; i < nchild; i++) {
        const int32_t c = children[i];
        if (c < 0 || c >= (int32_t)s->meshes.size()) return fail(VGPU_ERR_ARG, "geom %d out of range", c);
        if (s->meshes[c].kind >= vgpu::kXfInstance)
            return fail(VGPU_ERR_ARG, "geom %d is an instance or proc (they hold meshes and spheres only)", c);
    }
    int rc = check_transform(mats, nkeys, kind == vgpu::kXfProc);
    if (!rc) rc = check_box(bmin, bmax);
    if (rc) return rc;
    MeshHost m;
    m.kind = kind;
    m.built = true;
    m.children.assign(children, children + nchild);
    if (nkeys) m.xmats.assign(mats, mats + (size_t)nkeys * 16);
    else {
        m.xmats.assign(16, 0.0f);
        m.xmats[0] = m.xmats[5] = m.xmats[10] = m.xmats[15] = 1.0f;
    }
    set_user_box(m, bmin, bmax);
    s->meshes.push_back(std::move(m));
    if (geom_out) *geom_out = (int32_t)s->meshes.size() - 1;
    return VGPU_OK;
}
}  // namespace

int vgpu_scene_add_instance(vgpu_scene* s, int32_t geom, const float* mats, uint32_t nkeys, const float* bmin,
                            const float* bmax, int32_t* geom_out) {
    return add_wrapper(s, vgpu::kXfInstance, &geom, 1, mats, nkeys, bmin, bmax, geom_out);
}

int vgpu_scene_add_proc(vgpu_scene* s, const int32_t* geoms, uint32_t ngeoms, const float* mats, uint32_t nkeys,
                        const float* bmin, const float* bmax, int32_t* geom_out) {
    return add_wrapper(s, vgpu::kXfProc, geoms, ngeoms, mats, nkeys, bmin, bmax, geom_out);
}

int vgpu_scene_add_mesh_prebuilt(vgpu_scene* s, const float* verts, uint32_t nverts, const uint32_t* idxp_sorted,
                                 uint32_t ntris, const vgpu_qbvh_node* nodes, uint32_t nnodes, const float* bounds,
                                 float ray_bias, int32_t* geom_out) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    if (s->committed) return fail(VGPU_ERR_STATE, "scene already committed");
    if (!verts || !idxp_sorted || !nodes || !bounds || nverts == 0 || ntris == 0 || nnodes == 0)
        return fail(VGPU_ERR_ARG, "empty mesh");
    for (uint64_t i = 0; i < (uint64_t)ntris * 3; i++)
        if (idxp_sorted[i] >= nverts) return fail(VGPU_ERR_ARG, "index out of range");
    for (uint32_t i = 0; i < nnodes; i++)
        for (int c = 0; c < 4; c++) {
            const int32_t ch = nodes[i].children[c];
            if (ch >= (int32_t)nnodes) return fail(VGPU_ERR_ARG, "node %u child %d out of range", i, ch);
            // BuildAccel numbers nodes in depth-first pre-order, so a child always follows its parent;
            // requiring it rules out cycles, which would never terminate on the device
            if (ch >= 0 && ch <= (int32_t)i) return fail(VGPU_ERR_ARG, "node %u child %d does not follow its parent", i, ch);
            if (ch < -1) {
                const uint32_t b = (uint32_t)((ch & 0x7ffffff) >> 4), cnt = (uint32_t)((ch & 0xf) + 1);
                if (b + cnt > ntris) return fail(VGPU_ERR_ARG, "node %u leaf [%u,+%u) beyond %u triangles", i, b, cnt, ntris);
            }
        }
    MeshHost m;
    m.nverts = nverts; m.nkeys = 1; m.ntris = ntris; m.ray_bias = ray_bias;
    m.verts.assign(verts, verts + (size_t)nverts * 3);
    m.idxp_sorted.assign(idxp_sorted, idxp_sorted + (size_t)ntris * 3);
    m.nodes.assign(nodes, nodes + nnodes);
    std::memcpy(m.bounds, bounds, sizeof(m.bounds));
    m.built = true;
    s->meshes.push_back(std::move(m));
    if (geom_out) *geom_out = (int32_t)s->meshes.size() - 1;
    return VGPU_OK;
}

int vgpu_scene_add_motion_mesh_prebuilt(vgpu_scene* s, const float* verts, uint32_t nverts, uint32_t nkeys,
                                        const uint32_t* idxp, uint32_t ntris, const int32_t* accel_idx,
                                        const vgpu_mqbvh_node* nodes, uint32_t nnodes, const float* boxes,
                                        float ray_bias, uint32_t flags, int32_t* geom_out) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    if (s->committed) return fail(VGPU_ERR_STATE, "scene already committed");
    if (nkeys < 2) return fail(VGPU_ERR_ARG, "a moving mesh has at least 2 keys (vgpu_scene_add_mesh_prebuilt for static)");
    if ((flags & ~VGPU_MESH_MOTION_REFERENCE_INDEXING) != 0) return fail(VGPU_ERR_ARG, "unknown mesh flags 0x%x", flags);
    if (!accel_idx || !nodes || !boxes || nnodes == 0 || ntris == 0) return fail(VGPU_ERR_ARG, "empty accel");
    for (uint32_t i = 0; i < ntris; i++)
        if (accel_idx[i] < 0 || (uint32_t)accel_idx[i] >= ntris)
            return fail(VGPU_ERR_ARG, "accel index %d at %u out of range", accel_idx[i], i);
    for (uint32_t i = 0; i < nnodes; i++)
        for (int c = 0; c < 4; c++) {
            const int32_t ch = nodes[i].children[c];
            if (ch >= (int32_t)nnodes) return fail(VGPU_ERR_ARG, "node %u child %d out of range", i, ch);
            // pre-order numbering (BuildAccelMotion's recursion): a child follows its parent, so no cycles
            if (ch >= 0 && ch <= (int32_t)i) return fail(VGPU_ERR_ARG, "node %u child %d does not follow its parent", i, ch);
            if (ch < -1) {
                const uint32_t b = (uint32_t)((ch & 0x7ffffff) >> 4), cnt = (uint32_t)((ch & 0xf) + 1);
                if (b + cnt > ntris) return fail(VGPU_ERR_ARG, "node %u leaf [%u,+%u) beyond %u triangles", i, b, cnt, ntris);
            }
        }
    for (uint64_t i = 0; i < (uint64_t)nkeys * nnodes * 24; i++)
        if (std::isnan(boxes[i])) return fail(VGPU_ERR_ARG, "motion box value %llu is NaN", (unsigned long long)i);
    const int rc = vgpu_scene_add_mesh(s, verts, nverts, nkeys, idxp, ntris, ray_bias, flags, geom_out);
    if (rc) return rc;
    MeshHost& m = s->meshes.back();
    m.mnodes.assign(nodes, nodes + nnodes);
    m.mboxes.assign(boxes, boxes + (size_t)nkeys * nnodes * 24);
    m.perm.assign(accel_idx, accel_idx + ntris);
    // mesh.motionBounds[k] is the root's box in key k and mesh.bounds their union (initMotionBoxes,
    // buildqbvh.go:151-169): the stored child boxes of node 0 grown in child order
    m.motion_bounds.assign((size_t)nkeys * 6, 0.0f);
    Box6 full; full.reset();
    for (uint32_t k = 0; k < nkeys; k++) {
        Box6 rb; rb.reset();
        const float* nb = &m.mboxes[(size_t)k * nnodes * 24];
        for (int c = 0; c < 4; c++) {
            if (nodes[0].children[c] == -1) continue;
            float cb[6];
            for (int a = 0; a < 3; a++) { cb[a] = nb[c + a * 4]; cb[3 + a] = nb[c + 12 + a * 4]; }
            rb.grow_box(cb);
        }
        std::memcpy(&m.motion_bounds[(size_t)k * 6], rb.v, sizeof(rb.v));
        full.grow_box(rb.v);
    }
    std::memcpy(m.bounds, full.v, sizeof(full.v));
    m.built = true;
    return VGPU_OK;
}

int vgpu_scene_commit(vgpu_scene* s) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    if (s->committed) return fail(VGPU_ERR_STATE, "scene already committed");
    if (s->meshes.empty()) return fail(VGPU_ERR_STATE, "scene has no geometry");
    for (auto& m : s->meshes) {
        if (m.built) continue;
        int rc = m.nkeys == 1 ? build_static_mesh(m, s->leaf_max, s->device) : build_motion_mesh(m, s->device);
        if (rc) return rc;
    }
    // Scene.initAccel (scene.go:135-203) over the geoms the scene holds itself; hidden geoms (reached only
    // through instances and procs) take the slots after them, in insertion order.
    const int nall = (int)s->meshes.size();
    std::vector<int32_t> vis;
    for (int i = 0; i < nall; i++) if (!s->meshes[i].hidden) vis.push_back(i);
    if (vis.empty()) return fail(VGPU_ERR_STATE, "every geom is hidden");
    const int ng = (int)vis.size();
    // Geom.MotionKeys: a mesh's vertex keys (polymesh.go:102-112), an instance's geom's
    // (instance.go:151-156), a proc's transform keys (proc.go:188-193)
    auto motion_keys_of = [&](int i) -> int {
        const MeshHost& m = s->meshes[i];
        if (m.kind == vgpu::kXfInstance) return (int)s->meshes[m.children[0]].nkeys;
        if (m.kind == vgpu::kXfProc) return (int)(m.xmats.size() / 16);
        return (int)m.nkeys;
    };
    int maxKeys = 0;
    for (int i : vis) maxKeys = std::max(maxKeys, motion_keys_of(i));
    std::vector<float> boxes((size_t)ng * 6), cent((size_t)ng * 3);
    std::vector<int32_t> ids(ng);
    const float t_build = maxKeys == 1 ? 0.0f : 0.5f;
    for (int i = 0; i < ng; i++) {
        float b[6];
        s->meshes[vis[i]].bounds_at(t_build, b);
        std::memcpy(&boxes[(size_t)i * 6], b, sizeof(b));
        for (int k = 0; k < 3; k++) cent[(size_t)i * 3 + k] = (b[3 + k] + b[k]) * 0.5f;   // BoundingBox.Centroid
        ids[i] = i;
    }
    {
        Box6 all;
        all.reset();
        for (int i : vis) {
            const MeshHost& m = s->meshes[i];
            float b[6];
            if (m.kind >= vgpu::kXfInstance || m.nkeys == 1) { m.bounds_at(0.0f, b); all.grow_box(b); }
            else for (uint32_t k = 0; k < m.nkeys; k++) all.grow_box(&m.motion_bounds[(size_t)k * 6]);
        }
        std::memcpy(s->box, all.v, sizeof(s->box));
    }
    int32_t nn = 0;
    if (maxKeys == 1) {
        s->top_nodes.resize(std::max(ng, 1));
        int rc = vgpu::build_qbvh(boxes.data(), cent.data(), ids.data(), ng, 1, s->top_nodes.data(),
                                  (int32_t)s->top_nodes.size(), &nn, nullptr, 1);
        if (rc) return fail(VGPU_ERR_BUILD, "scene top-level build failed (coincident geom centroids?)");
        s->top_nodes.resize(nn);
        s->top_motion = false;
        s->top_keys = 1;
    } else {
        s->top_mnodes.resize(std::max(ng, 1));
        int rc = vgpu::build_qbvh_motion(boxes.data(), cent.data(), ids.data(), ng, 1, s->top_mnodes.data(),
                                         (int32_t)s->top_mnodes.size(), &nn, 1);
        if (rc) return fail(VGPU_ERR_BUILD, "scene top-level motion build failed");
        s->top_mnodes.resize(nn);
        s->top_motion = true;
        s->top_keys = maxKeys;
    }
    s->order.resize(nall);
    for (int i = 0; i < ng; i++) s->order[i] = vis[ids[i]];
    {
        int slot = ng;
        for (int i = 0; i < nall; i++) if (s->meshes[i].hidden) s->order[slot++] = i;
    }
    if (s->top_motion) {
        // scene.go:207-268
        s->top_mboxes.assign((size_t)maxKeys * nn * 24, 0.0f);
        for (int kk = 0; kk < maxKeys; kk++) {
            auto leaf = [&](int lb, int lc, Box6& b) {
                for (int i = lb; i < lb + lc; i++) {
                    const float time = (float)kk / (float)(maxKeys - 1);
                    float gb[6];
                    s->meshes[s->order[i]].bounds_at(time, gb);
                    b.grow_box(gb);
                }
            };
            Box6 rb;
            motion_boxes_rec(s->top_mnodes, 0, &s->top_mboxes[(size_t)kk * nn * 24], leaf, rb);
        }
    }
    s->any_motion = maxKeys > 1;
    // Default launch shape unless the caller chose one: static scenes run best at block 128 with a
    // 12-entry LDS stack and 24 waves per CU; the motion kernel (about 120 VGPRs, four waves per
    // SIMD) at block 256, 16 entries, 16 waves per CU.
    if (!s->cfg_set) s->cfg = s->any_motion ? vgpu_trace_config{64, 8, 20, 0, 0} : vgpu_trace_config{128, 12, 24, 0, 0};

    // Flatten into the device layout.
    std::vector<vgpu::DevNode> dnodes;
    std::vector<float4> dtris, dmnodes, dmtris, dspheres;
    std::vector<vgpu::DevGeom> dgeoms(nall);
    {
        size_t corners = 0;
        for (int slot = 0; slot < nall; slot++) {
            const MeshHost& m = s->meshes[s->order[slot]];
            if (m.kind < vgpu::kXfInstance && !m.sphere && m.nkeys == 1) corners += (size_t)m.ntris * 3;
        }
        dtris.reserve(corners);
    }
    if (!s->top_motion) {
        for (auto& nd : s->top_nodes) { vgpu::DevNode d; to_dev(nd, d); dnodes.push_back(d); }
    } else {
        s->view.top_mbase = 0;
        append_motion_nodes(dmnodes, s->top_mnodes, s->top_mboxes, maxKeys);
    }
    for (int slot = 0; slot < nall; slot++) {
        MeshHost& m = s->meshes[s->order[slot]];
        vgpu::DevGeom& g = dgeoms[slot];
        g.nkeys = (int32_t)m.nkeys;
        g.ray_bias = m.ray_bias;
        if (m.kind >= vgpu::kXfInstance) {   // instance or proc: nothing of its own to trace
            g.nkeys = -1;
            g.node_base = -1;
            g.tri_base = -1;
        } else if (m.sphere) {   // nkeys 0 marks an analytic sphere; tri_base indexes the sphere table
            g.nkeys = 0;
            g.node_base = -1;
            g.tri_base = (int32_t)dspheres.size();
            dspheres.push_back(make_float4(m.sphere_c[0], m.sphere_c[1], m.sphere_c[2], m.sphere_r));
        } else if (m.nkeys == 1) {
            g.node_base = (int32_t)dnodes.size();
            for (auto& nd : m.nodes) { vgpu::DevNode d; to_dev(nd, d); dnodes.push_back(d); }
            g.tri_base = (int32_t)(dtris.size() / 3);
            const size_t base = dtris.size();
            dtris.resize(base + (size_t)m.ntris * 3);
            float4* out = dtris.data() + base;
            parallel_ranges((size_t)m.ntris * 3, [&](size_t a, size_t b) {
                for (size_t k = a; k < b; k++) {
                    const float* v = m.vert(0, m.idxp_sorted[k]);
                    out[k] = make_float4(v[0], v[1], v[2], 0.0f);
                }
            });
        } else {
            g.node_base = (int32_t)dmnodes.size();
            append_motion_nodes(dmnodes, m.mnodes, m.mboxes, (int)m.nkeys);
            g.tri_base = (int32_t)dmtris.size();
            const bool ref_idx = (m.flags & VGPU_MESH_MOTION_REFERENCE_INDEXING) != 0;
            for (uint32_t i = 0; i < m.ntris; i++) {
                const uint32_t f = ref_idx ? i : (uint32_t)m.perm[i];
                for (uint32_t kk = 0; kk < m.nkeys; kk++)
                    for (int j = 0; j < 3; j++) {
                        const float* v = m.vert(kk, m.idx[(size_t)f * 3 + j]);
                        dmtris.push_back(make_float4(v[0], v[1], v[2], 0.0f));
                    }
            }
        }
        if (dtris.size() / 3 > 0x7fffffffull || dmtris.size() > 0x7fffffffull || dmnodes.size() > 0x7fffffffull)
            return fail(VGPU_ERR_ARG, "scene too large for 32-bit offsets");
    }
    DeviceGuard dg(s->device);
    int rc;
    if ((rc = upload(&s->d_nodes, dnodes))) return rc;
    if ((rc = upload(&s->d_tris, dtris))) return rc;
    if ((rc = upload(&s->d_mnodes, dmnodes))) return rc;
    if ((rc = upload(&s->d_mtris, dmtris))) return rc;
    if ((rc = upload(&s->d_geoms, dgeoms))) return rc;
    if ((rc = upload(&s->d_spheres, dspheres))) return rc;
    // Shading data per slot: per-triangle vertex normals (the mesh's normalidx equals its idxp,
    // init.go:85-91), per-triangle shader index, shader tables.
    {
        std::vector<float4> dn;
        std::vector<uint8_t> dsi;
        std::vector<int32_t> dsh;
        std::vector<vgpu::DevGeomShade> dgs(nall);
        for (int slot = 0; slot < nall; slot++) {
            MeshHost& m = s->meshes[s->order[slot]];
            vgpu::DevGeomShade& gsd = dgs[slot];
            gsd.nrm_base = -1;
            gsd.sidx_base = -1;
            if (!m.normals.empty() && m.nkeys == 1) {
                gsd.nrm_base = (int32_t)dn.size();
                const size_t base = dn.size();
                dn.resize(base + (size_t)m.ntris * 3);
                float4* out = dn.data() + base;
                parallel_ranges((size_t)m.ntris * 3, [&](size_t a, size_t b) {
                    for (size_t k = a; k < b; k++) {
                        const float* nv = &m.normals[(size_t)m.idxp_sorted[k] * 3];
                        out[k] = make_float4(nv[0], nv[1], nv[2], 0.0f);
                    }
                });
            }
            if (m.nkeys > 1 && !m.sphere) {
                // Moving mesh: leaf slot i shades face f = accel.idx[i], or f = i with the reference's
                // indexing (trace.go:527-545 reads idxp, shaderidx and Normals at the slot).  Normals
                // per slot: nkeys groups of the face's 3 corners, as the motion triangles are laid out.
                const bool ref_idx = (m.flags & VGPU_MESH_MOTION_REFERENCE_INDEXING) != 0;
                if (!m.mnormals.empty()) {
                    gsd.nrm_base = (int32_t)dn.size();
                    for (uint32_t i = 0; i < m.ntris; i++) {
                        const uint32_t f = ref_idx ? i : (uint32_t)m.perm[i];
                        for (uint32_t kk = 0; kk < m.nkeys; kk++)
                            for (int j = 0; j < 3; j++) {
                                const float* nv = &m.mnormals[(((size_t)kk * m.ntris + f) * 3 + j) * 3];
                                dn.push_back(make_float4(nv[0], nv[1], nv[2], 0.0f));
                            }
                    }
                }
                if (!m.shaderidx.empty()) {
                    gsd.sidx_base = (int32_t)dsi.size();
                    for (uint32_t i = 0; i < m.ntris; i++) dsi.push_back(m.shaderidx[ref_idx ? i : (uint32_t)m.perm[i]]);
                }
            } else if (!m.shaderidx_sorted.empty()) {
                gsd.sidx_base = (int32_t)dsi.size();
                dsi.insert(dsi.end(), m.shaderidx_sorted.begin(), m.shaderidx_sorted.end());
            }
            gsd.shader_base = (int32_t)dsh.size();
            gsd.nshaders = (int32_t)m.shaders.size();
            dsh.insert(dsh.end(), m.shaders.begin(), m.shaders.end());
            for (int32_t mid : m.shaders) s->max_material = std::max(s->max_material, mid);
        }
        if (dn.size() > 0x7fffffffull) return fail(VGPU_ERR_ARG, "scene normals too large for 32-bit offsets");
        if (dsh.empty()) dsh.push_back(-1);
        if ((rc = upload(&s->d_normals, dn))) return rc;
        if ((rc = upload(&s->d_shaderidx, dsi))) return rc;
        if ((rc = upload(&s->d_shaders, dsh))) return rc;
        if ((rc = upload(&s->d_gshade, dgs))) return rc;
    }
    // Transforms per slot (xform.h records; one key precomputed with its inverse)
    {
        bool any = false;
        for (auto& m : s->meshes) any = any || m.kind != vgpu::kXfNone;
        s->view.any_xform = any ? 1 : 0;
        if (any) {
            std::vector<int32_t> slot_of(nall);
            for (int slot = 0; slot < nall; slot++) slot_of[s->order[slot]] = slot;
            std::vector<vgpu::DevXform> dx(nall);
            std::vector<float> keys;
            std::vector<int32_t> kids;
            for (int slot = 0; slot < nall; slot++) {
                const MeshHost& m = s->meshes[s->order[slot]];
                vgpu::DevXform& x = dx[slot];
                x = vgpu::DevXform{m.kind, 0, -1, -1, 0, 0, 0, 0};
                if (m.kind == vgpu::kXfNone) continue;
                const int nk = (int)(m.xmats.size() / 16);
                x.nkeys = nk;
                if (nk == 1) {
                    float M[16], Minv[16];
                    if (m.kind == vgpu::kXfMesh) std::memcpy(M, m.xmats.data(), sizeof(M));   // trace.go:45-47: as given
                    else {   // TimeKey returns the one decomposition, recomposed (instance.go:17-34)
                        float rec[vgpu::kXfKeyFloats];
                        decompose_transform(m.xmats.data(), rec);
                        vgpu::xf_decomp_matrix(rec, rec, 0.0f, M);
                    }
                    if (!vgpu::xf_inverse(M, Minv)) std::fill(Minv, Minv + 16, 0.0f);
                    x.mat_base = (int32_t)keys.size();
                    keys.insert(keys.end(), M, M + 16);
                    keys.insert(keys.end(), Minv, Minv + 16);
                } else {
                    x.key_base = (int32_t)keys.size();
                    for (int k = 0; k < nk; k++) {
                        float rec[vgpu::kXfKeyFloats];
                        decompose_transform(m.xmats.data() + 16 * k, rec);
                        keys.insert(keys.end(), rec, rec + vgpu::kXfKeyFloats);
                    }
                }
                if (m.kind >= vgpu::kXfInstance) {
                    x.child_base = (int32_t)kids.size();
                    x.nchild = (int32_t)m.children.size();
                    for (int32_t c : m.children) kids.push_back(slot_of[c]);
                }
            }
            if (kids.empty()) kids.push_back(-1);
            if ((rc = upload(&s->d_xforms, dx))) return rc;
            if ((rc = upload(&s->d_xkeys, keys))) return rc;
            if ((rc = upload(&s->d_xchildren, kids))) return rc;
        }
        s->view.xforms = s->d_xforms;
        s->view.xkeys = s->d_xkeys;
        s->view.xchildren = s->d_xchildren;
    }
    s->view.nodes = s->d_nodes;
    s->view.tris = s->d_tris;
    s->view.mnodes = s->d_mnodes;
    s->view.mtris = s->d_mtris;
    s->view.geoms = s->d_geoms;
    s->view.spheres = s->d_spheres;
    s->view.top_motion = s->top_motion ? 1 : 0;
    s->view.top_keys = s->top_keys;
    s->view.nan_free_boxes = 1;
    for (const vgpu::DevNode& d : dnodes) {
        float bx[24];
        std::memcpy(bx, d.lox, sizeof(bx));
        for (int k = 0; k < 24; k++)
            if (std::isnan(bx[k])) s->view.nan_free_boxes = 0;
    }
    s->committed = true;
    return VGPU_OK;
}

int vgpu_scene_info(const vgpu_scene* s, int32_t* ngeoms, int32_t* top_nodes, int32_t* geom_order, int32_t* geom_nodes,
                    int32_t* geom_tris) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    if (!s->committed) return fail(VGPU_ERR_STATE, "scene not committed");
    const int ng = (int)s->meshes.size();
    if (ngeoms) *ngeoms = ng;
    if (top_nodes) *top_nodes = s->top_motion ? (int32_t)s->top_mnodes.size() : (int32_t)s->top_nodes.size();
    for (int i = 0; i < ng; i++) {
        const MeshHost& m = s->meshes[s->order[i]];
        if (geom_order) geom_order[i] = s->order[i];
        if (geom_nodes) geom_nodes[i] = m.nkeys == 1 ? (int32_t)m.nodes.size() : (int32_t)m.mnodes.size();
        if (geom_tris) geom_tris[i] = (int32_t)m.ntris;
    }
    return VGPU_OK;
}

int vgpu_scene_mesh_accel(const vgpu_scene* s, int32_t slot, vgpu_qbvh_node* nodes, uint32_t* idxp_sorted, int32_t* perm) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    if (!s->committed) return fail(VGPU_ERR_STATE, "scene not committed");
    if (slot < 0 || slot >= (int32_t)s->meshes.size()) return fail(VGPU_ERR_ARG, "slot out of range");
    const MeshHost& m = s->meshes[s->order[slot]];
    if (m.sphere) return fail(VGPU_ERR_ARG, "slot %d is a sphere", slot);
    if (m.nkeys != 1) return fail(VGPU_ERR_ARG, "slot %d is a motion mesh", slot);
    if (nodes) std::memcpy(nodes, m.nodes.data(), m.nodes.size() * sizeof(vgpu_qbvh_node));
    if (idxp_sorted) std::memcpy(idxp_sorted, m.idxp_sorted.data(), m.idxp_sorted.size() * sizeof(uint32_t));
    if (perm) {
        if (m.perm.empty()) return fail(VGPU_ERR_STATE, "prebuilt mesh carries no permutation");
        std::memcpy(perm, m.perm.data(), m.perm.size() * sizeof(int32_t));
    }
    return VGPU_OK;
}

int vgpu_trace(vgpu_scene* s, uint32_t flags, const vgpu_ray* rays, uint64_t n, vgpu_hit* hits, int32_t* geoms,
               int32_t* counters, void* stream) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    if (!s->committed) return fail(VGPU_ERR_STATE, "trace before commit");
    if (n == 0) return VGPU_OK;
    if (!rays || !hits) return fail(VGPU_ERR_ARG, "rays/hits are NULL");
    std::lock_guard<std::recursive_mutex> lk(s->mu);
    DeviceGuard dg(s->device);
    hipStream_t st = (hipStream_t)stream;
    if (flags & VGPU_TRACE_DEVICE_PTRS)
        return enqueue_trace(s, flags, rays, n, hits, geoms, counters, st, s->cfg);

    // host buffers: stage through device memory owned by the scene
    const size_t rb = n * sizeof(vgpu_ray), hb = n * sizeof(vgpu_hit), gb = geoms ? n * 4 : 0, cb = counters ? n * 16 : 0;
    const size_t need = rb + hb + gb + cb;
    if (need > s->stage_bytes) {
        if (s->d_stage) { HIP_TRY(hipFree(s->d_stage)); s->d_stage = nullptr; s->stage_bytes = 0; }
        HIP_TRY(hipMalloc(&s->d_stage, need));
        s->stage_bytes = need;
    }
    char* base = (char*)s->d_stage;
    void* d_rays = base;
    void* d_hits = base + rb;
    int32_t* d_geoms = geoms ? (int32_t*)(base + rb + hb) : nullptr;
    int32_t* d_cnt = counters ? (int32_t*)(base + rb + hb + gb) : nullptr;
    HIP_TRY(hipMemsetAsync(s->d_err, 0, sizeof(int32_t), st));
    HIP_TRY(hipMemcpyAsync(d_rays, rays, rb, hipMemcpyHostToDevice, st));
    int rc = enqueue_trace(s, flags, d_rays, n, d_hits, d_geoms, d_cnt, st, s->cfg);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(hits, d_hits, hb, hipMemcpyDeviceToHost, st));
    if (geoms) HIP_TRY(hipMemcpyAsync(geoms, d_geoms, gb, hipMemcpyDeviceToHost, st));
    if (counters) HIP_TRY(hipMemcpyAsync(counters, d_cnt, cb, hipMemcpyDeviceToHost, st));
    int32_t err = 0;
    HIP_TRY(hipMemcpyAsync(&err, s->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (err) return fail(VGPU_ERR_STACK, "a ray needed more than %d traversal stack entries", VGPU_MAX_STACK);
    return VGPU_OK;
}

int vgpu_trace_noalloc(vgpu_scene* s, uint32_t flags, const vgpu_ray* d_rays, uint64_t n, vgpu_hit* d_hits,
                       int32_t* d_geoms, int32_t* d_counters, int32_t* d_status, void* stream) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    if (!s->committed) return fail(VGPU_ERR_STATE, "trace before commit");
    if ((flags & ~(VGPU_TRACE_SHADOW | VGPU_TRACE_DEVICE_PTRS)) != 0) return fail(VGPU_ERR_ARG, "unknown trace flags 0x%x", flags);
    if (n > 0xffffffffull) return fail(VGPU_ERR_ARG, "%llu rays: one call traces at most 2^32-1", (unsigned long long)n);
    if (n > 0 && (!d_rays || !d_hits)) return fail(VGPU_ERR_ARG, "rays/hits are NULL");
    std::lock_guard<std::recursive_mutex> lk(s->mu);
    DeviceGuard dg(s->device);
    // The scene's launch shape with fewer resident waves where its spill columns would not fit the
    // scratch array; never sorted (the sort's buffers grow with n).
    vgpu::TraceLaunch L{s->cfg.block, s->cfg.lds_stack, s->cfg.waves_per_cu,
                        s->cfg.refill > 0 ? s->cfg.refill : vgpu::kRefillDefault};
    const uint64_t cap = vgpu::noalloc_max_threads(L.lds_depth);
    while (L.waves_per_cu > 1 && vgpu::trace_grid_threads(L, n, s->num_cus) > cap) L.waves_per_cu--;
    if (vgpu::trace_grid_threads(L, n, s->num_cus) > cap)
        return fail(VGPU_ERR_ARG, "the %llu-byte scratch holds the stack spill of fewer than one wavefront per CU on %d CUs",
                    (unsigned long long)vgpu::kNoAllocScratchBytes, s->num_cus);
    const hipError_t e = vgpu::launch_trace_noalloc(s->device, s->view, (flags & VGPU_TRACE_SHADOW) != 0, s->any_motion, L,
                                                    d_rays, n, d_hits, d_geoms, d_counters, d_status, s->num_cus,
                                                    (hipStream_t)stream);
    if (e != hipSuccess) return fail(VGPU_ERR_HIP, "allocation-free trace launch: %s", hipGetErrorString(e));
    return VGPU_OK;
}

int vgpu_scene_set_leaf_size(vgpu_scene* s, int32_t leaf_max) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    if (s->committed) return fail(VGPU_ERR_STATE, "leaf size must be set before commit");
    if (leaf_max < 1 || leaf_max > 16) return fail(VGPU_ERR_ARG, "leaf size %d outside [1,16]", leaf_max);
    s->leaf_max = leaf_max;
    return VGPU_OK;
}

int vgpu_scene_set_timing(vgpu_scene* s, int32_t enable) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    DeviceGuard dg(s->device);
    HIP_TRY(hipDeviceSynchronize());
    s->timing = enable != 0;
    s->ev_used = 0;
    return VGPU_OK;
}

int vgpu_scene_timing(vgpu_scene* s, float* trace_ms, float* sort_ms, int32_t* launches) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    DeviceGuard dg(s->device);
    double tt = 0.0, ts = 0.0;
    const size_t nl = s->ev_used / 3;
    for (size_t i = 0; i < nl; i++) {
        hipEvent_t* e = &s->ev_pool[3 * i];
        HIP_TRY(hipEventSynchronize(e[2]));
        float a = 0.0f, b = 0.0f;
        HIP_TRY(hipEventElapsedTime(&a, e[0], e[1]));
        HIP_TRY(hipEventElapsedTime(&b, e[1], e[2]));
        ts += a;
        tt += b;
    }
    s->ev_used = 0;
    if (trace_ms) *trace_ms = (float)tt;
    if (sort_ms) *sort_ms = (float)ts;
    if (launches) *launches = (int32_t)nl;
    return VGPU_OK;
}

int vgpu_scene_status(vgpu_scene* s, void* stream) {
    if (!s) return fail(VGPU_ERR_ARG, "scene is NULL");
    std::lock_guard<std::recursive_mutex> lk(s->mu);
    DeviceGuard dg(s->device);
    hipStream_t st = (hipStream_t)stream;
    int32_t err = 0;
    HIP_TRY(hipMemcpyAsync(&err, s->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipMemsetAsync(s->d_err, 0, sizeof(int32_t), st));
    HIP_TRY(hipStreamSynchronize(st));
    if (err) return fail(VGPU_ERR_STACK, "a ray needed more than %d traversal stack entries", VGPU_MAX_STACK);
    return VGPU_OK;
}

int vgpu_scene_set_trace_config(vgpu_scene* s, const vgpu_trace_config* cfg) {
    if (!s || !cfg) return fail(VGPU_ERR_ARG, "NULL argument");
    int rc = check_config(*cfg);
    if (rc) return rc;
    std::lock_guard<std::recursive_mutex> lk(s->mu);
    s->cfg = *cfg;
    s->cfg_set = true;
    return VGPU_OK;
}

int vgpu_scene_get_trace_config(const vgpu_scene* s, vgpu_trace_config* cfg) {
    if (!s || !cfg) return fail(VGPU_ERR_ARG, "NULL argument");
    *cfg = s->cfg;
    return VGPU_OK;
}

int vgpu_scene_autotune(vgpu_scene* s, uint32_t flags, const vgpu_ray* d_rays, uint64_t n, vgpu_hit* d_hits, void* stream,
                        vgpu_trace_config* best_out, float* ms_out) {
    if (!s || !d_rays || !d_hits) return fail(VGPU_ERR_ARG, "NULL argument");
    if (!s->committed) return fail(VGPU_ERR_STATE, "autotune before commit");
    std::lock_guard<std::recursive_mutex> lk(s->mu);
    DeviceGuard dg(s->device);
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    vgpu_trace_config bc = s->cfg;
    // one warm launch, then the mean of two timed ones; returns a status code
    auto time_config = [&](const vgpu_trace_config& c) -> int {
        int rc = enqueue_trace(s, flags | VGPU_TRACE_DEVICE_PTRS, d_rays, n, d_hits, nullptr, nullptr, st, c);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(e0, st));
        for (int rep = 0; rep < 2 && !rc; rep++)
            rc = enqueue_trace(s, flags | VGPU_TRACE_DEVICE_PTRS, d_rays, n, d_hits, nullptr, nullptr, st, c);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(e1, st));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        ms *= 0.5f;
        if (ms < best) { best = ms; bc = c; }
        return VGPU_OK;
    };
    int rc = VGPU_OK;
    const int sorts[] = {0, 4, 5, 6}, blocks[] = {64, 128, 256}, ldss[] = {8, 12, 16, 24, 32}, waves[] = {8, 16, 20, 24, 32};
    for (int so : sorts)
        for (int b : blocks)
            for (int l : ldss)
                for (int w : waves) {
                    const vgpu_trace_config c{b, l, w, so, 0};
                    if (rc || check_config(c)) continue;
                    rc = time_config(c);
                }
    // second stage: the refill threshold in the fastest shape
    const int refills[] = {16, 32, 40, 48, 56, 64};
    const vgpu_trace_config shape = bc;
    for (int r : refills) {
        vgpu_trace_config c = shape;
        c.refill = r;
        if (!rc) rc = time_config(c);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (bc.refill == 0) bc.refill = vgpu::kRefillDefault;   // report the threshold the shape ran with
    vgpu::clear_error();
    s->cfg = bc;
    s->cfg_set = true;
    if (best_out) *best_out = bc;
    if (ms_out) *ms_out = best;
    return vgpu_scene_status(s, stream);
}

}  // extern "C"
