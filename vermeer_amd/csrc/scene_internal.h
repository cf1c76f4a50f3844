// Scene internals shared by the C-ABI translation units (scene.cpp, render.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vermeer_gpu.h"
#include "qbvh_build.h"
#include "render_kernels.h"
#include "trace_kernels.h"

#define HIP_TRY(expr)                                                                                       \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) return vgpu::fail(VGPU_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

namespace vgpu {

// Records the message for vgpu_last_error() and returns code.
int fail(int code, const char* fmt, ...);
void clear_error();

}  // namespace vgpu

namespace vgpu_detail {

using vgpu::fail;

const float kInf = std::numeric_limits<float>::infinity();
inline float mins(float x, float y) { return (x < y) ? x : y; }
inline float maxs(float x, float y) { return (x > y) ? x : y; }

struct Box6 {
    float v[6];   // lo xyz, hi xyz
    void reset() { for (int k = 0; k < 3; k++) { v[k] = kInf; v[3 + k] = -kInf; } }
    void grow_point(const float* p) { for (int k = 0; k < 3; k++) { v[k] = mins(p[k], v[k]); v[3 + k] = maxs(p[k], v[3 + k]); } }
    void grow_box(const float* b) { for (int k = 0; k < 3; k++) { v[k] = mins(v[k], b[k]); v[3 + k] = maxs(v[3 + k], b[3 + k]); } }
};

// math/vec3.go:52-58
inline void lerp3(const float* a, const float* b, float t, float* o) {
    for (int k = 0; k < 3; k++) o[k] = (1.0f - t) * a[k] + t * b[k];
}

struct MeshHost {
    uint32_t nverts = 0, nkeys = 1, ntris = 0;
    std::vector<float> verts;          // nkeys * nverts * 3
    std::vector<uint32_t> idx;         // as given (static prebuilt: already post-sort)
    float ray_bias = 0.0f;
    uint32_t flags = 0;
    bool built = false;
    std::vector<vgpu_qbvh_node> nodes;
    std::vector<uint32_t> idxp_sorted;
    std::vector<int32_t> perm;
    std::vector<vgpu_mqbvh_node> mnodes;
    std::vector<float> mboxes;         // nkeys * nmnodes * 24
    float bounds[6];
    std::vector<float> motion_bounds;  // nkeys * 6
    std::vector<float> normals;        // nverts * 3 (shaded meshes), empty: N = Ng
    std::vector<float> mnormals;       // moving shaded meshes: nkeys * ntris * 3 corners * 3, empty: N = Ng
    std::vector<uint8_t> shaderidx;    // per triangle (insertion order; reordered at build)
    std::vector<uint8_t> shaderidx_sorted;
    std::vector<int32_t> shaders;      // shader table: material ids
    bool sphere = false;               // analytic sphere geom (vgpu_scene_add_sphere): sphere_c, sphere_r
    float sphere_c[3] = {0, 0, 0}, sphere_r = 0;
    // Transforms: a mesh's own Transform (kind kXfNone/kXfMesh), or a GeomInstance / Proc (kXfInstance,
    // kXfProc) over `children` (insertion indices) with bounds ubox (lo xyz, hi xyz).
    int kind = 0;
    std::vector<float> xmats;          // transform keys, 16 floats each (column-major)
    std::vector<int32_t> children;
    float ubox[6] = {0, 0, 0, 0, 0, 0};
    bool hidden = false;               // left out of the scene's top level

    const float* vert(uint32_t key, uint32_t i) const { return &verts[((size_t)key * nverts + i) * 3]; }

    // PolyMesh.Bounds (builtin/geom/polymesh/bounds.go:159-186; a mesh's Transform is ignored there,
    // bounds.go:26); an instance's or proc's given box whatever the time (instance.go:63-67)
    void bounds_at(float time, float* out) const {
        if (kind >= 2) { std::memcpy(out, ubox, sizeof(ubox)); return; }
        if (nkeys == 1) { std::memcpy(out, bounds, sizeof(bounds)); return; }
        const float k = time * (float)(nkeys - 1);
        const float t = k - std::floor(k);
        const int key = (int)std::floor(k), key2 = (int)std::ceil(k);
        for (int i = 0; i < 6; i++) out[i] = (1 - t) * motion_bounds[key * 6 + i] + t * motion_bounds[key2 * 6 + i];
    }
};

}  // namespace vgpu_detail

struct vgpu_scene {
    int device = 0;
    int num_cus = 256;
    std::vector<vgpu_detail::MeshHost> meshes;      // insertion order
    std::vector<int32_t> order;        // post-sort slot -> insertion index
    bool committed = false;
    bool top_motion = false;
    int top_keys = 1;
    std::vector<vgpu_qbvh_node> top_nodes;
    std::vector<vgpu_mqbvh_node> top_mnodes;
    std::vector<float> top_mboxes;
    bool any_motion = false;

    vgpu::DevNode* d_nodes = nullptr;
    float4* d_tris = nullptr;
    float4* d_mnodes = nullptr;
    float4* d_mtris = nullptr;
    vgpu::DevGeom* d_geoms = nullptr;
    float4* d_spheres = nullptr;
    int32_t* d_err = nullptr;
    float* d_spT = nullptr;
    int32_t* d_spN = nullptr;
    uint64_t spill_threads = 0;
    void* d_stage = nullptr;
    size_t stage_bytes = 0;
    vgpu::SceneView view{};
    // shading data (vgpu_scene_add_shaded_mesh)
    float4* d_normals = nullptr;
    uint8_t* d_shaderidx = nullptr;
    int32_t* d_shaders = nullptr;
    vgpu::DevGeomShade* d_gshade = nullptr;
    vgpu::DevXform* d_xforms = nullptr;  // transforms (scenes with instances, procs or mesh transforms)
    float* d_xkeys = nullptr;
    int32_t* d_xchildren = nullptr;
    int32_t max_material = -1;
    vgpu_trace_config cfg{256, 16, 16, 0, 0};
    bool cfg_set = false;               // the caller chose cfg (set_trace_config or autotune)
    int32_t leaf_max = 16;              // leafMax of the static meshes' BuildAccel (buildqbvh.go:32: 16)
    float box[6] = {0, 0, 0, 0, 0, 0};   // scene bounds over all keys (ray-sort grid)
    vgpu::RaySortBuffers sort{};
    // kernel timing (vgpu_scene_set_timing): events [sort start, trace start, trace end] per launch
    bool timing = false;
    // Serialises the calls that share this scene's device buffers (host staging, spill area,
    // ray-sort scratch, error flag) and may reallocate them.  Recursive: autotune ends in status.
    std::recursive_mutex mu;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;

    ~vgpu_scene() {
        int prev = 0;
        if (hipGetDevice(&prev) == hipSuccess) (void)hipSetDevice(device);
        void* ptrs[] = {d_nodes, d_tris, d_mnodes, d_mtris, d_geoms, d_spheres, d_err, d_spT, d_spN, d_stage, d_normals, d_shaderidx,
                        d_shaders, d_gshade, sort.keys, sort.perm, sort.rays, sort.hist, sort.bsum,
                        d_xforms, d_xkeys, d_xchildren};
        for (void* p : ptrs) if (p) (void)hipFree(p);
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        (void)hipSetDevice(prev);
    }
};


namespace vgpu_detail {

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) { if (hipGetDevice(&prev) != hipSuccess) prev = -1; (void)hipSetDevice(dev); }
    ~DeviceGuard() { if (prev >= 0) (void)hipSetDevice(prev); }
};

// Next free timing event of vgpu_scene_set_timing (the pool grows; vgpu_scene_timing recycles it);
// NULL when the runtime cannot create one.
hipEvent_t timing_event(vgpu_scene* s);

// Ensures the spill area covers `threads` traversal threads.
int ensure_spill(vgpu_scene* s, uint64_t threads);
// Host math of the camera (render.cpp), in the reference's float32 operation order.
void vec3_normalize(const float* a, float* out);
void vec3_cross(const float* a, const float* b, float* o);
// LocalToWorld keys of a LookAt camera (builtin/camera/camera.go:109-186), 16 floats each; returns the count.
int lookat_keys(const float* from, int nfrom, const float* to, int nto, const float* up, const float* roll, int nroll,
                std::vector<float>& out);
bool mat4_inverse(const float* m, float* out);
// TransformDecompMatrix4 (math/animdecomp.go:20-63) of a column-major matrix: one kXfKeyFloats record.
void decompose_transform(const float* m, float* rec);

// Ensures the ray-sort buffers hold n rays.
int ensure_sort(vgpu_scene* s, uint64_t n);

// A sequence of launches replayed as one HIP graph.  run() captures `body` on a private stream
// (thread-local capture mode: body may only enqueue work on the stream it is given, no allocation,
// no synchronisation, no hipMemsetAsync), then updates the executable graph in place when only
// kernel arguments changed since the last capture (re-instantiating it when the launch sequence
// differs), and launches it on `st`.
struct GraphCache {
    hipStream_t cap = nullptr;
    hipGraphExec_t exec = nullptr;
    int instantiations = 0;
    GraphCache() = default;
    GraphCache(const GraphCache&) = delete;
    GraphCache& operator=(const GraphCache&) = delete;
    ~GraphCache();
    int run(hipStream_t st, const std::function<int(hipStream_t)>& body);
};
// Enqueues one traversal (device pointers) with configuration cfg.  n_dev (nullable): device-resident
// count of rays to trace, at most n.
int enqueue_trace(vgpu_scene* s, uint32_t flags, const void* d_rays, uint64_t n, void* d_hits, int32_t* d_geoms,
                  int32_t* d_counters, hipStream_t st, const vgpu_trace_config& cfg, const uint32_t* n_dev = nullptr);

}  // namespace vgpu_detail
