// Internal interface between the scene object (scene.cpp) and the
// traversal kernels (trace.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "xform.h"

namespace vgpu {

// Idle lanes that trigger a refill of the wavefront in the traversal kernel; 48 of 64 measured
// fastest on incoherent bounce rays (64: refill only when the whole wavefront is idle).
constexpr int kRefillDefault = 48;

// Device node, 128 bytes = one cache line.  Same content as qbvh.Node
// (qbvh/qbvh.go:31-36) re-laid for aligned 16-byte loads: the six box rows
// (min x/y/z, max x/y/z; lane c = child c) are the reference's Boxes[24]
// verbatim, followed by the children and the three split axes.
struct alignas(16) DevNode {
    float lox[4], loy[4], loz[4];
    float hix[4], hiy[4], hiz[4];
    int32_t children[4];   // reference encoding, node indices relative to the geom's node base
    int32_t axes[4];       // axis0, axis1, axis2, 0
};
static_assert(sizeof(DevNode) == 128, "DevNode is one 128-byte line");

// Motion node (qbvh.MotionNode + its per-key MotionNodeBoxes): a 32-byte
// topology record followed by nkeys box blocks of 96 bytes (6 float4 rows as
// above).  Stride in float4 units: 2 + 6*nkeys.
struct alignas(16) DevMTopo {
    int32_t children[4];
    int32_t axes[4];
};

// Geom table entry.  Static geom: node_base indexes DevNode, tri_base indexes
// triangles (3 float4 each: v0, v1, v2 with w unused).  Motion geom:
// node_base/tri_base are float4 offsets into the motion arrays; a motion
// triangle is nkeys consecutive groups of 3 float4.
struct alignas(16) DevGeom {
    int32_t node_base;
    int32_t tri_base;
    int32_t nkeys;      // 1 = static
    float ray_bias;     // PolyMesh.RayBias (builtin/geom/polymesh/polymesh.go:19)
};

// Per geom slot: the transform a geom applies to rays before tracing what it holds.  kind:
//   kXfNone      a mesh or sphere without a transform
//   kXfMesh      a PolyMesh with its own Transform (builtin/geom/polymesh/trace.go:20-74)
//   kXfInstance  a GeomInstance of child 0 (builtin/geom/instance/instance.go:73-114)
//   kXfProc      a Proc over its children, in order (builtin/geom/proc/proc.go:88-139)
// One key: M and M^-1 precomputed at mat_base (32 floats; a mesh's single key is its raw matrix, an
// instance's or proc's the recomposed TransformDecomp).  More: nkeys TransformDecomp records
// (kXfKeyFloats each) at key_base, interpolated at the ray's time.
constexpr int kXfNone = 0, kXfMesh = 1, kXfInstance = 2, kXfProc = 3;
struct alignas(16) DevXform {
    int32_t kind;
    int32_t nkeys;
    int32_t key_base;      // float offset in SceneView::xkeys (more than one key)
    int32_t mat_base;      // float offset in SceneView::xkeys (one key)
    int32_t child_base;    // first entry in SceneView::xchildren (instances and procs: geom slots)
    int32_t nchild;
    int32_t pad0, pad1;
};

// The transform of a geom at a ray's time and its inverse (a singular matrix inverts to zero, as
// Matrix4Inverse leaves it).  Keys as polymesh/trace.go:27-44 and TransformSRTArray.TimeKey
// (instance.go:17-34) pick them; a time outside [0, 1] is held to the end keys.
__device__ inline void xform_at(const DevXform& x, const float* xkeys, float time, float* M, float* Minv) {
    if (x.nkeys == 1) {
        const float* m = xkeys + x.mat_base;
        for (int i = 0; i < 16; i++) { M[i] = m[i]; Minv[i] = m[16 + i]; }
        return;
    }
    const float k = time * (float)(x.nkeys - 1);
    const float fk = floorf(k);
    const float t = k - fk;
    const int k0 = min(max((int)fk, 0), x.nkeys - 1), k1 = min(max((int)ceilf(k), 0), x.nkeys - 1);
    xf_decomp_matrix(xkeys + x.key_base + k0 * kXfKeyFloats, xkeys + x.key_base + k1 * kXfKeyFloats, t, M);
    if (!xf_inverse(M, Minv))
        for (int i = 0; i < 16; i++) Minv[i] = 0.0f;
}

struct SceneView {
    const DevNode* nodes;     // static nodes: top level at [0, ntop) when the top is static
    const float4* tris;       // static triangles
    const float4* mnodes;     // motion nodes (topology + boxes)
    const float4* mtris;      // motion triangles
    const DevGeom* geoms;     // post-sort geom slots
    const float4* spheres;    // analytic sphere geoms (centre, radius), indexed by DevGeom::tri_base
    int32_t top_motion;       // 1: the top level is a motion QBVH at mnodes[top_mbase]
    int32_t top_keys;         // motion keys of the top level
    int32_t top_mbase;
    int32_t nan_free_boxes;   // 1: no static node box holds a NaN bound
    int32_t any_xform;        // 1: some geom applies a transform (the per-lane kernel traces every ray)
    const DevXform* xforms;   // per geom slot (any_xform only)
    const float* xkeys;       // transform records
    const int32_t* xchildren; // geom slots held by instances and procs
};

struct TraceLaunch {
    int block;          // 64, 128, 256
    int lds_depth;      // 8, 12, 16, 24, 32
    int waves_per_cu;   // resident waves per CU used to size the grid
    int refill_min = kRefillDefault;   // idle lanes that make a wavefront take new rays together (1..64)
};

// Launch shapes of launch_trace_ext (the shading passes' closest-hit traversal) for static and
// motion scenes (the motion shape runs the four-wave kernel; the five-wave one measured 3% slower
// in the motion-blur render); callers size the spill area for the larger grid.
constexpr TraceLaunch kExtLaunch{128, 12, 24};
constexpr TraceLaunch kExtLaunchMotion{256, 16, 16};

// Scratch of the ray sort (raysort.hip), sized for `capacity` rays.
struct RaySortBuffers {
    uint32_t* keys;      // capacity
    uint32_t* perm;      // capacity: sorted position -> original ray index
    float4* rays;        // 2 * capacity: sorted rays
    uint32_t* hist;      // ray_sort_buckets(bits) counters
    uint32_t* bsum;      // ray_sort_scan_blocks(bits)
    uint64_t capacity;
};
uint32_t ray_sort_buckets(int bits);
uint32_t ray_sort_scan_blocks(int bits);
// Sorts min(n, *n_dev) rays into sb.rays / sb.perm by (Morton cell of the origin in a 2^bits grid over
// box = lo xyz, hi xyz) x (direction octant); bits in [1,6].
hipError_t launch_ray_sort(const float* box, int bits, const void* rays, uint64_t n, const uint32_t* n_dev,
                           const RaySortBuffers& sb, int num_cus, hipStream_t stream);

// Enqueues one traversal over n rays (device pointers).  spill_T / spill_N
// hold (VGPU_MAX_STACK - lds_depth) entries per launched thread; the launch
// uses at most spill_threads threads.  n_dev (nullable): a device-resident
// ray count; the kernel traces min(n, *n_dev) rays (n sizes the grid).  perm (nullable):
// ray i's results go to slot perm[i] (rays reordered by launch_ray_sort).
// Returns hipSuccess or the launch error.
hipError_t launch_trace(const SceneView& sv, bool shadow, bool motion, const TraceLaunch& cfg,
                        const void* rays, uint64_t n, void* hits, int32_t* geoms, int32_t* counters,
                        float* spill_T, int32_t* spill_N, uint64_t spill_threads, int32_t* err_flag,
                        int num_cus, hipStream_t stream, const uint32_t* n_dev = nullptr,
                        const uint32_t* perm = nullptr);

// Closest hit with 32-byte records {prim, geom, t, U, V, W, 0, 0} (U, V, W: barycentric weights of the
// triangle's three vertices), for the shading kernels.  Launch shape kExtLaunch, or kExtLaunchMotion
// with motion.
hipError_t launch_trace_ext(const SceneView& sv, bool motion, const void* rays, uint64_t n, void* hits32, float* spill_T,
                            int32_t* spill_N, uint64_t spill_threads, int32_t* err_flag, int num_cus, hipStream_t stream,
                            const uint32_t* n_dev = nullptr);

// Threads the launch above will use (so the caller can size the spill area).
uint64_t trace_grid_threads(const TraceLaunch& cfg, uint64_t n, int num_cus);

// Allocation-free traversal (vgpu_trace_noalloc): the spill area and error flag live in a
// file-scope device array of kNoAllocScratchBytes in trace.hip, one copy per GPU.
constexpr uint64_t kNoAllocScratchBytes = 128ull << 20;
constexpr int kNoAllocMaxDevices = 64;
// Looks the array up on `device`, which must be current; once per device (later calls return at once).
hipError_t noalloc_init(int device);
// Threads whose spill columns fit the array at LDS depth lds_depth.
uint64_t noalloc_max_threads(int lds_depth);
// Clears the error flag (status, or the array's own flag word when status is NULL) with a kernel,
// then enqueues launch_trace with the array as its spill area.  The grid of cfg must fit
// noalloc_max_threads.  hipErrorNotInitialized: noalloc_init has not run on `device`.
hipError_t launch_trace_noalloc(int device, const SceneView& sv, bool shadow, bool motion, const TraceLaunch& cfg,
                                const void* rays, uint64_t n, void* hits, int32_t* geoms, int32_t* counters,
                                int32_t* status, int num_cus, hipStream_t stream);

}  // namespace vgpu
