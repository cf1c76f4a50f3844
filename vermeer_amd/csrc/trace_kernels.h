// Internal interface between the scene object (scene.cpp) and the
// traversal kernels (trace.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

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
};

struct TraceLaunch {
    int block;          // 64, 128, 256
    int lds_depth;      // 8, 12, 16, 24, 32
    int waves_per_cu;   // resident waves per CU used to size the grid
    int refill_min = kRefillDefault;   // idle lanes that make a wavefront take new rays together (1..64)
};

// Launch shapes of launch_trace_ext (the shading passes' closest-hit traversal) for static and
// motion scenes (the motion kernel holds four waves per SIMD); callers size the spill area for the
// larger grid.
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
