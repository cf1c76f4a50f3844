// Ray traversal kernels for gfx950 (MI355X).
//
// One ray per lane.  Each lane runs the reference's two-level traversal —
// scene QBVH over geoms, then the hit geom's triangle QBVH — with the same
// box test lane semantics, the same child push order, the same pop-time
// culling and the same watertight triangle test, so hit IDs, t and
// barycentrics match the reference bit for bit:
//   core/ray.go:103-148                 Ray.Setup
//   qbvh/intersect_amd64.s:13-100       intersectBoxes
//   qbvh/intersect.go:91-246            Trace
//   qbvh/motionintersect.go:22-127      TraceMotion
//   builtin/scene/scene.go:30-97        Scene.Trace / TraceElems
//   builtin/geom/polymesh/trace.go:108-194, :520-669   TraceElems / TraceMotionElems
//
// The traversal stack (the reference's RenderTask.Traversal.Stack, shared by
// both levels, core/ray.go:153-162) lives in LDS for its first LDS entries,
// laid out lane-minor ([depth][lane]) so lanes never conflict on a bank;
// deeper entries spill to a per-thread HBM region, also lane-minor so that
// lanes at equal depth coalesce.
//
// Compiled with -ffp-contract=off: every float operation rounds as the
// reference's unfused SSE code does.

#include "trace_kernels.h"
#include "../../include/vermeer_gpu.h"

namespace vgpu {

namespace {

struct RayS {
    float P[3];
    float Dinv[3];
    float S[3];
    int Kx, Ky, Kz;
    int neg;          // bit k set when D[k] < 0
    float Pkx, Pky, Pkz;
    float time;
    float tclosest;
};

__device__ __forceinline__ float sel3(float a, float b, float c, int k) { return k == 0 ? a : (k == 1 ? b : c); }

__device__ __forceinline__ float absf(float x) { return __int_as_float(__float_as_int(x) & 0x7fffffff); }

// core/ray.go:103-148
__device__ __forceinline__ void ray_setup(RayS& r, const float D[3]) {
    int Kz = 0;
    if (absf(D[1]) > absf(D[2])) {
        if (absf(D[1]) > absf(D[0])) Kz = 1;
    } else {
        if (absf(D[2]) > absf(D[0])) Kz = 2;
    }
    int Kx = Kz + 1; if (Kx == 3) Kx = 0;
    int Ky = Kx + 1; if (Ky == 3) Ky = 0;
    const float dkz = sel3(D[0], D[1], D[2], Kz);
    if (dkz < 0.0f) { int t = Kx; Kx = Ky; Ky = t; }
    const double z = (double)dkz;
    r.S[2] = (float)(1.0 / z);
    r.S[0] = (float)((double)sel3(D[0], D[1], D[2], Kx) / z);
    r.S[1] = (float)((double)sel3(D[0], D[1], D[2], Ky) / z);
    for (int k = 0; k < 3; k++) r.Dinv[k] = (float)(1.0 / (double)D[k]);
    r.Kx = Kx; r.Ky = Ky; r.Kz = Kz;
    r.neg = (D[0] < 0.0f ? 1 : 0) | (D[1] < 0.0f ? 2 : 0) | (D[2] < 0.0f ? 4 : 0);
    r.Pkx = sel3(r.P[0], r.P[1], r.P[2], Kx);
    r.Pky = sel3(r.P[0], r.P[1], r.P[2], Ky);
    r.Pkz = sel3(r.P[0], r.P[1], r.P[2], Kz);
}

// One axis slab of intersectBoxes (qbvh/intersect_amd64.s:26-81), NaN rules of MINPS/MAXPS.
__device__ __forceinline__ void slab(float lo, float hi, float o, float di, float& tmin, float& tmax, bool first) {
    const float t1 = (lo - o) * di;
    const float t2 = (hi - o) * di;
    const float m = (t2 < t1) ? t2 : t1;
    const float M = (t2 > t1) ? t2 : t1;
    if (first) { tmin = m; tmax = M; }
    else { tmin = (tmin > m) ? tmin : m; tmax = (tmax < M) ? tmax : M; }
}

struct Box4 {
    float t[4];
    bool h[4];
};

__device__ __forceinline__ void box4(const RayS& r, const float4& lx, const float4& ly, const float4& lz,
                                     const float4& hx, const float4& hy, const float4& hz, Box4& b) {
    const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, lya[4] = {ly.x, ly.y, ly.z, ly.w}, lza[4] = {lz.x, lz.y, lz.z, lz.w};
    const float hxa[4] = {hx.x, hx.y, hx.z, hx.w}, hya[4] = {hy.x, hy.y, hy.z, hy.w}, hza[4] = {hz.x, hz.y, hz.z, hz.w};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        float tmin, tmax;
        slab(lxa[c], hxa[c], r.P[0], r.Dinv[0], tmin, tmax, true);
        slab(lya[c], hya[c], r.P[1], r.Dinv[1], tmin, tmax, false);
        slab(lza[c], hza[c], r.P[2], r.Dinv[2], tmin, tmax, false);
        const float tn = (0.0f > tmin) ? 0.0f : tmin;   // MAXPS X6,X0 with X0 = 0
        b.t[c] = tn;
        b.h[c] = tn <= tmax;                             // CMPPS $2 (ordered <=)
    }
}

template <int BLOCK, int LDS>
struct Stack {
    float* sT;
    int32_t* sN;
    float* gT;
    int32_t* gN;
    uint64_t gstride;
    __device__ __forceinline__ void store(int i, float T, int32_t node) {
        if (i < LDS) { sT[i * BLOCK] = T; sN[i * BLOCK] = node; }
        else { const uint64_t o = (uint64_t)(i - LDS) * gstride; gT[o] = T; gN[o] = node; }
    }
    __device__ __forceinline__ void load(int i, float& T, int32_t& node) const {
        if (i < LDS) { T = sT[i * BLOCK]; node = sN[i * BLOCK]; }
        else { const uint64_t o = (uint64_t)(i - LDS) * gstride; T = gT[o]; node = gN[o]; }
    }
};

__device__ __forceinline__ int leaf_base(int32_t l) { return (int)((l & 0x7ffffff) >> 4); }
__device__ __forceinline__ int leaf_count(int32_t l) { return (int)((l & 0xf) + 1); }

// Watertight ray/triangle test (builtin/geom/polymesh/trace.go:123-194 static, :556-624 motion).
// Returns true when the triangle is accepted; updates tclosest, U, V.
template <bool MOTION_BOUND>
__device__ __forceinline__ bool tri_test(RayS& r, const float4& v0, const float4& v1, const float4& v2, float bias,
                                         float& U, float& V, float& W) {
    const float AKz = sel3(v0.x, v0.y, v0.z, r.Kz) - r.Pkz;
    const float BKz = sel3(v1.x, v1.y, v1.z, r.Kz) - r.Pkz;
    const float CKz = sel3(v2.x, v2.y, v2.z, r.Kz) - r.Pkz;
    const float Cx = (sel3(v2.x, v2.y, v2.z, r.Kx) - r.Pkx) - r.S[0] * CKz;
    const float By = (sel3(v1.x, v1.y, v1.z, r.Ky) - r.Pky) - r.S[1] * BKz;
    const float Cy = (sel3(v2.x, v2.y, v2.z, r.Ky) - r.Pky) - r.S[1] * CKz;
    const float Bx = (sel3(v1.x, v1.y, v1.z, r.Kx) - r.Pkx) - r.S[0] * BKz;
    const float Ax = (sel3(v0.x, v0.y, v0.z, r.Kx) - r.Pkx) - r.S[0] * AKz;
    const float Ay = (sel3(v0.x, v0.y, v0.z, r.Ky) - r.Pky) - r.S[1] * AKz;
    float fU = Cx * By - Cy * Bx;
    float fV = Ax * Cy - Ay * Cx;
    float fW = Bx * Ay - By * Ax;
    if (fU == 0.0f || fV == 0.0f || fW == 0.0f) {
        fU = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        fV = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        fW = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((fU < 0.0f || fV < 0.0f || fW < 0.0f) && (fU > 0.0f || fV > 0.0f || fW > 0.0f)) return false;
    const float det = fU + fV + fW;
    if (det == 0.0f) return false;
    const float T = r.S[2] * (fU * AKz + fV * BKz + fW * CKz);
    const int sgn = __float_as_int(det) & (int)0x80000000;
    const float Tx = __int_as_float(__float_as_int(T) ^ sgn);
    const float Dx = __int_as_float(__float_as_int(det) ^ sgn);
    if (MOTION_BOUND) {
        if (Tx <= bias * Dx || Tx > r.tclosest * Dx) return false;
    } else {
        if (Tx < (1.19209290E-07f + bias) * Dx || Tx > r.tclosest * Dx) return false;
    }
    const float rcp = 1.0f / det;
    U = fU * rcp;
    V = fV * rcp;
    W = fW * rcp;
    r.tclosest = T * rcp;
    return true;
}

__device__ __forceinline__ float4 lerp4(const float4& a, const float4& b, float t) {
    const float s = 1.0f - t;
    return make_float4(s * a.x + t * b.x, s * a.y + t * b.y, s * a.z + t * b.z, s * a.w + t * b.w);
}

struct NodeData {
    float4 lx, ly, lz, hx, hy, hz;
    int ch[4];
    int ax[3];
};

__device__ __forceinline__ void fetch_static(const DevNode* __restrict__ nodes, int idx, NodeData& nd) {
    const float4* p = reinterpret_cast<const float4*>(nodes + idx);
    nd.lx = p[0]; nd.ly = p[1]; nd.lz = p[2]; nd.hx = p[3]; nd.hy = p[4]; nd.hz = p[5];
    const int4 c = reinterpret_cast<const int4*>(p)[6];
    const int4 a = reinterpret_cast<const int4*>(p)[7];
    nd.ch[0] = c.x; nd.ch[1] = c.y; nd.ch[2] = c.z; nd.ch[3] = c.w;
    nd.ax[0] = a.x; nd.ax[1] = a.y; nd.ax[2] = a.z;
}

// Motion node: boxes lerped per key (qbvh/motionintersect.go:47-49).
__device__ __forceinline__ void fetch_motion(const float4* __restrict__ mnodes, int base, int node, int nkeys, int key,
                                             int key2, float time, NodeData& nd) {
    const float4* p = mnodes + base + (size_t)node * (2 + 6 * nkeys);
    const int4 c = reinterpret_cast<const int4*>(p)[0];
    const int4 a = reinterpret_cast<const int4*>(p)[1];
    nd.ch[0] = c.x; nd.ch[1] = c.y; nd.ch[2] = c.z; nd.ch[3] = c.w;
    nd.ax[0] = a.x; nd.ax[1] = a.y; nd.ax[2] = a.z;
    const float4* b0 = p + 2 + 6 * key;
    const float4* b1 = p + 2 + 6 * key2;
    nd.lx = lerp4(b0[0], b1[0], time); nd.ly = lerp4(b0[1], b1[1], time); nd.lz = lerp4(b0[2], b1[2], time);
    nd.hx = lerp4(b0[3], b1[3], time); nd.hy = lerp4(b0[4], b1[4], time); nd.hz = lerp4(b0[5], b1[5], time);
}

// ray.Time -> (key, key2, t) (builtin/geom/polymesh/trace.go:79-84, scene.go:49-54)
__device__ __forceinline__ void motion_keys(float rtime, int nkeys, int& key, int& key2, float& t) {
    const float k = rtime * (float)(nkeys - 1);
    const float fk = floorf(k);
    t = k - fk;
    key = (int)fk;
    key2 = (int)ceilf(k);
}

// EXT: hits are 32-byte records {prim, geom, t, U, V, W, 0, 0} for the shading kernels.
template <int BLOCK, int LDS, bool SHADOW, bool MOTION, bool COUNTERS, bool EXT = false>
__global__ __launch_bounds__(BLOCK) void k_trace(SceneView sv, const float4* __restrict__ rays, uint64_t n,
                                                 int4* __restrict__ hits, int32_t* __restrict__ geoms_out,
                                                 int4* __restrict__ counters, float* __restrict__ spT,
                                                 int32_t* __restrict__ spN, int32_t* __restrict__ err,
                                                 const uint32_t* __restrict__ n_dev,
                                                 const uint32_t* __restrict__ perm) {
    __shared__ float sT[LDS * BLOCK];
    __shared__ int32_t sN[LDS * BLOCK];
    const int tid = threadIdx.x;
    const uint64_t gtid = (uint64_t)blockIdx.x * BLOCK + tid;
    const uint64_t nthr = (uint64_t)gridDim.x * BLOCK;
    Stack<BLOCK, LDS> st;
    st.sT = sT + tid; st.sN = sN + tid;
    st.gT = spT + gtid; st.gN = spN + gtid; st.gstride = nthr;

    if (n_dev) n = min(n, (uint64_t)*n_dev);   // device-resident count (compacted queues)
    for (uint64_t ri = gtid; ri < n; ri += nthr) {
        const float4 ra = rays[2 * ri], rb = rays[2 * ri + 1];
        RayS r;
        r.P[0] = ra.x; r.P[1] = ra.y; r.P[2] = ra.z; r.tclosest = ra.w;
        const float D[3] = {rb.x, rb.y, rb.z};
        r.time = rb.w;
        ray_setup(r, D);

        int nodesT = 0, leafsT = 0, trisT = 0, maxDepth = 0;
        int best_geom = -1, best_prim = -1;
        float bu = 0.0f, bv = 0.0f, bw = 0.0f;

        // top level
        const bool top_motion = MOTION && sv.top_motion;
        int tkey = 0, tkey2 = 0; float ttime = 0.0f;
        if (top_motion) motion_keys(r.time, sv.top_keys, tkey, tkey2, ttime);

        bool in_mesh = false;
        int mesh_base = 0, gcur = 0, gend = 0;
        DevGeom g = {0, 0, 1, 0.0f};
        int mkey = 0, mkey2 = 0; float mtime = 0.0f;
        bool overflow = false;

        st.store(0, r.tclosest, 0);
        int sp = 1;
        while (true) {
            if (in_mesh && sp == mesh_base) {
                // PolyMesh.Trace returned; Scene.TraceElems moves to the leaf's next geom (scene.go:65-75)
                gcur++;
                if (gcur < gend) {
                    g = sv.geoms[gcur];
                    if (MOTION && g.nkeys > 1) motion_keys(r.time, g.nkeys, mkey, mkey2, mtime);
                    st.store(mesh_base, r.tclosest, 0);
                    sp = mesh_base + 1;
                    continue;
                }
                in_mesh = false;
                sp = top_motion ? mesh_base : mesh_base - 1;
            }
            if (sp == 0) break;
            sp--;
            float T; int32_t node;
            st.load(sp, T, node);
            if (r.tclosest < T || node == -1) continue;
            if (node >= 0) {
                if (COUNTERS) nodesT++;
                NodeData nd;
                if (in_mesh) {
                    if (MOTION && g.nkeys > 1) fetch_motion(sv.mnodes, g.node_base, node, g.nkeys, mkey, mkey2, mtime, nd);
                    else fetch_static(sv.nodes, g.node_base + node, nd);
                } else {
                    if (top_motion) fetch_motion(sv.mnodes, sv.top_mbase, node, sv.top_keys, tkey, tkey2, ttime, nd);
                    else fetch_static(sv.nodes, node, nd);
                }
                Box4 b;
                box4(r, nd.lx, nd.ly, nd.lz, nd.hx, nd.hy, nd.hz, b);
                // Push order of qbvh/intersect.go:137-216 (same as motionintersect.go:53-97): the
                // children are pushed as two pairs; rank(c) is child c's position in that sequence.
                // Hit children are stored compacted in rank order, i.e. exactly the entries the
                // reference leaves below its stack top.
                const int n0 = (r.neg >> nd.ax[0]) & 1;
                const int n1 = (r.neg >> nd.ax[1]) & 1;
                const int n2 = (r.neg >> nd.ax[2]) & 1;
                const int gA = n0 ? 0 : 2, gB = 2 - gA;
                const int rk0 = gA + (n1 ^ 1), rk1 = gA + n1, rk2 = gB + (n2 ^ 1), rk3 = gB + n2;
                const int m = (b.h[0] ? 1 << rk0 : 0) | (b.h[1] ? 1 << rk1 : 0) | (b.h[2] ? 1 << rk2 : 0) |
                              (b.h[3] ? 1 << rk3 : 0);
                if (sp + 4 > VGPU_MAX_STACK) {   // conservative: the reference writes up to 4 slots here
                    overflow = true;
                    break;
                }
                if (b.h[0]) st.store(sp + __popc(m & ((1 << rk0) - 1)), b.t[0], nd.ch[0]);
                if (b.h[1]) st.store(sp + __popc(m & ((1 << rk1) - 1)), b.t[1], nd.ch[1]);
                if (b.h[2]) st.store(sp + __popc(m & ((1 << rk2) - 1)), b.t[2], nd.ch[2]);
                if (b.h[3]) st.store(sp + __popc(m & ((1 << rk3) - 1)), b.t[3], nd.ch[3]);
                sp += __popc(m);
                if (COUNTERS) maxDepth = max(maxDepth, sp);
            } else {
                const int base = leaf_base(node), count = leaf_count(node);
                if (COUNTERS) leafsT++;
                if (in_mesh) {
                    if (COUNTERS) trisT += count;
                    int idx = -1;
                    float U = 0.0f, V = 0.0f, W = 0.0f;
                    if (MOTION && g.nkeys > 1) {
                        const float4* tp = sv.mtris + g.tri_base + (size_t)base * 3 * g.nkeys;
                        for (int i = 0; i < count; i++) {
                            const float4* q = tp + (size_t)i * 3 * g.nkeys;
                            const float4* q0 = q + 3 * mkey;
                            const float4* q1 = q + 3 * mkey2;
                            const float4 v0 = lerp4(q0[0], q1[0], mtime);
                            const float4 v1 = lerp4(q0[1], q1[1], mtime);
                            const float4 v2 = lerp4(q0[2], q1[2], mtime);
                            if (tri_test<true>(r, v0, v1, v2, g.ray_bias, U, V, W)) idx = base + i;
                        }
                    } else {
                        const float4* tp = sv.tris + ((size_t)g.tri_base + base) * 3;
                        for (int i = 0; i < count; i++) {
                            const float4 v0 = tp[3 * i], v1 = tp[3 * i + 1], v2 = tp[3 * i + 2];
                            if (tri_test<false>(r, v0, v1, v2, g.ray_bias, U, V, W)) idx = base + i;
                        }
                    }
                    if (idx >= 0) {
                        best_geom = gcur; best_prim = idx; bu = U; bv = V; bw = W;
                        if (SHADOW) break;   // qbvh/intersect.go:232-236 at both levels
                    }
                } else {
                    // Scene.TraceElems: geoms [base, base+count) (scene.go:61-78)
                    mesh_base = top_motion ? sp : sp + 1;
                    if (mesh_base >= VGPU_MAX_STACK) { overflow = true; break; }
                    gcur = base; gend = base + count;
                    g = sv.geoms[gcur];
                    if (MOTION && g.nkeys > 1) motion_keys(r.time, g.nkeys, mkey, mkey2, mtime);
                    st.store(mesh_base, r.tclosest, 0);
                    sp = mesh_base + 1;
                    in_mesh = true;
                }
            }
        }
        if (overflow) {
            atomicOr(err, 1);
            best_geom = -1; best_prim = -1;
        }
        const uint64_t oi = perm ? (uint64_t)perm[ri] : ri;
        if (EXT) {
            hits[2 * oi] = make_int4(best_prim, best_geom, __float_as_int(r.tclosest), __float_as_int(bu));
            hits[2 * oi + 1] = make_int4(__float_as_int(bv), __float_as_int(bw), 0, 0);
        } else {
            hits[oi] = make_int4(best_prim, __float_as_int(r.tclosest), __float_as_int(bu), __float_as_int(bv));
            if (geoms_out) geoms_out[oi] = best_geom;
        }
        if (COUNTERS) counters[oi] = make_int4(nodesT, leafsT, trisT, maxDepth);
    }
}

template <int BLOCK, int LDS, bool SHADOW, bool MOTION>
hipError_t launch_cfg(const SceneView& sv, const void* rays, uint64_t n, void* hits, int32_t* geoms, int32_t* counters,
                      float* spT, int32_t* spN, uint64_t grid, int32_t* err, const uint32_t* n_dev,
                      const uint32_t* perm, hipStream_t stream) {
    if (counters)
        hipLaunchKernelGGL((k_trace<BLOCK, LDS, SHADOW, MOTION, true>), dim3((unsigned)grid), dim3(BLOCK), 0, stream, sv,
                           (const float4*)rays, n, (int4*)hits, geoms, (int4*)counters, spT, spN, err, n_dev, perm);
    else
        hipLaunchKernelGGL((k_trace<BLOCK, LDS, SHADOW, MOTION, false>), dim3((unsigned)grid), dim3(BLOCK), 0, stream, sv,
                           (const float4*)rays, n, (int4*)hits, geoms, (int4*)nullptr, spT, spN, err, n_dev, perm);
    return hipGetLastError();
}

template <int BLOCK, int LDS>
hipError_t launch_mode(const SceneView& sv, bool shadow, bool motion, const void* rays, uint64_t n, void* hits,
                       int32_t* geoms, int32_t* counters, float* spT, int32_t* spN, uint64_t grid, int32_t* err,
                       const uint32_t* n_dev, const uint32_t* perm, hipStream_t stream) {
    if (shadow) {
        if (motion)
            return launch_cfg<BLOCK, LDS, true, true>(sv, rays, n, hits, geoms, counters, spT, spN, grid, err, n_dev, perm, stream);
        return launch_cfg<BLOCK, LDS, true, false>(sv, rays, n, hits, geoms, counters, spT, spN, grid, err, n_dev, perm, stream);
    }
    if (motion)
        return launch_cfg<BLOCK, LDS, false, true>(sv, rays, n, hits, geoms, counters, spT, spN, grid, err, n_dev, perm, stream);
    return launch_cfg<BLOCK, LDS, false, false>(sv, rays, n, hits, geoms, counters, spT, spN, grid, err, n_dev, perm, stream);
}

}  // namespace

hipError_t launch_trace_ext(const SceneView& sv, bool motion, const void* rays, uint64_t n, void* hits32, float* spT,
                            int32_t* spN, uint64_t spill_threads, int32_t* err, int num_cus, hipStream_t stream,
                            const uint32_t* n_dev) {
    if (n == 0) return hipSuccess;
    const TraceLaunch cfg{256, 16, 16};
    const uint64_t threads = trace_grid_threads(cfg, n, num_cus);
    if (threads > spill_threads) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)(threads / cfg.block);
    if (motion)
        hipLaunchKernelGGL((k_trace<256, 16, false, true, false, true>), dim3(grid), dim3(256), 0, stream, sv,
                           (const float4*)rays, n, (int4*)hits32, (int32_t*)nullptr, (int4*)nullptr, spT, spN, err, n_dev,
                           (const uint32_t*)nullptr);
    else
        hipLaunchKernelGGL((k_trace<256, 16, false, false, false, true>), dim3(grid), dim3(256), 0, stream, sv,
                           (const float4*)rays, n, (int4*)hits32, (int32_t*)nullptr, (int4*)nullptr, spT, spN, err, n_dev,
                           (const uint32_t*)nullptr);
    return hipGetLastError();
}

uint64_t trace_grid_threads(const TraceLaunch& cfg, uint64_t n, int num_cus) {
    const uint64_t resident = (uint64_t)num_cus * (uint64_t)cfg.waves_per_cu * 64ull;
    uint64_t blocks = (resident + cfg.block - 1) / cfg.block;
    const uint64_t need = (n + cfg.block - 1) / cfg.block;
    if (blocks > need) blocks = need;
    if (blocks < 1) blocks = 1;
    return blocks * (uint64_t)cfg.block;
}

hipError_t launch_trace(const SceneView& sv, bool shadow, bool motion, const TraceLaunch& cfg, const void* rays, uint64_t n,
                        void* hits, int32_t* geoms, int32_t* counters, float* spT, int32_t* spN, uint64_t spill_threads,
                        int32_t* err, int num_cus, hipStream_t stream, const uint32_t* n_dev, const uint32_t* perm) {
    if (n == 0) return hipSuccess;
    const uint64_t threads = trace_grid_threads(cfg, n, num_cus);
    if (threads > spill_threads) return hipErrorInvalidValue;
    const uint64_t grid = threads / cfg.block;
#define VGPU_CASE(B, L)                                                                                         \
    if (cfg.block == B && cfg.lds_depth == L)                                                                   \
        return launch_mode<B, L>(sv, shadow, motion, rays, n, hits, geoms, counters, spT, spN, grid, err, n_dev, perm, \
                                 stream);
    VGPU_CASE(64, 8) VGPU_CASE(64, 12) VGPU_CASE(64, 16) VGPU_CASE(64, 24) VGPU_CASE(64, 32)
    VGPU_CASE(128, 8) VGPU_CASE(128, 12) VGPU_CASE(128, 16) VGPU_CASE(128, 24) VGPU_CASE(128, 32)
    VGPU_CASE(256, 8) VGPU_CASE(256, 12) VGPU_CASE(256, 16) VGPU_CASE(256, 24) VGPU_CASE(256, 32)
#undef VGPU_CASE
    return hipErrorInvalidValue;
}

}  // namespace vgpu
