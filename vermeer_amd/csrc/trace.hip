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
//   builtin/geom/sphere/trace.go:13-103  Sphere.Trace (a SphereLight's analytic sphere)
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

#include <mutex>

#include "trace_kernels.h"
#include "xform.h"
#include "../../include/vermeer_gpu.h"

namespace vgpu {

namespace {

struct RayS {
    float P[3];
    float Dinv[3];
    float S[3];
    int Kx, Ky, Kz;
    int neg;          // bit k set when D[k] < 0
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
    // The reference divides in float64 and rounds the quotient to float32.  For float32
    // operands that double rounding is innocuous (53 >= 2 * 24 + 2 bits: the float64 quotient
    // never lands on a float32 rounding midpoint unless the exact quotient does), so a
    // correctly rounded float32 division gives the same bits.
    r.S[2] = 1.0f / dkz;
    r.S[0] = sel3(D[0], D[1], D[2], Kx) / dkz;
    r.S[1] = sel3(D[0], D[1], D[2], Ky) / dkz;
    for (int k = 0; k < 3; k++) r.Dinv[k] = 1.0f / D[k];
    r.Kx = Kx; r.Ky = Ky; r.Kz = Kz;
    r.neg = (D[0] < 0.0f ? 1 : 0) | (D[1] < 0.0f ? 2 : 0) | (D[2] < 0.0f ? 4 : 0);
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

// FAST: every slab value is known not to be NaN (finite origin, finite non-zero 1/D, no NaN in
// the node boxes).  MINPS/MAXPS then agree with IEEE min/max except for which zero of +0/-0
// they return, and the entry distances are only ever compared, where +0 == -0, so the
// hardware min/max (v_min3/v_max3) give the same hit flags and the same traversal.
template <bool FAST>
__device__ __forceinline__ void box4(const RayS& r, const float4& lx, const float4& ly, const float4& lz,
                                     const float4& hx, const float4& hy, const float4& hz, Box4& b) {
    const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, lya[4] = {ly.x, ly.y, ly.z, ly.w}, lza[4] = {lz.x, lz.y, lz.z, lz.w};
    const float hxa[4] = {hx.x, hx.y, hx.z, hx.w}, hya[4] = {hy.x, hy.y, hy.z, hy.w}, hza[4] = {hz.x, hz.y, hz.z, hz.w};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        if (FAST) {
            const float x1 = (lxa[c] - r.P[0]) * r.Dinv[0], x2 = (hxa[c] - r.P[0]) * r.Dinv[0];
            const float y1 = (lya[c] - r.P[1]) * r.Dinv[1], y2 = (hya[c] - r.P[1]) * r.Dinv[1];
            const float z1 = (lza[c] - r.P[2]) * r.Dinv[2], z2 = (hza[c] - r.P[2]) * r.Dinv[2];
            const float tn = fmaxf(fmaxf(fmaxf(fminf(x1, x2), fminf(y1, y2)), fminf(z1, z2)), 0.0f);
            const float tx = fminf(fminf(fmaxf(x1, x2), fmaxf(y1, y2)), fmaxf(z1, z2));
            b.t[c] = tn;
            b.h[c] = tn <= tx;
            continue;
        }
        float tmin, tmax;
        slab(lxa[c], hxa[c], r.P[0], r.Dinv[0], tmin, tmax, true);
        slab(lya[c], hya[c], r.P[1], r.Dinv[1], tmin, tmax, false);
        slab(lza[c], hza[c], r.P[2], r.Dinv[2], tmin, tmax, false);
        const float tn = (0.0f > tmin) ? 0.0f : tmin;   // MAXPS X6,X0 with X0 = 0
        b.t[c] = tn;
        b.h[c] = tn <= tmax;                             // CMPPS $2 (ordered <=)
    }
}

// The spill area is addressed from its (wavefront-uniform) bases plus one 32-bit per-lane offset,
// so the lane keeps a single register for it.
template <int BLOCK, int LDS>
struct Stack {
    float* sT;
    int32_t* sN;
    float* gT;          // spill bases (uniform)
    int32_t* gN;
    uint64_t gstride;
    uint32_t gofs;      // this thread's column
    __device__ __forceinline__ void store(int i, float T, int32_t node) {
        if (i < LDS) { sT[i * BLOCK] = T; sN[i * BLOCK] = node; }
        else { const uint64_t o = (uint64_t)(i - LDS) * gstride + gofs; gT[o] = T; gN[o] = node; }
    }
    __device__ __forceinline__ void load(int i, float& T, int32_t& node) const {
        if (i < LDS) { T = sT[i * BLOCK]; node = sN[i * BLOCK]; }
        else { const uint64_t o = (uint64_t)(i - LDS) * gstride + gofs; T = gT[o]; node = gN[o]; }
    }
};

__device__ __forceinline__ int leaf_base(int32_t l) { return (int)((l & 0x7ffffff) >> 4); }
__device__ __forceinline__ int leaf_count(int32_t l) { return (int)((l & 0xf) + 1); }

// Ray quantities the watertight test reads: the shear constants and the permuted origin.
struct TriRay {
    float S0, S1, S2, Pkx, Pky, Pkz;
    int Kx, Ky, Kz;
};


struct TriPre {
    float T, det, fU, fV, fW;
};

// The direction-sign bits and the permutation packed in one register: neg (bits 0-2, bit k set when
// D[k] < 0), Kx (3-4), Ky (5-6), Kz (7-8).  The wide kernel keeps this instead of the separate fields
// and the permuted origin, which it rebuilds from P when a leaf is tested.
__device__ __forceinline__ int pack_perm(const RayS& r) { return r.neg | (r.Kx << 3) | (r.Ky << 5) | (r.Kz << 7); }
__device__ __forceinline__ TriRay tri_ray_packed(const RayS& r, int kp) {
    TriRay q;
    q.S0 = r.S[0]; q.S1 = r.S[1]; q.S2 = r.S[2];
    q.Kx = (kp >> 3) & 3; q.Ky = (kp >> 5) & 3; q.Kz = (kp >> 7) & 3;
    q.Pkx = sel3(r.P[0], r.P[1], r.P[2], q.Kx);
    q.Pky = sel3(r.P[0], r.P[1], r.P[2], q.Ky);
    q.Pkz = sel3(r.P[0], r.P[1], r.P[2], q.Kz);
    return q;
}

// Watertight ray/triangle test (builtin/geom/polymesh/trace.go:123-194 static, :556-624 motion),
// the part that does not read the closest distance so far: edge functions (retried in double
// precision when one is zero), the sign test, det != 0 and the near bound.
template <bool MOTION_BOUND>
__device__ __forceinline__ bool tri_pre(const TriRay& q, const float4& v0, const float4& v1, const float4& v2, float bias,
                                        TriPre& p) {
    const float AKz = sel3(v0.x, v0.y, v0.z, q.Kz) - q.Pkz;
    const float BKz = sel3(v1.x, v1.y, v1.z, q.Kz) - q.Pkz;
    const float CKz = sel3(v2.x, v2.y, v2.z, q.Kz) - q.Pkz;
    const float Cx = (sel3(v2.x, v2.y, v2.z, q.Kx) - q.Pkx) - q.S0 * CKz;
    const float By = (sel3(v1.x, v1.y, v1.z, q.Ky) - q.Pky) - q.S1 * BKz;
    const float Cy = (sel3(v2.x, v2.y, v2.z, q.Ky) - q.Pky) - q.S1 * CKz;
    const float Bx = (sel3(v1.x, v1.y, v1.z, q.Kx) - q.Pkx) - q.S0 * BKz;
    const float Ax = (sel3(v0.x, v0.y, v0.z, q.Kx) - q.Pkx) - q.S0 * AKz;
    const float Ay = (sel3(v0.x, v0.y, v0.z, q.Ky) - q.Pky) - q.S1 * AKz;
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
    const float T = q.S2 * (fU * AKz + fV * BKz + fW * CKz);
    const int sgn = __float_as_int(det) & (int)0x80000000;
    const float Tx = __int_as_float(__float_as_int(T) ^ sgn);
    const float Dx = __int_as_float(__float_as_int(det) ^ sgn);
    if (MOTION_BOUND) {
        if (Tx <= bias * Dx) return false;
    } else {
        if (Tx < (1.19209290E-07f + bias) * Dx) return false;
    }
    p.T = T; p.det = det; p.fU = fU; p.fV = fV; p.fW = fW;
    return true;
}

// The rest of the test: far bound against the closest distance so far, then the hit update.
__device__ __forceinline__ bool tri_accept(RayS& r, const TriPre& p, float& U, float& V, float& W) {
    const int sgn = __float_as_int(p.det) & (int)0x80000000;
    const float Tx = __int_as_float(__float_as_int(p.T) ^ sgn);
    const float Dx = __int_as_float(__float_as_int(p.det) ^ sgn);
    if (Tx > r.tclosest * Dx) return false;
    const float rcp = 1.0f / p.det;
    U = p.fU * rcp;
    V = p.fV * rcp;
    W = p.fW * rcp;
    r.tclosest = p.T * rcp;
    return true;
}

// The whole test with the ray quantities given.
template <bool MOTION_BOUND>
__device__ __forceinline__ bool tri_test_q(RayS& r, const TriRay& q, const float4& v0, const float4& v1, const float4& v2,
                                           float bias, float& U, float& V, float& W) {
    TriPre p;
    return tri_pre<MOTION_BOUND>(q, v0, v1, v2, bias, p) && tri_accept(r, p, U, V, W);
}




__device__ __forceinline__ float4 lerp4(const float4& a, const float4& b, float t) {
    const float s = 1.0f - t;
    return make_float4(s * a.x + t * b.x, s * a.y + t * b.y, s * a.z + t * b.z, s * a.w + t * b.w);
}

// Vertices of a motion triangle (its nkeys groups of 3 float4 at q) at time t between keys k0 and k1
// (trace.go:547-554).
__device__ __forceinline__ void motion_tri(const float4* __restrict__ q, int k0, int k1, float t, float4& v0, float4& v1,
                                           float4& v2) {
    const float4* a = q + 3 * k0;
    const float4* b = q + 3 * k1;
    v0 = lerp4(a[0], b[0], t);
    v1 = lerp4(a[1], b[1], t);
    v2 = lerp4(a[2], b[2], t);
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

// Motion node (qbvh/motionintersect.go:47-49): the boxes lerped per key and tested one axis at a
// time, in the slab order of box4<false>, so only one axis of both keys is live at once.
__device__ __forceinline__ void box_motion(const RayS& r, const float4* __restrict__ mnodes, int base, int node,
                                           int nkeys, int key, int key2, float time, NodeData& nd, Box4& b) {
    const float4* p = mnodes + base + (size_t)node * (2 + 6 * nkeys);
    const int4 c = reinterpret_cast<const int4*>(p)[0];
    const int4 a = reinterpret_cast<const int4*>(p)[1];
    nd.ch[0] = c.x; nd.ch[1] = c.y; nd.ch[2] = c.z; nd.ch[3] = c.w;
    nd.ax[0] = a.x; nd.ax[1] = a.y; nd.ax[2] = a.z;
    const float4* b0 = p + 2 + 6 * key;
    const float4* b1 = p + 2 + 6 * key2;
    float tmin[4], tmax[4];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        const float4 lo = lerp4(b0[ax], b1[ax], time), hi = lerp4(b0[3 + ax], b1[3 + ax], time);
        const float la[4] = {lo.x, lo.y, lo.z, lo.w}, ha[4] = {hi.x, hi.y, hi.z, hi.w};
#pragma unroll
        for (int ch = 0; ch < 4; ch++) slab(la[ch], ha[ch], r.P[ax], r.Dinv[ax], tmin[ch], tmax[ch], ax == 0);
    }
#pragma unroll
    for (int ch = 0; ch < 4; ch++) {
        const float tn = (0.0f > tmin[ch]) ? 0.0f : tmin[ch];
        b.t[ch] = tn;
        b.h[ch] = tn <= tmax[ch];
    }
}

// ray.Time -> (key, key2, t) (builtin/geom/polymesh/trace.go:79-84, scene.go:49-54)
// A time outside [0, 1] (the reference indexes out of range and panics) is held to the end keys,
// so no key index leaves the node or triangle.
__device__ __forceinline__ void motion_keys(float rtime, int nkeys, int& key, int& key2, float& t) {
    const float k = rtime * (float)(nkeys - 1);
    const float fk = floorf(k);
    t = k - fk;
    key = min(max((int)fk, 0), nkeys - 1);
    key2 = min(max((int)ceilf(k), 0), nkeys - 1);
}

// Sphere.Trace (builtin/geom/sphere/trace.go:13-52): the closed-form quadratic (solveQuadratic,
// raySphereIntersect :54-103), accepted when t < Tclosest, with no ray bias.
__device__ bool sphere_geom_hit(const float* P, const float4& d, const float4& sp, float tclosest, float& t) {
    const float Lx = P[0] - sp.x, Ly = P[1] - sp.y, Lz = P[2] - sp.z;
    const float a = d.x * d.x + d.y * d.y + d.z * d.z;
    const float b = 2 * (d.x * Lx + d.y * Ly + d.z * Lz);
    const float c = (Lx * Lx + Ly * Ly + Lz * Lz) - sp.w * sp.w;
    float x0, x1;
    const float discr = b * b - 4 * a * c;
    if (discr < 0) return false;
    if (discr == 0) {
        x1 = -0.5f * b / a;
        x0 = x1;
    } else {
        const float q = b > 0 ? -0.5f * (b + sqrtf(discr)) : -0.5f * (b - sqrtf(discr));
        x0 = q / a;
        x1 = c / q;
    }
    if (x0 > x1) { const float tt = x0; x0 = x1; x1 = tt; }
    if (x0 < 0) {
        x0 = x1;
        if (x0 < 0) return false;
    }
    if (!(x0 < tclosest)) return false;
    t = x0;
    return true;
}

// Scene.TraceElems over geoms [gcur, gend) of a top-level leaf (scene.go:61-78): analytic spheres
// (DevGeom nkeys 0) are tested here, in order, up to the first mesh.  Returns 1 with g = that mesh,
// 2 when a shadow ray hit a sphere (the traversal ends, qbvh/intersect.go:232-236), 0 at the leaf's end.
template <bool SHADOW>
__device__ __forceinline__ int next_geom(const SceneView& sv, const float4* __restrict__ rays, uint64_t ri, RayS& r,
                                         int& gcur, int gend, DevGeom& g, int& best_geom, int& best_prim, float& bu,
                                         float& bv, float& bw) {
    for (; gcur < gend; gcur++) {
        g = sv.geoms[gcur];
        if (g.nkeys != 0) return 1;
        float t;
        if (sphere_geom_hit(r.P, rays[2 * ri + 1], sv.spheres[g.tri_base], r.tclosest, t)) {
            r.tclosest = t;
            best_geom = gcur; best_prim = 0; bu = 0.0f; bv = 0.0f; bw = 0.0f;
            if (SHADOW) return 2;
        }
    }
    return 0;
}

// The traversal, one ray per lane and one traversal step (a node or a leaf) per loop iteration.
// The triangles of the leaves popped in a step are tested by the whole wavefront: lane l
// owns a leaf of c_l triangles, the sum of the c_l triangle tests is dealt out over the 64
// lanes (consecutive lanes take consecutive triangles of one leaf, so the loads coalesce),
// and each lane reads the owning ray's shear constants with a cross-lane shuffle.  That part
// of the watertight test does not read the closest distance; the owner then runs the full
// test, in leaf order, on the triangles that passed it, which accepts the same triangles in
// the same order as testing every triangle of the leaf in turn.  Lanes whose ray is finished
// stay in the loop as workers until the whole wavefront is done.
// MOTION: the scene has motion keys (a motion top level and/or moving meshes); node boxes
// (qbvh/motionintersect.go:47-49) and triangles of moving meshes (trace.go:547-554) are
// interpolated at the ray's time.
// EXT: hits are 32-byte records {prim, geom, t, U, V, W, 0, 0} for the shading kernels.
template <int BLOCK, int LDS, bool SHADOW, bool MOTION, bool COUNTERS, bool EXT>
__device__ __forceinline__ void trace_wide(SceneView sv, const float4* __restrict__ rays, uint64_t n,
                                           int4* __restrict__ hits, int32_t* __restrict__ geoms_out,
                                           int4* __restrict__ counters, float* __restrict__ spT,
                                           int32_t* __restrict__ spN, int32_t* __restrict__ err,
                                           const uint32_t* __restrict__ n_dev, const uint32_t* __restrict__ perm,
                                           int refill_min) {
    static_assert(BLOCK % 64 == 0, "whole wavefronts");
    __shared__ float sT[LDS * BLOCK];
    __shared__ int32_t sN[LDS * BLOCK];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const uint64_t gtid = (uint64_t)blockIdx.x * BLOCK + tid;
    const uint64_t nthr = (uint64_t)gridDim.x * BLOCK;
    Stack<BLOCK, LDS> st;
    st.sT = sT + tid; st.sN = sN + tid;
    st.gT = spT; st.gN = spN; st.gstride = nthr; st.gofs = (uint32_t)gtid;

    if (n_dev) n = min(n, (uint64_t)*n_dev);
    // Rays of this wavefront: v = 0, 1, 2, ... -> ray wave0 + (v / 64) * nthr + v % 64 (increasing
    // in v).  A lane whose ray is finished idles (and serves as a worker in the leaf tests) until
    // refill_min lanes are idle; then the idle lanes write their results and take the next rays
    // of the wavefront together, so the ray set-up runs once for many lanes.
    const uint64_t wave0 = gtid - (uint64_t)lane;
    uint32_t cursor = 0;
    bool done = true, live = false, ok = true, fast = false;
    uint32_t ri = 0;   // launches hold fewer than 2^32 rays
    int kp = 0;   // pack_perm of the current ray
    RayS r;
    r.P[0] = 0.0f; r.P[1] = 0.0f; r.P[2] = 0.0f; r.tclosest = 0.0f; r.time = 0.0f;
    {
        const float D1[3] = {1.0f, 1.0f, 1.0f};
        ray_setup(r, D1);
    }
    int nodesT = 0, leafsT = 0, trisT = 0, maxDepth = 0;
    int best_geom = -1, best_prim = -1;
    float bu = 0.0f, bv = 0.0f, bw = 0.0f;
    int mesh_base = 0, gcur = 0, gend = 0;   // mesh_base 0: in the scene level, else inside geom gcur
    // A motion top level starts a geom's stack at the popped leaf's own slot, which can be 0, so
    // motion scenes keep a separate flag (qbvh/motionintersect.go:99-103 against intersect.go:219-226).
    const bool top_motion = MOTION && sv.top_motion;
    bool in_mesh_m = false;
#define VGPU_IN_MESH (MOTION ? in_mesh_m : mesh_base != 0)
    DevGeom g = {0, 0, 1, 0.0f};
    int sp = 0;
    while (true) {
        const uint64_t dm = __ballot(done);
        const uint64_t next = wave0 + (uint64_t)(cursor >> 6) * nthr + (cursor & 63u);
        const bool more = next < n;
        if (more ? __popcll(dm) >= refill_min : dm == ~0ull) {
            if (done) {
                if (live) {
                    if (best_prim == -2) {   // the stack overflowed
                        atomicOr(err, 1);
                        best_geom = -1; best_prim = -1;
                    }
                    const uint64_t oi = perm ? (uint64_t)perm[ri] : (uint64_t)ri;
                    if (EXT) {
                        hits[2 * oi] = make_int4(best_prim, best_geom, __float_as_int(r.tclosest), __float_as_int(bu));
                        hits[2 * oi + 1] = make_int4(__float_as_int(bv), __float_as_int(bw), 0, 0);
                    } else {
                        hits[oi] = make_int4(best_prim, __float_as_int(r.tclosest), __float_as_int(bu), __float_as_int(bv));
                        if (geoms_out) geoms_out[oi] = best_geom;
                    }
                    if (COUNTERS) counters[oi] = make_int4(nodesT, leafsT, trisT, maxDepth);
                    live = false;
                }
                if (more) {
                    const uint32_t v = cursor + (uint32_t)__popcll(dm & ((1ull << lane) - 1ull));
                    const uint64_t rnext = wave0 + (uint64_t)(v >> 6) * nthr + (v & 63u);
                    live = rnext < n;
                    ri = (uint32_t)rnext;
                    if (live) {
                        const float4 ra = rays[2 * ri], rb4 = rays[2 * ri + 1];
                        r.P[0] = ra.x; r.P[1] = ra.y; r.P[2] = ra.z; r.tclosest = ra.w;
                        const float D[3] = {rb4.x, rb4.y, rb4.z};
                        r.time = rb4.w;
                        ray_setup(r, D);
                        kp = pack_perm(r);
                        ok = true;
                        for (int k = 0; k < 3; k++)
                            ok = ok && absf(r.P[k]) < INFINITY && absf(r.Dinv[k]) < INFINITY && r.Dinv[k] != 0.0f;
                        nodesT = 0; leafsT = 0; trisT = 0; maxDepth = 0;
                        best_geom = -1; best_prim = -1;
                        bu = 0.0f; bv = 0.0f; bw = 0.0f;
                        mesh_base = 0;
                        in_mesh_m = false;
                        st.store(0, r.tclosest, 0);
                        sp = 1;
                        done = false;
                    } else {
                        ok = true;
                    }
                }
            }
            if (!more) break;
            cursor += (uint32_t)__popcll(dm);
            // motion boxes are interpolated, and 0 * inf there gives NaN: the exact form only
            fast = !MOTION && sv.nan_free_boxes && __all(ok ? 1 : 0) != 0;
        }
        bool pend = false;   // this lane popped a triangle leaf: [lbase, lbase + lcount) of geom gcur
        int lbase = 0, lcount = 0;
        if (!done) {
            do {
                if (VGPU_IN_MESH && sp == mesh_base) {
                    // PolyMesh.Trace returned; Scene.TraceElems moves to the leaf's next geom (scene.go:65-75)
                    gcur++;
                    const int ng = next_geom<SHADOW>(sv, rays, ri, r, gcur, gend, g, best_geom, best_prim, bu, bv, bw);
                    if (ng == 2) { done = true; break; }
                    if (ng == 1) {
                        st.store(mesh_base, r.tclosest, 0);
                        sp = mesh_base + 1;
                        break;
                    }
                    sp = top_motion ? mesh_base : mesh_base - 1;
                    mesh_base = 0;
                    in_mesh_m = false;
                }
                if (sp == 0) { done = true; break; }
                sp--;
                float T; int32_t node;
                st.load(sp, T, node);
                if (r.tclosest < T || node == -1) break;
                if (node >= 0) {
                    if (COUNTERS) nodesT++;
                    NodeData nd;
                    Box4 b;
                    if constexpr (MOTION) {
                        const bool mnode = in_mesh_m ? g.nkeys > 1 : top_motion;
                        if (mnode) {
                            const int nk = in_mesh_m ? g.nkeys : sv.top_keys;
                            int k0, k1;
                            float t;
                            motion_keys(r.time, nk, k0, k1, t);
                            box_motion(r, sv.mnodes, in_mesh_m ? g.node_base : sv.top_mbase, node, nk, k0, k1, t, nd, b);
                        } else {
                            fetch_static(sv.nodes, in_mesh_m ? g.node_base + node : node, nd);
                            box4<false>(r, nd.lx, nd.ly, nd.lz, nd.hx, nd.hy, nd.hz, b);
                        }
                    } else {
                        fetch_static(sv.nodes, mesh_base != 0 ? g.node_base + node : node, nd);
                        if (fast) box4<true>(r, nd.lx, nd.ly, nd.lz, nd.hx, nd.hy, nd.hz, b);
                        else box4<false>(r, nd.lx, nd.ly, nd.lz, nd.hx, nd.hy, nd.hz, b);
                    }
                    // Push order of qbvh/intersect.go:137-216 (same as motionintersect.go:53-97): the
                    // children are pushed as two pairs; rank(c) is child c's position in that sequence.
                    // Hit children are stored compacted in rank order, i.e. exactly the entries the
                    // reference leaves below its stack top.
                    const int n0 = (kp >> nd.ax[0]) & 1;
                    const int n1 = (kp >> nd.ax[1]) & 1;
                    const int n2 = (kp >> nd.ax[2]) & 1;
                    const int gA = n0 ? 0 : 2, gB = 2 - gA;
                    const int rk0 = gA + (n1 ^ 1), rk1 = gA + n1, rk2 = gB + (n2 ^ 1), rk3 = gB + n2;
                    const int m = (b.h[0] ? 1 << rk0 : 0) | (b.h[1] ? 1 << rk1 : 0) | (b.h[2] ? 1 << rk2 : 0) |
                                  (b.h[3] ? 1 << rk3 : 0);
                    if (sp + 4 > VGPU_MAX_STACK) {   // conservative: the reference writes up to 4 slots here
                        best_prim = -2;
                        done = true;
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
                    if (VGPU_IN_MESH) {
                        if (COUNTERS) trisT += count;
                        pend = true;
                        lbase = base;
                        lcount = count;
                    } else {
                        // Scene.TraceElems: geoms [base, base+count) (scene.go:61-78)
                        const int mb = top_motion ? sp : sp + 1;
                        if (mb >= VGPU_MAX_STACK) { best_prim = -2; done = true; break; }
                        gcur = base; gend = base + count;
                        const int ng = next_geom<SHADOW>(sv, rays, ri, r, gcur, gend, g, best_geom, best_prim, bu, bv, bw);
                        if (ng == 2) { done = true; break; }
                        if (ng == 1) {
                            mesh_base = mb;
                            in_mesh_m = true;
                            st.store(mesh_base, r.tclosest, 0);
                            sp = mesh_base + 1;
                        }
                    }
                }
            } while (false);
        }

        const uint64_t pm = __ballot(pend);
        if (pm != 0 && __ballot(pend && lcount > 2) == 0) {
            // Leaves of one or two triangles only: each owner tests its own (dealing them out
            // costs more than two tests).
            if (pend) {
                int idx = -1;
                float U = 0.0f, V = 0.0f, W = 0.0f;
                const TriRay q = tri_ray_packed(r, kp);
                if (MOTION && g.nkeys > 1) {
                    int k0, k1;
                    float t;
                    motion_keys(r.time, g.nkeys, k0, k1, t);
                    const float4* tp = sv.mtris + g.tri_base + (size_t)lbase * 3 * g.nkeys;
                    for (int i = 0; i < lcount; i++) {
                        float4 v0, v1, v2;
                        motion_tri(tp + (size_t)i * 3 * g.nkeys, k0, k1, t, v0, v1, v2);
                        if (tri_test_q<true>(r, q, v0, v1, v2, g.ray_bias, U, V, W)) idx = lbase + i;
                    }
                } else {
                    const float4* tp = sv.tris + ((size_t)g.tri_base + lbase) * 3;
                    for (int i = 0; i < lcount; i++)
                        if (tri_test_q<false>(r, q, tp[3 * i], tp[3 * i + 1], tp[3 * i + 2], g.ray_bias, U, V, W))
                            idx = lbase + i;
                }
                if (idx >= 0) {
                    best_geom = gcur; best_prim = idx; bu = U; bv = V; bw = W;
                    if (SHADOW) done = true;
                }
            }
        } else if (pm != 0) {
            // Deal the leaves' triangle tests out over the wavefront: inclusive prefix sum of the
            // counts; work item w belongs to the first lane whose prefix exceeds w.
            const int c = pend ? lcount : 0;
            int pin = c;
            for (int d = 1; d < 64; d <<= 1) {
                const int y = __shfl_up(pin, d);
                if (lane >= d) pin += y;
            }
            const int total = __shfl(pin, 63);
            const int pex = pin - c;
            // first triangle of the leaf: a triangle index (static) or a float4 offset (moving mesh)
            const bool mv = MOTION && g.nkeys > 1;
            const int tb = mv ? g.tri_base + lbase * 3 * g.nkeys : g.tri_base + lbase;
            const TriRay q = tri_ray_packed(r, kp);
            const int kpack = q.Kx | (q.Ky << 2) | (q.Kz << 4);
            const float bias = g.ray_bias;
            uint32_t cand = 0;   // owner: leaf triangles that passed the distance-free part of the test
            for (int k0 = 0; k0 < total; k0 += 64) {
                const int w = k0 + lane;
                int o = 0;
#pragma unroll
                for (int s = 32; s >= 1; s >>= 1)
                    if (__shfl(pin, o + s - 1) <= w) o += s;
                const int i = w - __shfl(pex, o);
                TriRay oq;
                oq.S0 = __shfl(q.S0, o); oq.S1 = __shfl(q.S1, o); oq.S2 = __shfl(q.S2, o);
                oq.Pkx = __shfl(q.Pkx, o); oq.Pky = __shfl(q.Pky, o); oq.Pkz = __shfl(q.Pkz, o);
                const int ok = __shfl(kpack, o);
                oq.Kx = ok & 3; oq.Ky = (ok >> 2) & 3; oq.Kz = (ok >> 4) & 3;
                const float obias = __shfl(bias, o);
                const int otb = __shfl(tb, o);
                bool pass = false;
                if constexpr (MOTION) {
                    const int onk = __shfl(mv ? g.nkeys : 1, o);
                    const float otime = __shfl(r.time, o);
                    if (w < total) {
                        TriPre pr;
                        if (onk > 1) {
                            int k0, k1;
                            float t;
                            motion_keys(otime, onk, k0, k1, t);
                            float4 v0, v1, v2;
                            motion_tri(sv.mtris + otb + (size_t)i * 3 * onk, k0, k1, t, v0, v1, v2);
                            pass = tri_pre<true>(oq, v0, v1, v2, obias, pr);
                        } else {
                            const float4* tp = sv.tris + (size_t)(otb + i) * 3;
                            pass = tri_pre<false>(oq, tp[0], tp[1], tp[2], obias, pr);
                        }
                    }
                } else if (w < total) {
                    const float4* tp = sv.tris + (size_t)(otb + i) * 3;
                    TriPre pr;
                    pass = tri_pre<false>(oq, tp[0], tp[1], tp[2], obias, pr);
                }
                const uint64_t cm = __ballot(pass);
                if (pend && cm != 0) {
                    const int s0 = max(pex, k0), e0 = min(pin, k0 + 64);
                    if (e0 > s0)
                        cand |= (uint32_t)((cm >> (s0 - k0)) & ((1ull << (e0 - s0)) - 1ull)) << (s0 - pex);
                }
            }
            if (cand != 0) {
                // the owner's sequential pass over the candidates, in leaf order (trace.go:108-122)
                int idx = -1;
                float U = 0.0f, V = 0.0f, W = 0.0f;
                if (mv) {
                    int k0, k1;
                    float t;
                    motion_keys(r.time, g.nkeys, k0, k1, t);
                    while (cand != 0) {
                        const int i = __ffs(cand) - 1;
                        cand &= cand - 1;
                        float4 v0, v1, v2;
                        motion_tri(sv.mtris + tb + (size_t)i * 3 * g.nkeys, k0, k1, t, v0, v1, v2);
                        if (tri_test_q<true>(r, q, v0, v1, v2, g.ray_bias, U, V, W)) idx = lbase + i;
                    }
                } else {
                    const float4* tp = sv.tris + (size_t)tb * 3;
                    while (cand != 0) {
                        const int i = __ffs(cand) - 1;
                        cand &= cand - 1;
                        if (tri_test_q<false>(r, q, tp[3 * i], tp[3 * i + 1], tp[3 * i + 2], g.ray_bias, U, V, W))
                            idx = lbase + i;
                    }
                }
                if (idx >= 0) {
                    best_geom = gcur; best_prim = idx; bu = U; bv = V; bw = W;
                    if (SHADOW) done = true;   // qbvh/intersect.go:232-236 at both levels
                }
            }
        }
    }
}

#undef VGPU_IN_MESH

#define VGPU_TRACE_PARAMS                                                                                      \
    SceneView sv, const float4 *__restrict__ rays, uint64_t n, int4 *__restrict__ hits,                       \
        int32_t *__restrict__ geoms_out, int4 *__restrict__ counters, float *__restrict__ spT,                \
        int32_t *__restrict__ spN, int32_t *__restrict__ err, const uint32_t *__restrict__ n_dev,              \
        const uint32_t *__restrict__ perm, int refill_min
#define VGPU_TRACE_ARGS sv, rays, n, hits, geoms_out, counters, spT, spN, err, n_dev, perm, refill_min

template <int BLOCK, int LDS, bool SHADOW, bool MOTION, bool COUNTERS, bool EXT>
__global__ __launch_bounds__(BLOCK) void k_trace_wide(VGPU_TRACE_PARAMS) {
    trace_wide<BLOCK, LDS, SHADOW, MOTION, COUNTERS, EXT>(VGPU_TRACE_ARGS);
}

// Motion scenes with a short LDS stack (8 or 12 entries): the box and triangle interpolation leave
// the plain kernel at 103-107 VGPRs, four waves per SIMD.  Held to 96 (a few dwords of spill), five
// waves fit, and with a grid of five waves per SIMD (waves_per_cu 20) the 2000-moving-sphere camera
// and shadow rays run 8-10% faster than the best four-wave shape.
template <int BLOCK, int LDS, bool SHADOW, bool COUNTERS, bool EXT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_trace_wide_m5(VGPU_TRACE_PARAMS) {
    static_assert(LDS <= 12, "the five-wave motion kernel is built for short LDS stacks");
    trace_wide<BLOCK, LDS, SHADOW, true, COUNTERS, EXT>(VGPU_TRACE_ARGS);
}

#undef VGPU_TRACE_PARAMS
#undef VGPU_TRACE_ARGS


// ---------------------------------------------------------------- scenes with transforms
//
// Scenes holding GeomInstances, Procs or meshes with their own Transform run this kernel: one ray
// per lane, the reference's recursion written out (Scene.Trace -> Geom.Trace -> qbvh.Trace ->
// TraceElems), each level on the same stack above the level that called it.  A geom with a
// transform moves the ray into its object space (P and D by the inverse transform at the ray's time,
// D not renormalised, then Setup), traces what it holds and restores the ray, Tclosest excepted
// (instance.go:78-104, proc.go:93-131, polymesh/trace.go:20-74).

// The ray in object space: P and D moved by Minv, then Ray.Setup.
__device__ __forceinline__ void ray_to_object(const RayS& r, const float* D, const float* Minv, RayS& o, float* Do) {
    xf_point(Minv, r.P, o.P);
    xf_vec(Minv, D, Do);
    ray_setup(o, Do);
    o.time = r.time;
    o.tclosest = r.tclosest;
}

struct XHit {
    int prim;
    float U, V, W;
};

// qbvh.Trace / TraceMotion over one mesh (or the analytic sphere test) with the ray in the mesh's
// space, its stack starting at entry `base`; true when it improved the hit.  -1 in *over: the stack
// overflowed.
template <bool SHADOW>
__device__ bool trace_mesh(const SceneView& sv, const DevGeom& g, RayS& r, const float* D, Stack<64, 8>& st, int base,
                           XHit& h, int* over) {
    if (g.nkeys == 0) {   // analytic sphere (builtin/geom/sphere/trace.go:13-52)
        float t;
        const float4 d = make_float4(D[0], D[1], D[2], 0.0f);
        if (!sphere_geom_hit(r.P, d, sv.spheres[g.tri_base], r.tclosest, t)) return false;
        r.tclosest = t;
        h.prim = 0; h.U = 0.0f; h.V = 0.0f; h.W = 0.0f;
        return true;
    }
    const int kp = pack_perm(r);
    const TriRay q = tri_ray_packed(r, kp);
    int k0 = 0, k1 = 0;
    float mt = 0.0f;
    if (g.nkeys > 1) motion_keys(r.time, g.nkeys, k0, k1, mt);
    bool hit = false;
    st.store(base, r.tclosest, 0);
    int sp = base + 1;
    while (sp > base) {
        sp--;
        float T; int32_t node;
        st.load(sp, T, node);
        if (r.tclosest < T || node == -1) continue;
        if (node >= 0) {
            NodeData nd;
            Box4 b;
            if (g.nkeys > 1) {
                box_motion(r, sv.mnodes, g.node_base, node, g.nkeys, k0, k1, mt, nd, b);
            } else {
                fetch_static(sv.nodes, g.node_base + node, nd);
                box4<false>(r, nd.lx, nd.ly, nd.lz, nd.hx, nd.hy, nd.hz, b);
            }
            const int n0 = (kp >> nd.ax[0]) & 1, n1 = (kp >> nd.ax[1]) & 1, n2 = (kp >> nd.ax[2]) & 1;
            const int gA = n0 ? 0 : 2, gB = 2 - gA;
            const int rk[4] = {gA + (n1 ^ 1), gA + n1, gB + (n2 ^ 1), gB + n2};
            if (sp + 4 > VGPU_MAX_STACK) { *over = 1; return hit; }
            int m = 0;
            for (int c = 0; c < 4; c++) if (b.h[c]) m |= 1 << rk[c];
            for (int c = 0; c < 4; c++)
                if (b.h[c]) st.store(sp + __popc(m & ((1 << rk[c]) - 1)), b.t[c], nd.ch[c]);
            sp += __popc(m);
        } else {
            const int lb = leaf_base(node), lc = leaf_count(node);
            bool leaf_hit = false;
            for (int i = 0; i < lc; i++) {
                float U, V, W;
                bool ok;
                if (g.nkeys > 1) {
                    float4 v0, v1, v2;
                    motion_tri(sv.mtris + g.tri_base + (size_t)(lb + i) * 3 * g.nkeys, k0, k1, mt, v0, v1, v2);
                    ok = tri_test_q<true>(r, q, v0, v1, v2, g.ray_bias, U, V, W);
                } else {
                    const float4* tp = sv.tris + ((size_t)g.tri_base + lb + i) * 3;
                    ok = tri_test_q<false>(r, q, tp[0], tp[1], tp[2], g.ray_bias, U, V, W);
                }
                if (ok) { h.prim = lb + i; h.U = U; h.V = V; h.W = W; leaf_hit = true; }
            }
            if (leaf_hit) {
                hit = true;
                if (SHADOW) return true;   // qbvh/intersect.go:232-236
            }
        }
    }
    return hit;
}

// Geom.Trace of a mesh or sphere slot, through its own Transform when it has one.
template <bool SHADOW>
__device__ bool trace_leaf_geom(const SceneView& sv, int slot, RayS& r, const float* D, Stack<64, 8>& st, int base,
                                XHit& h, int* over) {
    const DevGeom g = sv.geoms[slot];
    const DevXform x = sv.xforms[slot];
    if (x.kind != kXfMesh) return trace_mesh<SHADOW>(sv, g, r, D, st, base, h, over);
    float M[16], Minv[16];
    xform_at(x, sv.xkeys, r.time, M, Minv);
    RayS o;
    float Do[3];
    ray_to_object(r, D, Minv, o, Do);
    const bool hit = trace_mesh<SHADOW>(sv, g, o, Do, st, base, h, over);
    r.tclosest = o.tclosest;
    return hit;
}

// Geom.Trace of a top-level slot: a mesh or sphere, or an instance or proc over its geoms.  *leaf:
// the slot whose element was hit.  A proc stops at the first geom reporting a hit unless the ray is
// a shadow ray (proc.go:112-120).
template <bool SHADOW>
__device__ bool trace_geom(const SceneView& sv, int slot, RayS& r, const float* D, Stack<64, 8>& st, int base, XHit& h,
                           int& leaf, int* over) {
    const DevXform x = sv.xforms[slot];
    if (x.kind != kXfInstance && x.kind != kXfProc) {
        const bool hit = trace_leaf_geom<SHADOW>(sv, slot, r, D, st, base, h, over);
        if (hit) leaf = slot;
        return hit;
    }
    float M[16], Minv[16];
    xform_at(x, sv.xkeys, r.time, M, Minv);
    RayS o;
    float Do[3];
    ray_to_object(r, D, Minv, o, Do);
    bool hit = false;
    for (int c = 0; c < x.nchild && !*over; c++) {
        const int child = sv.xchildren[x.child_base + c];
        if (trace_leaf_geom<SHADOW>(sv, child, o, Do, st, base, h, over)) {
            hit = true;
            leaf = child;
            if (x.kind == kXfInstance || !SHADOW) break;
        }
    }
    r.tclosest = o.tclosest;
    return hit;
}

// The whole traversal for one ray per lane.  EXT: 32-byte records {prim, geom, t, U, V, W, leaf slot, 0}.
template <bool SHADOW, bool EXT>
__global__ __launch_bounds__(64) void k_trace_xform(SceneView sv, const float4* __restrict__ rays, uint64_t n,
                                                    int4* __restrict__ hits, int32_t* __restrict__ geoms_out,
                                                    float* __restrict__ spT, int32_t* __restrict__ spN,
                                                    int32_t* __restrict__ err, const uint32_t* __restrict__ n_dev,
                                                    const uint32_t* __restrict__ perm) {
    __shared__ float sT[8 * 64];
    __shared__ int32_t sN[8 * 64];
    const int tid = threadIdx.x;
    const uint64_t gtid = (uint64_t)blockIdx.x * 64 + tid;
    const uint64_t nthr = (uint64_t)gridDim.x * 64;
    Stack<64, 8> st;
    st.sT = sT + tid; st.sN = sN + tid;
    st.gT = spT; st.gN = spN; st.gstride = nthr; st.gofs = (uint32_t)gtid;
    if (n_dev) n = min(n, (uint64_t)*n_dev);
    for (uint64_t ri = gtid; ri < n; ri += nthr) {
        const float4 ra = rays[2 * ri], rb = rays[2 * ri + 1];
        RayS r;
        r.P[0] = ra.x; r.P[1] = ra.y; r.P[2] = ra.z;
        r.tclosest = ra.w;
        r.time = rb.w;
        const float D[3] = {rb.x, rb.y, rb.z};
        ray_setup(r, D);
        const int kp = pack_perm(r);
        XHit h{-1, 0.0f, 0.0f, 0.0f};
        int best_geom = -1, best_leaf = -1, over = 0;
        const bool top_motion = sv.top_motion != 0;
        int k0 = 0, k1 = 0;
        float mt = 0.0f;
        if (top_motion) motion_keys(r.time, sv.top_keys, k0, k1, mt);
        st.store(0, r.tclosest, 0);
        int sp = 1;
        bool done = false;
        while (sp > 0 && !done && !over) {
            sp--;
            float T; int32_t node;
            st.load(sp, T, node);
            if (r.tclosest < T || node == -1) continue;
            if (node >= 0) {
                NodeData nd;
                Box4 b;
                if (top_motion) {
                    box_motion(r, sv.mnodes, sv.top_mbase, node, sv.top_keys, k0, k1, mt, nd, b);
                } else {
                    fetch_static(sv.nodes, node, nd);
                    box4<false>(r, nd.lx, nd.ly, nd.lz, nd.hx, nd.hy, nd.hz, b);
                }
                const int n0 = (kp >> nd.ax[0]) & 1, n1 = (kp >> nd.ax[1]) & 1, n2 = (kp >> nd.ax[2]) & 1;
                const int gA = n0 ? 0 : 2, gB = 2 - gA;
                const int rk[4] = {gA + (n1 ^ 1), gA + n1, gB + (n2 ^ 1), gB + n2};
                if (sp + 4 > VGPU_MAX_STACK) { over = 1; break; }
                int m = 0;
                for (int c = 0; c < 4; c++) if (b.h[c]) m |= 1 << rk[c];
                for (int c = 0; c < 4; c++)
                    if (b.h[c]) st.store(sp + __popc(m & ((1 << rk[c]) - 1)), b.t[c], nd.ch[c]);
                sp += __popc(m);
            } else {
                // Scene.TraceElems over the leaf's geoms (scene.go:61-78)
                const int lb = leaf_base(node), lc = leaf_count(node);
                for (int gi = lb; gi < lb + lc && !over; gi++) {
                    int leaf = -1;
                    if (trace_geom<SHADOW>(sv, gi, r, D, st, sp, h, leaf, &over)) {
                        best_geom = gi;
                        best_leaf = leaf;
                        if (SHADOW) { done = true; break; }
                    }
                }
            }
        }
        if (over) {
            atomicOr(err, 1);
            best_geom = -1; best_leaf = -1; h.prim = -1;
        }
        if (best_geom < 0) h.prim = -1;
        const uint64_t oi = perm ? (uint64_t)perm[ri] : ri;
        if (EXT) {
            hits[2 * oi] = make_int4(h.prim, best_geom, __float_as_int(r.tclosest), __float_as_int(h.U));
            hits[2 * oi + 1] = make_int4(__float_as_int(h.V), __float_as_int(h.W), best_leaf, 0);
        } else {
            hits[oi] = make_int4(h.prim, __float_as_int(r.tclosest), __float_as_int(h.U), __float_as_int(h.V));
            if (geoms_out) geoms_out[oi] = best_geom;
        }
    }
}

template <bool SHADOW, bool EXT>
hipError_t launch_xform(const SceneView& sv, const void* rays, uint64_t n, void* hits, int32_t* geoms, float* spT,
                        int32_t* spN, uint64_t threads, int32_t* err, const uint32_t* n_dev, const uint32_t* perm,
                        hipStream_t stream) {
    hipLaunchKernelGGL((k_trace_xform<SHADOW, EXT>), dim3((unsigned)(threads / 64)), dim3(64), 0, stream, sv,
                       (const float4*)rays, n, (int4*)hits, geoms, spT, spN, err, n_dev, perm);
    return hipGetLastError();
}

template <int BLOCK, int LDS, bool SHADOW, bool MOTION>
hipError_t launch_cfg(const SceneView& sv, const void* rays, uint64_t n, void* hits, int32_t* geoms, int32_t* counters,
                      float* spT, int32_t* spN, uint64_t grid, int32_t* err, const uint32_t* n_dev,
                      const uint32_t* perm, hipStream_t stream, int refill_min) {
    if constexpr (MOTION && LDS <= 12) {
        if (counters)
            hipLaunchKernelGGL((k_trace_wide_m5<BLOCK, LDS, SHADOW, true, false>), dim3((unsigned)grid), dim3(BLOCK), 0,
                               stream, sv, (const float4*)rays, n, (int4*)hits, geoms, (int4*)counters, spT, spN, err,
                               n_dev, perm, refill_min);
        else
            hipLaunchKernelGGL((k_trace_wide_m5<BLOCK, LDS, SHADOW, false, false>), dim3((unsigned)grid), dim3(BLOCK), 0,
                               stream, sv, (const float4*)rays, n, (int4*)hits, geoms, (int4*)nullptr, spT, spN, err,
                               n_dev, perm, refill_min);
        return hipGetLastError();
    }
    if (counters)
        hipLaunchKernelGGL((k_trace_wide<BLOCK, LDS, SHADOW, MOTION, true, false>), dim3((unsigned)grid), dim3(BLOCK), 0,
                           stream, sv, (const float4*)rays, n, (int4*)hits, geoms, (int4*)counters, spT, spN, err,
                           n_dev, perm, refill_min);
    else
        hipLaunchKernelGGL((k_trace_wide<BLOCK, LDS, SHADOW, MOTION, false, false>), dim3((unsigned)grid), dim3(BLOCK), 0,
                           stream, sv, (const float4*)rays, n, (int4*)hits, geoms, (int4*)nullptr, spT, spN, err,
                           n_dev, perm, refill_min);
    return hipGetLastError();
}

template <int BLOCK, int LDS>
hipError_t launch_mode(const SceneView& sv, bool shadow, bool motion, const void* rays, uint64_t n, void* hits,
                       int32_t* geoms, int32_t* counters, float* spT, int32_t* spN, uint64_t grid, int32_t* err,
                       const uint32_t* n_dev, const uint32_t* perm, hipStream_t stream, int refill_min) {
    if (shadow) {
        if (motion)
            return launch_cfg<BLOCK, LDS, true, true>(sv, rays, n, hits, geoms, counters, spT, spN, grid, err, n_dev, perm, stream, refill_min);
        return launch_cfg<BLOCK, LDS, true, false>(sv, rays, n, hits, geoms, counters, spT, spN, grid, err, n_dev, perm, stream, refill_min);
    }
    if (motion)
        return launch_cfg<BLOCK, LDS, false, true>(sv, rays, n, hits, geoms, counters, spT, spN, grid, err, n_dev, perm, stream, refill_min);
    return launch_cfg<BLOCK, LDS, false, false>(sv, rays, n, hits, geoms, counters, spT, spN, grid, err, n_dev, perm, stream, refill_min);
}

}  // namespace

hipError_t launch_trace_ext(const SceneView& sv, bool motion, const void* rays, uint64_t n, void* hits32, float* spT,
                            int32_t* spN, uint64_t spill_threads, int32_t* err, int num_cus, hipStream_t stream,
                            const uint32_t* n_dev) {
    if (n == 0) return hipSuccess;
    if (n > 0xffffffffull) return hipErrorInvalidValue;
    // the block size is the kernel's launch bound: each shape is a template constant
    constexpr int kB = kExtLaunch.block, kL = kExtLaunch.lds_depth;
    constexpr int kBM = kExtLaunchMotion.block, kLM = kExtLaunchMotion.lds_depth;
    const TraceLaunch cfg = motion ? kExtLaunchMotion : kExtLaunch;
    const uint64_t threads = trace_grid_threads(cfg, n, num_cus);
    if (threads > spill_threads) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)(threads / cfg.block);
    if (sv.any_xform)
        return launch_xform<false, true>(sv, rays, n, hits32, nullptr, spT, spN, threads, err, n_dev, nullptr, stream);
    if (motion)
        hipLaunchKernelGGL((k_trace_wide<kBM, kLM, false, true, false, true>), dim3(grid), dim3(kBM), 0, stream, sv,
                           (const float4*)rays, n, (int4*)hits32, (int32_t*)nullptr, (int4*)nullptr, spT, spN, err, n_dev,
                           (const uint32_t*)nullptr, kRefillDefault);
    else
        hipLaunchKernelGGL((k_trace_wide<kB, kL, false, false, false, true>), dim3(grid), dim3(kB), 0, stream, sv,
                           (const float4*)rays, n, (int4*)hits32, (int32_t*)nullptr, (int4*)nullptr, spT, spN, err, n_dev,
                           (const uint32_t*)nullptr, kRefillDefault);
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
    if (n > 0xffffffffull) return hipErrorInvalidValue;
    const uint64_t threads = trace_grid_threads(cfg, n, num_cus);
    if (threads > spill_threads) return hipErrorInvalidValue;
    const uint64_t grid = threads / cfg.block;
    if (sv.any_xform) {   // the per-lane kernel (block 64; threads is a multiple of 64); no counters
        if (counters) return hipErrorInvalidValue;
        return shadow ? launch_xform<true, false>(sv, rays, n, hits, geoms, spT, spN, threads, err, n_dev, perm, stream)
                      : launch_xform<false, false>(sv, rays, n, hits, geoms, spT, spN, threads, err, n_dev, perm, stream);
    }
#define VGPU_CASE(B, L)                                                                                         \
    if (cfg.block == B && cfg.lds_depth == L)                                                                   \
        return launch_mode<B, L>(sv, shadow, motion, rays, n, hits, geoms, counters, spT, spN, grid, err, n_dev, perm, \
                                 stream, cfg.refill_min);
    VGPU_CASE(64, 8) VGPU_CASE(64, 12) VGPU_CASE(64, 16) VGPU_CASE(64, 24) VGPU_CASE(64, 32)
    VGPU_CASE(128, 8) VGPU_CASE(128, 12) VGPU_CASE(128, 16) VGPU_CASE(128, 24) VGPU_CASE(128, 32)
    VGPU_CASE(256, 8) VGPU_CASE(256, 12) VGPU_CASE(256, 16) VGPU_CASE(256, 24) VGPU_CASE(256, 32)
#undef VGPU_CASE
    return hipErrorInvalidValue;
}

// ------------------------------------------------------------- allocation-free traversal
//
// vgpu_trace_noalloc keeps its traversal spill area in this file-scope array instead of device
// memory the scene allocates.  Layout: the error flag in the first 256 bytes, then the spill
// depths T and nodes N, each (VGPU_MAX_STACK - lds) entries per launched thread, depth-major as
// the Stack struct addresses them.  The array's host-side name is not a device address: host code
// uses only the pointer hipGetSymbolAddress returned for the current device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

__global__ void k_clear_flag(int32_t* flag) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *flag = 0;
}

}  // namespace

hipError_t noalloc_init(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;   // the caller makes `device` current
    void* ptr = nullptr;
    e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

uint64_t noalloc_max_threads(int lds_depth) {
    const uint64_t per = (uint64_t)(VGPU_MAX_STACK - lds_depth) * (sizeof(float) + sizeof(int32_t));
    return (kNoAllocScratchBytes - 256) / per;
}

hipError_t launch_trace_noalloc(int device, const SceneView& sv, bool shadow, bool motion, const TraceLaunch& cfg,
                                const void* rays, uint64_t n, void* hits, int32_t* geoms, int32_t* counters,
                                int32_t* status, int num_cus, hipStream_t stream) {
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    unsigned char* base;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = (unsigned char*)g_scratch_ptr[device];
    }
    if (!base) return hipErrorNotInitialized;
    const uint64_t threads = trace_grid_threads(cfg, n, num_cus);
    if (threads > noalloc_max_threads(cfg.lds_depth)) return hipErrorInvalidValue;
    int32_t* flag = status ? status : (int32_t*)base;
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, stream, flag);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || n == 0) return e;
    const uint64_t per = (uint64_t)(VGPU_MAX_STACK - cfg.lds_depth);
    float* spT = (float*)(base + 256);
    int32_t* spN = (int32_t*)(base + 256 + per * threads * sizeof(float));
    return launch_trace(sv, shadow, motion, cfg, rays, n, hits, geoms, counters, spT, spN, threads, flag, num_cus,
                        stream, nullptr, nullptr);
}

}  // namespace vgpu
