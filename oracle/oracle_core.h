// ORACLE — test infrastructure only (see oracle_accel.cpp for the scope rules).
// Shared restatement of the reference traversal path, included by the oracle's
// translation units.
#pragma once

#include <cstdint>
#include <cstring>
#include <cmath>
#include <cstdio>
#include <vector>
#include <limits>
#include <xmmintrin.h>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------- scalar math

static inline float bits2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static inline uint32_t f2bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

// MAXSS X1,X0 : X0 = (x > y) ? x : y   (math/dim_amd64.s:8-13)
static inline float Max(float x, float y) { return (x > y) ? x : y; }
// MINSS X1,X0 : X0 = (x < y) ? x : y   (math/dim_amd64.s:16-21)
static inline float Min(float x, float y) { return (x < y) ? x : y; }
// math/abs.go:13-15
static inline float Abs(float x) { return bits2f(f2bits(x) & 0x7fffffffu); }
// math/unsafe.go:15-37
static inline uint32_t SignMask(float v) { return f2bits(v) & 0x80000000u; }
static inline float Xorf(float v, uint32_t i) { return bits2f(f2bits(v) ^ i); }
// ROUNDSS $1 / $2 (math/floor_amd64.s:12-21)
static inline float Floor(float x) { return std::floor(x); }
static inline float Ceil(float x) { return std::ceil(x); }

static const float kInf = std::numeric_limits<float>::infinity();
// math/ferror.go:16,19 (decimal constants rounded to float32: 2^-23 and 2^-24)
static const float EpsilonFloat32 = 1.19209290E-07f;
static const float MachineEpsilon32 = 5.9604645e-08f;
// math/ferror.go:22-24
static inline float Gamma(int32_t n) {
    float ne = (float)n * MachineEpsilon32;
    return ne / (1.0f - ne);
}

// Normalize estimate mode (see Mesh::Normalize); process-wide.
inline bool& normalize_ieee() { static bool v = false; return v; }

struct Vec3 { float v[3]; float& operator[](int i) { return v[i]; } float operator[](int i) const { return v[i]; } };

// math/vec3.go:52-58
static inline Vec3 Lerp(const Vec3& a, const Vec3& b, float t) {
    Vec3 r;
    for (int k = 0; k < 3; k++) r[k] = (1.0f - t) * a[k] + t * b[k];
    return r;
}

// ------------------------------------------------------------ bounding boxes

struct Box {
    float b[2][3];
    void Reset() { for (int i = 0; i < 3; i++) { b[0][i] = kInf; b[1][i] = -kInf; } }
    // math/boundingbox.go:107-116 (note the operand order differs from GrowBox)
    void Grow(float X, float Y, float Z) {
        b[0][0] = Min(X, b[0][0]); b[1][0] = Max(X, b[1][0]);
        b[0][1] = Min(Y, b[0][1]); b[1][1] = Max(Y, b[1][1]);
        b[0][2] = Min(Z, b[0][2]); b[1][2] = Max(Z, b[1][2]);
    }
    void GrowVec3(const Vec3& p) { Grow(p[0], p[1], p[2]); }
    // math/boundingbox.go:99-104
    void GrowBox(const Box& p) {
        for (int k = 0; k < 3; k++) { b[0][k] = Min(b[0][k], p.b[0][k]); b[1][k] = Max(b[1][k], p.b[1][k]); }
    }
    float Dim(int a) const { return b[1][a] - b[0][a]; }
    // math/boundingbox.go:17-19
    float SurfaceArea() const { return Dim(0) * Dim(1) * 2.0f + Dim(1) * Dim(2) * 2.0f + Dim(0) * Dim(2) * 2.0f; }
    // math/boundingbox.go:54-70
    int LongestAxis() const {
        if (Dim(0) < Dim(1)) return (Dim(1) < Dim(2)) ? 2 : 1;
        return (Dim(0) < Dim(2)) ? 2 : 0;
    }
    // math/boundingbox.go:27-32
    Vec3 Centroid() const {
        Vec3 o;
        for (int k = 0; k < 3; k++) o[k] = (b[1][k] + b[0][k]) * 0.5f;
        return o;
    }
};

static Box InfBox() { Box b; for (int i = 0; i < 2; i++) for (int k = 0; k < 3; k++) b.b[i][k] = kInf; return b; }

// math/boundingbox.go:119-127
static Box BoxLerp(const Box& a, const Box& c, float t) {
    Box o;
    for (int i = 0; i < 2; i++) for (int k = 0; k < 3; k++) o.b[i][k] = (1 - t) * a.b[i][k] + t * c.b[i][k];
    return o;
}

// --------------------------------------------------------------- QBVH nodes

// qbvh/qbvh.go:31-36 — 128-byte node, boxes SoA at idx + i*12 + k*4.
struct Node {
    float boxes[24];
    uint32_t axis0, axis1, axis2;
    int32_t children[4];
    int32_t parent;
};
static_assert(sizeof(Node) == 128, "node layout");

// qbvh/mqbvh.go:24-29 (topology only; boxes per key are separate)
struct MNode {
    int32_t axis0, axis1, axis2;
    int32_t children[4];
    int32_t parent;
    uint32_t pad[2];
};
typedef float MBoxes[24];

static inline void SetBounds(float* boxes, int idx, const Box& bb) {
    for (int i = 0; i < 2; i++) for (int k = 0; k < 3; k++) boxes[idx + i * 12 + k * 4] = bb.b[i][k];
}
static inline Box GetBounds(const float* boxes, int idx) {
    Box bb;
    for (int i = 0; i < 2; i++) for (int k = 0; k < 3; k++) bb.b[i][k] = boxes[idx + i * 12 + k * 4];
    return bb;
}
// qbvh/qbvh.go:61-64
static inline int LeafCount(int32_t l) { return (int)((l & 0xf) + 1); }
static inline int LeafBase(int32_t l) { return (int)((l & 0x7ffffff) >> 4); }
static inline int32_t LeafCode(uint32_t first, uint32_t count) {
    uint32_t v = (1u << 31) | ((first << 4) & 0xfffffff0u) | ((count - 1) & 0xfu);
    return (int32_t)v;
}
// qbvh/qbvh.go:67-93
static void NodeSetLeaf(Node& n, int idx, uint32_t first, uint32_t count) {
    if (count == 0) { n.children[idx] = -1; SetBounds(n.boxes, idx, InfBox()); return; }
    n.children[idx] = LeafCode(first, count);
}
// qbvh/mqbvh.go:62-78
static void MNodeSetLeaf(MNode& n, int idx, uint32_t first, uint32_t count) {
    if (count == 0) { n.children[idx] = -1; return; }
    n.children[idx] = LeafCode(first, count);
}

// ----------------------------------------------------------------- builder

struct BuildError {};

static const int nbins = 8;

// qbvh/build.go:28-136
static void calcMinCost(const Box& bounds, Vec3* centroids, Box* boxes, int32_t* indxs, int n, int& axis_out, int& pivot_out) {
    Box binBounds[nbins];
    int32_t binN[nbins] = {0};
    int axis = bounds.LongestAxis();
    if (bounds.b[1][axis] == bounds.b[0][axis]) { axis_out = axis; pivot_out = n / 2 + 1; return; }
    const float k1 = 7.99952f / (bounds.b[1][axis] - bounds.b[0][axis]);   // float32(8)*(1-0.00006) folded exactly, then rounded
    const float k0 = bounds.b[0][axis];
    for (int i = 0; i < nbins; i++) binBounds[i].Reset();
    for (int i = 0; i < n; i++) {
        float c = centroids[i][axis];
        int bin = (int)(k1 * (c - k0));
        if (bin < 0 || bin > nbins - 1) throw BuildError();
        binN[bin]++;
        binBounds[bin].GrowBox(boxes[i]);
    }
    Box lbox[nbins], rbox[nbins];
    int32_t lN[nbins], rN[nbins];
    Box box; box.Reset();
    int32_t cnt = 0;
    for (int i = 0; i < nbins; i++) { box.GrowBox(binBounds[i]); cnt += binN[i]; lbox[i] = box; lN[i] = cnt; }
    box.Reset(); cnt = 0;
    for (int i = 0; i < nbins; i++) {
        box.GrowBox(binBounds[nbins - 1 - i]); cnt += binN[nbins - 1 - i];
        rbox[nbins - 1 - i] = box; rN[nbins - 1 - i] = cnt;
    }
    int binMinCost = -1;
    float minCost = kInf;
    for (int i = 1; i < nbins; i++) {
        float cost = lbox[i - 1].SurfaceArea() * (float)lN[i - 1] + rbox[i].SurfaceArea() * (float)rN[i];
        if (cost < minCost) { binMinCost = i; minCost = cost; }
    }
    int left = 0, right = n - 1;
    while (left <= right) {
        float c = centroids[left][axis];
        int bin = (int)(k1 * (c - k0));
        if (bin < binMinCost) {
            left++;
        } else {
            std::swap(indxs[left], indxs[right]);
            std::swap(centroids[left], centroids[right]);
            std::swap(boxes[left], boxes[right]);
            right--;
        }
    }
    axis_out = axis; pivot_out = left;
}

// qbvh/build.go:138-149
static Box calcBox(const Box* boxes, int n) {
    Box b; b.Reset();
    for (int i = 0; i < n; i++) b.GrowBox(boxes[i]);
    return b;
}

// qbvh/build.go:151-178
static void binarySplit(Box* boxes, Vec3* centroids, int leafMax, int32_t* indxs, int n, int& axis, int& pivot) {
    Box bounds; bounds.Reset();
    for (int i = 0; i < n; i++) bounds.GrowVec3(centroids[i]);
    if (n <= leafMax) { axis = 0; pivot = n; return; }
    calcMinCost(bounds, centroids, boxes, indxs, n, axis, pivot);
}

// qbvh/build.go:180-287
static int32_t buildAccelRec(std::vector<Node>& nodes, Box* boxes, Vec3* centroids, int32_t* indxs, int n,
                             int leafMax, int baseidx, Box& nodebox_out, int depth) {
    if (depth > 4096) throw BuildError();   // the reference recurses forever on degenerate input
    int axis0, pivot0, axis1, pivot1, axis2, pivot2;
    binarySplit(boxes, centroids, leafMax, indxs, n, axis0, pivot0);
    binarySplit(boxes, centroids, leafMax, indxs, pivot0, axis1, pivot1);
    binarySplit(boxes + pivot0, centroids + pivot0, leafMax, indxs + pivot0, n - pivot0, axis2, pivot2);
    int32_t nodei = (int32_t)nodes.size();
    nodes.push_back(Node());
    std::memset(&nodes[nodei], 0, sizeof(Node));
    nodes[nodei].axis0 = (uint32_t)axis0; nodes[nodei].axis1 = (uint32_t)axis1; nodes[nodei].axis2 = (uint32_t)axis2;
    const int lo[4] = {0, pivot1, pivot0, pivot0 + pivot2};
    const int hi[4] = {pivot1, pivot0, pivot0 + pivot2, n};
    for (int c = 0; c < 4; c++) {
        int cn = hi[c] - lo[c];
        if (cn <= leafMax) {
            SetBounds(nodes[nodei].boxes, c, calcBox(boxes + lo[c], cn));
            NodeSetLeaf(nodes[nodei], c, (uint32_t)(baseidx + lo[c]), (uint32_t)cn);
        } else {
            Box cb;
            int32_t child = buildAccelRec(nodes, boxes + lo[c], centroids + lo[c], indxs + lo[c], cn, leafMax, baseidx + lo[c], cb, depth + 1);
            SetBounds(nodes[nodei].boxes, c, cb);
            nodes[nodei].children[c] = child;
        }
    }
    Box nb; nb.Reset();
    for (int i = 0; i < 4; i++)
        if (nodes[nodei].children[i] != -1) nb.GrowBox(GetBounds(nodes[nodei].boxes, i));
    nodebox_out = nb;
    return nodei;
}

// qbvh/build.go:293-307
static Box BuildAccel(std::vector<Node>& nodes, Box* boxes, Vec3* centroids, int32_t* indxs, int n, int leafMax) {
    if (leafMax > 16) leafMax = 16;
    if (leafMax < 1) leafMax = 1;
    Box bounds;
    buildAccelRec(nodes, boxes, centroids, indxs, n, leafMax, 0, bounds, 0);
    return bounds;
}

// qbvh/motionbuild.go:12-100
static int32_t buildAccelMotionRec(std::vector<MNode>& nodes, Box* boxes, Vec3* centroids, int32_t* indxs, int n,
                                   int leafMax, int baseidx, int depth) {
    if (depth > 4096) throw BuildError();
    int axis0, pivot0, axis1, pivot1, axis2, pivot2;
    binarySplit(boxes, centroids, leafMax, indxs, n, axis0, pivot0);
    binarySplit(boxes, centroids, leafMax, indxs, pivot0, axis1, pivot1);
    binarySplit(boxes + pivot0, centroids + pivot0, leafMax, indxs + pivot0, n - pivot0, axis2, pivot2);
    int32_t nodei = (int32_t)nodes.size();
    nodes.push_back(MNode());
    std::memset(&nodes[nodei], 0, sizeof(MNode));
    nodes[nodei].axis0 = axis0; nodes[nodei].axis1 = axis1; nodes[nodei].axis2 = axis2;
    const int lo[4] = {0, pivot1, pivot0, pivot0 + pivot2};
    const int hi[4] = {pivot1, pivot0, pivot0 + pivot2, n};
    for (int c = 0; c < 4; c++) {
        int cn = hi[c] - lo[c];
        if (cn <= leafMax) {
            MNodeSetLeaf(nodes[nodei], c, (uint32_t)(baseidx + lo[c]), (uint32_t)cn);
        } else {
            int32_t child = buildAccelMotionRec(nodes, boxes + lo[c], centroids + lo[c], indxs + lo[c], cn, leafMax, baseidx + lo[c], depth + 1);
            nodes[nodei].children[c] = child;
        }
    }
    return nodei;
}

// qbvh/motionbuild.go:108-122
static void BuildAccelMotion(std::vector<MNode>& nodes, Box* boxes, Vec3* centroids, int32_t* indxs, int n, int leafMax) {
    if (leafMax > 16) leafMax = 16;
    if (leafMax < 1) leafMax = 1;
    buildAccelMotionRec(nodes, boxes, centroids, indxs, n, leafMax, 0, 0);
}

// ---------------------------------------------------------------------- rays

enum { RayTypeCamera = 1, RayTypeShadow = 2, RayTypeReflected = 4 };

struct StackEntry { float T; int32_t Node; };

struct Task {
    float T[4];
    int32_t Hits[4];
    float Boxes[24];
    int32_t StackTop;
    StackEntry Stack[256];     // the reference holds 90 entries (core/ray.go:158); deeper use is reported
    int32_t maxDepth;
};

struct Ray {
    Vec3 P, D, Dinv;
    float Tclosest;
    float S[3];
    int32_t Kx, Ky, Kz;
    float Time;
    uint32_t Type;
    int64_t NodesT, LeafsT, TrisT;
    Task* task;
};

// core/ray.go:103-148
static void RaySetup(Ray& r) {
    r.Kz = 0;
    if (Abs(r.D[1]) > Abs(r.D[2])) {
        if (Abs(r.D[1]) > Abs(r.D[0])) r.Kz = 1;
    } else {
        if (Abs(r.D[2]) > Abs(r.D[0])) r.Kz = 2;
    }
    r.Kx = r.Kz + 1; if (r.Kx == 3) r.Kx = 0;
    r.Ky = r.Kx + 1; if (r.Ky == 3) r.Ky = 0;
    if (r.D[r.Kz] < 0.0f) { int32_t t = r.Kx; r.Kx = r.Ky; r.Ky = t; }
    double z = (double)r.D[r.Kz];
    r.S[2] = (float)(1.0 / z);
    r.S[0] = (float)((double)r.D[r.Kx] / z);
    r.S[1] = (float)((double)r.D[r.Ky] / z);
    for (int k = 0; k < 3; k++) r.Dinv[k] = (float)(1.0 / (double)r.D[k]);
}

// qbvh/intersect_amd64.s:13-100 — lane semantics of the SSE sequence.
static void intersectBoxes(const Ray& ray, const float* boxes, int32_t* hits, float* t) {
    for (int c = 0; c < 4; c++) {
        float o, di, t1, t2, tmin, tmax, m, M;
        o = ray.P[0]; di = ray.Dinv[0];
        t1 = (boxes[c + 0] - o) * di;          // boxmin.x
        t2 = (boxes[c + 12] - o) * di;         // boxmax.x
        tmin = (t2 < t1) ? t2 : t1;            // MINPS X2,X6
        tmax = (t2 > t1) ? t2 : t1;            // MAXPS X2,X7
        for (int k = 1; k < 3; k++) {
            o = ray.P[k]; di = ray.Dinv[k];
            t1 = (boxes[c + 4 * k] - o) * di;
            t2 = (boxes[c + 12 + 4 * k] - o) * di;
            m = (t2 < t1) ? t2 : t1;           // MINPS X2,X1
            M = (t2 > t1) ? t2 : t1;           // MAXPS X2,X0
            tmin = (tmin > m) ? tmin : m;      // MAXPS X1,X6
            tmax = (tmax < M) ? tmax : M;      // MINPS X0,X7
        }
        float tNear = (0.0f > tmin) ? 0.0f : tmin;   // XORPS; MAXPS X6,X0
        t[c] = tNear;
        hits[c] = (tNear <= tmax) ? -1 : 0;          // CMPPS $2 (ordered LE)
    }
}

// float32 trigonometry through the float64 C library (the reference calls Go's float64 math)
static inline float fSin(float x) { return (float)std::sin((double)x); }
static inline float fCos(float x) { return (float)std::cos((double)x); }
static inline float fAcos(float x) { return (float)std::acos((double)x); }
static inline Vec3 V3(float x, float y, float z) { Vec3 v; v[0] = x; v[1] = y; v[2] = z; return v; }

// --------------------------------------------------------------- matrices

typedef float M4[16];   // column-major, Elt(i,j) = m[j*4+i] (math/matrix4.go:50-57)

static void M4Identity(M4 m) { std::memset(m, 0, sizeof(M4)); m[0] = m[5] = m[10] = m[15] = 1.0f; }
static void M4Mul(const M4 a, const M4 b, M4 c) {
    float t[16] = {0};
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) for (int k = 0; k < 4; k++) t[j * 4 + i] += a[k * 4 + i] * b[j * 4 + k];
    std::memcpy(c, t, sizeof(t));
}
static void M4Transpose(const M4 a, M4 c) { float t[16]; for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) t[i * 4 + j] = a[j * 4 + i]; std::memcpy(c, t, sizeof(t)); }
static void M4Cof(const float* m, float* inv) {
    inv[0] = m[5]*m[10]*m[15] - m[5]*m[11]*m[14] - m[9]*m[6]*m[15] + m[9]*m[7]*m[14] + m[13]*m[6]*m[11] - m[13]*m[7]*m[10];
    inv[4] = -m[4]*m[10]*m[15] + m[4]*m[11]*m[14] + m[8]*m[6]*m[15] - m[8]*m[7]*m[14] - m[12]*m[6]*m[11] + m[12]*m[7]*m[10];
    inv[8] = m[4]*m[9]*m[15] - m[4]*m[11]*m[13] - m[8]*m[5]*m[15] + m[8]*m[7]*m[13] + m[12]*m[5]*m[11] - m[12]*m[7]*m[9];
    inv[12] = -m[4]*m[9]*m[14] + m[4]*m[10]*m[13] + m[8]*m[5]*m[14] - m[8]*m[6]*m[13] - m[12]*m[5]*m[10] + m[12]*m[6]*m[9];
    inv[1] = -m[1]*m[10]*m[15] + m[1]*m[11]*m[14] + m[9]*m[2]*m[15] - m[9]*m[3]*m[14] - m[13]*m[2]*m[11] + m[13]*m[3]*m[10];
    inv[5] = m[0]*m[10]*m[15] - m[0]*m[11]*m[14] - m[8]*m[2]*m[15] + m[8]*m[3]*m[14] + m[12]*m[2]*m[11] - m[12]*m[3]*m[10];
    inv[9] = -m[0]*m[9]*m[15] + m[0]*m[11]*m[13] + m[8]*m[1]*m[15] - m[8]*m[3]*m[13] - m[12]*m[1]*m[11] + m[12]*m[3]*m[9];
    inv[13] = m[0]*m[9]*m[14] - m[0]*m[10]*m[13] - m[8]*m[1]*m[14] + m[8]*m[2]*m[13] + m[12]*m[1]*m[10] - m[12]*m[2]*m[9];
    inv[2] = m[1]*m[6]*m[15] - m[1]*m[7]*m[14] - m[5]*m[2]*m[15] + m[5]*m[3]*m[14] + m[13]*m[2]*m[7] - m[13]*m[3]*m[6];
    inv[6] = -m[0]*m[6]*m[15] + m[0]*m[7]*m[14] + m[4]*m[2]*m[15] - m[4]*m[3]*m[14] - m[12]*m[2]*m[7] + m[12]*m[3]*m[6];
    inv[10] = m[0]*m[5]*m[15] - m[0]*m[7]*m[13] - m[4]*m[1]*m[15] + m[4]*m[3]*m[13] + m[12]*m[1]*m[7] - m[12]*m[3]*m[5];
    inv[14] = -m[0]*m[5]*m[14] + m[0]*m[6]*m[13] + m[4]*m[1]*m[14] - m[4]*m[2]*m[13] - m[12]*m[1]*m[6] + m[12]*m[2]*m[5];
    inv[3] = -m[1]*m[6]*m[11] + m[1]*m[7]*m[10] + m[5]*m[2]*m[11] - m[5]*m[3]*m[10] - m[9]*m[2]*m[7] + m[9]*m[3]*m[6];
    inv[7] = m[0]*m[6]*m[11] - m[0]*m[7]*m[10] - m[4]*m[2]*m[11] + m[4]*m[3]*m[10] + m[8]*m[2]*m[7] - m[8]*m[3]*m[6];
    inv[11] = -m[0]*m[5]*m[11] + m[0]*m[7]*m[9] + m[4]*m[1]*m[11] - m[4]*m[3]*m[9] - m[8]*m[1]*m[7] + m[8]*m[3]*m[5];
    inv[15] = m[0]*m[5]*m[10] - m[0]*m[6]*m[9] - m[4]*m[1]*m[10] + m[4]*m[2]*m[9] + m[8]*m[1]*m[6] - m[8]*m[2]*m[5];
}
// math/matrix4.go:116-149
static bool M4Inverse(const M4 m, M4 c) {
    float inv[16];
    M4Cof(m, inv);
    float det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0f) return false;
    det = 1.0f / det;
    for (int i = 0; i < 16; i++) c[i] = inv[i] * det;
    return true;
}
static float M4Det(const M4 m) { float inv[16]; M4Cof(m, inv); return m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12]; }
// math/matrix4.go:343-363
static bool M4PolarFactor(const M4 m, M4 Q) {
    std::memcpy(Q, m, sizeof(M4));
    for (int it = 0; it < 10; it++) {
        M4 Qinv, QinvT, Qnew;
        if (!M4Inverse(Q, Qinv)) return false;
        M4Transpose(Qinv, QinvT);
        for (int i = 0; i < 16; i++) Qnew[i] = 0.5f * (Q[i] + QinvT[i]);
        bool eq = true;
        for (int i = 0; i < 16; i++) if (Qnew[i] - Q[i] > 0.000001f) { eq = false; break; }
        if (eq) { std::memcpy(Q, Qnew, sizeof(M4)); return true; }
        std::memcpy(Q, Qnew, sizeof(M4));
    }
    return false;
}
struct Quat { float X, Y, Z, W; };
// math/quat.go:133-163
static Quat M4ToQuat(const M4 m) {
    auto E = [&](int i, int j) { return m[j * 4 + i]; };
    Quat q;
    float tr = E(0, 0) + E(1, 1) + E(2, 2);
    if (tr > 0.0f) {
        float S = std::sqrt(tr + 1.0f) * 2;
        q.W = 0.25f * S; q.X = (E(2, 1) - E(1, 2)) / S; q.Y = (E(0, 2) - E(2, 0)) / S; q.Z = (E(1, 0) - E(0, 1)) / S;
    } else if ((E(0, 0) > E(1, 1)) && (E(0, 0) > E(2, 2))) {
        float S = std::sqrt(1.0f + E(0, 0) - E(1, 1) - E(2, 2)) * 2;
        q.W = (E(2, 1) - E(1, 2)) / S; q.X = 0.25f * S; q.Y = (E(0, 1) + E(1, 0)) / S; q.Z = (E(0, 2) + E(2, 0)) / S;
    } else if (E(1, 1) > E(2, 2)) {
        float S = std::sqrt(1.0f + E(1, 1) - E(0, 0) - E(2, 2)) * 2;
        q.W = (E(0, 2) - E(2, 0)) / S; q.X = (E(0, 1) + E(1, 0)) / S; q.Y = 0.25f * S; q.Z = (E(1, 2) + E(2, 1)) / S;
    } else {
        float S = std::sqrt(1.0f + E(2, 2) - E(0, 0) - E(1, 1)) * 2;
        q.W = (E(1, 0) - E(0, 1)) / S; q.X = (E(0, 2) + E(2, 0)) / S; q.Y = (E(1, 2) + E(2, 1)) / S; q.Z = 0.25f * S;
    }
    return q;
}
// math/quat.go:80-104
static void QuatToM4(const Quat& q, M4 m) {
    std::memset(m, 0, sizeof(M4));
    float n = 1.0f / std::sqrt(q.X * q.X + q.Y * q.Y + q.Z * q.Z + q.W * q.W);
    float X = q.X * n, Y = q.Y * n, Z = q.Z * n, W = q.W * n;
    auto S = [&](int i, int j, float v) { m[j * 4 + i] = v; };
    S(0, 0, 1 - 2 * Y * Y - 2 * Z * Z); S(0, 1, 2 * X * Y - 2 * W * Z); S(0, 2, 2 * X * Z + 2 * W * Y);
    S(1, 0, 2 * X * Y + 2 * W * Z); S(1, 1, 1 - 2 * X * X - 2 * Z * Z); S(1, 2, 2 * Y * Z - 2 * W * X);
    S(2, 0, 2 * X * Z - 2 * W * Y); S(2, 1, 2 * Y * Z + 2 * W * X); S(2, 2, 1 - 2 * X * X - 2 * Y * Y);
    S(3, 3, 1.0f);
}
// math/quat.go:30-67
static Quat QuatSlerp(const Quat& a, const Quat& b, float t) {
    float c = a.W * b.W + a.X * b.X + a.Y * b.Y + a.Z * b.Z;
    if (Abs(c) >= 1.0f) return a;
    float half = fAcos(c);
    float s = std::sqrt(1.0f - c * c);
    if (Abs(s) < 0.001f) return Quat{a.X * 0.5f + b.X * 0.5f, a.Y * 0.5f + b.Y * 0.5f, a.Z * 0.5f + b.Z * 0.5f, a.W * 0.5f + b.W * 0.5f};
    float ra = fSin((1 - t) * half) / s, rb = fSin(t * half) / s;
    return Quat{a.X * ra + b.X * rb, a.Y * ra + b.Y * rb, a.Z * ra + b.Z * rb, a.W * ra + b.W * rb};
}
struct Decomp { Vec3 T; Quat R; M4 S; };
// math/animdecomp.go:20-63 (octanormal.go:16 sign)
static Decomp DecompM4(const M4 min) {
    Decomp d;
    M4 m; std::memcpy(m, min, sizeof(M4));
    float sgn = M4Det(m) >= 0.0f ? 1.0f : -1.0f;
    for (int i = 0; i < 3; i++) d.T[i] = m[3 * 4 + i];
    m[12] = 0; m[13] = 0; m[14] = 0;
    // Matrix4Mul(Matrix4Scale(-1, Identity), m) as the reference multiplies it: the zeros of -I are -0,
    // and the sums start from +0, which fixes the sign of the zero entries.
    M4 negI; M4Identity(negI);
    for (int i = 0; i < 16; i++) negI[i] = -1.0f * negI[i];
    if (sgn < 0.0f) M4Mul(negI, m, m);
    M4 Q;
    if (!M4PolarFactor(m, Q)) { d.R = Quat{0, 0, 0, 1}; M4Identity(d.S); return d; }
    M4 QT; M4Transpose(Q, QT);
    M4 S; M4Mul(QT, m, S);
    d.R = M4ToQuat(Q);
    if (sgn < 0.0f) { M4Mul(negI, S, d.S); d.S[15] = 1; }
    else std::memcpy(d.S, S, sizeof(M4));
    return d;
}
// TransformDecompLerp then TransformDecompToMatrix4 (animdecomp.go:66-80); a static camera has a == b
static void DecompToM4(const Decomp& a, const Decomp& b, float t, M4 out) {
    Decomp l;
    l.T = Lerp(a.T, b.T, t);
    l.R = QuatSlerp(a.R, b.R, t);
    for (int i = 0; i < 16; i++) l.S[i] = (1.0f - t) * a.S[i] + t * b.S[i];   // Matrix4Lerp
    M4 Tm; M4Identity(Tm); Tm[12] = l.T[0]; Tm[13] = l.T[1]; Tm[14] = l.T[2];
    M4 R; QuatToM4(l.R, R);
    M4 RS; M4Mul(R, l.S, RS);
    M4Mul(Tm, RS, out);
}
static Vec3 MulVec(const M4 a, const Vec3& b) {
    return V3(a[0] * b[0] + a[4] * b[1] + a[8] * b[2], a[1] * b[0] + a[5] * b[1] + a[9] * b[2], a[2] * b[0] + a[6] * b[1] + a[10] * b[2]);
}
static Vec3 MulPoint(const M4 a, const Vec3& b) {
    return V3(a[0] * b[0] + a[4] * b[1] + a[8] * b[2] + a[12], a[1] * b[0] + a[5] * b[1] + a[9] * b[2] + a[13],
              a[2] * b[0] + a[6] * b[1] + a[10] * b[2] + a[14]);
}

// ------------------------------------------------------------------- meshes

struct HitRecord {
    int32_t elem;          // ElemID (post-sort triangle index, trace.go:513)
    float U, V, W;         // barycentric weights of v0, v1, v2
    Vec3 Po;               // motion: point accumulated during the test loop (trace.go:625-629)
    Vec3 Ng;               // motion path sets these inside the loop (trace.go:650-665)
    Vec3 Poffset;
    Vec3 DdPdu, DdPdv;
    float Bu, Bv;
    Vec3 N;                // shading normal (interpolated normals, else Ng) (trace.go:324-338)
    int32_t shaderIdx;     // mesh.shaderidx[idx] (trace.go:300-306)
    const struct Mesh* leaf;   // the mesh or sphere whose element was hit (inside an instance or proc)
    bool hasXform;         // sg.Transform / sg.InvTransform set by the hit's geoms
    bool meshXform;        // the hit mesh's own Transform applied to N in PolyMesh.Trace (trace.go:65-70)
    M4 xform, xformInv;    // sg.Transform, sg.InvTransform
    M4 meshInv;            // the inverse of the hit mesh's own transform
};

enum { GEOM_MESH = 0, GEOM_INSTANCE = 1, GEOM_PROC = 2 };

struct Mesh {
    int nkeys = 1;
    int elemsPerKey = 0;
    std::vector<Vec3> verts;          // nkeys * elemsPerKey
    std::vector<uint32_t> idxp;       // post-sort for static meshes; original order for motion meshes
    std::vector<int32_t> accelIdx;    // accel.idx: leaf slot -> original face
    std::vector<Vec3> normals;        // Normals.Elems (per vertex), nullable
    std::vector<uint32_t> normalidx;  // reordered with idxp
    std::vector<Vec3> mnormals;       // moving meshes: Normals.Elems per face corner, normalKeys keys
    int normalKeys = 0;
    std::vector<uint8_t> shaderidx;   // per triangle, reordered with idxp
    std::vector<int32_t> shaders;     // shader table: material id per shader index
    int facecount = 0;
    float RayBias = 0.0f;
    bool motionBug = true;            // reproduce trace.go:532-537 (idxp indexed by leaf slot)
    std::vector<Node> qbvh;
    std::vector<MNode> mnodes;
    std::vector<std::vector<float>> mboxes;   // per key: 24 floats per node
    Box bounds;
    std::vector<Box> motionBounds;
    bool sphere = false;              // analytic sphere geom (builtin/geom/sphere), no triangles
    Vec3 sphereC; float sphereR = 0;
    // Geoms that wrap others under a transform: GeomInstance (builtin/geom/instance/instance.go) wraps
    // one geom, Proc (builtin/geom/proc/proc.go) a list; a mesh may carry its own Transform
    // (builtin/geom/polymesh/trace.go:20-74).  xform: the raw matrices (column-major); xformSRT their
    // TransformDecompMatrix4 (animdecomp.go:20-63).  userBounds: Bounds of an instance or proc, the
    // box of BMin[0], BMax[0] whatever the time (instance.go:63-67, proc.go:76-80).  hidden: not in
    // the scene's top level (a proc's or instance's geoms that the scene does not hold itself).
    int kind = 0;                     // GEOM_MESH, GEOM_INSTANCE, GEOM_PROC
    std::vector<std::vector<float>> xform;
    std::vector<Decomp> xformSRT;
    std::vector<Mesh*> children;
    Box userBounds;
    bool hidden = false;

    // builtin/geom/polymesh/buildqbvh.go:14-144
    int leafMax = 16;                 // BuildAccel leafMax (buildqbvh.go:32)
    void initAccel() {
        if (sphere) return;
        int n = facecount;
        std::vector<Box> boxes(n);
        std::vector<Vec3> centroids(n);
        std::vector<int32_t> idxs(n);
        if (nkeys > 1) {
            float k = 0.5f * (float)(nkeys - 1);
            float time = k - Floor(k);
            int key = (int)Floor(k), key2 = (int)Ceil(k);
            for (int i = 0; i < n; i++) {
                Vec3 V0 = Lerp(verts[idxp[i * 3 + 0] + elemsPerKey * key], verts[idxp[i * 3 + 0] + elemsPerKey * key2], time);
                Vec3 V1 = Lerp(verts[idxp[i * 3 + 1] + elemsPerKey * key], verts[idxp[i * 3 + 1] + elemsPerKey * key2], time);
                Vec3 V2 = Lerp(verts[idxp[i * 3 + 2] + elemsPerKey * key], verts[idxp[i * 3 + 2] + elemsPerKey * key2], time);
                boxes[i].Reset(); boxes[i].GrowVec3(V0); boxes[i].GrowVec3(V1); boxes[i].GrowVec3(V2);
                for (int c = 0; c < 3; c++) centroids[i][c] = (V0[c] + V1[c] + V2[c]) / 3.0f;
                idxs[i] = i;
            }
            BuildAccelMotion(mnodes, boxes.data(), centroids.data(), idxs.data(), n, 16);
            accelIdx = idxs;
            initMotionBoxes();
            return;
        }
        for (int i = 0; i < n; i++) {
            const Vec3& V0 = verts[idxp[i * 3 + 0]];
            const Vec3& V1 = verts[idxp[i * 3 + 1]];
            const Vec3& V2 = verts[idxp[i * 3 + 2]];
            boxes[i].Reset(); boxes[i].GrowVec3(V0); boxes[i].GrowVec3(V1); boxes[i].GrowVec3(V2);
            for (int c = 0; c < 3; c++) centroids[i][c] = (V0[c] + V1[c] + V2[c]) / 3.0f;
            idxs[i] = i;
        }
        bounds = BuildAccel(qbvh, boxes.data(), centroids.data(), idxs.data(), n, leafMax);
        accelIdx = idxs;
        std::vector<uint32_t> nidx(idxp.size());
        for (int i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) nidx[i * 3 + c] = idxp[idxs[i] * 3 + c];
        idxp.swap(nidx);
        if (!shaderidx.empty()) {   // buildqbvh.go:98-104
            std::vector<uint8_t> ns(shaderidx.size());
            for (int i = 0; i < n; i++) ns[i] = shaderidx[idxs[i]];
            shaderidx.swap(ns);
        }
        if (!normalidx.empty()) {   // buildqbvh.go:118-127
            std::vector<uint32_t> nn(normalidx.size());
            for (int i = 0; i < n; i++)
                for (int c = 0; c < 3; c++) nn[i * 3 + c] = normalidx[idxs[i] * 3 + c];
            normalidx.swap(nn);
        }
    }

    // buildqbvh.go:148-212
    void initMotionBoxes() {
        mboxes.assign(nkeys, std::vector<float>(mnodes.size() * 24, 0.0f));
        Box full; full.Reset();
        for (int k = 0; k < nkeys; k++) {
            Box b = initMotionBoxesRec(k, 0);
            motionBounds.push_back(b);
            full.GrowBox(b);
        }
        bounds = full;
    }
    Box initMotionBoxesRec(int key, int32_t node) {
        Box nodebox; nodebox.Reset();
        for (int c = 0; c < 4; c++) {
            int32_t ch = mnodes[node].children[c];
            if (ch < 0) {
                if (ch == -1) continue;
                Box box; box.Reset();
                int lb = LeafBase(ch), lc = LeafCount(ch);
                for (int i = lb; i < lb + lc; i++) {
                    int f = accelIdx[i];
                    box.GrowVec3(verts[idxp[f * 3 + 0] + elemsPerKey * key]);
                    box.GrowVec3(verts[idxp[f * 3 + 1] + elemsPerKey * key]);
                    box.GrowVec3(verts[idxp[f * 3 + 2] + elemsPerKey * key]);
                }
                SetBounds(&mboxes[key][node * 24], c, box);
                nodebox.GrowBox(box);
            } else {
                Box box = initMotionBoxesRec(key, ch);
                SetBounds(&mboxes[key][node * 24], c, box);
                nodebox.GrowBox(box);
            }
        }
        return nodebox;
    }

    // polymesh/bounds.go:159-186 (a mesh's Transform is ignored: bounds.go:26 reads `if false && ...`);
    // instance and proc: their given box
    Box Bounds(float time) const {
        if (kind != GEOM_MESH) return userBounds;
        if (nkeys == 1) return bounds;
        float k = time * (float)(motionBounds.size() - 1);
        float t = k - Floor(k);
        int key = (int)Floor(k), key2 = (int)Ceil(k);
        return BoxLerp(motionBounds[key], motionBounds[key2], t);
    }
    // polymesh.go:102-112; instance.go:151-156 (the wrapped geom's); proc.go:188-193 (its transform keys)
    int MotionKeys() const {
        if (kind == GEOM_INSTANCE) return children[0]->MotionKeys();
        if (kind == GEOM_PROC) return (int)xformSRT.size();
        return nkeys == 1 ? 1 : (int)mboxes.size();
    }
    // The transform at a ray's time: TransformSRTArray.TimeKey then TransformDecompToMatrix4
    // (instance.go:17-34, proc.go:20-37; a single key recomposed too); a mesh's own Transform with one
    // key is its raw matrix (trace.go:33-47).  Times outside [0, 1] are held to the end keys.
    void TransformAt(float time, M4 out) const {
        const int nk = (int)xformSRT.size();
        if (nk == 1) {
            if (kind == GEOM_MESH) std::memcpy(out, xform[0].data(), sizeof(M4));
            else DecompToM4(xformSRT[0], xformSRT[0], 0.0f, out);
            return;
        }
        float k = time * (float)(nk - 1);
        float t = k - Floor(k);
        int key = std::min(std::max((int)Floor(k), 0), nk - 1), key2 = std::min(std::max((int)Ceil(k), 0), nk - 1);
        DecompToM4(xformSRT[key], xformSRT[key2], t, out);
    }

    // builtin/geom/polymesh/trace.go:108-194 (idxp branch) — returns true on an improving hit.
    bool TraceElems(Ray& ray, HitRecord& hr, int base, int count) const {
        int32_t idx = -1;
        float U = 0, V = 0, W = 0;
        const int Kx = ray.Kx, Ky = ray.Ky, Kz = ray.Kz;
        for (int i = base; i < base + count; i++) {
            ray.TrisT++;
            const Vec3& v0 = verts[idxp[i * 3 + 0]];
            const Vec3& v1 = verts[idxp[i * 3 + 1]];
            const Vec3& v2 = verts[idxp[i * 3 + 2]];
            float AKz = v0[Kz] - ray.P[Kz];
            float BKz = v1[Kz] - ray.P[Kz];
            float CKz = v2[Kz] - ray.P[Kz];
            float Cx = (v2[Kx] - ray.P[Kx]) - ray.S[0] * CKz;
            float By = (v1[Ky] - ray.P[Ky]) - ray.S[1] * BKz;
            float Cy = (v2[Ky] - ray.P[Ky]) - ray.S[1] * CKz;
            float Bx = (v1[Kx] - ray.P[Kx]) - ray.S[0] * BKz;
            float Ax = (v0[Kx] - ray.P[Kx]) - ray.S[0] * AKz;
            float Ay = (v0[Ky] - ray.P[Ky]) - ray.S[1] * AKz;
            float fU = Cx * By - Cy * Bx;
            float fV = Ax * Cy - Ay * Cx;
            float fW = Bx * Ay - By * Ax;
            if (fU == 0.0f || fV == 0.0f || fW == 0.0f) {
                fU = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
                fV = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
                fW = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
            }
            if ((fU < 0.0f || fV < 0.0f || fW < 0.0f) && (fU > 0.0f || fV > 0.0f || fW > 0.0f)) continue;
            float det = fU + fV + fW;
            if (det == 0.0f) continue;
            float T = ray.S[2] * (fU * AKz + fV * BKz + fW * CKz);
            uint32_t detSign = SignMask(det);
            if (Xorf(T, detSign) < (EpsilonFloat32 + RayBias) * Xorf(det, detSign) ||
                Xorf(T, detSign) > ray.Tclosest * Xorf(det, detSign)) continue;
            float rcpDet = 1.0f / det;
            U = fU * rcpDet; V = fV * rcpDet; W = fW * rcpDet;
            ray.Tclosest = T * rcpDet;
            idx = i;
        }
        if (idx == -1) return false;
        hr.elem = idx; hr.U = U; hr.V = V; hr.W = W; hr.Bu = U; hr.Bv = V;
        // hit attributes, trace.go:292-342
        const Vec3& v0 = verts[idxp[idx * 3 + 0]];
        const Vec3& v1 = verts[idxp[idx * 3 + 1]];
        const Vec3& v2 = verts[idxp[idx * 3 + 2]];
        float xs = Abs(U * v0[0]) + Abs(V * v1[0]) + Abs(W * v2[0]);
        float ys = Abs(U * v0[1]) + Abs(V * v1[1]) + Abs(W * v2[1]);
        float zs = Abs(U * v0[2]) + Abs(V * v1[2]) + Abs(W * v2[2]);
        float e00 = v1[0] - v0[0], e01 = v1[1] - v0[1], e02 = v1[2] - v0[2];
        float e10 = v2[0] - v0[0], e11 = v2[1] - v0[1], e12 = v2[2] - v0[2];
        Vec3 Ng; Ng[0] = e01 * e12 - e02 * e11; Ng[1] = e02 * e10 - e00 * e12; Ng[2] = e00 * e11 - e01 * e10;
        Ng = Normalize(Ng);
        hr.Ng = Ng;
        float d = Gamma(7) * xs * Abs(Ng[0]) + Gamma(7) * ys * Abs(Ng[1]) + Gamma(7) * zs * Abs(Ng[2]);
        for (int k = 0; k < 3; k++) hr.Poffset[k] = Ng[k] * d;
        for (int k = 0; k < 3; k++) hr.Po[k] = U * v0[k] + V * v1[k] + W * v2[k];
        hr.shaderIdx = shaderidx.empty() ? 0 : shaderidx[idx];
        if (!normals.empty()) {   // trace.go:326-334
            Vec3 Nn;
            for (int k = 0; k < 3; k++)
                Nn[k] = U * normals[normalidx[idx * 3 + 0]][k] + V * normals[normalidx[idx * 3 + 1]][k] +
                        W * normals[normalidx[idx * 3 + 2]][k];
            hr.N = Normalize(Nn);
        } else {
            hr.N = Ng;
        }
        // trace.go:504-511
        Vec3 axisu;
        {
            float dx = 1 * Ng[0] + 0 * Ng[1] + 0 * Ng[2];
            for (int k = 0; k < 3; k++) axisu[k] = (k == 0 ? 1.0f : 0.0f) - Ng[k] * dx;
            float l2 = axisu[0] * axisu[0] + axisu[1] * axisu[1] + axisu[2] * axisu[2];
            float dn = axisu[0] * Ng[0] + axisu[1] * Ng[1] + axisu[2] * Ng[2];
            if (l2 < 0.1f || Abs(dn) > 0.3f) {
                float dz = 0 * Ng[0] + 0 * Ng[1] + 1 * Ng[2];
                for (int k = 0; k < 3; k++) axisu[k] = (k == 2 ? 1.0f : 0.0f) - Ng[k] * dz;
            }
        }
        hr.DdPdu = Normalize(axisu);
        const Vec3& a = Ng; const Vec3& b = hr.DdPdu;
        hr.DdPdv[0] = a[1] * b[2] - a[2] * b[1]; hr.DdPdv[1] = a[2] * b[0] - a[0] * b[2]; hr.DdPdv[2] = a[0] * b[1] - a[1] * b[0];
        return true;
    }

    // trace.go:520-669 — the motion variant.
    bool TraceMotionElems(float time, int key, int key2, Ray& ray, HitRecord& hr, int base, int count) const {
        int32_t idx = -1;
        float U = 0, V = 0, W = 0;
        const int Kx = ray.Kx, Ky = ray.Ky, Kz = ray.Kz;
        for (int i = base; i < base + count; i++) {
            ray.TrisT++;
            int32_t faceidx = i;
            int f = motionBug ? i : accelIdx[i];
            int i0 = (int)idxp[f * 3 + 0], i1 = (int)idxp[f * 3 + 1], i2 = (int)idxp[f * 3 + 2];
            Vec3 V0 = Lerp(verts[i0 + elemsPerKey * key], verts[i0 + elemsPerKey * key2], time);
            Vec3 V1 = Lerp(verts[i1 + elemsPerKey * key], verts[i1 + elemsPerKey * key2], time);
            Vec3 V2 = Lerp(verts[i2 + elemsPerKey * key], verts[i2 + elemsPerKey * key2], time);
            float AKz = V0[Kz] - ray.P[Kz];
            float BKz = V1[Kz] - ray.P[Kz];
            float CKz = V2[Kz] - ray.P[Kz];
            float Cx = (V2[Kx] - ray.P[Kx]) - ray.S[0] * CKz;
            float By = (V1[Ky] - ray.P[Ky]) - ray.S[1] * BKz;
            float Cy = (V2[Ky] - ray.P[Ky]) - ray.S[1] * CKz;
            float Bx = (V1[Kx] - ray.P[Kx]) - ray.S[0] * BKz;
            float Ax = (V0[Kx] - ray.P[Kx]) - ray.S[0] * AKz;
            float Ay = (V0[Ky] - ray.P[Ky]) - ray.S[1] * AKz;
            float fU = Cx * By - Cy * Bx;
            float fV = Ax * Cy - Ay * Cx;
            float fW = Bx * Ay - By * Ax;
            if (fU == 0.0f || fV == 0.0f || fW == 0.0f) {
                fU = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
                fV = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
                fW = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
            }
            if ((fU < 0.0f || fV < 0.0f || fW < 0.0f) && (fU > 0.0f || fV > 0.0f || fW > 0.0f)) continue;
            float det = fU + fV + fW;
            if (det == 0.0f) continue;
            float T = ray.S[2] * (fU * AKz + fV * BKz + fW * CKz);
            uint32_t detSign = SignMask(det);
            if (Xorf(T, detSign) <= RayBias * Xorf(det, detSign) ||
                Xorf(T, detSign) > ray.Tclosest * Xorf(det, detSign)) continue;
            float rcpDet = 1.0f / det;
            U = fU * rcpDet; V = fV * rcpDet; W = fW * rcpDet;
            ray.Tclosest = T * rcpDet;
            idx = faceidx;
            for (int k = 0; k < 3; k++) hr.Po[k] = U * V0[k] + V * V1[k] + W * V2[k];
            float xAbsSum = Abs(U * V0[0]) + Abs(V * V1[0]) + Abs(W * V2[0]);
            float yAbsSum = Abs(U * V0[1]) + Abs(V * V1[1]) + Abs(W * V2[1]);
            float zAbsSum = Abs(U * V0[2]) + Abs(V * V1[2]) + Abs(W * V2[2]);
            float e00 = V1[0] - V0[0], e01 = V1[1] - V0[1], e02 = V1[2] - V0[2];
            float e10 = V2[0] - V0[0], e11 = V2[1] - V0[1], e12 = V2[2] - V0[2];
            Vec3 Ng; Ng[0] = e01 * e12 - e02 * e11; Ng[1] = e02 * e10 - e00 * e12; Ng[2] = e00 * e11 - e01 * e10;
            hr.Ng = Normalize(Ng);
            hr.DdPdu[0] = e00; hr.DdPdu[1] = e01; hr.DdPdu[2] = e02;   // trace.go:653-659
            hr.DdPdv[0] = e10; hr.DdPdv[1] = e11; hr.DdPdv[2] = e12;
            float d = Gamma(7) * xAbsSum * Abs(hr.Ng[0]) + Gamma(7) * yAbsSum * Abs(hr.Ng[1]) +
                      Gamma(7) * zAbsSum * Abs(hr.Ng[2]);
            for (int k = 0; k < 3; k++) hr.Poffset[k] = d * hr.Ng[k];
            hr.Bu = U; hr.Bv = V;
        }
        if (idx == -1) return false;
        hr.elem = idx; hr.U = U; hr.V = V; hr.W = W;
        // trace.go:673-717: shader index and normals of the hit face (the leaf slot itself with the
        // reference's indexing; its accel.idx face otherwise)
        const int sf = motionBug ? idx : accelIdx[idx];
        hr.shaderIdx = shaderidx.empty() ? 0 : shaderidx[sf];
        if (!mnormals.empty() && normalKeys == nkeys) {
            const size_t perKey = (size_t)facecount * 3;   // Normals.ElemsPerKey
            Vec3 N0 = Lerp(mnormals[sf * 3 + 0 + perKey * key], mnormals[sf * 3 + 0 + perKey * key2], time);
            Vec3 N1 = Lerp(mnormals[sf * 3 + 1 + perKey * key], mnormals[sf * 3 + 1 + perKey * key2], time);
            Vec3 N2 = Lerp(mnormals[sf * 3 + 2 + perKey * key], mnormals[sf * 3 + 2 + perKey * key2], time);
            Vec3 Nn;
            for (int k = 0; k < 3; k++) Nn[k] = U * N0[k] + V * N1[k] + W * N2[k];
            hr.N = Normalize(Nn);
        } else {
            hr.N = hr.Ng;
        }
        return true;
    }

    // math/vec3_amd64.s:11-42 — reciprocal-square-root estimate + one Newton-Raphson step.
    // The estimate is RSQRTSS (12-bit, its bits differ between CPU vendors) in the default mode, or
    // the IEEE value 1/sqrt(x) when normalize_ieee() is set: the same Newton step on an estimate
    // every machine computes identically.
    static Vec3 Normalize(const Vec3& a) {
        float x2 = a[0] * a[0], y2 = a[1] * a[1], z2 = a[2] * a[2];
        float l2 = (y2 + x2) + z2;
        float g = normalize_ieee() ? 1.0f / std::sqrt(l2) : _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(l2)));
        float h = l2 * 0.5f;
        float gg = g * g;
        gg = gg * h;
        float r = (1.5f - gg) * g;
        Vec3 o; o[0] = a[0] * r; o[1] = a[1] * r; o[2] = a[2] * r;
        return o;
    }
};

// ------------------------------------------------------------------ traversal

// qbvh/intersect.go:91-246 with the Primitive call expressed as a callback.
template <class Leaf>
static bool TraceQBVH(const Node* qbvh, Ray& ray, Leaf&& traceElems) {
    Task& tk = *ray.task;
    int32_t stackTop = tk.StackTop;
    tk.Stack[stackTop].Node = 0;
    tk.Stack[stackTop].T = ray.Tclosest;
    bool hit = false;
    stackTop++;
    while (stackTop > tk.StackTop) {
        stackTop--;
        int32_t node = tk.Stack[stackTop].Node;
        if (ray.Tclosest < tk.Stack[stackTop].T || node == -1) continue;
        if (node >= 0) {
            ray.NodesT++;
            const Node& nd = qbvh[node];
            intersectBoxes(ray, nd.boxes, tk.Hits, tk.T);
            int order[4];
            if (ray.D[nd.axis0] < 0) {
                if (ray.D[nd.axis1] < 0) { order[0] = 0; order[1] = 1; } else { order[0] = 1; order[1] = 0; }
                if (ray.D[nd.axis2] < 0) { order[2] = 2; order[3] = 3; } else { order[2] = 3; order[3] = 2; }
            } else {
                if (ray.D[nd.axis2] < 0) { order[0] = 2; order[1] = 3; } else { order[0] = 3; order[1] = 2; }
                if (ray.D[nd.axis1] < 0) { order[2] = 0; order[3] = 1; } else { order[2] = 1; order[3] = 0; }
            }
            for (int j = 0; j < 4; j++) {
                int k = order[j];
                tk.Stack[stackTop].Node = nd.children[k];
                tk.Stack[stackTop].T = tk.T[k];
                stackTop -= tk.Hits[k];
            }
            if (stackTop > tk.maxDepth) tk.maxDepth = stackTop;
        } else {
            ray.LeafsT++;
            int32_t tmp = tk.StackTop;
            tk.StackTop = stackTop + 1;
            if (traceElems(LeafBase(node), LeafCount(node))) {
                hit = true;
                if (ray.Type & RayTypeShadow) { tk.StackTop = tmp; return true; }
            }
            tk.StackTop = tmp;
        }
    }
    return hit;
}

// qbvh/motionintersect.go:22-127
template <class Leaf>
static bool TraceMotionQBVH(const MNode* nodes, const std::vector<std::vector<float>>& boxes, float time, int key, int key2,
                            Ray& ray, Leaf&& traceElems) {
    Task& tk = *ray.task;
    int32_t stackTop = tk.StackTop;
    tk.Stack[stackTop].Node = 0;
    tk.Stack[stackTop].T = ray.Tclosest;
    bool hit = false;
    while (stackTop >= tk.StackTop) {
        int32_t node = tk.Stack[stackTop].Node;
        float T = tk.Stack[stackTop].T;
        stackTop--;
        if (ray.Tclosest < T) node = -1;
        if (node >= 0) {
            ray.NodesT++;
            const MNode& pn = nodes[node];
            const float* b0 = &boxes[key][node * 24];
            const float* b1 = &boxes[key2][node * 24];
            for (int i = 0; i < 24; i++) tk.Boxes[i] = (1.0f - time) * b0[i] + time * b1[i];
            intersectBoxes(ray, tk.Boxes, tk.Hits, tk.T);
            int order[4] = {0, 1, 2, 3};
            if (ray.D[pn.axis0] < 0) {
                if (ray.D[pn.axis2] < 0) { order[3] = 3; order[2] = 2; } else { order[3] = 2; order[2] = 3; }
                if (ray.D[pn.axis1] < 0) { order[1] = 1; order[0] = 0; } else { order[1] = 0; order[0] = 1; }
            } else {
                if (ray.D[pn.axis2] < 0) { order[1] = 3; order[0] = 2; } else { order[1] = 2; order[0] = 3; }
                if (ray.D[pn.axis1] < 0) { order[3] = 1; order[2] = 0; } else { order[3] = 0; order[2] = 1; }
            }
            for (int j = 0; j < 4; j++) {
                int k = order[j];
                if (tk.Hits[k] != 0) {
                    stackTop++;
                    tk.Stack[stackTop].Node = pn.children[k];
                    tk.Stack[stackTop].T = tk.T[k];
                }
            }
            if (stackTop + 1 > tk.maxDepth) tk.maxDepth = stackTop + 1;
        } else if (node < -1) {
            ray.LeafsT++;
            int32_t tmp = tk.StackTop;
            tk.StackTop = stackTop + 1;
            if (traceElems(LeafBase(node), LeafCount(node))) {
                hit = true;
                if (ray.Type & RayTypeShadow) { tk.StackTop = tmp; return true; }
            }
            tk.StackTop = tmp;
        }
    }
    return hit;
}

// -------------------------------------------------------------------- scene

struct SceneHit {
    int32_t geom;
    HitRecord hr;
};

struct Scene {
    std::vector<Mesh*> geoms;        // insertion order; after commit the top level's geoms in post-sort order
    std::vector<Mesh*> all;          // every geom in insertion order (owned), set at commit
    std::vector<int32_t> geomOrder;  // post-sort slot -> insertion index
    std::vector<Node> qbvh;
    std::vector<MNode> mnodes;
    std::vector<std::vector<float>> mboxes;
    bool motion = false;
    Box bounds;

    ~Scene() { for (auto* g : (all.empty() ? geoms : all)) delete g; }

    // builtin/scene/scene.go:135-203
    void initAccel() {
        all = geoms;
        std::vector<int32_t> vis;   // insertion indices of the top level's geoms
        for (int i = 0; i < (int)all.size(); i++) if (!all[i]->hidden) vis.push_back(i);
        geoms.clear();
        for (int i : vis) geoms.push_back(all[i]);
        int n = (int)geoms.size();
        std::vector<Box> boxes; std::vector<int32_t> indices; std::vector<Vec3> centroids;
        int maxKeys = 0;
        for (auto* g : geoms) if (g->MotionKeys() > maxKeys) maxKeys = g->MotionKeys();
        if (maxKeys == 1) {
            for (int i = 0; i < n; i++) { Box b = geoms[i]->Bounds(0); boxes.push_back(b); indices.push_back(i); centroids.push_back(b.Centroid()); }
            bounds = BuildAccel(qbvh, boxes.data(), centroids.data(), indices.data(), n, 1);
        } else {
            motion = true;
            for (int i = 0; i < n; i++) { Box b = geoms[i]->Bounds(0.5f); boxes.push_back(b); indices.push_back(i); centroids.push_back(b.Centroid()); }
            BuildAccelMotion(mnodes, boxes.data(), centroids.data(), indices.data(), n, 1);
        }
        std::vector<Mesh*> ng(n);
        for (int i = 0; i < n; i++) ng[i] = geoms[indices[i]];
        geoms.swap(ng);
        geomOrder.resize(n);
        for (int i = 0; i < n; i++) geomOrder[i] = vis[indices[i]];
        if (motion) {
            // scene.go:207-268
            mboxes.assign(maxKeys, std::vector<float>(mnodes.size() * 24, 0.0f));
            Box full; full.Reset();
            for (int k = 0; k < maxKeys; k++) full.GrowBox(initMotionBoxesRec(k, 0, maxKeys));
            bounds = full;
        }
    }
    Box initMotionBoxesRec(int key, int32_t node, int keys) {
        Box nodebox; nodebox.Reset();
        for (int c = 0; c < 4; c++) {
            int32_t ch = mnodes[node].children[c];
            if (ch < 0) {
                if (ch == -1) continue;
                Box box; box.Reset();
                int lb = LeafBase(ch), lc = LeafCount(ch);
                for (int i = lb; i < lb + lc; i++) {
                    float time = (float)key / (float)(keys - 1);
                    box.GrowBox(geoms[i]->Bounds(time));
                }
                SetBounds(&mboxes[key][node * 24], c, box);
                nodebox.GrowBox(box);
            } else {
                Box box = initMotionBoxesRec(key, ch, keys);
                SetBounds(&mboxes[key][node * 24], c, box);
                nodebox.GrowBox(box);
            }
        }
        return nodebox;
    }

    // builtin/geom/sphere/trace.go:54-103 (solveQuadratic, raySphereIntersect)
    static bool RaySphereIntersect(const Vec3& Ro, const Vec3& Rd, const Vec3& C, float radius, float& t) {
        Vec3 L; for (int k = 0; k < 3; k++) L[k] = Ro[k] - C[k];
        float a = Rd[0] * Rd[0] + Rd[1] * Rd[1] + Rd[2] * Rd[2];
        float b = 2 * (Rd[0] * L[0] + Rd[1] * L[1] + Rd[2] * L[2]);
        float c = (L[0] * L[0] + L[1] * L[1] + L[2] * L[2]) - radius * radius;
        float x0, x1;
        float discr = b * b - 4 * a * c;
        if (discr < 0) return false;
        if (discr == 0) { x1 = -0.5f * b / a; x0 = x1; }
        else {
            float q = b > 0 ? -0.5f * (b + std::sqrt(discr)) : -0.5f * (b - std::sqrt(discr));
            x0 = q / a; x1 = c / q;
        }
        if (x0 > x1) std::swap(x0, x1);
        if (x0 > x1) std::swap(x0, x1);
        if (x0 < 0) { x0 = x1; if (x0 < 0) return false; }
        t = x0;
        return true;
    }
    // Sphere.Trace (builtin/geom/sphere/trace.go:13-52): accepted when t < Tclosest, no ray bias.
    // The trace API reports prim 0 and u = v = 0 (the reference's Bu/Bv here are the sphere's UV
    // parameters, which the direct-lighting path never reads).
    static bool SphereTrace(const Mesh& m, Ray& ray, HitRecord& hr) {
        float t;
        if (!RaySphereIntersect(ray.P, ray.D, m.sphereC, m.sphereR, t)) return false;
        if (!(t < ray.Tclosest)) return false;
        ray.Tclosest = t;
        Vec3 P; for (int k = 0; k < 3; k++) P[k] = ray.P[k] + (ray.D[k] * t);
        Vec3 d; for (int k = 0; k < 3; k++) d[k] = P[k] - m.sphereC[k];
        Vec3 N = Mesh::Normalize(d);
        for (int k = 0; k < 3; k++) { hr.Poffset[k] = 0.001f * N[k]; hr.Ng[k] = N[k]; hr.N[k] = N[k]; hr.Po[k] = P[k]; }
        hr.DdPdu[0] = 1; hr.DdPdu[1] = 0; hr.DdPdu[2] = 0;
        hr.DdPdv[0] = 0; hr.DdPdv[1] = 0; hr.DdPdv[2] = 1;
        hr.elem = 0; hr.U = hr.V = hr.W = 0; hr.Bu = hr.Bv = 0; hr.shaderIdx = 0;
        return true;
    }

    // The ray moved into a geom's object space (instance.go:90-94, trace.go:49-53): P and D by the
    // inverse transform (D not renormalised, so t keeps its meaning), then Setup.  Restore puts the
    // ray back, Tclosest excepted (instance.go:98-104).
    struct SavedRay { Vec3 P, D, Dinv; float S[3]; int32_t Kx, Ky, Kz; };
    static SavedRay ToObject(Ray& ray, const M4 inv) {
        SavedRay sv{ray.P, ray.D, ray.Dinv, {ray.S[0], ray.S[1], ray.S[2]}, ray.Kx, ray.Ky, ray.Kz};
        ray.P = MulPoint(inv, sv.P);
        ray.D = MulVec(inv, sv.D);
        RaySetup(ray);
        return sv;
    }
    static void Restore(Ray& ray, const SavedRay& sv) {
        ray.P = sv.P; ray.D = sv.D; ray.Dinv = sv.Dinv;
        for (int k = 0; k < 3; k++) ray.S[k] = sv.S[k];
        ray.Kx = sv.Kx; ray.Ky = sv.Ky; ray.Kz = sv.Kz;
    }

    // builtin/geom/polymesh/trace.go:15-102 (and the analytic sphere).  A mesh's own Transform takes
    // the ray into object space; after a hit sg.Transform = that matrix, sg.P = M Po and sg.N =
    // M^-T N (ApplyTransform then transforms N once more).  The reference sets those after every
    // transformed mesh it traces, hit or not, so a transformed mesh traced after the closest hit
    // re-transforms that hit's frame; here they follow the mesh that holds the closest hit.
    bool MeshTrace(const Mesh& mesh, Ray& ray, SceneHit& sh) const {
        HitRecord& hr = sh.hr;
        if (mesh.sphere) {
            if (!SphereTrace(mesh, ray, hr)) return false;
            hr.leaf = &mesh; hr.hasXform = false; hr.meshXform = false;
            return true;
        }
        M4 M, Minv;
        SavedRay sv{};
        const bool xf = !mesh.xform.empty();
        if (xf) {
            mesh.TransformAt(ray.Time, M);
            M4Inverse(M, Minv);
            sv = ToObject(ray, Minv);
        }
        bool hit;
        if (mesh.nkeys == 1) {
            hit = TraceQBVH(mesh.qbvh.data(), ray, [&](int base, int count) { return mesh.TraceElems(ray, hr, base, count); });
        } else {
            float k = ray.Time * (float)(mesh.nkeys - 1);
            float time = k - Floor(k);
            int key = (int)Floor(k), key2 = (int)Ceil(k);
            hit = TraceMotionQBVH(mesh.mnodes.data(), mesh.mboxes, time, key, key2, ray,
                                  [&](int base, int count) { return mesh.TraceMotionElems(time, key, key2, ray, hr, base, count); });
        }
        if (xf) Restore(ray, sv);
        if (hit) {
            hr.leaf = &mesh;
            hr.hasXform = xf;
            hr.meshXform = xf;
            if (xf) {
                std::memcpy(hr.xform, M, sizeof(M4));
                std::memcpy(hr.xformInv, Minv, sizeof(M4));
                std::memcpy(hr.meshInv, Minv, sizeof(M4));
            }
        }
        return hit;
    }

    // Geom.Trace: meshes and spheres, GeomInstance.Trace (instance.go:73-114) and Proc.Trace
    // (proc.go:88-139).  Both move the ray by the inverse of their transform at the ray's time and set
    // sg.Transform only on a hit.  A proc tries its geoms in order and, for any ray but a shadow ray,
    // stops at the first one that reports a hit (proc.go:112-120), closer or not.
    bool GeomTrace(const Mesh& g, Ray& ray, SceneHit& sh) const {
        if (g.kind == GEOM_MESH) return MeshTrace(g, ray, sh);
        M4 M, Minv;
        g.TransformAt(ray.Time, M);
        M4Inverse(M, Minv);
        SavedRay sv = ToObject(ray, Minv);
        bool hit = false;
        for (const Mesh* c : g.children) {
            if (GeomTrace(*c, ray, sh)) {
                hit = true;
                if (g.kind == GEOM_INSTANCE || (ray.Type & RayTypeShadow) == 0) break;
            }
        }
        Restore(ray, sv);
        if (hit) {
            sh.hr.hasXform = true;
            std::memcpy(sh.hr.xform, M, sizeof(M4));
            std::memcpy(sh.hr.xformInv, Minv, sizeof(M4));
        }
        return hit;
    }

    // scene.go:61-97
    bool TraceElems(Ray& ray, SceneHit& sh, int base, int count) const {
        bool hit = false;
        for (int i = base; i < base + count; i++) {
            if (GeomTrace(*geoms[i], ray, sh)) {
                sh.geom = i;
                hit = true;
                if (ray.Type & RayTypeShadow) return true;
            }
        }
        return hit;
    }

    // scene.go:30-58
    bool Trace(Ray& ray, SceneHit& sh) const {
        if (!motion)
            return TraceQBVH(qbvh.data(), ray, [&](int base, int count) { return TraceElems(ray, sh, base, count); });
        float k = ray.Time * (float)(mboxes.size() - 1);
        float time = k - Floor(k);
        int key = (int)Floor(k), key2 = (int)Ceil(k);
        return TraceMotionQBVH(mnodes.data(), mboxes, time, key, key2, ray,
                               [&](int base, int count) { return TraceElems(ray, sh, base, count); });
    }
};

}  // namespace orc
