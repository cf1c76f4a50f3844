// Host-side QBVH builders: a parallel builder whose output equals the
// reference's single-threaded qbvh.BuildAccel / BuildAccelMotion node for
// node (qbvh/build.go:28-307, qbvh/motionbuild.go:12-122).
//
// Why it must be identical: hit IDs are post-sort triangle indices
// (builtin/geom/polymesh/trace.go:513), so the leaf order — and with it the
// binning arithmetic, the strict first-minimum SAH choice and the unstable
// two-pointer partition — has to match the reference bit for bit.
//
// What is different: elements are kept as one 40-byte record (box, centroid,
// index) so a partition swap moves one record, and subtrees above a size
// threshold are built on separate threads into private node arrays that are
// appended in child order afterwards, which keeps the depth-first node
// numbering of the sequential recursion.
//
// Must be compiled with -ffp-contract=off (float32 results have to round
// exactly as the reference's unfused SSE arithmetic).

#include "qbvh_build.h"

#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace vgpu {

namespace {

const float kInf = std::numeric_limits<float>::infinity();

// MINSS/MAXSS lane semantics (math/dim_amd64.s:8-21)
inline float mins(float x, float y) { return (x < y) ? x : y; }
inline float maxs(float x, float y) { return (x > y) ? x : y; }

struct Elem {
    float lo[3], hi[3];   // triangle (or geom) box
    float c[3];           // centroid used for binning
    int32_t idx;
};
static_assert(sizeof(Elem) == 40, "element record");

struct BBox {
    float lo[3], hi[3];
    void reset() { for (int k = 0; k < 3; k++) { lo[k] = kInf; hi[k] = -kInf; } }
    // BoundingBox.Grow (math/boundingbox.go:107-116): Min(X, b), Max(X, b)
    void grow_point(const float* p) { for (int k = 0; k < 3; k++) { lo[k] = mins(p[k], lo[k]); hi[k] = maxs(p[k], hi[k]); } }
    // BoundingBox.GrowBox (math/boundingbox.go:99-104): Min(b, p), Max(b, p)
    void grow_box(const float* plo, const float* phi) { for (int k = 0; k < 3; k++) { lo[k] = mins(lo[k], plo[k]); hi[k] = maxs(hi[k], phi[k]); } }
    void grow_box(const BBox& b) { grow_box(b.lo, b.hi); }
    float dim(int a) const { return hi[a] - lo[a]; }
    float area() const { return dim(0) * dim(1) * 2.0f + dim(1) * dim(2) * 2.0f + dim(0) * dim(2) * 2.0f; }
    int longest() const {
        if (dim(0) < dim(1)) return (dim(1) < dim(2)) ? 2 : 1;
        return (dim(0) < dim(2)) ? 2 : 0;
    }
};

struct BuildFailure {};

struct Split { int axis; int pivot; };

// binarySplit + calcMinCost on elements [0, n) of e.
Split split_set(Elem* e, int n, int leafMax) {
    BBox cb; cb.reset();
    for (int i = 0; i < n; i++) cb.grow_point(e[i].c);
    if (n <= leafMax) return {0, n};
    const int axis = cb.longest();
    if (cb.hi[axis] == cb.lo[axis]) return {axis, n / 2 + 1};
    const float k1 = 7.99952f / (cb.hi[axis] - cb.lo[axis]);
    const float k0 = cb.lo[axis];
    BBox bins[8];
    int32_t cnt[8] = {0};
    for (int b = 0; b < 8; b++) bins[b].reset();
    for (int i = 0; i < n; i++) {
        const int bin = (int)(k1 * (e[i].c[axis] - k0));
        if (bin < 0 || bin > 7) throw BuildFailure();
        cnt[bin]++;
        bins[bin].grow_box(e[i].lo, e[i].hi);
    }
    BBox lb[8], rb[8];
    int32_t ln[8], rn[8];
    BBox acc; acc.reset();
    int32_t m = 0;
    for (int b = 0; b < 8; b++) { acc.grow_box(bins[b]); m += cnt[b]; lb[b] = acc; ln[b] = m; }
    acc.reset(); m = 0;
    for (int b = 7; b >= 0; b--) { acc.grow_box(bins[b]); m += cnt[b]; rb[b] = acc; rn[b] = m; }
    int best = -1;
    float bestCost = kInf;
    for (int b = 1; b < 8; b++) {
        const float cost = lb[b - 1].area() * (float)ln[b - 1] + rb[b].area() * (float)rn[b];
        if (cost < bestCost) { best = b; bestCost = cost; }
    }
    int left = 0, right = n - 1;
    while (left <= right) {
        const int bin = (int)(k1 * (e[left].c[axis] - k0));
        if (bin < best) {
            left++;
        } else {
            Elem t = e[left]; e[left] = e[right]; e[right] = t;
            right--;
        }
    }
    return {axis, left};
}

BBox leaf_box(const Elem* e, int n) {
    BBox b; b.reset();
    for (int i = 0; i < n; i++) b.grow_box(e[i].lo, e[i].hi);
    return b;
}

inline void set_child_box(vgpu_qbvh_node& nd, int c, const BBox& b) {
    for (int k = 0; k < 3; k++) { nd.boxes[c + k * 4] = b.lo[k]; nd.boxes[c + 12 + k * 4] = b.hi[k]; }
}
inline BBox child_box(const vgpu_qbvh_node& nd, int c) {
    BBox b;
    for (int k = 0; k < 3; k++) { b.lo[k] = nd.boxes[c + k * 4]; b.hi[k] = nd.boxes[c + 12 + k * 4]; }
    return b;
}
inline int32_t leaf_code(uint32_t first, uint32_t count) {
    return (int32_t)((1u << 31) | ((first << 4) & 0xfffffff0u) | ((count - 1) & 0xfu));
}

const int kParallelMin = 1 << 15;   // subtrees at least this large may get their own thread

struct Builder {
    int leafMax;
    bool motion;
    std::atomic<int> spare_threads;

    template <class NodeT>
    struct Out { std::vector<NodeT> nodes; };

    // Builds the subtree over e[0,n) (global element offset base) into `out`, appending in
    // depth-first pre-order; returns the subtree root's index within `out` and its box.
    template <class NodeT>
    int32_t build(std::vector<NodeT>& out, Elem* e, int n, int base, BBox& box_out, int depth) {
        if (depth > 4096) throw BuildFailure();
        const Split s0 = split_set(e, n, leafMax);
        const Split s1 = split_set(e, s0.pivot, leafMax);
        const Split s2 = split_set(e + s0.pivot, n - s0.pivot, leafMax);
        const int32_t me = (int32_t)out.size();
        out.push_back(NodeT());
        std::memset(&out[me], 0, sizeof(NodeT));
        set_axes(out[me], s0.axis, s1.axis, s2.axis);
        const int lo[4] = {0, s1.pivot, s0.pivot, s0.pivot + s2.pivot};
        const int hi[4] = {s1.pivot, s0.pivot, s0.pivot + s2.pivot, n};

        // Large children: build into private arrays (possibly on other threads), append in order.
        bool big = false;
        for (int c = 0; c < 4; c++) if (hi[c] - lo[c] > leafMax && hi[c] - lo[c] >= kParallelMin) big = true;

        BBox cbox[4];
        if (!big) {
            for (int c = 0; c < 4; c++) {
                const int cn = hi[c] - lo[c];
                if (cn <= leafMax) {
                    cbox[c] = leaf_box(e + lo[c], cn);
                    set_leaf(out[me], c, (uint32_t)(base + lo[c]), (uint32_t)cn, cbox[c]);
                } else {
                    const int32_t ch = build(out, e + lo[c], cn, base + lo[c], cbox[c], depth + 1);
                    set_child(out[me], c, ch, cbox[c]);
                }
            }
        } else {
            std::vector<NodeT> sub[4];
            std::thread th[4];
            bool spawned[4] = {false, false, false, false};
            bool failed[4] = {false, false, false, false};
            for (int c = 0; c < 4; c++) {
                const int cn = hi[c] - lo[c];
                if (cn <= leafMax) continue;
                if (cn >= kParallelMin && spare_threads.fetch_sub(1) > 0) {
                    spawned[c] = true;
                    th[c] = std::thread([this, &sub, &cbox, &failed, e, lo, cn, base, depth, c]() {
                        try { build(sub[c], e + lo[c], cn, base + lo[c], cbox[c], depth + 1); }
                        catch (...) { failed[c] = true; }
                        spare_threads.fetch_add(1);
                    });
                } else {
                    if (cn >= kParallelMin) spare_threads.fetch_add(1);
                    // A failure here must not unwind past the joinable sibling threads (and the
                    // sub/cbox/failed arrays they write): record it and rethrow after every join.
                    try { build(sub[c], e + lo[c], cn, base + lo[c], cbox[c], depth + 1); }
                    catch (...) { failed[c] = true; }
                }
            }
            for (int c = 0; c < 4; c++) if (spawned[c]) th[c].join();
            for (int c = 0; c < 4; c++) if (failed[c]) throw BuildFailure();
            for (int c = 0; c < 4; c++) {
                const int cn = hi[c] - lo[c];
                if (cn <= leafMax) {
                    cbox[c] = leaf_box(e + lo[c], cn);
                    set_leaf(out[me], c, (uint32_t)(base + lo[c]), (uint32_t)cn, cbox[c]);
                } else {
                    const int32_t off = (int32_t)out.size();
                    for (auto& nd : sub[c]) {
                        for (int k = 0; k < 4; k++) if (nd.children[k] >= 0) nd.children[k] += off;
                        out.push_back(nd);
                    }
                    set_child(out[me], c, off, cbox[c]);
                }
            }
        }
        BBox nb; nb.reset();
        for (int c = 0; c < 4; c++) if (out[me].children[c] != -1) nb.grow_box(cbox_of(out[me], c, cbox[c]));
        box_out = nb;
        return me;
    }

    static void set_axes(vgpu_qbvh_node& nd, int a0, int a1, int a2) { nd.axis0 = a0; nd.axis1 = a1; nd.axis2 = a2; }
    static void set_axes(vgpu_mqbvh_node& nd, int a0, int a1, int a2) { nd.axis0 = a0; nd.axis1 = a1; nd.axis2 = a2; }
    // qbvh.Node.SetBounds + SetLeaf (empty leaves get InfBox, qbvh/qbvh.go:67-79)
    static void set_leaf(vgpu_qbvh_node& nd, int c, uint32_t first, uint32_t count, const BBox& b) {
        if (count == 0) {
            nd.children[c] = -1;
            BBox inf; for (int k = 0; k < 3; k++) { inf.lo[k] = kInf; inf.hi[k] = kInf; }
            set_child_box(nd, c, inf);
            return;
        }
        set_child_box(nd, c, b);
        nd.children[c] = leaf_code(first, count);
    }
    static void set_leaf(vgpu_mqbvh_node& nd, int c, uint32_t first, uint32_t count, const BBox&) {
        nd.children[c] = count == 0 ? -1 : leaf_code(first, count);
    }
    static void set_child(vgpu_qbvh_node& nd, int c, int32_t ch, const BBox& b) { set_child_box(nd, c, b); nd.children[c] = ch; }
    static void set_child(vgpu_mqbvh_node& nd, int c, int32_t ch, const BBox&) { nd.children[c] = ch; }
    // The node box unions the stored child boxes (qbvh/build.go:277-284).
    static BBox cbox_of(const vgpu_qbvh_node& nd, int c, const BBox&) { return child_box(nd, c); }
    static BBox cbox_of(const vgpu_mqbvh_node&, int, const BBox& b) { return b; }
};

template <class NodeT>
int run_build(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leafMax, NodeT* nodes_out,
              int32_t max_nodes, int32_t* nnodes_out, float* bounds_out, int32_t nthreads, bool motion) {
    if (n < 0 || !nnodes_out || (n > 0 && (!boxes || !centroids || !idx))) return -1;
    if (leafMax > 16) leafMax = 16;
    if (leafMax < 1) leafMax = 1;
    std::vector<Elem> e((size_t)n);
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            e[i].lo[k] = boxes[i * 6 + k]; e[i].hi[k] = boxes[i * 6 + 3 + k]; e[i].c[k] = centroids[i * 3 + k];
        }
        e[i].idx = idx[i];
    }
    Builder b;
    b.leafMax = leafMax;
    b.motion = motion;
    int hw = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
    b.spare_threads = hw > 1 ? hw - 1 : 0;
    std::vector<NodeT> nodes;
    BBox root;
    try {
        b.build(nodes, e.data(), n, 0, root, 0);
    } catch (...) {
        return -3;
    }
    *nnodes_out = (int32_t)nodes.size();
    if ((int32_t)nodes.size() > max_nodes || !nodes_out) return -1;
    std::memcpy(nodes_out, nodes.data(), nodes.size() * sizeof(NodeT));
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            boxes[i * 6 + k] = e[i].lo[k]; boxes[i * 6 + 3 + k] = e[i].hi[k]; centroids[i * 3 + k] = e[i].c[k];
        }
        idx[i] = e[i].idx;
    }
    if (bounds_out) for (int k = 0; k < 3; k++) { bounds_out[k] = root.lo[k]; bounds_out[3 + k] = root.hi[k]; }
    return 0;
}

}  // namespace

int build_qbvh(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leafMax, vgpu_qbvh_node* nodes_out,
               int32_t max_nodes, int32_t* nnodes_out, float* bounds_out, int32_t nthreads) {
    return run_build(boxes, centroids, idx, n, leafMax, nodes_out, max_nodes, nnodes_out, bounds_out, nthreads, false);
}

int build_qbvh_motion(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leafMax, vgpu_mqbvh_node* nodes_out,
                      int32_t max_nodes, int32_t* nnodes_out, int32_t nthreads) {
    return run_build(boxes, centroids, idx, n, leafMax, nodes_out, max_nodes, nnodes_out, nullptr, nthreads, true);
}

}  // namespace vgpu
