// QBVH build on the GPU with the reference's output: qbvh.BuildAccel (qbvh/build.go:28-307) node for
// node, with boxes / centroids / indxs permuted into the same leaf order.
//
// The reference recurses depth first.  Each node takes three binary splits (the whole range, then
// its two halves); a split bins the centroids of the longest axis of their bounds into 8 bins
// (k1 = 7.99952 / extent), picks the first strict minimum of the SAH cost and partitions with an
// unstable two-pointer loop; children with at most leafMax elements become leaves.  Every float32
// rounding and every swap decides the leaf order, so the arithmetic here is the reference's, and the
// two-pointer loop is replaced by a closed form of its result that each element computes alone:
//
//   The loop examines elements from two streams, positions 0, 1, ... (the left stream) and n-1,
//   n-2, ... (the right stream): it reads the left stream until it meets a right element (bin >=
//   best), then the right stream until it meets a left element, and so on; left elements land at
//   0, 1, ... and right elements at n-1, n-2, ... in the order examined.  With nl left elements the
//   left stream covers [0, s), s = nl + 1 if position nl holds a right element, else nl.  Then for
//   p < s a left element stays at p, and the k-th right element from position 0 goes to
//   n - 2 - lr[k-2] (k = 1: n - 1), where lr[j] = n - 1 - (position of the (j+1)-th left element
//   counted from n - 1); for p >= s a right element goes to p - 1 and the k-th left element counted
//   from n - 1 goes to rl[k-1], the position of the k-th right element counted from 0.
//
// Box arithmetic follows MINSS / MAXSS (math/dim_amd64.s): x < y ? x : y.  The reference grows boxes
// element by element; reductions here combine partial boxes in element order with the operand order
// that keeps the same element on ties, so the sign of zero bounds matches as well.
//
// Schedule.  Ranges above kSubtreeMax elements are split level by level with device-wide kernels
// over 4096-element chunks (all nodes of a level in one batch of launches; the host reads the pivots
// back after each batch).  A range of at most kSubtreeMax elements becomes a subtree that one
// wavefront builds completely, depth first with an explicit stack, into a scratch node area at its
// element offset.  The host then numbers the nodes in the reference's depth-first pre-order and one
// kernel moves every subtree into place.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <limits>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "qbvh_build.h"

namespace vgpu {

namespace {

constexpr int kSegThreads = 256;                 // threads of a chunk kernel
constexpr int kSegPer = 16;                      // contiguous elements per thread
constexpr int kChunk = kSegThreads * kSegPer;    // elements per chunk
constexpr int kSubtreeMax = 2048;                // largest range one wavefront builds
constexpr int kStackMax = 192;                   // subtree depth limit

// One element: a = (lo xyz, index bits), b = (hi xyz, 0), c = (centroid xyz, 0).
struct alignas(16) GElem {
    float4 a, b, c;
};

struct Box {
    float lo[3], hi[3];
};

__device__ __forceinline__ float mins(float x, float y) { return (x < y) ? x : y; }
__device__ __forceinline__ float maxs(float x, float y) { return (x > y) ? x : y; }

__device__ __forceinline__ Box box_empty() {
    const float inf = __builtin_inff();
    return Box{{inf, inf, inf}, {-inf, -inf, -inf}};
}
// BoundingBox.GrowBox (math/boundingbox.go:99-104): Min(acc, b), Max(acc, b); ties take b.  Two
// partial folds L (earlier elements) and R (later) combine as grow_box(L, R).
__device__ __forceinline__ void grow_box(Box& acc, const Box& b) {
    for (int k = 0; k < 3; k++) { acc.lo[k] = mins(acc.lo[k], b.lo[k]); acc.hi[k] = maxs(acc.hi[k], b.hi[k]); }
}
// BoundingBox.Grow (math/boundingbox.go:107-116): Min(p, acc), Max(p, acc); ties keep acc.
__device__ __forceinline__ void grow_point(Box& acc, const float4& p) {
    const float q[3] = {p.x, p.y, p.z};
    for (int k = 0; k < 3; k++) { acc.lo[k] = mins(q[k], acc.lo[k]); acc.hi[k] = maxs(q[k], acc.hi[k]); }
}
// Two partial point folds, L earlier and R later: the earlier one wins ties.
__device__ __forceinline__ void join_points(Box& L, const Box& R) {
    for (int k = 0; k < 3; k++) { L.lo[k] = mins(R.lo[k], L.lo[k]); L.hi[k] = maxs(R.hi[k], L.hi[k]); }
}
__device__ __forceinline__ float dim(const Box& b, int a) { return b.hi[a] - b.lo[a]; }
__device__ __forceinline__ float area(const Box& b) {
    return dim(b, 0) * dim(b, 1) * 2.0f + dim(b, 1) * dim(b, 2) * 2.0f + dim(b, 0) * dim(b, 2) * 2.0f;
}
__device__ __forceinline__ int longest(const Box& b) {
    if (dim(b, 0) < dim(b, 1)) return (dim(b, 1) < dim(b, 2)) ? 2 : 1;
    return (dim(b, 0) < dim(b, 2)) ? 2 : 0;
}
__device__ __forceinline__ Box elem_box(const GElem& e) {
    return Box{{e.a.x, e.a.y, e.a.z}, {e.b.x, e.b.y, e.b.z}};
}
__device__ __forceinline__ float cen(const float4& c, int axis) { return axis == 0 ? c.x : (axis == 1 ? c.y : c.z); }

// The binning of one split: bin(c) = int(k1 * (c[axis] - k0)); right side: bin >= best.
struct Cut {
    int axis, best;
    float k0, k1;
    __device__ __forceinline__ bool right(const float4& c) const { return !((int)(k1 * (cen(c, axis) - k0)) < best); }
};

// ---------------------------------------------------------------- wavefront and block reductions

template <bool POINTS>
__device__ __forceinline__ Box wave_fold(Box v) {
    const int lane = threadIdx.x & 63;
    for (int d = 1; d < 64; d <<= 1) {
        Box o;
        for (int k = 0; k < 3; k++) { o.lo[k] = __shfl_down(v.lo[k], d); o.hi[k] = __shfl_down(v.hi[k], d); }
        if ((lane & (2 * d - 1)) == 0) {
            if (POINTS) join_points(v, o);
            else grow_box(v, o);
        }
    }
    for (int k = 0; k < 3; k++) { v.lo[k] = __shfl(v.lo[k], 0); v.hi[k] = __shfl(v.hi[k], 0); }
    return v;
}

__device__ __forceinline__ int wave_scan(int x) {   // inclusive
    const int lane = threadIdx.x & 63;
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    return x;
}
__device__ __forceinline__ int wave_sum(int x) { return __shfl(wave_scan(x), 63); }

constexpr int kWaves = kSegThreads / 64;

template <bool POINTS>
__device__ Box block_fold(Box v, Box* sh) {
    v = wave_fold<POINTS>(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    Box r = sh[0];
    for (int w = 1; w < kWaves; w++) {
        if (POINTS) join_points(r, sh[w]);
        else grow_box(r, sh[w]);
    }
    __syncthreads();
    return r;
}

// Exclusive scan over the block (thread order); *total: the sum.
__device__ int block_scan_excl(int x, int* sh, int* total) {
    const int inc = wave_scan(x);
    if ((threadIdx.x & 63) == 63) sh[threadIdx.x >> 6] = inc;
    __syncthreads();
    int base = 0, tot = 0;
    for (int w = 0; w < kWaves; w++) {
        if (w < (int)(threadIdx.x >> 6)) base += sh[w];
        tot += sh[w];
    }
    __syncthreads();
    *total = tot;
    return base + inc - x;
}

// Per-thread bins (counts and boxes grown in element order), updated without dynamic indexing.
struct Bins {
    Box b[8];
    int n[8];
};
__device__ __forceinline__ void bins_init(Bins& B) {
#pragma unroll
    for (int k = 0; k < 8; k++) { B.b[k] = box_empty(); B.n[k] = 0; }
}
// Adds e to its bin; sets err bit 1 for a bin outside [0, 7] (the reference's index panics there).
__device__ __forceinline__ void bins_add(Bins& B, const GElem& e, const Cut& cut, int* err) {
    const float f = cut.k1 * (cen(e.c, cut.axis) - cut.k0);
    const int bin = (int)f;
    if (f != f || bin < 0 || bin > 7) { atomicOr(err, 1); return; }
    const Box eb = elem_box(e);
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (bin == k) { grow_box(B.b[k], eb); B.n[k]++; }
}

// calcMinCost's SAH choice (qbvh/build.go:60-107) over the folded bins: left boxes and counts grown
// from bin 0, right ones from bin 7, the first strict minimum of area(L) * nL + area(R) * nR.
__device__ void sah(const Bins& B, int* best_out, int* nl_out) {
    Box rb[8];
    int rn[8];
    Box acc = box_empty();
    int m = 0;
#pragma unroll
    for (int b = 7; b >= 0; b--) { grow_box(acc, B.b[b]); m += B.n[b]; rb[b] = acc; rn[b] = m; }
    acc = box_empty();
    m = 0;
    int best = -1;
    float bestCost = __builtin_inff();
#pragma unroll
    for (int b = 1; b < 8; b++) {
        grow_box(acc, B.b[b - 1]);
        m += B.n[b - 1];
        const float cost = area(acc) * (float)m + area(rb[b]) * (float)rn[b];
        if (cost < bestCost) { best = b; bestCost = cost; }
    }
    int nl = 0;
#pragma unroll
    for (int b = 0; b < 8; b++) if (b < best) nl += B.n[b];
    *best_out = best;
    *nl_out = nl;
}

// Destination of position p (rank r = right elements in [0, p], inclusive) under the closed form;
// -1 for an index outside [0, n), which only inconsistent bins (an error already flagged) produce.
__device__ __forceinline__ int dest_of(int p, bool right, int r, int n, int nl, int s, const int* rl, const int* lr) {
    int d;
    if (p < s) {
        if (!right) return p;
        if (r < 2) return n - 1;
        if (r - 2 >= n) return -1;
        d = n - 2 - lr[r - 2];
    } else {
        if (right) return p - 1;
        const int k = nl - (p - r) - 1;
        if (k < 0 || k >= n) return -1;
        d = rl[k];
    }
    return (d >= 0 && d < n) ? d : -1;
}
// Records position i in rl / lr (r: right elements in [0, i], inclusive); false if out of range.
__device__ __forceinline__ bool record_rank(int i, bool right, int r, int n, int nl, int* rl, int* lr) {
    if (right) {
        if (r - 1 >= n) return false;
        rl[r - 1] = i;
        return true;
    }
    const int k = nl - (i - r) - 1;
    if (k < 0 || k >= n) return false;
    lr[k] = n - 1 - i;
    return true;
}

// ---------------------------------------------------------------- one split by one wavefront

struct SplitOut {
    int axis, pivot;
};

__device__ __forceinline__ void wave_copy(const GElem* src, GElem* dst, int n) {
    for (int i = threadIdx.x; i < n; i += 64) dst[i] = src[i];
}

// binarySplit (qbvh/build.go:151-178) of src[0, n) into dst[0, n) by one wavefront (blockDim 64).
// rl / lr: n ints each (LDS).  Lane l takes the contiguous slice [l * per, (l + 1) * per).
__device__ SplitOut split_wave(const GElem* src, GElem* dst, int n, int leafMax, int* rl, int* lr, int* err) {
    if (n <= leafMax) {
        wave_copy(src, dst, n);
        __syncthreads();
        return SplitOut{0, n};
    }
    const int lane = threadIdx.x;
    const int per = (n + 63) >> 6;
    const int a = min(n, lane * per), b = min(n, a + per);
    Box cb = box_empty();
    for (int i = a; i < b; i++) grow_point(cb, src[i].c);
    cb = wave_fold<true>(cb);
    const int axis = longest(cb);
    if (cb.hi[axis] == cb.lo[axis]) {
        wave_copy(src, dst, n);
        __syncthreads();
        return SplitOut{axis, n / 2 + 1};
    }
    Cut cut;
    cut.axis = axis;
    cut.k1 = 7.99952f / (cb.hi[axis] - cb.lo[axis]);
    cut.k0 = cb.lo[axis];
    Bins B;
    bins_init(B);
    for (int i = a; i < b; i++) bins_add(B, src[i], cut, err);
#pragma unroll
    for (int k = 0; k < 8; k++) { B.b[k] = wave_fold<false>(B.b[k]); B.n[k] = wave_sum(B.n[k]); }
    int nl;
    sah(B, &cut.best, &nl);
    int cnt = 0;
    for (int i = a; i < b; i++) cnt += cut.right(src[i].c) ? 1 : 0;
    const int base = wave_scan(cnt) - cnt;
    int r = base;
    for (int i = a; i < b; i++) {
        const bool right = cut.right(src[i].c);
        r += right ? 1 : 0;
        if (!record_rank(i, right, r, n, nl, rl, lr)) atomicOr(err, 1);
    }
    __syncthreads();
    const int s = nl + ((nl < n && cut.right(src[nl].c)) ? 1 : 0);
    r = base;
    for (int i = a; i < b; i++) {
        const GElem e = src[i];
        const bool right = cut.right(e.c);
        r += right ? 1 : 0;
        const int d = dest_of(i, right, r, n, nl, s, rl, lr);
        if (d >= 0) dst[d] = e;
        else atomicOr(err, 1);
    }
    __syncthreads();
    return SplitOut{axis, nl};
}

__device__ __forceinline__ void set_child_box(vgpu_qbvh_node& nd, int c, const Box& b) {
    for (int k = 0; k < 3; k++) { nd.boxes[c + k * 4] = b.lo[k]; nd.boxes[c + 12 + k * 4] = b.hi[k]; }
}
__device__ __forceinline__ Box child_box(const vgpu_qbvh_node& nd, int c) {
    Box b;
    for (int k = 0; k < 3; k++) { b.lo[k] = nd.boxes[c + k * 4]; b.hi[k] = nd.boxes[c + 12 + k * 4]; }
    return b;
}
__device__ __forceinline__ int32_t leaf_code(uint32_t first, uint32_t count) {
    return (int32_t)((1u << 31) | ((first << 4) & 0xfffffff0u) | ((count - 1) & 0xfu));
}

// calcBox of a leaf (GrowBox in element order), by one thread.
__device__ Box leaf_box(const GElem* E, int n) {
    Box b = box_empty();
    for (int i = 0; i < n; i++) grow_box(b, elem_box(E[i]));
    return b;
}

// SetBounds + SetLeaf; an empty leaf is child -1 with the InfBox (qbvh/qbvh.go:67-79).
__device__ void set_leaf(vgpu_qbvh_node& nd, int c, uint32_t first, uint32_t count, const Box& b) {
    if (count == 0) {
        const float inf = __builtin_inff();
        nd.children[c] = -1;
        set_child_box(nd, c, Box{{inf, inf, inf}, {inf, inf, inf}});
        return;
    }
    set_child_box(nd, c, b);
    nd.children[c] = leaf_code(first, count);
}

struct Frame {
    int start, n, node, next;   // next: -1 before the splits, then the child to visit
    int p1, p0, p02;            // child c covers [lo(c), lo(c + 1)) with lo = {0, p1, p0, p02, n}
};
__device__ __forceinline__ int frame_lo(const Frame& f, int c) {
    return c == 0 ? 0 : (c == 1 ? f.p1 : (c == 2 ? f.p0 : (c == 3 ? f.p02 : f.n)));
}

// buildAccelRec (qbvh/build.go:180-287) of one subtree per workgroup of one wavefront.  Elements of
// subtree s: E[roots[s].x, + roots[s].y), built in place with T as the second buffer; nodes in local
// pre-order at nodes[roots[s].x ..] (the subtree has at most max(n, 1) nodes); counts[s]: the node
// count; root_boxes[s]: the root's box.  err bit 2: node area exceeded (only a range whose recursion
// never ends, as two coincident centroids with leafMax 1); bit 8: deeper than kStackMax levels.
__global__ __launch_bounds__(64) void k_subtree(GElem* E_all, GElem* T_all, const int2* roots,
                                                vgpu_qbvh_node* nodes_all, int* counts, float* root_boxes,
                                                int leafMax, int* err) {
    __shared__ int rl[kSubtreeMax], lr[kSubtreeMax];
    __shared__ Frame st[kStackMax];
    __shared__ int sh_sp, sh_nn;
    __shared__ float sh_box[6];
    const int2 root = roots[blockIdx.x];
    GElem* E = E_all + root.x;
    GElem* T = T_all + root.x;
    vgpu_qbvh_node* nodes = nodes_all + root.x;
    const int cap = max(root.y, 1);
    const int t = threadIdx.x;
    if (t == 0) {
        st[0] = Frame{0, root.y, -1, -1, 0, 0, 0};
        sh_sp = 1;
        sh_nn = 0;
        const float inf = __builtin_inff();
        for (int k = 0; k < 3; k++) { sh_box[k] = inf; sh_box[3 + k] = -inf; }
    }
    __syncthreads();
    while (true) {
        const int top = sh_sp - 1;
        if (top < 0) break;
        const Frame f = st[top];
        __syncthreads();
        if (f.next < 0) {
            // the node's three binary splits: E -> T over the range, then each half T -> E
            const SplitOut s0 = split_wave(E + f.start, T + f.start, f.n, leafMax, rl, lr, err);
            const SplitOut s1 = split_wave(T + f.start, E + f.start, s0.pivot, leafMax, rl, lr, err);
            const SplitOut s2 = split_wave(T + f.start + s0.pivot, E + f.start + s0.pivot, f.n - s0.pivot, leafMax, rl,
                                           lr, err);
            if (t == 0) {
                const int me = sh_nn;
                if (me >= cap) {
                    atomicOr(err, 2);
                    sh_sp = 0;
                } else {
                    sh_nn = me + 1;
                    vgpu_qbvh_node& nd = nodes[me];
                    for (int k = 0; k < 24; k++) nd.boxes[k] = 0.0f;
                    for (int c = 0; c < 4; c++) nd.children[c] = 0;
                    nd.parent = 0;
                    nd.axis0 = (uint32_t)s0.axis;
                    nd.axis1 = (uint32_t)s1.axis;
                    nd.axis2 = (uint32_t)s2.axis;
                    st[top] = Frame{f.start, f.n, me, 0, s1.pivot, s0.pivot, s0.pivot + s2.pivot};
                }
            }
        } else if (f.next < 4) {
            if (t == 0) {
                const int c = f.next;
                const int lo = frame_lo(f, c), cn = frame_lo(f, c + 1) - lo;
                st[top].next = c + 1;
                if (cn <= leafMax) {
                    set_leaf(nodes[f.node], c, (uint32_t)(root.x + f.start + lo), (uint32_t)cn,
                             leaf_box(E + f.start + lo, cn));
                } else if (top + 1 >= kStackMax) {
                    atomicOr(err, 8);
                    sh_sp = 0;
                } else {
                    st[top + 1] = Frame{f.start + lo, cn, -1, -1, 0, 0, 0};
                    sh_sp = top + 2;
                }
            }
        } else if (t == 0) {
            // the node is complete: its box unions the stored non-empty child boxes
            // (qbvh/build.go:277-284), and it fills its slot in the parent
            const vgpu_qbvh_node& nd = nodes[f.node];
            Box nb = box_empty();
            for (int c = 0; c < 4; c++)
                if (nd.children[c] != -1) grow_box(nb, child_box(nd, c));
            sh_sp = top;
            if (top > 0) {
                const Frame& p = st[top - 1];
                const int c = p.next - 1;
                set_child_box(nodes[p.node], c, nb);
                nodes[p.node].children[c] = f.node;
            } else {
                for (int k = 0; k < 3; k++) { sh_box[k] = nb.lo[k]; sh_box[3 + k] = nb.hi[k]; }
            }
        }
        __syncthreads();
    }
    if (t == 0) counts[blockIdx.x] = sh_nn;
    if (t < 6) root_boxes[blockIdx.x * 6 + t] = sh_box[t];
}

// ---------------------------------------------------------------- device-wide splits of large ranges

// A split of src[start, start + n) into dst, over chunks [chunk0, chunk0 + nchunk) of the batch.
struct SegTask {
    int start, n, chunk0, nchunk;
};
// mode 0: partition; 1: n <= leafMax (no binning); 2: flat axis.  pivot: the split's result.
struct SegParam {
    int mode, axis, best, nl, s, pivot;
    float k0, k1;
};
struct BinPart {
    Box b[8];
    int n[8];
};

struct ChunkRange {
    int task, first, end;   // task index and the chunk's element range within the task
};
__device__ __forceinline__ ChunkRange chunk_range(const SegTask* tasks, const int* chunk_task) {
    ChunkRange c;
    c.task = chunk_task[blockIdx.x];
    const SegTask tk = tasks[c.task];
    c.first = ((int)blockIdx.x - tk.chunk0) * kChunk;
    c.end = min(tk.n, c.first + kChunk);
    return c;
}
// The thread's contiguous slice of a chunk.
__device__ __forceinline__ void thread_slice(const ChunkRange& c, int* a, int* b) {
    *a = min(c.end, c.first + (int)threadIdx.x * kSegPer);
    *b = min(c.end, *a + kSegPer);
}
// The thread's contiguous share of m items (the per-task folds over chunk partials).
__device__ __forceinline__ void block_share(int m, int* a, int* b) {
    const int per = (m + kSegThreads - 1) / kSegThreads;
    *a = min(m, (int)threadIdx.x * per);
    *b = min(m, *a + per);
}

__global__ __launch_bounds__(kSegThreads) void k_seg_cb(const GElem* src, const SegTask* tasks, const int* chunk_task,
                                                         int leafMax, Box* part) {
    __shared__ Box sh[kWaves];
    const ChunkRange c = chunk_range(tasks, chunk_task);
    if (tasks[c.task].n <= leafMax) return;
    const GElem* E = src + tasks[c.task].start;
    int a, b;
    thread_slice(c, &a, &b);
    Box v = box_empty();
    for (int i = a; i < b; i++) grow_point(v, E[i].c);
    v = block_fold<true>(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ __launch_bounds__(kSegThreads) void k_seg_axis(const SegTask* tasks, const Box* part, int leafMax,
                                                           SegParam* params) {
    __shared__ Box sh[kWaves];
    const SegTask tk = tasks[blockIdx.x];
    SegParam P{};
    if (tk.n <= leafMax) {
        P.mode = 1;
        P.pivot = tk.n;
        if (threadIdx.x == 0) params[blockIdx.x] = P;
        return;
    }
    int a, b;
    block_share(tk.nchunk, &a, &b);
    Box v = box_empty();
    for (int i = a; i < b; i++) join_points(v, part[tk.chunk0 + i]);
    const Box cb = block_fold<true>(v, sh);
    P.axis = longest(cb);
    if (cb.hi[P.axis] == cb.lo[P.axis]) {
        P.mode = 2;
        P.pivot = tk.n / 2 + 1;
    } else {
        P.mode = 0;
        P.k1 = 7.99952f / (cb.hi[P.axis] - cb.lo[P.axis]);
        P.k0 = cb.lo[P.axis];
    }
    if (threadIdx.x == 0) params[blockIdx.x] = P;
}

__global__ __launch_bounds__(kSegThreads) void k_seg_bins(const GElem* src, const SegTask* tasks, const int* chunk_task,
                                                           const SegParam* params, BinPart* part, int* err) {
    __shared__ Box sh[kWaves];
    __shared__ int shi[kWaves];
    const ChunkRange c = chunk_range(tasks, chunk_task);
    const SegParam P = params[c.task];
    if (P.mode != 0) return;
    const GElem* E = src + tasks[c.task].start;
    const Cut cut{P.axis, 0, P.k0, P.k1};
    int a, b;
    thread_slice(c, &a, &b);
    Bins B;
    bins_init(B);
    for (int i = a; i < b; i++) bins_add(B, E[i], cut, err);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const Box r = block_fold<false>(B.b[k], sh);
        int tot;
        (void)block_scan_excl(B.n[k], shi, &tot);
        if (threadIdx.x == 0) { part[blockIdx.x].b[k] = r; part[blockIdx.x].n[k] = tot; }
    }
}

__global__ __launch_bounds__(kSegThreads) void k_seg_sah(const GElem* src, const SegTask* tasks, const BinPart* part,
                                                          SegParam* params) {
    __shared__ Box sh[kWaves];
    __shared__ int shi[kWaves];
    const SegTask tk = tasks[blockIdx.x];
    SegParam P = params[blockIdx.x];
    if (P.mode != 0) return;
    int a, b;
    block_share(tk.nchunk, &a, &b);
    Bins B;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        Box v = box_empty();
        int m = 0;
        for (int i = a; i < b; i++) { grow_box(v, part[tk.chunk0 + i].b[k]); m += part[tk.chunk0 + i].n[k]; }
        B.b[k] = block_fold<false>(v, sh);
        (void)block_scan_excl(m, shi, &B.n[k]);
    }
    Cut cut{P.axis, 0, P.k0, P.k1};
    sah(B, &cut.best, &P.nl);
    P.best = cut.best;
    P.pivot = P.nl;
    P.s = P.nl + ((P.nl < tk.n && cut.right(src[tk.start + P.nl].c)) ? 1 : 0);
    if (threadIdx.x == 0) params[blockIdx.x] = P;
}

__global__ __launch_bounds__(kSegThreads) void k_seg_rcount(const GElem* src, const SegTask* tasks,
                                                             const int* chunk_task, const SegParam* params, int* part) {
    __shared__ int shi[kWaves];
    const ChunkRange c = chunk_range(tasks, chunk_task);
    const SegParam P = params[c.task];
    if (P.mode != 0) return;
    const GElem* E = src + tasks[c.task].start;
    const Cut cut{P.axis, P.best, P.k0, P.k1};
    int a, b;
    thread_slice(c, &a, &b);
    int cnt = 0;
    for (int i = a; i < b; i++) cnt += cut.right(E[i].c) ? 1 : 0;
    int tot;
    (void)block_scan_excl(cnt, shi, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// Exclusive scan of a task's chunk counts, in place.
__global__ __launch_bounds__(kSegThreads) void k_seg_rscan(const SegTask* tasks, const SegParam* params, int* part) {
    __shared__ int shi[kWaves];
    const SegTask tk = tasks[blockIdx.x];
    if (params[blockIdx.x].mode != 0) return;
    int a, b;
    block_share(tk.nchunk, &a, &b);
    int m = 0;
    for (int i = a; i < b; i++) m += part[tk.chunk0 + i];
    int tot;
    int run = block_scan_excl(m, shi, &tot);
    for (int i = a; i < b; i++) {
        const int x = part[tk.chunk0 + i];
        part[tk.chunk0 + i] = run;
        run += x;
    }
}

// rl / lr of the closed form, at the task's element offset.
__global__ __launch_bounds__(kSegThreads) void k_seg_rank(const GElem* src, const SegTask* tasks, const int* chunk_task,
                                                           const SegParam* params, const int* part, int* rl_all,
                                                           int* lr_all, int* err) {
    __shared__ int shi[kWaves];
    const ChunkRange c = chunk_range(tasks, chunk_task);
    const SegParam P = params[c.task];
    if (P.mode != 0) return;
    const SegTask tk = tasks[c.task];
    const GElem* E = src + tk.start;
    int* rl = rl_all + tk.start;
    int* lr = lr_all + tk.start;
    const Cut cut{P.axis, P.best, P.k0, P.k1};
    int a, b;
    thread_slice(c, &a, &b);
    int cnt = 0;
    for (int i = a; i < b; i++) cnt += cut.right(E[i].c) ? 1 : 0;
    int tot;
    int r = part[blockIdx.x] + block_scan_excl(cnt, shi, &tot);
    for (int i = a; i < b; i++) {
        const bool right = cut.right(E[i].c);
        r += right ? 1 : 0;
        if (!record_rank(i, right, r, tk.n, P.nl, rl, lr)) atomicOr(err, 1);
    }
}

__global__ __launch_bounds__(kSegThreads) void k_seg_scatter(const GElem* src, GElem* dst, const SegTask* tasks,
                                                              const int* chunk_task, const SegParam* params,
                                                              const int* part, const int* rl_all, const int* lr_all,
                                                              int* err) {
    __shared__ int shi[kWaves];
    const ChunkRange c = chunk_range(tasks, chunk_task);
    const SegParam P = params[c.task];
    const SegTask tk = tasks[c.task];
    const GElem* S = src + tk.start;
    GElem* D = dst + tk.start;
    if (P.mode != 0) {
        for (int i = c.first + (int)threadIdx.x; i < c.end; i += kSegThreads) D[i] = S[i];
        return;
    }
    const int* rl = rl_all + tk.start;
    const int* lr = lr_all + tk.start;
    const Cut cut{P.axis, P.best, P.k0, P.k1};
    int a, b;
    thread_slice(c, &a, &b);
    int cnt = 0;
    for (int i = a; i < b; i++) cnt += cut.right(S[i].c) ? 1 : 0;
    int tot;
    int r = part[blockIdx.x] + block_scan_excl(cnt, shi, &tot);
    for (int i = a; i < b; i++) {
        const GElem e = S[i];
        const bool right = cut.right(e.c);
        r += right ? 1 : 0;
        const int d = dest_of(i, right, r, tk.n, P.nl, P.s, rl, lr);
        if (d >= 0) D[d] = e;
        else atomicOr(err, 1);
    }
}

// ---------------------------------------------------------------- packing, leaves, placement

// Packs the caller's arrays; err bit 4 for a non-finite centroid coordinate (the reference's binning
// fails on it) or a NaN box coordinate (box folds in another grouping would differ on it).
__global__ void k_pack(const float* boxes, const float* cent, const int32_t* idx, int n, GElem* E, int* err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    GElem e;
    e.a = make_float4(boxes[i * 6], boxes[i * 6 + 1], boxes[i * 6 + 2], __int_as_float(idx[i]));
    e.b = make_float4(boxes[i * 6 + 3], boxes[i * 6 + 4], boxes[i * 6 + 5], 0.0f);
    e.c = make_float4(cent[i * 3], cent[i * 3 + 1], cent[i * 3 + 2], 0.0f);
    bool ok = true;
    for (int k = 0; k < 6; k++) ok = ok && !isnan(boxes[i * 6 + k]);
    for (int k = 0; k < 3; k++) ok = ok && isfinite(cent[i * 3 + k]);
    if (!ok) atomicOr(err, 4);
    E[i] = e;
}

__global__ void k_unpack(const GElem* E, int n, float* boxes, float* cent, int32_t* idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GElem e = E[i];
    boxes[i * 6] = e.a.x; boxes[i * 6 + 1] = e.a.y; boxes[i * 6 + 2] = e.a.z;
    boxes[i * 6 + 3] = e.b.x; boxes[i * 6 + 4] = e.b.y; boxes[i * 6 + 5] = e.b.z;
    cent[i * 3] = e.c.x; cent[i * 3 + 1] = e.c.y; cent[i * 3 + 2] = e.c.z;
    idx[i] = __float_as_int(e.a.w);
}

// Boxes of the leaves under large-range nodes ({start, count}), one thread each.
__global__ void k_leaf_boxes(const GElem* E, const int2* leaves, int n, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Box b = leaf_box(E + leaves[i].x, leaves[i].y);
    for (int k = 0; k < 3; k++) { out[i * 6 + k] = b.lo[k]; out[i * 6 + 3 + k] = b.hi[k]; }
}

// Moves subtree s's nodes from the scratch area (offset src[s]) to their pre-order place dst[s],
// offsetting child node indices.
__global__ void k_place(const vgpu_qbvh_node* scratch, const int* src, const int* dst, const int* counts,
                        vgpu_qbvh_node* out) {
    const int s = blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < counts[s]; i += gridDim.x * blockDim.x) {
        vgpu_qbvh_node nd = scratch[src[s] + i];
        for (int c = 0; c < 4; c++)
            if (nd.children[c] >= 0) nd.children[c] += dst[s];
        out[dst[s] + i] = nd;
    }
}

// ---------------------------------------------------------------- host side

// ---------------------------------------------------------------- pinned staging of the host arrays
//
// A pageable hipMemcpy moves the builder's host arrays (40 bytes per element each way) at about
// 21 GB/s to the device and 14 GB/s back.  Here they go through kStageSlots pinned slots of
// kStageSlot bytes: the host side of each chunk is copied by kCopyThreads threads while the DMA of
// the neighbouring chunk runs.  The slots (128 MiB of pinned host memory per device) are allocated
// by the first build of at least kStageMin bytes on that device and kept until the process ends;
// builds that run at once take turns on them.  Smaller transfers use plain hipMemcpyAsync.
constexpr size_t kStageSlot = (size_t)32 << 20, kStageMin = (size_t)4 << 20;
constexpr int kStageSlots = 4, kCopyThreads = 8, kStageDevices = 64;

struct HostSeg {
    char* host;
    char* dev;
    size_t bytes;
};

struct PinnedStage {
    std::mutex mu;
    char* buf = nullptr;
    hipEvent_t ev[kStageSlots] = {};
    bool failed = false;   // the pinned allocation failed once: plain copies from then on

    bool ready() {   // under mu, with the stage's device current
        if (buf) return true;
        if (failed) return false;
        if (hipHostMalloc((void**)&buf, kStageSlot * kStageSlots, hipHostMallocDefault) != hipSuccess) {
            buf = nullptr;
            failed = true;
            return false;
        }
        for (int i = 0; i < kStageSlots; i++)
            if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) {
                failed = true;
                return false;   // buf stays allocated and unused
            }
        return true;
    }
};

PinnedStage* stage_for_current_device() {
    static PinnedStage stages[kStageDevices];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kStageDevices) return nullptr;
    return &stages[dev];
}

// memcpy split over up to kCopyThreads threads (64-byte aligned pieces).
void par_copy(char* dst, const char* src, size_t bytes) {
    const int T = bytes >= ((size_t)1 << 22) ? kCopyThreads : 1;
    const size_t per = ((bytes + T - 1) / T + 63) & ~(size_t)63;
    std::thread th[kCopyThreads];
    for (int t = 1; t < T; t++) {
        const size_t a = std::min(bytes, per * t), b = std::min(bytes, a + per);
        if (b > a) th[t] = std::thread([=]() { std::memcpy(dst + a, src + a, b - a); });
    }
    std::memcpy(dst, src, std::min(bytes, per));
    for (int t = 1; t < T; t++)
        if (th[t].joinable()) th[t].join();
}

std::vector<HostSeg> chunks_of(const HostSeg* segs, int nseg, size_t& total) {
    std::vector<HostSeg> c;
    total = 0;
    for (int i = 0; i < nseg; i++) {
        total += segs[i].bytes;
        for (size_t off = 0; off < segs[i].bytes; off += kStageSlot)
            c.push_back(HostSeg{segs[i].host + off, segs[i].dev + off, std::min(kStageSlot, segs[i].bytes - off)});
    }
    return c;
}

// Host -> device; the copies are complete on the host side when it returns, and ordered before later
// work on st.
hipError_t stage_h2d(const HostSeg* segs, int nseg, hipStream_t st) {
    size_t total = 0;
    const std::vector<HostSeg> c = chunks_of(segs, nseg, total);
    PinnedStage* S = total >= kStageMin ? stage_for_current_device() : nullptr;
    std::unique_lock<std::mutex> lk;
    if (S) {
        lk = std::unique_lock<std::mutex>(S->mu);
        if (!S->ready()) S = nullptr;
    }
    if (!S) {
        for (int i = 0; i < nseg; i++) {
            if (segs[i].bytes == 0) continue;
            const hipError_t e = hipMemcpyAsync(segs[i].dev, segs[i].host, segs[i].bytes, hipMemcpyHostToDevice, st);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    for (size_t k = 0; k < c.size(); k++) {
        const int slot = (int)(k % kStageSlots);
        char* sb = S->buf + (size_t)slot * kStageSlot;
        // the slot's previous DMA (this call's, or an earlier call's that returned early) must be done
        // before it is refilled; an event never recorded completes at once
        const hipError_t w = hipEventSynchronize(S->ev[slot]);
        if (w != hipSuccess) return w;
        par_copy(sb, c[k].host, c[k].bytes);
        hipError_t e = hipMemcpyAsync(c[k].dev, sb, c[k].bytes, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipEventRecord(S->ev[slot], st);
        if (e != hipSuccess) return e;
    }
    // the slots are free for the next user only once their DMAs are done
    for (size_t k = c.size() > (size_t)kStageSlots ? c.size() - kStageSlots : 0; k < c.size(); k++) {
        const hipError_t e = hipEventSynchronize(S->ev[k % kStageSlots]);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Device -> host after the work already on st; the host arrays are complete when it returns.
hipError_t stage_d2h(const HostSeg* segs, int nseg, hipStream_t st) {
    size_t total = 0;
    const std::vector<HostSeg> c = chunks_of(segs, nseg, total);
    PinnedStage* S = total >= kStageMin ? stage_for_current_device() : nullptr;
    std::unique_lock<std::mutex> lk;
    if (S) {
        lk = std::unique_lock<std::mutex>(S->mu);
        if (!S->ready()) S = nullptr;
    }
    if (!S) {
        for (int i = 0; i < nseg; i++) {
            if (segs[i].bytes == 0) continue;
            const hipError_t e = hipMemcpyAsync(segs[i].host, segs[i].dev, segs[i].bytes, hipMemcpyDeviceToHost, st);
            if (e != hipSuccess) return e;
        }
        return hipStreamSynchronize(st);
    }
    // chunk k's DMA is issued before chunk k - 1 is copied out on the host; a slot is refilled only
    // after its previous chunk was copied out (kStageSlots >= 2)
    auto drain = [&](size_t k) -> hipError_t {
        const int slot = (int)(k % kStageSlots);
        const hipError_t e = hipEventSynchronize(S->ev[slot]);
        if (e != hipSuccess) return e;
        par_copy(c[k].host, S->buf + (size_t)slot * kStageSlot, c[k].bytes);
        return hipSuccess;
    };
    for (size_t k = 0; k < c.size(); k++) {
        const int slot = (int)(k % kStageSlots);
        if (k < (size_t)kStageSlots) {   // a DMA into the slot left by an earlier call that returned early
            const hipError_t w = hipEventSynchronize(S->ev[slot]);
            if (w != hipSuccess) return w;
        }
        char* sb = S->buf + (size_t)slot * kStageSlot;
        hipError_t e = hipMemcpyAsync(sb, c[k].dev, c[k].bytes, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(S->ev[slot], st);
        if (e == hipSuccess && k >= 1) e = drain(k - 1);
        if (e != hipSuccess) return e;
    }
    if (!c.empty()) return drain(c.size() - 1);
    return hipSuccess;
}


template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t n) {
        n = std::max<size_t>(n, 1);
        if (n <= cap) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; cap = 0; }
        const hipError_t e = hipMalloc((void**)&p, n * sizeof(T));
        if (e == hipSuccess) cap = n;
        return e;
    }
};

// A node above the subtrees, assembled on the host from the splits of its range.
struct TopNode {
    int start, n;
    int axis[3];
    int lo[5];      // child c covers [lo[c], lo[c + 1]) within the node's range
    int kind[4];    // 0 leaf, 1 top node, 2 subtree
    int ref[4];     // leaf, top node or subtree index
    int index;      // pre-order index
};

#define BUILD_TRY(x)                         \
    do {                                     \
        if ((x) != hipSuccess) return -2;    \
    } while (0)

// One batch of device-wide splits, src -> dst over `ranges` ({start, n}); (axis, pivot) per range
// into `out`.
struct SegRunner {
    hipStream_t st;
    int leafMax;
    DevBuf<SegTask> d_tasks;
    DevBuf<int> d_chunk_task, d_rpart;
    DevBuf<SegParam> d_params;
    DevBuf<Box> d_cbpart;
    DevBuf<BinPart> d_binpart;
    std::vector<SegTask> tasks;
    std::vector<int> chunk_task;
    std::vector<SegParam> params;

    int run(const GElem* src, GElem* dst, const std::vector<int2>& ranges, int* rl, int* lr, int* err,
            std::vector<int2>& out) {
        tasks.clear();
        chunk_task.clear();
        for (size_t i = 0; i < ranges.size(); i++) {
            const int nch = std::max(1, (ranges[i].y + kChunk - 1) / kChunk);
            tasks.push_back(SegTask{ranges[i].x, ranges[i].y, (int)chunk_task.size(), nch});
            for (int k = 0; k < nch; k++) chunk_task.push_back((int)i);
        }
        const unsigned nt = (unsigned)tasks.size(), nc = (unsigned)chunk_task.size();
        BUILD_TRY(d_tasks.alloc(nt));
        BUILD_TRY(d_params.alloc(nt));
        BUILD_TRY(d_chunk_task.alloc(nc));
        BUILD_TRY(d_rpart.alloc(nc));
        BUILD_TRY(d_cbpart.alloc(nc));
        BUILD_TRY(d_binpart.alloc(nc));
        BUILD_TRY(hipMemcpyAsync(d_tasks.p, tasks.data(), nt * sizeof(SegTask), hipMemcpyHostToDevice, st));
        BUILD_TRY(hipMemcpyAsync(d_chunk_task.p, chunk_task.data(), nc * sizeof(int), hipMemcpyHostToDevice, st));
        const dim3 blk(kSegThreads);
        hipLaunchKernelGGL(k_seg_cb, dim3(nc), blk, 0, st, src, d_tasks.p, d_chunk_task.p, leafMax, d_cbpart.p);
        hipLaunchKernelGGL(k_seg_axis, dim3(nt), blk, 0, st, d_tasks.p, d_cbpart.p, leafMax, d_params.p);
        hipLaunchKernelGGL(k_seg_bins, dim3(nc), blk, 0, st, src, d_tasks.p, d_chunk_task.p, d_params.p, d_binpart.p, err);
        hipLaunchKernelGGL(k_seg_sah, dim3(nt), blk, 0, st, src, d_tasks.p, d_binpart.p, d_params.p);
        hipLaunchKernelGGL(k_seg_rcount, dim3(nc), blk, 0, st, src, d_tasks.p, d_chunk_task.p, d_params.p, d_rpart.p);
        hipLaunchKernelGGL(k_seg_rscan, dim3(nt), blk, 0, st, d_tasks.p, d_params.p, d_rpart.p);
        hipLaunchKernelGGL(k_seg_rank, dim3(nc), blk, 0, st, src, d_tasks.p, d_chunk_task.p, d_params.p, d_rpart.p, rl,
                           lr, err);
        hipLaunchKernelGGL(k_seg_scatter, dim3(nc), blk, 0, st, src, dst, d_tasks.p, d_chunk_task.p, d_params.p,
                           d_rpart.p, rl, lr, err);
        BUILD_TRY(hipGetLastError());
        params.resize(nt);
        BUILD_TRY(hipMemcpyAsync(params.data(), d_params.p, nt * sizeof(SegParam), hipMemcpyDeviceToHost, st));
        BUILD_TRY(hipStreamSynchronize(st));
        out.resize(nt);
        for (unsigned i = 0; i < nt; i++) out[i] = make_int2(params[i].axis, params[i].pivot);
        return 0;
    }
};

}  // namespace

hipError_t stage_upload(void* dev, const void* host, size_t bytes, hipStream_t st) {
    const HostSeg seg{(char*)host, (char*)dev, bytes};
    return stage_h2d(&seg, 1, st);
}

int build_qbvh_gpu(float* boxes, float* centroids, int32_t* idx, int32_t n, int32_t leafMax,
                   vgpu_qbvh_node* nodes_out, int32_t max_nodes, int32_t* nnodes_out, float* bounds_out,
                   hipStream_t st) {
    if (n < 0 || !nnodes_out || (n > 0 && (!boxes || !centroids || !idx))) return -1;
    if (n > (1 << 28)) return -1;   // leaf codes hold 28 bits of element offset
    if (leafMax > 16) leafMax = 16;
    if (leafMax < 1) leafMax = 1;
    const size_t N = (size_t)std::max(n, 1);
    DevBuf<float> d_boxes, d_cent, d_rootbox, d_leafbox;
    DevBuf<int32_t> d_idx;
    DevBuf<GElem> d_E, d_T;
    DevBuf<int> d_rl, d_lr, d_err, d_counts, d_src, d_dst;
    DevBuf<int2> d_roots, d_leaves;
    DevBuf<vgpu_qbvh_node> d_scratch, d_nodes;
    BUILD_TRY(d_boxes.alloc(N * 6));
    BUILD_TRY(d_cent.alloc(N * 3));
    BUILD_TRY(d_idx.alloc(N));
    BUILD_TRY(d_E.alloc(N));
    BUILD_TRY(d_T.alloc(N));
    BUILD_TRY(d_err.alloc(1));
    BUILD_TRY(d_scratch.alloc(N));
    BUILD_TRY(hipMemsetAsync(d_err.p, 0, sizeof(int), st));
    if (n > 0) {
        const HostSeg in[3] = {{(char*)boxes, (char*)d_boxes.p, sizeof(float) * 6 * n},
                               {(char*)centroids, (char*)d_cent.p, sizeof(float) * 3 * n},
                               {(char*)idx, (char*)d_idx.p, sizeof(int32_t) * n}};
        BUILD_TRY(stage_h2d(in, 3, st));
        hipLaunchKernelGGL(k_pack, dim3((n + 255) / 256), dim3(256), 0, st, d_boxes.p, d_cent.p, d_idx.p, n, d_E.p,
                           d_err.p);
        BUILD_TRY(hipGetLastError());
        int bad = 0;
        BUILD_TRY(hipMemcpyAsync(&bad, d_err.p, sizeof(int), hipMemcpyDeviceToHost, st));
        BUILD_TRY(hipStreamSynchronize(st));
        if (bad) return -3;
    }
    std::vector<TopNode> top;
    std::vector<int2> subtrees, leaves;   // {start, n}
    std::vector<int> level;
    auto add_child = [&](int start, int cn, int& kind, int& ref) {
        if (cn <= leafMax) {
            kind = 0;
            ref = (int)leaves.size();
            leaves.push_back(make_int2(start, cn));
        } else if (cn <= kSubtreeMax) {
            kind = 2;
            ref = (int)subtrees.size();
            subtrees.push_back(make_int2(start, cn));
        } else {
            kind = 1;
            ref = (int)top.size();
            TopNode tn{};
            tn.start = start;
            tn.n = cn;
            top.push_back(tn);
            level.push_back(ref);
        }
    };
    // the root is a top node when its range is large; otherwise one subtree builds the whole tree
    // (a root of at most leafMax elements included: the reference still makes a node for it)
    const int root_kind = n > kSubtreeMax ? 1 : 2;
    if (root_kind == 1) {
        TopNode r{};
        r.start = 0;
        r.n = n;
        top.push_back(r);
        level.push_back(0);
        BUILD_TRY(d_rl.alloc(N));
        BUILD_TRY(d_lr.alloc(N));
    } else {
        subtrees.push_back(make_int2(0, n));
    }
    SegRunner seg;
    seg.st = st;
    seg.leafMax = leafMax;
    std::vector<int2> ranges, r0, r12;
    while (!level.empty()) {
        const std::vector<int> cur = level;
        level.clear();
        ranges.clear();
        for (int i : cur) ranges.push_back(make_int2(top[i].start, top[i].n));
        int rc = seg.run(d_E.p, d_T.p, ranges, d_rl.p, d_lr.p, d_err.p, r0);
        if (rc) return rc;
        ranges.clear();
        for (size_t k = 0; k < cur.size(); k++) {
            const TopNode& tn = top[cur[k]];
            ranges.push_back(make_int2(tn.start, r0[k].y));
            ranges.push_back(make_int2(tn.start + r0[k].y, tn.n - r0[k].y));
        }
        rc = seg.run(d_T.p, d_E.p, ranges, d_rl.p, d_lr.p, d_err.p, r12);
        if (rc) return rc;
        for (size_t k = 0; k < cur.size(); k++) {
            const int ti = cur[k];
            const int p0 = r0[k].y, p1 = r12[2 * k].y, p2 = r12[2 * k + 1].y;
            top[ti].axis[0] = r0[k].x;
            top[ti].axis[1] = r12[2 * k].x;
            top[ti].axis[2] = r12[2 * k + 1].x;
            const int lo[5] = {0, p1, p0, p0 + p2, top[ti].n};
            for (int c = 0; c < 5; c++) top[ti].lo[c] = lo[c];
            for (int c = 0; c < 4; c++) {
                // a valid split leaves both sides of the first split non-empty; a child as large as
                // its parent only comes from bins already flagged as broken
                if (lo[c + 1] - lo[c] >= top[ti].n || lo[c + 1] < lo[c]) return -3;
                int kind, ref;
                add_child(top[ti].start + lo[c], lo[c + 1] - lo[c], kind, ref);
                top[ti].kind[c] = kind;
                top[ti].ref[c] = ref;
            }
        }
    }
    const int nsub = (int)subtrees.size();
    std::vector<int> counts(std::max(nsub, 1), 0);
    std::vector<float> rootbox((size_t)std::max(nsub, 1) * 6, 0.0f);
    if (nsub > 0) {
        BUILD_TRY(d_roots.alloc(nsub));
        BUILD_TRY(d_counts.alloc(nsub));
        BUILD_TRY(d_rootbox.alloc((size_t)nsub * 6));
        BUILD_TRY(hipMemcpyAsync(d_roots.p, subtrees.data(), nsub * sizeof(int2), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_subtree, dim3(nsub), dim3(64), 0, st, d_E.p, d_T.p, d_roots.p, d_scratch.p, d_counts.p,
                           d_rootbox.p, leafMax, d_err.p);
        BUILD_TRY(hipGetLastError());
        BUILD_TRY(hipMemcpyAsync(counts.data(), d_counts.p, nsub * sizeof(int), hipMemcpyDeviceToHost, st));
        BUILD_TRY(hipMemcpyAsync(rootbox.data(), d_rootbox.p, (size_t)nsub * 6 * sizeof(float), hipMemcpyDeviceToHost,
                                 st));
    }
    const int nleaf = (int)leaves.size();
    std::vector<float> leafbox((size_t)std::max(nleaf, 1) * 6, 0.0f);
    if (nleaf > 0) {
        BUILD_TRY(d_leaves.alloc(nleaf));
        BUILD_TRY(d_leafbox.alloc((size_t)nleaf * 6));
        BUILD_TRY(hipMemcpyAsync(d_leaves.p, leaves.data(), nleaf * sizeof(int2), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_leaf_boxes, dim3((nleaf + 255) / 256), dim3(256), 0, st, d_E.p, d_leaves.p, nleaf,
                           d_leafbox.p);
        BUILD_TRY(hipGetLastError());
        BUILD_TRY(hipMemcpyAsync(leafbox.data(), d_leafbox.p, (size_t)nleaf * 6 * sizeof(float), hipMemcpyDeviceToHost,
                                 st));
    }
    int err = 0;
    BUILD_TRY(hipMemcpyAsync(&err, d_err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    BUILD_TRY(hipStreamSynchronize(st));
    if (err & 7) return -3;
    if (err) return -4;
    // pre-order numbering: a top node takes the next index, then its children in order; a subtree
    // takes a block of its node count
    std::vector<int> sub_dst(std::max(nsub, 1), 0);
    int next = 0;
    std::vector<std::pair<int, int>> stack{{root_kind, 0}};
    while (!stack.empty()) {
        const std::pair<int, int> it = stack.back();
        stack.pop_back();
        if (it.first == 2) {
            sub_dst[it.second] = next;
            next += counts[it.second];
            continue;
        }
        top[it.second].index = next++;
        for (int c = 3; c >= 0; c--)
            if (top[it.second].kind[c] != 0) stack.push_back({top[it.second].kind[c], top[it.second].ref[c]});
    }
    const int total = next;
    *nnodes_out = total;
    if (total > max_nodes || !nodes_out) return -1;
    BUILD_TRY(d_nodes.alloc(total));
    if (nsub > 0) {
        std::vector<int> src(nsub);
        for (int s = 0; s < nsub; s++) src[s] = subtrees[s].x;
        BUILD_TRY(d_src.alloc(nsub));
        BUILD_TRY(d_dst.alloc(nsub));
        BUILD_TRY(hipMemcpyAsync(d_src.p, src.data(), nsub * sizeof(int), hipMemcpyHostToDevice, st));
        BUILD_TRY(hipMemcpyAsync(d_dst.p, sub_dst.data(), nsub * sizeof(int), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_place, dim3(8, nsub), dim3(256), 0, st, d_scratch.p, d_src.p, d_dst.p, d_counts.p,
                           d_nodes.p);
        BUILD_TRY(hipGetLastError());
    }
    if (n > 0) {
        hipLaunchKernelGGL(k_unpack, dim3((n + 255) / 256), dim3(256), 0, st, d_E.p, n, d_boxes.p, d_cent.p, d_idx.p);
        BUILD_TRY(hipGetLastError());
    }
    const HostSeg out[4] = {{(char*)nodes_out, (char*)d_nodes.p, (size_t)total * sizeof(vgpu_qbvh_node)},
                            {(char*)boxes, (char*)d_boxes.p, n > 0 ? sizeof(float) * 6 * n : 0},
                            {(char*)centroids, (char*)d_cent.p, n > 0 ? sizeof(float) * 3 * n : 0},
                            {(char*)idx, (char*)d_idx.p, n > 0 ? sizeof(int32_t) * n : 0}};
    BUILD_TRY(stage_d2h(out, 4, st));
    BUILD_TRY(hipStreamSynchronize(st));
    // the top nodes, children before parents (reverse creation order), with MINSS / MAXSS unions
    const float kInf = std::numeric_limits<float>::infinity();
    auto fmin_ss = [](float x, float y) { return (x < y) ? x : y; };
    auto fmax_ss = [](float x, float y) { return (x > y) ? x : y; };
    std::vector<float> topbox(top.size() * 6);
    for (int ti = (int)top.size() - 1; ti >= 0; ti--) {
        const TopNode& tn = top[ti];
        vgpu_qbvh_node nd;
        std::memset(&nd, 0, sizeof(nd));
        nd.axis0 = (uint32_t)tn.axis[0];
        nd.axis1 = (uint32_t)tn.axis[1];
        nd.axis2 = (uint32_t)tn.axis[2];
        float nb[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
        for (int c = 0; c < 4; c++) {
            const int cn = tn.lo[c + 1] - tn.lo[c];
            const float* b;
            if (tn.kind[c] == 0) {
                if (cn == 0) {
                    nd.children[c] = -1;
                    for (int k = 0; k < 3; k++) { nd.boxes[c + k * 4] = kInf; nd.boxes[c + 12 + k * 4] = kInf; }
                    continue;
                }
                b = &leafbox[(size_t)tn.ref[c] * 6];
                const uint32_t first = (uint32_t)(tn.start + tn.lo[c]);
                nd.children[c] = (int32_t)((1u << 31) | ((first << 4) & 0xfffffff0u) | ((uint32_t)(cn - 1) & 0xfu));
            } else if (tn.kind[c] == 1) {
                b = &topbox[(size_t)tn.ref[c] * 6];
                nd.children[c] = top[tn.ref[c]].index;
            } else {
                b = &rootbox[(size_t)tn.ref[c] * 6];
                nd.children[c] = sub_dst[tn.ref[c]];
            }
            for (int k = 0; k < 3; k++) { nd.boxes[c + k * 4] = b[k]; nd.boxes[c + 12 + k * 4] = b[3 + k]; }
            for (int k = 0; k < 3; k++) { nb[k] = fmin_ss(nb[k], b[k]); nb[3 + k] = fmax_ss(nb[3 + k], b[3 + k]); }
        }
        std::memcpy(&topbox[(size_t)ti * 6], nb, sizeof(nb));
        nodes_out[tn.index] = nd;
    }
    if (bounds_out) std::memcpy(bounds_out, root_kind == 1 ? &topbox[0] : &rootbox[0], 6 * sizeof(float));
    return 0;
}

}  // namespace vgpu
