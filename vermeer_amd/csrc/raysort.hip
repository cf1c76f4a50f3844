// Ray reordering before traversal: a counting sort of the rays by bucket,
// bucket = (Morton code of the origin's cell in a 2^b x 2^b x 2^b grid over the
// scene box) x (direction octant).  Rays in one wavefront then start in the same
// region heading the same way, so they fetch the same nodes and triangles and
// the caches serve most of them.  Order inside a bucket is whatever order the
// atomics hand out; the traversal writes each hit back to its ray's original
// slot through the permutation, so results do not depend on it.
//
//   k_sort_clear    zero the bucket histogram (a kernel, not a memset, so the
//                   sequence can be captured in a graph)
//   k_sort_keys     bucket key per ray + histogram (global atomics)
//   k_scan_block    exclusive scan of 2048-bucket blocks, block totals
//   k_scan_top      exclusive scan of the block totals (one workgroup)
//   k_sort_scatter  rank within the bucket by atomic increment; scatter ray + index

#include "trace_kernels.h"

namespace vgpu {

namespace {

constexpr int SCAN_BLOCK = 2048;   // buckets per scan workgroup (256 threads x 8)

__device__ __forceinline__ uint32_t spread3(uint32_t x) {   // 7 bits -> every third bit
    x &= 0x7f;
    x = (x | (x << 8)) & 0x0000f00fu;
    x = (x | (x << 4)) & 0x000c30c3u;
    x = (x | (x << 2)) & 0x00249249u;
    return x;
}

// atomicAdd(ctr + key, 1) for the lanes with act set, lanes with equal keys sharing one atomic:
// rays next to each other in a batch often share a bucket (a pinhole camera's rays share one origin
// cell), and per-lane atomics on one counter serialise.  Returns the lane's old count plus its rank
// among the wavefront's lanes with the same key.  Every lane of the wavefront must call it.  The loop
// only groups the lanes (no memory traffic); the group leaders then issue their atomics together, so
// the wavefront waits for one round trip however many keys it holds.  After kWaveGroups groups the
// remaining lanes count alone: a wavefront of scattered keys gains nothing from grouping, and the
// grouping loop made such batches' sort 20% slower when it ran to the end.
constexpr int kWaveGroups = 8;
__device__ __forceinline__ uint32_t wave_count(uint32_t* __restrict__ ctr, uint32_t key, bool act) {
    const int lane = (int)__lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    uint64_t todo = __ballot(act);
    int my_lead = lane;          // lanes left after the loop lead themselves
    uint32_t rank = 0, cnt = 1;
    for (int it = 0; it < kWaveGroups && todo; it++) {
        const int lead = __ffsll((unsigned long long)todo) - 1;
        const uint32_t lk = __shfl(key, lead);
        const uint64_t grp = __ballot(act && key == lk) & todo;
        if ((grp >> lane) & 1ull) { my_lead = lead; rank = (uint32_t)__popcll(grp & below); }
        if (lane == lead) cnt = (uint32_t)__popcll(grp);
        todo &= ~grp;
    }
    uint32_t base = 0;
    if (act && lane == my_lead) base = atomicAdd(ctr + key, cnt);
    return __shfl(base, my_lead) + rank;
}

__global__ void k_sort_clear(uint32_t* __restrict__ hist, uint32_t nb) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) hist[i] = 0;
}

__global__ __launch_bounds__(256) void k_sort_keys(const float4* __restrict__ rays, uint64_t n,
                                                   const uint32_t* __restrict__ n_dev, float3 lo, float3 scale,
                                                   int bits, uint32_t* __restrict__ keys, uint32_t* __restrict__ hist) {
    if (n_dev) n = min(n, (uint64_t)*n_dev);
    const float cmax = (float)((1 << bits) - 1);
    // whole wavefronts step together (wave_count needs every lane): the loop bound is per wavefront
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t w0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); w0 < n; w0 += stride) {
        const uint64_t i = w0 + (threadIdx.x & 63u);
        const bool act = i < n;
        uint32_t key = 0;
        if (act) {
            const float4 a = rays[2 * i], b = rays[2 * i + 1];
            const uint32_t qx = (uint32_t)fminf(fmaxf((a.x - lo.x) * scale.x, 0.0f), cmax);
            const uint32_t qy = (uint32_t)fminf(fmaxf((a.y - lo.y) * scale.y, 0.0f), cmax);
            const uint32_t qz = (uint32_t)fminf(fmaxf((a.z - lo.z) * scale.z, 0.0f), cmax);
            const uint32_t m = spread3(qx) | (spread3(qy) << 1) | (spread3(qz) << 2);
            const uint32_t oct = (b.x < 0.0f ? 1u : 0u) | (b.y < 0.0f ? 2u : 0u) | (b.z < 0.0f ? 4u : 0u);
            key = (m << 3) | oct;
            keys[i] = key;
        }
        (void)wave_count(hist, key, act);
    }
}

__global__ __launch_bounds__(256) void k_scan_block(uint32_t* __restrict__ hist, uint32_t nb, uint32_t* __restrict__ bsum) {
    __shared__ uint32_t part[256];
    const uint32_t base = blockIdx.x * SCAN_BLOCK + threadIdx.x * 8;
    uint32_t v[8], s = 0;
    for (int k = 0; k < 8; k++) {
        v[k] = base + k < nb ? hist[base + k] : 0u;
        s += v[k];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {   // inclusive Hillis-Steele scan of the thread sums
        const uint32_t t = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (int k = 0; k < 8; k++) {
        if (base + k < nb) hist[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) bsum[blockIdx.x] = part[255];
}

__global__ __launch_bounds__(1024) void k_scan_top(uint32_t* __restrict__ bsum, uint32_t nblk) {
    __shared__ uint32_t part[1024];
    const uint32_t per = (nblk + 1023) / 1024;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t s = 0;
    for (uint32_t k = 0; k < per; k++) if (b0 + k < nblk) s += bsum[b0 + k];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t t = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (uint32_t k = 0; k < per; k++) {
        if (b0 + k < nblk) {
            const uint32_t x = bsum[b0 + k];
            bsum[b0 + k] = run;
            run += x;
        }
    }
}

__global__ __launch_bounds__(256) void k_sort_scatter(const float4* __restrict__ rays, uint64_t n,
                                                      const uint32_t* __restrict__ n_dev,
                                                      const uint32_t* __restrict__ keys, uint32_t* __restrict__ offs,
                                                      const uint32_t* __restrict__ bsum, float4* __restrict__ out,
                                                      uint32_t* __restrict__ perm) {
    if (n_dev) n = min(n, (uint64_t)*n_dev);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t w0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); w0 < n; w0 += stride) {
        const uint64_t i = w0 + (threadIdx.x & 63u);
        const bool act = i < n;
        const uint32_t key = act ? keys[i] : 0u;
        const uint32_t rank = wave_count(offs, key, act);
        if (act) {
            const uint32_t pos = rank + bsum[key / SCAN_BLOCK];
            out[2 * (uint64_t)pos] = rays[2 * i];
            out[2 * (uint64_t)pos + 1] = rays[2 * i + 1];
            perm[pos] = (uint32_t)i;
        }
    }
}

}  // namespace

uint32_t ray_sort_buckets(int bits) { return 1u << (3 * bits + 3); }

uint32_t ray_sort_scan_blocks(int bits) { return (ray_sort_buckets(bits) + SCAN_BLOCK - 1) / SCAN_BLOCK; }

hipError_t launch_ray_sort(const float* box, int bits, const void* rays, uint64_t n, const uint32_t* n_dev,
                           const RaySortBuffers& sb, int num_cus, hipStream_t st) {
    if (bits < 1 || bits > 6 || n > sb.capacity || n > 0xffffffffull) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    const uint32_t nb = ray_sort_buckets(bits), nblk = ray_sort_scan_blocks(bits);
    const float cells = (float)(1 << bits);
    float3 lo = make_float3(box[0], box[1], box[2]);
    float3 sc;
    sc.x = box[3] > box[0] ? cells / (box[3] - box[0]) : 0.0f;
    sc.y = box[4] > box[1] ? cells / (box[4] - box[1]) : 0.0f;
    sc.z = box[5] > box[2] ? cells / (box[5] - box[2]) : 0.0f;
    uint64_t g = (n + 255) / 256;
    const uint64_t gmax = (uint64_t)num_cus * 16;
    if (g > gmax) g = gmax;
    hipLaunchKernelGGL(k_sort_clear, dim3((nb + 255) / 256 < 4096 ? (nb + 255) / 256 : 4096), dim3(256), 0, st,
                       sb.hist, nb);
    hipLaunchKernelGGL(k_sort_keys, dim3((unsigned)g), dim3(256), 0, st, (const float4*)rays, n, n_dev, lo, sc, bits,
                       sb.keys, sb.hist);
    hipLaunchKernelGGL(k_scan_block, dim3(nblk), dim3(256), 0, st, sb.hist, nb, sb.bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, sb.bsum, nblk);
    hipLaunchKernelGGL(k_sort_scatter, dim3((unsigned)g), dim3(256), 0, st, (const float4*)rays, n, n_dev, sb.keys,
                       sb.hist, sb.bsum, sb.rays, sb.perm);
    return hipGetLastError();
}

}  // namespace vgpu
