// Path-guiding network kernels for gfx950: bf16 matrix-core (MFMA 16x16x32)
// inference and training with fp32 accumulation.
//
// Layout.  The network is evaluated "transposed": for each layer Y = W X,
// the weights are the A operand (rows = outputs) and the activations of 16
// samples are the B operand (columns = samples).  A 16x16 fp32 result tile
// leaves lane l holding rows 4(l/16)..4(l/16)+3 of sample column l%16; two
// such tiles (features 32j..32j+31) are exactly the 8 values per lane that the
// next layer's B operand needs for its k-chunk j, provided the k-order of the
// next layer's weight fragment is permuted to match (hid_feature below).  So
// activations stay in registers from layer to layer, with no transpose.
// Weight fragments are packed once per parameter update (k_pack) and staged
// into LDS by every workgroup (50 KB forward, +32 KB transposed for backward).
//
// Training (one step over B samples, two launches): k_train runs forward, the loss gradient and
// the backward pass in registers, with each sample's loss weighted by its own c; after each
// backward layer it stages that layer's input X and output gradient dY for its 256 samples in LDS
// and forms dW = dY X^T on the matrix cores (samples as the k dimension) into one fp32 partial per
// workgroup, beside the workgroup's sum of c.  k_adam reduces the partials and the c sums in a
// fixed order, divides by the batch's c sum (the loss is sum c_i (-log q_i) / sum c_i), applies
// Adam and writes each updated weight's bf16 copy straight into its fragment slots.

#include <atomic>

#include "guide_kernels.h"

namespace vgpu {
namespace guide {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int F_L1 = 0, F_L2 = 20, F_L3 = 28, F_L4 = 36, F_L5 = 44;
constexpr int FB_L5 = 0, FB_L4 = 8, FB_L3 = 16, FB_L2 = 24;
static_assert(F_L5 + 6 == NF_FWD, "forward fragment count");
static_assert(FB_L2 + 8 == NF_BWD, "backward fragment count");

// Hidden feature held in k-slot i of lane group g for k-chunk j (the result-tile order).
__device__ __forceinline__ int hid_feature(int j, int g, int i) {
    return i < 4 ? 32 * j + 4 * g + i : 32 * j + 16 + 4 * g + (i - 4);
}
// Output-layer rows are permuted so that lane group g holds lobes 2g and 2g+1 in its
// slots s = 4T + r (T = result tile, r = row in the lane's quad); slots 10, 11 are empty.
__device__ __forceinline__ int out5_index(int g, int s) { return s < 10 ? 5 * (2 * g + s / 5) + s % 5 : -1; }

__device__ __forceinline__ uint16_t bf16_bits(__bf16 b) { return __builtin_bit_cast(uint16_t, b); }
__device__ __forceinline__ float bf16_val(__bf16 b) { return (float)b; }

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 frag(const u32x4* s, int f, int lane) {
    return __builtin_bit_cast(bf16x8, s[f * 64 + lane]);
}

// ---------------------------------------------------------------- packing

// Parameter held in element i of lane `lane` of forward fragment f (-1: a zero pad).  Every weight
// of the five layers is held exactly once.
__device__ int pack_fwd_index(int f, int lane, int i) {
    const int r = lane & 15, g = lane >> 4;
    if (f < F_L2) {
        const int o = f / 5, j = f % 5, q = 32 * j + 8 * g + i;
        return q < NENC ? OW1 + (16 * o + r) * NENC + q : -1;
    }
    if (f < F_L5) {
        const int l = (f - F_L2) / 8, ff = (f - F_L2) % 8, o = ff / 2, j = ff % 2;
        const int ow = l == 0 ? OW2 : (l == 1 ? OW3 : OW4);
        return ow + (16 * o + r) * NH + hid_feature(j, g, i);
    }
    const int ff = f - F_L5, T = ff / 2, j = ff % 2;
    const int out = out5_index(r >> 2, 4 * T + (r & 3));
    return out >= 0 ? OW5 + out * NH + hid_feature(j, g, i) : -1;
}

// The same for the transposed (backward) fragments of layers 5..2; each of their weights once.
__device__ int pack_bwd_index(int f, int lane, int i) {
    const int r = lane & 15, g = lane >> 4;
    const int o = (f % 8) / 2, j = f % 2, row = 16 * o + r;
    if (f < FB_L4) {
        const int out = out5_index(g, 8 * j + i);
        return out >= 0 ? OW5 + out * NH + row : -1;
    }
    const int l = (f - FB_L4) / 8;
    const int ow = l == 0 ? OW4 : (l == 1 ? OW3 : OW2);
    return ow + hid_feature(j, g, i) * NH + row;
}

__device__ float pack_fwd(const float* P, int f, int lane, int i) {
    const int k = pack_fwd_index(f, lane, i);
    return k >= 0 ? P[k] : 0.0f;
}
__device__ float pack_bwd(const float* P, int f, int lane, int i) {
    const int k = pack_bwd_index(f, lane, i);
    return k >= 0 ? P[k] : 0.0f;
}

// Inverse of the packing: slot_fwd[p] / slot_bwd[p] = the bf16 element (index into the fragment
// array as uint16) that holds parameter p, or -1 (biases; layer 1 has no backward fragment).
// Both arrays are -1 on entry.
__global__ void k_slot_map(int32_t* __restrict__ slot_fwd, int32_t* __restrict__ slot_bwd) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (NF_FWD + NF_BWD) * 64 * 8) return;
    const int e = t & 7, fl = t >> 3;
    const bool fwd = fl < NF_FWD * 64;
    const int idx = fwd ? fl : fl - NF_FWD * 64;
    const int k = fwd ? pack_fwd_index(idx >> 6, idx & 63, e) : pack_bwd_index(idx >> 6, idx & 63, e);
    if (k >= 0) (fwd ? slot_fwd : slot_bwd)[k] = idx * 8 + e;
}

__global__ void k_pack(const float* __restrict__ P, u32x4* __restrict__ ffwd, u32x4* __restrict__ fbwd) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (NF_FWD + NF_BWD) * 64) return;
    const bool fwd = t < NF_FWD * 64;
    const int idx = fwd ? t : t - NF_FWD * 64;
    const int f = idx >> 6, lane = idx & 63;
    uint32_t h[8];
    for (int i = 0; i < 8; i++) {
        const float v = fwd ? pack_fwd(P, f, lane, i) : pack_bwd(P, f, lane, i);
        h[i] = bf16_bits((__bf16)v);
    }
    u32x4 w;
    w.x = h[0] | (h[1] << 16); w.y = h[2] | (h[3] << 16); w.z = h[4] | (h[5] << 16); w.w = h[6] | (h[7] << 16);
    (fwd ? ffwd : fbwd)[idx] = w;
}

// ---------------------------------------------------------------- forward pieces

__device__ __forceinline__ float sel6(const float x[6], int d) {
    return d == 0 ? x[0] : d == 1 ? x[1] : d == 2 ? x[2] : d == 3 ? x[3] : d == 4 ? x[4] : x[5];
}

// Encoded features 32j + 8g .. +7 (canonical order: 24 d + 2 f (+1 for cos)).
// sin(pi 2^f x) is evaluated on the exactly reduced argument r = 2^f x - 2 round(2^(f-1) x),
// in revolutions (r / 2) for the hardware sine.
__device__ __forceinline__ bf16x8 enc_chunk(const float x[6], int j, int g) {
    bf16x8 b;
    const int q0 = 32 * j + 8 * g;
    if (q0 >= NENC) {
        for (int i = 0; i < 8; i++) b[i] = (__bf16)0.0f;
        return b;
    }
    const int d = q0 / 24, f0 = (q0 % 24) / 2;
    const float xd = sel6(x, d);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float t = ldexpf(xd, f0 + k);
        const float r = t - 2.0f * rintf(0.5f * t);
        const float rev = 0.5f * r;
        b[2 * k] = (__bf16)__builtin_amdgcn_sinf(rev);
        b[2 * k + 1] = (__bf16)__builtin_amdgcn_cosf(rev);
    }
    return b;
}

__device__ __forceinline__ void normalise(const float4& p, const float4& n, const float* bounds, float x[6]) {
    const float P[3] = {p.x, p.y, p.z};
    for (int k = 0; k < 3; k++) {
        const float lo = bounds[k], hi = bounds[3 + k];
        const float c = 0.5f * (lo + hi), h = 0.5f * (hi - lo);
        x[k] = (P[k] - c) / h;
    }
    x[3] = n.x; x[4] = n.y; x[5] = n.z;
}

template <int NT>
__device__ __forceinline__ void bias_hidden(f32x4 (&acc)[4][NT], const float* b, int g) {
#pragma unroll
    for (int o = 0; o < 4; o++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float v = b[16 * o + 4 * g + r];
#pragma unroll
            for (int t = 0; t < NT; t++) acc[o][t][r] = v;
        }
}

template <int NT>
__device__ __forceinline__ void relu_pack(const f32x4 (&acc)[4][NT], bf16x8 (&X)[2][NT]) {
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const float v = acc[2 * j + (i >> 2)][t][i & 3];
                X[j][t][i] = (__bf16)(v > 0.0f ? v : 0.0f);
            }
}

template <int NT>
__device__ __forceinline__ void layer_hidden(const u32x4* sF, int fbase, const float* bias, int g, int lane,
                                             const bf16x8 (&Xin)[2][NT], f32x4 (&acc)[4][NT]) {
    bias_hidden<NT>(acc, bias, g);
#pragma unroll
    for (int o = 0; o < 4; o++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const bf16x8 a = frag(sF, fbase + 2 * o + j, lane);
#pragma unroll
            for (int t = 0; t < NT; t++) acc[o][t] = mfma(a, Xin[j][t], acc[o][t]);
        }
}

template <int NT>
__device__ __forceinline__ void layer_first(const u32x4* sF, const float* bias, int g, int lane, const float (&x)[NT][6],
                                            f32x4 (&acc)[4][NT]) {
    bias_hidden<NT>(acc, bias, g);
#pragma unroll
    for (int j = 0; j < 5; j++) {
        bf16x8 b[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) b[t] = enc_chunk(x[t], j, g);
#pragma unroll
        for (int o = 0; o < 4; o++) {
            const bf16x8 a = frag(sF, F_L1 + 5 * o + j, lane);
#pragma unroll
            for (int t = 0; t < NT; t++) acc[o][t] = mfma(a, b[t], acc[o][t]);
        }
    }
}

template <int NT>
__device__ __forceinline__ void layer_out(const u32x4* sF, const float* b5, int g, int lane, const bf16x8 (&Xin)[2][NT],
                                          f32x4 (&acc)[3][NT]) {
#pragma unroll
    for (int T = 0; T < 3; T++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int out = out5_index(g, 4 * T + r);
            const float v = out >= 0 ? b5[out] : 0.0f;
#pragma unroll
            for (int t = 0; t < NT; t++) acc[T][t][r] = v;
        }
#pragma unroll
    for (int T = 0; T < 3; T++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const bf16x8 a = frag(sF, F_L5 + 2 * T + j, lane);
#pragma unroll
            for (int t = 0; t < NT; t++) acc[T][t] = mfma(a, Xin[j][t], acc[T][t]);
        }
}

// ---------------------------------------------------------------- lobes

struct LobeD {
    float mu[3], kappa, logit, ulen;
    bool kin;
};

__device__ __forceinline__ void decode_lobe(const float* v, LobeD& L) {
    const float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    L.ulen = len;
    if (len > 0.0f) {
        const float inv = 1.0f / len;
        L.mu[0] = v[0] * inv; L.mu[1] = v[1] * inv; L.mu[2] = v[2] * inv;
    } else {
        L.mu[0] = 0.0f; L.mu[1] = 0.0f; L.mu[2] = 1.0f;
    }
    L.kin = v[3] > KAPPA_RAW_MIN && v[3] < KAPPA_RAW_MAX;
    L.kappa = expf(fminf(fmaxf(v[3], KAPPA_RAW_MIN), KAPPA_RAW_MAX));
    L.logit = v[4];
}

__device__ __forceinline__ float quad_max(float v) {
    v = fmaxf(v, __shfl_xor(v, 16));
    return fmaxf(v, __shfl_xor(v, 32));
}
__device__ __forceinline__ float quad_sum(float v) {
    v += __shfl_xor(v, 16);
    return v + __shfl_xor(v, 32);
}

// decode_lobe, log_vmf_norm and dlog_norm for the training kernel's loss gradient, on the hardware
// exp2 / log2 / reciprocal / square-root approximations (about 1 ulp; the gradient is compared
// with the FP64 oracle at a relative tolerance far above that).
__device__ __forceinline__ void decode_lobe_fast(const float* v, LobeD& L) {
    const float len = __builtin_amdgcn_sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    L.ulen = len;
    if (len > 0.0f) {
        const float inv = __builtin_amdgcn_rcpf(len);
        L.mu[0] = v[0] * inv; L.mu[1] = v[1] * inv; L.mu[2] = v[2] * inv;
    } else {
        L.mu[0] = 0.0f; L.mu[1] = 0.0f; L.mu[2] = 1.0f;
    }
    L.kin = v[3] > KAPPA_RAW_MIN && v[3] < KAPPA_RAW_MAX;
    L.kappa = __expf(fminf(fmaxf(v[3], KAPPA_RAW_MIN), KAPPA_RAW_MAX));
    L.logit = v[4];
}
__device__ __forceinline__ float log_vmf_norm_fast(float k) {
    return __logf(k) - 1.8378770664093453f - __logf(-expm1f(-2.0f * k));
}
__device__ __forceinline__ float dlog_norm_fast(float k) {
    if (k < 0.05f) return 1.0f - k * (1.0f / 3.0f) + k * k * k * (1.0f / 45.0f);
    return __builtin_amdgcn_rcpf(k) - 2.0f * __builtin_amdgcn_rcpf(expm1f(2.0f * k));
}
// log of the vMF normalisation kappa / (2 pi (1 - exp(-2 kappa)))
__device__ __forceinline__ float log_vmf_norm(float k) {
    return logf(k) - 1.8378770664093453f - logf(-expm1f(-2.0f * k));
}

// 1/kappa - 2/(exp(2 kappa) - 1), with its series below 0.05
__device__ __forceinline__ float dlog_norm(float k) {
    if (k < 0.05f) return 1.0f - k * (1.0f / 3.0f) + k * k * k * (1.0f / 45.0f);
    return 1.0f / k - 2.0f / expm1f(2.0f * k);
}

// ---------------------------------------------------------------- inference

template <bool RAW>
__global__ __launch_bounds__(256) void k_infer(Net net, const float4* __restrict__ q, uint64_t n,
                                               const uint32_t* __restrict__ n_dev, float* __restrict__ out) {
    constexpr int NT = 2;   // 16-query tiles per wave and iteration: 2 keeps three waves per SIMD
    __shared__ u32x4 sF[NF_FWD * 64];
    for (int k = threadIdx.x; k < NF_FWD * 64; k += 256) sF[k] = reinterpret_cast<const u32x4*>(net.frag_fwd)[k];
    __syncthreads();
    uint64_t nn = n;
    if (n_dev) nn = min(nn, (uint64_t)*n_dev);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, c = lane & 15;
    const uint64_t nw = (uint64_t)gridDim.x * 4;
    const float* P = net.params;
    for (uint64_t base = ((uint64_t)blockIdx.x * 4 + wave) * (16 * NT); base < nn; base += nw * (16 * NT)) {
        // keeps the weight fragments' LDS reads inside the loop: hoisted, they would pin 200 registers
        __asm__ volatile("" ::: "memory");
        float x[NT][6];
        uint64_t s[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            s[t] = base + 16 * t + c;
            if (s[t] < nn) {
                normalise(q[2 * s[t]], q[2 * s[t] + 1], net.bounds, x[t]);
            } else {
                for (int k = 0; k < 6; k++) x[t][k] = 0.0f;
            }
        }
        f32x4 acc[4][NT];
        bf16x8 X[2][NT];
        layer_first<NT>(sF, P + OB1, g, lane, x, acc);
        relu_pack<NT>(acc, X);
        layer_hidden<NT>(sF, F_L2, P + OB2, g, lane, X, acc);
        relu_pack<NT>(acc, X);
        layer_hidden<NT>(sF, F_L3, P + OB3, g, lane, X, acc);
        relu_pack<NT>(acc, X);
        layer_hidden<NT>(sF, F_L4, P + OB4, g, lane, X, acc);
        relu_pack<NT>(acc, X);
        f32x4 o5[3][NT];
        layer_out<NT>(sF, P + OB5, g, lane, X, o5);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            float v[12];
#pragma unroll
            for (int T = 0; T < 3; T++)
#pragma unroll
                for (int r = 0; r < 4; r++) v[4 * T + r] = o5[T][t][r];
            if (RAW) {
                if (s[t] < nn)
                    for (int sl = 0; sl < 10; sl++) out[s[t] * NOUT + out5_index(g, sl)] = v[sl];
            } else {
                LobeD A, B;
                decode_lobe(v, A);
                decode_lobe(v + 5, B);
                const float m = quad_max(fmaxf(A.logit, B.logit));
                const float ea = expf(A.logit - m), eb = expf(B.logit - m);
                const float z = quad_sum(ea + eb);
                if (s[t] < nn) {
                    float* o = out + s[t] * NOUT + 10 * g;
                    o[0] = A.mu[0]; o[1] = A.mu[1]; o[2] = A.mu[2]; o[3] = A.kappa; o[4] = ea / z;
                    o[5] = B.mu[0]; o[6] = B.mu[1]; o[7] = B.mu[2]; o[8] = B.kappa; o[9] = eb / z;
                }
            }
        }
    }
}

// ---------------------------------------------------------------- training

__device__ __forceinline__ uint32_t lowbias32(uint32_t x) {   // C. Wellons' integer hash
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

// Batch slot i of a training step (SampleSource in guide_kernels.h).
__device__ __forceinline__ TrainSample fetch_sample(const SampleSource& src, uint32_t i) {
    if (!src.pick_count) return src.S[i];
    uint32_t n = *src.pick_count;
    if (n > src.pick_cap) n = src.pick_cap;
    TrainSample t{};
    if (n > 0) t = src.S[lowbias32(i ^ lowbias32(src.seed)) % n];
    return t;
}

// Sum of v[0 .. n): lane l adds v[l], v[l + 64], ... and the 64 lane sums are added pairwise across
// the wavefront, the same order in every wavefront that calls it.
__device__ __forceinline__ float wave_sum_fixed(const float* v, uint32_t n, int lane) {
    float a = 0.0f;
    for (uint32_t i = (uint32_t)lane; i < n; i += 64) a += v[i];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) a += __shfl_xor(a, m);
    return a;
}

// Masks the back-propagated tile with the layer input (ReLU derivative) and repacks it as the next
// B operand.
template <int NT>
__device__ __forceinline__ void mask_pack(const f32x4 (&acc)[4][NT], const bf16x8 (&Xf)[2][NT], bf16x8 (&D)[2][NT]) {
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const float v = acc[2 * j + (i >> 2)][t][i & 3];
                D[j][t][i] = (__bf16)(bf16_val(Xf[j][t][i]) > 0.0f ? v : 0.0f);
            }
}

template <int NT>
__device__ __forceinline__ void layer_back(const u32x4* sB, int fbase, int lane, const bf16x8 (&Din)[2][NT],
                                           f32x4 (&acc)[4][NT]) {
#pragma unroll
    for (int o = 0; o < 4; o++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[o][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int o = 0; o < 4; o++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const bf16x8 a = frag(sB, fbase + 2 * o + j, lane);
#pragma unroll
            for (int t = 0; t < NT; t++) acc[o][t] = mfma(a, Din[j][t], acc[o][t]);
        }
}

// ---- weight gradients inside the training kernel
//
// A workgroup trains TRAIN_WG samples as TRAIN_CHUNKS chunks of TRAIN_CHUNK (8 waves x 16 samples).
// After each backward layer it stages that layer's input X and output gradient dY for the chunk in
// LDS, feature-major with rows of STG bf16 (the 8-element pad makes the 16-byte operand reads of 16
// consecutive rows conflict-free), and adds dY X^T to its dW tiles on the matrix cores with the
// samples as the k dimension.  The tiles of every layer stay in the waves' accumulators over the
// chunks; the workgroup then writes one fp32 partial of every parameter, which k_adam adds in a fixed
// order.  X and dY never reach HBM.
constexpr int TRAIN_CHUNK = 128;
constexpr int TRAIN_CHUNKS = TRAIN_WG / TRAIN_CHUNK;
static_assert(TRAIN_CHUNKS * TRAIN_CHUNK == TRAIN_WG, "chunks tile the workgroup's samples");
constexpr int STG = TRAIN_CHUNK + 8;           // staged row stride, bf16 (272 bytes)
constexpr size_t FWD_BYTES = (size_t)NF_FWD * 64 * 16, BWD_BYTES = (size_t)NF_BWD * 64 * 16;
constexpr size_t STAGE_BYTES = (size_t)(NENC + NH) * STG * 2;   // X1 (144 rows) and dY1 (64 rows): the largest
constexpr size_t SAMPLE_BYTES = (size_t)TRAIN_WG * sizeof(TrainSample);   // the workgroup's batch slots
constexpr size_t TRAIN_LDS = FWD_BYTES + BWD_BYTES + STAGE_BYTES + SAMPLE_BYTES;
static_assert(STAGE_BYTES % 16 == 0 && TRAIN_LDS <= 160 * 1024, "LDS layout");

// Tiles of one layer's dW held by a wave: tiles wave, wave + 8, ... of NOR x NIT.
template <int NOR, int NIT>
struct DwTiles {
    static constexpr int N = (NOR * NIT + 7) / 8;
    f32x4 acc[N];
};

// Writes this wave's 16 sample columns of a hidden-layer tensor (result-tile order, hid_feature).
__device__ __forceinline__ void stage_hidden(uint16_t* rows, const bf16x8 (&V)[2][1], int g, int col) {
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
        for (int i = 0; i < 8; i++) rows[hid_feature(j, g, i) * STG + col] = bf16_bits(V[j][0][i]);
}

// Adds the staged chunk's dY X^T to the wave's tiles of dW_l and the chunk's dY row sums to bsum
// (8 threads per row, 16 samples each).
template <int NOR, int NIT>
__device__ __forceinline__ void dw_chunk(const uint16_t* Xs, const uint16_t* Ys, DwTiles<NOR, NIT>& T, float& bsum,
                                         int nout) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, c = lane & 15;
#pragma unroll
    for (int q = 0; q < DwTiles<NOR, NIT>::N; q++) {
        const int tile = wave + 8 * q;
        if (tile < NOR * NIT) {
            const int o = tile / NIT, it = tile % NIT;
            const uint16_t* a0 = Ys + (16 * o + c) * STG + 8 * g;
            const uint16_t* b0 = Xs + (16 * it + c) * STG + 8 * g;
#pragma unroll
            for (int k0 = 0; k0 < TRAIN_CHUNK; k0 += 32)
                T.acc[q] = mfma(__builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(a0 + k0)),
                                __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(b0 + k0)), T.acc[q]);
        }
    }
    const int row = threadIdx.x >> 3, part = threadIdx.x & 7;
    if (row < nout) {
        const uint16_t* y = Ys + row * STG + 16 * part;
        for (int k = 0; k < 16; k++) bsum += bf16_val(__builtin_bit_cast(__bf16, y[k]));
    }
}

// The wave's tiles and the bias sums of dW_l into the workgroup's partial Pp.
template <int NOR, int NIT>
__device__ __forceinline__ void dw_write(const DwTiles<NOR, NIT>& T, float bsum, float* Pp, int ow, int ob, int nin,
                                         int nout) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, c = lane & 15;
#pragma unroll
    for (int q = 0; q < DwTiles<NOR, NIT>::N; q++) {
        const int tile = wave + 8 * q;
        if (tile < NOR * NIT) {
            const int o = tile / NIT, it = tile % NIT;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int out = 16 * o + 4 * g + r;
                if (out < nout) Pp[ow + out * nin + 16 * it + c] = T.acc[q][r];
            }
        }
    }
    float b = bsum;
    b += __shfl_xor(b, 1);
    b += __shfl_xor(b, 2);
    b += __shfl_xor(b, 4);
    const int row = threadIdx.x >> 3;
    if (row < nout && (threadIdx.x & 7) == 0) Pp[ob + row] = b;
}

template <int NOR, int NIT>
__device__ __forceinline__ void dw_zero(DwTiles<NOR, NIT>& T) {
#pragma unroll
    for (int q = 0; q < DwTiles<NOR, NIT>::N; q++) T.acc[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
}

// One training step's forward, loss gradient, backward and weight-gradient partials for TRAIN_WG
// samples per workgroup; the workgroup's sum of c goes to tb.csum[1 + workgroup], and workgroup 0
// copies the step count to tb.step[1] for k_adam.
__global__ __launch_bounds__(512) void k_train(Net net, SampleSource src, TrainBuffers tb) {
    constexpr int NT = 1;
    extern __shared__ u32x4 lds[];
    u32x4* sF = lds;
    u32x4* sB = lds + FWD_BYTES / 16;
    uint16_t* Xs = reinterpret_cast<uint16_t*>(lds + (FWD_BYTES + BWD_BYTES) / 16);
    uint16_t* Ys = Xs + NENC * STG;
    TrainSample* sS = reinterpret_cast<TrainSample*>(lds + (FWD_BYTES + BWD_BYTES + STAGE_BYTES) / 16);
    {
        // the batch slots and both fragment sets (adjacent in LDS), every load issued before the
        // first store so that their latencies overlap
        constexpr int NFRAG = (NF_FWD + NF_BWD) * 64, NI = (NFRAG + 511) / 512;
        TrainSample ts{};
        if (threadIdx.x < TRAIN_WG) ts = fetch_sample(src, blockIdx.x * TRAIN_WG + threadIdx.x);
        u32x4 fr[NI];
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const int k = (int)threadIdx.x + 512 * i;
            fr[i] = u32x4{0u, 0u, 0u, 0u};
            if (k < NF_FWD * 64) fr[i] = reinterpret_cast<const u32x4*>(net.frag_fwd)[k];
            else if (k < NFRAG) fr[i] = reinterpret_cast<const u32x4*>(net.frag_bwd)[k - NF_FWD * 64];
        }
        if (threadIdx.x < TRAIN_WG) sS[threadIdx.x] = ts;
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const int k = (int)threadIdx.x + 512 * i;
            if (k < NFRAG) lds[k] = fr[i];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, c = lane & 15;
    const int col = wave * 16 + c;                          // this lane's staged column
    const float* P = net.params;
    if (wave == 0) {
        float a = 0.0f;
#pragma unroll
        for (int i = 0; i < TRAIN_WG / 64; i++) a += sS[lane + 64 * i].c;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) a += __shfl_xor(a, m);
        if (lane == 0) tb.csum[1 + blockIdx.x] = a;
        if (blockIdx.x == 0 && lane == 0) tb.step[1] = tb.step[0];   // the count k_adam reads
    }

    DwTiles<4, 9> T1;
    DwTiles<4, 4> T2, T3, T4;
    DwTiles<3, 4> T5;
    dw_zero(T1); dw_zero(T2); dw_zero(T3); dw_zero(T4); dw_zero(T5);
    float bs1 = 0.0f, bs2 = 0.0f, bs3 = 0.0f, bs4 = 0.0f, bs5 = 0.0f;

    for (int chunk = 0; chunk < TRAIN_CHUNKS; chunk++) {
        // keeps the weight fragments' LDS reads inside the loop (hoisted, they would pin ~330 registers)
        __asm__ volatile("" ::: "memory");
        const uint64_t s0 = (uint64_t)blockIdx.x * TRAIN_WG + (uint64_t)chunk * TRAIN_CHUNK + wave * 16 + c;
        float x[NT][6], dir[3], cw;
        {
            const TrainSample ts = sS[chunk * TRAIN_CHUNK + wave * 16 + c];
            normalise(make_float4(ts.px, ts.py, ts.pz, 0.0f), make_float4(ts.nx, ts.ny, ts.nz, 0.0f), net.bounds, x[0]);
            dir[0] = ts.dx; dir[1] = ts.dy; dir[2] = ts.dz;
            cw = ts.c;
        }
        f32x4 acc[4][NT];
        bf16x8 X2[2][NT], X3[2][NT], X4[2][NT], X5[2][NT];
        layer_first<NT>(sF, P + OB1, g, lane, x, acc);
        relu_pack<NT>(acc, X2);
        layer_hidden<NT>(sF, F_L2, P + OB2, g, lane, X2, acc);
        relu_pack<NT>(acc, X3);
        layer_hidden<NT>(sF, F_L3, P + OB3, g, lane, X3, acc);
        relu_pack<NT>(acc, X4);
        layer_hidden<NT>(sF, F_L4, P + OB4, g, lane, X4, acc);
        relu_pack<NT>(acc, X5);
        f32x4 o5[3][NT];
        layer_out<NT>(sF, P + OB5, g, lane, X5, o5);

        // loss gradient d(-(c/cbar) log q)/d raw for this lane's two lobes, staged as dY5 rows (out5_index)
        bf16x8 G5[2][NT];
        {
            float v[12];
#pragma unroll
            for (int T = 0; T < 3; T++)
#pragma unroll
                for (int r = 0; r < 4; r++) v[4 * T + r] = o5[T][0][r];
            LobeD L[2];
            decode_lobe_fast(v, L[0]);
            decode_lobe_fast(v + 5, L[1]);
            const float mlog = quad_max(fmaxf(L[0].logit, L[1].logit));
            const float e0 = __expf(L[0].logit - mlog), e1 = __expf(L[1].logit - mlog);
            const float z = quad_sum(e0 + e1);
            const float iz = __builtin_amdgcn_rcpf(z);
            const float w[2] = {e0 * iz, e1 * iz};
            float a[2], cosv[2];
            for (int k = 0; k < 2; k++) {
                cosv[k] = L[k].mu[0] * dir[0] + L[k].mu[1] * dir[1] + L[k].mu[2] * dir[2];
                a[k] = __logf(w[k]) + log_vmf_norm_fast(L[k].kappa) + L[k].kappa * (cosv[k] - 1.0f);
            }
            const float am = quad_max(fmaxf(a[0], a[1]));
            const float ex0 = __expf(a[0] - am), ex1 = __expf(a[1] - am);
            const float se = quad_sum(ex0 + ex1);
            const float logq = am + __logf(se);
            const float ise = __builtin_amdgcn_rcpf(se);
            const float rr[2] = {ex0 * ise, ex1 * ise};
            const float sc = cw;
            float gr[12];
            for (int k = 0; k < 2; k++) {
                const float kap = L[k].kappa;
                gr[5 * k + 4] = sc * (w[k] - rr[k]);
                const float dl = dlog_norm_fast(kap) + (cosv[k] - 1.0f);
                gr[5 * k + 3] = L[k].kin ? -sc * rr[k] * dl * kap : 0.0f;
                const float iul = L[k].ulen > 0.0f ? __builtin_amdgcn_rcpf(L[k].ulen) : 0.0f;
                for (int jj = 0; jj < 3; jj++)
                    gr[5 * k + jj] = -sc * rr[k] * kap * (dir[jj] - cosv[k] * L[k].mu[jj]) * iul;
            }
            gr[10] = 0.0f; gr[11] = 0.0f;
            if (g == 0) tb.loss[s0] = -sc * logq;
#pragma unroll
            for (int i = 0; i < 8; i++) G5[0][0][i] = (__bf16)gr[i];
#pragma unroll
            for (int i = 0; i < 8; i++) G5[1][0][i] = (__bf16)(i < 4 ? gr[8 + i] : 0.0f);
        }

        // backward: dX_l = W_l^T dY_l, dY_{l-1} = dX_l * (X_l > 0); dW_l from the staged X_l and dY_l
        __syncthreads();   // the previous chunk's layer-1 tiles are done with the staging area
        for (int sl = 0; sl < 10; sl++)
            Ys[out5_index(g, sl) * STG + col] = bf16_bits(sl < 8 ? G5[0][0][sl] : G5[1][0][sl - 8]);
        stage_hidden(Xs, X5, g, col);
        bf16x8 D[2][NT];
        layer_back<NT>(sB, FB_L5, lane, G5, acc);
        mask_pack<NT>(acc, X5, D);
        __syncthreads();
        dw_chunk(Xs, Ys, T5, bs5, NOUT);
        __syncthreads();
        stage_hidden(Xs, X4, g, col);
        stage_hidden(Ys, D, g, col);
        layer_back<NT>(sB, FB_L4, lane, D, acc);
        mask_pack<NT>(acc, X4, D);
        __syncthreads();
        dw_chunk(Xs, Ys, T4, bs4, NH);
        __syncthreads();
        stage_hidden(Xs, X3, g, col);
        stage_hidden(Ys, D, g, col);
        layer_back<NT>(sB, FB_L3, lane, D, acc);
        mask_pack<NT>(acc, X3, D);
        __syncthreads();
        dw_chunk(Xs, Ys, T3, bs3, NH);
        __syncthreads();
        stage_hidden(Xs, X2, g, col);
        stage_hidden(Ys, D, g, col);
        layer_back<NT>(sB, FB_L2, lane, D, acc);
        mask_pack<NT>(acc, X2, D);
        __syncthreads();
        dw_chunk(Xs, Ys, T2, bs2, NH);
        __syncthreads();
        // layer 1: its input is the encoding, recomputed (the forward pass keeps it in no register)
#pragma unroll
        for (int j = 0; j < 5; j++) {
            const int q0 = 32 * j + 8 * g;
            if (q0 < NENC) {
                const bf16x8 b = enc_chunk(x[0], j, g);
#pragma unroll
                for (int i = 0; i < 8; i++) Xs[(q0 + i) * STG + col] = bf16_bits(b[i]);
            }
        }
        stage_hidden(Ys, D, g, col);
        __syncthreads();
        dw_chunk(Xs, Ys, T1, bs1, NH);
    }
    float* Pp = tb.partial + (uint64_t)blockIdx.x * NPARAM;
    dw_write(T1, bs1, Pp, OW1, OB1, NENC, NH);
    dw_write(T2, bs2, Pp, OW2, OB2, NH, NH);
    dw_write(T3, bs3, Pp, OW3, OB3, NH, NH);
    dw_write(T4, bs4, Pp, OW4, OB4, NH, NH);
    dw_write(T5, bs5, Pp, OW5, OB5, NH, NOUT);
}

// k_adam: 64 parameters per workgroup as ADAM_COLS float4 columns.  Each of ADAM_RUNS groups of
// ADAM_COLS threads adds a contiguous run of the workgroup partials (16-byte loads, a wavefront
// reading four 256-byte rows), and the run sums are then added in run order, so the reduction order
// is fixed.
constexpr int ADAM_COLS = 16, ADAM_RUNS = 16, ADAM_THREADS = ADAM_COLS * ADAM_RUNS;
constexpr int ADAM_PARAMS = 4 * ADAM_COLS;
static_assert(NPARAM % 4 == 0, "partial rows are whole float4 columns");

// Adds the weight-gradient partials and the workgroup c sums (nwg of each), divides by the batch's
// c sum, applies Adam and writes the updated parameter's bf16 copy into its forward and backward
// fragment slots (so no separate repacking pass).  Every workgroup reads the step count from
// step[1], which k_train copied from step[0] before this launch, so workgroup 0 can advance step[0]
// without a race (and without device-wide atomics or fences); it also stores the c sum in csum[0].
__global__ __launch_bounds__(ADAM_THREADS) void k_adam(
        float* __restrict__ P, float* __restrict__ m, float* __restrict__ v, float* __restrict__ grad,
        const float* __restrict__ partial, uint32_t nwg, float lr, float* __restrict__ csum,
        int32_t* __restrict__ step, int update, const int32_t* __restrict__ slot_fwd,
        const int32_t* __restrict__ slot_bwd, uint16_t* __restrict__ ffwd, uint16_t* __restrict__ fbwd) {
    __shared__ float4 run_sum[ADAM_RUNS][ADAM_COLS];
    __shared__ float s_csum;
    const int colw = threadIdx.x % ADAM_COLS, run = threadIdx.x / ADAM_COLS;
    const int q4 = blockIdx.x * ADAM_COLS + colw;
    const uint32_t per = (nwg + ADAM_RUNS - 1) / ADAM_RUNS;
    const uint32_t s0 = min(nwg, (uint32_t)run * per), s1 = min(nwg, s0 + per);
    // the Adam operands, the step count and the c sum are loaded first, so that their latency
    // overlaps the partial reads
    const int p = blockIdx.x * ADAM_PARAMS + (int)threadIdx.x;
    const bool owner = threadIdx.x < ADAM_PARAMS && p < NPARAM;
    float m0 = 0.0f, v0 = 0.0f, p0 = 0.0f;
    int sf = -1, sb = -1;
    if (owner) { m0 = m[p]; v0 = v[p]; p0 = P[p]; sf = slot_fwd[p]; sb = slot_bwd[p]; }
    const int t = step[1] + 1;
    if (threadIdx.x < 64) {
        const float a = wave_sum_fixed(csum + 1, nwg, (int)threadIdx.x);
        if (threadIdx.x == 0) s_csum = a;
    }
    float4 gs = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (q4 < NPARAM / 4) {
        const float4* p4 = reinterpret_cast<const float4*>(partial) + q4;
        uint32_t s = s0;
        for (; s + 16 <= s1; s += 16) {   // 16 loads in flight, then the adds in order
            float4 t[16];
#pragma unroll
            for (int j = 0; j < 16; j++) t[j] = p4[(uint64_t)(s + j) * (NPARAM / 4)];
#pragma unroll
            for (int j = 0; j < 16; j++) { gs.x += t[j].x; gs.y += t[j].y; gs.z += t[j].z; gs.w += t[j].w; }
        }
        for (; s < s1; s++) {
            const float4 t = p4[(uint64_t)s * (NPARAM / 4)];
            gs.x += t.x; gs.y += t.y; gs.z += t.z; gs.w += t.w;
        }
    }
    run_sum[run][colw] = gs;
    __syncthreads();
    const float cs = s_csum;
    const bool upd = update && cs > 0.0f;
    if (owner) {
        const float* rs = reinterpret_cast<const float*>(&run_sum[0][0]) + threadIdx.x;
        float gsum = rs[0];
#pragma unroll
        for (int r = 1; r < ADAM_RUNS; r++) gsum += rs[r * ADAM_PARAMS];
        const float gr = cs > 0.0f ? gsum / cs : 0.0f;
        grad[p] = gr;
        if (upd) {
            const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
            const float mm = b1 * m0 + (1.0f - b1) * gr;
            const float vv = b2 * v0 + (1.0f - b2) * gr * gr;
            m[p] = mm;
            v[p] = vv;
            const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
            const float np = p0 - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
            P[p] = np;
            const uint16_t h = bf16_bits((__bf16)np);
            if (sf >= 0) ffwd[sf] = h;
            if (sb >= 0) fbwd[sb] = h;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (upd) step[0] = t;
        csum[0] = cs;
    }
}


}  // namespace

hipError_t launch_slot_map(int32_t* slot_fwd, int32_t* slot_bwd, hipStream_t st) {
    const int total = (NF_FWD + NF_BWD) * 64 * 8;
    hipLaunchKernelGGL(k_slot_map, dim3((total + 255) / 256), dim3(256), 0, st, slot_fwd, slot_bwd);
    return hipGetLastError();
}

hipError_t launch_pack(const float* params, uint4* ffwd, uint4* fbwd, hipStream_t st) {
    const int total = (NF_FWD + NF_BWD) * 64;
    hipLaunchKernelGGL(k_pack, dim3((total + 255) / 256), dim3(256), 0, st, params, (u32x4*)ffwd, (u32x4*)fbwd);
    return hipGetLastError();
}

hipError_t launch_infer(const Net& net, const float4* queries, uint64_t n, const uint32_t* n_dev, float* out, bool raw,
                        int num_cus, hipStream_t st) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    const uint64_t cap = (uint64_t)num_cus * 3;
    if (blocks > cap) blocks = cap;
    if (raw)
        hipLaunchKernelGGL(k_infer<true>, dim3((unsigned)blocks), dim3(256), 0, st, net, queries, n, n_dev, out);
    else
        hipLaunchKernelGGL(k_infer<false>, dim3((unsigned)blocks), dim3(256), 0, st, net, queries, n, n_dev, out);
    return hipGetLastError();
}

constexpr int kMaxDevices = 64;


hipError_t launch_train(const Net& net, float* params, float* m, float* v, const SampleSource& src, uint32_t B,
                        float lr, bool update, const TrainBuffers& tb, hipStream_t st) {
    if (B == 0 || B % KSPLIT != 0 || B > tb.max_batch) return hipErrorInvalidValue;
    // k_train's dynamic LDS (staging plus the backward fragments) exceeds the default limit: raise it once
    // per device
    static std::atomic<bool> attr_set[kMaxDevices];
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    if (!attr_set[dev].load(std::memory_order_acquire)) {
        e = hipFuncSetAttribute((const void*)k_train, hipFuncAttributeMaxDynamicSharedMemorySize, (int)TRAIN_LDS);
        if (e != hipSuccess) return e;
        attr_set[dev].store(true, std::memory_order_release);
    }
    if (!src.S) return hipErrorInvalidValue;
    const uint32_t nwg = B / TRAIN_WG;
    hipLaunchKernelGGL(k_train, dim3(nwg), dim3(512), TRAIN_LDS, st, net, src, tb);
    hipLaunchKernelGGL(k_adam, dim3((NPARAM + ADAM_PARAMS - 1) / ADAM_PARAMS), dim3(ADAM_THREADS), 0, st, params, m, v,
                       tb.grad, tb.partial, nwg, lr, tb.csum, tb.step, update ? 1 : 0, (const int32_t*)tb.slot_fwd,
                       (const int32_t*)tb.slot_bwd, (uint16_t*)net.frag_fwd, (uint16_t*)net.frag_bwd);
    return hipGetLastError();
}

}  // namespace guide
}  // namespace vgpu
