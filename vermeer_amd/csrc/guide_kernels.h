// Internal interface of the path-guiding network kernels (guide.hip).
//
// Network: 144 encoded inputs -> 4 x 64 hidden (ReLU) -> 40 outputs = 8 von
// Mises-Fisher lobes (u[3], kappa_raw, logit).  Canonical parameter order:
// W1 (64x144, row = output), b1, W2, b2, W3, b3, W4, b4, W5 (40x64), b5.
// Weights are multiplied in bf16 on the matrix cores with fp32 accumulation;
// the fp32 master copy, the Adam moments and the biases stay in fp32.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace vgpu {
namespace guide {

constexpr int NENC = 144, NH = 64, NOUT = 40, NLOBE = 8;
constexpr int OW1 = 0, OB1 = OW1 + NH * NENC, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NH, OW4 = OB3 + NH, OB4 = OW4 + NH * NH, OW5 = OB4 + NH, OB5 = OW5 + NOUT * NH,
              NPARAM = OB5 + NOUT;
static_assert(NPARAM == 24360, "parameter count");

// bf16 operand fragments of the 16x16x32 MFMA, 64 lanes x 16 bytes each.
constexpr int NF_FWD = 4 * 5 + 3 * 4 * 2 + 3 * 2;   // W1 (4 out tiles x 5 k-chunks), W2..W4, W5 (3 x 2)
constexpr int NF_BWD = 4 * 2 * 4;                   // W5^T, W4^T, W3^T, W2^T (4 in tiles x 2 k-chunks)
constexpr int KSPLIT = 1024;                        // batch granularity (a batch is a multiple of it)
// Samples per training workgroup: each writes one fp32 partial of every parameter's gradient.
constexpr int TRAIN_WG = 256;
static_assert(KSPLIT % TRAIN_WG == 0, "workgroups tile the batch");
constexpr float KAPPA_RAW_MIN = -5.0f, KAPPA_RAW_MAX = 7.0f;

// One training sample: hit position + c, shading normal, sampled direction (unit).
struct alignas(16) TrainSample {
    float px, py, pz, c;
    float nx, ny, nz, pad0;
    float dx, dy, dz, pad1;
};
static_assert(sizeof(TrainSample) == 48, "TrainSample is 48 bytes");

struct Net {
    const float* params;       // fp32 master (NPARAM)
    const uint4* frag_fwd;     // NF_FWD x 64
    const uint4* frag_bwd;     // NF_BWD x 64
    float bounds[6];           // scene box used to normalise positions
};

struct TrainBuffers {
    float* partial;     // [batch / TRAIN_WG][NPARAM]: each training workgroup's gradient sums
    float* loss;        // per sample: c_i (-log q_i)
    float* csum;        // [0]: sum of c over the last batch; [1 .. 1 + max_batch/TRAIN_WG): per-workgroup sums
    int32_t* step;      // [0]: Adam step count; [1]: its copy made by k_train, which k_adam reads
    float* grad;        // last mean gradient (NPARAM)
    int32_t* slot_fwd;  // NPARAM: bf16 element of the forward fragments holding each parameter, -1: none
    int32_t* slot_bwd;  // NPARAM: the same for the backward fragments
    uint32_t max_batch; // largest batch the buffers hold
};

// Rebuilds the bf16 fragments from the fp32 master copy.
hipError_t launch_pack(const float* params, uint4* frag_fwd, uint4* frag_bwd, hipStream_t st);
// Fills the parameter -> fragment slot maps (both arrays set to -1 beforehand).
hipError_t launch_slot_map(int32_t* slot_fwd, int32_t* slot_bwd, hipStream_t st);

// Inference on n queries (2 float4 each: position, normal).  When n_dev is not null the
// count is min(n, *n_dev).  raw = true writes the 40 raw outputs in canonical order,
// otherwise 8 decoded lobes of (mu[3], kappa, weight).
hipError_t launch_infer(const Net& net, const float4* queries, uint64_t n, const uint32_t* n_dev, float* out, bool raw,
                        int num_cus, hipStream_t st);

// Where a training step reads its batch: S[i] for slot i < B; or, when pick_count is set, a
// random draw from the first n = min(*pick_count, pick_cap) records of S: slot i takes record
// lowbias32(i ^ lowbias32(seed)) mod n (no records: an all-zero sample, which trains nothing).
struct SampleSource {
    const TrainSample* S;
    const uint32_t* pick_count;   // device; null: S is the batch itself
    uint32_t pick_cap;
    uint32_t seed;
};

// One training step on a batch of B samples (B a multiple of KSPLIT).  update = false computes
// the mean gradient only.
hipError_t launch_train(const Net& net, float* params, float* adam_m, float* adam_v, const SampleSource& src,
                        uint32_t B, float lr, bool update, const TrainBuffers& tb, hipStream_t st);

}  // namespace guide
}  // namespace vgpu
