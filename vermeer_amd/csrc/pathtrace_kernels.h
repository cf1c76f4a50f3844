// Internal interface of the path-mode kernels (pathtrace.hip).
//
// Path mode has no reference counterpart: the reference shades one camera hit
// with direct light and mirror recursion (core/render.go, builtin/shader/std.go).
// Path mode traces full paths up to max_depth segments with next-event
// estimation and one-sample MIS between the cosine-weighted BSDF sample and
// the guiding network's vMF mixture.  Its definition (the oracle in
// oracle/oracle_pt.cpp restates it):
//   random numbers  u(pixel, pass, dim) from the lowbias32 integer hash
//   camera          the reference's pinhole (builtin/camera/camera.go:244-284),
//                   pixel position x + u0, y + u1
//   surfaces        Lambertian albedo rho, emission Le (one- or two-sided),
//                   flat shading normal = geometric normal facing the ray
//   lights          every emissive triangle; NEE picks one with probability
//                   proportional to area x luminance(Le), then a uniform point
//   mixture         p(w) = (1 - alpha) cos/pi + alpha q(w); alpha = 0: plain
//   MIS             balance heuristic between NEE and the mixture sample
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "guide_kernels.h"
#include "render_kernels.h"
#include "trace_kernels.h"

namespace vgpu {

struct alignas(16) PTMat {
    float4 a;   // albedo xyz, w: 1 = one-sided emitter
    float4 e;   // emission xyz, w: NEE area density of this material's emitters (lum(Le) / sum A lum)
};

struct alignas(16) PTLight {
    float4 v0;  // w: area
    float4 e1;  // v1 - v0, w: cumulative selection probability
    float4 e2;  // v2 - v0, w: 1 = one-sided
    float4 n;   // unit geometric normal, w: unused
    float4 Le;  // emission, w: area density (selection probability / area)
};

// Per-vertex training record (6 float4): see k_pt_gather in pathtrace.hip.
constexpr int PT_REC_F4 = 6;
constexpr int PT_DIMS_PER_DEPTH = 8;

struct PTView {
    SceneView sv;
    const DevGeomShade* gshade;
    const uint8_t* shaderidx;
    const int32_t* shaders;
    const PTMat* mats;
    const PTLight* lights;
    int32_t nlights;
    int32_t width, height, max_depth;
    float M[16];
    float tan_theta_focal, aspect, focal;
    uint32_t seed_lo, seed_hi;
    float alpha;            // probability of sampling the guide (0: plain)
    uint32_t rec_stride;    // record paths of pixels with hash % stride == pass % stride (0: none)
};

// Per-pass device state (all arrays sized for width*height paths).
struct PTBuffers {
    float4* qray[2];        // ray queues (2 float4 per path)
    float4* qstate[2];      // beta xyz, pdf of the sampled direction; path meta in qmeta
    int4* qmeta[2];         // pixel, depth, 0, 0
    uint32_t* count;        // [0], [1]: queue lengths, [2]: shadow queue length, [3]: training samples
    unsigned long long* rays;  // rays traced (closest + shadow), accumulated
    int4* hits;             // 32-byte closest-hit records (2 int4 per path)
    float4* shade;          // per queue slot: P, alive | N, 0 | albedo, 0  (3 float4)
    float4* query;          // guide queries (2 float4 per slot)
    float* lobes;           // guide output, 40 floats per slot
    float4* sray;           // shadow rays (2 float4)
    float4* scontrib;       // beta * E xyz, pixel (as float bits) | E xyz, record slot (as float bits)
    int4* shits;            // shadow hits (16 bytes)
    float4* lpass;          // this pass's radiance per pixel
    float4* rec;            // training records [pixel][depth][PT_REC_F4]
    int32_t* plen;          // recorded path length per pixel (-1: not recorded this pass)
    guide::TrainSample* tsamples;  // gathered training samples
    uint32_t tcap;          // capacity of tsamples
};

hipError_t pt_launch_camera(const PTView& v, uint32_t pass, const PTBuffers& b, hipStream_t st);
hipError_t pt_launch_hit(const PTView& v, uint32_t pass, int cur, const PTBuffers& b, hipStream_t st);
hipError_t pt_launch_scatter(const PTView& v, uint32_t pass, int cur, const PTBuffers& b, hipStream_t st);
hipError_t pt_launch_shadow(const PTView& v, const PTBuffers& b, hipStream_t st);
hipError_t pt_launch_accumulate(const PTView& v, const PTBuffers& b, double* accum, hipStream_t st);
hipError_t pt_launch_gather(const PTView& v, uint32_t pass, const PTBuffers& b, hipStream_t st);

}  // namespace vgpu
