// Internal interface between the render host code (render.cpp) and the
// render kernels (render.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/vermeer_gpu.h"
#include "trace_kernels.h"
#include "xform.h"

namespace vgpu {

// Per post-sort geom slot: where its shading data lives.
struct alignas(16) DevGeomShade {
    int32_t nrm_base;      // first float4 of its per-triangle vertex normals (3 per triangle), -1: none
    int32_t sidx_base;     // first byte of its per-triangle shader index, -1: none (index 0)
    int32_t shader_base;   // first entry of its shader table (material ids)
    int32_t nshaders;
};

// A light node: Tri (builtin/light/triangle.go:18-28), Disk (disk.go:17-31) or Sphere (sphere.go:16-28).
struct alignas(16) DevLight {
    float4 P0, P1, P2;         // tri: vertices; disk: P (w = radius), Up, LookAt; sphere: centre (w = radius)
    int32_t material;
    int32_t samples;           // NumSamples = 1 << samples
    int32_t geom;              // post-sort geom slot of the light's own geometry
    int32_t type;              // VGPU_LIGHT_TRI / _DISK / _SPHERE
};

// Per pixel and recursion level, what the level's shading left for the final combine.
struct alignas(16) LevelRec {
    float4 a;                  // xyz: (0 + emission) + diffuse; w: kind (0 miss, 1 rough/no spec, 2 mirror ray traced, 3 mirror not traced)
    float4 s;                  // kind 1: specular direct lighting; kind 2: mirror factor rho.ToRGB() * Spec1Colour; w: spec weight
    float4 c;                  // path mode: continuation factor (Eval / pdf).ToRGB() * lobe colour; w: 0 none traced,
                               // kPathDelta (perfect mirror) or kPathSmooth (Oren-Nayar or GGX sample)
};
constexpr float kPathDelta = 2.0f;
constexpr float kPathSmooth = 4.0f;

struct RenderView {
    SceneView sv;
    const float4* normals;         // per-triangle vertex normals
    const uint8_t* shaderidx;      // per-triangle shader index
    const int32_t* shaders;        // shader tables (material ids)
    const DevGeomShade* gshade;
    const vgpu_material* materials;
    const DevLight* lights;
    int32_t nlights;
    int32_t width, height;
    int32_t slots;                 // shadow-ray slots per pixel
    int32_t groups;                // (light, lobe) groups per pixel = 2 * nlights
    int32_t max_half;              // max over lights of NSamples/2 (slots per group = 2 * max_half)
    float M[16];                   // camera matrix (column-major)
    float tan_theta_focal, aspect, focal;
    float lens_radius;             // thin lens (camera Radius); 0: pinhole
    const float* cam_keys;         // camera motion keys: cam_nkeys records of kCamKeyFloats (T, R, S)
    int32_t cam_nkeys;             // 0: the matrix M for every sample
    uint64_t raster_fwd[52];       // RasterXY enumeration tables for m = 12
    uint64_t raster_inv[24];
    const double* filt;            // pixel filter sampler: cdfV (filt_n), then cdfVU (filt_n x filt_n); null: box
    int32_t filt_n;
    int32_t max_level;             // deepest recursion level shaded (0 without mirrors, else 3; path mode: depth - 1)
    int32_t path_depth;            // 0: the reference's shading; D > 0: path mode with D segments per path
    unsigned long long* counts;    // rays traced since creation: [0] closest hit (camera, mirror, continuation), [1] shadow
    int32_t rec_levels;            // LevelRec entries per pixel
    double filt_w;
};

// A camera motion key: TransformDecomp (math/animdecomp.go:13-17) as T[3], R (X, Y, Z, W), S[16], one pad.
constexpr int kCamKeyFloats = kXfKeyFloats;
constexpr int kMaxCamKeys = 64;

constexpr int kMaxLevel = 3;       // ShaderStd.Eval returns at once above level 3 (std.go:97-99)
constexpr int kMaxPathDepth = 16;  // path mode: segments per path

// Launch wrappers (all enqueue on `stream`).  Levels: rays_in holds the level's rays (P, tmax; D,
// time), rays_out receives the next level's mirror rays (inactive pixels get tmax -1).
hipError_t launch_camera(const RenderView& rv, const uint64_t* scrambles, uint32_t iter, float4* rays, float4* pix,
                         hipStream_t stream);
hipError_t launch_shade(const RenderView& rv, const uint64_t* scrambles, uint32_t iter, int level, const float4* rays_in,
                        const int4* hits32, const float4* pix, float4* slot_rays, float4* slot_rgb, float4* groups,
                        float4* pixshade, float4* rays_out, LevelRec* recs, hipStream_t stream);
// film null: record the level for launch_combine; otherwise (no mirrors) update the film directly.
hipError_t launch_resolve(const RenderView& rv, int level, uint32_t iter, const float4* slot_rgb, const int4* slot_hits,
                          const float4* groups, const float4* pixshade, LevelRec* recs, float* film, float* sample_out,
                          hipStream_t stream);
hipError_t launch_combine(const RenderView& rv, uint32_t iter, const LevelRec* recs, float* film, float* sample_out,
                          hipStream_t stream);

}  // namespace vgpu
