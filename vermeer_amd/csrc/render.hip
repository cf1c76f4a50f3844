// Wavefront renderer kernels for gfx950: the reference's per-pixel render /
// shade loop (reference semantics: direct lighting with MIS, no indirect
// bounces) split into passes over all pixels:
//
//   k_camera   RasterXY / VdC sample generation and the pinhole camera ray
//              (core/render.go:75-125, builtin/camera/camera.go:221-319)
//   (trace)    closest hit with barycentric weights (trace.hip, EXT records)
//   k_shade    hit attributes (builtin/geom/polymesh/trace.go:280-513),
//              ShaderStd.Eval (builtin/shader/std.go:77-296) with Oren-Nayar and
//              GGX lobes (Dielectric or Conductor Fresnel), TriLight ValidSample / SampleArea
//              (builtin/light/triangle.go:136-343) and the MIS weights of
//              EvaluateLightSamples (core/shader.go:203-402).  Every sample that
//              would trace a shadow ray gets a slot: the shadow ray and the
//              colour it adds if unoccluded.
//   (trace)    shadow any-hit over all slots (trace.hip)
//   k_resolve  sums the unoccluded slot colours in the reference's order,
//              applies the per-light / per-lobe scales and updates the film as
//              fb = (fb*iter + c)/(iter + 1) (core/render.go:127-129).
//
// Compiled with -ffp-contract=off.  Float64 transcendentals (the reference
// calls Go's float64 math package and rounds to float32) use the device's
// double-precision library.  Vec3Normalize keeps the reference's Newton-Raphson
// step (math/vec3_amd64.s:11-42) on the IEEE estimate 1/sqrt(x).

#include "render_kernels.h"
#include "xform.h"

#include <cstdint>

namespace vgpu {

namespace {

// ------------------------------------------------------------- tables

// CIE 1931 2-degree standard observer, 5 nm steps from 360 nm (95 samples): public CIE data, with
// the 715-830 nm tail as the reference transcribes it (6 significant digits; 715 nm reads 0.004106457).
__constant__ float kCieX[95] = {
    0.0001299f, 0.0002321f, 0.0004149f, 0.0007416f, 0.001368f, 0.002236f, 0.004243f, 0.00765f,
    0.01431f, 0.02319f, 0.04351f, 0.07763f, 0.13438f, 0.21477f, 0.2839f, 0.3285f,
    0.34828f, 0.34806f, 0.3362f, 0.3187f, 0.2908f, 0.2511f, 0.19536f, 0.1421f,
    0.09564f, 0.05795001f, 0.03201f, 0.0147f, 0.0049f, 0.0024f, 0.0093f, 0.0291f,
    0.06327f, 0.1096f, 0.1655f, 0.2257499f, 0.2904f, 0.3597f, 0.4334499f, 0.5120501f,
    0.5945f, 0.6784f, 0.7621f, 0.8425f, 0.9163f, 0.9786f, 1.0263f, 1.0567f,
    1.0622f, 1.0456f, 1.0026f, 0.9384f, 0.8544499f, 0.7514f, 0.6424f, 0.5419f,
    0.4479f, 0.3608f, 0.2835f, 0.2187f, 0.1649f, 0.1212f, 0.0874f, 0.0636f,
    0.04677f, 0.0329f, 0.0227f, 0.01584f, 0.01135916f, 0.008110916f, 0.005790346f, 0.004106457f,
    0.002899327f, 0.00204919f, 0.001439971f, 0.000999949f, 0.000690079f, 0.000476021f, 0.000332301f, 0.000234826f,
    0.000166151f, 0.000117413f, 8.30753E-05f, 5.87065E-05f, 4.15099E-05f, 2.93533E-05f, 2.06738E-05f, 1.45598E-05f,
    1.0254E-05f, 7.22146E-06f, 5.08587E-06f, 3.58165E-06f, 2.52253E-06f, 1.77651E-06f, 1.25114E-06f};
__constant__ float kCieY[95] = {
    0.000003917f, 0.000006965f, 0.00001239f, 0.00002202f, 0.000039f, 0.000064f, 0.00012f, 0.000217f,
    0.000396f, 0.00064f, 0.00121f, 0.00218f, 0.004f, 0.0073f, 0.0116f, 0.01684f,
    0.023f, 0.0298f, 0.038f, 0.048f, 0.06f, 0.0739f, 0.09098f, 0.1126f,
    0.13902f, 0.1693f, 0.20802f, 0.2586f, 0.323f, 0.4073f, 0.503f, 0.6082f,
    0.71f, 0.7932f, 0.862f, 0.9148501f, 0.954f, 0.9803f, 0.9949501f, 1.0f,
    0.995f, 0.9786f, 0.952f, 0.9154f, 0.87f, 0.8163f, 0.757f, 0.6949f,
    0.631f, 0.5668f, 0.503f, 0.4412f, 0.381f, 0.321f, 0.265f, 0.217f,
    0.175f, 0.1382f, 0.107f, 0.0816f, 0.061f, 0.04458f, 0.032f, 0.0232f,
    0.017f, 0.01192f, 0.00821f, 0.005723f, 0.004102f, 0.002929f, 0.002091f, 0.001484f,
    0.001047f, 0.00074f, 0.00052f, 0.0003611f, 0.0002492f, 0.0001719f, 0.00012f, 0.0000848f,
    0.00006f, 0.0000424f, 0.00003f, 0.0000212f, 0.00001499f, 0.0000106f, 7.4657E-06f, 5.2578E-06f,
    3.7029E-06f, 2.6078E-06f, 1.8366E-06f, 1.2934E-06f, 9.1093E-07f, 6.4153E-07f, 4.5181E-07f};
__constant__ float kCieZ[95] = {
    0.0006061f, 0.001086f, 0.001946f, 0.003486f, 0.006450001f, 0.01054999f, 0.02005001f, 0.03621f, 0.06785001f, 0.1102f,
    0.2074f, 0.3713f, 0.6456f, 1.0390501f, 1.3856f, 1.62296f, 1.74706f, 1.7826f, 1.77211f, 1.7441f,
    1.6692f, 1.5281f, 1.28764f, 1.0419f, 0.8129501f, 0.6162f, 0.46518f, 0.3533f, 0.272f, 0.2123f,
    0.1582f, 0.1117f, 0.07824999f, 0.05725001f, 0.04216f, 0.02984f, 0.0203f, 0.0134f, 0.008749999f, 0.005749999f,
    0.0039f, 0.002749999f, 0.0021f, 0.0018f, 0.001650001f, 0.0014f, 0.0011f, 0.001f, 0.0008f, 0.0006f,
    0.00034f, 0.00024f, 0.00019f, 0.0001f, 0.00005f, 0.00003f, 0.00002f, 0.00001f, 0, 0,
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

// Smits 1999 RGB-to-spectrum basis, 10 bins on [380,720): white, cyan, magenta, yellow, red, green, blue.
__constant__ float kSmits[7][10] = {
    {1, 1, 0.999f, 0.9993f, 0.9992f, 0.9998f, 1, 1, 1, 1},
    {0.9710f, 0.9426f, 1.0007f, 1.0007f, 1.0007f, 0.1564f, 0, 0, 0, 0},
    {1, 1, 0.9685f, 0.2229f, 0, 0.0458f, 0.8369f, 1, 1, 0.9959f},
    {0.0001f, 0, 0.1088f, 0.6651f, 1, 1, 0.9996f, 0.9586f, 0.9685f, 0.9840f},
    {0.1012f, 0.0515f, 0, 0, 0, 0, 0.8325f, 1.0149f, 1.0149f, 1.0149f},
    {0, 0, 0.0273f, 0.7937f, 1, 0.9418f, 0.1719f, 0, 0, 0.0025f},
    {1, 1, 0.8916f, 0.3323f, 0, 0, 0.0003f, 0.0369f, 0.0483f, 0.0496f}};
enum { SM_WHITE, SM_CYAN, SM_MAGENTA, SM_YELLOW, SM_RED, SM_GREEN, SM_BLUE };

// ------------------------------------------------------------- vectors

struct V3 { float x, y, z; };
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 scale(float s, V3 a) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float fabs32(float x) { return __int_as_float(__float_as_int(x) & 0x7fffffff); }
__device__ __forceinline__ float dotabs(V3 a, V3 b) { return fabs32(a.x * b.x + a.y * b.y + a.z * b.z); }
__device__ __forceinline__ float len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
__device__ __forceinline__ V3 mad(V3 a, V3 b, float s) { return v3(a.x + (b.x * s), a.y + (b.y * s), a.z + (b.z * s)); }
__device__ __forceinline__ float maxs(float x, float y) { return (x > y) ? x : y; }
__device__ __forceinline__ float mins(float x, float y) { return (x < y) ? x : y; }
__device__ __forceinline__ V3 from4(float4 a) { return v3(a.x, a.y, a.z); }
// math/vec3_amd64.s:11-42: squared length (y2 + x2) + z2, reciprocal-square-root estimate, one
// Newton-Raphson step.  The reference's estimate is RSQRTSS, whose bits differ between CPU vendors;
// here it is the IEEE value 1/sqrt(x), so the result is reproducible on any machine.
__device__ __forceinline__ V3 normalize(V3 a) {
    const float x2 = a.x * a.x, y2 = a.y * a.y, z2 = a.z * a.z;
    const float l2 = (y2 + x2) + z2;
    const float g = 1.0f / sqrtf(l2);
    float gg = g * g;
    gg = gg * (l2 * 0.5f);
    const float r = (1.5f - gg) * g;
    return v3(a.x * r, a.y * r, a.z * r);
}
// math/vec3_amd64.s:45-56
__device__ __forceinline__ float length(V3 a) {
    const float x2 = a.x * a.x, y2 = a.y * a.y, z2 = a.z * a.z;
    return sqrtf((y2 + x2) + z2);
}
__device__ __forceinline__ V3 basis_project(V3 U, V3 V, V3 W, V3 S) { return v3(dot(U, S), dot(V, S), dot(W, S)); }
__device__ __forceinline__ V3 basis_expand(V3 U, V3 V, V3 W, V3 S) {
    return v3(U.x * S.x + V.x * S.y + W.x * S.z, U.y * S.x + V.y * S.y + W.y * S.z, U.z * S.x + V.z * S.y + W.z * S.z);
}
__constant__ float kPi32 = 3.14159265358f;
#define PI32 3.14159274101257324f   // float32(3.14159265358)
__device__ __forceinline__ float sqr32(float x) { return x * x; }
__device__ __forceinline__ float fsin(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float fcos(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float ftan(float x) { return (float)tan((double)x); }
__device__ __forceinline__ float facos(float x) { return (float)acos((double)x); }
__device__ __forceinline__ float fatan(float x) { return (float)atan((double)x); }
__device__ __forceinline__ float fatan2(float y, float x) { return (float)atan2((double)y, (double)x); }

// math/ferror.go:27-67 (v == -0.0 in Go compares with 0; the uint32 arithmetic wraps)
__device__ __forceinline__ float next_up(float v) {
    if (__builtin_isinf(v) && v > 0) return v;
    if (v == 0.0f) v = 0.0f;
    unsigned u = __float_as_uint(v);
    if (v >= 0.0f) u++; else u--;
    return __uint_as_float(u);
}
__device__ __forceinline__ float next_down(float v) {
    if (__builtin_isinf(v) && v < 0) return v;
    if (v == 0.0f) v = 0.0f;
    unsigned u = __float_as_uint(v);
    if (v >= 0.0f) u--; else u++;
    return __uint_as_float(u);
}

// ------------------------------------------------------------- colour

struct RGB { float r, g, b; };
__device__ __forceinline__ RGB rgb0() { return RGB{0.0f, 0.0f, 0.0f}; }
__device__ __forceinline__ void rgb_add(RGB& a, RGB b) { a.r += b.r; a.g += b.g; a.b += b.b; }
__device__ __forceinline__ void rgb_mul(RGB& a, RGB b) { a.r *= b.r; a.g *= b.g; a.b *= b.b; }
__device__ __forceinline__ void rgb_scale(RGB& a, float s) { a.r *= s; a.g *= s; a.b *= s; }

__device__ __forceinline__ float smits_eval(int s, float lambda) {
    if (lambda < 380.0f || lambda >= 720.0f) return 0.0f;
    const int bin = (int)(((lambda - 380.0f) / (720.0f - 380.0f)) * 10.0f);
    return kSmits[s][bin];
}
// colour/spectrum_smits9.go:48-84
__device__ float smits(float r, float g, float b, float lambda) {
    float c = 0;
    if (r <= g && r <= b) {
        c += r * smits_eval(SM_WHITE, lambda);
        if (g <= b) { c += (g - r) * smits_eval(SM_CYAN, lambda); c += (b - g) * smits_eval(SM_BLUE, lambda); }
        else { c += (b - r) * smits_eval(SM_CYAN, lambda); c += (g - b) * smits_eval(SM_GREEN, lambda); }
    } else if (g <= r && g <= b) {
        c += g * smits_eval(SM_WHITE, lambda);
        if (r <= b) { c += (r - g) * smits_eval(SM_MAGENTA, lambda); c += (b - r) * smits_eval(SM_BLUE, lambda); }
        else { c += (b - g) * smits_eval(SM_MAGENTA, lambda); c += (r - b) * smits_eval(SM_RED, lambda); }
    } else {
        c += b * smits_eval(SM_WHITE, lambda);
        if (r <= g) { c += (r - b) * smits_eval(SM_YELLOW, lambda); c += (g - r) * smits_eval(SM_GREEN, lambda); }
        else { c += (g - b) * smits_eval(SM_YELLOW, lambda); c += (r - g) * smits_eval(SM_RED, lambda); }
    }
    return c;
}
__device__ __forceinline__ float cie(const float* tab, float lambda) {
    if (lambda < 360.0f || lambda >= 830.0f) return -1.0f;
    const int bin = (int)(((lambda - 360.0f) / (830.0f - 360.0f)) * 95.0f);
    return tab[bin];
}

// Hero-wavelength spectrum (colour/spectrum.go:27-112)
struct Spec { float c[4]; float lambda; };
__device__ __forceinline__ float wavelength(float lambda, int j) {
    float v = (lambda - 450.0f + ((float)j / 4.0f) * 300.0f);
    if (v >= 300.0f) v -= 300.0f;
    v += 450.0f;
    return v;
}
__device__ __forceinline__ Spec spec_from_rgb(float lambda, RGB c) {
    Spec s;
    s.lambda = lambda;
#pragma unroll
    for (int k = 0; k < 4; k++) s.c[k] = smits(c.r, c.g, c.b, wavelength(lambda, k));
    return s;
}
__device__ __forceinline__ void spec_mul(Spec& a, const Spec& b) { for (int k = 0; k < 4; k++) a.c[k] *= b.c[k]; }
__device__ __forceinline__ void spec_scale(Spec& a, float s) { for (int k = 0; k < 4; k++) a.c[k] *= s; }
__device__ RGB spec_to_rgb(const Spec& s) {
    float x = 0, y = 0, z = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float l = wavelength(s.lambda, i);
        x += s.c[i] * cie(kCieX, l);
        y += s.c[i] * cie(kCieY, l);
        z += s.c[i] * cie(kCieZ, l);
    }
    // colour/space_srgb.go:8-14 through colourspace.go XYZToRGB
    return RGB{x * 3.2404542f + y * -1.5371385f + z * -0.4985314f,
               x * -0.9692660f + y * 1.8760108f + z * 0.0415560f,
               x * 0.0556434f + y * -0.2040259f + z * 1.0572252f};
}

// ------------------------------------------------------------- QMC

// math/ldseq/ldseq.go:55-75
__device__ __forceinline__ uint64_t vdc_bits(uint64_t i, uint64_t scramble) {
    uint64_t bits = (i << 32) | (i >> 32);
    bits = ((bits & 0x0000ffff0000ffffull) << 16) | ((bits & 0xffff0000ffff0000ull) >> 16);
    bits = ((bits & 0x00ff00ff00ff00ffull) << 8) | ((bits & 0xff00ff00ff00ff00ull) >> 8);
    bits = ((bits & 0x0f0f0f0f0f0f0f0full) << 4) | ((bits & 0xf0f0f0f0f0f0f0f0ull) >> 4);
    bits = ((bits & 0x3333333333333333ull) << 2) | ((bits & 0xccccccccccccccccull) >> 2);
    bits = ((bits & 0x5555555555555555ull) << 1) | ((bits & 0xaaaaaaaaaaaaaaaaull) >> 1);
    return (scramble ^ bits) >> (64 - 52);
}
__device__ __forceinline__ double vdc(uint64_t i, uint64_t s) { return (double)vdc_bits(i, s) / 4503599627370496.0; }
// ldseq.go:83-96
__device__ __forceinline__ uint64_t sobol_bits(uint64_t i, uint64_t scramble) {
    uint64_t r = scramble >> (64 - 52);
    for (uint64_t v = 1ull << 51; i != 0; i >>= 1) {
        if (i & 1) r ^= v;
        v ^= v >> 1;
    }
    return r;
}
__device__ __forceinline__ double sobol(uint64_t i, uint64_t s) { return (double)sobol_bits(i, s) / 4503599627370496.0; }

// Path mode: the scramble of QMC dimension pair k at recursion level `level`, derived from one of the
// pixel's two shading scrambles by the SplitMix64 finaliser (Steele, Lea, Flood 2014).  Level 0's light
// samples keep the pixel's own scrambles, as the reference's renderer uses them; every other (level, k)
// gets an independent random XOR scramble, so the dimensions of different bounces are not correlated.
__device__ __forceinline__ uint64_t level_scramble(uint64_t s, int level, int k) {
    uint64_t z = s ^ ((uint64_t)(8 * level + k + 1) * 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// ------------------------------------------------------------- BSDFs

struct Frame { V3 U, V, N; };

// builtin/shader/bsdf/orennayar.go
struct OrenNayar {
    float lambda; V3 wi; float rough; Frame f;
    __device__ V3 sample(double r0, double r1) const {
        const double r = sqrt(1 - r0), th = 2 * M_PI * r1;    // math/sample/sample.go:18-27
        const V3 s = v3((float)(r * cos(th)), (float)(r * sin(th)), (float)sqrt(r0));
        return basis_expand(f.U, f.V, f.N, s);
    }
    __device__ double pdf(V3 o_) const {
        const V3 o = basis_project(f.U, f.V, f.N, o_);
        return (double)maxs(0.0f, o.z) / M_PI;
    }
    __device__ Spec eval(V3 o_) const {
        const V3 o = basis_project(f.U, f.V, f.N, o_);
        const float sg = rough;
        const float A = 1 - (0.5f * (sg * sg) / ((sg * sg) + 0.57f));
        const float B = 0.45f * (sg * sg) / ((sg * sg) + 0.09f);
        const float phiI = fatan2(wi.y, wi.x), phiO = fatan2(o.y, o.x);
        const float thI = facos(wi.z), thO = facos(o.z);
        const float alpha = maxs(thI, thO), beta = mins(thI, thO);
        const float C = fsin(alpha) * ftan(beta);
        const float gamma = fcos(phiO - phiI);
        const float sc = o.z * (A + (B * maxs(0.0f, gamma) * C));
        Spec rho = spec_from_rgb(lambda, RGB{1.0f, 1.0f, 1.0f});
        spec_scale(rho, sc / PI32);
        return rho;
    }
};

__device__ __forceinline__ float chi(float x) { return x > 0.0f ? 1.0f : 0.0f; }
__device__ float ggx_g1(V3 w, V3 m, float alpha) {
    const float odn = dot(w, m);
    const float denom = 1 + sqrtf(1 + (alpha * alpha) * ((1.0f / (w.z * w.z)) - 1));
    return chi(odn / w.z) * 2 / denom;
}
__device__ float ggx_d(V3 m, float alpha) {
    const float numer = alpha * alpha * chi(m.z);
    if (m.z == 1.0f) return 1.0f / (PI32 * alpha * alpha);
    const float denom = PI32 * sqr32(m.z * m.z) * sqr32(alpha * alpha + ((1.0f / (m.z * m.z)) - 1));
    return numer / denom;
}
__device__ __forceinline__ float signf(float v) { return v < 0 ? -1.0f : 1.0f; }
// builtin/shader/fresnel/dielectric.go:34-47
__device__ float dielectric_kr(float eta, float c) {
    float g = (eta * eta) - 1 + (c * c);
    if (g < 0.0f) return 1.0f;
    g = sqrtf(g);
    return 0.5f * sqr32((g - c) / (g + c)) * (1 + sqr32((c * (g + c) - 1) / (c * (g - c) + 1)));
}
// builtin/shader/fresnel/conductor.go:31-81 ("Artist Friendly Metallic Fresnel", Gulbrandsen 2014):
// index n and extinction k^2 from reflectivity r and edge tint g, then the unpolarised conductor term.
__device__ __forceinline__ float cond_nmin(float r) { return (1 - r) / (1 + r); }
__device__ __forceinline__ float cond_nmax(float r) { return (1 + sqrtf(r)) / (1 - sqrtf(r)); }
__device__ float conductor_fresnel(float r, float g, float c) {
    const float nr = maxs(0.0f, mins(r, 0.99f));   // math.Clamp = Max(min, Min(x, max))
    const float n = cond_nmin(nr) * g + (1 - g) * cond_nmax(nr);
    const float k2 = ((n + 1) * (n + 1) * nr - (n - 1) * (n - 1)) / (1 - nr);
    const float rs = (n * n + k2 - 2 * n * c + c * c) / (n * n + k2 + 2 * n * c + c * c);
    const float rp = ((n * n + k2) * c * c - 2 * n * c + 1) / ((n * n + k2) * c * c + 2 * n * c + 1);
    return 0.5f * (rs + rp);
}
// The core.Fresnel ShaderStd builds (std.go:171-190): Dielectric(ior) or Conductor(refl, edge).
struct Fresnel {
    bool metal; float eta; RGB refl, edge;
    __device__ RGB kr(float c) const {
        if (!metal) { const float f = dielectric_kr(eta, c); return RGB{f, f, f}; }
        return RGB{conductor_fresnel(refl.r, edge.r, c), conductor_fresnel(refl.g, edge.g, c),
                   conductor_fresnel(refl.b, edge.b, c)};
    }
};
__device__ Fresnel make_fresnel(const vgpu_material& m, float ior) {
    Fresnel f;
    f.metal = m.fresnel_model == VGPU_FRESNEL_METAL;
    f.eta = ior;
    f.refl = RGB{0.5f, 0.5f, 0.5f};
    f.edge = RGB{0.5f, 0.5f, 0.5f};
    if (f.metal) {
        if (m.has_spec1_fresnel_refl) f.refl = RGB{m.spec1_fresnel_refl[0], m.spec1_fresnel_refl[1], m.spec1_fresnel_refl[2]};
        // std.go:185-187: a set Spec1FresnelEdge overwrites the reflectivity; the edge tint stays grey
        if (m.has_spec1_fresnel_edge) f.refl = RGB{m.spec1_fresnel_edge[0], m.spec1_fresnel_edge[1], m.spec1_fresnel_edge[2]};
    }
    return f;
}
// builtin/shader/bsdf/microfacetggx.go
struct GGX {
    float lambda; V3 wr; float rough; Fresnel fres; Frame f;
    __device__ V3 sample(double r0, double r1) const {
        const float alpha = sqr32(rough);
        const double thM = atan2((double)alpha * sqrt(r0), sqrt(1 - r0));
        const double phM = 2.0 * M_PI * r1;
        const V3 m = v3(fsin((float)thM) * fcos((float)phM), fsin((float)thM) * fsin((float)phM), fcos((float)thM));
        const V3 o = sub(scale(2.0f * dotabs(m, wr), m), wr);
        return basis_expand(f.U, f.V, f.N, normalize(o));
    }
    __device__ double pdf(V3 o_) const {
        const V3 o = basis_project(f.U, f.V, f.N, o_);
        const float alpha = sqr32(rough);
        const V3 m = scale(signf(wr.z), normalize(add(wr, o)));
        const double p = (double)(ggx_d(m, alpha) * m.z);
        return __builtin_isnan(p) ? 0.0 : p;
    }
    // Density of the direction sample() returns: pdf() is the half vector's density D(m) m_z
    // (microfacetggx.go:108-122); reflecting about m scales it by 1 / (4 |wr . m|).  Path mode weighs
    // its GGX continuation with this (the reference never continues a GGX sample).
    __device__ double pdf_dir(V3 o_) const {
        const V3 o = basis_project(f.U, f.V, f.N, o_);
        const float alpha = sqr32(rough);
        const V3 m = scale(signf(wr.z), normalize(add(wr, o)));
        const double p = (double)(ggx_d(m, alpha) * m.z / (4.0f * dotabs(wr, m)));
        return __builtin_isnan(p) ? 0.0 : p;
    }
    __device__ Spec eval(V3 o_) const {
        const V3 wi = basis_project(f.U, f.V, f.N, o_);
        const float alpha = sqr32(rough);
        const V3 h = scale(signf(wr.z), normalize(add(wr, wi)));
        const RGB F = fres.kr(dotabs(wr, h));
        const float numer = ggx_g1(wr, h, alpha) * ggx_g1(wi, h, alpha) * ggx_d(h, alpha);
        const float denom = 4 * fabs32(wr.z) * fabs32(wi.z);
        Spec rho = spec_from_rgb(lambda, F);
        spec_scale(rho, fabs32(wi.z) * numer / denom);
        for (int k = 0; k < 4; k++) if (rho.c[k] < 0 || __builtin_isnan(rho.c[k])) rho.c[k] = 0;
        return rho;
    }
};

// ------------------------------------------------------------- lights

struct SC {
    V3 P, Poffset, N, Ng;
    float lambda;
    uint64_t I;
    uint64_t scr0, scr1;
};

__device__ RGB emission(const vgpu_material& m) {   // std.go:299-316
    RGB e = rgb0();
    if (m.has_emission_colour) e = RGB{m.emission[0], m.emission[1], m.emission[2]};
    if (!m.has_emission_strength) return rgb0();
    rgb_scale(e, m.emission_strength);
    return e;
}

// Moller-Trumbore (triangle.go:79-129)
__device__ bool ray_triangle(V3 Ro, V3 Rd, V3 P0, V3 P1, V3 P2, V3& out) {
    const V3 e1 = sub(P1, P0), e2 = sub(P2, P0);
    const V3 P = cross(Rd, e2);
    const float det = dot(e1, P);
    if (det > -1e-6f && det < 1e-6f) return false;
    const float inv = 1 / det;
    const V3 T = sub(Ro, P0);
    const float u = dot(T, P) * inv;
    if (u < 0 || u > 1) return false;
    const V3 Q = cross(T, e1);
    const float v = dot(Rd, Q) * inv;
    if (v < 0 || u + v > 1) return false;
    const float t = dot(e2, Q) * inv;
    if (t > 1e-6f) {
        const V3 a = scale(1 - u - v, P0), b = scale(u, P1), c = scale(v, P2);
        out = v3(a.x + b.x + c.x, a.y + b.y + c.y, a.z + b.z + c.z);
        return true;
    }
    return false;
}
__device__ __forceinline__ float tri_area(V3 P0, V3 P1, V3 P2) { return 0.5f * length(cross(sub(P1, P0), sub(P2, P0))); }
__device__ void sides_to_angles(float as, float bs, float cs, float& a, float& b, float& c) {
    const float ssu = (as + bs + cs) / 2;
    const float sa = fsin(ssu - as), sb = fsin(ssu - bs), scc = fsin(ssu - cs), s = fsin(ssu);
    const float tA = sqrtf(sb * scc / (s * sa));
    const float tB = sqrtf(sa * scc / (s * sb));
    const float tC = sqrtf(sa * sb / (s * scc));
    a = 2 * fatan(tA); b = 2 * fatan(tB); c = 2 * fatan(tC);
}
__device__ __forceinline__ void verts_to_sides(V3 pa, V3 pb, V3 pc, float& a, float& b, float& c) {
    a = facos(dot(pb, pc)); b = facos(dot(pc, pa)); c = facos(dot(pa, pb));
}
// triangle.go:474-535
__device__ V3 sample_sph_tri(V3 p0, V3 p1, V3 p2, V3 p, double r0, double r1, double& pdf) {
    const V3 pa = normalize(sub(p0, p)), pb = normalize(sub(p1, p)), pc = normalize(sub(p2, p));
    float a, b, c; verts_to_sides(pa, pb, pc, a, b, c);
    float al, be, ga; sides_to_angles(a, b, c, al, be, ga);
    const float area = al + be + ga - PI32;
    const float areaHat = (float)r0 * area;
    const float s = fsin(areaHat - al), t = fcos(areaHat - al);
    const float sinA = fsin(al), cosA = fcos(al);
    const float u = t - cosA;
    const float v = s + sinA * fcos(c);
    float q = ((v * t - u * s) * cosA - v) / ((v * s + u * t) * sinA);
    q = maxs(-1.0f, mins(q, 1.0f));
    float w = dot(pc, pa);
    V3 v31 = normalize(v3(pc.x - w * pa.x, pc.y - w * pa.y, pc.z - w * pa.z));
    const float sq = sqrtf(1 - q * q);
    const V3 v4 = v3(q * pa.x + sq * v31.x, q * pa.y + sq * v31.y, q * pa.z + sq * v31.z);
    const float z = 1 - (float)r1 * (1 - dot(v4, pb));
    w = dot(v4, pb);
    const V3 v42 = normalize(v3(v4.x - w * pb.x, v4.y - w * pb.y, v4.z - w * pb.z));
    pdf = 1 / (double)area;
    return add(scale(z, pb), scale(sqrtf(1 - z * z), v42));
}

struct LSample { float pdf; Spec liu; V3 ld; float ldist; bool ok; };
struct BSample { V3 d; double pdf; float pdf_light; Spec liu; V3 ld; float ldist; bool ok; };

// Tri.ValidSample (triangle.go:136-230)
__device__ bool tri_valid_sample(const DevLight& L, const vgpu_material& lm, const SC& sg, BSample& s) {
    const V3 P0 = from4(L.P0), P1 = from4(L.P1), P2 = from4(L.P2);
    V3 P;
    if (!ray_triangle(sg.P, s.d, P0, P1, P2, P)) return false;
    const V3 N = normalize(cross(sub(P1, P0), sub(P2, P0)));
    if (dot(sg.Ng, sub(P0, P)) < 0 || dot(sg.Ng, sub(P1, P)) < 0 || dot(sg.Ng, sub(P2, P)) < 0) {
        const double pdf = (double)(1 / tri_area(P0, P1, P2));
        const V3 D = sub(P, sg.P);
        s.ldist = length(D); s.ld = normalize(D);
        if (dot(s.ld, N) > 0 || dot(s.ld, sg.Ng) < 0) return false;
        s.liu = spec_from_rgb(sg.lambda, emission(lm));
        s.pdf_light = (float)pdf * sqr32(s.ldist) / dotabs(s.ld, N);
    } else {
        const V3 pa = normalize(sub(P0, sg.P)), pb = normalize(sub(P1, sg.P)), pc = normalize(sub(P2, sg.P));
        float a, b, c; verts_to_sides(pa, pb, pc, a, b, c);
        float al, be, ga; sides_to_angles(a, b, c, al, be, ga);
        const float area = al + be + ga - PI32;
        const double pdf = (double)(1 / area);
        const V3 D = sub(P, sg.P);
        s.ldist = length(D); s.ld = normalize(D);
        if (dot(s.ld, N) > 0 || dot(s.ld, sg.Ng) < 0) return false;
        s.liu = spec_from_rgb(sg.lambda, emission(lm));
        s.pdf_light = (float)pdf;
    }
    return true;
}

// Tri.SampleArea for sample i of n (triangle.go:232-343); ok=false when the sample is skipped.
__device__ LSample tri_sample_area(const DevLight& L, const vgpu_material& lm, const SC& sg, int n, int i) {
    LSample ls; ls.ok = false;
    const V3 P0 = from4(L.P0), P1 = from4(L.P1), P2 = from4(L.P2);
    const V3 N = normalize(cross(sub(P1, P0), sub(P2, P0)));
    const uint64_t idx = sg.I * (uint64_t)n + (uint64_t)i;
    const double r0 = vdc(idx, sg.scr0), r1 = sobol(idx, sg.scr1);
    V3 P;
    float pdfv;
    bool area_branch = dot(sg.Ng, sub(P0, sg.P)) < 0 || dot(sg.Ng, sub(P1, sg.P)) < 0 || dot(sg.Ng, sub(P2, sg.P)) < 0;
    if (area_branch) {
        const double pdf = (double)(1 / tri_area(P0, P1, P2));
        const V3 a = scale((float)(r1 * sqrt(1 - r0)), sub(P1, P0));
        const V3 b = scale((float)(1 - sqrt(1 - r0)), sub(P2, P0));
        P = v3(P0.x + a.x + b.x, P0.y + a.y + b.y, P0.z + a.z + b.z);
        pdfv = (float)pdf;
    } else {
        double pdf;
        const V3 x = sample_sph_tri(P0, P1, P2, sg.P, r0, r1, pdf);
        // rayPlaneIntersect (disk.go:38-51); t stays 0 when the plane is parallel
        const float denom = dot(N, x);
        float t = 0.0f;
        if (fabs32(denom) > 1e-6f) t = dot(sub(P0, sg.P), N) / denom;
        P = mad(sg.P, x, t);
        pdfv = (float)pdf;
    }
    const V3 D = sub(P, sg.P);
    ls.ldist = length(D); ls.ld = normalize(D);
    if (dot(ls.ld, N) > 0 || dot(ls.ld, sg.Ng) < 0) return ls;
    ls.liu = spec_from_rgb(sg.lambda, emission(lm));
    ls.pdf = area_branch ? pdfv * sqr32(ls.ldist) / dotabs(ls.ld, N) : pdfv;
    ls.ok = true;
    return ls;
}


// ---- Disk (builtin/light/disk.go)

// rayPlaneIntersect (disk.go:38-51)
__device__ __forceinline__ bool ray_plane(V3 Ro, V3 Rd, V3 P, V3 N, float& t) {
    const float denom = dot(N, Rd);
    t = 0.0f;
    if (fabs32(denom) > 1e-6f) {
        t = dot(sub(P, Ro), N) / denom;
        return t >= 0;
    }
    return false;
}
// Disk.PreRender frame (disk.go:84-86): N towards LookAt, T = N x Up, B = N x T
__device__ __forceinline__ void disk_frame(const DevLight& L, V3& N, V3& T, V3& B) {
    N = normalize(sub(from4(L.P2), from4(L.P0)));
    T = normalize(cross(N, from4(L.P1)));
    B = cross(N, T);
}
// Disk.ValidSample (disk.go:120-168)
__device__ bool disk_valid_sample(const DevLight& L, const vgpu_material& lm, const SC& sg, BSample& s) {
    const V3 P = from4(L.P0);
    const float R = L.P0.w;
    V3 N, T, B;
    disk_frame(L, N, T, B);
    const double pdf = (double)(1.0f / (PI32 * R * R));
    float t;
    if (!ray_plane(sg.P, s.d, P, N, t)) return false;    // rayDiskIntersect (disk.go:54-68)
    {
        const V3 v = sub(mad(sg.P, s.d, t), P);
        if (!(sqrtf(dot(v, v)) <= R)) return false;
    }
    const V3 p = mad(sg.P, s.d, t);
    const V3 V = sub(p, sg.P);
    if (dot(V, sg.Ng) <= 0.0f || dot(V, N) >= 0.0f) return false;
    s.ldist = length(V);
    s.ld = normalize(V);
    s.liu = spec_from_rgb(sg.lambda, emission(lm));
    s.pdf_light = (float)pdf * (s.ldist * s.ldist) / fabs32(dot(s.ld, N));
    return true;
}
// Disk.SampleArea for sample i of n (disk.go:171-215)
__device__ LSample disk_sample_area(const DevLight& L, const vgpu_material& lm, const SC& sg, int n, int i) {
    LSample ls; ls.ok = false;
    const V3 P0 = from4(L.P0);
    const float R = L.P0.w;
    V3 N, T, B;
    disk_frame(L, N, T, B);
    const uint64_t idx = sg.I * (uint64_t)n + (uint64_t)i;
    const double r0 = vdc(idx, sg.scr0), r1 = sobol(idx, sg.scr1);
    const float u = R * sqrtf((float)r0) * fcos(2.0f * PI32 * (float)r1);
    const float v = R * sqrtf((float)r0) * fsin(2.0f * PI32 * (float)r1);
    const double pdf = (double)(1.0f / (PI32 * R * R));
    const V3 su = scale(u, B), sv = scale(v, T);
    const V3 P = v3(P0.x + su.x + sv.x, P0.y + su.y + sv.y, P0.z + su.z + sv.z);   // Vec3Add3
    const V3 V = sub(P, sg.P);
    if (!(dot(V, sg.Ng) > 0.0f && dot(V, N) < 0.0f)) return ls;
    ls.ldist = length(V);
    ls.ld = normalize(V);
    ls.liu = spec_from_rgb(sg.lambda, emission(lm));
    ls.pdf = (float)pdf * (ls.ldist * ls.ldist) / fabs32(dot(ls.ld, N));
    ls.ok = true;
    return ls;
}

// ---- Sphere (builtin/light/sphere.go; the same quadratic as builtin/geom/sphere/trace.go:54-103)

__device__ bool ray_sphere(V3 Ro, V3 Rd, V3 C, float radius, float& t) {
    const V3 L = sub(Ro, C);
    const float a = dot(Rd, Rd);
    const float b = 2 * dot(Rd, L);
    const float c = dot(L, L) - radius * radius;
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
    if (x0 > x1) { const float tt = x0; x0 = x1; x1 = tt; }   // (raySphereIntersect's second swap is a no-op)
    if (x0 < 0) {
        x0 = x1;
        if (x0 < 0) return false;
    }
    t = x0;
    return true;
}
// Shirley 1996 cone pdf q_2 (sphere.go:170, :253)
__device__ __forceinline__ float sphere_cone_pdf(float R, float l) {
    return 1 / (2 * PI32 * (1 - sqrtf(1 - sqr32(R / l))));
}
// Sphere.ValidSample (sphere.go:129-177)
__device__ bool sphere_valid_sample(const DevLight& L, const vgpu_material& lm, const SC& sg, BSample& s) {
    const V3 C = from4(L.P0);
    const float R = L.P0.w;
    float t;
    if (!ray_sphere(sg.P, s.d, C, R, t)) return false;
    const V3 x = mad(sg.P, s.d, t);
    const V3 D = sub(x, sg.P);
    s.ldist = length(D);
    s.ld = normalize(D);
    s.liu = spec_from_rgb(sg.lambda, emission(lm));
    const float l = length(sub(C, sg.P));
    s.pdf_light = sphere_cone_pdf(R, l);
    return true;
}
// Sphere.SampleArea for sample i of n (sphere.go:180-256); sg.Sample is 0 here
__device__ LSample sphere_sample_area(const DevLight& L, const vgpu_material& lm, const SC& sg, int n, int i) {
    LSample ls; ls.ok = false;
    const V3 C = from4(L.P0);
    const float R = L.P0.w;
    const V3 V = sub(C, sg.P);
    const float l = length(V);
    const V3 w = normalize(V);
    const V3 v = normalize(cross(w, sg.Ng));
    const V3 u = cross(w, v);
    const uint64_t idx = sg.I * (uint64_t)n + (uint64_t)i;
    const double r0 = vdc(idx, sg.scr0), r1 = sobol(idx, sg.scr1);
    const float theta = facos(1 - (float)r0 + (float)r0 * sqrtf(1 - sqr32(R / l)));
    const float phi = 2 * PI32 * (float)r1;
    const V3 a = v3(fcos(phi) * fsin(theta), fsin(phi) * fsin(theta), fcos(theta));
    const V3 omega = basis_expand(u, v, w, a);
    if (dot(omega, sg.Ng) < 0) return ls;
    float t;
    if (!ray_sphere(sg.P, omega, C, R, t)) return ls;
    const V3 x = mad(sg.P, omega, t);
    const V3 D = sub(x, sg.P);
    ls.ldist = length(D);
    ls.ld = normalize(D);
    ls.liu = spec_from_rgb(sg.lambda, emission(lm));
    ls.pdf = sphere_cone_pdf(R, l);
    ls.ok = true;
    return ls;
}

__device__ bool valid_sample(const DevLight& L, const vgpu_material& lm, const SC& sg, BSample& s) {
    if (L.type == VGPU_LIGHT_DISK) return disk_valid_sample(L, lm, sg, s);
    if (L.type == VGPU_LIGHT_SPHERE) return sphere_valid_sample(L, lm, sg, s);
    return tri_valid_sample(L, lm, sg, s);
}
__device__ LSample sample_area(const DevLight& L, const vgpu_material& lm, const SC& sg, int n, int i) {
    if (L.type == VGPU_LIGHT_DISK) return disk_sample_area(L, lm, sg, n, i);
    if (L.type == VGPU_LIGHT_SPHERE) return sphere_sample_area(L, lm, sg, n, i);
    return tri_sample_area(L, lm, sg, n, i);
}

// core/shader.go:139-163
__device__ V3 offset_p(const SC& sc, int dir) {
    const V3 pofs = dir < 0 ? neg(sc.Poffset) : sc.Poffset;
    V3 po = add(sc.P, pofs);
    po.x = pofs.x > 0 ? next_up(po.x) : (pofs.x < 0 ? next_down(po.x) : po.x);
    po.y = pofs.y > 0 ? next_up(po.y) : (pofs.y < 0 ? next_down(po.y) : po.y);
    po.z = pofs.z > 0 ? next_up(po.z) : (pofs.z < 0 ? next_down(po.z) : po.z);
    return po;
}

struct SlotOut {
    float4* rays;   // 2 float4 per slot
    float4* rgb;    // 1 float4 per slot: colour if unoccluded, w = 1 when a shadow ray is traced
    float time;     // ray time of the shaded hit, inherited by its shadow rays (core/ray.go:65)
};

__device__ __forceinline__ void slot_off(const SlotOut& so, int s) {
    so.rays[2 * s] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    so.rays[2 * s + 1] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    so.rgb[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
__device__ __forceinline__ void slot_on(const SlotOut& so, int s, V3 o, V3 d, RGB c) {
    so.rays[2 * s] = make_float4(o.x, o.y, o.z, 1.0f);   // tmax 1: D spans the distance to the light
    so.rays[2 * s + 1] = make_float4(d.x, d.y, d.z, so.time);
    so.rgb[s] = make_float4(c.r, c.g, c.b, 1.0f);
}

// EvaluateLightSamples (core/shader.go:203-402) up to the shadow tests: fills the group's slots
// [base, base + sg_slots) and returns the group record {scale, mode}.  Not inlined: k_shade then
// needs 214 instead of 256 VGPRs + 86 AGPRs, runs two waves per SIMD instead of one, and the
// box-and-spheres pass is 5% faster.
template <class B>
__device__ __noinline__ float4 light_samples(const DevLight& L, const vgpu_material& lm, const SC& sc, const B& bsdf, int NS,
                                const SlotOut& so, int base, int sg_slots) {
    const float k = 1.0f - 0.0001f;   // 1 - ShadowRayEpsilon (core/ray.go:22)
    int on = 0;                       // shadow rays set up (returned in .z)
    for (int s = 0; s < sg_slots; s++) slot_off(so, base + s);
    if (NS > 1) {
        const int half = NS / 2;
        int nB = 0, nL = 0;
        // BSDF samples (validity only; contributions after the totals are known)
        for (int i = 0; i < half; i++) {
            const uint64_t idx = sc.I * (uint64_t)half + (uint64_t)i;
            BSample b; b.ok = false;
            b.d = normalize(bsdf.sample(vdc(idx, sc.scr0), sobol(idx, sc.scr1)));
            b.pdf = bsdf.pdf(b.d);
            if (!(b.pdf <= 0) && valid_sample(L, lm, sc, b)) nB = half;
        }
        for (int i = 0; i < half; i++) if (sample_area(L, lm, sc, half, i).ok) nL = half;
        const int total = nB + nL;
        if (total == 0) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (int i = 0; i < half; i++) {
            const LSample l = sample_area(L, lm, sc, half, i);
            if (!l.ok || dot(l.ld, sc.N) <= 0) continue;
            const V3 o = dot(l.ld, sc.Ng) < 0 ? offset_p(sc, -1) : offset_p(sc, 1);
            Spec rho = bsdf.eval(l.ld);
            spec_mul(rho, l.liu);
            float p_hat = (float)nB * (float)bsdf.pdf(l.ld) / (float)total;
            p_hat += (float)nL * l.pdf / (float)total;
            spec_scale(rho, 1.0f / p_hat);
            RGB c = spec_to_rgb(rho);
            if (c.r < 0) c.r = 0;
            if (c.g < 0) c.g = 0;
            if (c.b < 0) c.b = 0;
            slot_on(so, base + i, o, scale(l.ldist * k, l.ld), c);
            on++;
        }
        for (int i = 0; i < half; i++) {
            const uint64_t idx = sc.I * (uint64_t)half + (uint64_t)i;
            BSample b; b.ok = false;
            b.d = normalize(bsdf.sample(vdc(idx, sc.scr0), sobol(idx, sc.scr1)));
            b.pdf = bsdf.pdf(b.d);
            if (b.pdf <= 0 || !valid_sample(L, lm, sc, b)) continue;
            if (dot(b.ld, sc.N) <= 0) continue;
            const V3 o = dot(b.ld, sc.Ng) < 0 ? offset_p(sc, -1) : offset_p(sc, 1);
            Spec rho = bsdf.eval(b.ld);
            spec_mul(rho, b.liu);
            float p_hat = (float)nB * (float)b.pdf / (float)total;
            p_hat += (float)nL * b.pdf_light / (float)total;
            spec_scale(rho, 1.0f / p_hat);
            RGB c = spec_to_rgb(rho);
            if (c.r < 0) c.r = 0;
            if (c.g < 0) c.g = 0;
            if (c.b < 0) c.b = 0;
            slot_on(so, base + half + i, o, scale(b.ldist * k, b.ld), c);
            on++;
        }
        return make_float4(1.0f / (float)total, 1.0f, (float)on, 0.0f);
    }
    for (int i = 0; i < NS; i++) {
        const LSample l = sample_area(L, lm, sc, NS, i);
        if (!l.ok || dot(l.ld, sc.N) <= 0) continue;
        const V3 o = dot(l.ld, sc.Ng) < 0 ? offset_p(sc, -1) : offset_p(sc, 1);
        Spec rho = bsdf.eval(l.ld);
        spec_mul(rho, l.liu);
        spec_scale(rho, 1.0f / l.pdf);
        slot_on(so, base + i, o, scale(l.ldist * k, l.ld), spec_to_rgb(rho));
        on++;
    }
    return make_float4(1.0f, 2.0f, (float)on, 0.0f);
}

// ------------------------------------------------------------- kernels

// Sampler.WarpSample (builtin/filter/filter.go:40-85): linear search of the marginal CDF, then of the
// chosen row's conditional CDF, with the reference's lerps (including its i == 0 form).  A row search
// that finds nothing returns (0, 0).
__device__ void warp_sample(const RenderView& rv, double r0, double r1, double& u, double& v) {
    const int n = rv.filt_n;
    const double w = rv.filt_w;
    const double* cdfV = rv.filt;
    u = 0.0;
    v = 0.0;
    int uI = -1;
    for (int i = 0; i < n; i++) {
        uI = i;
        if (r0 < cdfV[i]) {
            if (i == 0) {
                const double du = r0 / cdfV[i];
                u = (-w / 2) + (du / w);
            } else {
                const double du = (r0 - cdfV[i - 1]) / (cdfV[i] - cdfV[i - 1]);
                u = (-w / 2) + w * ((double)i + du) / (double)(n - 1);
            }
            break;
        }
    }
    const double* row = rv.filt + n + (size_t)uI * n;
    for (int i = 0; i < n; i++) {
        if (r1 < row[i]) {
            if (i == 0) {
                const double dv = r1 / row[i];
                v = (-w / 2) + (dv / w);
            } else {
                const double dv = (r1 - row[i - 1]) / (row[i] - row[i - 1]);
                v = (-w / 2) + w * ((double)i + dv) / (double)(n - 1);
            }
            return;
        }
    }
    u = 0.0;
    v = 0.0;
}

// Camera matrix at a sample's time from the motion keys (camera.go:224-236): TransformDecompLerp and
// TransformDecompToMatrix4 (math/animdecomp.go:66-80) with QuatSlerp (math/quat.go:30-67),
// QuatToMatrix4 (quat.go:80-104) and Matrix4Mul (math/matrix4.go:92-102, sums from +0).
__device__ void camera_key_matrix(const RenderView& rv, float time, float* M) {
    const float k = time * (float)(rv.cam_nkeys - 1);
    const float t = k - floorf(k);
    const float* a = rv.cam_keys + min(max((int)floorf(k), 0), rv.cam_nkeys - 1) * kCamKeyFloats;
    const float* b = rv.cam_keys + min(max((int)ceilf(k), 0), rv.cam_nkeys - 1) * kCamKeyFloats;
    xf_decomp_matrix(a, b, t, M);
}

__global__ __launch_bounds__(256) void k_camera(RenderView rv, const uint64_t* __restrict__ scr, uint32_t iter,
                                                float4* __restrict__ rays, float4* __restrict__ pix) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && rv.counts) atomicAdd(rv.counts, (unsigned long long)rv.width * (unsigned long long)rv.height);
    if (p >= rv.width * rv.height) return;
    const uint32_t x = (uint32_t)(p % rv.width), y = (uint32_t)(p / rv.width);
    const uint64_t* s = scr + (size_t)p * 6;
    // ldseq.RasterXY(12, iter, x, y, 0, 0) (raster.go:10-57)
    const uint32_t m = 12;
    uint64_t index = (uint64_t)iter << (2 * m);
    uint64_t delta = 0;
    {
        uint32_t f = iter;
        for (uint32_t c = 0; f != 0; f >>= 1, c++) if (f & 1) delta ^= rv.raster_fwd[c];
    }
    uint64_t b = (((uint64_t)x << m) | (uint64_t)y) ^ delta;
    for (uint32_t c = 0; b != 0; b >>= 1, c++) if (b & 1) index ^= rv.raster_inv[c];
    double rx = (double)vdc_bits(index, 0) / (double)(1ull << (52 - m));
    double ry = (double)sobol_bits(index, 0) / (double)(1ull << (52 - m));
    if (rv.filt_n > 0) {   // core/render.go:100-108
        const double pixu = rx - floor(rx), pixv = ry - floor(ry);
        double u, v;
        warp_sample(rv, pixu, pixv, u, v);
        rx = floor(rx) + 0.5 + u;
        ry = floor(ry) + 0.5 + v;
    }
    const double lambda = (720 - 450) * vdc(iter, s[3]) + 450;
    const float time = (float)vdc(iter, s[2]);   // core/render.go:93,118
    const float Sx = (float)(-1.0 + 2.0 * (rx / (double)rv.width));
    const float Sy = -(float)(-1.0 + 2.0 * (ry / (double)rv.height));
    // camera.ComputeRay (camera.go:221-300): pinhole, or a thin lens when Radius > 0
    const float camu = Sx * rv.tan_theta_focal;
    const float camv = Sy * (rv.tan_theta_focal / rv.aspect);
    const V3 sv = sub(add(scale(camu, v3(1, 0, 0)), scale(camv, v3(0, 1, 0))), scale(rv.focal, v3(0, 0, 1)));
    V3 e = v3(0.0f, 0.0f, 0.0f);
    if (rv.lens_radius > 0.0f) {
        // lensU / lensV (core/render.go:91-92), sample.UniformDisk2D (math/sample/sample.go:105-129)
        const float r0 = (float)vdc(iter, s[0]), r1 = (float)sobol(iter, s[1]);
        const float x = -1 + 2 * r0, y = -1 + 2 * r1;
        float r = 0.0f, th = 0.0f;
        if (x > -y && x > y) { r = x; th = (PI32 / 4) * y / x; }
        else if (x > -y && x < y) { r = y; th = (PI32 / 4) * (2 - x / y); }
        else if (x < y && x < -y) { r = -x; th = (PI32 / 4) * (4 + y / x); }
        else if (x > y && x < -y) { r = -y; th = (PI32 / 4) * (6 - x / y); }
        const float xo = rv.lens_radius * r * fcos(th), yo = rv.lens_radius * r * fsin(th);
        e = add(scale(xo, v3(1, 0, 0)), scale(yo, v3(0, 1, 0)));
    }
    const V3 se = sub(sv, e);
    float M[16];
    if (rv.cam_nkeys >= 2) camera_key_matrix(rv, time, M);
    else for (int i = 0; i < 16; i++) M[i] = rv.M[i];
    const V3 d = v3(M[0] * se.x + M[4] * se.y + M[8] * se.z, M[1] * se.x + M[5] * se.y + M[9] * se.z,
                    M[2] * se.x + M[6] * se.y + M[10] * se.z);
    const V3 D = normalize(d);
    const V3 P = v3(M[0] * e.x + M[4] * e.y + M[8] * e.z + M[12], M[1] * e.x + M[5] * e.y + M[9] * e.z + M[13],
                    M[2] * e.x + M[6] * e.y + M[10] * e.z + M[14]);
    rays[2 * p] = make_float4(P.x, P.y, P.z, __builtin_inff());
    rays[2 * p + 1] = make_float4(D.x, D.y, D.z, time);
    pix[p] = make_float4(D.x, D.y, D.z, (float)lambda);
}

// m.Vec3Lerp (math/vec3.go:52-58): (1 - t) a + t b per component
__device__ __forceinline__ V3 lerp3(const float4& a, const float4& b, float t) {
    const float s = 1.0f - t;
    return v3(s * a.x + t * b.x, s * a.y + t * b.y, s * a.z + t * b.z);
}

// builtin/shader/bsdf/specular.go: perfect mirror, reflect(omegaR, (0,0,1)) = 2 (N . omegaR) N - omegaR
__device__ __forceinline__ V3 reflect_z(V3 wr) {
    const V3 N = v3(0.0f, 0.0f, 1.0f);
    return sub(scale(2.0f * dot(N, wr), N), wr);
}
struct Specular {
    float lambda; V3 wr; Fresnel fres; Frame f;
    __device__ V3 sample() const { return basis_expand(f.U, f.V, f.N, normalize(reflect_z(wr))); }
    __device__ double pdf(V3 o_) const {
        const V3 o = basis_project(f.U, f.V, f.N, o_);
        return dot(o, reflect_z(wr)) < 0.9999f ? 0.0 : 1.0;
    }
    __device__ Spec eval(V3 o_) const {
        const V3 o = basis_project(f.U, f.V, f.N, o_);
        Spec rho;
        rho.c[0] = rho.c[1] = rho.c[2] = rho.c[3] = 0.0f;
        rho.lambda = 0.0f;
        if (dot(o, reflect_z(wr)) < 0.9999f) return rho;
        rho = spec_from_rgb(lambda, fres.kr(wr.z));
        spec_scale(rho, dotabs(o, v3(0.0f, 0.0f, 1.0f)));
        return rho;
    }
};

// Path mode: a continuation ray's origin moved off the surface along the unit geometric normal n by 256
// ulps per component, or by 2^-16 n within 1/32 of zero (Waechter and Binder, Ray Tracing Gems, 2019).
// offset_p's error bound is zero for a surface lying exactly in a coordinate plane through the origin,
// where a continuation ray would otherwise find its own surface again at t ~ 0.
__device__ __forceinline__ float offset_along1(float p, float n) {
    const int of = (int)(256.0f * n);
    const float pi = __int_as_float(__float_as_int(p) + (p < 0.0f ? -of : of));
    return fabs32(p) < 0.03125f ? p + 1.52587891e-05f * n : pi;
}
__device__ __forceinline__ V3 offset_along(V3 p, V3 n) {
    return v3(offset_along1(p.x, n.x), offset_along1(p.y, n.y), offset_along1(p.z, n.z));
}

__device__ __forceinline__ void ray_off(float4* rays, int p) {
    rays[2 * p] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    rays[2 * p + 1] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
}

// One recursion level of core.Trace + ShaderStd.Eval for one pixel whose level-`level` ray hit; returns the
// shadow rays it sets up, and in `next` the rays it sends to the next level (a mirror ray or path mode's
// continuation: 0 or 1).
__device__ __forceinline__ int shade_pixel(int& next, const RenderView& rv, const uint64_t* __restrict__ scr, uint32_t iter, int level,
                                               const float4* __restrict__ rays_in, const int4* __restrict__ hits32,
                                               const float4* __restrict__ pix, float4* __restrict__ slot_rays,
                                               float4* __restrict__ slot_rgb, float4* __restrict__ groups,
                                               float4* __restrict__ pixshade, float4* __restrict__ rays_out,
                                               LevelRec* __restrict__ recs) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    next = 0;
    if (p >= rv.width * rv.height) return 0;
    const int SGS = 2 * rv.max_half;   // slots per group
    const float4 ra = rays_in[2 * p], rb = rays_in[2 * p + 1];
    SlotOut so{slot_rays + (size_t)p * rv.slots * 2, slot_rgb + (size_t)p * rv.slots, rb.w};
    float4* grp = groups + (size_t)p * rv.groups;
    float4* ps = pixshade + (size_t)p * 4;
    LevelRec* rec = recs + (size_t)p * rv.rec_levels + level;
    const int4 h0 = hits32[2 * p], h1 = hits32[2 * p + 1];
    const int prim = h0.x, g = h0.y;
    // the mesh or sphere whose element was hit: the hit geom itself, or the geom an instance or proc holds
    const int lg = rv.sv.any_xform ? h1.z : g;
    int mat = -1;
    DevGeom geo = {0, 0, 1, 0.0f};
    DevGeomShade gs = {-1, -1, 0, 0};
    if (prim >= 0) {
        geo = rv.sv.geoms[lg];
        gs = rv.gshade[lg];
        const int sidx = gs.sidx_base >= 0 ? (int)rv.shaderidx[gs.sidx_base + prim] : 0;
        mat = gs.nshaders > 0 ? rv.shaders[gs.shader_base + sidx] : -1;
    }
    const bool path = rv.path_depth > 0;
    if (mat < 0) {   // a miss, or a geom without a shader: core.Trace reports no hit (trace.go:62-64)
        for (int s = 0; s < rv.slots; s++) slot_off(so, s);
        for (int q = 0; q < rv.groups; q++) grp[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        ps[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        rec->a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        rec->s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (path) rec->c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (rays_out) ray_off(rays_out, p);
        return 0;
    }
    const V3 Ro = v3(ra.x, ra.y, ra.z), Rd = v3(rb.x, rb.y, rb.z);
    V3 Ro_obj = Ro, Rd_obj = Rd;   // the ray in the space of the hit geom (for the analytic sphere)
    const float thit = __int_as_float(h0.z);
    // Transforms of the hit (instance.go:108-111, polymesh/trace.go:57-70): M / Minv end as
    // sg.Transform / sg.InvTransform (an instance's or proc's over its mesh's own); Mmi is the mesh's own
    // inverse, which PolyMesh.Trace applies to N before ApplyTransform applies Minv^T again.
    bool xf = false, mesh_xf = false;
    float M[16], Minv[16], Mmi[16];
    if (rv.sv.any_xform) {
        const DevXform xl = rv.sv.xforms[lg], xt = rv.sv.xforms[g];
        if (xl.kind == kXfMesh) {
            xform_at(xl, rv.sv.xkeys, rb.w, M, Minv);
            for (int i = 0; i < 16; i++) Mmi[i] = Minv[i];
            xf = mesh_xf = true;
        }
        if (xt.kind == kXfInstance || xt.kind == kXfProc) {
            xform_at(xt, rv.sv.xkeys, rb.w, M, Minv);
            xf = true;
            // the analytic sphere's hit point is found in the instance's space (sphere/trace.go:30-35)
            float o[3], d[3];
            const float Pw[3] = {Ro.x, Ro.y, Ro.z}, Dw[3] = {Rd.x, Rd.y, Rd.z};
            xf_point(Minv, Pw, o);
            xf_vec(Minv, Dw, d);
            Ro_obj = v3(o[0], o[1], o[2]);
            Rd_obj = v3(d[0], d[1], d[2]);
        }
    }
    SC sc;
    V3 N, Ng, dPdu, dPdv;
    if (geo.nkeys == 0) {
        // analytic sphere (builtin/geom/sphere/trace.go:13-52)
        const float4 sp = rv.sv.spheres[geo.tri_base];
        sc.P = mad(Ro_obj, Rd_obj, thit);
        N = normalize(sub(sc.P, v3(sp.x, sp.y, sp.z)));
        Ng = N;
        sc.Poffset = scale(0.001f, N);
        dPdu = v3(1.0f, 0.0f, 0.0f);
        dPdv = v3(0.0f, 0.0f, 1.0f);
    } else if (geo.nkeys > 1) {
        // moving mesh: attributes at the ray's time (builtin/geom/polymesh/trace.go:547-554, 625-703)
        const float U = __int_as_float(h0.w), V = __int_as_float(h1.x), W = __int_as_float(h1.y);
        int key, key2;
        float mt;
        {
            const float k = rb.w * (float)(geo.nkeys - 1);   // trace.go:79-84
            const float fk = floorf(k);
            mt = k - fk;
            key = min(max((int)fk, 0), geo.nkeys - 1);   // as the traversal holds a time outside [0, 1]
            key2 = min(max((int)ceilf(k), 0), geo.nkeys - 1);
        }
        const float4* q = rv.sv.mtris + geo.tri_base + (size_t)prim * 3 * geo.nkeys;
        const V3 a = lerp3(q[3 * key], q[3 * key2], mt), b = lerp3(q[3 * key + 1], q[3 * key2 + 1], mt),
                 c = lerp3(q[3 * key + 2], q[3 * key2 + 2], mt);
        sc.P = v3(U * a.x + V * b.x + W * c.x, U * a.y + V * b.y + W * c.y, U * a.z + V * b.z + W * c.z);
        const float xs = fabs32(U * a.x) + fabs32(V * b.x) + fabs32(W * c.x);
        const float ys = fabs32(U * a.y) + fabs32(V * b.y) + fabs32(W * c.y);
        const float zs = fabs32(U * a.z) + fabs32(V * b.z) + fabs32(W * c.z);
        const float e00 = b.x - a.x, e01 = b.y - a.y, e02 = b.z - a.z;
        const float e10 = c.x - a.x, e11 = c.y - a.y, e12 = c.z - a.z;
        Ng = normalize(v3(e01 * e12 - e02 * e11, e02 * e10 - e00 * e12, e00 * e11 - e01 * e10));
        dPdu = v3(e00, e01, e02);
        dPdv = v3(e10, e11, e12);
        const float g7 = (7.0f * 5.9604645e-08f) / (1.0f - 7.0f * 5.9604645e-08f);   // Gamma(7), math/ferror.go:22-24
        const float dd = g7 * xs * fabs32(Ng.x) + g7 * ys * fabs32(Ng.y) + g7 * zs * fabs32(Ng.z);
        sc.Poffset = scale(dd, Ng);
        N = Ng;
        if (gs.nrm_base >= 0) {   // per face corner and key, as trace.go:694-703 indexes them
            const float4* nq = rv.normals + gs.nrm_base + (size_t)prim * 3 * geo.nkeys;
            const V3 n0 = lerp3(nq[3 * key], nq[3 * key2], mt), n1 = lerp3(nq[3 * key + 1], nq[3 * key2 + 1], mt),
                     n2 = lerp3(nq[3 * key + 2], nq[3 * key2 + 2], mt);
            N = normalize(v3(U * n0.x + V * n1.x + W * n2.x, U * n0.y + V * n1.y + W * n2.y, U * n0.z + V * n1.z + W * n2.z));
        }
    } else {
        const float U = __int_as_float(h0.w), V = __int_as_float(h1.x), W = __int_as_float(h1.y);
        const float4* tp = rv.sv.tris + ((size_t)geo.tri_base + prim) * 3;
        const V3 a = from4(tp[0]), b = from4(tp[1]), c = from4(tp[2]);
        // hit attributes (builtin/geom/polymesh/trace.go:292-513)
        const float xs = fabs32(U * a.x) + fabs32(V * b.x) + fabs32(W * c.x);
        const float ys = fabs32(U * a.y) + fabs32(V * b.y) + fabs32(W * c.y);
        const float zs = fabs32(U * a.z) + fabs32(V * b.z) + fabs32(W * c.z);
        const float e00 = b.x - a.x, e01 = b.y - a.y, e02 = b.z - a.z;
        const float e10 = c.x - a.x, e11 = c.y - a.y, e12 = c.z - a.z;
        Ng = normalize(v3(e01 * e12 - e02 * e11, e02 * e10 - e00 * e12, e00 * e11 - e01 * e10));
        N = Ng;
        if (gs.nrm_base >= 0) {
            const float4* np = rv.normals + gs.nrm_base + (size_t)prim * 3;
            const V3 n0 = from4(np[0]), n1 = from4(np[1]), n2 = from4(np[2]);
            N = normalize(v3(U * n0.x + V * n1.x + W * n2.x, U * n0.y + V * n1.y + W * n2.y, U * n0.z + V * n1.z + W * n2.z));
        }
        const float g7 = (7.0f * 5.9604645e-08f) / (1.0f - 7.0f * 5.9604645e-08f);   // Gamma(7), math/ferror.go:22-24
        const float dd = g7 * xs * fabs32(Ng.x) + g7 * ys * fabs32(Ng.y) + g7 * zs * fabs32(Ng.z);
        sc.Poffset = scale(dd, Ng);
        sc.P = v3(U * a.x + V * b.x + W * c.x, U * a.y + V * b.y + W * c.y, U * a.z + V * b.z + W * c.z);
        V3 axisu;
        {
            const float dx = 1 * Ng.x + 0 * Ng.y + 0 * Ng.z;
            axisu = v3(1.0f - Ng.x * dx, 0.0f - Ng.y * dx, 0.0f - Ng.z * dx);
            if (len2(axisu) < 0.1f || fabs32(dot(axisu, Ng)) > 0.3f) {
                const float dz = 0 * Ng.x + 0 * Ng.y + 1 * Ng.z;
                axisu = v3(0.0f - Ng.x * dz, 0.0f - Ng.y * dz, 1.0f - Ng.z * dz);
            }
        }
        dPdu = normalize(axisu);
        dPdv = cross(Ng, dPdu);
    }
    if (xf) {
        // ApplyTransform (core/shader.go:129-135) with the hit's transform; Poffset stays in object space
        float a[3], b[3];
        const float Po[3] = {sc.P.x, sc.P.y, sc.P.z};
        xf_point(M, Po, a);
        sc.P = v3(a[0], a[1], a[2]);
        float n1[3] = {N.x, N.y, N.z};
        if (mesh_xf) { xf_vec_t(Mmi, n1, b); n1[0] = b[0]; n1[1] = b[1]; n1[2] = b[2]; }
        xf_vec_t(Minv, n1, a);
        sc.N = normalize(v3(a[0], a[1], a[2]));
        const float ng[3] = {Ng.x, Ng.y, Ng.z};
        xf_vec_t(Minv, ng, a);
        sc.Ng = normalize(v3(a[0], a[1], a[2]));
        const float du[3] = {dPdu.x, dPdu.y, dPdu.z}, dv[3] = {dPdv.x, dPdv.y, dPdv.z};
        xf_vec(M, du, a);
        dPdu = normalize(v3(a[0], a[1], a[2]));
        xf_vec(M, dv, a);
        dPdv = normalize(v3(a[0], a[1], a[2]));
    } else {
        // ApplyTransform with identity matrices re-normalises (core/shader.go:129-135)
        sc.N = normalize(N);
        sc.Ng = normalize(Ng);
        dPdu = normalize(dPdu);
        dPdv = normalize(dPdv);
    }
    sc.lambda = pix[p].w;
    sc.I = iter;
    const uint64_t scr4 = scr[(size_t)p * 6 + 4], scr5 = scr[(size_t)p * 6 + 5];
    sc.scr0 = (path && level > 0) ? level_scramble(scr4, level, 0) : scr4;
    sc.scr1 = (path && level > 0) ? level_scramble(scr5, level, 1) : scr5;
    const vgpu_material mt = rv.materials[mat];
    // ShaderStd.Eval (std.go:99-285)
    V3 Vv = cross(sc.N, dPdu);
    if (len2(Vv) < 0.1f) Vv = cross(sc.N, dPdv);
    Vv = normalize(Vv);
    const V3 Uu = normalize(cross(sc.N, Vv));
    const Frame fr{Uu, Vv, sc.N};
    const float drough = mt.has_diffuse_roughness ? mt.diffuse_roughness : 0.5f;
    float dw = mt.has_diffuse_strength ? mt.diffuse_strength : 0.0f;
    float sw = mt.has_spec1_strength ? mt.spec1_strength : 0.0f;
    const float total = dw + sw;
    dw /= total;
    sw /= total;
    const float srough = mt.has_spec1_roughness ? mt.spec1_roughness : 0.5f;
    const float ior = mt.has_ior ? mt.ior : 1.7f;
    const Fresnel fres = make_fresnel(mt, ior);
    const bool diff_on = dw > 0.0f, spec_on = sw > 0.0f && srough > 0.0f, mirror = sw > 0.0f && srough == 0.0f;
    const V3 wr = basis_project(Uu, Vv, sc.N, neg(Rd));
    const OrenNayar on{sc.lambda, wr, drough * drough, fr};
    const GGX gx{sc.lambda, wr, srough * srough, fres, fr};
    int nrays = 0;
    for (int l = 0; l < rv.nlights; l++) {
        const DevLight L = rv.lights[l];
        const int NS = level == 0 ? 1 << L.samples : 1;   // NextLight: one sample per light below level 0 (shader.go:186-191)
        const bool skip = L.geom == g;   // LightsPrepare drops the light owning the hit geom (scene.go:100-116)
        const vgpu_material lm = rv.materials[L.material];
        const int b0 = (2 * l) * SGS, b1 = (2 * l + 1) * SGS;
        float4 g0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), g1 = g0;
        if (!skip && diff_on) g0 = light_samples(L, lm, sc, on, NS, so, b0, SGS);
        else for (int s = 0; s < SGS; s++) slot_off(so, b0 + s);
        if (!skip && spec_on) g1 = light_samples(L, lm, sc, gx, NS, so, b1, SGS);
        else for (int s = 0; s < SGS; s++) slot_off(so, b1 + s);
        grp[2 * l] = g0;
        grp[2 * l + 1] = g1;
        nrays += (int)g0.z + (int)g1.z;
    }
    RGB e = emission(mt);
    if (path && level > 0 && rec[-1].c.w == kPathSmooth) {
        // Reached by an Oren-Nayar or GGX sample: a light's own geometry was sampled at the previous
        // vertex by EvaluateLightSamples (with its BSDF samples), so its emission is not added twice.
        for (int l = 0; l < rv.nlights; l++) if (rv.lights[l].geom == g) e = rgb0();
    }
    ps[0] = make_float4(e.r, e.g, e.b, 1.0f);
    ps[1] = mt.has_diffuse_colour ? make_float4(mt.diffuse[0], mt.diffuse[1], mt.diffuse[2], dw)
                                  : make_float4(0.0f, 0.0f, 0.0f, dw);
    ps[2] = mt.has_spec1_colour ? make_float4(mt.spec1[0], mt.spec1[1], mt.spec1[2], 0.0f)
                                : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    ps[3] = make_float4(diff_on ? 1.0f : 0.0f, spec_on ? 1.0f : 0.0f, 0.0f, 0.0f);
    // Perfect mirror (std.go:209-261): one reflected ray at level + 1.  A level-4 hit shades to zero
    // (std.go:97-99), so a level-3 mirror adds +0 whatever its ray hits, as an untraced one does.
    float kind = 1.0f;
    float4 srec = make_float4(0.0f, 0.0f, 0.0f, sw);
    bool traced = false;
    if (path) {
        // Path mode: one continuation ray per vertex.  The lobe is picked with probability equal to its
        // normalised weight (diffuse dw, specular sw), so the weight over the pick probability is 1 and
        // the factor is (Eval / pdf).ToRGB() * colour, as the reference weighs its mirror ray
        // (std.go:226-258).  Not traced at the last segment, below the geometric surface, or at pdf 0.
        float4 crec = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (level + 1 < rv.path_depth && rays_out) {
            const double u = vdc(iter, level_scramble(scr4, level, 2));
            const double r0 = vdc(iter, level_scramble(scr4, level, 3)), r1 = sobol(iter, level_scramble(scr5, level, 4));
            V3 o = v3(0.0f, 0.0f, 0.0f);
            double pdf = 0.0;
            Spec rho;
            RGB col = rgb0();
            float flag = 0.0f;
            if (u < (double)dw) {
                o = normalize(on.sample(r0, r1));
                pdf = on.pdf(o);
                rho = on.eval(o);
                if (mt.has_diffuse_colour) col = RGB{mt.diffuse[0], mt.diffuse[1], mt.diffuse[2]};
                flag = kPathSmooth;
            } else if (mirror) {
                const Specular spc{sc.lambda, wr, fres, fr};
                o = spc.sample();
                pdf = spc.pdf(o);
                rho = spc.eval(o);
                if (mt.has_spec1_colour) col = RGB{mt.spec1[0], mt.spec1[1], mt.spec1[2]};
                flag = kPathDelta;
            } else if (spec_on) {
                o = normalize(gx.sample(r0, r1));
                pdf = gx.pdf_dir(o);
                rho = gx.eval(o);
                if (mt.has_spec1_colour) col = RGB{mt.spec1[0], mt.spec1[1], mt.spec1[2]};
                flag = kPathSmooth;
            }
            if (flag != 0.0f && pdf > 0.0 && dot(o, sc.Ng) > 0.0f) {
                spec_scale(rho, 1.0f / (float)pdf);
                RGB f = spec_to_rgb(rho);
                rgb_mul(f, col);
                crec = make_float4(f.r, f.g, f.b, flag);
                const V3 po = offset_along(offset_p(sc, 1), sc.Ng);
                rays_out[2 * p] = make_float4(po.x, po.y, po.z, __builtin_inff());
                rays_out[2 * p + 1] = make_float4(o.x, o.y, o.z, rb.w);
                traced = true;
            }
        }
        rec->c = crec;
    } else if (mirror) {
        kind = 3.0f;
        const Specular spc{sc.lambda, wr, fres, fr};
        const V3 o = spc.sample();
        const double pdf = spc.pdf(o);
        if (dot(o, sc.Ng) > 0.0f && level < kMaxLevel && rays_out) {
            Spec rho = spc.eval(o);
            spec_scale(rho, 1.0f / (float)pdf);
            RGB f = spec_to_rgb(rho);
            const RGB scol = mt.has_spec1_colour ? RGB{mt.spec1[0], mt.spec1[1], mt.spec1[2]} : rgb0();
            rgb_mul(f, scol);
            srec = make_float4(f.r, f.g, f.b, sw);
            const V3 po = offset_p(sc, 1);
            rays_out[2 * p] = make_float4(po.x, po.y, po.z, __builtin_inff());
            rays_out[2 * p + 1] = make_float4(o.x, o.y, o.z, rb.w);
            kind = 2.0f;
            traced = true;
        }
    }
    if (rays_out && !traced) ray_off(rays_out, p);
    rec->a = make_float4(0.0f, 0.0f, 0.0f, kind);
    rec->s = srec;
    next = traced ? 1 : 0;
    return nrays;
}

// Adds every lane's count to *counter with one atomic per wavefront (all 64 lanes call it).
__device__ __forceinline__ void wave_count(unsigned long long* counter, int n) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) n += __shfl_xor(n, d);
    if ((threadIdx.x & 63) == 0 && n != 0) atomicAdd(counter, (unsigned long long)n);
}

__global__ __launch_bounds__(128) void k_shade(RenderView rv, const uint64_t* __restrict__ scr, uint32_t iter, int level,
                                               const float4* __restrict__ rays_in, const int4* __restrict__ hits32,
                                               const float4* __restrict__ pix, float4* __restrict__ slot_rays,
                                               float4* __restrict__ slot_rgb, float4* __restrict__ groups,
                                               float4* __restrict__ pixshade, float4* __restrict__ rays_out,
                                               LevelRec* __restrict__ recs) {
    int next;
    const int nshadow = shade_pixel(next, rv, scr, iter, level, rays_in, hits32, pix, slot_rays, slot_rgb, groups, pixshade,
                                    rays_out, recs);
    if (rv.counts) {
        wave_count(rv.counts, next);
        wave_count(rv.counts + 1, nshadow);
    }
}


// Sums one level's unoccluded slot colours in the reference's order and applies the per-light / per-lobe
// scales: a = (0 + emission) + diffuse, and for a rough specular lobe s = its direct lighting.
// Without mirrors (max_level 0) the single level is the sample: the film is updated here directly.
__device__ __forceinline__ void film_update(float* __restrict__ film, float* __restrict__ sample_out, int p,
                                            uint32_t iter, RGB contrib) {
    float* f = film + (size_t)p * 3;
    f[0] = (f[0] * (float)iter + contrib.r) / (float)(iter + 1);
    f[1] = (f[1] * (float)iter + contrib.g) / (float)(iter + 1);
    f[2] = (f[2] * (float)iter + contrib.b) / (float)(iter + 1);
    if (sample_out) {
        sample_out[(size_t)p * 3] = contrib.r;
        sample_out[(size_t)p * 3 + 1] = contrib.g;
        sample_out[(size_t)p * 3 + 2] = contrib.b;
    }
}

__global__ __launch_bounds__(256) void k_resolve(RenderView rv, int level, uint32_t iter, const float4* __restrict__ slot_rgb,
                                                 const int4* __restrict__ slot_hits, const float4* __restrict__ groups,
                                                 const float4* __restrict__ pixshade, LevelRec* __restrict__ recs,
                                                 float* __restrict__ film, float* __restrict__ sample_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= rv.width * rv.height) return;
    const float4* ps = pixshade + (size_t)p * 4;
    const float4 e = ps[0];
    if (e.w == 0.0f) {
        if (film) film_update(film, sample_out, p, iter, rgb0());
        return;
    }
    LevelRec* rec = recs + (size_t)p * rv.rec_levels + level;
    const int SGS = 2 * rv.max_half;
    const float4 dc = ps[1], spc = ps[2], fl = ps[3];
    const float4* rgb = slot_rgb + (size_t)p * rv.slots;
    const int4* hit = slot_hits + (size_t)p * rv.slots;
    const float4* grp = groups + (size_t)p * rv.groups;
    RGB diff = rgb0(), spec = rgb0();
    for (int lobe = 0; lobe < 2; lobe++) {
        if ((lobe == 0 ? fl.x : fl.y) == 0.0f) continue;
        RGB& acc = lobe == 0 ? diff : spec;
        const RGB colour = lobe == 0 ? RGB{dc.x, dc.y, dc.z} : RGB{spc.x, spc.y, spc.z};
        for (int l = 0; l < rv.nlights; l++) {
            const float4 gr = grp[2 * l + lobe];
            RGB col = rgb0();
            if (gr.y != 0.0f) {
                const int base = (2 * l + lobe) * SGS;
                for (int s = 0; s < SGS; s++) {
                    const float4 c = rgb[base + s];
                    if (c.w != 0.0f && hit[base + s].x < 0) rgb_add(col, RGB{c.x, c.y, c.z});
                }
                if (gr.y == 1.0f) rgb_scale(col, gr.x);
            }
            if (rv.lights[l].geom == -1) continue;
            rgb_mul(col, colour);
            rgb_add(acc, col);
        }
        if (lobe == 0) rgb_scale(diff, dc.w);
    }
    RGB a = rgb0();
    rgb_add(a, RGB{e.x, e.y, e.z});
    rgb_add(a, diff);
    if (film) {   // no mirrors: contrib = ((0 + emission) + diffuse) + specular
        rgb_add(a, spec);
        film_update(film, sample_out, p, iter, a);
        return;
    }
    const float kind = rec->a.w;
    rec->a = make_float4(a.r, a.g, a.b, kind);
    if (kind == 1.0f) rec->s = make_float4(spec.r, spec.g, spec.b, rec->s.w);
}

// Folds the levels bottom-up (each mirror's colour is f * (the next level's colour), clamped at 0 and
// NaN, scaled by its weight; std.go:226-258) and updates the film (core/render.go:127-129).
__global__ __launch_bounds__(256) void k_combine(RenderView rv, uint32_t iter, const LevelRec* __restrict__ recs,
                                                 float* __restrict__ film, float* __restrict__ sample_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= rv.width * rv.height) return;
    const LevelRec* r = recs + (size_t)p * rv.rec_levels;
    RGB out = rgb0();
    bool next_hit = false;
    if (rv.path_depth > 0) {
        // path mode: ((0 + emission + diffuse) + specular direct) + max(0, f * next level's colour)
        for (int L = rv.max_level; L >= 0; L--) {
            const float4 a = r[L].a, sv = r[L].s, c = r[L].c;
            if (a.w == 0.0f) { next_hit = false; continue; }
            RGB o{a.x, a.y, a.z};
            rgb_add(o, RGB{sv.x, sv.y, sv.z});
            if (c.w != 0.0f && next_hit) {
                RGB col{c.x, c.y, c.z};
                rgb_mul(col, out);
                if (col.r < 0 || __builtin_isnan(col.r)) col.r = 0;
                if (col.g < 0 || __builtin_isnan(col.g)) col.g = 0;
                if (col.b < 0 || __builtin_isnan(col.b)) col.b = 0;
                rgb_add(o, col);
            }
            out = o;
            next_hit = true;
        }
        film_update(film, sample_out, p, iter, (r[0].a.w != 0.0f) ? out : rgb0());
        return;
    }
    for (int L = rv.max_level; L >= 0; L--) {
        const float4 a = r[L].a, sv = r[L].s;
        if (a.w == 0.0f) { next_hit = false; continue; }
        RGB S = rgb0();
        if (a.w == 1.0f) {
            S = RGB{sv.x, sv.y, sv.z};
        } else if (a.w == 2.0f && next_hit) {
            RGB col{sv.x, sv.y, sv.z};
            rgb_mul(col, out);
            if (col.r < 0 || __builtin_isnan(col.r)) col.r = 0;
            if (col.g < 0 || __builtin_isnan(col.g)) col.g = 0;
            if (col.b < 0 || __builtin_isnan(col.b)) col.b = 0;
            rgb_add(S, col);
            rgb_scale(S, sv.w / 1.0f);
        }
        RGB o{a.x, a.y, a.z};
        rgb_add(o, S);
        out = o;
        next_hit = true;
    }
    film_update(film, sample_out, p, iter, (r[0].a.w != 0.0f) ? out : rgb0());
}

}  // namespace

hipError_t launch_camera(const RenderView& rv, const uint64_t* scr, uint32_t iter, float4* rays, float4* pix,
                         hipStream_t stream) {
    const int n = rv.width * rv.height;
    hipLaunchKernelGGL(k_camera, dim3((n + 255) / 256), dim3(256), 0, stream, rv, scr, iter, rays, pix);
    return hipGetLastError();
}

hipError_t launch_shade(const RenderView& rv, const uint64_t* scr, uint32_t iter, int level, const float4* rays_in,
                        const int4* hits32, const float4* pix, float4* slot_rays, float4* slot_rgb, float4* groups,
                        float4* pixshade, float4* rays_out, LevelRec* recs, hipStream_t stream) {
    const int n = rv.width * rv.height;
    hipLaunchKernelGGL(k_shade, dim3((n + 127) / 128), dim3(128), 0, stream, rv, scr, iter, level, rays_in, hits32, pix,
                       slot_rays, slot_rgb, groups, pixshade, rays_out, recs);
    return hipGetLastError();
}

hipError_t launch_resolve(const RenderView& rv, int level, uint32_t iter, const float4* slot_rgb, const int4* slot_hits,
                          const float4* groups, const float4* pixshade, LevelRec* recs, float* film, float* sample_out,
                          hipStream_t stream) {
    const int n = rv.width * rv.height;
    hipLaunchKernelGGL(k_resolve, dim3((n + 255) / 256), dim3(256), 0, stream, rv, level, iter, slot_rgb, slot_hits,
                       groups, pixshade, recs, film, sample_out);
    return hipGetLastError();
}

hipError_t launch_combine(const RenderView& rv, uint32_t iter, const LevelRec* recs, float* film, float* sample_out,
                          hipStream_t stream) {
    const int n = rv.width * rv.height;
    hipLaunchKernelGGL(k_combine, dim3((n + 255) / 256), dim3(256), 0, stream, rv, iter, recs, film, sample_out);
    return hipGetLastError();
}

}  // namespace vgpu
