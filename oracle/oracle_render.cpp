// ORACLE — test infrastructure only (same scope rules as oracle_accel.cpp).
//
// CPU restatement of the reference's per-pixel render/shade loop in its
// reference semantics (direct lighting with MIS; indirect light only through perfect mirrors),
// and of path mode (set_path_depth; no reference counterpart): the same shading at every vertex
// of a path plus one continuation ray per vertex (Renderer::Continue, LevelScramble):
//   core/render.go:66-137          render: RasterXY / VdC / Sobol sample generation, film update
//   core/trace.go:40-83            Trace: ShaderContext setup, ApplyTransform, Shader.Eval
//   core/shader.go:139-402         OffsetP, LightsPrepare, NextLight, EvaluateLightSamples
//   core/ray.go:56-88              Ray.Init
//   builtin/camera/camera.go:80-323  PreRender (TanThetaFocal), calcLookatMatrices, ComputeRay (pinhole, thin lens)
//   math/sample/sample.go:105-129   UniformDisk2D
//   builtin/shader/std.go:77-316     ShaderStd.Eval / EvalEmission, perfect mirrors by recursion to level 3
//   builtin/shader/bsdf/orennayar.go:24-73, microfacetggx.go:26-146, specular.go:13-101
//   builtin/shader/fresnel/dielectric.go:34-47, conductor.go:31-89
//   builtin/light/triangle.go:64-535 Tri light: ValidSample, SampleArea, spherical triangle sampling
//   builtin/light/disk.go:38-215     Disk light; sphere.go:66-256 Sphere light
//   builtin/geom/sphere/trace.go:13-103  analytic sphere geometry (oracle_core.h)
//   builtin/filter/filter.go:40-173, gauss.go:34-58, airy.go:35-106  pixel filters
//   builtin/geom/polymesh/trace.go:292-513  hit attributes (N, Poffset, P, DdPdu, DdPdv)
//   colour/spectrum.go:49-112, spectrum_smits9.go:16-84, cie1931_2deg.go:62-95, space_srgb.go:8-14
//   math/ldseq/ldseq.go:50-96, raster.go:10-57   VanDerCorput, Sobol, RasterXY
//   math/sample/sample.go:18-27     CosineHemisphere
//   math/ferror.go:27-67            NextFloatUp / NextFloatDown
//   math/matrix4.go, quat.go, animdecomp.go   camera matrix decomposition/recomposition
//
// Float64 transcendentals go through the C library (the reference calls Go's
// math package, both are within 1 ulp of the true value; the float32 results
// they feed agree except on rare rounding ties).
//
// Not restated (absent from the workloads): camera motion keys, textures, quad lights.

#include "oracle_core.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

// --------------------------------------------------------------- colour

struct RGB {
    float c[3] = {0, 0, 0};
    void Add(const RGB& o) { for (int k = 0; k < 3; k++) c[k] += o.c[k]; }
    void Mul(const RGB& o) { for (int k = 0; k < 3; k++) c[k] *= o.c[k]; }
    void Scale(float s) { for (int k = 0; k < 3; k++) c[k] *= s; }
};

// CIE 1931 2-degree standard observer, 5 nm from 360 nm (95 samples): public CIE data, with the
// 715-830 nm tail as transcribed in colour/cie1931_2deg.go (6 significant digits; 715 nm reads 0.004106457).
static const float cieX[95] = {
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
static const float cieY[95] = {
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
static const float cieZ[95] = {
    0.0006061f, 0.001086f, 0.001946f, 0.003486f, 0.006450001f, 0.01054999f, 0.02005001f, 0.03621f, 0.06785001f, 0.1102f,
    0.2074f, 0.3713f, 0.6456f, 1.0390501f, 1.3856f, 1.62296f, 1.74706f, 1.7826f, 1.77211f, 1.7441f,
    1.6692f, 1.5281f, 1.28764f, 1.0419f, 0.8129501f, 0.6162f, 0.46518f, 0.3533f, 0.272f, 0.2123f,
    0.1582f, 0.1117f, 0.07824999f, 0.05725001f, 0.04216f, 0.02984f, 0.0203f, 0.0134f, 0.008749999f, 0.005749999f,
    0.0039f, 0.002749999f, 0.0021f, 0.0018f, 0.001650001f, 0.0014f, 0.0011f, 0.001f, 0.0008f, 0.0006f,
    0.00034f, 0.00024f, 0.00019f, 0.0001f, 0.00005f, 0.00003f, 0.00002f, 0.00001f, 0, 0,
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

// colour/cie1931_2deg.go:62-95
static float cieLookup(const float* tab, float lambda) {
    if (lambda < 360.0f || lambda >= 830.0f) return -1;
    int bin = (int)(((lambda - 360.0f) / (830.0f - 360.0f)) * 95.0f);
    return tab[bin];
}

// Smits 1999 RGB-to-spectrum basis spectra, 10 bins on [380,720) (colour/spectrum_smits9.go:34-42).
static const float smWhite[10] = {1, 1, 0.999f, 0.9993f, 0.9992f, 0.9998f, 1, 1, 1, 1};
static const float smCyan[10] = {0.9710f, 0.9426f, 1.0007f, 1.0007f, 1.0007f, 0.1564f, 0, 0, 0, 0};
static const float smMagenta[10] = {1, 1, 0.9685f, 0.2229f, 0, 0.0458f, 0.8369f, 1, 1, 0.9959f};
static const float smYellow[10] = {0.0001f, 0, 0.1088f, 0.6651f, 1, 1, 0.9996f, 0.9586f, 0.9685f, 0.9840f};
static const float smRed[10] = {0.1012f, 0.0515f, 0, 0, 0, 0, 0.8325f, 1.0149f, 1.0149f, 1.0149f};
static const float smGreen[10] = {0, 0, 0.0273f, 0.7937f, 1, 0.9418f, 0.1719f, 0, 0, 0.0025f};
static const float smBlue[10] = {1, 1, 0.8916f, 0.3323f, 0, 0, 0.0003f, 0.0369f, 0.0483f, 0.0496f};

static float smEval(const float* s, float lambda) {
    if (lambda < 380.0f || lambda >= 720.0f) return 0;
    int bin = (int)(((lambda - 380.0f) / (720.0f - 380.0f)) * 10.0f);
    return s[bin];
}

// colour/spectrum_smits9.go:48-84
static float Smits(float r, float g, float b, float lambda) {
    float c = 0;
    if (r <= g && r <= b) {
        c += r * smEval(smWhite, lambda);
        if (g <= b) { c += (g - r) * smEval(smCyan, lambda); c += (b - g) * smEval(smBlue, lambda); }
        else { c += (b - r) * smEval(smCyan, lambda); c += (g - b) * smEval(smGreen, lambda); }
    } else if (g <= r && g <= b) {
        c += g * smEval(smWhite, lambda);
        if (r <= b) { c += (r - g) * smEval(smMagenta, lambda); c += (b - r) * smEval(smBlue, lambda); }
        else { c += (b - g) * smEval(smMagenta, lambda); c += (r - b) * smEval(smRed, lambda); }
    } else {
        c += b * smEval(smWhite, lambda);
        if (r <= g) { c += (r - b) * smEval(smYellow, lambda); c += (g - r) * smEval(smGreen, lambda); }
        else { c += (g - b) * smEval(smYellow, lambda); c += (r - g) * smEval(smRed, lambda); }
    }
    return c;
}

// colour/spectrum.go:27-112 (hero wavelength, 4 samples over [450,750))
struct Spectrum {
    float C[4] = {0, 0, 0, 0};
    float Lambda = 0;
    float Wavelength(int j) const {
        float v = (Lambda - 450.0f + ((float)j / 4.0f) * 300.0f);
        if (v >= 300.0f) v -= 300.0f;
        v += 450.0f;
        return v;
    }
    void FromRGB(const RGB& rgb) { for (int k = 0; k < 4; k++) C[k] = Smits(rgb.c[0], rgb.c[1], rgb.c[2], Wavelength(k)); }
    RGB ToRGB() const {
        float x = 0, y = 0, z = 0;
        for (int i = 0; i < 4; i++) {
            x += C[i] * cieLookup(cieX, Wavelength(i));
            y += C[i] * cieLookup(cieY, Wavelength(i));
            z += C[i] * cieLookup(cieZ, Wavelength(i));
        }
        // colour/space_srgb.go:8-14, colourspace.go XYZToRGB
        RGB o;
        o.c[0] = x * 3.2404542f + y * -1.5371385f + z * -0.4985314f;
        o.c[1] = x * -0.9692660f + y * 1.8760108f + z * 0.0415560f;
        o.c[2] = x * 0.0556434f + y * -0.2040259f + z * 1.0572252f;
        return o;
    }
    void Mul(const Spectrum& o) { for (int k = 0; k < 4; k++) C[k] *= o.C[k]; }
    void Scale(float s) { for (int k = 0; k < 4; k++) C[k] *= s; }
};

// ------------------------------------------------------------- QMC (ldseq)

// math/ldseq/ldseq.go:55-75
static uint64_t vdcBits(uint64_t i, uint64_t scramble) {
    uint64_t bits = (i << 32) | (i >> 32);
    bits = ((bits & 0x0000ffff0000ffffull) << 16) | ((bits & 0xffff0000ffff0000ull) >> 16);
    bits = ((bits & 0x00ff00ff00ff00ffull) << 8) | ((bits & 0xff00ff00ff00ff00ull) >> 8);
    bits = ((bits & 0x0f0f0f0f0f0f0f0full) << 4) | ((bits & 0xf0f0f0f0f0f0f0f0ull) >> 4);
    bits = ((bits & 0x3333333333333333ull) << 2) | ((bits & 0xccccccccccccccccull) >> 2);
    bits = ((bits & 0x5555555555555555ull) << 1) | ((bits & 0xaaaaaaaaaaaaaaaaull) >> 1);
    return (scramble ^ bits) >> (64 - 52);
}
static double VanDerCorput(uint64_t i, uint64_t s) { return (double)vdcBits(i, s) / (double)(1ull << 52); }
// ldseq.go:83-96
static uint64_t sobolBits(uint64_t i, uint64_t scramble) {
    uint64_t r = scramble >> (64 - 52);
    for (uint64_t v = 1ull << 51; i != 0; i >>= 1) {
        if (i & 1) r ^= v;
        v ^= v >> 1;
    }
    return r;
}
static double Sobol(uint64_t i, uint64_t s) { return (double)sobolBits(i, s) / (double)(1ull << 52); }

// Enumeration matrices for RasterXY (Gruenschloss, Raab, Keller 2012), derived from the two
// generator matrices above: image of index bit k on the top m bits of (x, y), and the inverse of
// the 2m x 2m block of the low index bits.  (math/ldseq/vdc_sobol_matrices52.go holds the same
// tables as literals.)
struct RasterTables {
    uint64_t fwd[52];   // fwd[c]: image of index bit 2m+c, packed (px << m) | py
    uint64_t inv[52];   // inv[c]: low-index bits produced by bit c of the packed (px,py)
};
static uint64_t indexBitImage(int m, int k) {
    uint64_t x = vdcBits(1ull << k, 0) >> (52 - m);
    uint64_t y = sobolBits(1ull << k, 0) >> (52 - m);
    return (x << m) | y;
}
static RasterTables makeRasterTables(int m) {
    RasterTables t;
    std::memset(&t, 0, sizeof(t));
    const int n = 2 * m;
    for (int c = 0; c + n < 52; c++) t.fwd[c] = indexBitImage(m, n + c);
    // Gauss-Jordan over GF(2): columns A[k] = image of low index bit k (n x n)
    uint64_t A[52], I[52];
    for (int k = 0; k < n; k++) { A[k] = indexBitImage(m, k); I[k] = 1ull << k; }
    // solve for each unit vector e_c: find combination of columns giving e_c
    // Work on rows: build row-major matrix R[r] = bits over k of A[k] bit r.
    uint64_t R[52], Rinv[52];
    for (int r = 0; r < n; r++) {
        R[r] = 0;
        for (int k = 0; k < n; k++) if ((A[k] >> r) & 1) R[r] |= 1ull << k;
        Rinv[r] = 1ull << r;
    }
    for (int col = 0; col < n; col++) {
        int piv = -1;
        for (int r = col; r < n; r++) if ((R[r] >> col) & 1) { piv = r; break; }
        if (piv < 0) continue;
        std::swap(R[piv], R[col]); std::swap(Rinv[piv], Rinv[col]);
        for (int r = 0; r < n; r++)
            if (r != col && ((R[r] >> col) & 1)) { R[r] ^= R[col]; Rinv[r] ^= Rinv[col]; }
    }
    // R is now identity; Rinv = inverse (row r of inverse: which image bits give index bit r)
    for (int c = 0; c < n; c++) {
        uint64_t v = 0;
        for (int r = 0; r < n; r++) if ((Rinv[r] >> c) & 1) v |= 1ull << r;
        t.inv[c] = v;
    }
    (void)I;
    return t;
}
static const RasterTables& rasterTables(int m) {
    // built once, before any thread reads them (a function-local static is initialised thread-safely)
    struct All {
        RasterTables t[26];
        All() { for (int k = 1; k <= 25; k++) t[k] = makeRasterTables(k); }
    };
    static const All all;
    return all.t[m];
}
// math/ldseq/raster.go:10-40
static uint64_t rasterLookup(uint32_t m, uint32_t frame, uint32_t px, uint32_t py, uint64_t sx, uint64_t sy) {
    const RasterTables& T = rasterTables((int)m);
    uint32_t m2 = m << 1;
    uint64_t index = (uint64_t)frame << m2;
    uint64_t delta = 0;
    for (uint32_t c = 0; frame != 0; frame >>= 1) {
        if (frame & 1) delta ^= T.fwd[c];
        c++;
    }
    px ^= (uint32_t)(sx >> (64 - m));
    py ^= (uint32_t)(sy >> (64 - m));
    uint64_t b = (((uint64_t)px << m) | (uint64_t)py) ^ delta;
    for (uint32_t c = 0; b != 0; b >>= 1) {
        if (b & 1) index ^= T.inv[c];
        c++;
    }
    return index;
}
// raster.go:50-57
static void RasterXY(uint32_t m, uint32_t frame, uint32_t px, uint32_t py, uint64_t sx, uint64_t sy, uint64_t& index,
                     double& rx, double& ry) {
    index = rasterLookup(m, frame, px, py, sx, sy);
    rx = (double)vdcBits(index, sx) / (double)(1ull << (52 - m));
    ry = (double)sobolBits(index, sy) / (double)(1ull << (52 - m));
}

// ----------------------------------------------------------------- vectors

static inline Vec3 Sub(const Vec3& a, const Vec3& b) { return V3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
static inline Vec3 Add(const Vec3& a, const Vec3& b) { return V3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
static inline Vec3 Scale(float s, const Vec3& a) { return V3(a[0] * s, a[1] * s, a[2] * s); }
static inline Vec3 Neg(const Vec3& a) { return V3(-a[0], -a[1], -a[2]); }
static inline float Dot(const Vec3& a, const Vec3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline float DotAbs(const Vec3& a, const Vec3& b) { return Abs(a[0] * b[0] + a[1] * b[1] + a[2] * b[2]); }
static inline float Length2(const Vec3& a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }
static inline Vec3 Cross(const Vec3& a, const Vec3& b) {
    return V3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
static inline Vec3 Mad(const Vec3& a, const Vec3& b, float s) { return V3(a[0] + (b[0] * s), a[1] + (b[1] * s), a[2] + (b[2] * s)); }
static inline Vec3 Normalize(const Vec3& a) { return Mesh::Normalize(a); }
// math/vec3_amd64.s:45-56
static inline float Length(const Vec3& a) {
    float x2 = a[0] * a[0], y2 = a[1] * a[1], z2 = a[2] * a[2];
    return std::sqrt((y2 + x2) + z2);
}
static inline Vec3 BasisProject(const Vec3& U, const Vec3& V, const Vec3& W, const Vec3& S) { return V3(Dot(U, S), Dot(V, S), Dot(W, S)); }
static inline Vec3 BasisExpand(const Vec3& U, const Vec3& V, const Vec3& W, const Vec3& S) {
    return V3(U[0] * S[0] + V[0] * S[1] + W[0] * S[2], U[1] * S[0] + V[1] * S[1] + W[1] * S[2], U[2] * S[0] + V[2] * S[1] + W[2] * S[2]);
}
static const float Pi32 = 3.14159265358f;
static inline float sqr32(float x) { return x * x; }
static inline float fTan(float x) { return (float)std::tan((double)x); }
static inline float fAtan(float x) { return (float)std::atan((double)x); }
static inline float fAtan2(float y, float x) { return (float)std::atan2((double)y, (double)x); }

// math/ferror.go:27-67
static float NextFloatUp(float v) {
    if (std::isinf(v) && v > 0) return v;
    if (v == 0.0f) v = 0.0f;
    uint32_t ui = f2bits(v);
    if (v >= 0.0f) ui++; else ui--;
    return bits2f(ui);
}
static float NextFloatDown(float v) {
    if (std::isinf(v) && v < 0) return v;
    if (v == 0.0f) v = 0.0f;
    uint32_t ui = f2bits(v);
    if (v >= 0.0f) ui--; else ui++;
    return bits2f(ui);
}

// --------------------------------------------------------------- materials

struct Material {              // ShaderStd parameters (builtin/shader/std.go:25-47); has_* = parameter set
    float emission[3]; float emission_strength; int32_t has_emission_colour, has_emission_strength;
    float diffuse[3]; int32_t has_diffuse_colour;
    float diffuse_strength; int32_t has_diffuse_strength;
    float diffuse_roughness; int32_t has_diffuse_roughness;
    float spec1[3]; int32_t has_spec1_colour;
    float spec1_strength; int32_t has_spec1_strength;
    float spec1_roughness; int32_t has_spec1_roughness;
    float ior; int32_t has_ior;
    int32_t fresnel_model;     // 0 / 1 = Dielectric, 2 = Metal (conductor)
    float spec1_fresnel_refl[3]; int32_t has_spec1_fresnel_refl;
    float spec1_fresnel_edge[3]; int32_t has_spec1_fresnel_edge;
};

// A light node: Tri (builtin/light/triangle.go:18-28), Disk (disk.go:17-31: P0 = P, P1 = Up,
// P2 = LookAt) or Sphere (sphere.go:16-28: P0 = centre).
struct Light {
    int32_t type;              // 0 tri, 1 disk, 2 sphere
    Vec3 P0, P1, P2;
    float radius;
    int32_t material;
    int32_t samples;
    int32_t geom;              // post-sort geom slot of the light's geometry
};

// ----------------------------------------------------------------- shading

struct SC {                    // the subset of core.ShaderContext the shading path reads
    int32_t geom;
    Vec3 P, Poffset, N, Ng, DdPdu, DdPdv, Rd;
    float Lambda;
    float Time;                // sample time, carried by every ray the hit spawns (core/ray.go:65)
    uint64_t I;
    uint64_t Scramble[2];      // the scrambles EvaluateLightSamples reads at this vertex
    uint64_t PixScramble[2];   // the pixel's scramble0/1 (path mode derives its other dimensions from them)
    int32_t shaderIdx;
};

// Path mode (not in the reference): a continuation ray's origin moved off the surface along the unit
// geometric normal n by 256 ulps per component, or by 2^-16 n within 1/32 of zero (Waechter and Binder,
// "A Fast and Robust Method for Avoiding Self-Intersection", Ray Tracing Gems, 2019).  OffsetP's error bound
// is zero for a surface lying exactly in a coordinate plane through the origin, where a continuation ray
// would otherwise find its own surface again at t ~ 0.
static float OffsetAlong1(float p, float n) {
    const int32_t of = (int32_t)(256.0f * n);
    int32_t bits;
    std::memcpy(&bits, &p, 4);
    bits += p < 0.0f ? -of : of;
    float pi;
    std::memcpy(&pi, &bits, 4);
    return std::fabs(p) < 0.03125f ? p + 1.52587891e-05f * n : pi;
}
static Vec3 OffsetAlong(const Vec3& p, const Vec3& n) {
    return V3(OffsetAlong1(p[0], n[0]), OffsetAlong1(p[1], n[1]), OffsetAlong1(p[2], n[2]));
}

// Path mode (not in the reference): the scramble of dimension pair k at recursion level `level`, the
// SplitMix64 finaliser (Steele, Lea, Flood 2014) of a pixel scramble xor a per-(level, k) constant.
static uint64_t LevelScramble(uint64_t s, int level, int k) {
    uint64_t z = s ^ ((uint64_t)(8 * level + k + 1) * 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct OrenNayar {             // builtin/shader/bsdf/orennayar.go
    float Lambda; Vec3 OmegaI; float Roughness; Vec3 U, V, N;
    Vec3 Sample(double r0, double r1) const {
        double r = std::sqrt(1 - r0), theta = 2 * M_PI * r1;   // sample.go:18-27
        Vec3 s = V3((float)(r * std::cos(theta)), (float)(r * std::sin(theta)), (float)std::sqrt(r0));
        return BasisExpand(U, V, N, s);
    }
    double PDF(const Vec3& o_) const {
        Vec3 o = BasisProject(U, V, N, o_);
        double odn = (double)Max(0, o[2]);
        return odn / M_PI;
    }
    Spectrum Eval(const Vec3& o_) const {
        Vec3 o = BasisProject(U, V, N, o_);
        float sigma = Roughness;
        float A = 1 - (0.5f * (sigma * sigma) / ((sigma * sigma) + 0.57f));
        float B = 0.45f * (sigma * sigma) / ((sigma * sigma) + 0.09f);
        float phiI = fAtan2(OmegaI[1], OmegaI[0]);
        float phiO = fAtan2(o[1], o[0]);
        float thetaI = fAcos(OmegaI[2]);
        float thetaO = fAcos(o[2]);
        float alpha = Max(thetaI, thetaO);
        float beta = Min(thetaI, thetaO);
        float C = fSin(alpha) * fTan(beta);
        float gamma = fCos(phiO - phiI);
        float scale = o[2] * (A + (B * Max(0, gamma) * C));
        Spectrum rho; rho.Lambda = Lambda;
        RGB one; one.c[0] = one.c[1] = one.c[2] = 1;
        rho.FromRGB(one);
        rho.Scale(scale / Pi32);
        return rho;
    }
};

static float chi(float x) { return x > 0.0f ? 1 : 0; }
static float ggxG1(const Vec3& w, const Vec3& m, float alpha) {
    float odn = Dot(w, m);
    float denom = 1 + std::sqrt(1 + (alpha * alpha) * ((1.0f / (w[2] * w[2])) - 1));
    return chi(odn / w[2]) * 2 / denom;
}
static float ggxD(const Vec3& m, float alpha) {
    float numer = alpha * alpha * chi(m[2]);
    if (m[2] == 1.0f) return 1.0f / (Pi32 * alpha * alpha);
    float denom = Pi32 * sqr32(m[2] * m[2]) * sqr32(alpha * alpha + ((1.0f / (m[2] * m[2])) - 1));
    return numer / denom;
}
static float signf(float v) { return v < 0 ? -1 : 1; }
// builtin/shader/fresnel/dielectric.go:34-47
static RGB DielectricKr(float eta, float c) {
    RGB o;
    float g = (eta * eta) - 1 + (c * c);
    if (g < 0.0f) { o.c[0] = o.c[1] = o.c[2] = 1; return o; }
    g = std::sqrt(g);
    float fr = 0.5f * sqr32((g - c) / (g + c)) * (1 + sqr32((c * (g + c) - 1) / (c * (g - c) + 1)));
    o.c[0] = o.c[1] = o.c[2] = fr;
    return o;
}
// builtin/shader/fresnel/conductor.go:31-81 ("Artist Friendly Metallic Fresnel", Gulbrandsen 2014):
// refractive index n and extinction k^2 from reflectivity r and edge tint g, then the unpolarised
// conductor Fresnel term.  Min/Max are MINSS/MAXSS (math/dim_amd64.s).
static float mins32(float x, float y) { return x < y ? x : y; }
static float maxs32(float x, float y) { return x > y ? x : y; }
static float condNMin(float r) { return (1 - r) / (1 + r); }
static float condNMax(float r) { return (1 + std::sqrt(r)) / (1 - std::sqrt(r)); }
static float ConductorFresnel(float r, float g, float cosTheta) {
    float nr = maxs32(0, mins32(r, 0.99f));
    float n = condNMin(nr) * g + (1 - g) * condNMax(nr);
    float k2 = ((n + 1) * (n + 1) * nr - (n - 1) * (n - 1)) / (1 - nr);
    float rs = (n * n + k2 - 2 * n * cosTheta + cosTheta * cosTheta) / (n * n + k2 + 2 * n * cosTheta + cosTheta * cosTheta);
    float rp = ((n * n + k2) * cosTheta * cosTheta - 2 * n * cosTheta + 1) / ((n * n + k2) * cosTheta * cosTheta + 2 * n * cosTheta + 1);
    return 0.5f * (rs + rp);
}
// The core.Fresnel a ShaderStd builds (std.go:171-190): Dielectric(ior), or Conductor(refl, edge).
struct FresnelModel {
    bool metal = false; float eta = 1.7f; float refl[3] = {0.5f, 0.5f, 0.5f}, edge[3] = {0.5f, 0.5f, 0.5f};
    RGB Kr(float c) const {
        if (!metal) return DielectricKr(eta, c);
        RGB o;
        for (int k = 0; k < 3; k++) o.c[k] = ConductorFresnel(refl[k], edge[k], c);
        return o;
    }
};
static FresnelModel MakeFresnel(const Material& sh, float ior) {
    FresnelModel f;
    f.eta = ior;
    if (sh.fresnel_model == 2) {
        f.metal = true;
        if (sh.has_spec1_fresnel_refl) for (int k = 0; k < 3; k++) f.refl[k] = sh.spec1_fresnel_refl[k];
        if (sh.has_spec1_fresnel_edge) for (int k = 0; k < 3; k++) f.refl[k] = sh.spec1_fresnel_edge[k];   // std.go:185-187
    }
    return f;
}
struct GGX {                   // builtin/shader/bsdf/microfacetggx.go
    float Lambda; Vec3 OmegaR; float Roughness; FresnelModel fresnel; Vec3 U, V, N;
    Vec3 Sample(double r0, double r1) const {
        float alpha = sqr32(Roughness);
        double thetaM = std::atan2((double)alpha * std::sqrt(r0), std::sqrt(1 - r0));
        double phiM = 2.0 * M_PI * r1;
        Vec3 m = V3(fSin((float)thetaM) * fCos((float)phiM), fSin((float)thetaM) * fSin((float)phiM), fCos((float)thetaM));
        Vec3 o = Sub(Scale(2.0f * DotAbs(m, OmegaR), m), OmegaR);
        return BasisExpand(U, V, N, Normalize(o));
    }
    double PDF(const Vec3& o_) const {
        Vec3 o = BasisProject(U, V, N, o_);
        float alpha = sqr32(Roughness);
        Vec3 m = Scale(signf(OmegaR[2]), Normalize(Add(OmegaR, o)));
        double pdf = (double)(ggxD(m, alpha) * m[2]);
        if (std::isnan(pdf)) return 0;
        return pdf;
    }
    // Path mode: density of the direction Sample returns.  PDF above is the half vector's density
    // D(m) m_z (microfacetggx.go:108-122); the reflection about m divides it by 4 |omegaR . m|.
    double PDFDir(const Vec3& o_) const {
        Vec3 o = BasisProject(U, V, N, o_);
        float alpha = sqr32(Roughness);
        Vec3 m = Scale(signf(OmegaR[2]), Normalize(Add(OmegaR, o)));
        double pdf = (double)(ggxD(m, alpha) * m[2] / (4.0f * DotAbs(OmegaR, m)));
        if (std::isnan(pdf)) return 0;
        return pdf;
    }
    Spectrum Eval(const Vec3& o_) const {
        Vec3 wi = BasisProject(U, V, N, o_);
        float alpha = sqr32(Roughness);
        Vec3 h = Scale(signf(OmegaR[2]), Normalize(Add(OmegaR, wi)));
        RGB F = fresnel.Kr(DotAbs(OmegaR, h));
        float numer = ggxG1(OmegaR, h, alpha) * ggxG1(wi, h, alpha) * ggxD(h, alpha);
        float denom = 4 * Abs(OmegaR[2]) * Abs(wi[2]);
        Spectrum rho; rho.Lambda = Lambda;
        rho.FromRGB(F);
        rho.Scale(Abs(wi[2]) * numer / denom);
        for (int k = 0; k < 4; k++) if (rho.C[k] < 0 || std::isnan(rho.C[k])) rho.C[k] = 0;
        return rho;
    }
};

// builtin/shader/bsdf/specular.go: perfect mirror; reflect(omegaR, N) = 2 (N . omegaR) N - omegaR
static Vec3 ReflectZ(const Vec3& wr) {
    Vec3 N = V3(0, 0, 1);
    return Sub(Scale(2.0f * Dot(N, wr), N), wr);
}
struct Specular {
    float Lambda; Vec3 OmegaR; FresnelModel fresnel; Vec3 U, V, N;
    Vec3 Sample() const { return BasisExpand(U, V, N, Normalize(ReflectZ(OmegaR))); }
    double PDF(const Vec3& o_) const {
        Vec3 o = BasisProject(U, V, N, o_);
        return Dot(o, ReflectZ(OmegaR)) < 0.9999f ? 0 : 1;
    }
    Spectrum Eval(const Vec3& o_) const {
        Vec3 o = BasisProject(U, V, N, o_);
        Spectrum rho;
        if (Dot(o, ReflectZ(OmegaR)) < 0.9999f) return rho;
        rho.Lambda = Lambda;
        rho.FromRGB(fresnel.Kr(OmegaR[2]));
        rho.Scale(DotAbs(o, V3(0, 0, 1)));
        return rho;
    }
};

struct LightSample { float Pdf; Spectrum Liu; Vec3 Ld; float Ldist; };
struct BSDFSample { Vec3 D; double Pdf; float PdfLight; Spectrum Liu; Vec3 Ld; float Ldist; };

// builtin/light/triangle.go helpers
static bool rayTriangleIntersect(const Vec3& Ro, const Vec3& Rd, const Vec3& P0, const Vec3& P1, const Vec3& P2, Vec3& pout) {
    Vec3 e1 = Sub(P1, P0), e2 = Sub(P2, P0);
    Vec3 P = Cross(Rd, e2);
    float det = Dot(e1, P);
    if (det > -1e-6f && det < 1e-6f) return false;
    float inv_det = 1 / det;
    Vec3 T = Sub(Ro, P0);
    float u = Dot(T, P) * inv_det;
    if (u < 0 || u > 1) return false;
    Vec3 Q = Cross(T, e1);
    float v = Dot(Rd, Q) * inv_det;
    if (v < 0 || u + v > 1) return false;
    float t = Dot(e2, Q) * inv_det;
    if (t > 1e-6f) {
        // Vec3Add3(Scale(1-u-v, P0), Scale(u, P1), Scale(v, P2))
        Vec3 a = Scale(1 - u - v, P0), b = Scale(u, P1), c = Scale(v, P2);
        pout = V3(a[0] + b[0] + c[0], a[1] + b[1] + c[1], a[2] + b[2] + c[2]);
        return true;
    }
    return false;
}
static float triangleArea(const Vec3& P0, const Vec3& P1, const Vec3& P2) { return 0.5f * Length(Cross(Sub(P1, P0), Sub(P2, P0))); }
static void sidesToAngles(float as, float bs, float cs, float& a, float& b, float& c) {
    float ssu = (as + bs + cs) / 2;
    float sa = fSin(ssu - as), sb = fSin(ssu - bs), scc = fSin(ssu - cs), s = fSin(ssu);
    float tA = std::sqrt(sb * scc / (s * sa));
    float tB = std::sqrt(sa * scc / (s * sb));
    float tC = std::sqrt(sa * sb / (s * scc));
    a = 2 * fAtan(tA); b = 2 * fAtan(tB); c = 2 * fAtan(tC);
}
static void verticesToSides(const Vec3& pa, const Vec3& pb, const Vec3& pc, float& a, float& b, float& c) {
    a = fAcos(Dot(pb, pc)); b = fAcos(Dot(pc, pa)); c = fAcos(Dot(pa, pb));
}
static Vec3 sampleSphericalTriangle(const Vec3& p0, const Vec3& p1, const Vec3& p2, const Vec3& p, double r0, double r1, double& pdf) {
    Vec3 pa = Normalize(Sub(p0, p)), pb = Normalize(Sub(p1, p)), pc = Normalize(Sub(p2, p));
    float a, b, c; verticesToSides(pa, pb, pc, a, b, c);
    float alpha, beta, gamma; sidesToAngles(a, b, c, alpha, beta, gamma);
    float area = alpha + beta + gamma - Pi32;
    float areaHat = (float)r0 * area;
    float s = fSin(areaHat - alpha), t = fCos(areaHat - alpha);
    float sinAlpha = fSin(alpha), cosAlpha = fCos(alpha);
    float u = t - cosAlpha;
    float v = s + sinAlpha * fCos(c);
    float q = ((v * t - u * s) * cosAlpha - v) / ((v * s + u * t) * sinAlpha);
    q = Max(-1, Min(q, 1));
    float w = Dot(pc, pa);
    Vec3 v31; for (int k = 0; k < 3; k++) v31[k] = pc[k] - w * pa[k];
    v31 = Normalize(v31);
    Vec3 v4; for (int k = 0; k < 3; k++) v4[k] = q * pa[k] + std::sqrt(1 - q * q) * v31[k];
    float z = 1 - (float)r1 * (1 - Dot(v4, pb));
    w = Dot(v4, pb);
    Vec3 v42; for (int k = 0; k < 3; k++) v42[k] = v4[k] - w * pb[k];
    v42 = Normalize(v42);
    Vec3 x = Add(Scale(z, pb), Scale(std::sqrt(1 - z * z), v42));
    pdf = 1 / (double)area;
    return x;
}
static bool rayPlaneIntersect(const Vec3& Ro, const Vec3& Rd, const Vec3& P, const Vec3& N, float& t) {
    float denom = Dot(N, Rd);
    if (Abs(denom) > 1e-6f) { Vec3 p0l0 = Sub(P, Ro); t = Dot(p0l0, N) / denom; return t >= 0; }
    t = 0;
    return false;
}

// builtin/filter/airy.go:35-79 (Numerical Recipes bessj1, as the reference copies it from gipsy bessel.c)
static double BesselJ1(double x) {
    double ax = std::fabs(x);
    if (ax < 8.0) {
        double y = x * x;
        double ans1 = x * (72362614232.0 + y * (-7895059235.0 + y * (242396853.1 + y * (-2972611.439 +
                      y * (15704.48260 + y * (-30.16036606))))));
        double ans2 = 144725228442.0 + y * (2300535178.0 + y * (18583304.74 + y * (99447.43394 +
                      y * (376.9991397 + y * 1.0))));
        return ans1 / ans2;
    }
    double z = 8.0 / ax, y = z * z, xx = ax - 2.356194491;
    double ans1 = 1.0 + y * (0.183105e-2 + y * (-0.3516396496e-4 + y * (0.2457520174e-5 + y * (-0.240337019e-6))));
    double ans2 = 0.04687499995 + y * (-0.2002690873e-3 + y * (0.8449199096e-5 + y * (-0.88228987e-6 + y * 0.105787412e-6)));
    double ans = std::sqrt(0.636619772 / ax) * (std::cos(xx) * ans1 - z * std::sin(xx) * ans2);
    if (x < 0.0) ans = -ans;
    return ans;
}
// filter.CreateSampler (filter.go:87-173); out = cdfV then cdfVU
template <class F>
static void CreateSampler(int n, double w, F f, std::vector<double>& out) {
    std::vector<double> filter((size_t)n * n);
    double du = w / (double)(n - 1), u = -w / 2, Fsum = 0;
    for (int j = 0; j < n; j++) {
        double dv = w / (double)(n - 1), v = -w / 2;
        for (int i = 0; i < n; i++) {
            double fuv = f(u, v);
            filter[(size_t)j + (size_t)i * n] = fuv;
            Fsum += fuv;
            v += dv;
        }
        u += du;
    }
    std::vector<std::vector<double>> pdf(n, std::vector<double>(n)), pVU(n, std::vector<double>(n));
    std::vector<double> pV(n, 0.0);
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) pdf[j][i] = filter[(size_t)j + (size_t)i * n] / Fsum;
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) pV[j] += pdf[j][i];
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) pVU[j][i] = pdf[j][i] / pV[j];
    out.assign((size_t)n + (size_t)n * n, 0.0);
    double p = 0;
    for (int i = 0; i < n; i++) { p += pV[i]; out[i] = p; }
    for (int j = 0; j < n; j++) {
        double q = 0;
        for (int i = 0; i < n; i++) { q += pVU[j][i]; out[(size_t)n + (size_t)j * n + i] = q; }
    }
}

struct Renderer {
    Scene* scene;
    int width, height;
    M4 M;
    std::vector<Decomp> camKeys;  // camera.decomp with two or more motion keys; empty: M for every sample
    float tanThetaFocal, aspect, focal;
    std::vector<Material> mats;
    std::vector<Light> lights;
    std::vector<double> filt;     // pixel filter sampler: cdfV (filtN) then cdfVU (filtN x filtN); empty: box
    int filtN = 0;
    double filtW = 0;
    float lensRadius = 0;         // camera Radius (thin lens when > 0)
    std::vector<std::vector<int32_t>> geomShaders;   // per post-sort geom slot: material id per shader index
    int shadowRays = 0;
    int pathDepth = 0;            // path mode (not in the reference): segments per path; 0: the reference's shading

    // EvalEmission (std.go:299-316)
    RGB Emission(const Material& m) const {
        RGB e;
        if (m.has_emission_colour) for (int k = 0; k < 3; k++) e.c[k] = m.emission[k];
        if (!m.has_emission_strength) return RGB();
        e.Scale(m.emission_strength);
        return e;
    }

    bool TraceShadow(Task* tk, const Vec3& P, const Vec3& D, float tmax, float time) const {
        Ray r{};
        r.P = P; r.D = D; r.Tclosest = tmax; r.Type = RayTypeShadow; r.Time = time;
        RaySetup(r);
        tk->StackTop = 0; r.task = tk;
        SceneHit sh{}; sh.geom = -1;
        return scene->Trace(r, sh);
    }

    // core/shader.go:139-163
    static Vec3 OffsetP(const SC& sc, int dir) {
        Vec3 pofs = dir < 0 ? Neg(sc.Poffset) : sc.Poffset;
        Vec3 po = Add(sc.P, pofs);
        for (int i = 0; i < 3; i++) {
            if (pofs[i] > 0) po[i] = NextFloatUp(po[i]);
            else if (pofs[i] < 0) po[i] = NextFloatDown(po[i]);
        }
        return po;
    }

    // Tri.ValidSample (triangle.go:136-230)
    bool TriValidSample(const Light& L, const SC& sg, BSDFSample& s) const {
        Vec3 P;
        if (!rayTriangleIntersect(sg.P, s.D, L.P0, L.P1, L.P2, P)) return false;
        Vec3 N = Normalize(Cross(Sub(L.P1, L.P0), Sub(L.P2, L.P0)));
        const Material& lm = mats[L.material];
        if (Dot(sg.Ng, Sub(L.P0, P)) < 0 || Dot(sg.Ng, Sub(L.P1, P)) < 0 || Dot(sg.Ng, Sub(L.P2, P)) < 0) {
            double pdf = (double)(1 / triangleArea(L.P0, L.P1, L.P2));
            Vec3 D = Sub(P, sg.P);
            s.Ldist = Length(D); s.Ld = Normalize(D);
            if (Dot(s.Ld, N) > 0 || Dot(s.Ld, sg.Ng) < 0) return false;
            s.Liu.Lambda = sg.Lambda; s.Liu.FromRGB(Emission(lm));
            s.PdfLight = (float)pdf * sqr32(s.Ldist) / DotAbs(s.Ld, N);
        } else {
            Vec3 pa = Normalize(Sub(L.P0, sg.P)), pb = Normalize(Sub(L.P1, sg.P)), pc = Normalize(Sub(L.P2, sg.P));
            float a, b, c; verticesToSides(pa, pb, pc, a, b, c);
            float al, be, ga; sidesToAngles(a, b, c, al, be, ga);
            float area = al + be + ga - Pi32;
            double pdf = (double)(1 / area);
            Vec3 D = Sub(P, sg.P);
            s.Ldist = Length(D); s.Ld = Normalize(D);
            if (Dot(s.Ld, N) > 0 || Dot(s.Ld, sg.Ng) < 0) return false;
            s.Liu.Lambda = sg.Lambda; s.Liu.FromRGB(Emission(lm));
            s.PdfLight = (float)pdf;
        }
        return true;
    }

    // Tri.SampleArea (triangle.go:232-343)
    void TriSampleArea(const Light& L, const SC& sg, int n, std::vector<LightSample>& out) const {
        const Material& lm = mats[L.material];
        Vec3 N = Normalize(Cross(Sub(L.P1, L.P0), Sub(L.P2, L.P0)));
        if (Dot(sg.Ng, Sub(L.P0, sg.P)) < 0 || Dot(sg.Ng, Sub(L.P1, sg.P)) < 0 || Dot(sg.Ng, Sub(L.P2, sg.P)) < 0) {
            double pdf = (double)(1 / triangleArea(L.P0, L.P1, L.P2));
            for (int i = 0; i < n; i++) {
                uint64_t idx = sg.I * (uint64_t)n + (uint64_t)i;
                double r0 = VanDerCorput(idx, sg.Scramble[0]), r1 = Sobol(idx, sg.Scramble[1]);
                Vec3 a = Scale((float)(r1 * std::sqrt(1 - r0)), Sub(L.P1, L.P0));
                Vec3 b = Scale((float)(1 - std::sqrt(1 - r0)), Sub(L.P2, L.P0));
                Vec3 P = V3(L.P0[0] + a[0] + b[0], L.P0[1] + a[1] + b[1], L.P0[2] + a[2] + b[2]);
                Vec3 D = Sub(P, sg.P);
                LightSample ls;
                ls.Ldist = Length(D); ls.Ld = Normalize(D);
                if (Dot(ls.Ld, N) > 0 || Dot(ls.Ld, sg.Ng) < 0) continue;
                ls.Liu.Lambda = sg.Lambda; ls.Liu.FromRGB(Emission(lm));
                ls.Pdf = (float)pdf * sqr32(ls.Ldist) / DotAbs(ls.Ld, N);
                out.push_back(ls);
            }
            return;
        }
        for (int i = 0; i < n; i++) {
            uint64_t idx = sg.I * (uint64_t)n + (uint64_t)i;
            double r0 = VanDerCorput(idx, sg.Scramble[0]), r1 = Sobol(idx, sg.Scramble[1]);
            double pdf;
            Vec3 x = sampleSphericalTriangle(L.P0, L.P1, L.P2, sg.P, r0, r1, pdf);
            float t; rayPlaneIntersect(sg.P, x, L.P0, N, t);
            Vec3 P = Mad(sg.P, x, t);
            Vec3 D = Sub(P, sg.P);
            LightSample ls;
            ls.Ldist = Length(D); ls.Ld = Normalize(D);
            if (Dot(ls.Ld, N) > 0 || Dot(ls.Ld, sg.Ng) < 0) continue;
            ls.Liu.Lambda = sg.Lambda; ls.Liu.FromRGB(Emission(lm));
            ls.Pdf = (float)pdf;
            out.push_back(ls);
        }
    }

    // Disk.PreRender's frame (disk.go:84-86)
    static void DiskFrame(const Light& L, Vec3& N, Vec3& T, Vec3& B) {
        N = Normalize(Sub(L.P2, L.P0));
        T = Normalize(Cross(N, L.P1));
        B = Cross(N, T);
    }
    // Disk.ValidSample (disk.go:120-168) with rayDiskIntersect (disk.go:54-68)
    bool DiskValidSample(const Light& L, const SC& sg, BSDFSample& s) const {
        Vec3 N, T, B; DiskFrame(L, N, T, B);
        const float R = L.radius;
        double pdf = (double)(1.0f / (Pi32 * R * R));
        float t;
        if (!rayPlaneIntersect(sg.P, s.D, L.P0, N, t) || !(t >= 0)) return false;
        {
            Vec3 v = Sub(Mad(sg.P, s.D, t), L.P0);
            if (!(std::sqrt(Dot(v, v)) <= R)) return false;
        }
        Vec3 p = Mad(sg.P, s.D, t);
        Vec3 V = Sub(p, sg.P);
        if (Dot(V, sg.Ng) <= 0.0f || Dot(V, N) >= 0.0f) return false;
        s.Ldist = Length(V);
        s.Ld = Normalize(V);
        s.Liu.Lambda = sg.Lambda; s.Liu.FromRGB(Emission(mats[L.material]));
        s.PdfLight = (float)pdf * (s.Ldist * s.Ldist) / Abs(Dot(s.Ld, N));
        return true;
    }
    // Disk.SampleArea (disk.go:171-215)
    void DiskSampleArea(const Light& L, const SC& sg, int n, std::vector<LightSample>& out) const {
        Vec3 N, T, B; DiskFrame(L, N, T, B);
        const float R = L.radius;
        for (int i = 0; i < n; i++) {
            uint64_t idx = sg.I * (uint64_t)n + (uint64_t)i;
            double r0 = VanDerCorput(idx, sg.Scramble[0]), r1 = Sobol(idx, sg.Scramble[1]);
            float u = R * std::sqrt((float)r0) * fCos(2 * Pi32 * (float)r1);
            float v = R * std::sqrt((float)r0) * fSin(2 * Pi32 * (float)r1);
            double pdf = (double)(1.0f / (Pi32 * R * R));
            Vec3 su = Scale(u, B), sv = Scale(v, T);
            Vec3 P = V3(L.P0[0] + su[0] + sv[0], L.P0[1] + su[1] + sv[1], L.P0[2] + su[2] + sv[2]);
            Vec3 V = Sub(P, sg.P);
            if (!(Dot(V, sg.Ng) > 0.0f && Dot(V, N) < 0.0f)) continue;
            LightSample ls;
            ls.Ldist = Length(V);
            ls.Ld = Normalize(V);
            ls.Liu.Lambda = sg.Lambda; ls.Liu.FromRGB(Emission(mats[L.material]));
            ls.Pdf = (float)pdf * (ls.Ldist * ls.Ldist) / Abs(Dot(ls.Ld, N));
            out.push_back(ls);
        }
    }
    static float ConePdf(float R, float l) { return 1 / (2 * Pi32 * (1 - std::sqrt(1 - sqr32(R / l)))); }
    // Sphere.ValidSample (sphere.go:129-177)
    bool SphereValidSample(const Light& L, const SC& sg, BSDFSample& s) const {
        float t;
        if (!Scene::RaySphereIntersect(sg.P, s.D, L.P0, L.radius, t)) return false;
        Vec3 x = Mad(sg.P, s.D, t);
        Vec3 D = Sub(x, sg.P);
        s.Ldist = Length(D);
        s.Ld = Normalize(D);
        s.Liu.Lambda = sg.Lambda; s.Liu.FromRGB(Emission(mats[L.material]));
        float l = Length(Sub(L.P0, sg.P));
        s.PdfLight = ConePdf(L.radius, l);
        return true;
    }
    // Sphere.SampleArea (sphere.go:180-256), sg.Sample = 0
    void SphereSampleArea(const Light& L, const SC& sg, int n, std::vector<LightSample>& out) const {
        Vec3 V = Sub(L.P0, sg.P);
        float l = Length(V);
        Vec3 w = Normalize(V);
        Vec3 v = Normalize(Cross(w, sg.Ng));
        Vec3 u = Cross(w, v);
        for (int i = 0; i < n; i++) {
            uint64_t idx = sg.I * (uint64_t)n + (uint64_t)i;
            double r0 = VanDerCorput(idx, sg.Scramble[0]), r1 = Sobol(idx, sg.Scramble[1]);
            float theta = fAcos(1 - (float)r0 + (float)r0 * std::sqrt(1 - sqr32(L.radius / l)));
            float phi = 2 * Pi32 * (float)r1;
            Vec3 a = V3(fCos(phi) * fSin(theta), fSin(phi) * fSin(theta), fCos(theta));
            Vec3 omega = BasisExpand(u, v, w, a);
            if (Dot(omega, sg.Ng) < 0) continue;
            float t;
            if (!Scene::RaySphereIntersect(sg.P, omega, L.P0, L.radius, t)) continue;
            Vec3 x = Mad(sg.P, omega, t);
            Vec3 D = Sub(x, sg.P);
            LightSample ls;
            ls.Ldist = Length(D);
            ls.Ld = Normalize(D);
            ls.Liu.Lambda = sg.Lambda; ls.Liu.FromRGB(Emission(mats[L.material]));
            ls.Pdf = ConePdf(L.radius, l);
            out.push_back(ls);
        }
    }
    bool ValidSample(const Light& L, const SC& sg, BSDFSample& s) const {
        if (L.type == 1) return DiskValidSample(L, sg, s);
        if (L.type == 2) return SphereValidSample(L, sg, s);
        return TriValidSample(L, sg, s);
    }
    void SampleArea(const Light& L, const SC& sg, int n, std::vector<LightSample>& out) const {
        if (L.type == 1) DiskSampleArea(L, sg, n, out);
        else if (L.type == 2) SphereSampleArea(L, sg, n, out);
        else TriSampleArea(L, sg, n, out);
    }

    // core/shader.go:203-402
    template <class B>
    RGB EvaluateLightSamples(Task* tk, const SC& sc, const Light& L, int NSamples, const B& bsdf) {
        RGB col;
        const float k = 1.0f - 0.0001f;   // 1 - ShadowRayEpsilon (core/ray.go:22)
        if (NSamples > 1) {
            int half = NSamples / 2;
            std::vector<BSDFSample> bs;
            for (int i = 0; i < half; i++) {
                uint64_t idx = sc.I * (uint64_t)half + (uint64_t)i;
                double r0 = VanDerCorput(idx, sc.Scramble[0]), r1 = Sobol(idx, sc.Scramble[1]);
                Vec3 o = Normalize(bsdf.Sample(r0, r1));
                BSDFSample s{}; s.D = o; s.Pdf = bsdf.PDF(o);
                if (s.Pdf <= 0) continue;
                if (ValidSample(L, sc, s)) bs.push_back(s);
            }
            std::vector<LightSample> ls;
            SampleArea(L, sc, half, ls);
            int nB = half, nL = half;
            if (ls.empty()) nL = 0;
            if (bs.empty()) nB = 0;
            int total = nB + nL;
            if (total == 0) return RGB();
            for (const auto& l : ls) {
                if (Dot(l.Ld, sc.N) <= 0) continue;
                Vec3 o = Dot(l.Ld, sc.Ng) < 0 ? OffsetP(sc, -1) : OffsetP(sc, 1);
                shadowRays++;
                if (!TraceShadow(tk, o, Scale(l.Ldist * k, l.Ld), 1.0f, sc.Time)) {
                    Spectrum rho = bsdf.Eval(l.Ld);
                    rho.Mul(l.Liu);
                    float p_hat = (float)nB * (float)bsdf.PDF(l.Ld) / (float)total;
                    p_hat += (float)nL * l.Pdf / (float)total;
                    rho.Scale(1.0f / p_hat);
                    RGB rgb = rho.ToRGB();
                    for (int c = 0; c < 3; c++) if (rgb.c[c] < 0) rgb.c[c] = 0;
                    col.Add(rgb);
                }
            }
            for (const auto& b : bs) {
                if (Dot(b.Ld, sc.N) <= 0) continue;
                Vec3 o = Dot(b.Ld, sc.Ng) < 0 ? OffsetP(sc, -1) : OffsetP(sc, 1);
                shadowRays++;
                if (!TraceShadow(tk, o, Scale(b.Ldist * k, b.Ld), 1.0f, sc.Time)) {
                    Spectrum rho = bsdf.Eval(b.Ld);
                    rho.Mul(b.Liu);
                    float p_hat = (float)nB * (float)b.Pdf / (float)total;
                    p_hat += (float)nL * b.PdfLight / (float)total;
                    rho.Scale(1.0f / p_hat);
                    RGB rgb = rho.ToRGB();
                    for (int c = 0; c < 3; c++) if (rgb.c[c] < 0) rgb.c[c] = 0;
                    col.Add(rgb);
                }
            }
            col.Scale(1.0f / (float)total);
        } else {
            std::vector<LightSample> ls;
            SampleArea(L, sc, NSamples, ls);
            for (const auto& l : ls) {
                if (Dot(l.Ld, sc.N) <= 0) continue;
                Vec3 o = Dot(l.Ld, sc.Ng) < 0 ? OffsetP(sc, -1) : OffsetP(sc, 1);
                shadowRays++;
                if (!TraceShadow(tk, o, Scale(l.Ldist * k, l.Ld), 1.0f, sc.Time)) {
                    Spectrum rho = bsdf.Eval(l.Ld);
                    rho.Mul(l.Liu);
                    rho.Scale(1.0f / l.Pdf);
                    col.Add(rho.ToRGB());
                }
            }
        }
        return col;
    }

    // ShaderStd.Eval (std.go:77-296) at recursion level `level`
    RGB Shade(Task* tk, const SC& sg, const Material& sh, int level, bool smooth) {
        if (pathDepth == 0 && level > 3) return RGB();
        Vec3 V = Cross(sg.N, sg.DdPdu);
        if (Length2(V) < 0.1f) V = Cross(sg.N, sg.DdPdv);
        V = Normalize(V);
        Vec3 U = Normalize(Cross(sg.N, V));
        float diffRoughness = sh.has_diffuse_roughness ? sh.diffuse_roughness : 0.5f;
        OrenNayar diff{sg.Lambda, BasisProject(U, V, sg.N, Neg(sg.Rd)), diffRoughness * diffRoughness, U, V, sg.N};
        RGB diffContrib, diffColour;
        if (sh.has_diffuse_colour) for (int c = 0; c < 3; c++) diffColour.c[c] = sh.diffuse[c];
        float dw = sh.has_diffuse_strength ? sh.diffuse_strength : 0.0f;
        float sw = sh.has_spec1_strength ? sh.spec1_strength : 0.0f;
        float total = dw + sw;
        dw /= total; sw /= total;
        // NextLight: 1 << Samples samples per light at level 0, one below (shader.go:179-195); the lights
        // a shading point sees are all but the one owning the hit geom (scene.go:100-116)
        if (dw > 0.0f) {
            for (const auto& L : lights) {
                if (L.geom == sg.geom) continue;
                int NS = level == 0 ? 1 << L.samples : 1;
                RGB c = EvaluateLightSamples(tk, sg, L, NS, diff);
                c.Mul(diffColour);
                diffContrib.Add(c);
            }
            diffContrib.Scale(dw);
        }
        float ior = sh.has_ior ? sh.ior : 1.7f;
        FresnelModel fres = MakeFresnel(sh, ior);
        RGB specContrib;
        if (sw > 0.0f) {
            float rough = sh.has_spec1_roughness ? sh.spec1_roughness : 0.5f;
            RGB specColour;
            if (sh.has_spec1_colour) for (int c = 0; c < 3; c++) specColour.c[c] = sh.spec1[c];
            if (rough == 0.0f && pathDepth == 0) {
                // perfect mirror (std.go:209-261): one reflected ray, recursively shaded at level + 1
                Specular spec{sg.Lambda, BasisProject(U, V, sg.N, Neg(sg.Rd)), fres, U, V, sg.N};
                Vec3 o = spec.Sample();
                double pdf = spec.PDF(o);
                if (Dot(o, sg.Ng) > 0.0f) {
                    Ray r{};
                    r.P = OffsetP(sg, 1); r.D = o; r.Tclosest = kInf; r.Type = RayTypeCamera; r.Time = sg.Time;
                    RaySetup(r);
                    bool hit;
                    RGB sc = TraceShade(tk, r, level + 1, sg.Lambda, sg.I, sg.Scramble, hit);
                    if (hit) {
                        Spectrum rho = spec.Eval(o);
                        rho.Scale(1.0f / (float)pdf);
                        RGB col = rho.ToRGB();
                        col.Mul(specColour);
                        col.Mul(sc);
                        for (int k = 0; k < 3; k++) if (col.c[k] < 0 || std::isnan(col.c[k])) col.c[k] = 0;
                        specContrib.Add(col);
                    }
                }
                specContrib.Scale(sw / 1.0f);
            } else if (rough > 0.0f) {   // no direct lighting for other roughness values (std.go:263)
                GGX ggx{sg.Lambda, BasisProject(U, V, sg.N, Neg(sg.Rd)), rough * rough, fres, U, V, sg.N};
                for (const auto& L : lights) {
                    if (L.geom == sg.geom) continue;
                    int NS = level == 0 ? 1 << L.samples : 1;
                    RGB c = EvaluateLightSamples(tk, sg, L, NS, ggx);
                    c.Mul(specColour);
                    specContrib.Add(c);
                }
            }
        }
        RGB contrib;
        RGB em = Emission(sh);
        if (pathDepth > 0 && level > 0 && smooth) {
            // path mode: a light's own geometry reached by an Oren-Nayar or GGX sample was covered by the
            // previous vertex's EvaluateLightSamples
            for (const auto& L : lights) if (L.geom == sg.geom) em = RGB();
        }
        contrib.Add(em);
        contrib.Add(diffContrib);
        contrib.Add(specContrib);
        if (pathDepth > 0 && level + 1 < pathDepth) contrib.Add(Continue(tk, sg, sh, level, dw, sw, fres));
        return contrib;
    }

    // Path mode's continuation at a vertex (not in the reference): the diffuse, mirror or GGX lobe picked
    // with its normalised weight and sampled by the lobe's Sample; the next vertex's colour weighted by
    // (Eval / pdf).ToRGB() * lobe colour and clamped at 0 as the reference's mirror ray (std.go:226-258).
    RGB Continue(Task* tk, const SC& sg, const Material& sh, int level, float dw, float sw, const FresnelModel& fres) {
        Vec3 V = Cross(sg.N, sg.DdPdu);
        if (Length2(V) < 0.1f) V = Cross(sg.N, sg.DdPdv);
        V = Normalize(V);
        Vec3 U = Normalize(Cross(sg.N, V));
        const Vec3 wr = BasisProject(U, V, sg.N, Neg(sg.Rd));
        const double u = VanDerCorput(sg.I, LevelScramble(sg.PixScramble[0], level, 2));
        const double r0 = VanDerCorput(sg.I, LevelScramble(sg.PixScramble[0], level, 3));
        const double r1 = Sobol(sg.I, LevelScramble(sg.PixScramble[1], level, 4));
        const float rough = sh.has_spec1_roughness ? sh.spec1_roughness : 0.5f;
        Vec3 o = V3(0, 0, 0);
        double pdf = 0;
        Spectrum rho;
        RGB colour;
        bool smooth = true, ok = false;
        if (u < (double)dw) {
            float dr = sh.has_diffuse_roughness ? sh.diffuse_roughness : 0.5f;
            OrenNayar diff{sg.Lambda, wr, dr * dr, U, V, sg.N};
            o = Normalize(diff.Sample(r0, r1));
            pdf = diff.PDF(o);
            rho = diff.Eval(o);
            if (sh.has_diffuse_colour) for (int c = 0; c < 3; c++) colour.c[c] = sh.diffuse[c];
            ok = true;
        } else if (sw > 0.0f && rough == 0.0f) {
            Specular spec{sg.Lambda, wr, fres, U, V, sg.N};
            o = spec.Sample();
            pdf = spec.PDF(o);
            rho = spec.Eval(o);
            if (sh.has_spec1_colour) for (int c = 0; c < 3; c++) colour.c[c] = sh.spec1[c];
            smooth = false;
            ok = true;
        } else if (sw > 0.0f && rough > 0.0f) {
            GGX ggx{sg.Lambda, wr, rough * rough, fres, U, V, sg.N};
            o = Normalize(ggx.Sample(r0, r1));
            pdf = ggx.PDFDir(o);
            rho = ggx.Eval(o);
            if (sh.has_spec1_colour) for (int c = 0; c < 3; c++) colour.c[c] = sh.spec1[c];
            ok = true;
        }
        RGB out;
        if (!ok || !(pdf > 0) || !(Dot(o, sg.Ng) > 0.0f)) return out;
        Ray r{};
        r.P = OffsetAlong(OffsetP(sg, 1), sg.Ng); r.D = o; r.Tclosest = kInf; r.Type = RayTypeCamera; r.Time = sg.Time;
        RaySetup(r);
        bool hit;
        RGB next = TraceShade(tk, r, level + 1, sg.Lambda, sg.I, sg.PixScramble, hit, smooth);
        if (!hit) return out;
        rho.Scale(1.0f / (float)pdf);
        RGB col = rho.ToRGB();
        col.Mul(colour);
        col.Mul(next);
        for (int k = 0; k < 3; k++) if (col.c[k] < 0 || std::isnan(col.c[k])) col.c[k] = 0;
        return col;
    }

    // core.Trace (trace.go:40-83): closest hit, then the hit's shader; hit = false on a miss or a geom
    // without a shader.
    RGB TraceShade(Task* tk, Ray& r, int level, float lambda, uint64_t I, const uint64_t* scramble, bool& hit,
                   bool smooth = false) {
        tk->StackTop = 0; r.task = tk;
        SceneHit sh{}; sh.geom = -1; sh.hr.elem = -1;
        RGB out;
        hit = scene->Trace(r, sh);
        if (!hit) return out;
        const HitRecord& h = sh.hr;
        SC sc;
        sc.geom = sh.geom; sc.Lambda = lambda; sc.Time = r.Time; sc.I = I; sc.Scramble[0] = scramble[0]; sc.Scramble[1] = scramble[1];
        sc.PixScramble[0] = scramble[0]; sc.PixScramble[1] = scramble[1];
        if (pathDepth > 0 && level > 0) {
            sc.Scramble[0] = LevelScramble(scramble[0], level, 0);
            sc.Scramble[1] = LevelScramble(scramble[1], level, 1);
        }
        sc.Rd = r.D;
        if (h.hasXform) {
            // PolyMesh.Trace's own transform of N (trace.go:65-70), then ApplyTransform with sg.Transform
            // (core/shader.go:129-135); Poffset stays in object space
            M4 T;
            Vec3 N1 = h.N;
            if (h.meshXform) { M4Transpose(h.meshInv, T); N1 = MulVec(T, h.N); }
            M4Transpose(h.xformInv, T);
            sc.P = MulPoint(h.xform, h.Po);
            sc.N = Normalize(MulVec(T, N1));
            sc.Ng = Normalize(MulVec(T, h.Ng));
            sc.DdPdu = Normalize(MulVec(h.xform, h.DdPdu));
            sc.DdPdv = Normalize(MulVec(h.xform, h.DdPdv));
        } else {
            // ApplyTransform with identity matrices (core/shader.go:129-135)
            sc.P = h.Po;
            sc.N = Normalize(h.N); sc.Ng = Normalize(h.Ng);
            sc.DdPdu = Normalize(h.DdPdu); sc.DdPdv = Normalize(h.DdPdv);
        }
        sc.Poffset = h.Poffset;
        const auto& shaders = h.leaf ? h.leaf->shaders : geomShaders[sh.geom];
        int mat = shaders.empty() ? -1 : shaders[h.shaderIdx];
        if (mat < 0) { hit = false; return out; }
        return Shade(tk, sc, mats[mat], level, smooth);
    }

    // Sampler.WarpSample (builtin/filter/filter.go:40-85)
    void WarpSample(double r0, double r1, double& u, double& v) const {
        const int n = filtN;
        const double w = filtW;
        const double* cdfV = filt.data();
        u = 0; v = 0;
        int uI = -1;
        for (int i = 0; i < n; i++) {
            uI = i;
            if (r0 < cdfV[i]) {
                if (i == 0) { double du = r0 / cdfV[i]; u = (-w / 2) + (du / w); }
                else { double du = (r0 - cdfV[i - 1]) / (cdfV[i] - cdfV[i - 1]); u = (-w / 2) + w * ((double)i + du) / (double)(n - 1); }
                break;
            }
        }
        const double* row = filt.data() + n + (size_t)uI * n;
        for (int i = 0; i < n; i++) {
            if (r1 < row[i]) {
                if (i == 0) { double dv = r1 / row[i]; v = (-w / 2) + (dv / w); }
                else { double dv = (r1 - row[i - 1]) / (row[i] - row[i - 1]); v = (-w / 2) + w * ((double)i + dv) / (double)(n - 1); }
                return;
            }
        }
        u = 0; v = 0;
    }

    // core/render.go:75-130 for one pixel; returns the sample colour (and whether it hit)
    RGB Sample(Task* tk, uint32_t iter, uint32_t x, uint32_t y, const uint64_t* scr, bool& hit, int32_t& hitGeom) {
        uint64_t index; double rx, ry;
        RasterXY(12, iter, x, y, 0, 0, index, rx, ry);
        double lambda = (720 - 450) * VanDerCorput(iter, scr[3]) + 450;
        double time = VanDerCorput(iter, scr[2]);   // core/render.go:93
        if (filtN > 0) {
            double pixu = rx - std::floor(rx), pixv = ry - std::floor(ry);
            double u, v;
            WarpSample(pixu, pixv, u, v);
            rx = std::floor(rx) + 0.5 + u;
            ry = std::floor(ry) + 0.5 + v;
        }
        float Sx = (float)(-1.0 + 2.0 * (rx / (double)width));
        float Sy = -(float)(-1.0 + 2.0 * (ry / (double)height));
        // camera.ComputeRay, camera.go:221-300: pinhole, or a thin lens when Radius > 0
        float camu = Sx * tanThetaFocal;
        float camv = Sy * (tanThetaFocal / aspect);
        Vec3 s = Sub(Add(Scale(camu, V3(1, 0, 0)), Scale(camv, V3(0, 1, 0))), Scale(focal, V3(0, 0, 1)));
        Vec3 e = V3(0, 0, 0);
        if (lensRadius > 0.0f) {
            // lensU, lensV (core/render.go:91-92) through sample.UniformDisk2D (math/sample/sample.go:105-129)
            float r0 = (float)VanDerCorput(iter, scr[0]), r1 = (float)Sobol(iter, scr[1]);
            float x = -1 + 2 * r0, y = -1 + 2 * r1, rr = 0, theta = 0;
            if (x > -y && x > y) { rr = x; theta = (Pi32 / 4) * y / x; }
            else if (x > -y && x < y) { rr = y; theta = (Pi32 / 4) * (2 - x / y); }
            else if (x < y && x < -y) { rr = -x; theta = (Pi32 / 4) * (4 + y / x); }
            else if (x > y && x < -y) { rr = -y; theta = (Pi32 / 4) * (6 - x / y); }
            float xo = lensRadius * rr * fCos(theta), yo = lensRadius * rr * fSin(theta);
            e = Add(Scale(xo, V3(1, 0, 0)), Scale(yo, V3(0, 1, 0)));
        }
        // camera.go:224-236: the matrix of the lerped decomposition at the sample's time
        M4 Ms;
        std::memcpy(Ms, M, sizeof(M4));
        if (camKeys.size() >= 2) {
            float k = (float)time * (float)(camKeys.size() - 1);
            float t = k - Floor(k);
            int key = (int)Floor(k), key2 = (int)Ceil(k);
            DecompToM4(camKeys[key], camKeys[key2], t, Ms);
        }
        Vec3 d = MulVec(Ms, Sub(s, e));
        Ray r{};
        r.D = Normalize(d);
        r.P = MulPoint(Ms, e);
        r.Tclosest = kInf; r.Type = RayTypeCamera; r.Time = (float)time;   // sc.Time (render.go:118), Ray.Init
        RaySetup(r);
        RGB out = TraceShade(tk, r, 0, (float)lambda, iter, scr + 4, hit);
        hitGeom = -1;
        if (!hit) return RGB();
        return out;
    }
};

}  // namespace orc

using namespace orc;

extern "C" {

// Camera matrix as the reference computes it for a static LookAt camera: calcLookatMatrices
// (camera.go:149-192), TransformDecompMatrix4 and TransformDecompToMatrix4 of the lerp at t=0
// (camera.go:225-236).  Also returns TanThetaFocal (camera.go:89).
void orc_camera_lookat(const float* from, const float* to, const float* up, float fov_deg, float focal, float* M_out,
                       float* tan_theta_focal) {
    Vec3 F = V3(from[0], from[1], from[2]), T = V3(to[0], to[1], to[2]), Up = V3(up[0], up[1], up[2]);
    Vec3 W = Normalize(Sub(F, T));
    Vec3 u = Normalize(Cross(Up, W));
    Vec3 v = Normalize(Cross(u, W));
    float roll = 0;
    Vec3 U = Add(Scale(fCos(roll), u), Scale(fSin(roll), v));
    Vec3 V = Add(Scale(-fSin(roll), u), Scale(fCos(roll), v));
    M4 Tr; M4Identity(Tr); Tr[12] = F[0]; Tr[13] = F[1]; Tr[14] = F[2];
    M4 B; std::memset(B, 0, sizeof(M4));
    B[0] = U[0]; B[1] = U[1]; B[2] = U[2]; B[4] = V[0]; B[5] = V[1]; B[6] = V[2]; B[8] = W[0]; B[9] = W[1]; B[10] = W[2]; B[15] = 1;
    M4 mtx; M4Mul(Tr, B, mtx);
    Decomp d = DecompM4(mtx);
    DecompToM4(d, d, 0.0f, M_out);
    *tan_theta_focal = fTan((fov_deg / 2) * Pi32 / 180.0f) * focal;
}

// Camera.calcLookatMatrices (camera.go:109-193) for keyed From / To / Roll: LocalToWorld per key of
// the parameter with more keys (key i at time i / nkeys, as the reference samples the other one).
// Writes max(nfrom, nto) column-major matrices and returns their count.
int orc_camera_lookat_keys(const float* from, int32_t nfrom, const float* to, int32_t nto, const float* up,
                           const float* roll, int32_t nroll, float* mats_out) {
    Vec3 Up = V3(up[0], up[1], up[2]);
    const bool byTarget = nto > nfrom;
    const int n = byTarget ? nto : nfrom;
    for (int i = 0; i < n; i++) {
        float time = (float)i / (float)n;
        const float* other = byTarget ? from : to;
        int nother = byTarget ? nfrom : nto;
        float k = time * (float)(nother - 1);
        float t = k - Floor(k);
        int key = (int)Floor(k), key2 = (int)Ceil(k);
        Vec3 Pl = Lerp(V3(other[key * 3], other[key * 3 + 1], other[key * 3 + 2]),
                       V3(other[key2 * 3], other[key2 * 3 + 1], other[key2 * 3 + 2]), t);
        const float* own = byTarget ? to : from;
        Vec3 Pi = V3(own[i * 3], own[i * 3 + 1], own[i * 3 + 2]);
        Vec3 F = byTarget ? Pl : Pi, T = byTarget ? Pi : Pl;
        Vec3 W = Normalize(Sub(F, T));
        Vec3 u = Normalize(Cross(Up, W));
        Vec3 v = Normalize(Cross(u, W));
        float rl = 0;
        if (roll && nroll > 0) {
            float kr = time * (float)(nroll - 1);
            float tr = kr - Floor(kr);
            int r0 = (int)Floor(kr), r1 = (int)Ceil(kr);
            rl = (1 - tr) * roll[r0] + tr * roll[r1];
        }
        Vec3 U = Add(Scale(fCos(rl), u), Scale(fSin(rl), v));
        Vec3 V = Add(Scale(-fSin(rl), u), Scale(fCos(rl), v));
        M4 Tr; M4Identity(Tr); Tr[12] = F[0]; Tr[13] = F[1]; Tr[14] = F[2];
        M4 B; std::memset(B, 0, sizeof(M4));
        B[0] = U[0]; B[1] = U[1]; B[2] = U[2]; B[4] = V[0]; B[5] = V[1]; B[6] = V[2]; B[8] = W[0]; B[9] = W[1]; B[10] = W[2]; B[15] = 1;
        M4Mul(Tr, B, mats_out + (size_t)i * 16);
    }
    return n;
}

// The camera's LocalToWorld keys: TransformDecompMatrix4 of each (camera.go:188-192).  One key sets the
// static matrix (the decomposition recomposed at t = 0); two or more make the matrix per sample.
void orc_render_set_camera_keys(void* rp, const float* mats, int32_t nkeys) {
    Renderer* r = (Renderer*)rp;
    r->camKeys.clear();
    for (int i = 0; i < nkeys; i++) r->camKeys.push_back(DecompM4(mats + (size_t)i * 16));
    if (nkeys == 1) { DecompToM4(r->camKeys[0], r->camKeys[0], 0.0f, r->M); r->camKeys.clear(); }
}

// TransformDecompMatrix4 then TransformDecompToMatrix4 of the lerp of two decompositions at t
// (animdecomp.go:20-80), for known-answer tests.
void orc_decomp_lerp_matrix(const float* a, const float* b, float t, float* out) {
    Decomp da = DecompM4(a), db = DecompM4(b);
    DecompToM4(da, db, t, out);
}

// VdC / Sobol / RasterXY for known-answer tests
double orc_vdc(uint64_t i, uint64_t s) { return VanDerCorput(i, s); }
double orc_sobol(uint64_t i, uint64_t s) { return Sobol(i, s); }
void orc_raster_xy(uint32_t m, uint32_t frame, uint32_t px, uint32_t py, uint64_t* index, double* rx, double* ry) {
    RasterXY(m, frame, px, py, 0, 0, *index, *rx, *ry);
}
// enumeration matrices for m in [1,25]: fwd (52) and inv (52)
void orc_raster_tables(int32_t m, uint64_t* fwd, uint64_t* inv) {
    const RasterTables& t = rasterTables(m);
    std::memcpy(fwd, t.fwd, sizeof(t.fwd));
    std::memcpy(inv, t.inv, sizeof(t.inv));
}
// spectrum helpers for known-answer tests
float orc_smits(float r, float g, float b, float lambda) { return Smits(r, g, b, lambda); }
void orc_spectrum_to_rgb(const float* C4, float lambda, float* rgb) {
    Spectrum s; for (int k = 0; k < 4; k++) s.C[k] = C4[k]; s.Lambda = lambda;
    RGB o = s.ToRGB(); for (int k = 0; k < 3; k++) rgb[k] = o.c[k];
}
void orc_cie_tables(float* x, float* y, float* z) {
    std::memcpy(x, cieX, sizeof(cieX)); std::memcpy(y, cieY, sizeof(cieY)); std::memcpy(z, cieZ, sizeof(cieZ));
}

struct orc_material {
    float emission[3]; float emission_strength; int32_t has_emission_colour, has_emission_strength;
    float diffuse[3]; int32_t has_diffuse_colour;
    float diffuse_strength; int32_t has_diffuse_strength;
    float diffuse_roughness; int32_t has_diffuse_roughness;
    float spec1[3]; int32_t has_spec1_colour;
    float spec1_strength; int32_t has_spec1_strength;
    float spec1_roughness; int32_t has_spec1_roughness;
    float ior; int32_t has_ior;
    int32_t fresnel_model;
    float spec1_fresnel_refl[3]; int32_t has_spec1_fresnel_refl;
    float spec1_fresnel_edge[3]; int32_t has_spec1_fresnel_edge;
};
static_assert(sizeof(orc_material) == sizeof(Material), "material layout");

// Adds a mesh with shading data: normals (nverts*3, nullable), shader index per triangle (nullable),
// shader table (material ids, nshaders).  Returns the insertion index.
int orc_scene_add_shaded_mesh(void* sp, const float* verts, int32_t nverts, const uint32_t* idx, int32_t ntris,
                              const float* normals, const uint8_t* shaderidx, const int32_t* shaders, int32_t nshaders) {
    Scene* s = (Scene*)sp;
    Mesh* m = new Mesh();
    m->nkeys = 1; m->elemsPerKey = nverts;
    m->verts.resize(nverts);
    std::memcpy(m->verts.data(), verts, sizeof(float) * 3 * (size_t)nverts);
    m->idxp.assign(idx, idx + (size_t)ntris * 3);
    m->facecount = ntris;
    if (normals) {
        m->normals.resize(nverts);
        std::memcpy(m->normals.data(), normals, sizeof(float) * 3 * (size_t)nverts);
        m->normalidx.assign(idx, idx + (size_t)ntris * 3);
    }
    if (shaderidx) m->shaderidx.assign(shaderidx, shaderidx + ntris);
    m->shaders.assign(shaders, shaders + nshaders);
    s->geoms.push_back(m);
    return (int)s->geoms.size() - 1;
}

// Adds a moving mesh with shading data: verts nkeys*nverts*3 (key-major), normals normal_keys*ntris*3
// per face corner (nullable; trace.go:694-703 uses them only when normal_keys == nkeys), shader index per
// triangle (nullable), shader table.  motion_bug: the reference's slot indexing (trace.go:527-545).
int orc_scene_add_shaded_motion_mesh(void* sp, const float* verts, int32_t nverts, int32_t nkeys, const uint32_t* idx,
                                     int32_t ntris, const float* normals, int32_t normal_keys, const uint8_t* shaderidx,
                                     const int32_t* shaders, int32_t nshaders, int32_t motion_bug) {
    Scene* s = (Scene*)sp;
    Mesh* m = new Mesh();
    m->nkeys = nkeys; m->elemsPerKey = nverts;
    m->verts.resize((size_t)nverts * nkeys);
    std::memcpy(m->verts.data(), verts, sizeof(float) * 3 * (size_t)nverts * nkeys);
    m->idxp.assign(idx, idx + (size_t)ntris * 3);
    m->facecount = ntris;
    m->motionBug = motion_bug != 0;
    if (normals) {
        m->normalKeys = normal_keys;
        m->mnormals.resize((size_t)ntris * 3 * normal_keys);
        std::memcpy(m->mnormals.data(), normals, sizeof(float) * 9 * (size_t)ntris * normal_keys);
    }
    if (shaderidx) m->shaderidx.assign(shaderidx, shaderidx + ntris);
    m->shaders.assign(shaders, shaders + nshaders);
    s->geoms.push_back(m);
    return (int)s->geoms.size() - 1;
}

void* orc_render_new(void* sp, int32_t w, int32_t h, const float* M, float tan_theta_focal, float aspect, float focal,
                     const orc_material* mats, int32_t nmats) {
    Renderer* r = new Renderer();
    r->scene = (Scene*)sp;
    r->width = w; r->height = h;
    std::memcpy(r->M, M, sizeof(M4));
    r->tanThetaFocal = tan_theta_focal; r->aspect = aspect; r->focal = focal;
    r->mats.resize(nmats);
    std::memcpy(r->mats.data(), mats, sizeof(Material) * nmats);
    for (auto* g : r->scene->geoms) r->geomShaders.push_back(g->shaders);
    return r;
}
void orc_render_free(void* r) { delete (Renderer*)r; }
// the camera node's Radius (camera.go:265-283): a thin lens when > 0
void orc_render_set_lens(void* rp, float radius) { ((Renderer*)rp)->lensRadius = radius; }
void orc_render_set_path_depth(void* rp, int32_t depth) { ((Renderer*)rp)->pathDepth = depth; }
// sample.UniformDisk2D (math/sample/sample.go:105-129) for known-answer tests
void orc_uniform_disk(float radius, float r0, float r1, float* xy) {
    float x = -1 + 2 * r0, y = -1 + 2 * r1, rr = 0, theta = 0;
    if (x > -y && x > y) { rr = x; theta = (Pi32 / 4) * y / x; }
    else if (x > -y && x < y) { rr = y; theta = (Pi32 / 4) * (2 - x / y); }
    else if (x < y && x < -y) { rr = -x; theta = (Pi32 / 4) * (4 + y / x); }
    else if (x > y && x < -y) { rr = -y; theta = (Pi32 / 4) * (6 - x / y); }
    xy[0] = radius * rr * fCos(theta);
    xy[1] = radius * rr * fSin(theta);
}

// light_geom_insertion: the insertion index of the light's mesh (mapped to its post-sort slot here)
// type: 0 TriLight (P0, P1, P2 vertices), 1 DiskLight (P0 = P, P1 = Up, P2 = LookAt, radius),
// 2 SphereLight (P0 = centre, radius).  light_geom_insertion: the insertion index of the light's
// geometry (mapped to its post-sort slot here).
void orc_render_add_light(void* rp, int32_t type, const float* P0, const float* P1, const float* P2, float radius,
                          int32_t material, int32_t samples, int32_t light_geom_insertion) {
    Renderer* r = (Renderer*)rp;
    Light L;
    L.type = type;
    L.P0 = V3(P0[0], P0[1], P0[2]); L.P1 = V3(P1[0], P1[1], P1[2]); L.P2 = V3(P2[0], P2[1], P2[2]);
    L.radius = radius;
    L.material = material; L.samples = samples;
    L.geom = -1;
    for (size_t i = 0; i < r->scene->geomOrder.size(); i++) if (r->scene->geomOrder[i] == light_geom_insertion) L.geom = (int32_t)i;
    r->lights.push_back(L);
}

// type 0 box (no filter), 1 Gaussian (gauss.go:34-58; defaults Width 2, Res 17), 2 Airy (airy.go:82-106;
// Width 6, Res 49, Peak 4).  Values <= 0 take the defaults.
void orc_render_set_filter(void* rp, int32_t type, float width, int32_t res, float peak) {
    Renderer* r = (Renderer*)rp;
    r->filt.clear();
    r->filtN = 0;
    if (type == 0) return;
    const bool gauss = type == 1;
    const float W = width > 0 ? width : (gauss ? 2.0f : 6.0f);
    const int n = res > 0 ? res : (gauss ? 17 : 49);
    const float pk = peak > 0 ? peak : 4.0f;
    if (gauss) {
        CreateSampler(n, (double)W, [W](double x, double y) {
            double q = std::sqrt(x * x + y * y);
            if (q > (double)(W / 2)) return 0.0;
            double sigma = (double)(1.0 / std::sqrt((double)W));
            double G = (1 / (2 * M_PI * (sigma * sigma))) * std::exp(-(x * x + y * y) / 2 * (sigma * sigma));
            return G;
        }, r->filt);
    } else {
        CreateSampler(n, (double)W, [W, pk](double x, double y) {
            double q = std::sqrt(x * x + y * y);
            if (q > (double)(W / 2)) return 0.0;
            double lambda = 550, N = 5.6;
            double v = (20000 / (double)W) * (M_PI * q) / (lambda * N);
            double t = 2 * BesselJ1(v) / v;
            return (double)pk * (t * t);
        }, r->filt);
    }
    r->filtN = n;
    r->filtW = (double)W;
}

// Sampler.WarpSample on (r0, r1) for known-answer tests (box filter: (0, 0))
void orc_render_warp_sample(void* rp, double r0, double r1, double* uv) {
    Renderer* r = (Renderer*)rp;
    double u = 0, v = 0;
    if (r->filtN > 0) r->WarpSample(r0, r1, u, v);
    uv[0] = u; uv[1] = v;
}
// the sampler's tables: cdfV (n) then cdfVU (n x n); returns n (0: box)
int32_t orc_render_filter_tables(void* rp, double* out) {
    Renderer* r = (Renderer*)rp;
    if (out) std::memcpy(out, r->filt.data(), r->filt.size() * sizeof(double));
    return r->filtN;
}

// One pass (core/render.go render() for every pixel).  scrambles: 6 uint64 per pixel (lensU, lensV,
// time, lambda, scramble0, scramble1 as core/render.go:166-176).  iter: the 1-based pass number the
// reference passes to render().  film: w*h*3 floats updated as fb = (fb*iter + c)/(iter+1).
// sample_out (nullable): w*h*3 floats receiving this pass's sample colours.  Returns shadow rays traced.
int64_t orc_render_pass(void* rp, const uint64_t* scrambles, uint32_t iter, float* film, float* sample_out, int32_t nthreads) {
    Renderer* r = (Renderer*)rp;
    const int w = r->width, h = r->height;
    int64_t shadow = 0;
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt) reduction(+ : shadow)
#endif
    {
        Task* tk = new Task();
        Renderer local = *r;
        local.shadowRays = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int p = 0; p < w * h; p++) {
            int x = p % w, y = p / w;
            bool hit; int32_t g;
            RGB c = local.Sample(tk, iter, (uint32_t)x, (uint32_t)y, scrambles + (size_t)p * 6, hit, g);
            for (int k = 0; k < 3; k++) film[p * 3 + k] = (film[p * 3 + k] * (float)iter + c.c[k]) / (float)(iter + 1);
            if (sample_out) for (int k = 0; k < 3; k++) sample_out[p * 3 + k] = c.c[k];
        }
        shadow += local.shadowRays;
        delete tk;
    }
    (void)nthreads;
    return shadow;
}

}  // extern "C"

extern "C" {
// BSDF pieces for known-answer tests
float orc_fresnel_dielectric(float eta, float c) { return DielectricKr(eta, c).c[0]; }
float orc_fresnel_conductor(float r, float g, float c) { return ConductorFresnel(r, g, c); }
float orc_ggx_d(const float* m, float alpha) { return ggxD(V3(m[0], m[1], m[2]), alpha); }
// Oren-Nayar Eval (as a spectrum, returned as its 4 values) for omegaI/omegaO given in the tangent frame
void orc_oren_nayar_eval(float lambda, const float* wi, const float* wo, float roughness, float* out4) {
    OrenNayar b{lambda, V3(wi[0], wi[1], wi[2]), roughness * roughness, V3(1, 0, 0), V3(0, 1, 0), V3(0, 0, 1)};
    Spectrum s = b.Eval(V3(wo[0], wo[1], wo[2]));
    for (int k = 0; k < 4; k++) out4[k] = s.C[k];
}
}

extern "C" {
// Hooks for the tests that pin the light samplers, EvaluateLightSamples and the GGX lobe against
// closed forms and invariants.  They call the restated functions above unchanged.

// sampleSphericalTriangle (triangle.go:474-535) for n (r0, r1) pairs: unit directions and pdfs.
void orc_sample_spherical_triangle_n(const float* p0, const float* p1, const float* p2, const float* p, int64_t n,
                                     const double* r0, const double* r1, float* dirs, double* pdf) {
    const Vec3 a = V3(p0[0], p0[1], p0[2]), b = V3(p1[0], p1[1], p1[2]), c = V3(p2[0], p2[1], p2[2]);
    const Vec3 o = V3(p[0], p[1], p[2]);
    for (int64_t i = 0; i < n; i++) {
        Vec3 x = sampleSphericalTriangle(a, b, c, o, r0[i], r1[i], pdf[i]);
        for (int k = 0; k < 3; k++) dirs[i * 3 + k] = x[k];
    }
}

// Light.SampleArea of light `li` of a renderer at a shading point (P, Ng, with N = Ng) for the
// sample indices I = I0 .. I0+count-1, n samples each (ldseq VdC/Sobol with scrambles s0, s1).
// out: 5 floats per accepted sample (Ld xyz, Ldist, Pdf); returns the number accepted.
int64_t orc_light_sample_range(void* rp, int32_t li, const float* P, const float* Ng, float lambda, uint64_t s0,
                               uint64_t s1, uint64_t I0, int64_t count, int32_t n, float* out) {
    Renderer* r = (Renderer*)rp;
    SC sg{};
    sg.P = V3(P[0], P[1], P[2]);
    sg.Ng = V3(Ng[0], Ng[1], Ng[2]);
    sg.N = sg.Ng;
    sg.Lambda = lambda;
    sg.Scramble[0] = s0; sg.Scramble[1] = s1;
    int64_t m = 0;
    std::vector<LightSample> ls;
    for (int64_t i = 0; i < count; i++) {
        sg.I = I0 + (uint64_t)i;
        ls.clear();
        r->SampleArea(r->lights[li], sg, n, ls);
        for (const auto& l : ls) {
            for (int k = 0; k < 3; k++) out[m * 5 + k] = l.Ld[k];
            out[m * 5 + 3] = l.Ldist;
            out[m * 5 + 4] = l.Pdf;
            m++;
        }
    }
    return m;
}

// EvaluateLightSamples (core/shader.go:203-402) of light `li` with ShaderStd's diffuse lobe
// (Oren-Nayar with sigma = roughness^2, frame built as std.go:99-104) at a point P with normal
// N = Ng, tangent dPdu, view direction Rd, for the sample indices I0 .. I0+count-1 of one pixel
// (scrambles[0..1]); rgb: 3 floats per index.  The renderer's scene supplies the shadow rays.
void orc_eval_light_samples(void* rp, int32_t li, const float* P, const float* N, const float* dPdu, const float* Rd,
                            float lambda, float roughness, const uint64_t* scrambles, uint64_t I0, int64_t count,
                            int32_t nsamples, float* rgb) {
    Renderer* r = (Renderer*)rp;
    Task* tk = new Task();
    SC sg{};
    sg.geom = -1;
    sg.P = V3(P[0], P[1], P[2]);
    sg.N = V3(N[0], N[1], N[2]);
    sg.Ng = sg.N;
    sg.Poffset = V3(0, 0, 0);
    sg.DdPdu = V3(dPdu[0], dPdu[1], dPdu[2]);
    sg.DdPdv = Cross(sg.N, sg.DdPdu);
    sg.Rd = V3(Rd[0], Rd[1], Rd[2]);
    sg.Lambda = lambda;
    sg.Scramble[0] = scrambles[0]; sg.Scramble[1] = scrambles[1];
    Vec3 V = Normalize(Cross(sg.N, sg.DdPdu));
    Vec3 U = Normalize(Cross(sg.N, V));
    OrenNayar diff{sg.Lambda, BasisProject(U, V, sg.N, Neg(sg.Rd)), roughness * roughness, U, V, sg.N};
    for (int64_t i = 0; i < count; i++) {
        sg.I = I0 + (uint64_t)i;
        RGB c = r->EvaluateLightSamples(tk, sg, r->lights[li], nsamples, diff);
        for (int k = 0; k < 3; k++) rgb[i * 3 + k] = c.c[k];
    }
    delete tk;
}

// ToRGB(FromRGB(a) * FromRGB(b)) at hero wavelength lambda (colour/spectrum.go): the colour factor of a
// product of two RGB reflectances/emissions, which shading applies per light sample.
void orc_spectrum_product_rgb(const float* a, const float* b, float lambda, float* rgb) {
    Spectrum sa, sb;
    sa.Lambda = lambda; sb.Lambda = lambda;
    RGB ra, rb;
    for (int k = 0; k < 3; k++) { ra.c[k] = a[k]; rb.c[k] = b[k]; }
    sa.FromRGB(ra); sb.FromRGB(rb);
    sa.Mul(sb);
    RGB o = sa.ToRGB();
    for (int k = 0; k < 3; k++) rgb[k] = o.c[k];
}

// MicrofacetGGX (microfacetggx.go) in the tangent frame (U, V, N = x, y, z), dielectric Fresnel with eta:
// Sample for n (r0, r1) pairs, PDF and Eval (4 spectral values) of given directions.
void orc_ggx_sample_n(float roughness, const float* omegaR, int64_t n, const double* r0, const double* r1, float* dirs) {
    GGX g{500.0f, V3(omegaR[0], omegaR[1], omegaR[2]), roughness, FresnelModel(), V3(1, 0, 0), V3(0, 1, 0), V3(0, 0, 1)};
    for (int64_t i = 0; i < n; i++) {
        Vec3 o = g.Sample(r0[i], r1[i]);
        for (int k = 0; k < 3; k++) dirs[i * 3 + k] = o[k];
    }
}
void orc_ggx_pdf_eval_n(float roughness, float eta, float lambda, const float* omegaR, int64_t n, const float* dirs,
                        double* pdf, float* eval4) {
    FresnelModel f;
    f.eta = eta;
    GGX g{lambda, V3(omegaR[0], omegaR[1], omegaR[2]), roughness, f, V3(1, 0, 0), V3(0, 1, 0), V3(0, 0, 1)};
    for (int64_t i = 0; i < n; i++) {
        Vec3 o = V3(dirs[i * 3], dirs[i * 3 + 1], dirs[i * 3 + 2]);
        if (pdf) pdf[i] = g.PDF(o);
        if (eval4) { Spectrum s = g.Eval(o); for (int k = 0; k < 4; k++) eval4[i * 4 + k] = s.C[k]; }
    }
}
float orc_ggx_g1(const float* w, const float* m, float alpha) {
    return ggxG1(V3(w[0], w[1], w[2]), V3(m[0], m[1], m[2]), alpha);
}
}

extern "C" {
// ToRGB(FromRGB(rgb)) at each hero wavelength lambdas[i] (colour/spectrum.go): the RGB factor a white
// (or any) reflectance contributes at one bounce of a sample at that wavelength.
void orc_spectrum_rgb_n(const float* rgb, int64_t n, const float* lambdas, float* out) {
    RGB c;
    for (int k = 0; k < 3; k++) c.c[k] = rgb[k];
    for (int64_t i = 0; i < n; i++) {
        Spectrum s;
        s.Lambda = lambdas[i];
        s.FromRGB(c);
        RGB o = s.ToRGB();
        for (int k = 0; k < 3; k++) out[i * 3 + k] = o.c[k];
    }
}

// Path mode's continuation weight Eval(o) / pdf(o) (4 spectral values) of one lobe in the tangent frame
// (U, V, N = x, y, z) for n (r0, r1) pairs, as Renderer::Continue forms it: kind 0 Oren-Nayar (field
// Roughness = sigma) sampled by Sample with PDF; kind 1 GGX (field Roughness, alpha = Roughness^2,
// dielectric Fresnel eta) sampled by Sample with PDFDir.  0 where the pdf is 0 or o is below the surface.
void orc_path_lobe_weight_n(int32_t kind, float roughness, float eta, float lambda, const float* omega, int64_t n,
                            const double* r0, const double* r1, float* dirs, float* out4) {
    const Vec3 w = V3(omega[0], omega[1], omega[2]);
    FresnelModel f;
    f.eta = eta;
    OrenNayar on{lambda, w, roughness, V3(1, 0, 0), V3(0, 1, 0), V3(0, 0, 1)};
    GGX g{lambda, w, roughness, f, V3(1, 0, 0), V3(0, 1, 0), V3(0, 0, 1)};
    for (int64_t i = 0; i < n; i++) {
        Vec3 o = Normalize(kind == 0 ? on.Sample(r0[i], r1[i]) : g.Sample(r0[i], r1[i]));
        double pdf = kind == 0 ? on.PDF(o) : g.PDFDir(o);
        Spectrum e = kind == 0 ? on.Eval(o) : g.Eval(o);
        for (int k = 0; k < 3; k++) dirs[i * 3 + k] = o[k];
        for (int k = 0; k < 4; k++) out4[i * 4 + k] = (pdf > 0 && o[2] > 0) ? e.C[k] * (1.0f / (float)pdf) : 0.0f;
    }
}

// The GGX D and Smith G1 terms (microfacetggx.go:26-80) over n directions.
void orc_ggx_d_n(const float* m, int64_t n, float alpha, float* out) {
    for (int64_t i = 0; i < n; i++) out[i] = ggxD(V3(m[i * 3], m[i * 3 + 1], m[i * 3 + 2]), alpha);
}
void orc_ggx_g1_n(const float* w, const float* m, int64_t n, float alpha, float* out) {
    const Vec3 ww = V3(w[0], w[1], w[2]);
    for (int64_t i = 0; i < n; i++) out[i] = ggxG1(ww, V3(m[i * 3], m[i * 3 + 1], m[i * 3 + 2]), alpha);
}
}

extern "C" {
// A mesh's normals indexed by its own NormalIdx (normalidx, builtin/geom/polymesh/init.go:47-57),
// and its RayBias (polymesh.go:19); geom is the insertion index.
void orc_scene_mesh_set_normals(void* sp, int32_t geom, const float* normals, int32_t nnormals, const uint32_t* nidx) {
    Mesh* m = ((Scene*)sp)->geoms[geom];
    m->normals.resize(nnormals);
    std::memcpy(m->normals.data(), normals, sizeof(float) * 3 * (size_t)nnormals);
    m->normalidx.assign(nidx, nidx + (size_t)m->facecount * 3);
}
void orc_scene_mesh_set_ray_bias(void* sp, int32_t geom, float bias) { ((Scene*)sp)->geoms[geom]->RayBias = bias; }
}
