// Render object and its C-ABI entry points (vgpu_render_*).
//
// One pass on one stream: camera rays; then per recursion level (one level
// without perfect mirrors, four with) closest hit, shade (shadow-ray slots
// and the mirror rays of the next level), shadow any-hit over every slot and
// resolve; finally the levels are folded bottom-up into the film.  All
// buffers are allocated once at creation; a pass allocates nothing except,
// in host-pointer mode, a staging area the first time it is needed.

#include <array>
#include <cmath>
#include <new>

#include "scene_internal.h"

using vgpu::fail;
using namespace vgpu_detail;

namespace {

// RasterXY enumeration tables (math/ldseq/raster.go:10-40) for 2^m x 2^m elementary intervals.
// Index bit k of the (VdC, Sobol) sequence moves the top m bits of x and y by image(k); the
// lookup needs the images of bits 2m.. (fwd) and the GF(2) inverse of the images of bits 0..2m-1.
uint64_t vdc52(uint64_t i) {
    uint64_t r = 0;
    for (int b = 0; b < 64; b++) if ((i >> b) & 1) r |= 1ull << (63 - b);
    return r >> 12;
}
uint64_t sobol52(uint64_t i) {
    uint64_t r = 0;
    for (uint64_t v = 1ull << 51; i != 0; i >>= 1) {
        if (i & 1) r ^= v;
        v ^= v >> 1;
    }
    return r;
}
uint64_t bit_image(int m, int k) {
    const uint64_t x = vdc52(1ull << k) >> (52 - m), y = sobol52(1ull << k) >> (52 - m);
    return (x << m) | y;
}
bool raster_tables(int m, uint64_t* fwd, int nfwd, uint64_t* inv) {
    const int n = 2 * m;
    for (int c = 0; c < nfwd; c++) fwd[c] = (n + c < 64) ? bit_image(m, n + c) : 0;
    // columns col[k] = image of low bit k; invert the n x n GF(2) matrix by elimination on [A | I]
    std::vector<uint64_t> rowsA(n), rowsI(n);
    for (int r = 0; r < n; r++) {
        uint64_t a = 0;
        for (int k = 0; k < n; k++) if ((bit_image(m, k) >> r) & 1) a |= 1ull << k;
        rowsA[r] = a;
        rowsI[r] = 1ull << r;
    }
    for (int c = 0; c < n; c++) {
        int piv = -1;
        for (int r = c; r < n; r++) if ((rowsA[r] >> c) & 1) { piv = r; break; }
        if (piv < 0) return false;
        std::swap(rowsA[piv], rowsA[c]);
        std::swap(rowsI[piv], rowsI[c]);
        for (int r = 0; r < n; r++)
            if (r != c && ((rowsA[r] >> c) & 1)) { rowsA[r] ^= rowsA[c]; rowsI[r] ^= rowsI[c]; }
    }
    for (int c = 0; c < n; c++) {
        uint64_t v = 0;
        for (int r = 0; r < n; r++) if ((rowsI[r] >> c) & 1) v |= 1ull << r;
        inv[c] = v;
    }
    return true;
}

// builtin/filter/airy.go:35-79: J1 by the rational / asymptotic approximation the reference
// copies (Numerical Recipes bessj1, via gipsy bessel.c).
double bessel_j1(double x) {
    const double ax = std::fabs(x);
    if (ax < 8.0) {
        const double y = x * x;
        const double ans1 = x * (72362614232.0 + y * (-7895059235.0 + y * (242396853.1 + y * (-2972611.439 +
                                 y * (15704.48260 + y * (-30.16036606))))));
        const double ans2 = 144725228442.0 + y * (2300535178.0 + y * (18583304.74 + y * (99447.43394 +
                                 y * (376.9991397 + y * 1.0))));
        return ans1 / ans2;
    }
    const double z = 8.0 / ax, y = z * z, xx = ax - 2.356194491;
    const double ans1 = 1.0 + y * (0.183105e-2 + y * (-0.3516396496e-4 + y * (0.2457520174e-5 + y * (-0.240337019e-6))));
    const double ans2 = 0.04687499995 + y * (-0.2002690873e-3 + y * (0.8449199096e-5 + y * (-0.88228987e-6 +
                        y * 0.105787412e-6)));
    double ans = std::sqrt(0.636619772 / ax) * (std::cos(xx) * ans1 - z * std::sin(xx) * ans2);
    if (x < 0.0) ans = -ans;
    return ans;
}

// filter.CreateSampler (builtin/filter/filter.go:87-173): the filter on an n x n grid over
// [-w/2, w/2]^2 (u and v advanced by repeated addition), normalised to a pdf, its marginal over v
// and the conditionals, as running-sum CDFs.  out = cdfV (n) followed by cdfVU (n x n).
template <class F>
void create_sampler(int n, double w, F f, std::vector<double>& out) {
    std::vector<double> filter((size_t)n * n), pV(n, 0.0), cdfV(n), pdf((size_t)n * n), pVU((size_t)n * n);
    const double du = w / (double)(n - 1);
    double u = -w / 2;
    double F_ = 0;
    for (int j = 0; j < n; j++) {
        const double dv = w / (double)(n - 1);
        double v = -w / 2;
        for (int i = 0; i < n; i++) {
            const double fuv = f(u, v);
            filter[(size_t)j + (size_t)i * n] = fuv;
            F_ += fuv;
            v += dv;
        }
        u += du;
    }
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) pdf[(size_t)j * n + i] = filter[(size_t)j + (size_t)i * n] / F_;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) pV[j] += pdf[(size_t)j * n + i];
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) pVU[(size_t)j * n + i] = pdf[(size_t)j * n + i] / pV[j];
    out.assign((size_t)n + (size_t)n * n, 0.0);
    double pacc = 0;
    for (int i = 0; i < n; i++) {
        pacc += pV[i];
        out[i] = pacc;
    }
    for (int j = 0; j < n; j++) {
        double q = 0;
        for (int i = 0; i < n; i++) {
            q += pVU[(size_t)j * n + i];
            out[(size_t)n + (size_t)j * n + i] = q;
        }
    }
}

// ---- camera motion keys: TransformDecompMatrix4 (math/animdecomp.go:20-63) on the host, in float32
// with the reference's operation order (element (i, j) of a column-major matrix at [j * 4 + i]).

using Mat4 = std::array<float, 16>;

Mat4 mat_mul(const Mat4& a, const Mat4& b) {   // Matrix4Mul (math/matrix4.go:92-102): sums from +0
    Mat4 c{};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++) c[j * 4 + i] += a[k * 4 + i] * b[j * 4 + k];
    return c;
}

Mat4 mat_transpose(const Mat4& a) {
    Mat4 c{};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) c[i * 4 + j] = a[j * 4 + i];
    return c;
}

// Cofactor expansion of Matrix4Inverse / Matrix4Det (math/matrix4.go:116-180); the terms and their
// order fix the rounding, so they are written as the reference writes them.
Mat4 mat_cofactors(const Mat4& m) {
    Mat4 c{};
    vgpu::xf_cofactors(m.data(), c.data());
    return c;
}

float mat_det(const Mat4& m, const Mat4& c) { return m[0] * c[0] + m[1] * c[4] + m[2] * c[8] + m[3] * c[12]; }

// Matrix4PolarFactor (math/matrix4.go:343-363): Q <- (Q + Q^-T) / 2 until no entry grows by more than
// 1e-6 (the reference's one-sided test), at most 10 steps.
bool polar_factor(const Mat4& m, Mat4& Q) {
    Q = m;
    for (int it = 0; it < 10; it++) {
        const Mat4 c = mat_cofactors(Q);
        float det = mat_det(Q, c);
        if (det == 0.0f) return false;
        det = 1.0f / det;
        Mat4 inv{};
        for (int i = 0; i < 16; i++) inv[i] = c[i] * det;
        const Mat4 invT = mat_transpose(inv);
        Mat4 next{};
        for (int i = 0; i < 16; i++) next[i] = 0.5f * (Q[i] + invT[i]);
        bool same = true;
        for (int i = 0; i < 16 && same; i++) same = !(next[i] - Q[i] > 0.000001f);
        Q = next;
        if (same) return true;
    }
    return false;
}

// TransformDecompMatrix4: T from column 3; a negative determinant factored out as -I; polar factor Q;
// S = Q^T m; R = Matrix4ToQuat(Q) (math/quat.go:133-163).  Writes one kCamKeyFloats record.
void decompose_key(Mat4 m, float* rec) {
    const float sgn = mat_det(m, mat_cofactors(m)) >= 0.0f ? 1.0f : -1.0f;   // sign() (octanormal.go:16-23)
    for (int i = 0; i < vgpu::kCamKeyFloats; i++) rec[i] = 0.0f;
    rec[0] = m[12]; rec[1] = m[13]; rec[2] = m[14];
    m[12] = m[13] = m[14] = 0.0f;
    Mat4 negI{};
    for (int i = 0; i < 16; i++) negI[i] = -1.0f * ((i % 5 == 0) ? 1.0f : 0.0f);   // Matrix4Scale(-1, Identity)
    if (sgn < 0.0f) m = mat_mul(negI, m);
    Mat4 Q;
    float* R = rec + 3;
    float* S = rec + 7;
    if (!polar_factor(m, Q)) {   // the reference falls back to the identity rotation and scale
        R[3] = 1.0f;
        S[0] = S[5] = S[10] = S[15] = 1.0f;
        return;
    }
    Mat4 Sm = mat_mul(mat_transpose(Q), m);
    auto E = [&](int i, int j) { return Q[j * 4 + i]; };
    const float tr = E(0, 0) + E(1, 1) + E(2, 2);
    float& X = R[0]; float& Y = R[1]; float& Z = R[2]; float& W = R[3];
    if (tr > 0.0f) {
        const float q = std::sqrt(tr + 1.0f) * 2;
        W = 0.25f * q; X = (E(2, 1) - E(1, 2)) / q; Y = (E(0, 2) - E(2, 0)) / q; Z = (E(1, 0) - E(0, 1)) / q;
    } else if ((E(0, 0) > E(1, 1)) && (E(0, 0) > E(2, 2))) {
        const float q = std::sqrt(1.0f + E(0, 0) - E(1, 1) - E(2, 2)) * 2;
        W = (E(2, 1) - E(1, 2)) / q; X = 0.25f * q; Y = (E(0, 1) + E(1, 0)) / q; Z = (E(0, 2) + E(2, 0)) / q;
    } else if (E(1, 1) > E(2, 2)) {
        const float q = std::sqrt(1.0f + E(1, 1) - E(0, 0) - E(2, 2)) * 2;
        W = (E(0, 2) - E(2, 0)) / q; X = (E(0, 1) + E(1, 0)) / q; Y = 0.25f * q; Z = (E(1, 2) + E(2, 1)) / q;
    } else {
        const float q = std::sqrt(1.0f + E(2, 2) - E(0, 0) - E(1, 1)) * 2;
        W = (E(1, 0) - E(0, 1)) / q; X = (E(0, 2) + E(2, 0)) / q; Y = (E(1, 2) + E(2, 1)) / q; Z = 0.25f * q;
    }
    if (sgn < 0.0f) {
        Sm = mat_mul(negI, Sm);
        Sm[15] = 1.0f;
    }
    for (int i = 0; i < 16; i++) S[i] = Sm[i];
}

// TransformDecompLerp then TransformDecompToMatrix4 (math/animdecomp.go:66-80) of two key records, with
// QuatSlerp (math/quat.go:30-67) and QuatToMatrix4 (quat.go:80-104); k_camera runs the same steps.
Mat4 decomp_matrix(const float* a, const float* b, float t) {
    Mat4 M{};
    vgpu::xf_decomp_matrix(a, b, t, M.data());
    return M;
}


// Vec3Normalize (math/vec3_amd64.s:11-42) with the IEEE 1/sqrt estimate under the reference's
// Newton-Raphson step, as the renderer's device code computes it.
void normalize3(const float* a, float* out) {
    const float l2 = (a[1] * a[1] + a[0] * a[0]) + a[2] * a[2];
    const float g = 1.0f / std::sqrt(l2);
    const float gg = (g * g) * (l2 * 0.5f);
    const float r = (1.5f - gg) * g;
    for (int k = 0; k < 3; k++) out[k] = a[k] * r;
}
void cross3(const float* a, const float* b, float* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace

namespace vgpu_detail {

void vec3_normalize(const float* a, float* out) { normalize3(a, out); }
void vec3_cross(const float* a, const float* b, float* o) { cross3(a, b, o); }

int lookat_keys(const float* from, int nfrom, const float* to, int nto, const float* up, const float* roll, int nroll,
                std::vector<float>& out) {
    // Camera.calcLookatMatrices (builtin/camera/camera.go:109-186): one LocalToWorld key per key of the
    // array with more keys, the other one (and Roll) lerped at time i / n.
    const bool by_target = nto > nfrom;
    const int n = by_target ? nto : nfrom;
    out.assign((size_t)n * 16, 0.0f);
    auto key_lerp = [](const float* arr, int nk, float time, float* o, int width) {
        const float k = time * (float)(nk - 1);
        const float t = k - std::floor(k);
        const int k0 = (int)std::floor(k), k1 = (int)std::ceil(k);
        for (int c = 0; c < width; c++) o[c] = (1.0f - t) * arr[k0 * width + c] + t * arr[k1 * width + c];
    };
    for (int i = 0; i < n; i++) {
        const float time = (float)i / (float)n;
        float other[3], F[3], T[3];
        key_lerp(by_target ? from : to, by_target ? nfrom : nto, time, other, 3);
        const float* own = (by_target ? to : from) + (size_t)i * 3;
        for (int c = 0; c < 3; c++) { F[c] = by_target ? other[c] : own[c]; T[c] = by_target ? own[c] : other[c]; }
        float d[3] = {F[0] - T[0], F[1] - T[1], F[2] - T[2]}, W[3], u0[3], u[3], v0[3], v[3];
        normalize3(d, W);
        cross3(up, W, u0);
        normalize3(u0, u);
        cross3(u, W, v0);
        normalize3(v0, v);
        float rl = 0.0f;
        if (roll && nroll > 0) key_lerp(roll, nroll, time, &rl, 1);
        const float c = (float)std::cos((double)rl), sn = (float)std::sin((double)rl);
        float U[3], V[3];
        for (int k = 0; k < 3; k++) {
            U[k] = c * u[k] + sn * v[k];
            V[k] = -sn * u[k] + c * v[k];
        }
        Mat4 Tr{}, B{};
        Tr[0] = Tr[5] = Tr[10] = Tr[15] = 1.0f;
        Tr[12] = F[0]; Tr[13] = F[1]; Tr[14] = F[2];
        B[0] = U[0]; B[1] = U[1]; B[2] = U[2]; B[4] = V[0]; B[5] = V[1]; B[6] = V[2];
        B[8] = W[0]; B[9] = W[1]; B[10] = W[2]; B[15] = 1.0f;
        const Mat4 M = mat_mul(Tr, B);
        std::copy(M.begin(), M.end(), out.begin() + (size_t)i * 16);
    }
    return n;
}

void decompose_transform(const float* m, float* rec) {
    Mat4 a{};
    std::copy(m, m + 16, a.begin());
    decompose_key(a, rec);
}

bool mat4_inverse(const float* m, float* out) {
    // Matrix4Inverse (math/matrix4.go:116-176): cofactors scaled by 1 / det; false when det == 0
    Mat4 a{};
    std::copy(m, m + 16, a.begin());
    const Mat4 c = mat_cofactors(a);
    float det = mat_det(a, c);
    if (det == 0.0f) return false;
    det = 1.0f / det;
    for (int i = 0; i < 16; i++) out[i] = c[i] * det;
    return true;
}

}  // namespace vgpu_detail

struct vgpu_render {
    vgpu_scene* scene = nullptr;
    vgpu::RenderView rv{};
    uint32_t width = 0, height = 0;
    vgpu_material* d_materials = nullptr;
    vgpu::DevLight* d_lights = nullptr;
    float4* d_rays = nullptr;      // camera rays and even levels' mirror rays (2 float4 per pixel)
    float4* d_rays2 = nullptr;     // odd levels' mirror rays
    vgpu::LevelRec* d_recs = nullptr;   // per pixel, per level
    unsigned long long* d_counts = nullptr;   // rays traced: closest hit, shadow (RenderView::counts)
    double* d_filter = nullptr;    // pixel filter sampler tables
    float* d_camkeys = nullptr;    // camera motion keys (kCamKeyFloats each)
    float cam_M[16] = {};          // the vgpu_camera matrix given at creation
    float4* d_pix = nullptr;       // per pixel: D, lambda
    int4* d_hits32 = nullptr;      // extended hits (2 int4 per pixel)
    float4* d_slot_rays = nullptr;
    float4* d_slot_rgb = nullptr;
    int4* d_slot_hits = nullptr;
    float4* d_groups = nullptr;
    float4* d_pixshade = nullptr;
    void* d_stage = nullptr;       // host-pointer mode: scrambles, film, samples
    size_t stage_bytes = 0;
    vgpu_detail::GraphCache graph; // VGPU_GRAPH passes
    bool any_mirror = false;       // a material has a perfect-mirror lobe (the reference recurses to level 3)

    ~vgpu_render() {
        int prev = 0;
        if (hipGetDevice(&prev) == hipSuccess && scene) (void)hipSetDevice(scene->device);
        void* ptrs[] = {d_materials, d_lights, d_rays, d_rays2, d_recs, d_filter, d_camkeys, d_pix, d_hits32, d_slot_rays,
                        d_slot_rgb, d_slot_hits, d_groups, d_pixshade, d_stage, d_counts};
        for (void* p : ptrs) if (p) (void)hipFree(p);
        (void)hipSetDevice(prev);
    }
};

extern "C" {

int vgpu_render_create(vgpu_render** out, vgpu_scene* s, const vgpu_material* materials, uint32_t nmaterials,
                       const vgpu_light* lights, uint32_t nlights, const vgpu_camera* cam, uint32_t width,
                       uint32_t height) {
    if (!out || !s || !materials || !cam) return fail(VGPU_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (!s->committed) return fail(VGPU_ERR_STATE, "render_create before scene commit");
    if (width == 0 || height == 0 || width > 4096 || height > 4096)
        return fail(VGPU_ERR_ARG, "image size %ux%u outside [1,4096] (RasterXY uses 2^12 intervals)", width, height);
    if (s->max_material >= (int32_t)nmaterials)
        return fail(VGPU_ERR_ARG, "a mesh refers to material %d of %u", s->max_material, nmaterials);
    bool any_mirror = false;
    for (uint32_t i = 0; i < nmaterials; i++) {
        const vgpu_material& m = materials[i];
        const float rough = m.has_spec1_roughness ? m.spec1_roughness : 0.5f;
        const float sw = m.has_spec1_strength ? m.spec1_strength : 0.0f;
        if (sw > 0.0f && rough == 0.0f) any_mirror = true;
        if (m.fresnel_model < 0 || m.fresnel_model > VGPU_FRESNEL_METAL)
            return fail(VGPU_ERR_ARG, "material %u: fresnel_model %d is not 0, 1 (Dielectric) or 2 (Metal)", i,
                        m.fresnel_model);
        const float dw = m.has_diffuse_strength ? m.diffuse_strength : 0.0f;
        if (dw + sw == 0.0f) return fail(VGPU_ERR_ARG, "material %u has no weight (std.go:141-143 panics)", i);
    }
    int ngeoms = (int)s->meshes.size();
    std::vector<vgpu::DevLight> dl(nlights);
    int max_ns = 1;
    for (uint32_t l = 0; l < nlights; l++) {
        const vgpu_light& L = lights[l];
        if (L.material < 0 || L.material >= (int32_t)nmaterials) return fail(VGPU_ERR_ARG, "light %u material", l);
        if (L.samples < 0 || L.samples > 8) return fail(VGPU_ERR_ARG, "light %u samples must be in [0,8]", l);
        if (L.geom < 0 || L.geom >= ngeoms) return fail(VGPU_ERR_ARG, "light %u geom %d out of range", l, L.geom);
        if (L.type != VGPU_LIGHT_TRI && L.type != VGPU_LIGHT_DISK && L.type != VGPU_LIGHT_SPHERE)
            return fail(VGPU_ERR_ARG, "light %u type %d", l, L.type);
        if (L.type != VGPU_LIGHT_TRI && !(L.radius > 0.0f)) return fail(VGPU_ERR_ARG, "light %u radius must be > 0", l);
        if ((L.type == VGPU_LIGHT_SPHERE) != s->meshes[L.geom].sphere)
            return fail(VGPU_ERR_ARG, "light %u: a sphere light's geom must be its vgpu_scene_add_sphere geom", l);
        const float rad = L.type == VGPU_LIGHT_TRI ? 0.0f : L.radius;
        dl[l].P0 = make_float4(L.P0[0], L.P0[1], L.P0[2], rad);
        dl[l].P1 = make_float4(L.P1[0], L.P1[1], L.P1[2], 0.0f);
        dl[l].P2 = make_float4(L.P2[0], L.P2[1], L.P2[2], 0.0f);
        dl[l].material = L.material;
        dl[l].samples = L.samples;
        dl[l].type = L.type;
        dl[l].geom = -1;
        for (int slot = 0; slot < ngeoms; slot++) if (s->order[slot] == L.geom) dl[l].geom = slot;
        max_ns = std::max(max_ns, 1 << L.samples);
    }
    vgpu_render* r = new (std::nothrow) vgpu_render();
    if (!r) return fail(VGPU_ERR_NOMEM, "out of host memory");
    r->scene = s;
    r->width = width;
    r->height = height;
    vgpu::RenderView& rv = r->rv;
    rv.sv = s->view;
    rv.width = (int32_t)width;
    rv.height = (int32_t)height;
    rv.nlights = (int32_t)nlights;
    rv.groups = 2 * (int32_t)nlights;
    rv.max_half = std::max(1, max_ns / 2);
    rv.slots = std::max(1, rv.groups * 2 * rv.max_half);
    std::memcpy(rv.M, cam->M, sizeof(rv.M));
    std::memcpy(r->cam_M, cam->M, sizeof(r->cam_M));
    rv.tan_theta_focal = cam->tan_theta_focal;
    rv.aspect = cam->aspect;
    rv.focal = cam->focal;
    if (!(cam->radius >= 0.0f) || !std::isfinite(cam->radius)) return fail(VGPU_ERR_ARG, "camera radius must be >= 0");
    rv.lens_radius = cam->radius;
    rv.filt = nullptr;
    rv.filt_n = 0;
    rv.filt_w = 0.0;
    rv.max_level = any_mirror ? vgpu::kMaxLevel : 0;
    rv.path_depth = 0;
    rv.rec_levels = vgpu::kMaxLevel + 1;
    r->any_mirror = any_mirror;
    if (!raster_tables(12, rv.raster_fwd, 52, rv.raster_inv)) { delete r; return fail(VGPU_ERR_ARG, "RasterXY tables"); }
    std::lock_guard<std::recursive_mutex> lk(s->mu);
    DeviceGuard dg(s->device);
    const size_t np = (size_t)width * height, ns = np * (size_t)rv.slots;
    auto alloc = [&](void** p, size_t bytes) { return hipMalloc(p, std::max<size_t>(bytes, 16)); };
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = alloc((void**)&r->d_materials, sizeof(vgpu_material) * nmaterials);
    if (e == hipSuccess) e = alloc((void**)&r->d_lights, sizeof(vgpu::DevLight) * nlights);
    if (e == hipSuccess) e = alloc((void**)&r->d_rays, np * 32);
    if (e == hipSuccess && any_mirror) e = alloc((void**)&r->d_rays2, np * 32);
    if (e == hipSuccess) e = alloc((void**)&r->d_recs, np * sizeof(vgpu::LevelRec) * (vgpu::kMaxLevel + 1));
    if (e == hipSuccess) e = alloc((void**)&r->d_pix, np * 16);
    if (e == hipSuccess) e = alloc((void**)&r->d_hits32, np * 32);
    if (e == hipSuccess) e = alloc((void**)&r->d_slot_rays, ns * 32);
    if (e == hipSuccess) e = alloc((void**)&r->d_slot_rgb, ns * 16);
    if (e == hipSuccess) e = alloc((void**)&r->d_slot_hits, ns * 16);
    if (e == hipSuccess) e = alloc((void**)&r->d_groups, np * (size_t)rv.groups * 16);
    if (e == hipSuccess) e = alloc((void**)&r->d_pixshade, np * 64);
    if (e == hipSuccess) e = alloc((void**)&r->d_counts, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(r->d_counts, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemcpy(r->d_materials, materials, sizeof(vgpu_material) * nmaterials, hipMemcpyHostToDevice);
    if (e == hipSuccess && nlights) e = hipMemcpy(r->d_lights, dl.data(), sizeof(vgpu::DevLight) * nlights, hipMemcpyHostToDevice);
    if (e != hipSuccess) { delete r; return fail(VGPU_ERR_HIP, "render buffers: %s", hipGetErrorString(e)); }
    rv.normals = s->d_normals;
    rv.shaderidx = s->d_shaderidx;
    rv.shaders = s->d_shaders;
    rv.gshade = s->d_gshade;
    rv.materials = r->d_materials;
    rv.lights = r->d_lights;
    rv.counts = r->d_counts;
    // size the traversal spill area for the larger (shadow) launch
    vgpu::TraceLaunch L{s->cfg.block, s->cfg.lds_stack, s->cfg.waves_per_cu};
    const vgpu::TraceLaunch ext = s->any_motion ? vgpu::kExtLaunchMotion : vgpu::kExtLaunch;
    int rc = ensure_spill(s, std::max(vgpu::trace_grid_threads(L, ns, s->num_cus),
                                      vgpu::trace_grid_threads(ext, np, s->num_cus)));
    if (rc) { delete r; return rc; }
    *out = r;
    return VGPU_OK;
}

int vgpu_render_destroy(vgpu_render* r) {
    delete r;
    return VGPU_OK;
}

int vgpu_render_pass(vgpu_render* r, uint32_t flags, const uint64_t* scrambles, uint32_t iter, float* film,
                     float* sample_out, void* stream, uint64_t* shadow_rays_out) {
    if (!r || !scrambles || !film) return fail(VGPU_ERR_ARG, "NULL argument");
    if (iter == 0) return fail(VGPU_ERR_ARG, "iter is 1-based");
    vgpu_scene* s = r->scene;
    std::lock_guard<std::recursive_mutex> lk(s->mu);
    DeviceGuard dg(s->device);
    hipStream_t st = (hipStream_t)stream;
    const size_t np = (size_t)r->width * r->height, ns = np * (size_t)r->rv.slots;
    const uint64_t* d_scr = scrambles;
    float* d_film = film;
    float* d_samp = sample_out;
    const bool host = (flags & VGPU_TRACE_DEVICE_PTRS) == 0;
    if (host) {
        const size_t need = np * 48 + np * 12 + (sample_out ? np * 12 : 0);
        if (need > r->stage_bytes) {
            if (r->d_stage) { HIP_TRY(hipFree(r->d_stage)); r->d_stage = nullptr; r->stage_bytes = 0; }
            HIP_TRY(hipMalloc(&r->d_stage, need));
            r->stage_bytes = need;
        }
        char* b = (char*)r->d_stage;
        d_scr = (const uint64_t*)b;
        d_film = (float*)(b + np * 48);
        d_samp = sample_out ? (float*)(b + np * 60) : nullptr;
        HIP_TRY(hipMemcpyAsync((void*)d_scr, scrambles, np * 48, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_film, film, np * 12, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemsetAsync(s->d_err, 0, sizeof(int32_t), st));
    }
    const bool graph = (flags & VGPU_GRAPH) != 0;
    if (graph && host) return fail(VGPU_ERR_ARG, "VGPU_GRAPH needs VGPU_TRACE_DEVICE_PTRS");
    if (graph && s->timing) return fail(VGPU_ERR_STATE, "VGPU_GRAPH with kernel timing enabled");
    if (graph && s->cfg.sort > 0) { const int rc = ensure_sort(s, ns); if (rc) return rc; }
    uint64_t shadow_count = 0;
    // The pass's launches, on the caller's stream or on the capture stream of the pass's graph.
    auto enqueue = [&](hipStream_t st) -> int {
        hipError_t e = vgpu::launch_camera(r->rv, d_scr, iter, r->d_rays, r->d_pix, st);
        if (e != hipSuccess) return fail(VGPU_ERR_HIP, "camera: %s", hipGetErrorString(e));
        for (int level = 0; level <= r->rv.max_level; level++) {
            float4* rin = (level & 1) ? r->d_rays2 : r->d_rays;
            float4* rout = level < r->rv.max_level ? ((level & 1) ? r->d_rays : r->d_rays2) : nullptr;
            e = vgpu::launch_trace_ext(s->view, s->any_motion, rin, np, r->d_hits32, s->d_spT, s->d_spN, s->spill_threads, s->d_err,
                                       s->num_cus, st);
            if (e != hipSuccess) return fail(VGPU_ERR_HIP, "closest hit: %s", hipGetErrorString(e));
            e = vgpu::launch_shade(r->rv, d_scr, iter, level, rin, r->d_hits32, r->d_pix, r->d_slot_rays, r->d_slot_rgb,
                                   r->d_groups, r->d_pixshade, rout, r->d_recs, st);
            if (e != hipSuccess) return fail(VGPU_ERR_HIP, "shade: %s", hipGetErrorString(e));
            if (r->rv.groups > 0) {   // without lights no slot holds a shadow ray
                const int rc = enqueue_trace(s, VGPU_TRACE_SHADOW | VGPU_TRACE_DEVICE_PTRS, r->d_slot_rays, ns,
                                             r->d_slot_hits, nullptr, nullptr, st, s->cfg);
                if (rc) return rc;
            }
            const bool direct = r->rv.max_level == 0;
            e = vgpu::launch_resolve(r->rv, level, iter, r->d_slot_rgb, r->d_slot_hits, r->d_groups, r->d_pixshade, r->d_recs,
                                     direct ? d_film : nullptr, direct ? d_samp : nullptr, st);
            if (e != hipSuccess) return fail(VGPU_ERR_HIP, "resolve: %s", hipGetErrorString(e));
            if (host && shadow_rays_out && r->rv.groups > 0) {
                // count active slots (diagnostic, host mode only)
                std::vector<float4> rgb(ns);
                HIP_TRY(hipMemcpyAsync(rgb.data(), r->d_slot_rgb, ns * 16, hipMemcpyDeviceToHost, st));
                HIP_TRY(hipStreamSynchronize(st));
                for (const float4& v : rgb) shadow_count += v.w != 0.0f;
            }
        }
        if (r->rv.max_level > 0) {
            e = vgpu::launch_combine(r->rv, iter, r->d_recs, d_film, d_samp, st);
            if (e != hipSuccess) return fail(VGPU_ERR_HIP, "combine: %s", hipGetErrorString(e));
        }
        return VGPU_OK;
    };
    const int rc = graph ? r->graph.run(st, enqueue) : enqueue(st);
    if (rc) return rc;
    if (host) {
        HIP_TRY(hipMemcpyAsync(film, d_film, np * 12, hipMemcpyDeviceToHost, st));
        if (sample_out) HIP_TRY(hipMemcpyAsync(sample_out, d_samp, np * 12, hipMemcpyDeviceToHost, st));
        int32_t err = 0;
        HIP_TRY(hipMemcpyAsync(&err, s->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (err) return fail(VGPU_ERR_STACK, "a ray needed more than %d traversal stack entries", VGPU_MAX_STACK);
        if (shadow_rays_out) *shadow_rays_out = shadow_count;
    }
    return VGPU_OK;
}

int vgpu_render_stats(vgpu_render* r, uint64_t* closest_rays, uint64_t* shadow_rays) {
    if (!r) return fail(VGPU_ERR_ARG, "render is NULL");
    DeviceGuard dg(r->scene->device);
    HIP_TRY(hipDeviceSynchronize());
    unsigned long long c[2] = {0, 0};
    HIP_TRY(hipMemcpy(c, r->d_counts, sizeof(c), hipMemcpyDeviceToHost));
    if (closest_rays) *closest_rays = c[0];
    if (shadow_rays) *shadow_rays = c[1];
    return VGPU_OK;
}

int vgpu_render_set_path_depth(vgpu_render* r, uint32_t depth) {
    if (!r) return fail(VGPU_ERR_ARG, "render is NULL");
    if (depth > (uint32_t)vgpu::kMaxPathDepth)
        return fail(VGPU_ERR_ARG, "path depth %u outside [0,%d]", depth, vgpu::kMaxPathDepth);
    vgpu_scene* s = r->scene;
    std::lock_guard<std::recursive_mutex> lk(s->mu);
    DeviceGuard dg(s->device);
    const size_t np = (size_t)r->width * r->height;
    const int levels = std::max(vgpu::kMaxLevel + 1, (int)depth);
    if (levels > r->rv.rec_levels) {
        vgpu::LevelRec* recs = nullptr;
        HIP_TRY(hipMalloc((void**)&recs, np * sizeof(vgpu::LevelRec) * (size_t)levels));
        if (r->d_recs) (void)hipFree(r->d_recs);
        r->d_recs = recs;
        r->rv.rec_levels = levels;
    }
    if (depth > 1 && !r->d_rays2) HIP_TRY(hipMalloc((void**)&r->d_rays2, np * 32));
    r->rv.path_depth = (int32_t)depth;
    r->rv.max_level = depth > 0 ? (int32_t)depth - 1 : (r->any_mirror ? vgpu::kMaxLevel : 0);
    return VGPU_OK;
}

int vgpu_render_set_filter(vgpu_render* r, int32_t type, float width, int32_t res, float peak) {
    if (!r) return fail(VGPU_ERR_ARG, "render is NULL");
    if (type != VGPU_FILTER_BOX && type != VGPU_FILTER_GAUSSIAN && type != VGPU_FILTER_AIRY)
        return fail(VGPU_ERR_ARG, "filter type %d", type);
    DeviceGuard dg(r->scene->device);
    if (type == VGPU_FILTER_BOX) {
        r->rv.filt = nullptr;
        r->rv.filt_n = 0;
        return VGPU_OK;
    }
    const bool gauss = type == VGPU_FILTER_GAUSSIAN;
    // node defaults (filter.go:14-25)
    const float W = width > 0.0f ? width : (gauss ? 2.0f : 6.0f);
    const int n = res > 0 ? res : (gauss ? 17 : 49);
    const float pk = peak > 0.0f ? peak : 4.0f;
    if (n < 2 || n > 1024) return fail(VGPU_ERR_ARG, "filter resolution %d outside [2,1024]", n);
    std::vector<double> tab;
    if (gauss) {
        // GaussianFilter.PreRender (gauss.go:34-58); the exponent reads -(r^2)/2 * sigma^2 as written
        create_sampler(n, (double)W, [W](double x, double y) {
            const double q = std::sqrt(x * x + y * y);
            if (q > (double)(W / 2)) return 0.0;
            const double sigma = (double)(1.0 / std::sqrt((double)W));
            return (1 / (2 * M_PI * (sigma * sigma))) * std::exp(-(x * x + y * y) / 2 * (sigma * sigma));
        }, tab);
    } else {
        // AiryFilter.PreRender (airy.go:82-106); at q = 0 the value is 0/0 = NaN, as in the reference
        create_sampler(n, (double)W, [W, pk](double x, double y) {
            const double q = std::sqrt(x * x + y * y);
            if (q > (double)(W / 2)) return 0.0;
            const double lambda = 550.0, N = 5.6;
            const double v = (20000 / (double)W) * (M_PI * q) / (lambda * N);
            const double t = 2 * bessel_j1(v) / v;
            return (double)pk * (t * t);
        }, tab);
    }
    if (r->d_filter) { HIP_TRY(hipFree(r->d_filter)); r->d_filter = nullptr; }
    HIP_TRY(hipMalloc((void**)&r->d_filter, tab.size() * sizeof(double)));
    HIP_TRY(hipMemcpy(r->d_filter, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    r->rv.filt = r->d_filter;
    r->rv.filt_n = n;
    r->rv.filt_w = (double)W;
    return VGPU_OK;
}

int vgpu_render_set_camera_keys(vgpu_render* r, const float* local_to_world, uint32_t nkeys) {
    if (!r) return fail(VGPU_ERR_ARG, "render is NULL");
    if (nkeys > 0 && !local_to_world) return fail(VGPU_ERR_ARG, "camera keys are NULL");
    if (nkeys > (uint32_t)vgpu::kMaxCamKeys) return fail(VGPU_ERR_ARG, "%u camera keys (at most %d)", nkeys, vgpu::kMaxCamKeys);
    for (size_t i = 0; i < (size_t)nkeys * 16; i++)
        if (!std::isfinite(local_to_world[i])) return fail(VGPU_ERR_ARG, "camera key matrices must be finite");
    DeviceGuard dg(r->scene->device);
    std::vector<float> recs((size_t)std::max<uint32_t>(nkeys, 1) * vgpu::kCamKeyFloats);
    for (uint32_t k = 0; k < nkeys; k++) {
        Mat4 m;
        for (int i = 0; i < 16; i++) m[i] = local_to_world[(size_t)k * 16 + i];
        decompose_key(m, &recs[(size_t)k * vgpu::kCamKeyFloats]);
    }
    if (nkeys == 1) {   // one key: the matrix of every sample (the key lerped with itself, camera.go:224-236)
        const Mat4 M = decomp_matrix(recs.data(), recs.data(), 0.0f);
        for (int i = 0; i < 16; i++) r->rv.M[i] = M[i];
    } else {            // none: the vgpu_camera matrix; two or more: per sample from the keys
        std::memcpy(r->rv.M, r->cam_M, sizeof(r->rv.M));
    }
    if (r->d_camkeys) { HIP_TRY(hipFree(r->d_camkeys)); r->d_camkeys = nullptr; }
    r->rv.cam_keys = nullptr;
    r->rv.cam_nkeys = 0;
    if (nkeys >= 2) {
        HIP_TRY(hipMalloc((void**)&r->d_camkeys, recs.size() * sizeof(float)));
        HIP_TRY(hipMemcpy(r->d_camkeys, recs.data(), recs.size() * sizeof(float), hipMemcpyHostToDevice));
        r->rv.cam_keys = r->d_camkeys;
        r->rv.cam_nkeys = (int32_t)nkeys;
    }
    return VGPU_OK;
}

}  // extern "C"
