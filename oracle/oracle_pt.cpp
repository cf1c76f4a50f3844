// ORACLE — test infrastructure only.
//
// CPU restatement of path mode (plain sampling, and guided sampling with a fixed
// proposal given as lobes): the GPU path
// tracer's estimator, written independently and in the same order of float
// operations, so that plain-mode paths can be compared bit for bit.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may call it.
//
// Path mode has no reference counterpart (the reference shades one camera hit
// with direct light).  Its definition, restated here:
//   u(pixel, pass, dim)  lowbias32 hash chain: k = mix(mix(pixel ^ mix(s_lo)) ^ (pass + s_hi)),
//                        u = (mix(k + dim * 0x9e3779b9) >> 8) * 2^-24
//   dims                 0, 1: pixel jitter; vertex d uses 2 + 8d + {0: light pick,
//                        1, 2: point on the light, 3: mixture pick, 4, 5: direction}
//   camera               the reference's pinhole (builtin/camera/camera.go:244-284)
//                        through (x + u0, y + u1)
//   surfaces             Lambertian rho, emission Le (one- or two-sided); shading
//                        normal = geometric normal turned towards the ray
//   emission             found by a sampled direction: weight p / (p + pdf_light)
//   NEE                  one emissive triangle by area x lum(Le), uniform point,
//                        weight pdf_light / (pdf_light + p), shadow ray from the offset
//                        origin to 0.9999 of its distance to the point
//   continuation         cosine-weighted direction, throughput rho cos / (pi p)
//   depth                up to max_depth segments from the camera
// Closest-hit and shadow traversal go through the traversal oracle (oracle_core.h).

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_core.h"

namespace opt {

using orc::Scene;

struct F3 { float x, y, z; };
static inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
static inline F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline F3 mul(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline float dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline F3 cross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float lum(F3 c) { return (0.2126f * c.x + 0.7152f * c.y) + 0.0722f * c.z; }
static const float INV_PI = 0.318309886f;

static inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
static inline float rnd(uint32_t key, uint32_t dim) {
    return (float)(mix32(key + dim * 0x9e3779b9u) >> 8) * 5.9604645e-08f;
}

// sin/cos(2 pi u): quadrant, octant reflection, Taylor polynomials (Horner) on [0, pi/4]
static void sincos2pi(float u, float& s, float& c) {
    const float t = u * 4.0f;
    const int q = (int)t;
    const float f = t - (float)q;
    const bool sw = f > 0.5f;
    const float a = (sw ? 1.0f - f : f) * 1.57079637f;
    const float a2 = a * a;
    const float S[6] = {-2.50521084e-08f, 2.75573192e-06f, -1.98412698e-04f, 8.33333377e-03f, -1.66666672e-01f, 1.0f};
    const float Cc[7] = {2.08767570e-09f, -2.75573192e-07f, 2.48015876e-05f, -1.38888892e-03f, 4.16666679e-02f, -0.5f, 1.0f};
    float ps = S[0];
    for (int k = 1; k < 6; k++) ps = ps * a2 + S[k];
    const float sa = ps * a;
    float pc = Cc[0];
    for (int k = 1; k < 7; k++) pc = pc * a2 + Cc[k];
    const float sq = sw ? pc : sa, cq = sw ? sa : pc;
    switch (q) {
        case 0: s = sq; c = cq; break;
        case 1: s = cq; c = -sq; break;
        case 2: s = -sq; c = -cq; break;
        default: s = -cq; c = sq; break;
    }
}

static void onb(F3 n, F3& t1, F3& t2) {   // Duff et al. 2017
    const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    t1 = f3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    t2 = f3(b, sign + n.y * n.y * a, -n.y);
}

static float off1(float p, float n) {      // Waechter & Binder 2019
    const int of = (int)(256.0f * n);
    uint32_t u;
    std::memcpy(&u, &p, 4);
    u = (uint32_t)((int32_t)u + (p < 0.0f ? -of : of));
    float pi;
    std::memcpy(&pi, &u, 4);
    return std::fabs(p) < 0.03125f ? p + 1.52587891e-05f * n : pi;
}

static inline orc::Vec3 vec(F3 a) { orc::Vec3 v; v[0] = a.x; v[1] = a.y; v[2] = a.z; return v; }

// Guided sampling with a fixed proposal: the same 8 vMF lobes (mu xyz, kappa, weight) at every
// vertex (the GPU evaluates them per vertex from the network).  Definitions as in the GPU kernels:
// q(w) = sum_k w_k kappa_k / (2 pi (1 - exp(-2 kappa_k))) exp(kappa_k (mu_k . w - 1)); a lobe is
// picked by its weight (dim 6), then cos theta = 1 + log(u + (1 - u) exp(-2 kappa)) / kappa
// (Jakob 2012) around mu with the azimuth of dim 5.
static float vmf_c(float k) { return k / (6.28318548f * -std::expm1(-2.0f * k)); }
static float guide_pdf(const float* lob, F3 w) {
    float q = 0.0f;
    for (int k = 0; k < 8; k++) {
        const float* L = lob + 5 * k;
        const float c = (L[0] * w.x + L[1] * w.y) + L[2] * w.z;
        q += L[4] * (vmf_c(L[3]) * std::exp(L[3] * (c - 1.0f)));
    }
    return q;
}

struct Mat { F3 a, e; bool one_sided; float pdf_area; };
struct Light { F3 v0, e1, e2, n, Le; float cdf, pdf_area; bool one_sided; };

struct PT {
    Scene* scene = nullptr;
    std::vector<Mat> mats;
    std::vector<Light> lights;
    std::vector<double> lw;
    bool finalized = false;
    int W = 0, H = 0, D = 1;
    float M[16];
    float ttf = 0, aspect = 1, focal = 1;
    uint32_t s_lo = 0, s_hi = 0;
    float alpha = 0.0f;
    float lobes[40];
    bool no_nee = false;   // diagnostic: BSDF/guide sampling only, emission weight 1
    // traversal work of the traced segments (closest hit, then shadow): rays, node visits, triangle tests
    mutable std::atomic<int64_t> work[6] = {};

    void finalize() {
        if (finalized) return;
        double wsum = 0.0, acc = 0.0;
        for (double x : lw) wsum += x;
        for (size_t i = 0; i < lights.size(); i++) {
            acc += lw[i];
            lights[i].cdf = i + 1 == lights.size() ? 1.0f : (float)(acc / wsum);
            lights[i].pdf_area = (float)((double)lum(lights[i].Le) / wsum);
        }
        for (Mat& m : mats) {
            const float le = lum(m.e);
            m.pdf_area = (le > 0.0f && wsum > 0.0) ? (float)((double)le / wsum) : 0.0f;
        }
        finalized = true;
    }

    bool closest(orc::Task* tk, F3 P, F3 D, int& geom, int& elem, float& t, float& U, float& V, float& Wb) const {
        orc::Ray r{};
        r.P = vec(P); r.D = vec(D);
        r.Tclosest = orc::kInf; r.Type = orc::RayTypeCamera; r.Time = 0;
        orc::RaySetup(r);
        tk->StackTop = 0; r.task = tk;
        orc::SceneHit sh{}; sh.geom = -1; sh.hr.elem = -1;
        const bool hit = scene->Trace(r, sh);
        work[0]++; work[1] += r.NodesT; work[2] += r.TrisT;
        if (!hit) return false;
        geom = sh.geom; elem = sh.hr.elem; t = r.Tclosest; U = sh.hr.U; V = sh.hr.V; Wb = sh.hr.W;
        return true;
    }
    bool occluded(orc::Task* tk, F3 P, F3 D, float tmax) const {
        orc::Ray r{};
        r.P = vec(P); r.D = vec(D);
        r.Tclosest = tmax; r.Type = orc::RayTypeShadow; r.Time = 0;
        orc::RaySetup(r);
        tk->StackTop = 0; r.task = tk;
        orc::SceneHit sh{}; sh.geom = -1;
        const bool hit = scene->Trace(r, sh);
        work[3]++; work[4] += r.NodesT; work[5] += r.TrisT;
        return hit;
    }

    // One path (float arithmetic in the GPU's order); returns this pass's pixel radiance.
    F3 path(orc::Task* tk, int pixel, uint32_t pass, int64_t& rays) const {
        const uint32_t key = mix32(mix32((uint32_t)pixel ^ mix32(s_lo)) ^ (pass + s_hi));
        const float fx = (float)(pixel % W) + rnd(key, 0);
        const float fy = (float)(pixel / W) + rnd(key, 1);
        const float Sx = 2.0f * fx / (float)W - 1.0f;
        const float Sy = 1.0f - 2.0f * fy / (float)H;
        const float camu = Sx * ttf;
        const float camv = Sy * (ttf / aspect);
        const float sz = -focal;
        const F3 d = f3((M[0] * camu + M[4] * camv) + M[8] * sz, (M[1] * camu + M[5] * camv) + M[9] * sz,
                        (M[2] * camu + M[6] * camv) + M[10] * sz);
        const float l = std::sqrt(dot(d, d));
        F3 Dr = f3(d.x / l, d.y / l, d.z / l);
        F3 Pr = f3(M[12], M[13], M[14]);
        F3 beta = f3(1.0f, 1.0f, 1.0f);
        float pdf_prev = 0.0f;
        F3 L = f3(0.0f, 0.0f, 0.0f);
        for (int depth = 0; depth < D; depth++) {
            int g, elem;
            float t, U, V, Wb;
            rays++;
            if (!closest(tk, Pr, Dr, g, elem, t, U, V, Wb)) break;
            const orc::Mesh& mesh = *scene->geoms[g];
            const orc::Vec3& A = mesh.verts[mesh.idxp[3 * (size_t)elem]];
            const orc::Vec3& B = mesh.verts[mesh.idxp[3 * (size_t)elem + 1]];
            const orc::Vec3& Cv = mesh.verts[mesh.idxp[3 * (size_t)elem + 2]];
            const F3 a = f3(A[0], A[1], A[2]), b = f3(B[0], B[1], B[2]), c = f3(Cv[0], Cv[1], Cv[2]);
            const F3 P = f3((U * a.x + V * b.x) + Wb * c.x, (U * a.y + V * b.y) + Wb * c.y, (U * a.z + V * b.z) + Wb * c.z);
            const F3 ng = cross(sub(b, a), sub(c, a));
            const float nl = std::sqrt(dot(ng, ng));
            const F3 Ng = f3(ng.x / nl, ng.y / nl, ng.z / nl);
            const float cosi = dot(Ng, Dr);
            const bool front = cosi < 0.0f;
            const F3 N = front ? Ng : f3(-Ng.x, -Ng.y, -Ng.z);
            const int si = mesh.shaderidx.empty() ? 0 : mesh.shaderidx[elem];
            const Mat& m = mats[mesh.shaders[si]];
            if ((m.e.x > 0.0f || m.e.y > 0.0f || m.e.z > 0.0f) && (!m.one_sided || front)) {
                float wB = 1.0f;
                if (depth > 0 && !no_nee) {
                    const float pdfL = m.pdf_area * (t * t) / std::fabs(cosi);
                    wB = pdf_prev / (pdf_prev + pdfL);
                }
                L = f3(L.x + beta.x * m.e.x * wB, L.y + beta.y * m.e.y * wB, L.z + beta.z * m.e.z * wB);
            }
            const F3 rho = m.a;
            if (!(depth + 1 < D && (rho.x > 0.0f || rho.y > 0.0f || rho.z > 0.0f))) break;
            const uint32_t d0 = 2 + 8 * depth;
            if (!lights.empty() && !no_nee) {
                const float ul = rnd(key, d0);
                int j = 0;
                while (j + 1 < (int)lights.size() && !(lights[j].cdf > ul)) j++;
                const Light& Lt = lights[j];
                const float u1 = rnd(key, d0 + 1), u2 = rnd(key, d0 + 2);
                const float su = std::sqrt(u1);
                const float ca = su * (1.0f - u2), cb = su * u2;
                const F3 y = f3((Lt.v0.x + Lt.e1.x * ca) + Lt.e2.x * cb, (Lt.v0.y + Lt.e1.y * ca) + Lt.e2.y * cb,
                                (Lt.v0.z + Lt.e1.z * ca) + Lt.e2.z * cb);
                const F3 dv = sub(y, P);
                const float d2 = dot(dv, dv);
                const float dist = std::sqrt(d2);
                const F3 wl = f3(dv.x / dist, dv.y / dist, dv.z / dist);
                float cosl = -dot(Lt.n, wl);
                if (!Lt.one_sided) cosl = std::fabs(cosl);
                const float cosx = dot(N, wl);
                if (cosl > 0.0f && cosx > 0.0f) {
                    const float pdfL = Lt.pdf_area * d2 / cosl;
                    float pmL = cosx * INV_PI;
                    if (alpha > 0.0f) pmL = (1.0f - alpha) * pmL + alpha * guide_pdf(lobes, wl);
                    const float w = pdfL / (pdfL + pmL);
                    const float k = cosx * w / pdfL;
                    const F3 E = f3(rho.x * INV_PI * k * Lt.Le.x, rho.y * INV_PI * k * Lt.Le.y, rho.z * INV_PI * k * Lt.Le.z);
                    if (E.x > 0.0f || E.y > 0.0f || E.z > 0.0f) {
                        rays++;
                        // the shadow ray runs from the offset origin to y, stopping 1e-4 short of it
                        const F3 o = f3(off1(P.x, N.x), off1(P.y, N.y), off1(P.z, N.z));
                        const F3 sv = sub(y, o);
                        const float sd = std::sqrt(dot(sv, sv));
                        const F3 sdir = f3(sv.x / sd, sv.y / sd, sv.z / sd);
                        if (!occluded(tk, o, sdir, sd * 0.9999f)) {
                            const F3 cc = mul(beta, E);
                            L = f3(L.x + cc.x, L.y + cc.y, L.z + cc.z);
                        }
                    }
                }
            }
            const float u3 = rnd(key, d0 + 3), u4 = rnd(key, d0 + 4), u5 = rnd(key, d0 + 5);
            F3 wd;
            if (alpha > 0.0f && u3 < alpha) {
                const float ul = rnd(key, d0 + 6);
                int kk = 7;
                float acc = 0.0f;
                for (int k = 0; k < 8; k++) {
                    acc += lobes[5 * k + 4];
                    if (ul < acc) { kk = k; break; }
                }
                const float* Lb = lobes + 5 * kk;
                const F3 mu = f3(Lb[0], Lb[1], Lb[2]);
                const float kap = Lb[3];
                float wc = 1.0f + std::log(u4 + (1.0f - u4) * std::exp(-2.0f * kap)) / kap;
                wc = std::fmin(std::fmax(wc, -1.0f), 1.0f);
                const float st = std::sqrt(std::fmax(0.0f, 1.0f - wc * wc));
                float sp, cp;
                sincos2pi(u5, sp, cp);
                F3 t1, t2;
                onb(mu, t1, t2);
                const float lx = st * cp, ly = st * sp;
                wd = f3((t1.x * lx + t2.x * ly) + mu.x * wc, (t1.y * lx + t2.y * ly) + mu.y * wc,
                        (t1.z * lx + t2.z * ly) + mu.z * wc);
            } else {
                F3 t1, t2;
                onb(N, t1, t2);
                const float r = std::sqrt(u4);
                float sp, cp;
                sincos2pi(u5, sp, cp);
                const float lx = r * cp, ly = r * sp, lz = std::sqrt(std::fmax(0.0f, 1.0f - u4));
                wd = f3((t1.x * lx + t2.x * ly) + N.x * lz, (t1.y * lx + t2.y * ly) + N.y * lz,
                        (t1.z * lx + t2.z * ly) + N.z * lz);
            }
            const float cosw = dot(N, wd);
            if (!(cosw > 0.0f)) break;
            float pm = cosw * INV_PI;
            if (alpha > 0.0f) pm = (1.0f - alpha) * pm + alpha * guide_pdf(lobes, wd);
            const float k = cosw * INV_PI / pm;
            beta = mul(beta, f3(rho.x * k, rho.y * k, rho.z * k));
            pdf_prev = pm;
            Pr = f3(off1(P.x, N.x), off1(P.y, N.y), off1(P.z, N.z));
            Dr = wd;
        }
        return L;
    }
};

}  // namespace opt

using opt::PT;

extern "C" {

// materials: n records of {albedo[3], emission[3], one_sided, pad} (vgpu_pt_material).
void* orc_pt_new(void* scene, const float* mats, int32_t nmats, const float* M, float ttf, float aspect, float focal,
                 int32_t W, int32_t H, int32_t max_depth, uint64_t seed) {
    PT* p = new PT();
    p->scene = (orc::Scene*)scene;
    for (int i = 0; i < nmats; i++) {
        const float* m = mats + 8 * (size_t)i;
        int32_t os;
        std::memcpy(&os, m + 6, 4);
        p->mats.push_back(opt::Mat{opt::f3(m[0], m[1], m[2]), opt::f3(m[3], m[4], m[5]), os != 0, 0.0f});
    }
    std::memcpy(p->M, M, sizeof(p->M));
    p->ttf = ttf; p->aspect = aspect; p->focal = focal;
    p->W = W; p->H = H; p->D = max_depth;
    p->s_lo = (uint32_t)seed; p->s_hi = (uint32_t)(seed >> 32);
    return p;
}

void orc_pt_free(void* p) { delete (PT*)p; }

// Adds the emissive triangles of one mesh (insertion order; call once per mesh, in insertion order).
void orc_pt_add_mesh_lights(void* pp, const float* verts, const uint32_t* idx, int32_t ntris, const uint8_t* shaderidx,
                            const int32_t* shaders, int32_t nshaders) {
    PT* p = (PT*)pp;
    for (int t = 0; t < ntris; t++) {
        const int si = shaderidx ? shaderidx[t] : 0;
        if (si >= nshaders) continue;
        const opt::Mat& mt = p->mats[shaders[si]];
        const float le = opt::lum(mt.e);
        if (!(le > 0.0f)) continue;
        const float* a = verts + 3 * (size_t)idx[3 * (size_t)t];
        const float* b = verts + 3 * (size_t)idx[3 * (size_t)t + 1];
        const float* c = verts + 3 * (size_t)idx[3 * (size_t)t + 2];
        const opt::F3 e1 = opt::f3(b[0] - a[0], b[1] - a[1], b[2] - a[2]);
        const opt::F3 e2 = opt::f3(c[0] - a[0], c[1] - a[1], c[2] - a[2]);
        const opt::F3 n = opt::f3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
        const float nl = std::sqrt((n.x * n.x + n.y * n.y) + n.z * n.z);
        if (!(nl > 0.0f)) continue;
        opt::Light L;
        L.v0 = opt::f3(a[0], a[1], a[2]);
        L.e1 = e1; L.e2 = e2;
        L.n = opt::f3(n.x / nl, n.y / nl, n.z / nl);
        L.Le = mt.e;
        L.one_sided = mt.one_sided;
        L.cdf = 0.0f; L.pdf_area = 0.0f;
        p->lights.push_back(L);
        p->lw.push_back((double)(0.5f * nl) * (double)le);
    }
}

// One pass over every pixel: accum[3p + k] += (double)L_k.  Returns the rays traced.
int64_t orc_pt_pass(void* pp, uint32_t pass, double* accum, float* sample_out, int32_t nthreads) {
    PT* p = (PT*)pp;
    p->finalize();
    const int np = p->W * p->H;
    int64_t rays = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(+ : rays)
    {
        orc::Task* tk = new orc::Task();
#pragma omp for schedule(dynamic, 64)
        for (int px = 0; px < np; px++) {
            const opt::F3 L = p->path(tk, px, pass, rays);
            accum[3 * (size_t)px] += (double)L.x;
            accum[3 * (size_t)px + 1] += (double)L.y;
            accum[3 * (size_t)px + 2] += (double)L.z;
            if (sample_out) { sample_out[3 * (size_t)px] = L.x; sample_out[3 * (size_t)px + 1] = L.y; sample_out[3 * (size_t)px + 2] = L.z; }
        }
        delete tk;
    }
    return rays;
}

// Traversal work since the last call: closest-hit rays, node visits, triangle tests, then the same for
// shadow rays (the counters qbvh.Trace keeps, qbvh/intersect.go:114,222); resets them.
void orc_pt_work(void* pp, int64_t* out6) {
    PT* p = (PT*)pp;
    for (int k = 0; k < 6; k++) out6[k] = p->work[k].exchange(0);
}

// Fixed-proposal guided sampling (alpha > 0): lobes = 8 x (mu xyz, kappa, weight).
void orc_pt_set_fixed_guide(void* pp, float alpha, const float* lobes) {
    PT* p = (PT*)pp;
    p->alpha = alpha;
    if (lobes) std::memcpy(p->lobes, lobes, sizeof(p->lobes));
}

// Diagnostic: 1 = no next-event estimation (emission found by sampled directions counts fully).
void orc_pt_set_no_nee(void* pp, int32_t on) { ((PT*)pp)->no_nee = on != 0; }

void orc_pt_sincos2pi(float u, float* s, float* c) { opt::sincos2pi(u, *s, *c); }

float orc_pt_rnd(uint32_t pixel, uint32_t pass, uint32_t dim, uint64_t seed) {
    const uint32_t key = opt::mix32(opt::mix32(pixel ^ opt::mix32((uint32_t)seed)) ^ (pass + (uint32_t)(seed >> 32)));
    return opt::rnd(key, dim);
}

}  // extern "C"
