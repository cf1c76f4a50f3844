// Path-mode kernels (wavefront): one pass traces one path per pixel.
//
//   k_pt_camera    camera rays for every pixel (queue 0), clears the pass radiance
//   (trace)        closest hit over the current queue (device-resident length)
//   k_pt_hit       hit point, flat shading normal, material; emission with the
//                  MIS weight of the sample that found it; writes the guide query
//   (guide)        vMF mixture per live vertex (guide.hip), guided mode only
//   k_pt_scatter   next-event estimation (one shadow ray per vertex), one-sample
//                  MIS choice between the cosine lobe and the guide, throughput
//                  update, compaction of live paths into the next queue
//   (trace)        shadow any-hit over the compacted shadow queue
//   k_pt_shadow    adds unoccluded NEE contributions
//   k_pt_accumulate  adds the pass radiance to the FP64 accumulation
//   k_pt_gather     training samples for the guide (guided mode); each training step draws
//                   its batch from them inside the guide's kernels
//
// Every float operation is written out in a fixed order and the file is
// compiled with -ffp-contract=off, so that the CPU restatement
// (oracle/oracle_pt.cpp) reproduces plain-mode paths exactly.  Sines and
// cosines of 2 pi u use the polynomial below instead of a library call.

#include "pathtrace_kernels.h"

namespace vgpu {

namespace {

struct F3 { float x, y, z; };
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 f3(const float4& a) { return F3{a.x, a.y, a.z}; }
__device__ __forceinline__ F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 mul(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ F3 mul(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ F3 neg(F3 a) { return f3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float lum(F3 c) { return (0.2126f * c.x + 0.7152f * c.y) + 0.0722f * c.z; }
__device__ __forceinline__ float4 f4(F3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

constexpr float INV_PI = 0.318309886f;

__device__ __forceinline__ uint32_t mix32(uint32_t x) {   // lowbias32 (C. Wellons)
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
// The seed is hashed before it meets the pixel index: XOR with the raw seed would only permute the
// pixels, so two small seeds would share the same streams at different pixels.
__device__ __forceinline__ uint32_t path_key(const PTView& v, uint32_t pixel, uint32_t pass) {
    return mix32(mix32(pixel ^ mix32(v.seed_lo)) ^ (pass + v.seed_hi));
}
__device__ __forceinline__ float rnd(uint32_t key, uint32_t dim) {
    return (float)(mix32(key + dim * 0x9e3779b9u) >> 8) * 5.9604645e-08f;
}
__device__ __forceinline__ bool recorded(const PTView& v, uint32_t pixel, uint32_t pass) {
    return v.rec_stride != 0 && mix32(pixel ^ 0x5bd1e995u) % v.rec_stride == pass % v.rec_stride;
}

// sin and cos of 2 pi u, u in [0, 1): quadrant, then the nearer of the octant's two angles,
// then Taylor polynomials to degree 11 / 12 on [0, pi/4].
__device__ __forceinline__ void sincos2pi(float u, float& s, float& c) {
    const float t = u * 4.0f;
    const int q = (int)t;
    const float f = t - (float)q;
    const bool sw = f > 0.5f;
    const float a = (sw ? 1.0f - f : f) * 1.57079637f;
    const float a2 = a * a;
    float ps = -2.50521084e-08f;
    ps = ps * a2 + 2.75573192e-06f;
    ps = ps * a2 + -1.98412698e-04f;
    ps = ps * a2 + 8.33333377e-03f;
    ps = ps * a2 + -1.66666672e-01f;
    ps = ps * a2 + 1.0f;
    const float sa = ps * a;
    float pc = 2.08767570e-09f;
    pc = pc * a2 + -2.75573192e-07f;
    pc = pc * a2 + 2.48015876e-05f;
    pc = pc * a2 + -1.38888892e-03f;
    pc = pc * a2 + 4.16666679e-02f;
    pc = pc * a2 + -0.5f;
    pc = pc * a2 + 1.0f;
    const float sq = sw ? pc : sa, cq = sw ? sa : pc;
    if (q == 0) { s = sq; c = cq; }
    else if (q == 1) { s = cq; c = -sq; }
    else if (q == 2) { s = -sq; c = -cq; }
    else { s = -cq; c = sq; }
}

// Orthonormal basis around a unit vector (Duff et al. 2017).
__device__ __forceinline__ void onb(F3 n, F3& t1, F3& t2) {
    const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    t1 = f3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    t2 = f3(b, sign + n.y * n.y * a, -n.y);
}

__device__ __forceinline__ F3 expand(F3 t1, F3 t2, F3 n, float lx, float ly, float lz) {
    return f3((t1.x * lx + t2.x * ly) + n.x * lz, (t1.y * lx + t2.y * ly) + n.y * lz, (t1.z * lx + t2.z * ly) + n.z * lz);
}

// Ray origin pushed off the surface along n (Waechter & Binder, Ray Tracing Gems 2019, ch. 6).
__device__ __forceinline__ float off1(float p, float n) {
    const int of = (int)(256.0f * n);
    const float pi = __int_as_float(__float_as_int(p) + (p < 0.0f ? -of : of));
    return fabsf(p) < 0.03125f ? p + 1.52587891e-05f * n : pi;
}
__device__ __forceinline__ F3 offset(F3 p, F3 n) { return f3(off1(p.x, n.x), off1(p.y, n.y), off1(p.z, n.z)); }

// ------------------------------------------------------------------ guide mixture

__device__ __forceinline__ float vmf_c(float k) { return k / (6.28318548f * -expm1f(-2.0f * k)); }

__device__ float guide_pdf(const float* lob, F3 w) {
    float q = 0.0f;
    for (int k = 0; k < guide::NLOBE; k++) {
        const float* L = lob + 5 * k;
        const float c = (L[0] * w.x + L[1] * w.y) + L[2] * w.z;
        q += L[4] * (vmf_c(L[3]) * expf(L[3] * (c - 1.0f)));
    }
    return q;
}

// Picks a lobe by its weight and samples it (Jakob 2012, "Numerically stable sampling of the
// von Mises-Fisher distribution on S^2").
__device__ F3 guide_sample(const float* lob, float ul, float u1, float u2) {
    int k = guide::NLOBE - 1;
    float acc = 0.0f;
    for (int kk = 0; kk < guide::NLOBE; kk++) {
        acc += lob[5 * kk + 4];
        if (ul < acc) { k = kk; break; }
    }
    const float* L = lob + 5 * k;
    const F3 mu = f3(L[0], L[1], L[2]);
    const float kap = L[3];
    float wc = 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * kap)) / kap;
    wc = fminf(fmaxf(wc, -1.0f), 1.0f);
    const float st = sqrtf(fmaxf(0.0f, 1.0f - wc * wc));
    float sp, cp;
    sincos2pi(u2, sp, cp);
    F3 t1, t2;
    onb(mu, t1, t2);
    return expand(t1, t2, mu, st * cp, st * sp, wc);
}

// ------------------------------------------------------------------ kernels

__global__ __launch_bounds__(256) void k_pt_camera(PTView v, uint32_t pass, PTBuffers b) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int np = v.width * v.height;
    if (p == 0) { b.count[0] = (uint32_t)np; b.count[3] = 0; }
    if (p >= np) return;
    const uint32_t key = path_key(v, (uint32_t)p, pass);
    const float fx = (float)(p % v.width) + rnd(key, 0);
    const float fy = (float)(p / v.width) + rnd(key, 1);
    const float Sx = 2.0f * fx / (float)v.width - 1.0f;
    const float Sy = 1.0f - 2.0f * fy / (float)v.height;
    const float camu = Sx * v.tan_theta_focal;
    const float camv = Sy * (v.tan_theta_focal / v.aspect);
    const float sz = -v.focal;
    const float* M = v.M;
    const F3 d = f3((M[0] * camu + M[4] * camv) + M[8] * sz, (M[1] * camu + M[5] * camv) + M[9] * sz,
                    (M[2] * camu + M[6] * camv) + M[10] * sz);
    const float l = sqrtf(dot(d, d));
    const F3 D = f3(d.x / l, d.y / l, d.z / l);
    b.qray[0][2 * p] = make_float4(M[12], M[13], M[14], __builtin_inff());
    b.qray[0][2 * p + 1] = f4(D, 0.0f);
    b.qstate[0][p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    b.qmeta[0][p] = make_int4(p, 0, 0, 0);
    b.lpass[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (b.plen) b.plen[p] = recorded(v, (uint32_t)p, pass) ? 0 : -1;
}

__global__ __launch_bounds__(256) void k_pt_hit(PTView v, uint32_t pass, int cur, PTBuffers b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = b.count[cur];
    if (i == 0) { b.count[1 - cur] = 0; b.count[2] = 0; atomicAdd(b.rays, (unsigned long long)n); }
    if ((uint32_t)i >= n) return;
    const int4 meta = b.qmeta[cur][i];
    const int pixel = meta.x, depth = meta.y;
    const float4 st = b.qstate[cur][i];
    const F3 beta = f3(st);
    const float pdf_prev = st.w;
    const int4 h0 = b.hits[2 * i], h1 = b.hits[2 * i + 1];
    float4* sh = b.shade + 3 * (size_t)i;
    const bool rec = b.plen && b.plen[pixel] >= 0;
    float4* R = rec ? b.rec + ((size_t)pixel * v.max_depth + depth) * PT_REC_F4 : nullptr;
    if (h0.x < 0) {   // miss: the room has no environment light
        sh[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const int prim = h0.x, g = h0.y;
    const float t = __int_as_float(h0.z), U = __int_as_float(h0.w), V = __int_as_float(h1.x), W = __int_as_float(h1.y);
    const DevGeom geo = v.sv.geoms[g];
    const float4* tp = v.sv.tris + ((size_t)geo.tri_base + prim) * 3;
    const F3 a = f3(tp[0]), bb = f3(tp[1]), c = f3(tp[2]);
    const F3 P = f3((U * a.x + V * bb.x) + W * c.x, (U * a.y + V * bb.y) + W * c.y, (U * a.z + V * bb.z) + W * c.z);
    const F3 ng = cross(sub(bb, a), sub(c, a));
    const float nl = sqrtf(dot(ng, ng));
    const F3 Ng = f3(ng.x / nl, ng.y / nl, ng.z / nl);
    const F3 D = f3(b.qray[cur][2 * i + 1]);
    const float cosi = dot(Ng, D);
    const bool front = cosi < 0.0f;
    const F3 N = front ? Ng : neg(Ng);
    const DevGeomShade gs = v.gshade[g];
    const int sidx = gs.sidx_base >= 0 ? (int)v.shaderidx[gs.sidx_base + prim] : 0;
    const PTMat m = v.mats[v.shaders[gs.shader_base + sidx]];
    const F3 Le0 = f3(m.e);
    const bool emits = (Le0.x > 0.0f || Le0.y > 0.0f || Le0.z > 0.0f) && (m.a.w == 0.0f || front);
    F3 Le = f3(0.0f, 0.0f, 0.0f);
    float wB = 1.0f;
    if (emits) {
        Le = Le0;
        if (depth > 0) {
            const float pdfL = m.e.w * (t * t) / fabsf(cosi);
            wB = pdf_prev / (pdf_prev + pdfL);
        }
        const float4 lp = b.lpass[pixel];
        b.lpass[pixel] = make_float4(lp.x + beta.x * Le.x * wB, lp.y + beta.y * Le.y * wB, lp.z + beta.z * Le.z * wB, 0.0f);
    }
    if (R) {
        R[3] = f4(Le, wB);
        R[4] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        R[5] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        b.plen[pixel] = depth + 1;
    }
    const F3 rho = f3(m.a);
    const bool cont = depth + 1 < v.max_depth && (rho.x > 0.0f || rho.y > 0.0f || rho.z > 0.0f);
    sh[0] = f4(P, cont ? 1.0f : 0.0f);
    sh[1] = f4(N, 0.0f);
    sh[2] = f4(rho, 0.0f);
    b.query[2 * i] = f4(P, 0.0f);
    b.query[2 * i + 1] = f4(N, 0.0f);
}


// Workgroup size of the kernels that append to the path queues (k_pt_scatter, k_pt_gather).
constexpr int kAppendBlock = 1024;
constexpr int kAppendWaves = kAppendBlock / 64;

// Appends n entries per lane for the whole workgroup (kAppendBlock threads) with one atomic on
// `counter` (same-address device atomics are slow when every wavefront issues its own); returns
// this lane's first slot.  Lanes take consecutive slots in thread order.  Every thread of the
// workgroup calls it; s_tot holds kAppendWaves + 1 entries.
__device__ __forceinline__ uint32_t block_append(uint32_t* counter, uint32_t n, uint32_t* s_tot) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = n;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_tot[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < kAppendWaves; w++) {
            const uint32_t t = s_tot[w];
            s_tot[w] = run;
            run += t;
        }
        s_tot[kAppendWaves] = run ? atomicAdd(counter, run) : 0u;
    }
    __syncthreads();
    const uint32_t slot = s_tot[kAppendWaves] + s_tot[wave] + (x - n);
    __syncthreads();
    return slot;
}

__global__ __launch_bounds__(kAppendBlock) void k_pt_scatter(PTView v, uint32_t pass, int cur, PTBuffers b) {
    __shared__ uint32_t s_tot[kAppendWaves + 1];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = b.count[cur];
    bool alive = false, shadow = false;
    F3 P = f3(0, 0, 0), N = f3(0, 0, 0), rho = f3(0, 0, 0), beta = f3(0, 0, 0), nbeta = f3(0, 0, 0), wdir = f3(0, 0, 0);
    F3 sorg = f3(0, 0, 0), sdir = f3(0, 0, 0), E = f3(0, 0, 0);
    float pm = 0.0f, stmax = 0.0f;
    int pixel = 0, depth = 0;
    float4* R = nullptr;
    if ((uint32_t)i < n) {
        const float4 s0 = b.shade[3 * (size_t)i];
        if (s0.w != 0.0f) {
            P = f3(s0);
            N = f3(b.shade[3 * (size_t)i + 1]);
            rho = f3(b.shade[3 * (size_t)i + 2]);
            const int4 meta = b.qmeta[cur][i];
            pixel = meta.x; depth = meta.y;
            beta = f3(b.qstate[cur][i]);
            const uint32_t key = path_key(v, (uint32_t)pixel, pass);
            const uint32_t d0 = 2 + PT_DIMS_PER_DEPTH * depth;
            const float* lob = b.lobes + (size_t)i * guide::NOUT;
            const bool guided = v.alpha > 0.0f;
            if (b.plen && b.plen[pixel] >= 0) R = b.rec + ((size_t)pixel * v.max_depth + depth) * PT_REC_F4;
            // next-event estimation
            if (v.nlights > 0) {
                const float ul = rnd(key, d0 + 0);
                int lo = 0, hi = v.nlights - 1;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (v.lights[mid].e1.w > ul) hi = mid; else lo = mid + 1;
                }
                const PTLight L = v.lights[lo];
                const float u1 = rnd(key, d0 + 1), u2 = rnd(key, d0 + 2);
                const float su = sqrtf(u1);
                const float ca = su * (1.0f - u2), cb = su * u2;
                const F3 y = f3((L.v0.x + L.e1.x * ca) + L.e2.x * cb, (L.v0.y + L.e1.y * ca) + L.e2.y * cb,
                                (L.v0.z + L.e1.z * ca) + L.e2.z * cb);
                const F3 dv = sub(y, P);
                const float d2 = dot(dv, dv);
                const float dist = sqrtf(d2);
                const F3 wl = f3(dv.x / dist, dv.y / dist, dv.z / dist);
                float cosl = -dot(f3(L.n), wl);
                if (L.e2.w == 0.0f) cosl = fabsf(cosl);
                const float cosx = dot(N, wl);
                if (cosl > 0.0f && cosx > 0.0f) {
                    const float pdfL = L.Le.w * d2 / cosl;
                    float pmL = cosx * INV_PI;
                    if (guided) pmL = (1.0f - v.alpha) * pmL + v.alpha * guide_pdf(lob, wl);
                    const float w = pdfL / (pdfL + pmL);
                    const float k = cosx * w / pdfL;
                    E = f3(rho.x * INV_PI * k * L.Le.x, rho.y * INV_PI * k * L.Le.y, rho.z * INV_PI * k * L.Le.z);
                    if (E.x > 0.0f || E.y > 0.0f || E.z > 0.0f) {
                        // aimed from the offset origin: with the shading point's distance the light's own
                        // plane can come before tmax when the offset is large next to the distance
                        shadow = true;
                        sorg = offset(P, N);
                        const F3 sv = sub(y, sorg);
                        const float sd = sqrtf(dot(sv, sv));
                        sdir = f3(sv.x / sd, sv.y / sd, sv.z / sd);
                        stmax = sd * 0.9999f;
                    }
                }
            }
            // continuation: one-sample MIS between the cosine lobe and the guide
            const float u3 = rnd(key, d0 + 3), u4 = rnd(key, d0 + 4), u5 = rnd(key, d0 + 5);
            F3 t1, t2;
            if (guided && u3 < v.alpha) {
                wdir = guide_sample(lob, rnd(key, d0 + 6), u4, u5);
            } else {
                onb(N, t1, t2);
                const float r = sqrtf(u4);
                float sp, cp;
                sincos2pi(u5, sp, cp);
                wdir = expand(t1, t2, N, r * cp, r * sp, sqrtf(fmaxf(0.0f, 1.0f - u4)));
            }
            const float cosw = dot(N, wdir);
            if (cosw > 0.0f) {
                pm = cosw * INV_PI;
                if (guided) pm = (1.0f - v.alpha) * pm + v.alpha * guide_pdf(lob, wdir);
                const float k = cosw * INV_PI / pm;
                const F3 w = f3(rho.x * k, rho.y * k, rho.z * k);
                nbeta = mul(beta, w);
                alive = true;
                if (R) {
                    R[0] = f4(P, pm);
                    R[1] = f4(N, cosw);
                    R[2] = f4(wdir, 0.0f);
                    R[5] = f4(w, 1.0f);
                }
            }
        }
    }
    const uint32_t si = block_append(b.count + 2, shadow ? 1u : 0u, s_tot);
    if (shadow) {
        b.sray[2 * si] = f4(sorg, stmax);
        b.sray[2 * si + 1] = f4(sdir, 0.0f);
        const F3 c = mul(beta, E);
        b.scontrib[2 * si] = f4(c, __int_as_float(pixel));
        b.scontrib[2 * si + 1] = f4(E, __int_as_float(R ? pixel * v.max_depth + depth : -1));
    }
    const int nxt = 1 - cur;
    const uint32_t qi = block_append(b.count + nxt, alive ? 1u : 0u, s_tot);
    if (alive) {
        b.qray[nxt][2 * qi] = f4(offset(P, N), __builtin_inff());
        b.qray[nxt][2 * qi + 1] = f4(wdir, 0.0f);
        b.qstate[nxt][qi] = f4(nbeta, pm);
        b.qmeta[nxt][qi] = make_int4(pixel, depth + 1, 0, 0);
    }
}

__global__ __launch_bounds__(256) void k_pt_shadow(PTView v, PTBuffers b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = b.count[2];
    if (i == 0) atomicAdd(b.rays, (unsigned long long)n);
    if ((uint32_t)i >= n) return;
    if (b.shits[i].x >= 0) return;   // occluded
    const float4 c = b.scontrib[2 * i], e = b.scontrib[2 * i + 1];
    const int pixel = __float_as_int(c.w);
    const float4 lp = b.lpass[pixel];
    b.lpass[pixel] = make_float4(lp.x + c.x, lp.y + c.y, lp.z + c.z, 0.0f);
    const int r = __float_as_int(e.w);
    if (r >= 0) b.rec[(size_t)r * PT_REC_F4 + 4] = make_float4(e.x, e.y, e.z, 0.0f);
}

__global__ __launch_bounds__(256) void k_pt_accumulate(PTView v, PTBuffers b, double* __restrict__ acc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= v.width * v.height) return;
    const float4 l = b.lpass[p];
    acc[3 * (size_t)p] += (double)l.x;
    acc[3 * (size_t)p + 1] += (double)l.y;
    acc[3 * (size_t)p + 2] += (double)l.z;
}

// Training samples from the recorded paths.  With vertices j = 0..n-1 of one path, Le_j the
// emission found at j (unweighted), wB_j its MIS weight, E_j the NEE estimate at j and w_j the
// throughput factor of the direction sampled at j:
//   R_j = E_j + w_j (wB_{j+1} Le_{j+1} + R_{j+1})        (radiance leaving j, emission excluded)
//   L_in(j) = Le_{j+1} + R_{j+1}                         (radiance arriving at j along its sample)
// Vertex j yields the sample (P_j, N_j, dir_j) with c = lum(L_in(j)) cos_j / p_j when c > 0.
//
// gather_step: one step of the backward walk over record j (next: j + 1 < n, a later vertex
// exists); true when vertex j yields a sample, written to s.
__device__ __forceinline__ bool gather_step(const float4* __restrict__ r, bool next, F3& Rn, F3& Len, float& wBn,
                                            guide::TrainSample& s) {
    bool emit = false;
    const float4 r5 = r[5];
    if (next && r5.w != 0.0f) {
        const F3 Lin = add(Len, Rn);
        const float4 r0 = r[0], r1 = r[1], r2 = r[2];
        const float c = lum(Lin) * r1.w / r0.w;
        if (c > 0.0f && c < 3.0e38f) {
            emit = true;
            s.px = r0.x; s.py = r0.y; s.pz = r0.z; s.c = c;
            s.nx = r1.x; s.ny = r1.y; s.nz = r1.z;
            s.dx = r2.x; s.dy = r2.y; s.dz = r2.z;
        }
    }
    const F3 inner = next ? add(mul(Len, wBn), Rn) : f3(0, 0, 0);
    const float4 r3 = r[3], r4 = r[4];
    Rn = add(f3(r4), mul(f3(r5), inner));
    Len = f3(r3);
    wBn = r3.w;
    return emit;
}

// Walks each recorded path backwards twice: once to count its training samples, then, after one
// append for the whole workgroup, to write them to consecutive slots.
__global__ __launch_bounds__(kAppendBlock) void k_pt_gather(PTView v, uint32_t pass, PTBuffers b) {
    __shared__ uint32_t s_tot[kAppendWaves + 1];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = p < v.width * v.height ? b.plen[p] : 0;
    const float4* R = b.rec + (size_t)(p < v.width * v.height ? p : 0) * v.max_depth * PT_REC_F4;
    uint32_t cnt = 0;
    {
        F3 Rn = f3(0, 0, 0), Len = f3(0, 0, 0);
        float wBn = 0.0f;
        for (int j = n - 1; j >= 0; j--) {
            guide::TrainSample s{};
            if (gather_step(R + (size_t)j * PT_REC_F4, j + 1 < n, Rn, Len, wBn, s)) cnt++;
        }
    }
    uint32_t k = block_append(b.count + 3, cnt, s_tot);
    if (cnt == 0) return;
    F3 Rn = f3(0, 0, 0), Len = f3(0, 0, 0);
    float wBn = 0.0f;
    for (int j = n - 1; j >= 0; j--) {
        guide::TrainSample s{};
        if (gather_step(R + (size_t)j * PT_REC_F4, j + 1 < n, Rn, Len, wBn, s)) {
            if (k < b.tcap) b.tsamples[k] = s;
            k++;
        }
    }
}


}  // namespace

static unsigned grid_for(uint64_t n) { return (unsigned)((n + 255) / 256); }

hipError_t pt_launch_camera(const PTView& v, uint32_t pass, const PTBuffers& b, hipStream_t st) {
    hipLaunchKernelGGL(k_pt_camera, dim3(grid_for((uint64_t)v.width * v.height)), dim3(256), 0, st, v, pass, b);
    return hipGetLastError();
}
hipError_t pt_launch_hit(const PTView& v, uint32_t pass, int cur, const PTBuffers& b, hipStream_t st) {
    hipLaunchKernelGGL(k_pt_hit, dim3(grid_for((uint64_t)v.width * v.height)), dim3(256), 0, st, v, pass, cur, b);
    return hipGetLastError();
}
hipError_t pt_launch_scatter(const PTView& v, uint32_t pass, int cur, const PTBuffers& b, hipStream_t st) {
    const uint64_t np = (uint64_t)v.width * v.height;
    hipLaunchKernelGGL(k_pt_scatter, dim3((unsigned)((np + kAppendBlock - 1) / kAppendBlock)), dim3(kAppendBlock), 0, st, v,
                       pass, cur, b);
    return hipGetLastError();
}
hipError_t pt_launch_shadow(const PTView& v, const PTBuffers& b, hipStream_t st) {
    hipLaunchKernelGGL(k_pt_shadow, dim3(grid_for((uint64_t)v.width * v.height)), dim3(256), 0, st, v, b);
    return hipGetLastError();
}
hipError_t pt_launch_accumulate(const PTView& v, const PTBuffers& b, double* acc, hipStream_t st) {
    hipLaunchKernelGGL(k_pt_accumulate, dim3(grid_for((uint64_t)v.width * v.height)), dim3(256), 0, st, v, b, acc);
    return hipGetLastError();
}
hipError_t pt_launch_gather(const PTView& v, uint32_t pass, const PTBuffers& b, hipStream_t st) {
    const uint64_t np = (uint64_t)v.width * v.height;
    hipLaunchKernelGGL(k_pt_gather, dim3((unsigned)((np + kAppendBlock - 1) / kAppendBlock)), dim3(kAppendBlock), 0, st, v,
                       pass, b);
    return hipGetLastError();
}

}  // namespace vgpu
