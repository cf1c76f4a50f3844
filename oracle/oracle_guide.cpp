// ORACLE — test infrastructure only.
//
// FP64 restatement of the path-guiding network and its training step.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may call it,
// as the checker.
//
// The network has no counterpart in the reference: it is a new component of
// the GPU path tracer.  Its definition, which the GPU kernels implement:
//   input    hit position p, normalised to [-1,1]^3 by the scene bounds, and
//            shading normal n: x = (p, n), 6 values
//   encoding x_d, f = 0..11: sin(pi 2^f x_d), cos(pi 2^f x_d) at feature
//            index 24 d + 2 f (+1 for cos): 144 features
//   layers   144 -> 64 -> 64 -> 64 -> 64 (ReLU after each) -> 40
//   output   8 lobes of (u[3], kappa_raw, logit):
//            mu = u / |u|, kappa = exp(clamp(kappa_raw, -5, 7)), w = softmax(logit)
//   density  q(d) = sum_k w_k kappa_k / (2 pi (1 - exp(-2 kappa_k))) exp(kappa_k (mu_k . d - 1))
//   loss     L = (1/B) sum_i -(c_i / cbar) log q(d_i), cbar = (1/B) sum_i c_i,
//            i.e. sum_i c_i (-log q(d_i)) / sum_i c_i
//   update   Adam (Kingma & Ba 2015, Algorithm 1)
// Parameters are packed W1 (64x144, row = output), b1, W2, b2, W3, b3, W4, b4,
// W5 (40x64), b5: 24,360 floats.
//
// emulate = 1 rounds (to bf16, nearest even) what the GPU kernels round: the
// features, the weights and the hidden activations before each product in the
// forward pass, and the back-propagated gradients (sample i's scaled by c_i,
// not by c_i / cbar) before each product in the backward pass.  All sums are
// in double.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace og {

constexpr int NENC = 144, NH = 64, NOUT = 40, NLOBE = 8;
constexpr int OW1 = 0, OB1 = OW1 + NH * NENC, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NH, OW4 = OB3 + NH, OB4 = OW4 + NH * NH, OW5 = OB4 + NH, OB5 = OW5 + NOUT * NH,
              NPARAM = OB5 + NOUT;
static_assert(NPARAM == 24360, "parameter count");
constexpr double KMIN = -5.0, KMAX = 7.0;

static float bf16r(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&x, &u, 4);
    return x;
}

static double rnd(double x, bool em) { return em ? (double)bf16r((float)x) : x; }

struct Acts {
    double x0[NENC];
    double h[4][NH];      // post-ReLU activations as consumed by the next layer (rounded when emulating)
    double raw[NOUT];
};

static void encode(const float pn[6], bool em, double* x0) {
    for (int d = 0; d < 6; d++)
        for (int f = 0; f < 12; f++) {
            const double a = M_PI * std::ldexp((double)pn[d], f);
            x0[24 * d + 2 * f] = rnd(std::sin(a), em);
            x0[24 * d + 2 * f + 1] = rnd(std::cos(a), em);
        }
}

static void normalise_input(const float* pos, const float* nrm, const float* bounds, float pn[6]) {
    for (int k = 0; k < 3; k++) {
        const float lo = bounds[k], hi = bounds[3 + k];
        const float c = 0.5f * (lo + hi), h = 0.5f * (hi - lo);
        pn[k] = (pos[k] - c) / h;
        pn[3 + k] = nrm[k];
    }
}

static void forward(const float* P, const float pn[6], bool em, Acts& a) {
    encode(pn, em, a.x0);
    const double* in = a.x0;
    int nin = NENC;
    const int ow[4] = {OW1, OW2, OW3, OW4}, ob[4] = {OB1, OB2, OB3, OB4};
    for (int l = 0; l < 4; l++) {
        for (int o = 0; o < NH; o++) {
            double s = 0.0;
            for (int i = 0; i < nin; i++) s += rnd(P[ow[l] + o * nin + i], em) * in[i];
            s += (double)P[ob[l] + o];
            a.h[l][o] = rnd(s > 0.0 ? s : 0.0, em);
        }
        in = a.h[l];
        nin = NH;
    }
    for (int o = 0; o < NOUT; o++) {
        double s = 0.0;
        for (int i = 0; i < NH; i++) s += rnd(P[OW5 + o * NH + i], em) * in[i];
        a.raw[o] = s + (double)P[OB5 + o];
    }
}

struct Lobes {
    double mu[NLOBE][3], kappa[NLOBE], w[NLOBE], ulen[NLOBE];
    bool kin[NLOBE];   // kappa_raw strictly inside the clamp
};

static void decode(const double* raw, Lobes& L) {
    double mx = -1e300;
    for (int k = 0; k < NLOBE; k++) mx = std::fmax(mx, raw[5 * k + 4]);
    double z = 0.0;
    for (int k = 0; k < NLOBE; k++) {
        const double* u = raw + 5 * k;
        const double len = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        L.ulen[k] = len;
        if (len > 0.0) {
            L.mu[k][0] = u[0] / len; L.mu[k][1] = u[1] / len; L.mu[k][2] = u[2] / len;
        } else {
            L.mu[k][0] = 0.0; L.mu[k][1] = 0.0; L.mu[k][2] = 1.0;
        }
        const double kr = u[3];
        L.kin[k] = kr > KMIN && kr < KMAX;
        L.kappa[k] = std::exp(std::fmin(std::fmax(kr, KMIN), KMAX));
        L.w[k] = std::exp(u[4] - mx);
        z += L.w[k];
    }
    for (int k = 0; k < NLOBE; k++) L.w[k] /= z;
}

static double vmf_norm(double kappa) { return kappa / (2.0 * M_PI * -std::expm1(-2.0 * kappa)); }

static double lobe_pdf(const Lobes& L, int k, const double d[3]) {
    const double c = L.mu[k][0] * d[0] + L.mu[k][1] * d[1] + L.mu[k][2] * d[2];
    return vmf_norm(L.kappa[k]) * std::exp(L.kappa[k] * (c - 1.0));
}

static double mix_pdf(const Lobes& L, const double d[3]) {
    double q = 0.0;
    for (int k = 0; k < NLOBE; k++) q += L.w[k] * lobe_pdf(L, k, d);
    return q;
}

// d(-log q)/d raw, scaled by s.
static void nll_grad(const double* raw, const double d[3], double s, double* g) {
    Lobes L;
    decode(raw, L);
    double p[NLOBE], q = 0.0;
    for (int k = 0; k < NLOBE; k++) { p[k] = lobe_pdf(L, k, d); q += L.w[k] * p[k]; }
    for (int k = 0; k < NLOBE; k++) {
        const double r = L.w[k] * p[k] / q;
        const double kap = L.kappa[k];
        const double c = L.mu[k][0] * d[0] + L.mu[k][1] * d[1] + L.mu[k][2] * d[2];
        g[5 * k + 4] = s * (L.w[k] - r);
        const double dlogp = (1.0 / kap - 2.0 / std::expm1(2.0 * kap)) + (c - 1.0);
        g[5 * k + 3] = L.kin[k] ? -s * r * dlogp * kap : 0.0;
        for (int j = 0; j < 3; j++)
            g[5 * k + j] = L.ulen[k] > 0.0 ? -s * r * kap * (d[j] - c * L.mu[k][j]) / L.ulen[k] : 0.0;
    }
}

// Accumulates dL_i/dtheta for one sample whose output gradient is gout (already scaled).
static void backward(const float* P, const Acts& a, const double* gout, bool em, double* G) {
    double dy[NH], dx[NH];
    double go[NOUT];
    for (int o = 0; o < NOUT; o++) go[o] = rnd(gout[o], em);
    for (int o = 0; o < NOUT; o++) {
        G[OB5 + o] += go[o];
        for (int i = 0; i < NH; i++) G[OW5 + o * NH + i] += go[o] * a.h[3][i];
    }
    for (int i = 0; i < NH; i++) {
        double s = 0.0;
        for (int o = 0; o < NOUT; o++) s += rnd(P[OW5 + o * NH + i], em) * go[o];
        dx[i] = s;
    }
    const int ow[4] = {OW1, OW2, OW3, OW4}, ob[4] = {OB1, OB2, OB3, OB4};
    for (int l = 3; l >= 0; l--) {
        for (int o = 0; o < NH; o++) dy[o] = rnd(a.h[l][o] > 0.0 ? dx[o] : 0.0, em);
        const int nin = l == 0 ? NENC : NH;
        const double* in = l == 0 ? a.x0 : a.h[l - 1];
        for (int o = 0; o < NH; o++) {
            G[ob[l] + o] += dy[o];
            for (int i = 0; i < nin; i++) G[ow[l] + o * nin + i] += dy[o] * in[i];
        }
        if (l > 0)
            for (int i = 0; i < NH; i++) {
                double s = 0.0;
                for (int o = 0; o < NH; o++) s += rnd(P[ow[l] + o * NH + i], em) * dy[o];
                dx[i] = s;
            }
    }
}

}  // namespace og

extern "C" {

int ogd_nparams() { return og::NPARAM; }

// Forward pass: raw outputs (n x 40) and decoded lobes (n x 8 x (mu[3], kappa, w)).
void ogd_forward(const float* params, int64_t n, const float* pos, const float* nrm, const float* bounds, int emulate,
                 float* raw_out, float* lobes_out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        float pn[6];
        og::normalise_input(pos + 3 * i, nrm + 3 * i, bounds, pn);
        og::Acts a;
        og::forward(params, pn, emulate != 0, a);
        if (raw_out)
            for (int o = 0; o < og::NOUT; o++) raw_out[i * og::NOUT + o] = (float)a.raw[o];
        if (lobes_out) {
            og::Lobes L;
            og::decode(a.raw, L);
            for (int k = 0; k < og::NLOBE; k++) {
                float* q = lobes_out + (i * og::NLOBE + k) * 5;
                q[0] = (float)L.mu[k][0]; q[1] = (float)L.mu[k][1]; q[2] = (float)L.mu[k][2];
                q[3] = (float)L.kappa[k]; q[4] = (float)L.w[k];
            }
        }
    }
}

// Mixture density of directions dirs (n x 3) under raw outputs (n x 40).
void ogd_mix_pdf(const float* raw, int64_t n, const float* dirs, double* pdf_out) {
    for (int64_t i = 0; i < n; i++) {
        double r[og::NOUT];
        for (int o = 0; o < og::NOUT; o++) r[o] = raw[i * og::NOUT + o];
        og::Lobes L;
        og::decode(r, L);
        const double d[3] = {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
        pdf_out[i] = og::mix_pdf(L, d);
    }
}

// d(-log q(d_i))/d raw_i for given raw outputs (n x 40), in double.
void ogd_nll_grad_raw(const double* raw, int64_t n, const float* dirs, double* g_out) {
    for (int64_t i = 0; i < n; i++) {
        const double d[3] = {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
        og::nll_grad(raw + i * og::NOUT, d, 1.0, g_out + i * og::NOUT);
    }
}

// -log q(d) in double for given raw outputs.
void ogd_nll(const double* raw, int64_t n, const float* dirs, double* out) {
    for (int64_t i = 0; i < n; i++) {
        og::Lobes L;
        og::decode(raw + i * og::NOUT, L);
        const double d[3] = {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
        out[i] = -std::log(og::mix_pdf(L, d));
    }
}

// Batch loss sum_i c_i (-log q_i) / sum_i c_i and its gradient.  Each sample's output gradient is
// weighted by its own c_i (the GPU rounds it to bf16 at that scale) and the sums are divided by
// sum c_i at the end.  Returns 0, or 1 when sum c = 0 (no step).
int ogd_loss_grad(const float* params, int64_t n, const float* pos, const float* nrm, const float* dirs,
                  const float* c, const float* bounds, int emulate, double* loss_out, double* grad_out) {
    double csum = 0.0;
    for (int64_t i = 0; i < n; i++) csum += c[i];
    for (int p = 0; p < og::NPARAM; p++) grad_out[p] = 0.0;
    *loss_out = 0.0;
    if (!(csum > 0.0)) return 1;
    const bool em = emulate != 0;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<std::vector<double>> Gt(nt, std::vector<double>(og::NPARAM, 0.0));
    std::vector<double> lt(nt, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        if (!(c[i] > 0.0f)) continue;
        float pn[6];
        og::normalise_input(pos + 3 * i, nrm + 3 * i, bounds, pn);
        og::Acts a;
        og::forward(params, pn, em, a);
        const double d[3] = {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
        const double s = (double)c[i];
        og::Lobes L;
        og::decode(a.raw, L);
        lt[t] += -s * std::log(og::mix_pdf(L, d));
        double g[og::NOUT];
        og::nll_grad(a.raw, d, s, g);
        og::backward(params, a, g, em, Gt[t].data());
    }
    for (int t = 0; t < nt; t++) {
        *loss_out += lt[t];
        for (int p = 0; p < og::NPARAM; p++) grad_out[p] += Gt[t][p];
    }
    *loss_out /= csum;
    for (int p = 0; p < og::NPARAM; p++) grad_out[p] /= csum;
    return 0;
}

// One Adam step (step t >= 1) in double on float state.
void ogd_adam(float* params, float* m, float* v, const double* grad, int n, int t, double lr, double b1, double b2,
              double eps) {
    const double c1 = 1.0 - std::pow(b1, t), c2 = 1.0 - std::pow(b2, t);
    for (int p = 0; p < n; p++) {
        const double g = grad[p];
        const double mm = b1 * m[p] + (1.0 - b1) * g;
        const double vv = b2 * v[p] + (1.0 - b2) * g * g;
        m[p] = (float)mm;
        v[p] = (float)vv;
        params[p] = (float)((double)params[p] - lr * (mm / c1) / (std::sqrt(vv / c2) + eps));
    }
}

float ogd_bf16(float x) { return og::bf16r(x); }

}  // extern "C"
