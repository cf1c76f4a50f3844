"""Pins the guiding network's FP64 oracle (oracle/oracle_guide.cpp).

The network has no reference counterpart; its definition is written at the top
of oracle_guide.cpp.  What pins the restatement:
- the vMF lobe density integrates to 1 over the sphere (quadrature), so the
  normalisation constant is right;
- the analytic gradients (w.r.t. the raw outputs and w.r.t. every parameter
  block) agree with central finite differences of the loss;
- a hand-built network that routes one encoded feature to one output gives
  relu(sin/cos(pi 2^f x)) for the documented feature index;
- one and two Adam steps from a zero state agree with the closed form;
- bf16 rounding is round-to-nearest-even on hand-picked values.
"""
import math

import numpy as np
import pytest

import oracle

NP = 24360
OW1, OB1, OW2, OB2, OW3, OB3, OW4, OB4, OW5, OB5 = 0, 9216, 9280, 13376, 13440, 17536, 17600, 21696, 21760, 24320
BOUNDS = np.array([-4, 0, -4, 4, 4, 4], np.float32)


def _rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP, np.float32)
    blocks = [(OW1, 64 * 144, 144), (OW2, 4096, 64), (OW3, 4096, 64), (OW4, 4096, 64), (OW5, 2560, 64)]
    for off, n, fan in blocks:
        p[off:off + n] = rng.uniform(-1, 1, n) * math.sqrt(6.0 / fan) * scale
    for off, n in [(OB1, 64), (OB2, 64), (OB3, 64), (OB4, 64), (OB5, 40)]:
        p[off:off + n] = rng.uniform(-0.1, 0.1, n)
    return p


def _rand_inputs(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(BOUNDS[:3], BOUNDS[3:], (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = rng.uniform(0, 2, n).astype(np.float32)
    return pos, nrm.astype(np.float32), d.astype(np.float32), c


def _sphere_grid(n_th=400, n_ph=800):
    th = (np.arange(n_th) + 0.5) / n_th * math.pi
    ph = (np.arange(n_ph) + 0.5) / n_ph * 2 * math.pi
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(Pp), np.sin(T) * np.sin(Pp), np.cos(T)], -1).reshape(-1, 3)
    w = (np.sin(T) * (math.pi / n_th) * (2 * math.pi / n_ph)).reshape(-1)
    return d.astype(np.float32), w


@pytest.mark.parametrize("kraw", [-5.0, -1.0, 0.0, 1.5, 3.0])
def test_vmf_mixture_integrates_to_one(kraw):
    rng = np.random.default_rng(3)
    raw = rng.normal(size=40).astype(np.float32)
    raw[3::5] = kraw + rng.uniform(-0.5, 0.5, 8)
    d, w = _sphere_grid()
    pdf = oracle.guide_mix_pdf(np.repeat(raw[None], len(d), 0), d)
    assert abs(np.sum(pdf * w) - 1.0) < 2e-3


def test_decode_is_unit_mu_softmax_weights_clamped_kappa():
    raw = np.zeros((1, 40), np.float32)
    raw[0, 0:3] = [3, 0, 4]
    raw[0, 3] = 9.0            # clamped to 7
    raw[0, 8] = -9.0           # lobe 1 kappa_raw clamped to -5
    raw[0, 4::5] = np.arange(8, dtype=np.float32)
    params = np.zeros(NP, np.float32)
    params[OB5:OB5 + 40] = raw[0]
    _, lobes = oracle.guide_forward(params, np.zeros((1, 3)), np.array([[0, 0, 1.0]]), BOUNDS)
    L = lobes[0]
    assert np.allclose(L[0, :3], [0.6, 0, 0.8], atol=1e-7)
    assert abs(L[0, 3] - math.exp(7.0)) < 1e-3 and abs(L[1, 3] - math.exp(-5.0)) < 1e-8
    e = np.exp(np.arange(8) - 7.0)
    assert np.allclose(L[:, 4], e / e.sum(), rtol=1e-6)
    assert np.allclose(L[2:, :3], [0, 0, 1])   # zero direction -> +z


def test_nll_grad_raw_matches_finite_differences():
    rng = np.random.default_rng(11)
    raw = rng.normal(size=(6, 40))
    raw[:, 3::5] = rng.uniform(-4, 6, (6, 8))
    pos, nrm, d, c = _rand_inputs(6, 12)
    g = oracle.guide_nll_grad_raw(raw, d)
    h = 1e-6
    for i in range(6):
        for o in range(40):
            rp, rm = raw.copy(), raw.copy()
            rp[i, o] += h
            rm[i, o] -= h
            fd = (oracle.guide_nll(rp[i:i + 1], d[i:i + 1])[0] - oracle.guide_nll(rm[i:i + 1], d[i:i + 1])[0]) / (2 * h)
            assert abs(fd - g[i, o]) <= 1e-5 * max(1.0, abs(fd)), (i, o, fd, g[i, o])


def test_parameter_gradient_matches_finite_differences():
    params = _rand_params(5)
    pos, nrm, d, c = _rand_inputs(64, 6)
    loss, grad = oracle.guide_loss_grad(params, pos, nrm, d, c, BOUNDS)
    rng = np.random.default_rng(7)
    picks = []
    for off, n in [(OW1, 9216), (OB1, 64), (OW2, 4096), (OB2, 64), (OW3, 4096), (OW4, 4096), (OB4, 64), (OW5, 2560),
                   (OB5, 40)]:
        picks += list(off + rng.choice(n, 4, replace=False))
    h = 2e-3
    bad = []
    for p in picks:
        pp, pm = params.copy(), params.copy()
        pp[p] += h
        pm[p] -= h
        hp = float(pp[p]) - float(params[p])
        hm = float(params[p]) - float(pm[p])
        lp, _ = oracle.guide_loss_grad(pp, pos, nrm, d, c, BOUNDS)
        lm, _ = oracle.guide_loss_grad(pm, pos, nrm, d, c, BOUNDS)
        fd = (lp - lm) / (hp + hm)
        if abs(fd - grad[p]) > 2e-3 * max(abs(fd), 1e-2):
            bad.append((p, fd, grad[p]))
    # a ReLU kink inside the +-h interval can spoil a difference; allow one
    assert len(bad) <= 1, bad


def test_encoding_layout_routes_feature_to_output():
    # W1 row 0 selects feature q; hidden layers pass unit 0 through; W5 row 0 reads it.
    for q in (0, 1, 23, 24 + 2 * 5, 24 * 3 + 2 * 11 + 1, 143):
        p = np.zeros(NP, np.float32)
        p[OW1 + q] = 1.0
        for ow in (OW2, OW3, OW4):
            p[ow + 0] = 1.0
        p[OW5 + 0] = 1.0
        pos = np.array([[1.3, 2.9, -0.7], [-3.1, 0.2, 3.3]], np.float32)
        nrm = np.array([[0.36, 0.48, 0.8], [-0.6, 0.0, 0.8]], np.float32)
        raw, _ = oracle.guide_forward(p, pos, nrm, BOUNDS)
        d, f, cs = q // 24, (q % 24) // 2, q % 2
        for i in range(2):
            c0 = (BOUNDS[:3] + BOUNDS[3:]) / 2
            h = (BOUNDS[3:] - BOUNDS[:3]) / 2
            x = np.concatenate([(pos[i] - c0) / h, nrm[i]]).astype(np.float32)
            a = math.pi * math.ldexp(float(x[d]), f)
            want = max(0.0, math.cos(a) if cs else math.sin(a))
            assert abs(raw[i, 0] - want) < 1e-6, (q, i)


def test_adam_closed_form_two_steps():
    p = np.array([1.0, -2.0, 0.5], np.float32)
    m = np.zeros(3, np.float32)
    v = np.zeros(3, np.float32)
    g1 = np.array([0.3, -0.02, 5.0])
    oracle.guide_adam(p, m, v, g1, 1, lr=0.01)
    # step 1: m_hat = g, v_hat = g^2 -> step = lr * g / (|g| + eps)
    assert np.allclose(p, [1.0 - 0.01, -2.0 + 0.01, 0.5 - 0.01], atol=1e-6)
    g2 = np.array([0.3, 0.02, -5.0])
    oracle.guide_adam(p, m, v, g2, 2, lr=0.01)
    m2 = 0.9 * 0.1 * g1 + 0.1 * g2
    v2 = 0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2
    step2 = 0.01 * (m2 / (1 - 0.81)) / (np.sqrt(v2 / (1 - 0.999 ** 2)) + 1e-8)
    assert np.allclose(p, np.array([0.99, -1.99, 0.49]) - step2, atol=1e-6)


def test_bf16_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -3.0e-39, 65504.0], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, vals[4], 65536.0], np.float32)
    got = oracle.bf16(vals)
    assert np.array_equal(got[[0, 1, 2, 3, 5]], want[[0, 1, 2, 3, 5]])


def test_emulation_stays_within_bf16_bound_of_exact():
    params = _rand_params(9)
    pos, nrm, d, c = _rand_inputs(256, 10)
    r64, _ = oracle.guide_forward(params, pos, nrm, BOUNDS, emulate=False)
    r16, _ = oracle.guide_forward(params, pos, nrm, BOUNDS, emulate=True)
    scale = np.abs(r64).max()
    assert np.abs(r16 - r64).max() < 0.05 * scale
    assert np.abs(r16 - r64).max() > 0      # the emulation does round
