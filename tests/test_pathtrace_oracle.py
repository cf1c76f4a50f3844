"""Pins the path-mode oracle (oracle/oracle_pt.cpp) on the CPU.

- the sin/cos polynomial against the C library to 2e-7;
- the hashed random numbers are uniform (moments, bins) and decorrelated across dimensions;
- furnace: inside a closed box whose walls all have albedo rho and emit Le, paths of at most D
  segments see Le (1 - rho^D) / (1 - rho) exactly in expectation; the oracle's estimate (NEE + MIS
  + cosine sampling) matches it within its statistical error, so a wrong MIS weight, pdf or
  throughput factor shows up as a bias;
- the occluded-light room renders finite, non-negative radiance, deterministically.
"""
import math

import numpy as np
import pytest

import oracle
from vermeer_amd import pathtrace, scenes


def test_sincos2pi_polynomial_matches_libm():
    us = np.concatenate([np.linspace(0, 1, 20001, endpoint=False), [0.125, 0.25, 0.375, 0.5, 0.75, 0.99999994]])
    err = 0.0
    for u in us:
        s, c = oracle.pt_sincos2pi(float(np.float32(u)))
        a = 2 * math.pi * float(np.float32(u))
        err = max(err, abs(s - math.sin(a)), abs(c - math.cos(a)))
    assert err < 2e-7


def test_hash_random_numbers_are_uniform():
    u = np.array([oracle.pt_rnd(p, 3, d, 12345) for p in range(4000) for d in range(4)]).reshape(4000, 4)
    assert np.all((u >= 0) & (u < 1))
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.003
    h, _ = np.histogram(u[:, 0], bins=10, range=(0, 1))
    assert h.min() > 330 and h.max() < 470
    c = np.corrcoef(u.T)
    assert np.abs(c - np.eye(4)).max() < 0.05


def test_small_seeds_give_unrelated_streams():
    """Seeds 1 and 2 (or any two seeds) must not share streams: XOR-ing the raw seed into the pixel
    index would give pixel p under seed 2 the numbers of pixel p ^ 3 under seed 1."""
    W = 1 << 12
    for s1, s2 in ((1, 2), (1, 3), (7, 8)):
        a = np.array([oracle.pt_rnd(p, 0, 2, s1) for p in range(W)])
        b = np.array([oracle.pt_rnd(p, 0, 2, s2) for p in range(W)])
        delta = s1 ^ s2
        assert not np.any(b == a[np.arange(W) ^ delta])
        assert len(np.intersect1d(a, b)) < 8            # 2^24 possible values: chance matches only
        assert abs(np.corrcoef(a, b)[0, 1]) < 0.06


@pytest.mark.parametrize("rho,depth", [(0.5, 4), (0.8, 3)])
def test_furnace_closed_form(rho, depth):
    d = scenes.furnace_box_scene(albedo=rho, emission=1.0)
    W = H = 24
    cam = pathtrace.camera_for(d, W, H)
    o = oracle.OraclePathTracer(d, W, H, cam, max_depth=depth, seed=3)
    npass = 48
    per_pass = []
    for p in range(npass):
        before = o.accum.copy()
        o.render_pass(p)
        per_pass.append((o.accum - before).mean())
    want = (1 - rho ** depth) / (1 - rho)
    m, se = np.mean(per_pass), np.std(per_pass) / math.sqrt(npass)
    assert abs(m - want) < 4 * se + 1e-3, (m, want, se)


def test_room_render_is_finite_and_deterministic():
    d = scenes.guided_room_scene(res=5, nboxes=300)
    cam = pathtrace.camera_for(d, 32, 32)
    a = oracle.OraclePathTracer(d, 32, 32, cam, max_depth=8, seed=7)
    b = oracle.OraclePathTracer(d, 32, 32, cam, max_depth=8, seed=7)
    for p in range(3):
        a.render_pass(p, nthreads=1)
        b.render_pass(p, nthreads=4)
    assert np.array_equal(a.accum, b.accum)
    assert np.all(np.isfinite(a.accum)) and a.accum.min() >= 0 and a.accum.max() > 0


def _alcove_view():
    """The room without boxes, seen from inside the alcove (behind the one-sided light) at its
    back wall: at depth 2 the image is the back wall's direct light from a close, large light at
    coordinates near 4.8, where a ray offset is large next to the light distance."""
    d = scenes.guided_room_scene(res=5, nboxes=0)
    c = dict(d["camera"], **{"from": (4.1, 3.3, 0.5), "to": (4.8, 3.0, 0.0)})
    return dict(d, camera=c)


def _mean_and_se(o, passes):
    per = []
    for p in range(passes):
        before = o.accum.copy()
        o.render_pass(p)
        per.append((o.accum - before).mean())
    per = np.array(per)
    return per.mean(), per.std() / math.sqrt(passes)


def test_nee_with_mis_equals_sampling_alone():
    """Next-event estimation + MIS and plain direction sampling estimate the same radiance; a shadow
    ray that stops on its own light (the offset-origin bug) loses most of the direct light."""
    d = _alcove_view()
    W = 24
    cam = pathtrace.camera_for(d, W, W)
    a = oracle.OraclePathTracer(d, W, W, cam, max_depth=2, seed=1)
    b = oracle.OraclePathTracer(d, W, W, cam, max_depth=2, seed=2)
    b.set_no_nee(True)
    ma, sa = _mean_and_se(a, 400)
    mb, sb = _mean_and_se(b, 1200)
    assert abs(ma - mb) < 4 * math.hypot(sa, sb), (ma, sa, mb, sb)


def _random_lobes(seed, kappa_raw=6.0):
    rng = np.random.default_rng(seed)
    dirs = rng.normal(size=(8, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    L = np.zeros((8, 5), np.float32)
    L[:, :3] = dirs
    L[:, 3] = math.exp(kappa_raw)
    L[:, 4] = 0.125
    return L


@pytest.mark.parametrize("kappa_raw", [6.0, -5.0])
def test_fixed_proposal_mixture_is_unbiased(kappa_raw):
    """One-sample MIS with a fixed vMF-mixture proposal (sharp or nearly uniform) agrees with plain
    cosine sampling on the alcove view, NEE included."""
    d = _alcove_view()
    W = 24
    cam = pathtrace.camera_for(d, W, W)
    a = oracle.OraclePathTracer(d, W, W, cam, max_depth=3, seed=1)
    g = oracle.OraclePathTracer(d, W, W, cam, max_depth=3, seed=2)
    g.set_fixed_guide(0.5, _random_lobes(1, kappa_raw))
    ma, sa = _mean_and_se(a, 400)
    mg, sg = _mean_and_se(g, 400)
    assert abs(ma - mg) < 4 * math.hypot(sa, sg), (ma, sa, mg, sg)
