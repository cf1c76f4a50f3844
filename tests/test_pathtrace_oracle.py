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
