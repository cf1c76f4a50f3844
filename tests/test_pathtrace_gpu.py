"""GPU path mode against the path oracle and closed forms, through the C ABI.

Bars:
- plain mode (alpha = 0): every pixel's FP64 accumulation equals the oracle's bit for bit (same
  random numbers, same float operations, bit-exact traversal), and the ray counts agree;
- furnace: plain and guided (alpha = 0.5, with online training) estimates converge to
  Le (1 - rho^D) / (1 - rho) within 4 standard errors: the one-sample MIS mixture stays unbiased;
- guided and plain agree on the occluded-light room's mean radiance within their errors;
- after online training, the guided estimate of the room at equal samples has an RMSE below 0.8x
  the plain estimate's, against a high-sample plain render.
"""
import math

import numpy as np
import pytest

import oracle
from vermeer_amd import guide, pathtrace, scenes

pytestmark = pytest.mark.gpu


def _eq_nan(a, b):
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


@pytest.mark.parametrize("res,nboxes,W,depth", [(5, 2000, 64, 8), (10, 6000, 48, 5)])
def test_plain_matches_oracle_bitexact(res, nboxes, W, depth):
    print(f"\n[case] plain bit-exact res={res} boxes={nboxes} {W}x{W} depth {depth}", flush=True)
    d = scenes.guided_room_scene(res=res, nboxes=nboxes)
    g = pathtrace.PathTracer(d, W, W, max_depth=depth, seed=5)
    o = oracle.OraclePathTracer(d, W, W, g.cam, max_depth=depth, seed=5)
    rays_o = 0
    for p in range(4):
        g.render_pass(p)
        n, _ = o.render_pass(p)
        rays_o += n
        diff = np.argwhere(g.accum != o.accum)
        assert _eq_nan(g.accum, o.accum), f"pass {p}: {len(diff)} values differ, first {diff[:3]}"
    rays_g, _ = g.stats()
    assert rays_g == rays_o
    assert g.accum.max() > 0
    g.close()


def _furnace_run(rho, depth, guided, passes=64, W=32):
    d = scenes.furnace_box_scene(albedo=rho, emission=1.0)
    pt = pathtrace.PathTracer(d, W, W, max_depth=depth, seed=9)
    if guided:
        gd = guide.Guide(np.array([-1, -1, -1, 1, 1, 1], np.float32), seed=3, max_batch=1024)
        pt.set_guide(gd, alpha=0.5, record_stride=1, train_steps=2, batch=1024, lr=3e-3)
    per = []
    for p in range(passes):
        before = pt.accum.copy()
        pt.render_pass(p)
        per.append((pt.accum - before).mean())
    pt.close()
    return np.array(per)


@pytest.mark.parametrize("guided", [False, True])
def test_furnace_closed_form(guided):
    rho, depth = 0.5, 4
    per = _furnace_run(rho, depth, guided)
    want = (1 - rho ** depth) / (1 - rho)
    m, se = per.mean(), per.std() / math.sqrt(len(per))
    print(f"[case] furnace guided={guided}: {m:.5f} +- {se:.5f} (want {want:.5f})", flush=True)
    assert abs(m - want) < 4 * se + 1e-3


def test_frozen_guide_keeps_its_weights_and_stays_unbiased():
    """Train online, then attach the same guide again with no training steps and a larger sampling
    probability (0.65). Re-attaching keeps the trained weights, a frozen pass records no samples and
    leaves the weights unchanged, and the furnace estimate still meets the closed form."""
    rho, depth, W = 0.5, 4, 32
    print("\n[case] furnace, guide trained 6 passes then frozen at alpha 0.65", flush=True)
    d = scenes.furnace_box_scene(albedo=rho, emission=1.0)
    pt = pathtrace.PathTracer(d, W, W, max_depth=depth, seed=11)
    gd = guide.Guide(np.array([-1, -1, -1, 1, 1, 1], np.float32), seed=3, max_batch=1024)
    p0 = gd.params
    pt.set_guide(gd, alpha=0.5, record_stride=1, train_steps=2, batch=1024, lr=3e-3)
    per = []
    for p in range(64):
        if p == 6:
            trained = gd.params
            assert not np.array_equal(trained, p0)
            pt.set_guide(gd, alpha=0.65, record_stride=1, train_steps=0, batch=1024, lr=3e-3)
            assert np.array_equal(gd.params, trained)
        before = pt.accum.copy()
        pt.render_pass(p)
        per.append((pt.accum - before).mean())
        if p >= 6:
            assert pt.stats()[1] == 0
    assert np.array_equal(gd.params, trained)
    pt.close()
    gd.close()
    per = np.array(per)
    want = (1 - rho ** depth) / (1 - rho)
    m, se = per.mean(), per.std() / math.sqrt(len(per))
    print(f"[case] frozen guide furnace: {m:.5f} +- {se:.5f} (want {want:.5f})", flush=True)
    assert abs(m - want) < 4 * se + 1e-3


@pytest.fixture(scope="module")
def room():
    return scenes.guided_room_scene(res=10, nboxes=20000)


def _render(desc, W, passes, guided, seed, first_pass=0, train_passes=None):
    pt = pathtrace.PathTracer(desc, W, W, max_depth=8, seed=seed)
    if guided:
        lo = np.array([-4.1, -0.2, -4.1], np.float32)
        hi = np.array([4.9, 4.1, 4.1], np.float32)
        gd = guide.Guide(np.concatenate([lo, hi]), seed=1, max_batch=16384)
        pt.set_guide(gd, alpha=0.5, record_stride=4, train_steps=4, batch=16384, lr=2e-3)
        pt._gd = gd
    for p in range(passes):
        train = train_passes is None or p < train_passes
        pt.render_pass(first_pass + p, train=train)
    img = pt.image
    pt.close()
    return img


def _per_pass_means(desc, W, passes, guided, seed):
    pt = pathtrace.PathTracer(desc, W, W, max_depth=8, seed=seed, leaf_size=2)
    if guided:
        gd = guide.Guide(pathtrace.scene_bounds(desc), seed=1, max_batch=16384)
        pt.set_guide(gd, alpha=0.5, record_stride=1, train_steps=4, batch=16384, lr=2e-3)
        pt._gd = gd
    per = []
    for p in range(passes):
        before = pt.accum.copy()
        pt.render_pass(p)
        per.append((pt.accum - before).mean())
    pt.close()
    return np.array(per)


def test_guided_and_plain_agree_on_the_room(room):
    """The guided estimator (trained online from the first pass) and plain sampling estimate the same
    image mean: within 4 combined standard errors of their per-pass means."""
    print("[case] room mean guided vs plain", flush=True)
    a = _per_pass_means(room, 64, 128, False, seed=21)
    b = _per_pass_means(room, 64, 128, True, seed=22)
    se = math.hypot(a.std() / math.sqrt(len(a)), b.std() / math.sqrt(len(b)))
    print(f"[case] plain {a.mean():.4f} guided {b.mean():.4f} (combined se {se:.4f})", flush=True)
    assert abs(a.mean() - b.mean()) < 4 * se


def test_guiding_reduces_error_at_equal_samples(room):
    """After online training, the guided estimate at equal samples has a markedly lower RMSE against a
    high-sample plain render than plain sampling (the room's light reaches the camera's view only
    through the slit, which the guide learns to aim at)."""
    print("[case] guided vs plain RMSE", flush=True)
    W = 128
    scene = pathtrace.build_scene(room, leaf_size=2)

    def render(passes, guided, seed, train_passes=0):
        pt = pathtrace.PathTracer(room, W, W, max_depth=8, seed=seed, scene=scene)
        if guided:
            gd = guide.Guide(pathtrace.scene_bounds(room), seed=1, max_batch=16384)
            pt.set_guide(gd, alpha=0.5, record_stride=1, train_steps=16, batch=16384, lr=3e-3)
            pt._gd = gd
            for p in range(train_passes):
                pt.render_pass(10_000 + p)
            pt.accum[:] = 0
            pt.passes = 0
        for p in range(passes):
            pt.render_pass(p)
        img = pt.image
        pt.close()
        return img

    ref = render(512, False, 100)
    plain = render(32, False, 7)
    guided = render(32, True, 8, train_passes=48)
    e_p = math.sqrt(np.mean((plain - ref) ** 2))
    e_g = math.sqrt(np.mean((guided - ref) ** 2))
    print(f"[case] RMSE plain {e_p:.4f} guided {e_g:.4f} ratio {e_g / e_p:.3f}", flush=True)
    assert e_g < 0.8 * e_p
    scene.close()


def _fixed_sharp_guide(bounds, dirs, kappa_raw=6.0):
    """A guide whose output ignores its input: 8 lobes along `dirs`, log kappa = kappa_raw, equal weights."""
    gd = guide.Guide(bounds, seed=3, max_batch=1024)
    p = np.zeros(guide.NPARAMS, np.float32)
    b5 = np.zeros(40, np.float32)
    for k in range(8):
        b5[5 * k:5 * k + 3] = dirs[k]
        b5[5 * k + 3] = kappa_raw
    p[24320:24360] = b5
    gd.params = p
    return gd


def test_furnace_with_fixed_sharp_guide():
    """A sharp fixed proposal (8 lobes, kappa ~ 400) must not bias the estimate: furnace closed form."""
    rho, depth = 0.5, 4
    d = scenes.furnace_box_scene(albedo=rho, emission=1.0)
    rng = np.random.default_rng(4)
    dirs = rng.normal(size=(8, 3)).astype(np.float32)
    gd = _fixed_sharp_guide(np.array([-1, -1, -1, 1, 1, 1], np.float32), dirs)
    pt = pathtrace.PathTracer(d, 32, 32, max_depth=depth, seed=9)
    pt.set_guide(gd, alpha=0.5, record_stride=0, train_steps=0, batch=1024)
    per = []
    for p in range(96):
        before = pt.accum.copy()
        pt.render_pass(p)
        per.append((pt.accum - before).mean())
    pt.close()
    per = np.array(per)
    want = (1 - rho ** depth) / (1 - rho)
    m, se = per.mean(), per.std() / math.sqrt(len(per))
    print(f"[case] furnace fixed sharp guide: {m:.5f} +- {se:.5f} (want {want:.5f})", flush=True)
    assert abs(m - want) < 4 * se + 1e-3


def test_guided_fixed_proposal_matches_oracle():
    """Guided sampling with a fixed proposal against the oracle's restatement of it: the GPU guide's
    lobes (read back) drive the oracle.  The vMF density goes through the devices' own exp/log, so
    a path can change after a last-bit difference; most pixels still agree to 1e-4 and the image
    means agree within their errors."""
    d = scenes.guided_room_scene(res=5, nboxes=500)
    d = dict(d, camera=dict(d["camera"], **{"from": (4.1, 3.3, 0.5), "to": (4.8, 3.0, 0.0)}))
    W = 48
    rng = np.random.default_rng(5)
    dirs = rng.normal(size=(8, 3)).astype(np.float32)
    gd = _fixed_sharp_guide(pathtrace.scene_bounds(d), dirs)
    lobes = gd.infer(np.zeros((1, 3), np.float32), np.array([[0, 1, 0]], np.float32))[0]
    g = pathtrace.PathTracer(d, W, W, max_depth=4, seed=6)
    g.set_guide(gd, alpha=0.5, record_stride=0, train_steps=0, batch=1024)
    o = oracle.OraclePathTracer(d, W, W, g.cam, max_depth=4, seed=6)
    o.set_fixed_guide(0.5, lobes)
    for p in range(8):
        g.render_pass(p)
        o.render_pass(p)
    a, b = g.accum, o.accum
    close = np.abs(a - b) <= 1e-4 * np.maximum(np.abs(b), 1e-3)
    print(f"[case] guided fixed proposal: {close.mean():.4f} of values within 1e-4", flush=True)
    assert close.mean() > 0.9
    assert abs(a.mean() - b.mean()) < 0.02 * b.mean()
    g.close()
