"""GPU parity at the workloads' full sizes, through the C ABI.

- Motion blur: 2000 moving spheres (500,000 triangles, 2 keys), 1024^2 x 4 camera rays with van der
  Corput times and one point-light shadow ray per camera ray.  Hits (prim, geom slot, t, u, v) and
  per-ray traversal counters equal the oracle's bit for bit on a strided slice of 262,144 camera rays
  and of their shadow rays; over all 8,388,608 rays, any-hit occlusion equals closest-hit occlusion
  and repeated launches are identical.
- Occluded-light room (1,004,048 triangles), depth 8: plain path mode's FP64 accumulation equals the
  path oracle's bit for bit at 128^2 x 2 passes with the same ray counts; at 256^2 the guided
  renderer (trained online, batch 65,536) and the plain one agree on the mean radiance.
- Guide training at batch 65,536 (128 split-K partials of the weight gradient, summed by the Adam
  kernel): gradient and Adam step against the FP64 oracle in bf16 emulation.
"""
import math

import numpy as np
import pytest
import torch

import oracle
import vermeer_amd as va
from conftest import bits_equal
from vermeer_amd import guide, pathtrace, scenes

pytestmark = pytest.mark.gpu


def _fsame(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1).copy()).cuda()


def _trace_all(g, d_rays, n, shadow):
    st = torch.cuda.current_stream().cuda_stream
    d_hits = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
    d_geoms = torch.empty(n, dtype=torch.int32, device="cuda")
    g.trace_device(d_rays.data_ptr(), n, d_hits.data_ptr(), shadow=shadow, d_geoms=d_geoms.data_ptr(), stream=st)
    g.status(st)
    return d_hits, d_geoms


def _check_slice(g, o, rays, shadow, hits_all, geoms_all, stride, label):
    sl = rays[::stride]
    ref, rcnt = o.trace(sl, shadow=shadow, counters=True)
    h = hits_all[::stride]
    gm = geoms_all[::stride]
    bad = np.flatnonzero((h["prim"] != ref["prim"]) | (gm != ref["geom"]) | ~_fsame(h["t"], ref["t"]) |
                         ~_fsame(h["u"], ref["u"]) | ~_fsame(h["v"], ref["v"]))
    assert len(bad) == 0, f"{label}: {len(bad)} of {len(sl)} rays differ, first {bad[:4]}"
    # counters through the host path on the same slice
    _, _, cnt = g.trace(sl, shadow=shadow, want_counters=True)
    assert np.array_equal(cnt, rcnt), f"{label}: counters differ"


@pytest.fixture(scope="module")
def moving_spheres_2000():
    print("\n[case] 2000 moving spheres (500k triangles) on GPU and oracle", flush=True)
    verts, idx = scenes.moving_spheres()
    assert len(idx) == 500_000
    g = va.Scene(0)
    g.add_geom(va.PolyMesh(verts, idx))
    g.pre_render()
    o = oracle.OracleScene()
    o.add_mesh(verts, idx, motion_reference_indexing=False)
    o.commit()
    yield g, o
    g.close()


def test_motion_blur_workload_full_size(moving_spheres_2000):
    g, o = moving_spheres_2000
    rays = scenes.moving_spheres_camera_rays()
    n = len(rays)
    assert n == 4 * 1024 * 1024
    print(f"[case] motion camera rays {n}", flush=True)
    d_rays = _dev(rays)
    d_h1, d_g1 = _trace_all(g, d_rays, n, shadow=False)
    d_h2, _ = _trace_all(g, d_rays, n, shadow=False)
    assert torch.equal(d_h1, d_h2)
    hits = d_h1.cpu().numpy().view(va.HIT_DTYPE)
    geoms = d_g1.cpu().numpy()
    hit_frac = float((hits["prim"] >= 0).mean())
    assert 0.2 < hit_frac < 0.95, hit_frac
    _check_slice(g, o, rays, False, hits, geoms, 16, "camera")
    # camera rays with any-hit: occluded exactly where closest hit finds something
    d_ha, _ = _trace_all(g, d_rays, n, shadow=True)
    ha = d_ha.cpu().numpy().view(va.HIT_DTYPE)
    assert np.array_equal(ha["prim"] >= 0, hits["prim"] >= 0)

    srays = scenes.point_light_shadow_rays(rays, hits["prim"], hits["t"])
    print(f"[case] motion shadow rays {n}", flush=True)
    d_s = _dev(srays)
    d_sh, d_sg = _trace_all(g, d_s, n, shadow=True)
    sh = d_sh.cpu().numpy().view(va.HIT_DTYPE)
    sg = d_sg.cpu().numpy()
    _check_slice(g, o, srays, True, sh, sg, 16, "shadow")
    d_sc, _ = _trace_all(g, d_s, n, shadow=False)
    sc = d_sc.cpu().numpy().view(va.HIT_DTYPE)
    occluded = sh["prim"] >= 0
    assert np.array_equal(occluded, sc["prim"] >= 0)
    assert np.all(sh["t"][occluded] >= sc["t"][occluded])
    assert not occluded[hits["prim"] < 0].any()            # rays of misses can hit nothing
    frac = occluded[hits["prim"] >= 0].mean()
    print(f"[case] hit fraction {hit_frac:.3f}, shadowed fraction of hits {frac:.3f}", flush=True)
    assert 0.01 < frac < 0.99


@pytest.fixture(scope="module")
def room_full():
    d = scenes.guided_room_scene()
    assert sum(len(m["idx"]) for m in d["meshes"]) == 1_004_048
    return d


def test_room_plain_path_mode_bitexact_full_scene(room_full):
    W = 128
    print(f"\n[case] full room plain path mode {W}x{W} x 2 passes", flush=True)
    g = pathtrace.PathTracer(room_full, W, W, max_depth=8, seed=5)
    o = oracle.OraclePathTracer(room_full, W, W, g.cam, max_depth=8, seed=5)
    rays_o = 0
    for p in range(2):
        g.render_pass(p)
        n, _ = o.render_pass(p)
        rays_o += n
        same = (g.accum == o.accum) | (np.isnan(g.accum) & np.isnan(o.accum))
        assert same.all(), f"pass {p}: {int((~same).sum())} values differ"
    assert g.stats()[0] == rays_o
    assert (g.accum > 0).mean() > 0.01          # lit only by indirect paths through the slit
    g.close()


def test_room_guided_and_plain_agree_full_scene(room_full):
    """256^2, depth 8: per-pass mean radiance of plain and guided (48 training steps of 65,536 samples
    per pass for 6 passes, then frozen at alpha 0.65) agree within 5 standard errors."""
    W = 256
    print(f"\n[case] full room guided vs plain {W}x{W}", flush=True)
    lo = np.array([-4.1, -0.2, -4.1], np.float32)
    hi = np.array([4.9, 4.1, 4.1], np.float32)
    means = {}
    for guided in (False, True):
        pt = pathtrace.PathTracer(room_full, W, W, max_depth=8, seed=21 if guided else 22)
        gd = None
        if guided:
            gd = guide.Guide(np.concatenate([lo, hi]), seed=1, max_batch=65536)
            pt.set_guide(gd, alpha=0.5, record_stride=1, train_steps=48, batch=65536, lr=1e-3)
        per = []
        for p in range(40):
            if guided and p == 6:
                pt.set_guide(gd, alpha=0.65, record_stride=1, train_steps=0, batch=65536, lr=1e-3)
            before = pt.accum.copy()
            pt.render_pass(p)
            if p >= 6:
                per.append((pt.accum - before).mean())
        pt.close()
        if gd is not None:
            gd.close()
        per = np.array(per)
        means[guided] = (per.mean(), per.std(ddof=1) / math.sqrt(len(per)))
    (mp, sp), (mg, sg) = means[False], means[True]
    print(f"[case] plain {mp:.5f} +- {sp:.5f}, guided {mg:.5f} +- {sg:.5f}", flush=True)
    assert mp > 0 and abs(mp - mg) < 5 * math.hypot(sp, sg)


def _inputs(n, seed, bounds):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(bounds[:3], bounds[3:], (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = rng.uniform(0, 2, n).astype(np.float32)
    c[rng.uniform(size=n) < 0.3] = 0.0
    return pos, nrm.astype(np.float32), d.astype(np.float32), c


def test_guide_training_step_at_batch_65536():
    bounds = np.array([-4, 0, -4, 4, 4, 4], np.float32)
    print("\n[case] guide gradient + Adam at batch 65536", flush=True)
    net = guide.Guide(bounds, seed=5, max_batch=65536)
    p = net.params
    p[21760:24320] *= 8.0
    p[24320:24360] = np.random.default_rng(1).uniform(-1, 1, 40)
    net.params = p
    pos, nrm, d, c = _inputs(65536, 3, bounds)
    loss = net.train_step(pos, nrm, d, c, update=False)
    gr = net.grad()
    lo, want = oracle.guide_loss_grad(net.params, pos, nrm, d, c, bounds, emulate=True)
    assert abs(loss - lo) < 1e-3 * max(1.0, abs(lo))
    blocks = [(0, 9216), (9216, 9280), (9280, 13376), (13376, 13440), (13440, 17536), (17536, 17600),
              (17600, 21696), (21696, 21760), (21760, 24320), (24320, 24360)]
    for a, b in blocks:
        rel = np.linalg.norm(gr[a:b] - want[a:b]) / np.linalg.norm(want[a:b])
        print(f"[case] block {a}:{b} rel {rel:.2e}", flush=True)
        assert rel <= 2e-2, (a, b, rel)
    # one Adam step (the gradient the update uses is the same batch's) against the oracle's step
    net.params = p                       # resets Adam: step 1
    net.train_step(pos, nrm, d, c, lr=1e-3, update=True)
    g64 = net.grad().astype(np.float64)
    got = net.params
    want, m, v = p.copy(), np.zeros(len(p), np.float32), np.zeros(len(p), np.float32)
    oracle.guide_adam(want, m, v, g64, 1, lr=1e-3)
    assert np.abs(got - want).max() < 1e-6
    net.close()
