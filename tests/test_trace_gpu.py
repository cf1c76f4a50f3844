"""GPU parity: the HIP traversal against the CPU oracle, through the C ABI.

Bar: bit-exact prim IDs, geom slots, t, u and v on every ray (the north-star
allows a 1e-6 edge band and 1e-5 relative t/u/v; the kernel meets the
stricter bar), and identical per-ray traversal statistics (node visits, leaf
visits, triangle tests, stack depth), which shows the GPU walks the same
nodes in the same order as the reference.
"""
import numpy as np
import pytest

import oracle
import vermeer_amd as va
from conftest import bits_equal
from vermeer_amd import scenes

pytestmark = pytest.mark.gpu


def _fdiff(a, b):
    """Float32 arrays differ bitwise, except that any NaN equals any NaN: the payload and sign of a
    NaN produced by 0*inf differ between x86 (negative default NaN) and the GPU (positive)."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return (a.view(np.uint32) != b.view(np.uint32)) & ~(np.isnan(a) & np.isnan(b))


def _same(hits, geoms, ref, what=""):
    bad = np.flatnonzero((hits["prim"] != ref["prim"]) | (geoms != ref["geom"]) | _fdiff(hits["t"], ref["t"]) |
                         _fdiff(hits["u"], ref["u"]) | _fdiff(hits["v"], ref["v"]))
    if len(bad):
        rows = "; ".join(f"#{i}: gpu {hits[i]} geom {geoms[i]} vs oracle {ref[i]}" for i in bad[:4])
        raise AssertionError(f"{what}: {len(bad)} of {len(hits)} rays differ; {rows}")


def _pair(meshes, **kw):
    g = va.Scene(0)
    o = oracle.OracleScene()
    for v, f in meshes:
        g.add_geom(va.PolyMesh(v, f, **kw))
        o.add_mesh(v, f, motion_reference_indexing=kw.get("motion_reference_indexing", False))
    g.pre_render()
    o.commit()
    return g, o


@pytest.fixture(scope="module")
def sphere_ground():
    print("\n[case] sphere_ground 256x128 + ground quad", flush=True)
    return _pair([scenes.uv_sphere(), scenes.ground_quad()])


@pytest.mark.parametrize("shadow", [False, True])
def test_pinhole_1024_bitexact(sphere_ground, shadow):
    g, o = sphere_ground
    rays = scenes.pinhole_rays(1024, 1024)
    print(f"[case] pinhole 1024^2 shadow={shadow}", flush=True)
    hits, geoms, _ = g.trace(rays, shadow=shadow)
    ref, _ = o.trace(rays, shadow=shadow)
    _same(hits, geoms, ref, "pinhole")
    assert 0.3 < (hits["prim"] >= 0).mean() < 0.9


def test_sorted_pinhole_rays_bitexact(sphere_ground):
    """A pinhole camera's rays share one origin cell, so whole wavefronts fall into a few buckets of the
    ray sort; the sorted trace still returns every hit to its ray's slot."""
    g, o = sphere_ground
    rays = scenes.pinhole_rays(512, 509)
    print("[case] pinhole 512x509 sorted, bits 5", flush=True)
    c = g.trace_config()
    g.set_trace_config(c["block"], c["lds_stack"], c["waves_per_cu"], 5)
    try:
        hits, geoms, _ = g.trace(rays)
    finally:
        g.set_trace_config(c["block"], c["lds_stack"], c["waves_per_cu"], 0)
    ref, _ = o.trace(rays)
    _same(hits, geoms, ref, "sorted pinhole")


@pytest.mark.parametrize("shadow", [False, True])
def test_uniform_1m_bitexact(sphere_ground, shadow):
    g, o = sphere_ground
    rays = scenes.uniform_rays(1_000_000, seed=1)
    print(f"[case] uniform 1M shadow={shadow}", flush=True)
    hits, geoms, _ = g.trace(rays, shadow=shadow)
    ref, _ = o.trace(rays, shadow=shadow)
    _same(hits, geoms, ref, "uniform")


def test_traversal_counters_equal_oracle(sphere_ground):
    g, o = sphere_ground
    rays = np.concatenate([scenes.pinhole_rays(256, 256), scenes.uniform_rays(100_000, seed=4)])
    hits, geoms, cnt = g.trace(rays, want_counters=True)
    ref, rc = o.trace(rays, counters=True)
    _same(hits, geoms, ref, "counters run")
    assert np.array_equal(cnt, rc)


def test_prebuilt_accel_path_matches(sphere_ground):
    """The cgo path: accels built by the reference's builder (here the oracle's) uploaded as-is."""
    _, o = sphere_ground
    inf = o.info()
    g = va.Scene(0)
    meshes = [scenes.uv_sphere(), scenes.ground_quad()]
    for v, f in meshes:
        g.add_geom(va.PolyMesh(v, f))
    g.pre_render()
    gp = va.Scene(0)
    # slots are post-sort; add in insertion order with the oracle's per-mesh accels
    slot_of = {int(ins): s for s, ins in enumerate(inf["geom_order"])}
    for ins, (v, f) in enumerate(meshes):
        nodes, idxp, _ = o.mesh_accel(slot_of[ins])
        b = nodes_bounds(nodes)
        gp.add_prebuilt(v, idxp, nodes, b)
    gp.pre_render()
    rays = scenes.uniform_rays(200_000, seed=8)
    h1, g1, _ = g.trace(rays)
    h2, g2, _ = gp.trace(rays)
    assert h1.tobytes() == h2.tobytes() and np.array_equal(g1, g2)


def nodes_bounds(nodes):
    """Root box = union of the root's non-empty children (qbvh/build.go:277-284)."""
    lo = np.full(3, np.inf, np.float32)
    hi = np.full(3, -np.inf, np.float32)
    for c in range(4):
        if nodes[0]["children"][c] == -1:
            continue
        for k in range(3):
            lo[k] = min(lo[k], nodes[0]["boxes"][c + 4 * k])
            hi[k] = max(hi[k], nodes[0]["boxes"][c + 12 + 4 * k])
    return np.concatenate([lo, hi])


def test_builder_accel_equals_oracle_accel(sphere_ground):
    g, o = sphere_ground
    for slot in range(2):
        n1, i1, p1 = g.mesh_accel(slot)
        n2, i2, p2 = o.mesh_accel(slot)
        assert n1.tobytes() == n2.tobytes() and np.array_equal(i1, i2) and np.array_equal(p1, p2)
    assert np.array_equal(g.info()["geom_order"], o.info()["geom_order"])


def test_edge_case_rays(sphere_ground):
    g, o = sphere_ground
    P = np.array([[0, 3, 0], [0, 3, 0], [0, 0, 0], [5, -1, 5], [0, 1, 0], [0.0, -1.0, 0.0], [0.3, 2, 0.1],
                  [0, 0, 0], [2, 0.5, 0]], np.float32)
    D = np.array([[0, -1, 0], [0, -0.0, 0], [0, 0, 0], [-1, 0, -1], [1, 0, 0], [1, 0, 0], [-0.0, -1, 0.0],
                  [1e-30, -1, 0], [-1, -1e-20, 0]], np.float32)
    rays = scenes.make_rays(P, D)
    rays["tmax"][4] = 0.5
    for shadow in (False, True):
        hits, geoms, _ = g.trace(rays, shadow=shadow)
        ref, _ = o.trace(rays, shadow=shadow)
        _same(hits, geoms, ref, "edge cases")


def test_edge_rays_mixed_into_a_large_batch(sphere_ground):
    """Zero, tiny and infinite direction components and non-finite origins scattered through a batch
    of ordinary rays: wavefronts holding one of them take the exact box test, the rest the
    hardware min/max form; ray refill mixes both kinds in one wavefront over time."""
    g, o = sphere_ground
    rng = np.random.default_rng(23)
    n = 1 << 16
    P = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    D = rng.normal(size=(n, 3)).astype(np.float32)
    special = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1], [1e-30, -1, 0], [-1, -1e-20, 0], [0, 0, 0],
                        [np.inf, -1, 0], [-0.0, -1, 0.0], [1e-38, 1e-38, -1]], np.float32)
    idx = np.arange(0, n, 53)
    D[idx] = special[np.arange(len(idx)) % len(special)]
    P[idx[::7]] = np.array([np.nan, 0, 0], np.float32)
    P[idx[3::7]] = np.array([0, np.inf, 0], np.float32)
    rays = scenes.make_rays(P, D)
    for shadow in (False, True):
        hits, geoms, _ = g.trace(rays, shadow=shadow)
        ref, _ = o.trace(rays, shadow=shadow)
        _same(hits, geoms, ref, f"mixed edge rays shadow={shadow}")


@pytest.mark.parametrize("motion", [False, True])
def test_scene_with_analytic_spheres_bitexact(motion):
    """Analytic sphere geoms (a SphereLight's, builtin/geom/sphere/trace.go) beside meshes, static and
    motion scenes: closest hit and shadow rays from outside, inside and grazing, against the oracle."""
    print(f"\n[case] analytic spheres motion={motion}", flush=True)
    gv, gf = scenes.ground_quad()
    sv, sf = scenes.uv_sphere(0.5, 32, 16)
    sv = sv + np.array([0.8, 0.6, 0.0], np.float32)
    g = va.Scene(0)
    o = oracle.OracleScene()
    g.add_geom(va.PolyMesh(gv, gf))
    o.add_mesh(gv, gf)
    if motion:
        keys = np.stack([sv, sv + np.array([0.0, 0.3, 0.0], np.float32)])
        g.add_geom(va.PolyMesh(keys, sf))
        o.add_mesh(keys, sf, motion_reference_indexing=False)
    else:
        g.add_geom(va.PolyMesh(sv, sf))
        o.add_mesh(sv, sf)
    for c, r in (((-0.7, 0.5, 0.2), 0.45), ((0.0, 1.6, -0.4), 0.25), ((-0.2, 0.35, 0.9), 0.3)):
        g.add_sphere(c, r, 0)
        o.add_sphere(c, r, 0)
    g.pre_render()
    o.commit()
    rng = np.random.default_rng(31)
    n = 1 << 15
    P = rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32)
    P[:, 1] = rng.uniform(0.05, 2.0, n)
    D = rng.normal(size=(n, 3)).astype(np.float32)
    P[:200] = np.array([-0.7, 0.5, 0.2], np.float32)          # inside a sphere
    rays = scenes.make_rays(P, D, time=rng.uniform(0, 1, n).astype(np.float32) if motion else 0.0)
    for shadow in (False, True):
        hits, geoms, _ = g.trace(rays, shadow=shadow)
        ref, _ = o.trace(rays, shadow=shadow)
        _same(hits, geoms, ref, f"spheres motion={motion} shadow={shadow}")
    assert (ref["prim"] == 0).sum() > 1000


def test_empty_batch(sphere_ground):
    g, _ = sphere_ground
    hits, geoms, _ = g.trace(np.zeros(0, va.RAY_DTYPE))
    assert len(hits) == 0


@pytest.mark.parametrize("cfg", [(64, 8, 8), (64, 12, 32), (128, 24, 16), (256, 32, 32), (256, 8, 4)])
def test_every_launch_config_gives_identical_hits(sphere_ground, cfg):
    g, o = sphere_ground
    rays = scenes.uniform_rays(300_000, seed=12)
    g.set_trace_config(*cfg)
    try:
        hits, geoms, _ = g.trace(rays)
    finally:
        g.set_trace_config(256, 16, 16)
    ref, _ = o.trace(rays)
    _same(hits, geoms, ref, f"config {cfg}")


def test_device_pointer_path_with_torch(sphere_ground):
    import torch
    g, o = sphere_ground
    rays = scenes.uniform_rays(100_000, seed=13)
    d_rays = torch.from_numpy(rays.view(np.uint8).copy()).cuda()
    d_hits = torch.empty(len(rays) * 16, dtype=torch.uint8, device="cuda")
    d_geoms = torch.empty(len(rays), dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    g.trace_device(d_rays.data_ptr(), len(rays), d_hits.data_ptr(), d_geoms=d_geoms.data_ptr(), stream=stream)
    g.status(stream)
    hits = d_hits.cpu().numpy().view(va.HIT_DTYPE)
    ref, _ = o.trace(rays)
    _same(hits, d_geoms.cpu().numpy(), ref, "device pointers")


@pytest.fixture(scope="module")
def icosphere_field():
    print("\n[case] icosphere field 4096 x 1280 tris", flush=True)
    v, f = scenes.icosphere_field()[:2]
    g, o = _pair([(v, f)])
    return g, o, v, f


def test_icosphere_field_bounce_rays_bitexact(icosphere_field):
    g, o, v, f = icosphere_field
    rays, _ = scenes.bounce_rays(v, f, 300_000, seed=11)
    print("[case] icosphere field 300k bounce rays", flush=True)
    hits, geoms, cnt = g.trace(rays, want_counters=True)
    ref, rc = o.trace(rays, counters=True)
    _same(hits, geoms, ref, "icosphere field")
    assert np.array_equal(cnt, rc)
    hits_s, geoms_s, _ = g.trace(rays, shadow=True)
    ref_s, _ = o.trace(rays, shadow=True)
    _same(hits_s, geoms_s, ref_s, "icosphere field shadow")


@pytest.mark.parametrize("sort", [1, 4, 6])
def test_sorted_traversal_is_bitexact(icosphere_field, sort):
    """Reordering rays by (origin cell, direction octant) changes only the order of work: every
    ray's hit, geom slot and traversal counters land in its own slot, equal to the oracle's."""
    g, o, v, f = icosphere_field
    rays, _ = scenes.bounce_rays(v, f, 200_003, seed=23)   # a partial last wavefront
    rays[::97]["P"] = np.nan          # NaN origins fall into cell 0 and still miss
    print(f"[case] icosphere field 200k bounce rays, sort bits {sort}", flush=True)
    g.set_trace_config(256, 16, 16, sort)
    try:
        hits, geoms, cnt = g.trace(rays, want_counters=True)
        hits_s, geoms_s, _ = g.trace(rays, shadow=True)
    finally:
        g.set_trace_config(256, 16, 16, 0)
    ref, rc = o.trace(rays, counters=True)
    _same(hits, geoms, ref, f"sorted {sort}")
    assert np.array_equal(cnt, rc)
    ref_s, _ = o.trace(rays, shadow=True)
    _same(hits_s, geoms_s, ref_s, f"sorted {sort} shadow")


def test_icosphere_field_full_16m_properties(icosphere_field):
    """Full size (16,777,216 rays): sizes the oracle cannot check in seconds are checked by
    properties: a sampled slice equals the oracle, two runs are bit-identical, every hit lies
    within the ray's range and its barycentrics are valid, any-hit agrees with closest-hit on
    occlusion."""
    import torch
    g, o, v, f = icosphere_field
    n = 1 << 24
    print("[case] icosphere field 16M rays", flush=True)
    d_rays, _ = scenes.bounce_rays_device(v, f, n, seed=17)
    d_h1 = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
    d_h2 = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
    d_hs = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    g.trace_device(d_rays.data_ptr(), n, d_h1.data_ptr(), stream=st)
    g.trace_device(d_rays.data_ptr(), n, d_h2.data_ptr(), stream=st)
    g.trace_device(d_rays.data_ptr(), n, d_hs.data_ptr(), shadow=True, stream=st)
    g.status(st)
    assert torch.equal(d_h1, d_h2)
    h = d_h1.cpu().numpy().view(va.HIT_DTYPE)
    hs = d_hs.cpu().numpy().view(va.HIT_DTYPE)
    hit = h["prim"] >= 0
    assert np.array_equal(hit, hs["prim"] >= 0)
    assert np.all(h["t"][hit] > 0) and np.all(np.isinf(h["t"][~hit]))
    w = 1.0 - h["u"][hit].astype(np.float64) - h["v"][hit].astype(np.float64)
    assert np.all(h["u"][hit] >= 0) and np.all(h["v"][hit] >= 0) and np.all(w > -1e-6)
    sl = slice(n - 50_000, n)
    ref, _ = o.trace(d_rays[sl.start * 32:].cpu().numpy().view(va.RAY_DTYPE))
    assert np.array_equal(h["prim"][sl], ref["prim"]) and bits_equal(h["t"][sl], ref["t"])


def test_motion_blur_scene_bitexact():
    verts, idx = scenes.moving_spheres(count=300)
    g, o = _pair([(verts, idx)])
    gs = scenes.rng(3)
    n = 200_000
    P = gs.uniform(-25, 25, (n, 3))
    D = scenes.uniform_sphere_dirs(gs, n)
    rays = scenes.make_rays(P, D, time=gs.random(n))
    print("[case] motion 300 spheres 200k rays", flush=True)
    for shadow in (False, True):
        hits, geoms, cnt = g.trace(rays, shadow=shadow, want_counters=True)
        ref, rc = o.trace(rays, shadow=shadow, counters=True)
        _same(hits, geoms, ref, f"motion shadow={shadow}")
        assert np.array_equal(cnt, rc)
    assert (hits["prim"] >= 0).mean() > 0.01


@pytest.fixture(scope="module")
def mixed_motion():
    """Static and moving meshes in one scene (one moving mesh with the reference's slot indexing) and an
    analytic sphere, under a motion top level, with 120k random rays and the oracle's answers."""
    print("\n[case] mixed motion scene", flush=True)
    mv, mi = scenes.moving_spheres(count=120, seed=5)
    mv2, mi2 = scenes.moving_spheres(count=60, seed=9, extent=10.0)
    sv_, sf = scenes.uv_sphere(3.0, 48, 24, center=(2.0, 1.0, -4.0))
    gv, gf = scenes.ground_quad(size=30.0, y=-22.0)
    g = va.Scene(0)
    o = oracle.OracleScene()
    for v, f, ref_idx in ((mv, mi, False), (sv_, sf, False), (mv2, mi2, True), (gv, gf, False)):
        g.add_geom(va.PolyMesh(v, f, motion_reference_indexing=ref_idx))
        o.add_mesh(v, f, motion_reference_indexing=ref_idx)
    g.add_sphere((-3.0, 2.0, 5.0), 2.5, 0)
    o.add_sphere((-3.0, 2.0, 5.0), 2.5)
    g.pre_render()
    o.commit()
    gs = scenes.rng(17)
    n = 120_000
    rays = scenes.make_rays(gs.uniform(-25, 25, (n, 3)), scenes.uniform_sphere_dirs(gs, n), time=gs.random(n))
    refs = {shadow: o.trace(rays, shadow=shadow, counters=True) for shadow in (False, True)}
    yield g, rays, refs
    g.close()


@pytest.mark.parametrize("cfg", [(64, 8, 32), (128, 12, 24), (256, 16, 16), (64, 32, 16), (64, 8, 20)])
def test_mixed_motion_scene_every_config_bitexact(mixed_motion, cfg):
    """Hits and per-ray counters equal the oracle's for closest and any hit in every launch shape."""
    g, rays, refs = mixed_motion
    print(f"[case] mixed motion scene, config {cfg}", flush=True)
    g.set_trace_config(*cfg)
    try:
        for shadow in (False, True):
            hits, geoms, cnt = g.trace(rays, shadow=shadow, want_counters=True)
            ref, rc = refs[shadow]
            _same(hits, geoms, ref, f"mixed motion cfg={cfg} shadow={shadow}")
            assert np.array_equal(cnt, rc)
            assert len(np.unique(geoms[geoms >= 0])) >= 4
    finally:
        g.set_trace_config(256, 16, 16)


def test_motion_times_outside_the_shutter_are_held_to_the_end_keys(mixed_motion):
    """A ray time below 0 traces as time 0 and one above 1 as time 1 (the reference would index its
    keys out of range); a NaN time traces without fault."""
    g, rays, _ = mixed_motion
    r = rays[:50_000].copy()
    out = {}
    for name, t in (("lo", -0.5), ("zero", 0.0), ("hi", 1.5), ("one", 1.0), ("nan", np.nan)):
        r["time"] = np.float32(t)
        print(f"[case] motion time {t}", flush=True)
        out[name] = g.trace(r)
    for a, b in (("lo", "zero"), ("hi", "one")):
        (ha, ga, _), (hb, gb, _) = out[a], out[b]
        assert np.array_equal(ha.view(np.uint32).reshape(len(ha), -1), hb.view(np.uint32).reshape(len(hb), -1))
        assert np.array_equal(ga, gb)
    assert (out["zero"][0]["prim"] >= 0).mean() > 0.01


def test_concurrent_host_traces_on_one_scene(sphere_ground):
    """Host-memory traces from several threads at once share the scene's staging buffer and spill
    area under its lock: every thread's result equals the oracle's for its own rays (different sizes,
    so the staging buffer is also reallocated while other threads wait)."""
    import threading
    g, o = sphere_ground
    sets = [scenes.uniform_rays(n, seed=40 + i) for i, n in enumerate((300_000, 50_000, 700_000, 120_000))]
    refs = [o.trace(r)[0] for r in sets]
    results = [None] * len(sets)
    errors = []

    def work(i):
        try:
            for _ in range(3):
                results[i] = g.trace(sets[i])
        except Exception as e:  # noqa: BLE001 (reported below)
            errors.append(e)

    print("[case] 4 host threads tracing one scene", flush=True)
    th = [threading.Thread(target=work, args=(i,)) for i in range(len(sets))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for i, (hits, geoms, _) in enumerate(results):
        _same(hits, geoms, refs[i], f"thread {i}")


def test_mesh_with_non_finite_vertex_is_rejected():
    v, f = scenes.uv_sphere(1.0, 16, 8)
    v = v.copy()
    v[5, 1] = np.nan
    g = va.Scene(0)
    with pytest.raises(va.VgpuError) as ei:
        g.add_geom(va.PolyMesh(v, f))
    assert ei.value.code == va._lib.VGPU_ERR_ARG
    g.close()


@pytest.mark.parametrize("ref_idx", [False, True])
def test_prebuilt_motion_mesh_equals_built_mesh_and_oracle(ref_idx):
    """vgpu_scene_add_motion_mesh_prebuilt fed with the oracle's BuildAccelMotion + initMotionBoxes output
    (topology, per-key boxes, accel.idx) and the unsorted idxp, as a Go host hands them over: traversal
    bit-identical to the same mesh built by the library and to the oracle, in both idxp indexings."""
    verts, idx = scenes.moving_spheres(count=300)
    o = oracle.OracleScene()
    o.add_mesh(verts, idx, motion_reference_indexing=ref_idx)
    o.commit()
    topo, boxes, perm = o.mesh_motion_accel(0)
    pre = va.Scene(0)
    pre.add_motion_prebuilt(verts, idx, perm, topo, boxes, reference_indexing=ref_idx)
    pre.pre_render()
    built = va.Scene(0)
    built.add_geom(va.PolyMesh(verts, idx, motion_reference_indexing=ref_idx))
    built.pre_render()
    gs = scenes.rng(4)
    n = 100_000
    rays = scenes.make_rays(gs.uniform(-25, 25, (n, 3)), scenes.uniform_sphere_dirs(gs, n), time=gs.random(n))
    print(f"[case] prebuilt motion mesh ref_idx={ref_idx}", flush=True)
    for shadow in (False, True):
        h1, g1, _ = pre.trace(rays, shadow=shadow)
        h2, g2, _ = built.trace(rays, shadow=shadow)
        ref, _ = o.trace(rays, shadow=shadow)
        _same(h1, g1, ref, f"prebuilt motion shadow={shadow}")
        assert h1.tobytes() == h2.tobytes() and np.array_equal(g1, g2)
    if not ref_idx:   # leaf-slot indexing tests other faces than the boxes hold: few rays hit
        assert (h1["prim"] >= 0).mean() > 0.01
    pre.close()
    built.close()


def test_prebuilt_motion_mesh_rejects_bad_accels():
    verts, idx = scenes.moving_spheres(count=4)
    o = oracle.OracleScene()
    o.add_mesh(verts, idx, motion_reference_indexing=False)
    o.commit()
    topo, boxes, perm = o.mesh_motion_accel(0)
    sc = va.Scene(0)
    with pytest.raises(va.VgpuError):            # one key, through the C ABI
        va._lib.check(sc._lib.vgpu_scene_add_motion_mesh_prebuilt(
            sc._h, va._lib.ptr(np.ascontiguousarray(verts[:1])), verts.shape[1], 1, va._lib.ptr(idx), len(idx),
            va._lib.ptr(perm), va._lib.ptr(topo), len(topo), va._lib.ptr(np.ascontiguousarray(boxes[:1])), 0.0, 0,
            None))
    bad = topo.copy()
    bad[0, 3] = len(topo)                        # child beyond the node count
    with pytest.raises(va.VgpuError):
        sc.add_motion_prebuilt(verts, idx, perm, bad, boxes)
    if len(topo) > 1:
        cyc = topo.copy()
        j = int(np.flatnonzero(cyc[0, 3:7] > 0)[0]) + 3
        child = cyc[0, j]
        k = int(np.flatnonzero(cyc[child, 3:7] >= -1)[0]) + 3 if (cyc[child, 3:7] >= -1).any() else 3
        cyc[child, k] = 0                        # points back at the root
        with pytest.raises(va.VgpuError):
            sc.add_motion_prebuilt(verts, idx, perm, cyc, boxes)
    bp = perm.copy()
    bp[0] = len(perm)
    with pytest.raises(va.VgpuError):
        sc.add_motion_prebuilt(verts, idx, bp, topo, boxes)
    sc.close()
