"""The oracle's motion rendering (reference semantics), pinned on the CPU.

Each sample's time is VanDerCorput(iter, time scramble) (core/render.go:93,118); camera, mirror and
shadow rays carry it (core/ray.go:65, core/trace.go:53), and a moving mesh is traced and shaded at it
(builtin/geom/polymesh/trace.go:520-719).  Pins:
- a quad that slides across the view is seen by a centre pixel exactly while the sample time lies in
  the window the geometry fixes (time source, key order, vertex lerp);
- a moving mesh whose keys are all equal shades like the same static mesh: P, Ng, N, the offset and the
  shader index agree, only dPdu / dPdv differ (raw edges), which rotates the shading frame about N and
  changes the result by rounding when every light takes one sample;
- normals with fewer keys than the vertices are ignored (N = Ng, trace.go:694).
"""
import numpy as np

import oracle
from vermeer_amd import scenes


def _render(d, w, h, passes, scr=None, want_sample=False):
    c = d["camera"]
    M, ttf = oracle.camera_lookat(c["from"], c["to"], c["up"], c["fov"], c["focal"])
    r = oracle.OracleRenderer(d, w, h, M, ttf, c["focal"])
    if scr is None:
        scr = scenes.pixel_scrambles(w, h, seed=11)
    samples = []
    for it in range(1, passes + 1):
        _, s = r.render_pass(scr, it, want_sample=want_sample)
        samples.append(None if s is None else s.copy())
    return r.film, samples


def _sliding_quad_scene():
    quad_v = np.array([[-3, -5, 0], [-1, -5, 0], [-1, 5, 0], [-3, 5, 0]], np.float64)
    keys = np.stack([quad_v, quad_v + np.array([4.0, 0.0, 0.0])]).astype(np.float32)
    mesh = {"keys": 2, "verts": keys, "idx": np.array([[0, 1, 2], [0, 2, 3]], np.uint32), "shaderidx": None,
            "shaders": np.array([0], np.int32), "normals": None}
    mats = np.array([scenes.material(emission=(1.0, 1.0, 1.0), emission_strength=1.0, diffuse=(0.0, 0.0, 0.0),
                                     diffuse_strength=1.0)], scenes.MATERIAL_DTYPE)
    camera = {"from": (0.0, 0.0, 5.0), "to": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0), "fov": 45.0, "focal": 12.0}
    return {"meshes": [mesh], "materials": mats, "lights": [], "camera": camera}


def test_sliding_quad_is_seen_while_the_sample_time_is_in_its_window():
    """The quad spans x in [-3 + 4t, -1 + 4t] on z = 0.  The two centre columns of a 32-pixel image see
    x within +-0.13 there, so they hit it for t in [0.25 +- 0.033, 0.75 +- 0.033] and miss outside."""
    w = h = 32
    d = _sliding_quad_scene()
    scr = scenes.pixel_scrambles(w, h, seed=5)
    _, samples = _render(d, w, h, 6, scr=scr, want_sample=True)
    checked = 0
    for it, s in enumerate(samples, start=1):
        for y in range(h):
            for x in (15, 16):
                t = oracle.vdc(it, int(scr[y, x, 2]))
                lit = bool(s[y, x].sum() > 0)
                if 0.3 < t < 0.7:
                    assert lit, (it, x, y, t)
                    checked += 1
                elif t < 0.2 or t > 0.8:
                    assert not lit, (it, x, y, t)
                    checked += 1
    assert checked > 200
    # the lambda scramble (index 3) would not predict the hits: the time index is pinned
    wrong = 0
    for it, s in enumerate(samples, start=1):
        for y in range(h):
            t = oracle.vdc(it, int(scr[y, 15, 3]))
            if (0.3 < t < 0.7) != bool(s[y, 15].sum() > 0) and (t < 0.2 or t > 0.8 or 0.3 < t < 0.7):
                wrong += 1
    assert wrong > 10


def _spheres_only(d, static_keys):
    """The moving scene without the blob; with static_keys, every key of the spheres is key 0."""
    box, sph, _blob, disk = d["meshes"]
    sph = dict(sph)
    if static_keys:
        sph["verts"] = np.repeat(sph["verts"][:1], sph["keys"], axis=0)
        sph["normals"] = np.repeat(sph["normals"][:1], sph["keys"], axis=0)
    lights = [dict(L, geom=2, samples=0) for L in d["lights"]]
    return {"meshes": [box, sph, disk], "materials": d["materials"], "lights": lights, "camera": d["camera"]}


def test_moving_mesh_with_equal_keys_shades_like_the_static_mesh():
    w = h = 40
    d = scenes.moving_scene(nseg=24, nring=12)
    moving = _spheres_only(d, static_keys=True)
    sph = moving["meshes"][1]
    ix = sph["idx"].astype(np.int64)
    nrm_v = np.zeros((sph["verts"].shape[1], 3), np.float32)
    nrm_v[ix.reshape(-1)] = sph["normals"][0]    # per-corner normals of key 0 are per-vertex normals here
    static = dict(moving)
    static["meshes"] = [moving["meshes"][0],
                        {"verts": sph["verts"][0], "idx": sph["idx"], "shaderidx": sph["shaderidx"],
                         "shaders": sph["shaders"], "normals": nrm_v},
                        moving["meshes"][2]]
    fm, _ = _render(moving, w, h, 3)
    fs, _ = _render(static, w, h, 3)
    assert np.all(np.isfinite(fm)) and fm.mean() > 0.01
    rel = np.abs(fm - fs).sum(axis=2) / np.maximum(np.abs(fs).sum(axis=2), 1e-3)
    assert (rel > 1e-4).mean() < 0.005, (rel > 1e-4).mean()
    assert np.median(rel) < 1e-6
    # a moving mesh whose keys differ does not render as the static one
    fk, _ = _render(_spheres_only(d, static_keys=False), w, h, 3)
    assert (np.abs(fk - fs).sum(axis=2) / np.maximum(np.abs(fs).sum(axis=2), 1e-3) > 1e-2).mean() > 0.02


def test_normals_with_other_key_count_are_ignored():
    w = h = 32
    d = scenes.moving_scene(nseg=16, nring=8)
    blob = d["meshes"][2]
    assert blob["normals"].shape[0] == 1 and blob["keys"] > 1
    no_nrm = dict(d)
    no_nrm["meshes"] = [d["meshes"][0], d["meshes"][1], dict(blob, normals=None), d["meshes"][3]]
    fa, _ = _render(d, w, h, 2)
    fb, _ = _render(no_nrm, w, h, 2)
    assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))
    # while normals with every key change the shading of a smooth sphere
    sph = d["meshes"][1]
    flat = dict(d)
    flat["meshes"] = [d["meshes"][0], dict(sph, normals=None), blob, d["meshes"][3]]
    fc, _ = _render(flat, w, h, 2)
    assert not np.array_equal(fa, fc)


def test_moving_scene_renders_finite_and_lit():
    film, _ = _render(scenes.moving_scene(nseg=16, nring=8), 40, 40, 2)
    assert np.all(np.isfinite(film)) and film.mean() > 0.02


# ------------------------------------------------------------ camera motion keys

def _rot_y(theta):
    c, s = np.cos(theta), np.sin(theta)
    m = np.eye(4, dtype=np.float64)
    m[0, 0], m[0, 2], m[2, 0], m[2, 2] = c, s, -s, c
    return m


def _col_major(m):
    return np.asarray(m, np.float32).T.reshape(16)


def test_decomposition_recomposes_the_key_and_slerps_rotations():
    A = _rot_y(0.3)
    A[:3, 3] = (1.0, 2.0, -3.0)
    A[:3, :3] = A[:3, :3] @ np.diag([2.0, 1.0, 0.5])     # rotation times a scale
    a = _col_major(A)
    np.testing.assert_allclose(oracle.decomp_lerp_matrix(a, a, 0.0), a, atol=2e-6)
    # two rigid keys: rotation angle and translation interpolate linearly
    B0, B1 = _rot_y(0.2), _rot_y(1.0)
    B0[:3, 3], B1[:3, 3] = (0.0, 0.0, 0.0), (4.0, -2.0, 1.0)
    for t in (0.25, 0.5, 0.8):
        M = oracle.decomp_lerp_matrix(_col_major(B0), _col_major(B1), t).reshape(4, 4).T
        E = _rot_y(0.2 + 0.8 * t)
        E[:3, 3] = np.array([4.0, -2.0, 1.0]) * t
        np.testing.assert_allclose(M, E, atol=3e-6)


def test_lookat_keys_build_the_basis_per_key():
    frm = [(0.0, 1.0, 3.4), (0.5, 1.1, 3.3), (1.0, 1.2, 3.0)]
    to = [(0.0, 0.8, 0.0)]
    mats = oracle.camera_lookat_keys(frm, to, (0.0, -1.0, 0.0))
    assert mats.shape == (3, 16)
    for k, f in enumerate(frm):
        m = mats[k].reshape(4, 4).T
        np.testing.assert_allclose(m[:3, 3], f, atol=0)
        w = np.asarray(f) - np.asarray(to[0])
        np.testing.assert_allclose(m[:3, 2], w / np.linalg.norm(w), atol=1e-6)
        assert abs(np.linalg.det(m[:3, :3]) + 1.0) < 1e-5      # U x V = -W: a LookAt basis has det -1
    # one key recomposed is the static LookAt matrix bit for bit
    M1, _ = oracle.camera_lookat(frm[0], to[0], (0.0, -1.0, 0.0), 45.0, 12.0)
    assert np.array_equal(oracle.decomp_lerp_matrix(mats[0], mats[0], 0.0).view(np.uint32),
                          np.asarray(M1, np.float32).reshape(16).view(np.uint32))
    # roll turns U towards V
    rolled = oracle.camera_lookat_keys(frm[:1], to, (0.0, -1.0, 0.0), roll=[0.5])
    m0, m1 = mats[0].reshape(4, 4).T, rolled[0].reshape(4, 4).T
    np.testing.assert_allclose(m1[:3, 0], np.cos(0.5) * m0[:3, 0] + np.sin(0.5) * m0[:3, 1], atol=1e-6)


def test_panning_camera_sees_a_quad_while_the_time_is_in_its_window():
    """The camera translates from x = -2 to x = 2 over the shutter, looking down -z at a static emissive
    quad spanning x in [-1, 1]: the centre columns see it for t in [0.25, 0.75] (+- 0.033)."""
    w = h = 32
    quad_v = np.array([[-1, -5, 0], [1, -5, 0], [1, 5, 0], [-1, 5, 0]], np.float32)
    mesh = {"verts": quad_v, "idx": np.array([[0, 1, 2], [0, 2, 3]], np.uint32), "shaderidx": None,
            "shaders": np.array([0], np.int32), "normals": None}
    mats = np.array([scenes.material(emission=(1.0, 1.0, 1.0), emission_strength=1.0, diffuse=(0.0, 0.0, 0.0),
                                     diffuse_strength=1.0)], scenes.MATERIAL_DTYPE)
    d = {"meshes": [mesh], "materials": mats, "lights": [],
         "camera": {"from": (0.0, 0.0, 5.0), "to": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0), "fov": 45.0,
                    "focal": 12.0}}
    # LocalToWorld keys of a "matrix" camera: no rotation (it looks down -z), translated along x.  (A keyed
    # LookAt would not do: calcLookatMatrices samples the other parameter at time i / nkeys.)
    keys = np.zeros((2, 16), np.float32)
    for k, x0 in enumerate((-2.0, 2.0)):
        m = np.eye(4)
        m[:3, 3] = (x0, 0.0, 5.0)
        keys[k] = _col_major(m)
    c = d["camera"]
    M, ttf = oracle.camera_lookat(c["from"], c["to"], c["up"], c["fov"], c["focal"])
    r = oracle.OracleRenderer(d, w, h, M, ttf, c["focal"])
    r.set_camera_keys(keys)
    scr = scenes.pixel_scrambles(w, h, seed=6)
    checked = 0
    for it in range(1, 7):
        _, s = r.render_pass(scr, it, want_sample=True)
        for y in range(h):
            for x in (15, 16):
                t = oracle.vdc(it, int(scr[y, x, 2]))
                lit = bool(s[y, x].sum() > 0)
                if 0.3 < t < 0.7:
                    assert lit, (it, x, y, t)
                    checked += 1
                elif t < 0.2 or t > 0.8:
                    assert not lit, (it, x, y, t)
                    checked += 1
    assert checked > 200


def test_equal_camera_keys_render_like_the_static_camera():
    w = h = 40
    d = scenes.feature_scene(nseg=16, nring=8)
    c = d["camera"]
    M, ttf = oracle.camera_lookat(c["from"], c["to"], c["up"], c["fov"], c["focal"])
    key = oracle.camera_lookat_keys([c["from"]], [c["to"]], c["up"])
    scr = scenes.pixel_scrambles(w, h, seed=4)
    films = []
    for keys in (None, np.repeat(key, 3, axis=0)):
        r = oracle.OracleRenderer(d, w, h, M, ttf, c["focal"])
        if keys is not None:
            r.set_camera_keys(keys)
        for it in (1, 2):
            r.render_pass(scr, it)
        films.append(r.film.copy())
    rel = np.abs(films[0] - films[1]).sum(axis=2) / np.maximum(np.abs(films[0]).sum(axis=2), 1e-3)
    assert (rel > 1e-4).mean() < 0.01 and np.median(rel) < 1e-6
