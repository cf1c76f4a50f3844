"""Pins the oracle's restatement of the renderer features beyond direct lighting of triangle lights:
conductor Fresnel (builtin/shader/fresnel/conductor.go), pixel filters (builtin/filter), Disk and
Sphere lights (builtin/light/disk.go, sphere.go), the analytic sphere geom (builtin/geom/sphere) and
perfect-mirror recursion (builtin/shader/std.go:209-261, bsdf/specular.go).

Each check is a closed form, an invariant or a brute-force comparison chosen so that a dropped term,
a wrong sign or a swapped operand fails it.
"""
import math

import numpy as np
import pytest

import oracle
from vermeer_amd import scenes


# ------------------------------------------------------------ conductor Fresnel

@pytest.mark.parametrize("r", [0.04, 0.3, 0.6, 0.9, 0.95])
@pytest.mark.parametrize("g", [0.0, 0.4, 1.0])
def test_conductor_normal_incidence_equals_reflectivity(r, g):
    # n and k^2 are chosen (Gulbrandsen 2014) so that ((n-1)^2 + k^2) / ((n+1)^2 + k^2) = r exactly
    assert oracle.fresnel_conductor(r, g, 1.0) == pytest.approx(r, rel=2e-6)


@pytest.mark.parametrize("r", [0.1, 0.5, 0.9])
def test_conductor_grazing_is_one_and_clamp(r):
    assert oracle.fresnel_conductor(r, 0.5, 0.0) == 1.0
    # reflectivities above 0.99 are clamped (math.Clamp(r, 0, 0.99))
    for c in (0.2, 0.7, 1.0):
        assert oracle.fresnel_conductor(1.0, 0.3, c) == oracle.fresnel_conductor(0.99, 0.3, c)


def test_conductor_against_float64_fresnel_equations():
    """The textbook unpolarised conductor reflectance with n(r, g), k^2(r, n) evaluated in float64."""
    for r in (0.2, 0.55, 0.9):
        for g in (0.1, 0.8):
            n = (1 - r) / (1 + r) * g + (1 - g) * (1 + math.sqrt(r)) / (1 - math.sqrt(r))
            k2 = ((n + 1) ** 2 * r - (n - 1) ** 2) / (1 - r)
            for c in np.linspace(0.05, 1.0, 9):
                rs = ((n - c) ** 2 + k2) / ((n + c) ** 2 + k2)
                rp = ((n * n + k2) * c * c - 2 * n * c + 1) / ((n * n + k2) * c * c + 2 * n * c + 1)
                assert oracle.fresnel_conductor(r, g, float(c)) == pytest.approx(0.5 * (rs + rp), rel=2e-5)


# ------------------------------------------------------------ pixel filters

def _renderer(w=8, h=8):
    d = scenes.feature_scene(nseg=8, nring=4)
    c = d["camera"]
    M, ttf = oracle.camera_lookat(c["from"], c["to"], c["up"], c["fov"], c["focal"])
    return oracle.OracleRenderer(d, w, h, M, ttf, c["focal"]), d


def test_airy_filter_warps_every_sample_to_the_pixel_centre():
    """AiryFilter's table holds 2 J1(v)/v at v = 0 = 0/0 (the grid passes exactly through the centre):
    the normalisation makes every CDF entry NaN, no comparison succeeds and WarpSample returns (0, 0)."""
    r, _ = _renderer()
    r.set_filter("airy")
    cdfV, cdfVU = r.filter_tables()
    assert np.isnan(cdfV).all() and np.isnan(cdfVU).all()
    for r0 in (0.0, 0.1, 0.5, 0.999):
        for r1 in (0.0, 0.3, 0.77):
            assert r.warp_sample(r0, r1) == (0.0, 0.0)


def test_gaussian_filter_tables_and_warp_statistics():
    r, _ = _renderer()
    r.set_filter("gaussian")
    cdfV, cdfVU = r.filter_tables()
    n, w = 17, 2.0
    assert len(cdfV) == n and np.all(np.diff(cdfV) >= 0) and abs(cdfV[-1] - 1) < 1e-12
    assert np.all(np.abs(cdfVU[:, -1] - 1) < 1e-12)
    # the tabulated filter, independently: sigma = 1/sqrt(W), cut off at radius W/2, on the n x n grid
    g = np.linspace(-w / 2, w / 2, n)
    U, V = np.meshgrid(g, g, indexing="ij")
    s2 = 1.0 / w
    f = np.exp(-(U * U + V * V) / 2 * s2) * (np.sqrt(U * U + V * V) <= w / 2 + 1e-12)
    pv = f.sum(axis=1) / f.sum()
    assert np.allclose(cdfV, np.cumsum(pv), atol=1e-9)
    # i == 0 branch of WarpSample: u = -w/2 + (r0 / cdfV[0]) / w (filter.go:49-52)
    r0 = cdfV[0] / 2
    u, _ = r.warp_sample(r0, 0.5)
    assert u == pytest.approx(-w / 2 + 0.5 / w, abs=1e-12)
    # warped samples of a stratified grid
    us, vs = [], []
    for a in (np.arange(64) + 0.5) / 64:
        for b in (np.arange(64) + 0.5) / 64:
            x, y = r.warp_sample(a, b)
            us.append(x)
            vs.append(y)
    us, vs = np.array(us), np.array(vs)
    # WarpSample spreads cell i's mass uniformly over [g_i, g_i + h] (h = w/(n-1)), except cell 0, which
    # lands on [-w/2, -w/2 + 1/w): the warped distribution is that piecewise-uniform density, whose mean is
    # about h/2 (the reference's filtered samples sit half a cell off centre) and whose moments follow
    hcell = w / (n - 1)
    for x in (us, vs):
        assert x.min() >= -w / 2 and x.max() <= w / 2 + hcell
    lo0, hi0 = -w / 2, -w / 2 + 1 / w
    mean = pv[0] * (lo0 + hi0) / 2 + float((pv[1:] * (g[1:] + hcell / 2)).sum())
    m2 = pv[0] * (lo0 * lo0 + lo0 * hi0 + hi0 * hi0) / 3 + float((pv[1:] * (g[1:] ** 2 + g[1:] * hcell + hcell ** 2 / 3)).sum())
    assert abs(us.mean() - mean) < 2e-3 and abs(vs.mean() - mean) < 2e-3
    assert abs(mean - hcell / 2) < 0.01
    assert float((us * us).mean()) == pytest.approx(m2, rel=0.02)
    assert float((vs * vs).mean()) == pytest.approx(m2, rel=0.02)


def test_box_filter_is_identity():
    r, _ = _renderer()
    assert r.filter_tables() is None
    assert r.warp_sample(0.3, 0.6) == (0.0, 0.0)


# ------------------------------------------------------------ disk light mesh

def test_disk_light_mesh_lies_on_the_disk_rim_and_faces_away_from_lookat():
    P, up, look, R = (0.2, 1.9, -0.1), (1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.3
    v, idx, nrm = scenes.disk_light_mesh(P, up, look, R, 20)
    assert v.shape == (60, 3) and idx.shape == (20, 3)
    N = np.asarray(look) - np.asarray(P)
    N = N / np.linalg.norm(N)
    P = np.asarray(P)
    rim = v.reshape(20, 3, 3)[:, 1:, :].reshape(-1, 3).astype(np.float64)
    assert np.allclose(np.linalg.norm(rim - P, axis=1), R, rtol=1e-5)
    assert np.all(np.abs((rim - P) @ N) < 1e-5)
    assert np.allclose(v.reshape(20, 3, 3)[:, 0, :], P.astype(np.float32))
    # consecutive triangles share their rim vertices up to float32 angle accumulation
    t = v.reshape(20, 3, 3).astype(np.float64)
    assert np.allclose(t[:-1, 2], t[1:, 1], atol=1e-6)
    # fan winding: (v1 - P) x (v2 - P) = sin(dv) (B x T) = -sin(dv) N
    e = np.cross(t[:, 1] - P, t[:, 2] - P)
    assert np.all(e @ N < 0)
    assert np.allclose(nrm, N.astype(np.float32), atol=1e-6)


# ------------------------------------------------------------ analytic sphere geom

def test_oracle_sphere_geom_hits_at_the_closed_form_distance():
    sc = oracle.OracleScene()
    gv, gf = scenes.ground_quad()
    sc.add_mesh(gv, gf)
    C, R = np.array([0.3, 1.5, -0.2]), 0.4
    sc.add_sphere(C, R, 0)
    sc.commit()
    slot_of = {int(ins): s for s, ins in enumerate(sc.info()["geom_order"])}
    rng = np.random.default_rng(3)
    d = rng.normal(size=(400, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    O = C + 3.0 * d                                  # outside, aimed at the centre
    rays = scenes.make_rays(O, -d)
    hits, _ = sc.trace(rays)
    sphere = hits["geom"] == slot_of[1]
    # rays from below the ground plane hit the ground first; all others hit the sphere at 3 - R
    assert sphere.sum() > 100
    assert np.all(hits["prim"][sphere] == 0) and np.all(hits["u"][sphere] == 0)
    assert np.allclose(hits["t"][sphere], 3.0 - R, rtol=1e-5)
    # from inside, the far side at R
    rays_in = scenes.make_rays(np.repeat(C[None], 50, 0), d[:50])
    h_in, _ = sc.trace(rays_in)
    upward = d[:50, 1] > -0.2
    assert np.allclose(h_in["t"][upward], R, rtol=1e-5)
    # tmax short of the surface: no hit; shadow rays through it: occluded
    short = scenes.make_rays(O[:20], -d[:20], tmax=2.0)
    hs, _ = sc.trace(short)
    assert np.all(hs["geom"][:20] != slot_of[1])
    sh, _ = sc.trace(rays[:50], shadow=True)
    assert np.all(sh["geom"][sphere[:50]] >= 0)


# ------------------------------------------------------------ rendering with the features

def _render(d, w, h, passes, filt="box", scr_seed=11):
    c = d["camera"]
    M, ttf = oracle.camera_lookat(c["from"], c["to"], c["up"], c["fov"], c["focal"])
    r = oracle.OracleRenderer(d, w, h, M, ttf, c["focal"])
    r.set_filter(filt)
    scr = scenes.pixel_scrambles(w, h, seed=scr_seed)
    for it in range(1, passes + 1):
        r.render_pass(scr, it)
    return r.film


def test_feature_scene_renders_finite_and_lit():
    d = scenes.feature_scene(nseg=16, nring=8)
    film = _render(d, 48, 48, 2)
    assert np.all(np.isfinite(film)) and film.mean() > 0.02


def test_mirror_recursion_shows_the_scene_in_the_mirrors():
    """With perfect mirrors the mirror spheres' pixels carry light from other surfaces (level-1 and deeper
    hits); with roughness 0.05 instead they show only direct highlights.  The mirror sphere pixels differ."""
    w = h = 40
    with_m = _render(scenes.feature_scene(nseg=16, nring=8, mirrors=True), w, h, 1)
    without = _render(scenes.feature_scene(nseg=16, nring=8, mirrors=False), w, h, 1)
    diff = np.abs(with_m - without).sum(axis=2)
    assert (diff > 1e-3).mean() > 0.02


def test_mirror_of_an_emitter_reflects_it_with_the_fresnel_weight():
    """A camera ray that hits a pure mirror (dielectric, IOR 1.5, SpecStrength 1) and then an emitter of
    RGB radiance L returns ToRGB(FromRGB(F, F, F) |o.z|) * L (std.go:226-258): seen head-on, cos ~ 1, so
    F = ((1.5 - 1) / 2.5)^2 = 0.04 and |o.z| ~ 1, at the pass's hero wavelength."""
    L = 10.0
    mats = np.array([scenes.material(spec=(1.0, 1.0, 1.0), spec_strength=1.0, spec_roughness=0.0, ior=1.5),
                     scenes.material(emission=(1.0, 1.0, 1.0), emission_strength=L, diffuse=(0.0, 0.0, 0.0),
                                     diffuse_strength=1.0)])
    mirror = {"verts": np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32),
              "idx": np.array([[0, 1, 2], [0, 2, 3]], np.uint32), "shaderidx": None,
              "shaders": np.array([0], np.int32), "normals": None}
    # the emitter is a big quad behind the camera, facing the mirror
    emit = {"verts": np.array([[-5, -5, 6], [5, -5, 6], [5, 5, 6], [-5, 5, 6]], np.float32),
            "idx": np.array([[0, 2, 1], [0, 3, 2]], np.uint32), "shaderidx": None,
            "shaders": np.array([1], np.int32), "normals": None}
    d = {"meshes": [mirror, emit], "materials": mats, "lights": [],
         "camera": {"from": (0.0, 0.0, 3.0), "to": (0.0, 0.0, 0.0), "up": (0.0, -1.0, 0.0), "fov": 20.0,
                    "focal": 12.0}}
    film = _render(d, 9, 9, 1, scr_seed=11)
    scr = scenes.pixel_scrambles(9, 9, seed=11)
    lam = (720 - 450) * oracle.vdc(1, int(scr[4, 4, 3])) + 450
    F = 0.04
    C4 = np.array([oracle.smits(F, F, F, lam + 75.0 * j if lam + 75.0 * j < 750 else lam + 75.0 * j - 300)
                   for j in range(4)], np.float32)
    # negative components are clamped to 0 (std.go:243-247); F and |o.z| are ~0.04 and ~1 off-axis
    want = np.maximum(np.asarray(oracle.spectrum_to_rgb(C4, np.float32(lam))) * L, 0)
    assert np.allclose(film[4, 4] * 2, want, rtol=2e-3, atol=1e-4)   # film = (0 * 1 + c) / 2 after pass 1
    # no light sources: everything else in the image is the mirror's reflection or the emitter
    assert np.all(np.isfinite(film))


# ------------------------------------------------------------ thin-lens camera

def test_uniform_disk_is_the_concentric_map():
    """Shirley-Chiu concentric square-to-disk map (sample.go:105-129): points inside the radius, the
    square's edges on the rim, equal areas (r^2 uniform, angle uniform), and the degenerate diagonals
    x = +-y, which no branch takes, at the centre."""
    R = 0.7
    pts = np.array([oracle.uniform_disk(R, a, b) for a in (np.arange(96) + 0.5) / 96 for b in (np.arange(96) + 0.5) / 96])
    rr = np.hypot(pts[:, 0], pts[:, 1])
    assert rr.max() <= R * (1 + 1e-6)
    assert abs(np.mean((rr / R) ** 2) - 0.5) < 0.01                  # uniform over the area
    ang = np.arctan2(pts[:, 1], pts[:, 0])
    h, _ = np.histogram(ang, bins=8, range=(-np.pi, np.pi))
    assert h.min() > 0.95 * h.mean()
    x, y = oracle.uniform_disk(R, 1.0, 0.5)                            # right edge midpoint -> (R, 0)
    assert abs(x - R) < 1e-6 and abs(y) < 1e-6
    assert oracle.uniform_disk(R, 0.75, 0.75) == (0.0, 0.0)           # x == y
    assert oracle.uniform_disk(R, 0.25, 0.75) == (0.0, 0.0)           # x == -y


def test_thin_lens_converges_and_reduces_to_pinhole():
    d = scenes.feature_scene(nseg=16, nring=8)
    c = d["camera"]
    M, ttf = oracle.camera_lookat(c["from"], c["to"], c["up"], c["fov"], c["focal"])
    scr = scenes.pixel_scrambles(32, 32, seed=4)
    films = {}
    for radius in (0.0, 0.08):
        r = oracle.OracleRenderer(d, 32, 32, M, ttf, c["focal"])
        r.set_lens(radius)
        for it in (1, 2, 3):
            r.render_pass(scr, it)
        films[radius] = r.film.copy()
    pin = oracle.OracleRenderer(d, 32, 32, M, ttf, c["focal"])
    for it in (1, 2, 3):
        pin.render_pass(scr, it)
    assert np.array_equal(films[0.0], pin.film)                        # radius 0 is the pinhole path
    assert not np.array_equal(films[0.08], pin.film)
    assert abs(films[0.08].mean() - pin.film.mean()) < 0.25 * pin.film.mean()
