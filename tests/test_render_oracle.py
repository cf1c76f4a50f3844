"""Pins the renderer half of the oracle (oracle_render.cpp) with known answers.

Closed forms for the QMC sequences, structural properties of RasterXY, the
reference's own literal tables (compared as text where the reference checkout
is present; these tests are skipped elsewhere), and closed forms for the
Fresnel term, the GGX distribution, the Oren-Nayar lobe and the LookAt camera.
"""
import math
import os
import re

import numpy as np
import pytest

import oracle
from vermeer_amd import scenes

REF = "/root/reference"


# ------------------------------------------------------------ ldseq (math/ldseq)

def test_van_der_corput_closed_form():
    # radical inverse in base 2: bit-reversed index over 52 bits
    for i, want in [(0, 0.0), (1, 0.5), (2, 0.25), (3, 0.75), (4, 0.125), (5, 0.625), (6, 0.375)]:
        assert oracle.vdc(i, 0) == want
    # scrambling XORs the top bits: scramble with only bit 63 set flips the first digit
    assert oracle.vdc(1, 1 << 63) == 0.0 and oracle.vdc(0, 1 << 63) == 0.5


def test_sobol_second_dimension_closed_form():
    # direction numbers v_{k+1} = v_k ^ (v_k >> 1) from v_0 = 1/2: 1/2, 3/4, 5/8, 15/16, ...
    assert oracle.sobol(1, 0) == 0.5
    assert oracle.sobol(2, 0) == 0.75
    assert oracle.sobol(3, 0) == 0.25
    assert oracle.sobol(4, 0) == 0.625
    # first 2^k points of (VdC, Sobol) form a (0,k,2)-net: one point per elementary box
    k = 6
    pts = [(oracle.vdc(i, 0), oracle.sobol(i, 0)) for i in range(1 << k)]
    for a in range(k + 1):
        cells = {(int(x * (1 << a)), int(y * (1 << (k - a)))) for x, y in pts}
        assert len(cells) == 1 << k


def test_raster_xy_lands_in_pixel_and_stratifies():
    m = 12
    for px, py in [(0, 0), (5, 9), (511, 511), (1023, 7), (300, 1000)]:
        idx = set()
        for frame in range(16):
            i, rx, ry = oracle.raster_xy(m, frame, px, py)
            assert px <= rx < px + 1 and py <= ry < py + 1
            idx.add(i)
        assert len(idx) == 16


def _parse_go_tables(name):
    text = open(os.path.join(REF, "math/ldseq/vdc_sobol_matrices52.go")).read()
    body = text[text.index("var " + name + " = "):]
    body = body[body.index("{") + 1:]
    tables = []
    depth = 0
    cur = None
    for tok in re.findall(r"\{|\}|0x[0-9a-fA-F]+", body):
        if tok == "{":
            depth += 1
            cur = []
        elif tok == "}":
            if depth == 0:
                break
            depth -= 1
            if cur is not None:
                tables.append(cur)
                cur = None
        else:
            cur.append(int(tok, 16))
    return tables


@pytest.mark.skipif(not os.path.exists(REF), reason="reference checkout not present")
def test_raster_tables_equal_reference_literals():
    fwd_ref = _parse_go_tables("vdcSobolMatrices")
    inv_ref = _parse_go_tables("vdcSobolMatricesInv")
    assert len(fwd_ref) == 25 and len(inv_ref) >= 25   # the inverse table also lists m = 26
    for m in range(1, 26):
        fwd, inv = oracle.raster_tables(m)
        f = list(fwd_ref[m - 1])
        assert list(fwd[:len(f)]) == f, m
        iv = list(inv_ref[m - 1])
        assert list(inv[:len(iv)]) == iv, m


# ------------------------------------------------------------ colour

@pytest.mark.skipif(not os.path.exists(REF), reason="reference checkout not present")
def test_cie_table_equals_reference_literal():
    text = open(os.path.join(REF, "colour/cie1931_2deg.go")).read()
    ours = oracle.cie_tables()
    for name, tab in zip(("xBar", "yBar", "zBar"), ours):
        seg = text[text.index(name + ": []float32{") + len(name) + 13:]
        seg = seg[:seg.index("}")]
        vals = np.array([float(v) for v in re.findall(r"[-+0-9.eE]+", seg)], np.float32)
        assert len(vals) == 95
        assert np.array_equal(vals, tab), name


def test_smits_white_is_flat_and_primaries_sum():
    # the basis covers [380, 720); the hero wavelengths also reach [720, 750), where it is 0
    for lam in np.linspace(450, 719.9, 40):
        assert abs(oracle.smits(1, 1, 1, lam) - 1.0) < 1.5e-3
        assert oracle.smits(0, 0, 0, lam) == 0.0
    assert oracle.smits(1, 1, 1, 725.0) == 0.0
    # linear in a common offset: (r+a, g+a, b+a) adds a * white
    for lam in (455.0, 520.0, 610.0, 700.0):
        a = oracle.smits(0.2, 0.5, 0.3, lam)
        b = oracle.smits(0.4, 0.7, 0.5, lam)
        assert abs((b - a) - 0.2 * oracle.smits(1, 1, 1, lam)) < 1e-6


def test_spectrum_to_rgb_is_cie_sum_times_srgb():
    x, y, z = oracle.cie_tables()
    lam = 470.0
    C4 = np.array([1.0, 2.0, 0.5, 0.25], np.float32)
    waves = [(lam - 450 + j / 4 * 300) % 300 + 450 for j in range(4)]
    bins = [int((w - 360) / 470 * 95) for w in waves]
    X = sum(C4[j] * x[bins[j]] for j in range(4))
    Y = sum(C4[j] * y[bins[j]] for j in range(4))
    Z = sum(C4[j] * z[bins[j]] for j in range(4))
    M = np.array([[3.2404542, -1.5371385, -0.4985314], [-0.9692660, 1.8760108, 0.0415560],
                  [0.0556434, -0.2040259, 1.0572252]])
    want = M @ np.array([X, Y, Z], np.float64)
    got = oracle.spectrum_to_rgb(C4, lam)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6)


# ------------------------------------------------------------ BSDFs

@pytest.mark.parametrize("eta", [1.33, 1.5, 1.7])
def test_fresnel_normal_incidence(eta):
    assert abs(oracle.fresnel_dielectric(eta, 1.0) - ((eta - 1) / (eta + 1)) ** 2) < 1e-6
    assert abs(oracle.fresnel_dielectric(eta, 1e-6) - 1.0) < 1e-3      # grazing -> total reflection limit


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.6])
def test_ggx_distribution_is_normalised(alpha):
    # integral over the hemisphere of D(m) cos(theta_m) = 1 (Walter et al. 2007)
    n = 4000
    th = (np.arange(n) + 0.5) / n * (math.pi / 2)
    vals = np.array([oracle.ggx_d([math.sin(t), 0.0, math.cos(t)], alpha) for t in th], np.float64)
    integral = np.sum(vals * np.cos(th) * np.sin(th)) * (math.pi / 2 / n) * 2 * math.pi
    assert abs(integral - 1.0) < 2e-3


def test_oren_nayar_at_zero_roughness_is_lambert():
    wi = np.array([0.3, -0.2, math.sqrt(1 - 0.13)], np.float32)
    for wo in ([0.0, 0.0, 1.0], [0.6, 0.0, 0.8], [-0.48, 0.6, 0.64]):
        wo = np.array(wo, np.float32)
        rho = oracle.oren_nayar_eval(500.0, wi, wo, 0.0)
        white = [oracle.smits(1, 1, 1, (500.0 - 450 + j / 4 * 300) % 300 + 450) for j in range(4)]
        assert np.allclose(rho, np.array(white) * wo[2] / math.pi, rtol=2e-6)


# ------------------------------------------------------------ camera

def test_lookat_camera_matches_closed_form_basis():
    frm, to, up = (0.0, 1.0, 3.4), (0.0, 1.0, 0.0), (0.0, -1.0, 0.0)
    M, ttf = oracle.camera_lookat(frm, to, up, 45.0, 12.0)
    W = np.subtract(frm, to)
    W = W / np.linalg.norm(W)
    u = np.cross(up, W)
    u /= np.linalg.norm(u)
    v = np.cross(u, W)
    v /= np.linalg.norm(v)
    want = np.eye(4)
    want[0, :3], want[1, :3], want[2, :3], want[3, :3] = u, v, W, frm
    assert np.allclose(M.reshape(4, 4), want, atol=1e-6)
    assert abs(ttf - math.tan(math.radians(22.5)) * 12.0) < 1e-5


def test_oracle_render_pass_is_deterministic_and_finite():
    d = scenes.box_spheres_scene(nseg=16, nring=8)
    c = d["camera"]
    M, ttf = oracle.camera_lookat(c["from"], c["to"], c["up"], c["fov"], c["focal"])
    scr = scenes.pixel_scrambles(48, 48)
    films = []
    for threads in (1, 4):
        r = oracle.OracleRenderer(d, 48, 48, M, ttf, c["focal"])
        for it in (1, 2):
            r.render_pass(scr, it, nthreads=threads)
        films.append(r.film.copy())
    assert np.array_equal(films[0], films[1])
    assert np.all(np.isfinite(films[0])) and films[0].mean() > 0
