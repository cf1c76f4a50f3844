"""GPU output drivers against the oracle, byte for byte: the .hdr file (RGBE encoded on the GPU) and
the raw float file, from host and device films, on a rendered film and on a film of special values
(NaN, infinities, negatives, values at the 1e-6 threshold, the largest and smallest floats)."""
import numpy as np
import pytest
import torch

import oracle
from vermeer_amd import output, render, scenes

pytestmark = pytest.mark.gpu


def _special_film(h=37, w=53, seed=3):
    g = np.random.default_rng(seed)
    film = (np.exp(g.uniform(-30, 30, (h, w, 3))) * g.choice([-1, 1, 1, 1], (h, w, 3))).astype(np.float32)
    specials = np.array([np.nan, np.inf, -np.inf, 0.0, -0.0, 1e-6, 9.99e-7, 1.0000001e-6, 3.4e38, 1.2e-38,
                         1.0, 0.5, 255.999, 1e30], np.float32)
    mask = g.random((h, w, 3)) < 0.3
    film[mask] = g.choice(specials, mask.sum())
    return film


@pytest.fixture(scope="module")
def rendered_film():
    desc = scenes.box_spheres_scene(nseg=32, nring=16)
    c = desc["camera"]
    W = 96
    cam = render.lookat_camera(c["from"], c["to"], c["up"], c["fov"], c["focal"], W, W)
    R = render.Renderer(desc, W, W, cam)
    scr = scenes.pixel_scrambles(W, W, seed=2)
    for it in range(1, 5):
        R.render_pass(scr, it)
    film = R.film.copy()
    R.close()
    return film


@pytest.mark.parametrize("which", ["rendered", "special"])
def test_hdr_and_float_files_equal_oracle(tmp_path, rendered_film, which):
    film = rendered_film if which == "rendered" else _special_film()
    print(f"[case] output drivers, {which} film {film.shape}", flush=True)
    h, w = film.shape[:2]
    assert np.array_equal(output.film_to_rgbe(film), oracle.film_to_rgbe(film))
    oracle.write_hdr(film, tmp_path / "o.hdr")
    oracle.write_float(film, tmp_path / "o.float")
    output.write_hdr(film, tmp_path / "g.hdr")
    output.write_float(film, tmp_path / "g.float")
    d_film = torch.from_numpy(film.reshape(-1).copy()).cuda()
    st = torch.cuda.current_stream().cuda_stream
    output.write_hdr_device(d_film.data_ptr(), w, h, tmp_path / "gd.hdr", stream=st)
    output.write_float_device(d_film.data_ptr(), w, h, tmp_path / "gd.float", stream=st)
    ref_hdr = (tmp_path / "o.hdr").read_bytes()
    ref_float = (tmp_path / "o.float").read_bytes()
    assert (tmp_path / "g.hdr").read_bytes() == ref_hdr
    assert (tmp_path / "gd.hdr").read_bytes() == ref_hdr
    assert (tmp_path / "g.float").read_bytes() == ref_float
    assert (tmp_path / "gd.float").read_bytes() == ref_float
