"""Scene files (.vnf): load, build and render them as the reference's `vermeer scene.vnf` does
(cmd/vermeer/main.go: nodes.Parse, core.PreRender, core.Render, core.PostRender).

    python -m vermeer_amd.vnf scene.vnf [--maxiter N] [--device 0] [--seed 1]

Parsing and building happen in libvermeer_gpu.so (vgpu_vnf_*); this module marshals the calls and
runs the pass loop.  The per-pixel scrambles come from a seeded generator: the reference draws them
from Go's math/rand global source, whose sequence depends on the Go version.
"""
import argparse
import ctypes as C
import sys

import numpy as np

from . import _lib
from ._lib import check


class VnfFile:
    def __init__(self, path):
        self._lib = _lib.load()
        h = C.c_void_p()
        check(self._lib.vgpu_vnf_load(str(path).encode(), C.byref(h)))
        self._h = h
        self.path = str(path)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vgpu_vnf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        w, h, mi, ne = C.c_uint32(), C.c_uint32(), C.c_int32(), C.c_int32()
        check(self._lib.vgpu_vnf_info(self._h, C.byref(w), C.byref(h), C.byref(mi), C.byref(ne)))
        return {"width": w.value, "height": h.value, "max_iter": mi.value, "errors": ne.value}

    def _text(self, what):
        need = C.c_uint64()
        check(self._lib.vgpu_vnf_text(self._h, what, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        check(self._lib.vgpu_vnf_text(self._h, what, buf, need.value, None))
        return buf.value.decode()

    def messages(self):
        return self._text(_lib.VGPU_VNF_MESSAGES)

    def nodes_text(self):
        return self._text(_lib.VGPU_VNF_NODES)

    def build(self, device=0):
        """core.PreRender: returns a render.Renderer over the built scene and renderer."""
        from .accel import Scene
        from .render import Renderer
        s, r = C.c_void_p(), C.c_void_p()
        check(self._lib.vgpu_vnf_build(self._h, int(device), C.byref(s), C.byref(r)))
        inf = self.info()
        return Renderer.adopt(Scene.adopt(s, device), r, inf["width"], inf["height"])

    def outputs(self):
        n = C.c_int32()
        check(self._lib.vgpu_vnf_output(self._h, -1, None, None, 0, C.byref(n)))
        out = []
        for i in range(n.value):
            t = C.c_int32()
            buf = C.create_string_buffer(4096)
            check(self._lib.vgpu_vnf_output(self._h, i, C.byref(t), buf, 4096, None))
            out.append(("hdr" if t.value == _lib.VGPU_OUTPUT_HDR else "float", buf.value.decode()))
        return out


def render_file(path, maxiter=None, device=0, seed=1, log=print):
    """Loads, builds and renders a .vnf file; writes its output nodes; returns the film."""
    from . import output, scenes
    f = VnfFile(path)
    inf = f.info()
    if inf["errors"]:
        log(f.messages(), end="")
    R = f.build(device)
    # core.Render (render.go:141-146, 184): maxiter > -1 overrides Globals.MaxIter; 0 renders until
    # interrupted (Ctrl-C finishes the current pass, then the outputs are written)
    n = inf["max_iter"] if maxiter is None or maxiter < 0 else maxiter
    scr = scenes.pixel_scrambles(inf["width"], inf["height"], seed=seed)
    it = 0
    try:
        while n == 0 or it < n:
            it += 1
            R.render_pass(scr, it)
    except KeyboardInterrupt:
        log(f"interrupted after pass {it}")
    for kind, name in f.outputs():
        (output.write_hdr if kind == "hdr" else output.write_float)(R.film, name, device=device)
        log(f"wrote {name}")
    film = R.film.copy()
    R.close()
    f.close()
    return film


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("file")
    ap.add_argument("--maxiter", type=int, default=-1,
                    help="passes; -1: the file's Globals MaxIter; 0: until interrupted")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--seed", type=int, default=1, help="seed of the per-pixel scrambles")
    a = ap.parse_args(argv)
    render_file(a.file, a.maxiter, a.device, a.seed)


if __name__ == "__main__":
    sys.exit(main())
