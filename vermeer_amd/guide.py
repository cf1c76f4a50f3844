"""Path-guiding network (no reference counterpart): a bf16 matrix-core MLP from
(hit position, normal) to a mixture of 8 von Mises-Fisher lobes, trained online
with Adam.  Thin mirror of the vgpu_guide_* entry points of include/vermeer_gpu.h.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, ptr

NPARAMS = _lib.VGPU_GUIDE_NPARAMS
# query: position xyz, pad, normal xyz, pad
QUERY_DTYPE = np.dtype([("p", "<f4", 3), ("pad0", "<f4"), ("n", "<f4", 3), ("pad1", "<f4")])
# training sample: position xyz, c, normal xyz, pad, direction xyz, pad
SAMPLE_DTYPE = np.dtype([("p", "<f4", 3), ("c", "<f4"), ("n", "<f4", 3), ("pad0", "<f4"), ("d", "<f4", 3),
                         ("pad1", "<f4")])


class Guide:
    def __init__(self, bounds, seed=1, max_batch=65536, device=0):
        lib = _lib.load()
        b = np.ascontiguousarray(bounds, np.float32).reshape(6)
        h = C.c_void_p()
        check(lib.vgpu_guide_create(C.byref(h), int(device), ptr(b), int(seed), int(max_batch)))
        self._h = h
        self._lib = lib
        self.bounds = b

    def close(self):
        if self._h:
            self._lib.vgpu_guide_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        out = np.empty(NPARAMS, np.float32)
        check(self._lib.vgpu_guide_get_params(self._h, ptr(out)))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (NPARAMS,)
        check(self._lib.vgpu_guide_set_params(self._h, ptr(p)))

    def grad(self):
        out = np.empty(NPARAMS, np.float32)
        check(self._lib.vgpu_guide_get_grad(self._h, ptr(out)))
        return out

    def infer(self, pos, nrm, raw=False):
        """Returns (n, 40): raw outputs, or 8 lobes of (mu xyz, kappa, weight) as (n, 8, 5)."""
        pos = np.asarray(pos, np.float32).reshape(-1, 3)
        q = np.zeros(len(pos), QUERY_DTYPE)
        q["p"] = pos
        q["n"] = np.asarray(nrm, np.float32).reshape(-1, 3)
        out = np.empty((len(pos), 40), np.float32)
        flags = _lib.VGPU_GUIDE_RAW if raw else 0
        check(self._lib.vgpu_guide_infer(self._h, flags, ptr(q), len(q), ptr(out), None))
        return out if raw else out.reshape(-1, 8, 5)

    def train_step(self, pos, nrm, dirs, c, lr=1e-3, update=True):
        s = np.zeros(len(c), SAMPLE_DTYPE)
        s["p"] = np.asarray(pos, np.float32).reshape(-1, 3)
        s["n"] = np.asarray(nrm, np.float32).reshape(-1, 3)
        s["d"] = np.asarray(dirs, np.float32).reshape(-1, 3)
        s["c"] = np.asarray(c, np.float32)
        loss = C.c_float()
        flags = 0 if update else _lib.VGPU_GUIDE_NO_UPDATE
        check(self._lib.vgpu_guide_train_step(self._h, flags, ptr(s), len(s), float(lr), None, C.byref(loss)))
        return float(loss.value)
