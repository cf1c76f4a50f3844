"""Output drivers: the film as a Radiance .hdr file or as raw float32, mirroring the reference's
OutputHDR and OutputFloat nodes (builtin/driver/outputhdr.go, outputfloat.go).  The RGBE encoding
runs on the GPU (vgpu_film_to_rgbe); these functions only marshal buffers across the C ABI.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, ptr


def _film(film):
    f = np.ascontiguousarray(film, np.float32)
    if f.ndim != 3 or f.shape[2] != 3:
        raise ValueError("film must be (height, width, 3) float32")
    return f


def film_to_rgbe(film, device=0):
    """(h, w, 3) host film -> (h, w, 4) uint8 RGBE in the file's scanline order (bottom row first)."""
    f = _film(film)
    h, w = f.shape[:2]
    out = np.zeros((h, w, 4), np.uint8)
    check(_lib.load().vgpu_film_to_rgbe(int(device), 0, ptr(f), w, h, ptr(out), None))
    return out


def write_hdr(film, path, device=0):
    """OutputHDR.PostRender: a host film written as the reference writes it."""
    f = _film(film)
    h, w = f.shape[:2]
    check(_lib.load().vgpu_write_hdr(int(device), 0, ptr(f), w, h, str(path).encode(), None))


def write_hdr_device(d_film, width, height, path, device=0, stream=0):
    """OutputHDR.PostRender for a device-resident film (an int data pointer)."""
    check(_lib.load().vgpu_write_hdr(int(device), _lib.VGPU_TRACE_DEVICE_PTRS, C.c_void_p(d_film), int(width),
                                     int(height), str(path).encode(), C.c_void_p(stream or None)))


def write_float(film, path, device=0):
    """OutputFloat.PostRender: the film's float32 values, little-endian."""
    f = _film(film)
    h, w = f.shape[:2]
    check(_lib.load().vgpu_write_float(int(device), 0, ptr(f), w, h, str(path).encode(), None))


def write_float_device(d_film, width, height, path, device=0, stream=0):
    check(_lib.load().vgpu_write_float(int(device), _lib.VGPU_TRACE_DEVICE_PTRS, C.c_void_p(d_film), int(width),
                                       int(height), str(path).encode(), C.c_void_p(stream or None)))
