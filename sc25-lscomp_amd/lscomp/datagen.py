"""Seeded synthetic detector volumes (liblsgen.so, built from csrc/lsgen.c).

Kinds: SPARSE (Poisson 0.05 + 0.1% hot pixels), CSSI (Poisson 0.3 background,
50 Gaussian peaks per frame, gap stripes), XPCS (Exp x Gamma speckle rates,
Poisson counts).  Values are a pure function of (seed, frame, pixel).
"""
from __future__ import annotations

import ctypes
import hashlib
import os

import numpy as np

SPARSE, CSSI, XPCS = 1, 2, 3
KINDS = {"sparse": SPARSE, "cssi": CSSI, "xpcs": XPCS}

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblsgen.so")
_lib = None


def _get():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise ImportError(f"{_LIB_PATH} is not built; run make -C sc25-lscomp_amd")
        lib = ctypes.CDLL(_LIB_PATH)
        u16p = ctypes.POINTER(ctypes.c_uint16)
        lib.lsgen_volume.restype = None
        lib.lsgen_volume.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                     ctypes.c_uint32, ctypes.c_uint32, u16p]
        lib.lsgen_frame.restype = None
        lib.lsgen_frame.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                    ctypes.c_uint32, u16p]
        _lib = lib
    return _lib


def volume(kind, seed: int, dims, first_frame: int = 0, out: np.ndarray | None = None) -> np.ndarray:
    """(x, y, z) uint16 volume; frame f has global index first_frame + f."""
    k = KINDS[kind] if isinstance(kind, str) else int(kind)
    x, y, z = (int(v) for v in dims)
    if out is None:
        out = np.empty((x, y, z), np.uint16)
    assert out.dtype == np.uint16 and out.flags.c_contiguous and out.size == x * y * z
    _get().lsgen_volume(k, seed, first_frame, x, y, z, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    return out


def frame(kind, seed: int, index: int, y: int, z: int) -> np.ndarray:
    k = KINDS[kind] if isinstance(kind, str) else int(kind)
    out = np.empty((y, z), np.uint16)
    _get().lsgen_frame(k, seed, index, y, z, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    return out


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(memoryview(np.ascontiguousarray(a)).cast("B")).hexdigest()
