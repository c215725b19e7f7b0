"""CPU oracle of the lsCOMP uint16 and uint32 codecs -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper over ``liblscomp_oracle.so`` (built from ``lscomp_oracle.c`` by
``make -C oracle``), a scalar C restatement of the reference's per-block logic
(lsCOMP/src/cuLSZ_kernel.cu:827-1646, lsCOMP/src/cuLSZ_entry.cu:93-166; uint32:
cuLSZ_kernel.cu:4-824, cuLSZ_entry.cu:18-91).
Only the tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the checker or the timed CPU baseline;
the product library never calls it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblscomp_oracle.so")

OK, ERR_ARG, ERR_CAPACITY, ERR_CORRUPT, ERR_RANGE = 0, 1, 2, 3, 4

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64


def build(mutant: int | None = None) -> str:
    """Compile the oracle (or a mutant with one planted mistake); returns the .so path."""
    import subprocess
    if mutant is None:
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
        return _LIB_PATH
    subprocess.run(["make", "-s", "-C", _HERE, "mutant", f"MUTANT={mutant}"], check=True)
    return os.path.join(_HERE, "_build", f"liblscomp_oracle_mutant{mutant}.so")


class Oracle:
    """One loaded oracle library (the real one, or a mutant for pin tests)."""

    def __init__(self, path: str = _LIB_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing; run `make -C oracle`")
        lib = ctypes.CDLL(path)
        lib.oracle_num_blocks.restype = _u64
        lib.oracle_num_blocks.argtypes = [_u32, _u32, _u32]
        lib.oracle_header_bytes.restype = _u64
        lib.oracle_header_bytes.argtypes = [_u64]
        lib.oracle_max_compressed_bytes.restype = _u64
        lib.oracle_max_compressed_bytes.argtypes = [_u32, _u32, _u32]
        lib.oracle_block_header.restype = ctypes.c_uint8
        lib.oracle_block_header.argtypes = [_u32p, _u32p, ctypes.c_float, _u32p]
        lib.oracle_encode_block.restype = _u32
        lib.oracle_encode_block.argtypes = [_u32p, ctypes.c_uint8, _u32p, _u8p]
        lib.oracle_decode_block.restype = None
        lib.oracle_decode_block.argtypes = [ctypes.c_uint8, _u8p, _u32p, _u32p]
        for name in ("oracle_compress_u16", "oracle_compress_u16_omp"):
            f = getattr(lib, name)
            f.restype = ctypes.c_int
            args = [_u16p, _u32, _u32, _u32, _u32p, ctypes.c_float, _u8p, _u64, _u64p]
            if name.endswith("_omp"):
                args.append(_u64p)
            f.argtypes = args
        for name in ("oracle_decompress_u16", "oracle_decompress_u16_omp"):
            f = getattr(lib, name)
            f.restype = ctypes.c_int
            args = [_u8p, _u64, _u32, _u32, _u32, _u32p, _u16p]
            if name.endswith("_omp"):
                args.append(_u64p)
            f.argtypes = args
        lib.oracle_max_compressed_bytes_u32.restype = _u64
        lib.oracle_max_compressed_bytes_u32.argtypes = [_u32, _u32, _u32]
        lib.oracle_compress_u32.restype = ctypes.c_int
        lib.oracle_compress_u32.argtypes = [_u32p, _u32, _u32, _u32, _u32p, ctypes.c_float, _u8p, _u64, _u64p]
        lib.oracle_decompress_u32.restype = ctypes.c_int
        lib.oracle_decompress_u32.argtypes = [_u8p, _u64, _u32, _u32, _u32, _u32p, _u32p]
        lib.oracle_max_threads.restype = ctypes.c_int
        lib.oracle_ssim_u16.restype = ctypes.c_double
        lib.oracle_ssim_u16.argtypes = [_u16p, _u16p, _u32, _u32, _u32]
        self.lib = lib

    # -- geometry ---------------------------------------------------------
    def num_blocks(self, dims) -> int:
        return int(self.lib.oracle_num_blocks(*map(int, dims)))

    def header_bytes(self, dims) -> int:
        return int(self.lib.oracle_header_bytes(self.num_blocks(dims)))

    def max_compressed_bytes(self, dims) -> int:
        return int(self.lib.oracle_max_compressed_bytes(*map(int, dims)))

    # -- single block -----------------------------------------------------
    def block_header(self, v64, bins, p: float):
        v = np.ascontiguousarray(v64, dtype=np.uint32).reshape(64)
        b = np.ascontiguousarray(bins, dtype=np.uint32)
        pb = _u32(0)
        h = self.lib.oracle_block_header(v.ctypes.data_as(_u32p), b.ctypes.data_as(_u32p),
                                         ctypes.c_float(p), ctypes.byref(pb))
        return int(h), int(pb.value)

    def encode_block(self, v64, hdr: int, bins) -> bytes:
        v = np.ascontiguousarray(v64, dtype=np.uint32).reshape(64)
        b = np.ascontiguousarray(bins, dtype=np.uint32)
        out = np.zeros(256, np.uint8)
        n = self.lib.oracle_encode_block(v.ctypes.data_as(_u32p), ctypes.c_uint8(hdr),
                                         b.ctypes.data_as(_u32p), out.ctypes.data_as(_u8p))
        return out[:n].tobytes()

    def decode_block(self, hdr: int, payload: bytes, bins) -> np.ndarray:
        pl = np.zeros(256, np.uint8)
        pl[: len(payload)] = np.frombuffer(payload, np.uint8)
        b = np.ascontiguousarray(bins, dtype=np.uint32)
        v = np.zeros(64, np.uint32)
        self.lib.oracle_decode_block(ctypes.c_uint8(hdr), pl.ctypes.data_as(_u8p),
                                     b.ctypes.data_as(_u32p), v.ctypes.data_as(_u32p))
        return v

    # -- whole volume -----------------------------------------------------
    def ssim(self, ori: np.ndarray, dec: np.ndarray) -> float:
        """Mean SSIM of two uint16 volumes (x, y, z), as the reference's computeSSIM
        (lsCOMP/src/cuLSZ_utility.cu:757-886) on their float values; NaN when a dimension is below 7."""
        a = np.ascontiguousarray(ori, dtype=np.uint16)
        b = np.ascontiguousarray(dec, dtype=np.uint16)
        assert a.shape == b.shape and a.ndim == 3
        x, y, z = a.shape
        return float(self.lib.oracle_ssim_u16(a.ctypes.data_as(_u16p), b.ctypes.data_as(_u16p), x, y, z))

    def compress(self, vol: np.ndarray, bins, p: float, omp: bool = False, capacity: int | None = None):
        """Returns (status, stream bytes as np.uint8 array, required size)."""
        vol = np.ascontiguousarray(vol, dtype=np.uint16)
        x, y, z = vol.shape
        cap = self.max_compressed_bytes((x, y, z)) if capacity is None else capacity
        out = np.zeros(max(cap, 1), np.uint8)
        b = np.ascontiguousarray(bins, dtype=np.uint32)
        n = _u64(0)
        args = [vol.ctypes.data_as(_u16p), x, y, z, b.ctypes.data_as(_u32p), ctypes.c_float(p),
                out.ctypes.data_as(_u8p), cap, ctypes.byref(n)]
        if omp:
            scratch = np.zeros(self.lib.oracle_max_threads() + 2, np.uint64)
            st = self.lib.oracle_compress_u16_omp(*args, scratch.ctypes.data_as(_u64p))
        else:
            st = self.lib.oracle_compress_u16(*args)
        size = int(n.value)
        return int(st), out[: min(size, cap)], size

    def decompress(self, stream, dims, bins, omp: bool = False, fill: int = 0):
        """Returns (status, volume uint16 (x, y, z))."""
        s = np.ascontiguousarray(np.frombuffer(bytes(stream), np.uint8) if not isinstance(stream, np.ndarray)
                                 else stream, dtype=np.uint8)
        x, y, z = map(int, dims)
        out = np.full((x, y, z), fill, np.uint16)
        b = np.ascontiguousarray(bins, dtype=np.uint32)
        args = [s.ctypes.data_as(_u8p), s.size, x, y, z, b.ctypes.data_as(_u32p), out.ctypes.data_as(_u16p)]
        if omp:
            scratch = np.zeros(self.lib.oracle_max_threads() + 2, np.uint64)
            st = self.lib.oracle_decompress_u16_omp(*args, scratch.ctypes.data_as(_u64p))
        else:
            st = self.lib.oracle_decompress_u16(*args)
        return int(st), out

    # -- uint32 volumes ----------------------------------------------------
    def max_compressed_bytes_u32(self, dims) -> int:
        return int(self.lib.oracle_max_compressed_bytes_u32(*map(int, dims)))

    def compress_u32(self, vol: np.ndarray, bins, p: float, capacity: int | None = None):
        """Returns (status, stream bytes as np.uint8 array, required size)."""
        vol = np.ascontiguousarray(vol, dtype=np.uint32)
        x, y, z = vol.shape
        cap = self.max_compressed_bytes_u32((x, y, z)) if capacity is None else capacity
        out = np.zeros(max(cap, 1), np.uint8)
        b = np.ascontiguousarray(bins, dtype=np.uint32)
        n = _u64(0)
        st = self.lib.oracle_compress_u32(vol.ctypes.data_as(_u32p), x, y, z, b.ctypes.data_as(_u32p),
                                          ctypes.c_float(p), out.ctypes.data_as(_u8p), cap, ctypes.byref(n))
        size = int(n.value)
        return int(st), out[: min(size, cap)], size

    def decompress_u32(self, stream, dims, bins, fill: int = 0):
        """Returns (status, volume uint32 (x, y, z))."""
        s = np.ascontiguousarray(np.frombuffer(bytes(stream), np.uint8) if not isinstance(stream, np.ndarray)
                                 else stream, dtype=np.uint8)
        x, y, z = map(int, dims)
        out = np.full((x, y, z), fill, np.uint32)
        b = np.ascontiguousarray(bins, dtype=np.uint32)
        st = self.lib.oracle_decompress_u32(s.ctypes.data_as(_u8p), s.size, x, y, z, b.ctypes.data_as(_u32p),
                                            out.ctypes.data_as(_u32p))
        return int(st), out

    def max_threads(self) -> int:
        return int(self.lib.oracle_max_threads())


_default = None


def get() -> Oracle:
    """The real oracle (built on demand)."""
    global _default
    if _default is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _default = Oracle(_LIB_PATH)
    return _default
