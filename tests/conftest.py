"""Shared test setup: import paths, the `gpu` marker, and common fixtures."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sc25-lscomp_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle.get()


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """The lscomp binding on cuda:0; fails loudly if the HIP library is absent."""
    if not _gpu_available():
        pytest.fail("GPU test collected without a GPU")
    import torch
    import lscomp  # raises ImportError when liblscomp.so is missing
    torch.cuda.set_device(0)
    return lscomp


@pytest.fixture(scope="session")
def block_host():
    """Host build of the kernels' per-block code (tests/csrc/block_host.cpp)."""
    import ctypes
    out_dir = os.path.join(ROOT, "tests", "_build")
    os.makedirs(out_dir, exist_ok=True)
    so = os.path.join(out_dir, "libblock_host.so")
    src = os.path.join(ROOT, "tests", "csrc", "block_host.cpp")
    hdrs = [os.path.join(PKG, "csrc", h) for h in ("lscomp_block.h", "lscomp_params.h")]
    newest = max(os.path.getmtime(f) for f in [src] + hdrs)
    if not os.path.exists(so) or os.path.getmtime(so) < newest:
        clang = "/opt/rocm/llvm/bin/clang++"
        if not os.path.exists(clang):
            pytest.skip("clang++ from ROCm not available")
        subprocess.run([clang, "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(PKG, "csrc"),
                        "-o", so + ".tmp", src], check=True)
        os.replace(so + ".tmp", so)
    lib = ctypes.CDLL(so)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    u16p = ctypes.POINTER(ctypes.c_uint16)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib.bh_encode.restype = ctypes.c_uint32
    lib.bh_encode.argtypes = [u16p, u32p, ctypes.c_float, u8p, u32p]
    lib.bh_set_planes_form.restype = None
    lib.bh_set_planes_form.argtypes = [ctypes.c_int]
    lib.bh_decode.restype = None
    lib.bh_decode.argtypes = [ctypes.c_uint32, u8p, u32p, u16p]
    lib.bh_encode_u32.restype = ctypes.c_uint32
    lib.bh_encode_u32.argtypes = [u32p, u32p, ctypes.c_float, u8p, u32p, u32p]
    lib.bh_decode_u32.restype = None
    lib.bh_decode_u32.argtypes = [ctypes.c_uint32, u8p, u32p, u32p]
    lib.bh_fdiv.restype = ctypes.c_uint32
    lib.bh_fdiv.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    lib.bh_quant_mismatches.restype = ctypes.c_uint64
    lib.bh_quant_mismatches.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    lib.bh_quant2_mismatches.restype = ctypes.c_uint64
    lib.bh_quant2_mismatches.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    lib.bh_q24_mismatches.restype = ctypes.c_uint64
    lib.bh_q24_mismatches.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    lib.bh_decide_flags.restype = ctypes.c_uint32
    lib.bh_decide_flags.argtypes = [u32p, ctypes.c_float]
    lib.bh_zth.restype = ctypes.c_uint32
    lib.bh_zth.argtypes = [ctypes.c_float]
    return lib
