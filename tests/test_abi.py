"""The C-ABI library (CPU): it loads, exports every declared symbol, computes
stream geometry, and rejects bad arguments before touching the GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_c_functions():
    names = set()
    for h in ("lscomp.h",):
        text = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"\b(lscomp_\w+)\s*\(", text))
    return sorted(names)


def test_exports_every_declared_symbol():
    import lscomp
    syms = subprocess.run(["nm", "-D", "--defined-only", lscomp.LIB_PATH], capture_output=True, text=True,
                          check=True).stdout
    exported = {line.split()[-1] for line in syms.splitlines() if line.strip()}
    declared = _declared_c_functions()
    assert declared, "no declarations parsed"
    missing = [n for n in declared if n not in exported]
    assert not missing, missing
    assert set(declared) == set(lscomp.EXPORTS)
    # the reference-named C++ shims (lscomp_culsz.h), C++ linkage
    dem = subprocess.run(["nm", "-DC", "--defined-only", lscomp.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    assert "cuLSZ_compression_uint16_bsize64(unsigned short*, unsigned char*, unsigned long*, HIP_vector_type" in dem
    assert "cuLSZ_decompression_uint16_bsize64(unsigned short*, unsigned char*, unsigned long, HIP_vector_type" in dem


@pytest.mark.parametrize("dims", [(1, 1, 1), (1, 8, 8), (1, 8, 9), (2, 13, 17), (600, 1813, 1558),
                                  (512, 1024, 1024), (16, 512, 512)])
def test_geometry(orc, dims):
    import lscomp
    assert lscomp.num_blocks(dims) == orc.num_blocks(dims)
    assert lscomp.header_bytes(dims) == orc.header_bytes(dims)
    assert lscomp.max_compressed_bytes(dims) == orc.max_compressed_bytes(dims)


def test_geometry_cfg2_numbers():
    import lscomp
    assert lscomp.num_blocks((600, 1813, 1558)) == 26_559_000
    assert lscomp.max_compressed_bytes((600, 1813, 1558)) == 3_426_111_000


def test_argument_validation_without_gpu():
    """Every domain check happens before any HIP call (this container has no GPU)."""
    import lscomp
    L = lscomp.lib()
    D, B = lscomp.Dim3, lscomp.Bins4
    n = ctypes.c_size_t(0)
    ok_in, ok_out = ctypes.c_void_p(0x10000), ctypes.c_void_p(0x20000)
    cap = 1 << 20

    def comp(din=ok_in, dout=ok_out, c=cap, dims=(1, 8, 8), bins=(1, 1, 1, 1), p=1.0):
        return L.lscomp_compress_u16(din, dout, c, ctypes.byref(n), D(*dims), B(*bins), ctypes.c_float(p), None)

    E = lscomp.ERR_INVALID_ARG
    assert comp(dims=(0, 8, 8)) == E
    assert comp(dims=(1, 0, 8)) == E
    assert comp(bins=(0, 1, 1, 1)) == E
    assert comp(bins=(3, 2, 5, 6)) == E            # x <= y <= z <= w
    assert comp(bins=(1, 2, 3, 2)) == E
    assert comp(p=-0.1) == E
    assert comp(p=1.5) == E
    assert comp(p=float("nan")) == E
    assert comp(din=ctypes.c_void_p(0)) == E
    assert comp(din=ctypes.c_void_p(0x10001)) == E    # odd address
    assert comp(dout=ctypes.c_void_p(0x20002)) == E   # stream must be 4-byte aligned
    assert comp(c=3) == lscomp.ERR_CAPACITY and n.value == 4
    assert comp(dims=(65536, 65536, 65536)) == lscomp.ERR_TOO_LARGE
    # y or z near 2^32: the block counts are formed in 64 bits (a 32-bit ceil
    # wrapped to 0 and divided by zero on the host); validation goes on to the
    # null pointer
    for dims in ((1, 0xFFFFFFFF, 8), (1, 8, 0xFFFFFFFF), (1, 0xFFFFFFF9, 16)):
        assert comp(dims=dims, din=ctypes.c_void_p(0)) == E, dims
        assert L.lscomp_num_blocks_u16(D(*dims)) == dims[0] * ((dims[1] + 7) // 8) * ((dims[2] + 7) // 8)
    assert L.lscomp_compress_u16_async(ok_in, ok_out, cap, None, D(1, 8, 8), B(1, 1, 1, 1),
                                       ctypes.c_float(1.0), None) == E
    dec = L.lscomp_decompress_u16
    assert dec(ok_out, ok_in, cap, D(1, 8, 8), B(1, 1, 1, 0), None) == E
    assert dec(ctypes.c_void_p(0x20001), ok_in, cap, D(1, 8, 8), B(1, 1, 1, 1), None) == E
    assert dec(ok_out, ok_in, 3, D(1, 8, 8), B(1, 1, 1, 1), None) == lscomp.ERR_CORRUPT  # shorter than headers


def test_status_strings():
    import lscomp
    for st in range(7):
        assert lscomp.lib().lscomp_status_string(st)
    assert lscomp.version() >= 100


def test_header_compiles_as_c(tmp_path):
    """include/lscomp.h is plain C (what a cgo/ctypes/FFI binding would parse)."""
    src = tmp_path / "t.c"
    src.write_text('#include "lscomp.h"\nint main(void){lscomp_dim3 d={1,8,8}; return (int)lscomp_num_blocks_u16(d) - 1;}\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
                    "-I", "/opt/rocm/include", "-c", str(src), "-o", str(tmp_path / "t.o")], check=True)
