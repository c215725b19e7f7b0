"""The kernels' per-block code (lscomp_block.h) built for the host, against the oracle.

The gfx950 kernels and this host build compile the same header; only
v_perm_b32 / v_alignbit_b32 are emulated on the host.  The GPU parity tests
cover the device build of the same functions.
"""
import ctypes

import numpy as np
import pytest

from lsc_helpers import MODES

u8p = ctypes.POINTER(ctypes.c_uint8)
u16p = ctypes.POINTER(ctypes.c_uint16)
u32p = ctypes.POINTER(ctypes.c_uint32)


def _blocks(rng, n):
    out = []
    for i in range(n):
        kind = i % 7
        if kind == 0:
            v = rng.poisson(0.3, 64)
        elif kind == 1:
            v = rng.integers(0, 65536, 64)
        elif kind == 2:
            v = rng.integers(0, 16, 64) * (rng.random(64) < 0.3)
        elif kind == 3:
            v = np.full(64, rng.integers(0, 65536))
        elif kind == 4:
            v = rng.integers(0, 2 ** int(rng.integers(1, 17)), 64)
        elif kind == 5:
            v = np.zeros(64)
            v[rng.integers(0, 64)] = rng.integers(0, 65536)
        else:
            v = rng.poisson(0.05, 64) * rng.integers(1, 4000)
        out.append(np.clip(v, 0, 65535).astype(np.uint16))
    return out


# planes form: by each block's plane count, one plane at a time, all 8 at once
@pytest.mark.parametrize("form", [0, 1, 2], ids=["by_count", "per_plane", "transpose"])
@pytest.mark.parametrize("mode", sorted(MODES))
def test_encode_decode_vs_oracle(block_host, orc, mode, form):
    bins, p = MODES[mode]
    b = np.array(bins, np.uint32)
    rng = np.random.default_rng(sum(map(ord, mode)))
    block_host.bh_set_planes_form(form)
    for v in _blocks(rng, 3000):
        h, _ = orc.block_header(v.astype(np.uint32), b, p)
        ref = orc.encode_block(v.astype(np.uint32), h, b)
        out = np.zeros(128, np.uint8)
        hh = ctypes.c_uint32(0)
        n = block_host.bh_encode(v.ctypes.data_as(u16p), b.ctypes.data_as(u32p), ctypes.c_float(p),
                                 out.ctypes.data_as(u8p), ctypes.byref(hh))
        assert hh.value == h, (v, hex(h), hex(hh.value))
        assert out[:n].tobytes() == ref
        pl = np.zeros(128, np.uint8)
        pl[: len(ref)] = np.frombuffer(ref, np.uint8)
        dec = np.zeros(64, np.uint16)
        block_host.bh_decode(h, pl.ctypes.data_as(u8p), b.ctypes.data_as(u32p), dec.ctypes.data_as(u16p))
        assert np.array_equal(dec, orc.decode_block(h, ref, b).astype(np.uint16))


@pytest.mark.parametrize("form", [0, 1, 2], ids=["by_count", "per_plane", "transpose"])
def test_decode_random_payloads(block_host, orc, form):
    """Arbitrary (also non-canonical) headers and payloads decode like the oracle, incl. mod-2^16 products."""
    rng = np.random.default_rng(99)
    block_host.bh_set_planes_form(form)
    for bins in [(1, 1, 1, 1), (3, 5, 10, 15), (7, 300, 5000, 70000)]:
        b = np.array(bins, np.uint32)
        for _ in range(2000):
            h = int(rng.integers(0, 256)) & 0xF0 | int(rng.integers(0, 17))
            h = (h & 0xE0) | min(h & 0x1F, 16)
            pl = rng.integers(0, 256, 128).astype(np.uint8)
            dec = np.zeros(64, np.uint16)
            block_host.bh_decode(h, pl.ctypes.data_as(u8p), b.ctypes.data_as(u32p), dec.ctypes.data_as(u16p))
            ref = orc.decode_block(h, pl.tobytes(), b).astype(np.uint64) & 0xFFFF
            assert np.array_equal(dec, ref.astype(np.uint16)), hex(h)


def test_pool_threshold_parameter(block_host):
    from oracle_pins import POOL_THRESHOLDS
    for p, zmin in POOL_THRESHOLDS.items():
        assert block_host.bh_zth(p) == (65 if zmin is None else zmin)


def test_fastdiv(block_host):
    rng = np.random.default_rng(0)
    ds = list(range(1, 2000)) + [int(x) for x in rng.integers(1, 2 ** 32, 2000)]
    ds += [2 ** 31, 2 ** 31 + 1, 2 ** 32 - 1, 2 ** 32 - 2, 65536, 65537, 195, 227 * 195]
    for d in ds:
        ns = [0, 1, d - 1, d, d + 1, 2 ** 32 - 1, 2 ** 32 - 2, 2 ** 31, 2 ** 31 - 1]
        ns += [int(x) for x in rng.integers(0, 2 ** 32, 16)]
        for n in ns:
            n %= 2 ** 32
            assert block_host.bh_fdiv(n, d) == n // d, (n, d)


def test_quantize_exhaustive(block_host):
    """q = trunc((float)v * f_d) == v // d for every v < 2^16 and every d <= 65536, and large d."""
    assert block_host.bh_quant_mismatches(1, 65536) == 0
    assert block_host.bh_quant_mismatches(65537, 66000) == 0
    assert block_host.bh_quant_mismatches(4294967000, 4294967295) == 0


def test_quantize_pooled_sums(block_host):
    """Pooled values come from the pair sum s < 2^17: trunc((float)s * f2_d) == s // 2d
    at both ends of every step, for every d <= 65536 and large d (the function is
    monotone in s, so this covers every s)."""
    assert block_host.bh_quant2_mismatches(1, 65536) == 0
    assert block_host.bh_quant2_mismatches(65537, 70000) == 0
    assert block_host.bh_quant2_mismatches(4294967000, 4294967295) == 0


def test_quantize_mulhi24_exhaustive(block_host):
    """The 24-bit multiply-high division (every bin in [2, 16384]) equals exact
    division for every plain value < 2^16 and every pooled pair sum < 2^17, run
    through the kernels' own quantize_plain24 / quantize_pooled24."""
    assert block_host.bh_q24_mismatches(2, 16384) == 0


def test_mulhi24_path_selection(block_host):
    """Bins in [2, 16384] take the multiply-high path; a bin of 1 or above 16384 keeps the float path."""
    q24 = 8
    for bins, on in [((3, 5, 10, 15), True), ((2, 2, 2, 2), True), ((16384,) * 4, True), ((3, 4, 5, 6), True),
                     ((1, 2, 4, 8), False), ((1, 1, 1, 1), False), ((2, 65535, 70000, 100000), False),
                     ((16385,) * 4, False)]:
        b = np.array(bins, np.uint32)
        flags = block_host.bh_decide_flags(b.ctypes.data_as(u32p), ctypes.c_float(0.5))
        assert bool(flags & q24) == on, bins


def test_header_bin_choice_edges(block_host, orc):
    """Blocks whose maximum sits at, just below and just above each bin edge: the
    bin choice skips its minimum search exactly when the maximum is below the
    second bin, and must still agree with the oracle's counts."""
    rng = np.random.default_rng(11)
    for bins in [(3, 5, 10, 15), (3, 3, 10, 15), (2, 65535, 70000, 100000), (1, 2, 4, 8), (5, 5, 5, 6)]:
        b = np.array(bins, np.uint32)
        edges = sorted({max(0, min(65535, e + k)) for e in bins for k in (-1, 0, 1)})
        for p in (0.0, 0.5, 1.0):
            for top in edges:
                for low in (0, 1, bins[0] - 1 if bins[0] > 0 else 0, bins[0]):
                    v = np.minimum(rng.integers(0, low + 1, 64), 65535).astype(np.uint16)
                    v[int(rng.integers(0, 64))] = top
                    h, _ = orc.block_header(v.astype(np.uint32), b, p)
                    out = np.zeros(128, np.uint8)
                    hh = ctypes.c_uint32(0)
                    n = block_host.bh_encode(v.ctypes.data_as(u16p), b.ctypes.data_as(u32p), ctypes.c_float(p),
                                             out.ctypes.data_as(u8p), ctypes.byref(hh))
                    assert hh.value == h, (bins, p, top, low, hex(h), hex(hh.value))
                    assert out[:n].tobytes() == orc.encode_block(v.astype(np.uint32), h, b)


def _blocks_u32(rng, n):
    out = []
    for i in range(n):
        kind = i % 6
        if kind == 0:
            v = rng.poisson(0.3, 64) * rng.integers(1, 1 << 24)
        elif kind == 1:
            v = rng.integers(0, 1 << 31, 64)
        elif kind == 2:   # pairs whose sum wraps modulo 2^32
            v = rng.integers((1 << 32) - 1000, 1 << 32, 64) * (rng.random(64) < 0.5)
        elif kind == 3:
            v = rng.integers(0, 2 ** int(rng.integers(1, 33)), 64)
        elif kind == 4:
            v = np.zeros(64, np.int64)
            v[rng.integers(0, 64)] = rng.integers(0, 1 << 32)
        else:
            v = rng.integers(0, 1 << 16, 64)
        out.append(np.asarray(v, np.uint64).astype(np.uint32))
    return out


@pytest.mark.parametrize("form", [0, 1, 2], ids=["by_count", "per_plane", "transpose"])
@pytest.mark.parametrize("mode", sorted(MODES))
def test_encode_decode_u32_vs_oracle(block_host, orc, mode, form):
    """uint32 blocks: header, payload and decoded values equal the oracle's (rate-32 blocks are flagged)."""
    bins, p = MODES[mode]
    rng = np.random.default_rng(7 + sum(map(ord, mode)))
    block_host.bh_set_planes_form(form)
    for bins_ in (bins, (1000, 5000, 70000, 1 << 20)):
        b = np.array(bins_, np.uint32)
        for v in _blocks_u32(rng, 1500):
            h, pb = orc.block_header(v, b, p)
            rate = pb // (4 if h & 0x80 else 8)
            out = np.zeros(256, np.uint8)
            hh, r32 = ctypes.c_uint32(0), ctypes.c_uint32(0)
            n = block_host.bh_encode_u32(v.ctypes.data_as(u32p), b.ctypes.data_as(u32p), ctypes.c_float(p),
                                         out.ctypes.data_as(u8p), ctypes.byref(hh), ctypes.byref(r32))
            assert bool(r32.value) == (rate == 32)
            if rate == 32:
                continue
            assert hh.value == h, (v, hex(h), hex(hh.value))
            ref = orc.encode_block(v, h, b)
            assert out[:n].tobytes() == ref
            pl = np.zeros(256, np.uint8)
            pl[: len(ref)] = np.frombuffer(ref, np.uint8)
            dec = np.zeros(64, np.uint32)
            block_host.bh_decode_u32(h, pl.ctypes.data_as(u8p), b.ctypes.data_as(u32p), dec.ctypes.data_as(u32p))
            assert np.array_equal(dec, orc.decode_block(h, ref, b))

