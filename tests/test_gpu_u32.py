"""uint32 volumes on the GPU against the oracle (lsCOMP/src/cuLSZ_kernel.cu:4-824):
byte-equal streams and equal decoded volumes, the rate-32 rejection, and the
uint16 equivalence (a uint32 volume of uint16 values gives the uint16 stream)."""
import numpy as np
import pytest

from lsc_helpers import synth

pytestmark = pytest.mark.gpu

U32_MODES = {
    "lossless": ((1, 1, 1, 1), 1.0),
    "b3-5-10-15_p0.5": ((3, 5, 10, 15), 0.5),
    "big_bins_p0.25": ((1000, 5000, 70000, 1 << 20), 0.25),
}


def synth_u32(kind, shape, seed):
    rng = np.random.default_rng(seed)
    if kind == "wide":        # values below 2^31 (rate <= 31 with bin 1), half zeros
        v = rng.integers(0, 1 << 31, size=shape, dtype=np.uint64)
        v[rng.random(shape) < 0.5] = 0
    elif kind == "counts":    # photon counts scaled into the high bits
        v = rng.poisson(0.4, size=shape).astype(np.uint64) * rng.integers(1, 1 << 22, size=shape, dtype=np.uint64)
    elif kind == "u16":       # uint16 range
        v = synth("cssi", shape, seed).astype(np.uint64)
    else:
        raise ValueError(kind)
    return v.astype(np.uint32)


def to_dev32(vol):
    import torch
    return torch.from_numpy(np.ascontiguousarray(vol).view(np.int32)).cuda()


def check_u32(gpu, orc, vol, bins, p):
    st, ref, refn = orc.compress_u32(vol, bins, p)
    assert st == 0
    out, n = gpu.compress_u32(to_dev32(vol), vol.shape, bins, p)
    assert n == refn, (n, refn)
    got = out[:n].cpu().numpy()
    if not np.array_equal(got, ref):
        bad = np.nonzero(got != ref)[0]
        pytest.fail(f"stream differs at {bad.size} bytes, first {bad[:8]}")
    dec = gpu.decompress_u32(out, n, vol.shape, bins).cpu().numpy().view(np.uint32)
    st, refdec = orc.decompress_u32(ref, vol.shape, bins)
    assert st == 0 and np.array_equal(dec, refdec)
    if tuple(bins) == (1, 1, 1, 1) and p >= 1.0:
        assert np.array_equal(dec, vol)


SHAPES = [(1, 8, 8), (1, 1, 1), (2, 13, 17), (1, 9, 10), (3, 64, 64), (1, 100, 101), (2, 50, 1558),
          (4, 136, 200)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("kind", ["wide", "counts", "u16"])
@pytest.mark.parametrize("mode", sorted(U32_MODES))
def test_u32_shapes(gpu, orc, shape, kind, mode):
    bins, p = U32_MODES[mode]
    seed = (shape[0] * 31 + shape[1] * 7 + shape[2] + len(kind)) % 1000
    check_u32(gpu, orc, synth_u32(kind, shape, seed), bins, p)


def test_u32_cssi_frame_shape(gpu, orc):
    """Sixteen 1813 x 1558 frames (the CSSI detector's frame: a partial block at
    every row end) of photon counts scaled past 16 bits, lossless and lossy."""
    vol = synth_u32("counts", (16, 1813, 1558), 11)
    for bins, p in (((1, 1, 1, 1), 1.0), ((3, 5, 10, 15), 0.5)):
        print(f"u32 frame-shape case bins={bins} p={p}", flush=True)
        check_u32(gpu, orc, vol, bins, p)


def test_u32_long_chain(gpu, orc):
    """128 and 512 tiles of 256 blocks: look-back and pre-pass over many tiles."""
    for shape in ((8, 512, 512), (32, 512, 512)):
        check_u32(gpu, orc, synth_u32("counts", shape, 3), (3, 5, 10, 15), 0.5)


def test_u32_wrapping_pairs(gpu, orc):
    """Pairs near 2^32 whose sum wraps (reference uint arithmetic), with bins large enough for rate <= 31."""
    rng = np.random.default_rng(9)
    vol = rng.integers((1 << 32) - 4096, 1 << 32, size=(2, 24, 40), dtype=np.uint64)
    vol[rng.random(vol.shape) < 0.6] = 0
    vol = vol.astype(np.uint32)
    for bins, p in (((2, 4, 8, 16), 0.5), ((2, 2, 2, 2), 0.0), ((4, 4, 4, 4), 1.0)):
        check_u32(gpu, orc, vol, bins, p)


def test_u32_rate32_rejected(gpu, orc):
    """A quantized maximum of 2^31 needs rate 32: ERR_RANGE, like the oracle; the next call is clean."""
    vol = synth_u32("counts", (2, 16, 16), 4)
    vol[1, 9, 3] = 1 << 31
    assert orc.compress_u32(vol, (1, 1, 1, 1), 1.0)[0] == 4
    with pytest.raises(gpu.LscompError) as e:
        gpu.compress_u32(to_dev32(vol), vol.shape, (1, 1, 1, 1), 1.0)
    assert e.value.status == gpu.ERR_RANGE
    check_u32(gpu, orc, vol, (2, 2, 2, 2), 1.0)   # with bin 2 it is rate 31


def test_u32_matches_u16_path(gpu):
    """uint16 values in a uint32 volume give the uint16 stream, byte for byte."""
    import torch
    vol = synth("cssi", (3, 40, 72), 5)
    for bins, p in (((1, 1, 1, 1), 1.0), ((3, 5, 10, 15), 0.5)):
        a, na = gpu.compress(torch.from_numpy(vol.view(np.int16)).cuda(), vol.shape, bins, p)
        b, nb = gpu.compress_u32(to_dev32(vol.astype(np.uint32)), vol.shape, bins, p)
        assert na == nb and torch.equal(a[:na], b[:nb])


def test_u32_misaligned_and_capacity(gpu, orc):
    """Volume at a 4-byte offset; too small an output buffer reports the required size;
    a truncated stream is corrupt."""
    import torch
    vol = synth_u32("wide", (2, 21, 30), 6)
    big = torch.zeros(vol.size + 4, dtype=torch.int32, device="cuda")
    big[1:1 + vol.size] = to_dev32(vol).reshape(-1)
    out, n = gpu.compress_u32(big[1:1 + vol.size], vol.shape, (1, 1, 1, 1), 1.0)
    st, ref, refn = orc.compress_u32(vol, (1, 1, 1, 1), 1.0)
    assert n == refn and np.array_equal(out[:n].cpu().numpy(), ref)
    small = torch.empty(refn - 64, dtype=torch.uint8, device="cuda")
    with pytest.raises(gpu.LscompError) as e:
        gpu.compress_u32(to_dev32(vol), vol.shape, (1, 1, 1, 1), 1.0, out=small)
    assert e.value.status == gpu.ERR_CAPACITY and e.value.required == refn
    with pytest.raises(gpu.LscompError) as e:
        gpu.decompress_u32(out, refn - 8, vol.shape)
    assert e.value.status == gpu.ERR_CORRUPT
    dec = gpu.decompress_u32(out, refn, vol.shape).cpu().numpy().view(np.uint32)
    assert np.array_equal(dec, vol)


@pytest.mark.parametrize("kind", ["rate24", "alternating", "u16x4099"])
def test_u32_many_tiles_per_workgroup(gpu, orc, kind):
    """4096 tiles of 256 blocks over 512 workgroups (8 each), lossless: payloads
    too large to stage beside the previous tile (24-bit values), alternating
    dense and sparse frames, and CSSI-like values scaled up (the staging ring
    wraps).  The decode equals the input; a slab of frames also equals the
    oracle's stream byte for byte."""
    import torch
    rng = np.random.default_rng(51)
    dims = (64, 1024, 1024)
    if kind == "rate24":
        vol = rng.integers(0, 1 << 24, size=dims, dtype=np.uint32)
    elif kind == "alternating":
        vol = rng.poisson(0.3, size=dims).astype(np.uint32)
        vol[::2] = rng.integers(0, 1 << 20, size=(dims[0] // 2,) + dims[1:], dtype=np.uint32)
    else:
        vol = synth("cssi", dims, seed=53).astype(np.uint32) * 4099
    d = to_dev32(vol)
    out, n = gpu.compress_u32(d, dims, (1, 1, 1, 1), 1.0)
    dec = gpu.decompress_u32(out, n, dims, (1, 1, 1, 1))
    assert bool(torch.equal(dec, d))
    check_u32(gpu, orc, np.ascontiguousarray(vol[:4]), (1, 1, 1, 1), 1.0)


@pytest.mark.parametrize("shape", [(2, 2048, 2048), (2, 2048, 2056)], ids=["256tiles", "257tiles"])
def test_u32_small_volume_decode_limit(gpu, orc, shape):
    """uint32 volumes of at most 256 decode tiles (of 512 blocks) decode with
    one workgroup per tile; 257 tiles take the offset-phase launch.  Both sides
    of the limit equal the oracle."""
    for bins, p in (((1, 1, 1, 1), 1.0), ((3, 5, 10, 15), 0.5)):
        check_u32(gpu, orc, synth_u32("counts", shape, 9), bins, p)
