"""Stream buffers at every legal alignment on the GPU.  The C ABI accepts
stream buffers at 4-byte alignment: compression's output (d_out) and
decompression's input (d_in) are placed at +0, +4, +8 and +12 bytes from a
16-byte boundary, for uint16 and uint32 volumes, lossless and lossy, through
the plain calls and through plans (each item at its own offset).  Streams and
decoded volumes are byte-equal to the oracle's, and the bytes around each
buffer are untouched (the 16-byte header stores, the offset phase's header
reads and the payload loads all run at dword alignment there)."""
import numpy as np
import pytest

from lsc_helpers import synth

pytestmark = pytest.mark.gpu

OFFSETS = (0, 4, 8, 12)
MODES = {"lossless": ((1, 1, 1, 1), 1.0), "lossy": ((3, 5, 10, 15), 0.5)}
GUARD = 0xC3
# several tiles (the look-back and the payload ring cross tiles), partial blocks;
# the last has more than 256 tiles, so its decode reads the stream through the
# offset phase and the pipelined (lossless) or sparse (lossy) decoder rather
# than with one workgroup per tile
SHAPES = [(3, 77, 130), (5, 300, 264), (10, 1000, 1030)]


def _window(n, off):
    """A uint8 device buffer of n bytes at `off` from a 16-byte boundary, with
    guard bytes on both sides; returns (whole buffer, view)."""
    import torch
    big = torch.full((n + off + 64,), GUARD, dtype=torch.uint8, device="cuda")
    assert big.data_ptr() % 16 == 0
    return big, big[off:off + n]


def _guards_intact(big, off, n):
    host = big.cpu().numpy()
    return bool(np.all(host[:off] == GUARD) and np.all(host[off + n:] == GUARD))


def _stream_at(ref, off):
    import torch
    big, view = _window(ref.size, off)
    view.copy_(torch.from_numpy(ref).cuda())
    return big, view


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("off", OFFSETS)
def test_u16_stream_offsets(gpu, orc, off, mode, shape):
    import torch
    bins, p = MODES[mode]
    vol = synth("cssi", shape, seed=21)
    st, ref, refn = orc.compress(vol, bins, p)
    assert st == 0
    d_in = torch.from_numpy(vol.view(np.int16)).cuda()
    cap = gpu.max_compressed_bytes(shape)
    big, out = _window(cap, off)
    assert out.data_ptr() % 16 == off
    _, n = gpu.compress(d_in, shape, bins, p, out=out)
    assert n == refn and np.array_equal(out[:n].cpu().numpy(), ref)
    assert _guards_intact(big, off, cap)
    # the same stream read from an offset buffer
    sbig, s = _stream_at(ref, off)
    dec = gpu.decompress(s, refn, shape, bins).cpu().numpy().view(np.uint16)
    _, refdec = orc.decompress(ref, shape, bins)
    assert np.array_equal(dec, refdec)
    assert _guards_intact(sbig, off, ref.size)


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("off", OFFSETS)
def test_u32_stream_offsets(gpu, orc, off, mode):
    import torch
    bins, p = MODES[mode]
    shape = (3, 90, 140)
    vol = (synth("cssi", shape, seed=22).astype(np.uint32) * 3001) & 0x7fffffff
    st, ref, refn = orc.compress_u32(vol, bins, p)
    assert st == 0
    d_in = torch.from_numpy(vol.view(np.int32)).cuda()
    cap = gpu.max_compressed_bytes_u32(shape)
    big, out = _window(cap, off)
    _, n = gpu.compress_u32(d_in, shape, bins, p, out=out)
    assert n == refn and np.array_equal(out[:n].cpu().numpy(), ref)
    assert _guards_intact(big, off, cap)
    sbig, s = _stream_at(ref, off)
    dec = gpu.decompress_u32(s, refn, shape, bins).cpu().numpy().view(np.uint32)
    _, refdec = orc.decompress_u32(ref, shape, bins)
    assert np.array_equal(dec, refdec)
    assert _guards_intact(sbig, off, ref.size)


@pytest.mark.parametrize("graph", [False, True])
def test_plan_stream_offsets(gpu, orc, graph):
    """One plan, items at +0/+4/+8/+12 in turn (compress outputs, then the
    decompress plan's inputs), lossless and lossy items mixed."""
    import torch
    shapes = [(2, 77, 130), (1, 136, 200), (3, 64, 66), (2, 250, 130), (1, 30, 1558), (2, 8, 8), (1, 99, 99),
              (4, 40, 72)]
    vols = [synth("cssi" if i % 2 else "xpcs", s, seed=30 + i) for i, s in enumerate(shapes)]
    params = [MODES["lossless"] if i % 2 == 0 else MODES["lossy"] for i in range(len(vols))]
    offs = [OFFSETS[i % 4] for i in range(len(vols))]
    refs = [orc.compress(v, b, p) for v, (b, p) in zip(vols, params)]
    d_in = [torch.from_numpy(v.view(np.int16)).cuda() for v in vols]
    caps = [gpu.max_compressed_bytes(v.shape) for v in vols]
    wins = [_window(c, o) for c, o in zip(caps, offs)]
    nblk = sum(gpu.num_blocks(v.shape) for v in vols)
    plan = gpu.Plan(gpu.COMPRESS, len(vols), nblk)
    plan.set_items([(d, v.shape, b, p, w[1], c) for d, v, (b, p), w, c in zip(d_in, vols, params, wins, caps)])
    plan.launch(graph=graph)
    sizes = plan.sizes()
    assert plan.status() == gpu.OK
    for (st, ref, refn), n, (big, view), c, o in zip(refs, sizes, wins, caps, offs):
        assert n == refn and np.array_equal(view[:n].cpu().numpy(), ref), o
        assert _guards_intact(big, o, c)
    plan.close()

    dplan = gpu.Plan(gpu.DECOMPRESS, len(vols), nblk)
    streams = [_stream_at(ref, o) for (st, ref, refn), o in zip(refs, offs)]
    outs = [torch.full(v.shape, 7, dtype=torch.int16, device="cuda") for v in vols]
    dplan.set_items([(out, v.shape, b, p, s[1], refn)
                     for out, v, (b, p), s, (st, ref, refn) in zip(outs, vols, params, streams, refs)])
    dplan.launch(graph=graph)
    assert dplan.status() == gpu.OK
    for out, v, (b, p), (st, ref, refn), (sbig, s), o in zip(outs, vols, params, refs, streams, offs):
        _, refdec = orc.decompress(ref, v.shape, b)
        assert np.array_equal(out.cpu().numpy().view(np.uint16), refdec), o
        assert _guards_intact(sbig, o, ref.size)
    dplan.close()
