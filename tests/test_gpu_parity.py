"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

Bit-exact streams and volumes on seeded inputs at sizes the oracle handles in
seconds, edge shapes (partial blocks on both axes, tiny volumes, odd widths,
misaligned pointers), multi-tile look-back chains, capacity and corruption
handling, and scratch reuse across launches and streams.
"""
import numpy as np
import pytest

from lsc_helpers import MODES, synth

pytestmark = pytest.mark.gpu


def to_dev(vol):
    import torch
    return torch.from_numpy(np.ascontiguousarray(vol).view(np.int16)).cuda()


def check_case(gpu, orc, vol, bins, p):
    dims = vol.shape
    d = to_dev(vol)
    out, n = gpu.compress(d, dims, bins, p)
    st, ref, refn = orc.compress(vol, bins, p)
    assert st == 0
    assert n == refn, f"stream size {n} != oracle {refn}"
    got = out[:n].cpu().numpy()
    if not np.array_equal(got, ref):
        bad = np.nonzero(got != ref)[0]
        pytest.fail(f"stream differs at {bad.size} bytes, first {bad[:8]} (R={orc.header_bytes(dims)})")
    dec = gpu.decompress(out, n, dims, bins)
    st, refdec = orc.decompress(ref, dims, bins)
    assert st == 0
    assert np.array_equal(dec.cpu().numpy().view(np.uint16), refdec)
    if tuple(bins) == (1, 1, 1, 1) and p >= 1.0:
        assert np.array_equal(refdec, vol)
    return n


# z even but not a multiple of 8 (1558, 10, 12, 2, 1030): blocks at the row
# end load 16 bytes past z and clear the excess, except on the volume's last row
SHAPES = [(1, 8, 8), (1, 1, 1), (1, 8, 9), (2, 13, 17), (1, 9, 9), (1, 8, 1), (3, 64, 64),
          (1, 100, 101), (2, 50, 1558), (1, 1024, 1024), (16, 512, 512),
          (1, 9, 10), (2, 16, 12), (1, 8, 2), (3, 17, 1030)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("kind", ["cssi", "uniform", "mixed"])
@pytest.mark.parametrize("mode", ["lossless", "b3-5-10-15_p0.5"])
def test_shapes(gpu, orc, shape, kind, mode):
    bins, p = MODES[mode]
    seed = (shape[0] * 31 + shape[1] * 7 + shape[2] + len(kind)) % 1000
    check_case(gpu, orc, synth(kind, shape, seed=seed), bins, p)


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("kind", ["sparse", "cssi", "xpcs", "mixed", "zeros", "max"])
def test_modes(gpu, orc, mode, kind):
    bins, p = MODES[mode]
    check_case(gpu, orc, synth(kind, (4, 136, 200), seed=7), bins, p)


def test_cfg1_sparse_lossless(gpu, orc):
    """16x512x512 Poisson(0.05) + hot pixels, lossless: exact round trip, bytes == oracle."""
    vol = synth("sparse", (16, 512, 512), seed=1)
    check_case(gpu, orc, vol, (1, 1, 1, 1), 1.0)


def test_long_lookback_chain(gpu, orc):
    """1024 tiles in one stream; the look-back must resolve across many probes."""
    vol = synth("xpcs", (64, 512, 512), seed=3)
    check_case(gpu, orc, vol, (1, 1, 1, 1), 1.0)
    check_case(gpu, orc, vol, (5, 5, 5, 5), 0.5)


@pytest.mark.parametrize("shape", [(2, 64, 96), (2, 21, 30), (3, 40, 67), (2, 17, 1557)],
                         ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("offset_elems", [1, 2, 8])
def test_misaligned_input(gpu, orc, offset_elems, shape):
    """Volumes at 2-, 4- and 16-byte offsets inside a larger buffer use other load
    paths; with z = 30 the row-end blocks are partial; with z odd (67, 1557)
    rows alternate between 4-byte and 2-byte alignment (shifted 16-byte loads,
    split stores).  Decoding leaves the bytes around the volume untouched."""
    import torch
    vol = synth("cssi", shape, seed=11)
    big = torch.zeros(vol.size + 16, dtype=torch.int16, device="cuda")
    big[offset_elems:offset_elems + vol.size] = to_dev(vol).reshape(-1)
    d = big[offset_elems:offset_elems + vol.size]
    for bins, p in (MODES["lossless"], MODES["b3-5-10-15_p0.5"]):
        out, n = gpu.compress(d, vol.shape, bins, p)
        st, ref, refn = orc.compress(vol, bins, p)
        assert n == refn and np.array_equal(out[:n].cpu().numpy(), ref)
        decbig = torch.full((vol.size + 16,), 7, dtype=torch.int16, device="cuda")
        dv = decbig[offset_elems:offset_elems + vol.size]
        gpu.decompress(out, n, vol.shape, bins, out=dv)
        _, refdec = orc.decompress(ref, vol.shape, bins)
        host = decbig.cpu().numpy().view(np.uint16)
        assert np.array_equal(host[offset_elems:offset_elems + vol.size].reshape(vol.shape), refdec)
        assert np.all(host[:offset_elems] == 7) and np.all(host[offset_elems + vol.size:] == 7)


def test_capacity_short(gpu, orc):
    """A short output buffer: error, required size reported, nothing written past capacity."""
    import torch
    vol = synth("uniform", (1, 64, 64), seed=5)
    st, ref, refn = orc.compress(vol, (1, 1, 1, 1), 1.0)
    cap = refn - 100
    buf = torch.full((refn + 64,), 0xAB, dtype=torch.uint8, device="cuda")
    with pytest.raises(gpu.LscompError) as ei:
        gpu.compress(to_dev(vol), vol.shape, (1, 1, 1, 1), 1.0, out=buf[:cap])
    assert ei.value.status == gpu.ERR_CAPACITY
    assert ei.value.required == refn
    assert np.all(buf[cap:].cpu().numpy() == 0xAB)
    # the exact size is enough
    out, n = gpu.compress(to_dev(vol), vol.shape, (1, 1, 1, 1), 1.0, out=buf[:refn])
    assert n == refn and np.array_equal(out[:n].cpu().numpy(), ref)


def test_corrupt_streams(gpu, orc):
    import torch
    vol = synth("cssi", (2, 64, 64), seed=9)
    dims = vol.shape
    st, ref, refn = orc.compress(vol, (1, 1, 1, 1), 1.0)
    d = torch.from_numpy(ref.copy()).cuda()
    with pytest.raises(gpu.LscompError) as ei:          # truncated payload
        gpu.decompress(d, refn - 8, dims)
    assert ei.value.status == gpu.ERR_CORRUPT
    bad = ref.copy()
    bad[3] = 0x1f                                        # rate 31 > 16
    with pytest.raises(gpu.LscompError) as ei:
        gpu.decompress(torch.from_numpy(bad).cuda(), refn, dims)
    assert ei.value.status == gpu.ERR_CORRUPT
    # the sticky error was cleared: a good stream decodes again
    dec = gpu.decompress(d, refn, dims)
    assert np.array_equal(dec.cpu().numpy().view(np.uint16), vol)


def test_repeated_launches_and_streams(gpu, orc):
    """Scratch reuse: many launches in a row and on two streams give identical bytes."""
    import torch
    vol = synth("xpcs", (8, 256, 256), seed=21)
    dims = vol.shape
    st, ref, refn = orc.compress(vol, (3, 5, 10, 15), 0.5)
    d = to_dev(vol)
    for _ in range(40):
        out, n = gpu.compress(d, dims, (3, 5, 10, 15), 0.5)
        assert n == refn
    assert np.array_equal(out[:n].cpu().numpy(), ref)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    outs = []
    for s in (s1, s2):
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            o = torch.empty(gpu.max_compressed_bytes(dims), dtype=torch.uint8, device="cuda")
            sz = torch.zeros(1, dtype=torch.int64, device="cuda")
            gpu.compress_async(d, dims, (3, 5, 10, 15), 0.5, o, sz, stream=s)
            outs.append((o, sz))
    torch.cuda.synchronize()
    for o, sz in outs:
        assert int(sz.item()) == refn
        assert np.array_equal(o[:refn].cpu().numpy(), ref)
    assert gpu.stream_status(s1) == gpu.OK and gpu.stream_status(s2) == gpu.OK


def test_decode_oracle_stream(gpu, orc):
    """The GPU decoder reads streams written by the oracle (not only its own)."""
    import torch
    vol = synth("mixed", (3, 40, 72), seed=4)
    for bins, p in MODES.values():
        st, ref, refn = orc.compress(vol, bins, p)
        dec = gpu.decompress(torch.from_numpy(ref.copy()).cuda(), refn, vol.shape, bins)
        _, refdec = orc.decompress(ref, vol.shape, bins)
        assert np.array_equal(dec.cpu().numpy().view(np.uint16), refdec)


@pytest.mark.parametrize("tile", [128, 256, 512])
@pytest.mark.parametrize("wg_per_cu", [0, 1])
def test_launch_shapes_give_identical_bytes(gpu, orc, tile, wg_per_cu):
    """The stream does not depend on the launch shape (tile size, persistent grid size).
    The 12 x 600 x 1558 volume (175,500 blocks) has more than 256 tiles at every
    tile size, so it also decodes through the offset-phase launch; the smaller
    one decodes with one workgroup per tile where it has at most 256."""
    try:
        for d in (gpu.COMPRESS, gpu.DECOMPRESS):
            gpu.set_launch_shape(d, tile, wg_per_cu)
        for shape in ((6, 300, 1558), (12, 600, 1558)):
            vol = synth("cssi", shape, seed=13)
            check_case(gpu, orc, vol, (1, 1, 1, 1), 1.0)
            check_case(gpu, orc, vol, (3, 5, 10, 15), 0.5)
    finally:
        for d in (gpu.COMPRESS, gpu.DECOMPRESS):
            gpu.set_launch_shape(d, 0, 0)


def test_autotune_applies_a_shape_and_keeps_bytes(gpu, orc):
    """Coordinate descent over tile, workgroups per CU, window and counters for
    both directions: every axis is measured, the applied shapes are ones it
    measured, and the streams stay the oracle's."""
    import torch
    vol = synth("xpcs", (8, 512, 512), seed=17)
    d = to_dev(vol)
    saved = [gpu.get_launch_shape_ex(k) for k in (gpu.COMPRESS, gpu.DECOMPRESS)]
    try:
        table = gpu.autotune(d, vol.shape, (1, 1, 1, 1), 1.0, reps=2)
        for k in (gpu.COMPRESS, gpu.DECOMPRESS):
            mine = [e for e in table if e["direction"] == k]
            assert all(e["ms"] > 0 for e in mine)
            assert {e["tile"] for e in mine} == {128, 256, 512}
            assert len({e["counters"] for e in mine}) == 5 and len({e["lb_window"] for e in mine}) >= 2
            cur = gpu.get_launch_shape_ex(k)
            assert any(all(e[f] == cur[f] for f in ("tile", "wg_per_cu", "lb_window", "counters")) for e in mine)
        assert gpu.autotune(d, vol.shape, (1, 1, 1, 1), 1.0) == []          # cached
        check_case(gpu, orc, vol, (1, 1, 1, 1), 1.0)
        check_case(gpu, orc, vol, (3, 5, 10, 15), 0.5)
    finally:
        for k, sh in zip((gpu.COMPRESS, gpu.DECOMPRESS), saved):
            gpu.set_launch_shape_ex(k, **sh)
    torch.cuda.synchronize()


@pytest.mark.parametrize("shape", [(128, 0, 128, 1), (256, 1, 256, 16), (512, 2, 512, 8), (512, 0, 64, 3)])
def test_every_window_and_counter_count_gives_identical_bytes(gpu, orc, shape):
    """The stream does not depend on the look-back window or the number of tile-id
    counters (the 12 x 600 x 1558 volume decodes through the offset-phase launch
    at every tile size)."""
    tile, w, win, nc = shape
    saved = [gpu.get_launch_shape_ex(k) for k in (gpu.COMPRESS, gpu.DECOMPRESS)]
    try:
        for k in (gpu.COMPRESS, gpu.DECOMPRESS):
            gpu.set_launch_shape_ex(k, tile, w, win, nc)
        for shape3 in ((5, 300, 1558), (12, 600, 1558)):
            vol = synth("cssi", shape3, seed=19)
            check_case(gpu, orc, vol, (1, 1, 1, 1), 1.0)
            check_case(gpu, orc, vol, (3, 5, 10, 15), 0.5)
    finally:
        for k, sh in zip((gpu.COMPRESS, gpu.DECOMPRESS), saved):
            gpu.set_launch_shape_ex(k, **sh)


def test_shape_changes_on_another_thread(gpu, orc):
    """A call reads the launch shape once (its tile count, scratch and grid all
    use that copy), so a shape changed by another thread mid-call (as the
    autotuner does) never leaves part of a volume unwritten."""
    import threading
    import torch
    vol = synth("cssi", (7, 200, 1558), seed=23)
    dims = vol.shape
    refs = {}
    for bins, p in ((1, 1, 1, 1), 1.0), ((3, 5, 10, 15), 0.5):
        st, ref, refn = orc.compress(vol, bins, p)
        refs[bins] = (p, ref, orc.decompress(ref, dims, bins)[1])
    d_in = torch.from_numpy(vol.view(np.int16)).cuda()
    saved = [gpu.get_launch_shape_ex(k) for k in (gpu.COMPRESS, gpu.DECOMPRESS)]
    stop = threading.Event()

    def flip():
        shapes = [(128, 0, 128, 2), (512, 0, 64, 4), (256, 1, 256, 1), (512, 2, 512, 16)]
        i = 0
        while not stop.is_set():
            for k in (gpu.COMPRESS, gpu.DECOMPRESS):
                gpu.set_launch_shape_ex(k, *shapes[i % len(shapes)])
            i += 1

    t = threading.Thread(target=flip)
    t.start()
    try:
        for it in range(24):
            bins = (1, 1, 1, 1) if it % 2 == 0 else (3, 5, 10, 15)
            p, ref, refdec = refs[bins]
            out, n = gpu.compress(d_in, dims, bins, p)
            assert n == ref.size and np.array_equal(out[:n].cpu().numpy(), ref), it
            dec = gpu.decompress(out, n, dims, bins).cpu().numpy().view(np.uint16)
            assert np.array_equal(dec, refdec), it
    finally:
        stop.set()
        t.join()
        for k, sh in zip((gpu.COMPRESS, gpu.DECOMPRESS), saved):
            gpu.set_launch_shape_ex(k, **sh)


@pytest.mark.parametrize("prepass", [0, 1, 2])
def test_decode_with_and_without_prepass(gpu, orc, prepass):
    """Decode tile offsets from the in-kernel look-back (0), the offset phase of
    the decode launch (1) or a separate pre-pass kernel (2): same volume."""
    old = gpu.set_decode_prepass(prepass)
    try:
        # the last two have more than 256 tiles (at most 256 decode with one
        # workgroup per tile in every mode)
        for shape, kind in (((3, 300, 1558), "cssi"), ((1, 8, 9), "uniform"), ((64, 128, 128), "xpcs"),
                            ((2, 2048, 2056), "mixed"), ((12, 600, 1558), "cssi")):
            vol = synth(kind, shape, seed=31)
            check_case(gpu, orc, vol, (1, 1, 1, 1), 1.0)
            check_case(gpu, orc, vol, (3, 5, 10, 15), 0.5)
    finally:
        gpu.set_decode_prepass(old)


@pytest.mark.parametrize("kind", ["uniform", "alternating", "sparse_dense"])
def test_decode_many_tiles_per_workgroup(gpu, orc, kind):
    """Decoding with one workgroup per CU, so each takes a dozen tiles or more:
    dense streams use the pipelined decode (payload staged beside the previous
    tile's in the LDS ring, wrapping round it; tiles too large to stage there
    are loaded after the barrier), sparse ones the other decode.  Lossless
    decodes equal the input; the lossy stream and decode equal the oracle's."""
    import torch
    rng = np.random.default_rng(41)
    dims = (48, 1024, 1000)   # 750k blocks, 1465 tiles of 512
    if kind == "uniform":          # 128 payload bytes per block: no tile fits beside another
        vol = rng.integers(0, 65536, size=dims, dtype=np.uint16)
    elif kind == "alternating":    # frames alternate incompressible and Poisson(0.3)
        vol = rng.poisson(0.3, size=dims).astype(np.uint16)
        vol[::2] = rng.integers(0, 65536, size=(dims[0] // 2,) + dims[1:], dtype=np.uint16)
    else:                          # CSSI-like frames: a few KB per tile, the ring wraps
        vol = synth("cssi", dims, seed=43)
    try:
        gpu.set_launch_shape(gpu.DECOMPRESS, 512, 1)
        d = to_dev(vol)
        out, n = gpu.compress(d, dims, (1, 1, 1, 1), 1.0)
        dec = gpu.decompress(out, n, dims, (1, 1, 1, 1))
        assert bool(torch.equal(dec, d))
        if kind == "sparse_dense":
            check_case(gpu, orc, vol, (3, 5, 10, 15), 0.5)
            check_case(gpu, orc, vol, (1, 1, 1, 1), 1.0)
    finally:
        gpu.set_launch_shape(gpu.DECOMPRESS, 0, 0)


@pytest.mark.parametrize("mode", ["lossless", "b3-5-10-15_p0.5"])
def test_decode_offset_chunks_beyond_the_grid(gpu, orc, mode):
    """The decode's offset phase takes chunks of 128 tiles; at tile 128 with one
    workgroup per CU a 280 x 1024 x 1024 volume (4.6 M blocks, 35840 tiles, 280
    chunks) has more chunks than a 256-CU grid has workgroups, so workgroups take
    further chunks after their first.  Lossless decodes equal the input, lossy
    ones the oracle's decode of the same stream, and the stream equals the
    default shape's."""
    import torch
    bins, p = MODES[mode]
    dims = (280, 1024, 1024)
    vol = synth("cssi", dims, seed=47)
    d = to_dev(vol)
    ref_out, ref_n = gpu.compress(d, dims, bins, p)
    try:
        for k in (gpu.COMPRESS, gpu.DECOMPRESS):
            gpu.set_launch_shape(k, 128, 1)
        out, n = gpu.compress(d, dims, bins, p)
        assert n == ref_n and bool(torch.equal(out[:n], ref_out[:ref_n]))
        dec = gpu.decompress(out, n, dims, bins)
        if tuple(bins) == (1, 1, 1, 1):
            assert bool(torch.equal(dec, d))
        else:
            _, refdec = orc.decompress(out[:n].cpu().numpy(), dims, bins)
            assert np.array_equal(dec.cpu().numpy().view(np.uint16), refdec)
    finally:
        for k in (gpu.COMPRESS, gpu.DECOMPRESS):
            gpu.set_launch_shape(k, 0, 0)
    torch.cuda.synchronize()


@pytest.mark.parametrize("shape", [(1, 4096, 2048), (1, 4096, 2056), (16, 512, 512), (1, 8, 9)],
                         ids=["256tiles", "257tiles", "128tiles", "1tile"])
@pytest.mark.parametrize("mode", ["lossless", "b3-5-10-15_p0.5"])
def test_small_volume_decode(gpu, orc, shape, mode):
    """Volumes of at most 256 tiles decode with one workgroup per tile that sums
    the header bytes before its tile itself; 257 tiles take the launch with the
    offset phase.  Both sides of the limit equal the oracle's stream and decode."""
    bins, p = MODES[mode]
    check_case(gpu, orc, synth("cssi", shape, seed=61 + shape[1] % 7), bins, p)
