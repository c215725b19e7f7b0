"""The compression entry point that allocates nothing
(lscomp_compress_u16_noalloc_async): its streams and sizes are byte-equal to
lscomp_compress_u16_async's on the same inputs, over calls in a row on the
one per-GPU scratch array (a large call, then a smaller one, which must not
see the first call's look-back words), and a volume beyond the array's limit
is refused before anything is launched."""
import numpy as np
import pytest

from lsc_helpers import synth

pytestmark = pytest.mark.gpu

LOSSY = ((3, 5, 10, 15), 0.5)
LOSSLESS = ((1, 1, 1, 1), 1.0)
NOALLOC_BYTES = 4 << 20


def _both(gpu, vol, dims, bins, p):
    """(stream, size) from the allocating entry point and from the no-alloc one."""
    import torch
    d_in = torch.from_numpy(vol.view(np.int16)).cuda()
    cap = gpu.max_compressed_bytes(dims)
    outs, sizes = [], []
    for fn in (gpu.compress_async, gpu.compress_noalloc_async):
        out = torch.full((cap,), 0xA5, dtype=torch.uint8, device="cuda")
        d_size = torch.zeros(1, dtype=torch.int64, device="cuda")
        fn(d_in, dims, bins, p, out, d_size)
        torch.cuda.synchronize()
        n = int(d_size.item())
        assert 0 < n <= cap
        outs.append(out[:n].cpu().numpy())
        sizes.append(n)
    assert gpu.stream_status() == gpu.OK
    return outs, sizes


@pytest.mark.parametrize("order", ["large_then_small", "small_then_large_then_small"])
def test_noalloc_matches_async(gpu, orc, order):
    big = ((40, 600, 700), LOSSY, 11)
    small = ((3, 77, 130), LOSSLESS, 12)
    calls = [big, small] if order == "large_then_small" else [small, big, small]
    for dims, (bins, p), seed in calls:
        print(f"noalloc call dims={dims} bins={bins} p={p}", flush=True)
        vol = synth("cssi", dims, seed=seed)
        (ref, got), (nref, ngot) = _both(gpu, vol, dims, bins, p)
        assert ngot == nref
        assert np.array_equal(got, ref)
        if vol.size < 100000:
            st, oref, orefn = orc.compress(vol, bins, p)
            assert st == 0 and orefn == ngot and np.array_equal(got, oref)


def test_noalloc_every_shape(gpu):
    """Tile sizes, counters and windows all share the one array."""
    dims = (6, 300, 260)
    vol = synth("xpcs", dims, seed=13)
    saved = gpu.get_launch_shape_ex(0)
    try:
        for shape in ((128, 0, 128, 8), (256, 1, 64, 1), (512, 0, 512, 16), (512, 0, 64, 4)):
            gpu.set_launch_shape_ex(0, *shape)
            (ref, got), (nref, ngot) = _both(gpu, vol, dims, *LOSSY)
            assert ngot == nref and np.array_equal(got, ref), shape
    finally:
        gpu.set_launch_shape_ex(0, *(saved[k] for k in ("tile", "wg_per_cu", "lb_window", "counters")))


def test_noalloc_refuses_beyond_limit(gpu):
    """More tiles than the array holds: LSCOMP_ERR_TOO_LARGE, nothing written.
    The buffers are real and the output capacity covers the header section, so
    the call would stay in bounds even if it were launched."""
    import torch
    saved = gpu.get_launch_shape_ex(0)
    max_tiles = (NOALLOC_BYTES - 8192) // 8
    dims = (1, 65536, 65536)                   # 2^26 blocks
    try:
        gpu.set_launch_shape_ex(0, 128)
        assert gpu.num_blocks(dims) > max_tiles * 128
        vol = torch.empty(dims, dtype=torch.int16, device="cuda")
        cap = gpu.header_bytes(dims) + 4096
        out = torch.full((cap,), 0x5A, dtype=torch.uint8, device="cuda")
        d_size = torch.full((1,), 7, dtype=torch.int64, device="cuda")
        with pytest.raises(gpu.LscompError) as ei:
            gpu.compress_noalloc_async(vol, dims, *LOSSLESS, out, d_size)
        assert ei.value.status == gpu.ERR_TOO_LARGE
        torch.cuda.synchronize()
        assert int(d_size.item()) == 7 and bool((out == 0x5A).all())
        del vol
    finally:
        gpu.set_launch_shape_ex(0, *(saved[k] for k in ("tile", "wg_per_cu", "lb_window", "counters")))
    torch.cuda.empty_cache()
