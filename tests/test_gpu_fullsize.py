"""Full-size parity at the benchmark shapes.

CSSI-shaped 600x1813x1558 (lossless, then bins 3 5 10 15 / pooling 0.5) and
XPCS-shaped 512x1024x1024 (lossless and bins 3/5/10 x4, pooling 0.5): the
whole GPU stream equals the OpenMP oracle's byte for byte, the GPU decode
equals the input (lossless) or the oracle's decode (lossy), and the batched
form over 512 uneven frames equals the oracle per frame.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _full_case(gpu, orc, vol, modes):
    import torch
    dims = vol.shape
    d_in = torch.from_numpy(vol.view(np.int16)).cuda()
    out = torch.empty(gpu.max_compressed_bytes(dims), dtype=torch.uint8, device="cuda")
    dec = torch.empty(dims, dtype=torch.int16, device="cuda")
    for bins, p in modes:
        print(f"full-size case dims={dims} bins={bins} p={p}", flush=True)
        _, n = gpu.compress(d_in, dims, bins, p, out=out)
        st, ref, refn = orc.compress(vol, bins, p, omp=True)
        assert st == 0 and n == refn, (n, refn)
        got = out[:n].cpu().numpy()
        assert np.array_equal(got, ref), f"stream differs at {np.count_nonzero(got != ref)} bytes"
        del got
        gpu.decompress(out, n, dims, bins, out=dec)
        if tuple(bins) == (1, 1, 1, 1) and p >= 1.0:
            assert bool(torch.equal(dec, d_in))
        else:
            _, refdec = orc.decompress(ref, dims, bins, omp=True)
            assert np.array_equal(dec.cpu().numpy().view(np.uint16), refdec)
            del refdec
        del ref


def test_cssi_600x1813x1558(gpu, orc):
    from lscomp import datagen
    vol = datagen.volume("cssi", 2, (600, 1813, 1558))
    _full_case(gpu, orc, vol, [((1, 1, 1, 1), 1.0), ((3, 5, 10, 15), 0.5)])


def test_xpcs_512x1024x1024(gpu, orc):
    from lscomp import datagen
    vol = datagen.volume("xpcs", 3, (512, 1024, 1024))
    _full_case(gpu, orc, vol, [((1, 1, 1, 1), 1.0), ((3, 3, 3, 3), 0.5), ((5, 5, 5, 5), 0.5),
                               ((10, 10, 10, 10), 0.5)])


def test_batched_512_uneven_frames(gpu, orc):
    import torch
    from lscomp import datagen
    rng = np.random.default_rng(4)
    sides = rng.choice([256, 512, 768, 1024, 1536, 2048], size=512)
    vols = [datagen.volume("xpcs", 4, (1, int(s), int(s)), first_frame=i) for i, s in enumerate(sides)]
    dev = [torch.from_numpy(v.view(np.int16)).cuda() for v in vols]
    outs = [torch.empty(gpu.max_compressed_bytes(v.shape), dtype=torch.uint8, device="cuda") for v in vols]
    plan = gpu.Plan(gpu.COMPRESS, len(vols), sum(gpu.num_blocks(v.shape) for v in vols))
    plan.set_items([(d, v.shape, (1, 1, 1, 1), 1.0, o, o.numel()) for d, v, o in zip(dev, vols, outs)])
    plan.launch(graph=True)
    sizes = plan.sizes()
    assert plan.status() == gpu.OK
    for v, o, n in zip(vols, outs, sizes):
        st, ref, refn = orc.compress(v, (1, 1, 1, 1), 1.0)
        assert n == refn and np.array_equal(o[:n].cpu().numpy(), ref)
    plan.close()
