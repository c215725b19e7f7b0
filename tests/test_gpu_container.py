"""Frames compressed on the GPU (uint16 and uint32), stored in one container and
decompressed from it: every frame comes back as compressed."""
import numpy as np
import pytest

from lsc_helpers import synth

pytestmark = pytest.mark.gpu


def to_dev(vol):
    import torch
    return torch.from_numpy(np.ascontiguousarray(vol).view(np.int16)).cuda()


def test_container_of_gpu_streams(gpu, orc):
    import torch
    entries, vols = [], []
    for i, shape in enumerate([(1, 256, 256), (1, 512, 384), (2, 100, 130)]):
        vol = synth("cssi", shape, seed=20 + i)
        bins, p = ((1, 1, 1, 1), 1.0) if i % 2 == 0 else ((3, 5, 10, 15), 0.5)
        out, n = gpu.compress(to_dev(vol), shape, bins, p)
        entries.append({"stream": out, "stream_bytes": n, "dims": shape, "elem_bytes": 2, "bins": bins,
                        "pooling": p, "tag": i})
        vols.append((vol, bins, p))
    v32 = synth("cssi", (2, 64, 96), seed=30).astype(np.uint32) * 3
    out, n = gpu.compress_u32(torch.from_numpy(v32.view(np.int32)).cuda(), v32.shape)
    entries.append({"stream": out, "stream_bytes": n, "dims": v32.shape, "elem_bytes": 4, "bins": (1, 1, 1, 1),
                    "pooling": 1.0, "tag": 3})
    buf = gpu.container_pack(entries)
    decs = gpu.container_decompress(buf)
    assert len(decs) == 4
    for (vol, bins, p), d in zip(vols, decs):
        st, ref, _ = orc.compress(vol, bins, p)
        _, refdec = orc.decompress(ref, vol.shape, bins)
        assert np.array_equal(d.cpu().numpy().view(np.uint16), refdec)
    assert np.array_equal(decs[3].cpu().numpy().view(np.uint32), v32)
