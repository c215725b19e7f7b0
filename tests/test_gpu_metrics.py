"""GPU error metrics against a float64 numpy restatement of the reference's
computePSNR definitions (lsCOMP/src/cuLSZ_utility.cu:887-946)."""
import numpy as np
import pytest

from lsc_helpers import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,offset", [(1, 0), (7, 0), (4096, 0), (1000003, 0), (1000003, 1)])
def test_metrics_vs_numpy(gpu, n, offset):
    import torch
    rng = np.random.default_rng(n)
    a = rng.integers(0, 65536, n + 1).astype(np.uint16)[offset:offset + n]
    d = np.clip(a.astype(np.int64) + rng.integers(-40, 41, n), 0, 65535).astype(np.uint16)
    ta = torch.from_numpy(np.concatenate([np.zeros(offset, np.uint16), a]).view(np.int16)).cuda()[offset:]
    td = torch.from_numpy(np.concatenate([np.zeros(offset, np.uint16), d]).view(np.int16)).cuda()[offset:]
    s = gpu.error_stats(ta, td)
    a64, d64 = a.astype(np.float64), d.astype(np.float64)
    err = np.abs(d64 - a64)
    assert s["n"] == n and s["max_abs_err"] == int(err.max())
    assert s["sum_sq_err"] == int((err.astype(np.int64) ** 2).sum())
    assert s["min_orig"] == int(a.min()) and s["max_orig"] == int(a.max())
    mse = (err ** 2).mean()
    rng_ = a64.max() - a64.min()
    if mse > 0 and rng_ > 0:
        assert s["psnr"] == pytest.approx(20 * np.log10(rng_) - 10 * np.log10(mse), rel=1e-12)
        assert s["nrmse"] == pytest.approx(np.sqrt(mse) / rng_, rel=1e-12)
    nz = a64 != 0
    if nz.any():
        assert s["max_rel_err"] == pytest.approx((err[nz] / a64[nz]).max(), rel=1e-6)
    # normErr = sqrt(sum err^2), normErr_norm = normErr / sqrt(sum dec^2)
    # (cuLSZ_utility.cu:900, 935-936: sum22 accumulates the decoded values)
    norm = np.sqrt((err ** 2).sum())
    assert s["norm_err"] == pytest.approx(norm, rel=1e-12)
    dsq = (d64 ** 2).sum()
    if dsq > 0:
        assert s["norm_err_norm"] == pytest.approx(norm / np.sqrt(dsq), rel=1e-12)


def test_metrics_of_a_lossy_round_trip(gpu):
    import torch
    vol = synth("cssi", (4, 256, 520), seed=3)
    d = torch.from_numpy(vol.view(np.int16)).cuda()
    out, n = gpu.compress(d, vol.shape, (1, 1, 1, 1), 1.0)
    s = gpu.error_stats(d, gpu.decompress(out, n, vol.shape))
    assert s["max_abs_err"] == 0 and s["sum_sq_err"] == 0
    out, n = gpu.compress(d, vol.shape, (3, 5, 10, 15), 1.0)        # no pooling: error < bin
    s = gpu.error_stats(d, gpu.decompress(out, n, vol.shape, (3, 5, 10, 15)))
    assert 0 < s["max_abs_err"] <= 14 and np.isfinite(s["psnr"])


@pytest.mark.gpu
@pytest.mark.parametrize("nbytes", [16, 1000, 4096 + 7, 3 * 1024 * 1024 + 5])
def test_copy_bytes(gpu, nbytes):
    """The calibration copy reproduces its source exactly, tail bytes included,
    and leaves the bytes after the destination alone."""
    import torch
    src = torch.randint(0, 256, (nbytes,), dtype=torch.uint8, device="cuda")
    dst = torch.full((nbytes + 32,), 7, dtype=torch.uint8, device="cuda")
    gpu.copy_bytes(dst[:nbytes], src)
    torch.cuda.synchronize()
    assert torch.equal(dst[:nbytes], src)
    assert bool((dst[nbytes:] == 7).all())
    with pytest.raises(ValueError):
        gpu.copy_bytes(dst[:nbytes - 1], src)
    with pytest.raises(gpu.LscompError):
        gpu.copy_bytes(dst[1:nbytes + 1], src)   # misaligned destination


# ---- SSIM (cuLSZ_utility.cu:757-886) against the oracle's restatement --------

def _noisy(a, rng, amp):
    return np.clip(a.astype(np.int64) + rng.integers(-amp, amp + 1, a.shape), 0, 65535).astype(np.uint16)


def _dev(v):
    import torch
    return torch.from_numpy(np.ascontiguousarray(v).view(np.int16)).cuda()


# one window; tiles of 64 x 4 windows along z, y and chunks of 64 along x,
# with odd and even remainders; a shape whose windows end one short of z
@pytest.mark.parametrize("shape", [(7, 7, 7), (9, 20, 140), (20, 33, 300), (140, 9, 9), (8, 14, 8), (11, 7, 270)],
                         ids=lambda s: "x".join(map(str, s)))
def test_ssim_vs_oracle(gpu, orc, shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.integers(0, 65536, shape).astype(np.uint16)
    b = _noisy(a, rng, 2000)
    got, want = gpu.ssim(_dev(a), _dev(b), shape), orc.ssim(a, b)
    assert got == pytest.approx(want, rel=1e-10, abs=1e-13)


def test_ssim_of_lossy_round_trips(gpu, orc):
    """CSSI- and XPCS-like frames through the codec: sparse data, windows whose
    original is constant (range 0), and the decoded volume's quantization."""
    for kind, shape, bins, p in (("cssi", (12, 200, 260), (3, 5, 10, 15), 0.5),
                                 ("xpcs", (9, 64, 520), (3, 3, 3, 3), 0.5),
                                 ("cssi", (10, 90, 77), (10, 10, 10, 10), 1.0)):
        vol = synth(kind, shape, seed=9)
        d = _dev(vol)
        out, n = gpu.compress(d, shape, bins, p)
        dec = gpu.decompress(out, n, shape, bins)
        got = gpu.ssim(d, dec, shape)
        want = orc.ssim(vol, dec.cpu().numpy().view(np.uint16))
        assert got == pytest.approx(want, rel=1e-10, abs=1e-13), kind


def test_ssim_medium_volume(gpu, orc):
    """40 x 300 x 400: 492 k windows, many workgroups and chunks."""
    shape = (40, 300, 400)
    rng = np.random.default_rng(40)
    a = synth("xpcs", shape, seed=40)
    b = _noisy(a, rng, 3)
    got, want = gpu.ssim(_dev(a), _dev(b), shape), orc.ssim(a, b)
    assert got == pytest.approx(want, rel=1e-10, abs=1e-13)


def test_ssim_identity_at_full_frame_size(gpu):
    """SSIM(v, v) = 1 and a lossless round trip gives 1, on 64 CSSI frames of
    1813 x 1558 (20 M windows)."""
    shape = (64, 1813, 1558)
    vol = synth("cssi", shape, seed=2)
    d = _dev(vol)
    assert abs(gpu.ssim(d, d, shape) - 1.0) < 1e-12
    out, n = gpu.compress(d, shape, (1, 1, 1, 1), 1.0)
    assert abs(gpu.ssim(d, gpu.decompress(out, n, shape), shape) - 1.0) < 1e-12


def test_ssim_rejects_volumes_without_a_window(gpu):
    import lscomp
    v = _dev(np.zeros((6, 10, 10), np.uint16))
    with pytest.raises(lscomp.LscompError):
        gpu.ssim(v, v, (6, 10, 10))
