"""Served form on the GPU: continuous batching gives every request exactly the
stream of a plain call on its frame alone (and the oracle's), also when a
small HBM budget makes admission wait for releases, and under a replayed
Poisson arrival trace."""
import ctypes
import time

import numpy as np
import pytest

from lsc_helpers import synth

pytestmark = pytest.mark.gpu

SIDES = (64, 136, 256, 520)


def _d2h(ptr, nbytes):
    """Copy device bytes at a raw pointer to the host (plumbing for arena streams)."""
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    host = np.empty(nbytes, np.uint8)
    assert hip.hipMemcpy(host.ctypes.data, ptr, nbytes, 2) == 0   # hipMemcpyDeviceToHost
    return host


def _requests(gpu, seed, n):
    import torch
    rng = np.random.default_rng(seed)
    frames, reqs, params = [], [], []
    for i in range(n):
        y, z = int(rng.choice(SIDES)), int(rng.choice(SIDES))
        vol = synth("cssi", (1, y, z), seed=seed * 100 + i)
        d = torch.from_numpy(vol.view(np.int16)).cuda()
        b, p = ((1, 1, 1, 1), 1.0) if i % 2 == 0 else ((3, 5, 10, 15), 0.5)
        frames.append((vol, d))
        params.append((b, p))
        reqs.append(gpu.Server.request(d, vol.shape, b, p, tag=i))
    return frames, reqs, params


def _check_one(gpu, orc, frame, param, c, plain=True):
    vol, d = frame
    b, p = param
    assert c.status == gpu.OK
    st, ref, refn = orc.compress(vol, b, p)
    assert c.stream_bytes == refn
    assert np.array_equal(_d2h(c.d_stream, refn), ref)
    if plain:
        out, n = gpu.compress(d, vol.shape, b, p)
        assert n == refn and np.array_equal(out[:n].cpu().numpy(), ref)
    assert c.t_submit <= c.t_admit <= c.t_launch <= c.t_done


def _serve_all(srv, reqs, on_done, timeout=60.0):
    for r in reqs:
        srv.submit(r)
    done, t0 = 0, time.time()
    while done < len(reqs):
        assert time.time() - t0 < timeout, f"{done} of {len(reqs)} finished"
        for c in srv.poll():
            on_done(c)
            srv.release(c.id)
            done += 1


def test_mixed_requests_match_alone(gpu, orc):
    frames, reqs, params = _requests(gpu, 1, 40)
    srv = gpu.Server(slots=3, max_items=16, max_blocks=1 << 16, budget_bytes=1 << 30)
    _serve_all(srv, reqs, lambda c: _check_one(gpu, orc, frames[c.tag], params[c.tag], c))
    assert srv.drain() == gpu.OK
    srv.close()


def test_small_budget_admission(gpu, orc):
    """Room for about two requests: the rest are admitted as outputs are released."""
    frames, reqs, params = _requests(gpu, 2, 16)
    biggest = max(gpu.max_compressed_bytes(v.shape) for v, _ in frames)
    srv = gpu.Server(slots=2, max_items=8, max_blocks=1 << 16, budget_bytes=2 * biggest + 1024)
    admits = []

    def on_done(c):
        _check_one(gpu, orc, frames[c.tag], params[c.tag], c, plain=False)
        admits.append((c.t_admit, c.t_submit))

    _serve_all(srv, reqs, on_done)
    # with room for two, some requests must have waited for a release
    assert sum(1 for a, s in admits if a - s > 0) > 0
    srv.close()


def test_replay_poisson(gpu, orc):
    frames, reqs, params = _requests(gpu, 3, 60)
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1 / 5000.0, len(reqs)))
    srv = gpu.Server(slots=4, max_items=32, max_blocks=1 << 16, budget_bytes=1 << 30)
    out = srv.replay(reqs, arrivals)
    for i, c in enumerate(out):
        assert c.tag == i and c.status == gpu.OK
        vol, _ = frames[i]
        st, ref, refn = orc.compress(vol, *params[i])
        assert c.stream_bytes == refn
    assert all(c.t_done > c.t_submit for c in out)
    srv.close()
