"""Served form on the GPU, both engines (the device-resident request loop and
the host-scheduled batches): every request gets exactly the stream of a plain
call on its frame alone and the oracle's, in a mix of sizes and settings,
when a small HBM budget makes admission wait for releases, under replayed
Poisson arrivals, and on the 512^2..2048^2 CSSI frame mix with a 64 GB budget.

References (oracle and plain calls) are computed before a server starts: the
device engine's persistent kernel holds most of each CU's LDS while an
instance runs, so a plain launch would wait for the instance to go idle."""
import ctypes
import time

import numpy as np
import pytest

from lsc_helpers import synth

pytestmark = pytest.mark.gpu

SIDES = (64, 136, 256, 520)
ENGINES = ("device", "host")


def _d2h(ptr, nbytes):
    """Copy device bytes at a raw pointer to the host (plumbing for arena streams)."""
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    host = np.empty(nbytes, np.uint8)
    assert hip.hipMemcpy(host.ctypes.data, ptr, nbytes, 2) == 0   # hipMemcpyDeviceToHost
    return host


def _server(gpu, engine, max_blocks, budget, idle_us=2000):
    if engine == "device":
        return gpu.Server.device(max_blocks, budget, idle_us)
    return gpu.Server(slots=3, max_items=16, max_blocks=max(max_blocks, 1 << 16), budget_bytes=budget)


class Refs:
    """Frames on the device with their oracle streams, checked against plain calls once."""

    def __init__(self, gpu, orc):
        self.gpu, self.orc, self.frames, self.ref = gpu, orc, [], {}

    def add(self, vol):
        import torch
        self.frames.append((vol, torch.from_numpy(vol.view(np.int16)).cuda()))
        return len(self.frames) - 1

    def expect(self, k, bins, p):
        key = (k, bins, p)
        if key not in self.ref:
            vol, d = self.frames[k]
            st, ref, refn = self.orc.compress(vol, bins, p)
            assert st == 0
            out, n = self.gpu.compress(d, vol.shape, bins, p)
            assert n == refn and np.array_equal(out[:n].cpu().numpy(), ref), "plain call differs from the oracle"
            self.ref[key] = ref
        return self.ref[key]

    def request(self, k, bins, p, tag):
        vol, d = self.frames[k]
        return self.gpu.Server.request(d, vol.shape, bins, p, tag=tag)


def _mixed(gpu, orc, seed, n):
    rng = np.random.default_rng(seed)
    refs, work = Refs(gpu, orc), []
    for i in range(n):
        y, z = int(rng.choice(SIDES)), int(rng.choice(SIDES))
        k = refs.add(synth("cssi", (1, y, z), seed=seed * 100 + i))
        b, p = ((1, 1, 1, 1), 1.0) if i % 2 == 0 else ((3, 5, 10, 15), 0.5)
        refs.expect(k, b, p)
        work.append((k, b, p))
    return refs, work


def _check(gpu, refs, work, c):
    k, b, p = work[c.tag]
    ref = refs.expect(k, b, p)
    assert c.status == gpu.OK, c.status
    assert c.stream_bytes == len(ref)
    assert np.array_equal(_d2h(c.d_stream, len(ref)), ref)
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


@pytest.mark.parametrize("engine", ENGINES)
def test_mixed_requests_match_alone(gpu, orc, engine):
    refs, work = _mixed(gpu, orc, 1, 40)
    reqs = [refs.request(k, b, p, i) for i, (k, b, p) in enumerate(work)]
    srv = _server(gpu, engine, 1 << 16, 1 << 30)
    _serve_all(srv, reqs, lambda c: _check(gpu, refs, work, c))
    assert srv.drain() == gpu.OK
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["device_loop"] == (engine == "device")
    srv.close()


@pytest.mark.parametrize("engine", ENGINES)
def test_small_budget_admission(gpu, orc, engine):
    """Room for about two requests: the rest are admitted as outputs are released,
    so arena ranges are rewritten by later requests (on the device engine by
    workgroups of other XCDs) and must still hold exactly their own stream."""
    refs, work = _mixed(gpu, orc, 2, 16)
    reqs = [refs.request(k, b, p, i) for i, (k, b, p) in enumerate(work)]
    biggest = max(gpu.max_compressed_bytes(v.shape) for v, _ in refs.frames)
    srv = _server(gpu, engine, 1 << 16, 2 * biggest + 1024)
    admits = []

    def on_done(c):
        _check(gpu, refs, work, c)
        admits.append((c.t_admit, c.t_submit))

    _serve_all(srv, reqs, on_done)
    assert sum(1 for a, s in admits if a - s > 0) > 0   # some requests waited for a release
    srv.close()


@pytest.mark.parametrize("engine", ENGINES)
def test_replay_poisson(gpu, orc, engine):
    """Replayed Poisson arrivals; outputs held, so every stream is byte-checked."""
    refs, work = _mixed(gpu, orc, 3, 60)
    reqs = [refs.request(k, b, p, i) for i, (k, b, p) in enumerate(work)]
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1 / 5000.0, len(reqs)))
    srv = _server(gpu, engine, 1 << 16, 1 << 30)
    out = srv.replay(reqs, arrivals, hold=True)
    for i, c in enumerate(out):
        assert c.tag == i
        _check(gpu, refs, work, c)
        srv.release(c.id)
    srv.close()


@pytest.mark.parametrize("engine", ENGINES)
def test_cssi_frame_mix_64gb(gpu, orc, engine):
    """The served workload: CSSI-style frames of 512^2 to 2048^2 pixels, half
    lossless and half bins 3 5 10 15 / pooling 0.5, Poisson arrivals at 20k
    frames/s, a 64 GB HBM budget.  Every completion is byte-equal to the
    oracle's stream and to a plain call on the frame alone."""
    from lscomp import datagen
    sides = (512, 768, 1024, 1536, 2048)
    refs = Refs(gpu, orc)
    keys = [refs.add(datagen.volume("cssi", 5, (1, s, s), first_frame=i)) for i, s in enumerate(sides * 2)]
    rng = np.random.default_rng(55)
    work = []
    for i in range(160):
        k = keys[int(rng.integers(0, len(keys)))]
        b, p = ((1, 1, 1, 1), 1.0) if i % 2 == 0 else ((3, 5, 10, 15), 0.5)
        refs.expect(k, b, p)
        work.append((k, b, p))
    reqs = [refs.request(k, b, p, i) for i, (k, b, p) in enumerate(work)]
    arrivals = np.cumsum(rng.exponential(1 / 20000.0, len(reqs)))
    srv = _server(gpu, engine, 1 << 17, 64 << 30) if engine == "device" else \
        gpu.Server(slots=4, max_items=256, max_blocks=1 << 22, budget_bytes=64 << 30)
    out = srv.replay(reqs, arrivals, hold=True)
    for i, c in enumerate(out):
        assert c.tag == i
        _check(gpu, refs, work, c)
        srv.release(c.id)
    assert srv.drain() == gpu.OK
    srv.close()


@pytest.mark.parametrize("engine", ENGINES)
def test_request_larger_than_budget_is_refused(gpu, engine):
    """A request whose worst case exceeds the whole budget is refused at submit
    (it could never be admitted and would hold up every later request)."""
    import torch
    d = torch.zeros((1, 256, 256), dtype=torch.int16, device="cuda")
    need = gpu.max_compressed_bytes((1, 256, 256))
    srv = _server(gpu, engine, 1 << 16, need // 2)
    with pytest.raises(gpu.LscompError) as e:
        srv.submit(gpu.Server.request(d, (1, 256, 256)))
    assert e.value.status == gpu.ERR_CAPACITY
    small = torch.zeros((1, 64, 64), dtype=torch.int16, device="cuda")
    rid = srv.submit(gpu.Server.request(small, (1, 64, 64)))   # still serves what fits
    assert srv.drain() == gpu.OK
    got = []
    while not got:
        got = srv.poll()
    assert got[0].id == rid and got[0].status == gpu.OK
    srv.close()


def test_device_loop_instances(gpu, orc):
    """The device loop's kernel instance ends after its idle time and the next
    request starts a new one; while requests keep arriving one instance serves
    them all (far fewer launches than requests)."""
    refs, work = _mixed(gpu, orc, 4, 48)
    reqs = [refs.request(k, b, p, i) for i, (k, b, p) in enumerate(work)]
    srv = gpu.Server.device(1 << 16, 1 << 30, idle_us=300)
    _serve_all(srv, reqs[:24], lambda c: _check(gpu, refs, work, c))
    first = srv.stats()["launches"]
    assert 1 <= first <= 4
    time.sleep(0.05)                       # the instance has gone idle and ended
    _serve_all(srv, reqs[24:], lambda c: _check(gpu, refs, work, c))
    st = srv.stats()
    assert st["launches"] > first and st["completed"] == len(reqs)
    assert srv.drain() == gpu.OK
    srv.close()
