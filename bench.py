#!/usr/bin/env python3
"""Benchmark of the gfx950 lsCOMP uint16 codec on one MI355X.

Workload (the headline `value`): the CSSI-shaped volume 600 x 1813 x 1558
uint16 (3.39 GB), Poisson(0.3) background + 50 Gaussian peaks per frame + gap
stripes, seed 2, generated on the host and resident in HBM before timing.  A
step is one compression of the whole volume (lossless: bins 1 1 1 1,
pooling 1); `value` is its throughput in GB/s (1e9 bytes of uncompressed input
per second).  The same K steps are timed for decompression and for the lossy
setting (bins 3 5 10 15, pooling 0.5) and reported in `modes`, with the
XPCS-shaped 512 x 1024 x 1024 volume (lossless and bins 3/5/10 x4, pooling 0.5),
the small 16 x 512 x 512 volume (checked byte for byte against the oracle) and
the uint32 path.  `forms` holds the batched form (512 uneven frames, one
autotuned plan replayed from a hipGraph, against 512 plain launches) and the
served form (512^2..2048^2 CSSI frames, half lossless and half lossy, under a
64 GB budget: the device-resident loop at the offered 20k frames/s and at
saturation, beside the host-scheduled engine and plain per-frame launches).

Each mode also carries the end-to-end time of one blocking call (the stream size
or status read back on the host, as the reference's CLI times its calls; median
and mean of `--steps` calls) and that rate in the reference's unit (input MiB per
ms).

Prints ONE JSON line.  `--gpus` above 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sc25-lscomp_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
CSSI_DIMS = (600, 1813, 1558)
XPCS_DIMS = (512, 1024, 1024)
SMALL_DIMS = (16, 512, 512)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


_SPIN_PER_MS = []


def gpu_busy(torch, ms=1.0):
    """Keep the stream busy for about `ms` (torch's spin kernel) so that the launches
    the host enqueues next queue up behind it: an event pair around a launch then
    measures the kernel, not a pause of the host before the launch was issued.  The
    spin's unit is calibrated once (its counter's rate depends on the platform)."""
    try:
        if not _SPIN_PER_MS:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            torch.cuda._sleep(100000)
            torch.cuda.synchronize()
            _SPIN_PER_MS.append(100000 / max((time.perf_counter() - t0) * 1e3, 1e-3))
        torch.cuda._sleep(max(1, int(ms * _SPIN_PER_MS[0])))
    except (AttributeError, RuntimeError):
        pass


def timed_loop(torch, fn, steps):
    """Wall-clock ms per launch of fn() over `steps` launches on the current stream
    (host enqueue included), then the mean and per-launch kernel times from event
    pairs around `steps` more launches queued behind a spin kernel."""
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    # torch creates a HIP event at its first record; a creation that stalls the
    # host while the stream has run dry would land between an event pair, so
    # every event is created (recorded once) before the timed launches
    for a, b in evs:
        a.record(stream)
        b.record(stream)
    torch.cuda.synchronize()
    gpu_busy(torch)
    for a, b in evs:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    per = [a.elapsed_time(b) for a, b in evs]
    return wall, sum(per) / len(per), per


def e2e_ms(torch, fn, steps, median=False):
    """Wall-clock ms of `steps` blocking calls (each returns after its stream size
    or status has reached the host): their mean, or with `median` the (median,
    mean) of the per-call times (one stalled call moves the mean, not the median)."""
    fn()
    torch.cuda.synchronize()
    per = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        fn()
        per.append((time.perf_counter() - t) * 1e3)
    mean = (time.perf_counter() - t0) * 1e3 / steps
    return (statistics.median(per), mean) if median else mean


def profiled_traffic(dims):
    """HBM bytes per compression launch from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE of the compress kernel, scaled by the same counters
    on a plain copy of known size); None when no summary matches this workload."""
    path = os.path.join(ROOT, "profiles", "compress_cssi_600x1813x1558_counters.json")
    try:
        with open(path) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None
    if list(prof.get("dims", [])) != list(dims):
        return None
    return prof.get("traffic_bytes")


def time_codec(torch, lscomp, d_in, dims, bins, pool, d_out, d_dec, d_size, args, in_bytes, check_lossless):
    """Compress and decompress timing of one setting (K steps each after W warm-ups)."""
    comp = lambda: lscomp.compress_async(d_in, dims, bins, pool, d_out, d_size)
    for _ in range(args.warmup):
        comp()
    torch.cuda.synchronize()
    c_wall, c_ev, c_per = timed_loop(torch, comp, args.steps)
    s_bytes = int(d_size.item())
    dec = lambda: lscomp.decompress_async(d_out, s_bytes, dims, bins, d_dec)
    for _ in range(args.warmup):
        dec()
    torch.cuda.synchronize()
    d_wall, d_ev, d_per = timed_loop(torch, dec, args.steps)
    ok = lscomp.stream_status() == lscomp.OK
    res = {
        "compress_GBps": in_bytes / (c_wall * 1e-3) / 1e9,
        "decompress_GBps": in_bytes / (d_wall * 1e-3) / 1e9,
        "compress_ms": c_wall, "decompress_ms": d_wall,
        "compress_kernel_ms": c_ev, "decompress_kernel_ms": d_ev,
        "compress_kernel_ms_min_max": [min(c_per), max(c_per)],
        "decompress_kernel_ms_min_max": [min(d_per), max(d_per)],
        "stream_bytes": s_bytes, "ratio": in_bytes / s_bytes,
        "compress_hbm_frac": (in_bytes + s_bytes) / (c_ev * 1e-3) / HBM_PEAK,
        "decompress_hbm_frac": (in_bytes + s_bytes) / (d_ev * 1e-3) / HBM_PEAK,
    }
    # end to end per blocking call, as the reference's CLI times it (lsCOMP_xpcs.cpp:144-174:
    # the call returns with the stream size on the host), and in the reference's unit
    # (input MiB per ms, i.e. 1.048576e9 B/s)
    c_e2e, c_e2e_mean = e2e_ms(torch, lambda: lscomp.compress(d_in, dims, bins, pool, out=d_out), args.steps, True)
    d_e2e, d_e2e_mean = e2e_ms(torch, lambda: lscomp.decompress(d_out, s_bytes, dims, bins, out=d_dec), args.steps,
                               True)
    res.update({"compress_e2e_ms": c_e2e, "decompress_e2e_ms": d_e2e,
                "compress_e2e_mean_ms": c_e2e_mean, "decompress_e2e_mean_ms": d_e2e_mean,
                "compress_e2e_ref_unit_GBps": in_bytes / 1048576.0 / c_e2e,
                "decompress_e2e_ref_unit_GBps": in_bytes / 1048576.0 / d_e2e})
    if check_lossless:
        ok = ok and bool(torch.equal(d_dec, d_in))
    else:
        diff = ((d_in.to(torch.int32) & 0xFFFF) - (d_dec.to(torch.int32) & 0xFFFF)).abs()
        res["max_abs_err"] = int(diff.max().item())
        del diff
    res["verified"] = ok
    return res


def bench_xpcs(torch, lscomp, datagen, np, args):
    """XPCS-shaped 512 x 1024 x 1024: lossless and bins 3/5/10 x4 with pooling 0.5."""
    dims = XPCS_DIMS
    vol = datagen.volume("xpcs", 3, dims)
    d_in = torch.from_numpy(vol.view(np.int16)).cuda()
    in_bytes = 2 * vol.size
    d_out = torch.empty(lscomp.max_compressed_bytes(dims), dtype=torch.uint8, device="cuda")
    d_dec = torch.empty(dims, dtype=torch.int16, device="cuda")
    d_size = torch.zeros(1, dtype=torch.int64, device="cuda")
    lscomp.reserve(dims)
    res = {}
    for name, bins, pool in (("xpcs_lossless", (1, 1, 1, 1), 1.0), ("xpcs_b3x4_p0.5", (3, 3, 3, 3), 0.5),
                             ("xpcs_b5x4_p0.5", (5, 5, 5, 5), 0.5), ("xpcs_b10x4_p0.5", (10, 10, 10, 10), 0.5)):
        res[name] = time_codec(torch, lscomp, d_in, dims, bins, pool, d_out, d_dec, d_size, args, in_bytes,
                               bins == (1, 1, 1, 1))
        log(f"{name}: {json.dumps(res[name])}")
    del d_in, d_out, d_dec
    return res


def bench_small(torch, lscomp, datagen, np, args):
    """The 16 x 512 x 512 sparse volume (Poisson 0.05 + 0.1 % hot pixels), lossless,
    byte-equal to the oracle's stream and round-tripped exactly."""
    import oracle
    dims = SMALL_DIMS
    vol = datagen.volume("sparse", 1, dims)
    d_in = torch.from_numpy(vol.view(np.int16)).cuda()
    in_bytes = 2 * vol.size
    d_out = torch.empty(lscomp.max_compressed_bytes(dims), dtype=torch.uint8, device="cuda")
    d_dec = torch.empty(dims, dtype=torch.int16, device="cuda")
    d_size = torch.zeros(1, dtype=torch.int64, device="cuda")
    r = time_codec(torch, lscomp, d_in, dims, (1, 1, 1, 1), 1.0, d_out, d_dec, d_size, args, in_bytes, True)
    _, ref, refn = oracle.get().compress(vol, (1, 1, 1, 1), 1.0)
    r["oracle_identical"] = bool(refn == r["stream_bytes"] and np.array_equal(d_out[:refn].cpu().numpy(), ref))
    r["verified"] = r["verified"] and r["oracle_identical"]
    log(f"small_lossless: {json.dumps(r)}")
    return {"small_16x512x512_lossless": r}


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def cpu_baseline(vol, bins, pooling, frames_all=600, frames_one=96, reps=5):
    """The oracle (C; OpenMP two-pass variant on all threads, serial variant on
    one) on the first frames of the volume: one warm-up, then the median of
    `reps` passes of the same in-memory buffer."""
    import numpy as np
    import oracle
    orc = oracle.get()

    def median_pass(sub, omp):
        orc.compress(sub, bins, pooling, omp=omp)           # warm-up
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            orc.compress(sub, bins, pooling, omp=omp)
            times.append(time.perf_counter() - t0)
        return statistics.median(times)

    sub = np.ascontiguousarray(vol[:frames_all])
    dt = median_pass(sub, True)
    one = np.ascontiguousarray(vol[:frames_one])
    dt1 = median_pass(one, False)
    threads = orc.max_threads()

    # the other configurations, each on a bounded sample (GB/s of input, decimal)
    def config(v, b, p, n_all, n_one):
        a_ = np.ascontiguousarray(v[:n_all])
        o_ = np.ascontiguousarray(v[:n_one])
        ta = median_pass_s(a_, b, p, True)
        to = median_pass_s(o_, b, p, False)
        return {"all_threads_GBps": a_.nbytes / ta / 1e9, "one_thread_GBps": o_.nbytes / to / 1e9,
                "frames_all_threads": int(a_.shape[0]), "frames_one_thread": int(o_.shape[0])}

    def median_pass_s(v, b, p, omp):
        orc.compress(v, b, p, omp=omp)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            orc.compress(v, b, p, omp=omp)
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts)

    from lscomp import datagen
    configs = {"cssi_600x1813x1558_b3-5-10-15_p0.5": config(vol, (3, 5, 10, 15), 0.5, frames_all, 48)}
    small = datagen.volume("sparse", 1, SMALL_DIMS)
    configs["small_16x512x512_lossless"] = config(small, (1, 1, 1, 1), 1.0, SMALL_DIMS[0], SMALL_DIMS[0])
    xp = datagen.volume("xpcs", 3, (256,) + XPCS_DIMS[1:])   # the first 256 frames of the XPCS volume
    configs["xpcs_512x1024x1024_lossless"] = config(xp, (1, 1, 1, 1), 1.0, 256, 64)
    configs["xpcs_512x1024x1024_b3x4_p0.5"] = config(xp, (3, 3, 3, 3), 0.5, 256, 64)
    del xp, small
    return {"value": sub.nbytes / dt / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"lossless compression of the first {frames_all} frames ({sub.nbytes / 1e9:.3f} GB) of the "
                      f"same volume, OpenMP two-pass oracle on {threads} threads, 1 warm-up + median of {reps} "
                      f"passes ({dt:.3f} s)",
            "value_1_thread": one.nbytes / dt1 / 1e9,
            "sample_1_thread": f"serial oracle, first {frames_one} frames ({one.nbytes / 1e9:.3f} GB), "
                               f"1 warm-up + median of {reps} ({dt1:.3f} s)",
            "configs": configs,
            "cpu_model": cpu_info(), "nproc": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def bench_batched(torch, lscomp, datagen, np, steps):
    """512 square frames (sides drawn from 256..2048, XPCS-style, seed 4): one persistent
    plan launch (shape autotuned on the batch, replayed from a hipGraph) vs 512 plain
    per-frame launches; identical bytes.  Then the same streams decoded both ways (a
    decompression plan in one graph launch vs 512 plain decodes), each equal to the input."""
    rng = np.random.default_rng(4)
    sides = rng.choice([256, 512, 768, 1024, 1536, 2048], size=512)
    frames = [torch.from_numpy(datagen.volume("xpcs", 4, (1, int(s), int(s)), first_frame=i).view(np.int16)).cuda()
              for i, s in enumerate(sides)]
    dims = [(1, int(s), int(s)) for s in sides]
    caps = [lscomp.max_compressed_bytes(d) for d in dims]
    bins, pool = (1, 1, 1, 1), 1.0
    out_plain = [torch.empty(c, dtype=torch.uint8, device="cuda") for c in caps]
    out_batch = [torch.empty(c, dtype=torch.uint8, device="cuda") for c in caps]
    sizes = torch.zeros(len(frames), dtype=torch.int64, device="cuda")
    in_bytes = sum(2 * d[1] * d[2] for d in dims)
    nblocks = sum(lscomp.num_blocks(d) for d in dims)
    lscomp.reserve(max(dims, key=lambda d: d[1]))

    def plain():
        for i, (f, d) in enumerate(zip(frames, dims)):
            lscomp.compress_async(f, d, bins, pool, out_plain[i], sizes[i:i + 1])

    plan = lscomp.Plan(lscomp.COMPRESS, len(frames), nblocks)
    plan.set_items([(f, d, bins, pool, o, o.numel()) for f, d, o in zip(frames, dims, out_batch)])
    batch = lambda: plan.launch(graph=True)
    for fn in (plain, batch):
        fn()
    torch.cuda.synchronize()
    _, default_ms, _ = timed_loop(torch, batch, steps)
    default_shape = plan.get_shape()
    table = plan.autotune(reps=3)
    tuned_shape = plan.get_shape()
    batch()
    torch.cuda.synchronize()
    _, plain_ms, _ = timed_loop(torch, plain, steps)
    _, batch_ms, _ = timed_loop(torch, batch, steps)
    bsz = plan.sizes()
    psz = sizes.cpu().tolist()
    same = bsz == psz and all(torch.equal(a[:n], b[:n]) for a, b, n in zip(out_plain, out_batch, bsz))
    ok = plan.status() == lscomp.OK and lscomp.stream_status() == lscomp.OK
    plan.close()
    del out_plain

    # the same 512 streams decoded: one plan launch (graph) vs 512 plain launches
    dec_plain = [torch.empty_like(f) for f in frames]
    dec_batch = [torch.empty_like(f) for f in frames]

    def dplain():
        for i, d in enumerate(dims):
            lscomp.decompress_async(out_batch[i], bsz[i], d, bins, dec_plain[i])

    dplan = lscomp.Plan(lscomp.DECOMPRESS, len(frames), nblocks)
    dplan.set_items([(o, d, bins, pool, sb, n) for o, d, sb, n in zip(dec_batch, dims, out_batch, bsz)])
    dbatch = lambda: dplan.launch(graph=True)
    for fn in (dplain, dbatch):
        fn()
    torch.cuda.synchronize()
    _, dplain_ms, _ = timed_loop(torch, dplain, steps)
    _, dbatch_ms, _ = timed_loop(torch, dbatch, steps)
    dsame = all(torch.equal(a, f) and torch.equal(b, f) for a, b, f in zip(dec_plain, dec_batch, frames))
    dok = dplan.status() == lscomp.OK and lscomp.stream_status() == lscomp.OK
    dplan.close()
    return {"frames": len(frames), "input_bytes": in_bytes,
            "plain_launches_ms": plain_ms, "plain_frames_per_s": len(frames) / (plain_ms * 1e-3),
            "batched_graph_ms": batch_ms, "batched_frames_per_s": len(frames) / (batch_ms * 1e-3),
            "batched_GBps": in_bytes / (batch_ms * 1e-3) / 1e9, "speedup": plain_ms / batch_ms,
            "tuned_shape": tuned_shape, "default_shape": default_shape, "default_shape_ms": default_ms,
            "tuner_shapes_measured": len(table),
            "identical_streams": bool(same), "verified": bool(ok),
            "decode_plain_launches_ms": dplain_ms, "decode_plain_frames_per_s": len(frames) / (dplain_ms * 1e-3),
            "decode_batched_graph_ms": dbatch_ms, "decode_batched_frames_per_s": len(frames) / (dbatch_ms * 1e-3),
            "decode_batched_GBps": in_bytes / (dbatch_ms * 1e-3) / 1e9, "decode_speedup": dplain_ms / dbatch_ms,
            "decode_identical_to_input": bool(dsame), "decode_verified": bool(dok)}


def bench_served(torch, lscomp, datagen, np, seconds, rate, n_sat):
    """512^2..2048^2 CSSI frames, half lossless and half bins 3 5 10 15 / pooling 0.5, 64 GB
    budget.  The device-resident loop at the offered rate (Poisson arrivals) and at
    saturation (every request offered at once), the host-scheduled engine at saturation,
    and the same requests as plain per-frame launches.  Served streams are compared with
    the plain ones."""
    rng = np.random.default_rng(5)
    pool_sides = [512, 768, 1024, 1536, 2048]
    pool = []
    for i in range(40):
        s = pool_sides[i % len(pool_sides)]
        v = datagen.volume("cssi", 5, (1, s, s), first_frame=i)
        pool.append((torch.from_numpy(v.view(np.int16)).cuda(), (1, s, s)))
    modes = [((1, 1, 1, 1), 1.0), ((3, 5, 10, 15), 0.5)]
    budget = 64 << 30

    def trace(n, seed):
        r = np.random.default_rng(seed)
        return [(int(k), i % 2) for i, k in enumerate(r.integers(0, len(pool), n))]

    def requests(tr):
        return [lscomp.Server.request(pool[k][0], pool[k][1], *modes[m], tag=i) for i, (k, m) in enumerate(tr)]

    def frame_bytes(tr):
        return sum(2 * pool[k][1][1] * pool[k][1][2] for k, _ in tr)

    # plain per-frame launches of the saturation trace, and the reference streams
    tr_sat = trace(n_sat, 51)
    outs = {}
    for k, m in set(tr_sat):
        outs[(k, m)] = torch.empty(lscomp.max_compressed_bytes(pool[k][1]), dtype=torch.uint8, device="cuda")
    size = torch.zeros(1, dtype=torch.int64, device="cuda")
    lscomp.reserve(pool[4][1])

    def plain():
        for k, m in tr_sat:
            lscomp.compress_async(pool[k][0], pool[k][1], *modes[m], outs[(k, m)], size)

    plain()
    torch.cuda.synchronize()
    _, plain_ms, _ = timed_loop(torch, plain, 3)
    ref = {}
    for (k, m), o in outs.items():
        _, n = lscomp.compress(pool[k][0], pool[k][1], *modes[m], out=o)
        ref[(k, m)] = o[:n].cpu().numpy()
    torch.cuda.synchronize()
    res = {"frame_mix": "CSSI-style 512^2..2048^2, half lossless / half bins 3 5 10 15 p0.5",
           "budget_bytes": budget,
           "plain": {"requests": n_sat, "frames_per_s": n_sat / (plain_ms * 1e-3),
                     "GBps": frame_bytes(tr_sat) / (plain_ms * 1e-3) / 1e9}}

    def run(srv, tr, arrivals, check):
        reqs = requests(tr)
        t0 = time.perf_counter()
        out = srv.replay(reqs, arrivals, hold=check)
        wall = time.perf_counter() - t0
        same = True
        if check:
            for c in out:
                k, m = tr[c.tag]
                r = ref.get((k, m))
                if r is not None and c.status == lscomp.OK:
                    got = _copy_dev(c.d_stream, c.stream_bytes)
                    same = same and c.stream_bytes == r.size and np.array_equal(got, r)
                srv.release(c.id)
        ok = srv.drain() == lscomp.OK and all(c.status == lscomp.OK for c in out)
        lat = np.array([c.t_done - c.t_submit for c in out]) * 1e3
        st = srv.stats()
        return {"requests": len(tr), "frames_per_s": len(tr) / wall, "GBps": frame_bytes(tr) / wall / 1e9,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "max_ms": float(lat.max()), "kernel_launches": st["launches"], "verified": bool(ok),
                **({"identical_to_plain": bool(same)} if check else {})}

    # device loop: offered rate (Poisson), then saturation (with a byte check of every stream)
    n_off = int(seconds * rate)
    tr_off = trace(n_off, 52)
    arr = np.cumsum(np.random.default_rng(53).exponential(1.0 / rate, n_off))
    srv = lscomp.Server.device(1 << 17, budget, idle_us=2000)
    res["device_offered"] = {"offered_frames_per_s": rate, **run(srv, tr_off, arr, False)}
    srv.close()
    srv = lscomp.Server.device(1 << 17, budget, idle_us=2000)
    run(srv, tr_sat[:200], np.zeros(200), False)   # warm-up
    res["device_saturation"] = run(srv, tr_sat, np.zeros(n_sat), True)
    srv.close()
    srv = lscomp.Server(slots=4, max_items=256, max_blocks=1 << 22, budget_bytes=budget)
    run(srv, tr_sat[:200], np.zeros(200), False)
    res["host_saturation"] = run(srv, tr_sat, np.zeros(n_sat), True)
    srv.close()
    res["device_vs_plain_frames_per_s"] = res["device_saturation"]["frames_per_s"] / res["plain"]["frames_per_s"]
    return res


def _copy_dev(ptr, nbytes):
    """Copy a stream out of the server's arena (a raw device pointer) to the host."""
    import ctypes
    import numpy as np
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    host = np.empty(nbytes, np.uint8)
    if hip.hipMemcpy(host.ctypes.data, ptr, nbytes, 2) != 0:   # hipMemcpyDeviceToHost
        raise RuntimeError("hipMemcpy failed")
    return host


def bench_u32(torch, lscomp, d_in16, dims, d_out16, args):
    """The uint32 path (the reference's lsCOMP_cssi path) on the same volume stored
    as uint32 (4 bytes per pixel).  Its stream must equal the uint16 stream of
    the same values."""
    res = {}
    vol32 = (d_in16.to(torch.int32) & 0xFFFF).contiguous()
    in_bytes = 4 * vol32.numel()
    d_out = torch.empty(lscomp.max_compressed_bytes_u32(dims), dtype=torch.uint8, device="cuda")
    d_dec = torch.empty_like(vol32)
    d_size = torch.zeros(1, dtype=torch.int64, device="cuda")
    for name, bins, pool in (("u32_lossless", (1, 1, 1, 1), 1.0), ("u32_lossy_b3-5-10-15_p0.5", (3, 5, 10, 15), 0.5)):
        comp = lambda: lscomp.compress_u32_async(vol32, dims, bins, pool, d_out, d_size)
        for _ in range(args.warmup):
            comp()
        torch.cuda.synchronize()
        c_wall, c_ev, _ = timed_loop(torch, comp, args.steps)
        s_bytes = int(d_size.item())
        dec = lambda: lscomp.decompress_u32_async(d_out, s_bytes, dims, bins, d_dec)
        for _ in range(args.warmup):
            dec()
        torch.cuda.synchronize()
        d_wall, d_ev, _ = timed_loop(torch, dec, args.steps)
        ok = lscomp.stream_status() == lscomp.OK
        # same values as the uint16 volume: same stream
        _, n16 = lscomp.compress(d_in16, dims, bins, pool, out=d_out16)
        ok = ok and n16 == s_bytes and bool(torch.equal(d_out16[:n16], d_out[:s_bytes]))
        if bins == (1, 1, 1, 1):
            ok = ok and bool(torch.equal(d_dec, vol32))
        res[name] = {
            "compress_GBps": in_bytes / (c_wall * 1e-3) / 1e9,
            "decompress_GBps": in_bytes / (d_wall * 1e-3) / 1e9,
            "compress_ms": c_wall, "decompress_ms": d_wall,
            "compress_kernel_ms": c_ev, "decompress_kernel_ms": d_ev,
            "stream_bytes": s_bytes, "ratio": in_bytes / s_bytes,
            "compress_hbm_frac": (in_bytes + s_bytes) / (c_ev * 1e-3) / HBM_PEAK,
            "decompress_hbm_frac": (in_bytes + s_bytes) / (d_ev * 1e-3) / HBM_PEAK,
            "verified": ok,
        }
        log(f"{name}: {json.dumps(res[name])}")
    del vol32, d_out, d_dec
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=CSSI_DIMS[0], help="frames of the CSSI volume (default 600)")
    ap.add_argument("--cpu-frames", type=int, default=CSSI_DIMS[0],
                    help="frames of the CPU baseline's threaded sample (default: the whole volume)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the batched and served measurements")
    ap.add_argument("--modes", default="lossless,lossy,xpcs,small,u32",
                    help="settings to time: lossless,lossy (CSSI volume), xpcs, small, u32")
    ap.add_argument("--served-seconds", type=float, default=2.0)
    ap.add_argument("--served-rate", type=float, default=20000.0)
    ap.add_argument("--served-saturation", type=int, default=8000, help="requests of the saturation runs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import lscomp
    from lscomp import datagen

    dims = (args.frames, CSSI_DIMS[1], CSSI_DIMS[2])
    n_elem = dims[0] * dims[1] * dims[2]
    in_bytes = 2 * n_elem
    torch.cuda.set_device(0)
    wanted = args.modes.split(",")

    t = time.perf_counter()
    vol = datagen.volume("cssi", 2, dims)
    vol_sha = hashlib.sha256(memoryview(vol).cast("B")).hexdigest()
    log(f"generated {dims} in {time.perf_counter() - t:.1f}s, sha256 {vol_sha}")
    d_in = torch.from_numpy(vol.view(np.int16)).cuda()
    cap = lscomp.max_compressed_bytes(dims)
    d_out = torch.empty(cap, dtype=torch.uint8, device="cuda")
    d_dec = torch.empty(dims, dtype=torch.int16, device="cuda")
    d_size = torch.zeros(1, dtype=torch.int64, device="cuda")
    lscomp.reserve(dims)

    results = {}
    for name, bins, pool in (("lossless", (1, 1, 1, 1), 1.0), ("lossy_b3-5-10-15_p0.5", (3, 5, 10, 15), 0.5)):
        if name.split("_")[0] not in wanted:
            continue
        results[name] = time_codec(torch, lscomp, d_in, dims, bins, pool, d_out, d_dec, d_size, args, in_bytes,
                                   bins == (1, 1, 1, 1))
        log(f"{name}: {json.dumps(results[name])}")

    if "u32" in wanted:
        results.update(bench_u32(torch, lscomp, d_in, dims, d_out, args))

    # measured device-copy bandwidth for context (read + write of the input).
    # The library's plain copy makes 16-byte accesses per lane like the codec
    # kernels, so its FETCH_SIZE/WRITE_SIZE calibrate the counters for them.
    cp = torch.empty_like(d_in)
    lscomp.copy_bytes(cp, d_in)   # first touch of the new buffer outside the timing
    torch.cuda.synchronize()
    _, cp_ms, _ = timed_loop(torch, lambda: lscomp.copy_bytes(cp, d_in), 5)
    copy_gbps = 2 * in_bytes / (cp_ms * 1e-3) / 1e9
    del cp
    # each mode's kernel rate also as a fraction of that measured copy rate
    for r in results.values():
        for d in ("compress", "decompress"):
            r[f"{d}_copy_frac"] = r[f"{d}_hbm_frac"] * (HBM_PEAK / 1e9) / copy_gbps

    del d_dec
    if "xpcs" in wanted:
        results.update(bench_xpcs(torch, lscomp, datagen, np, args))
    if "small" in wanted:
        results.update(bench_small(torch, lscomp, datagen, np, args))

    forms = {}
    if not args.no_forms:
        forms["batched"] = bench_batched(torch, lscomp, datagen, np, max(3, args.steps // 2))
        log(f"batched: {json.dumps(forms['batched'])}")
        forms["served"] = bench_served(torch, lscomp, datagen, np, args.served_seconds, args.served_rate,
                                       args.served_saturation)
        log(f"served: {json.dumps(forms['served'])}")

    ll = results["lossless"]
    s_bytes = ll["stream_bytes"]
    achieved = (in_bytes + s_bytes) / (ll["compress_kernel_ms"] * 1e-3)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(vol, (1, 1, 1, 1), 1.0, frames_all=min(args.cpu_frames, dims[0]),
                           frames_one=min(96, dims[0]))
        log(f"cpu: {json.dumps(cpu)}")
    rec = {
        "metric": "compress/decompress GB/s + % of HBM peak on 1 MI355X; p99 per-frame latency",
        "value": ll["compress_GBps"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ll["compress_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u16", "data": "synthetic",
        "config": {"workload": f"CSSI-shaped {dims[0]}x{dims[1]}x{dims[2]} uint16, Poisson 0.3 + 50 peaks/frame "
                               f"+ gap stripes, seed 2; lossless compression (bins 1 1 1 1, pooling 1)",
                   "dims": list(dims), "input_bytes": in_bytes, "volume_sha256": vol_sha},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": profiled_traffic(dims)},
        "cpu_baseline": cpu,
        "modes": results,
        "forms": forms,
        "device_copy_GBps": copy_gbps,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
