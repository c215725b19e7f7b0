#!/usr/bin/env python3
"""Benchmark of the gfx950 lsCOMP uint16 codec on one MI355X.

Workload (default): the CSSI-shaped volume 600 x 1813 x 1558 uint16 (3.39 GB),
Poisson(0.3) background + 50 Gaussian peaks per frame + gap stripes, seed 2,
generated on the host and resident in HBM before timing.  A step is one
compression of the whole volume (lossless: bins 1 1 1 1, pooling 1); the same
K steps are then timed for decompression and for the lossy setting
(bins 3 5 10 15, pooling 0.5).  `value` is the lossless compression
throughput in GB/s (1e9 bytes of uncompressed input per second).

Prints ONE JSON line.  `--gpus` above 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sc25-lscomp_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
CSSI_DIMS = (600, 1813, 1558)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed_loop(torch, fn, steps):
    """Average ms of fn() over `steps` launches on the current stream, plus per-launch event times."""
    stream = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    per = [a.elapsed_time(b) for a, b in evs]
    return wall, sum(per) / len(per), per


def cpu_baseline(vol, sample_frames, bins, pooling):
    """Oracle (C, OpenMP two-pass) on the first `sample_frames` frames, all host threads."""
    import numpy as np
    import oracle
    orc = oracle.get()
    sub = np.ascontiguousarray(vol[:sample_frames])
    orc.compress(sub, bins, pooling, omp=True)           # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        st, _, n = orc.compress(sub, bins, pooling, omp=True)
        reps += 1
        if time.perf_counter() - t0 > 5.0 or reps >= 5:
            break
    dt = (time.perf_counter() - t0) / reps
    return sub.nbytes / dt / 1e9, orc.max_threads(), sub.nbytes, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=CSSI_DIMS[0], help="frames of the CSSI volume (default 600)")
    ap.add_argument("--cpu-frames", type=int, default=48, help="frames of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
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

    t = time.perf_counter()
    vol = datagen.volume("cssi", 2, dims)
    log(f"generated {dims} in {time.perf_counter() - t:.1f}s")
    d_in = torch.from_numpy(vol.view(np.int16)).cuda()
    cap = lscomp.max_compressed_bytes(dims)
    d_out = torch.empty(cap, dtype=torch.uint8, device="cuda")
    d_dec = torch.empty(dims, dtype=torch.int16, device="cuda")
    d_size = torch.zeros(1, dtype=torch.int64, device="cuda")
    lscomp.reserve(dims)

    results = {}
    for name, bins, pool in (("lossless", (1, 1, 1, 1), 1.0), ("lossy_b3-5-10-15_p0.5", (3, 5, 10, 15), 0.5)):
        comp = lambda: lscomp.compress_async(d_in, dims, bins, pool, d_out, d_size)
        for _ in range(args.warmup):
            comp()
        torch.cuda.synchronize()
        c_wall, c_ev, _ = timed_loop(torch, comp, args.steps)
        s_bytes = int(d_size.item())
        dec = lambda: lscomp.decompress_async(d_out, s_bytes, dims, bins, d_dec)
        for _ in range(args.warmup):
            dec()
        torch.cuda.synchronize()
        d_wall, d_ev, _ = timed_loop(torch, dec, args.steps)
        st = lscomp.stream_status()
        ok = st == lscomp.OK
        if name == "lossless":
            ok = ok and bool(torch.equal(d_dec, d_in))
        else:
            diff = ((d_in.to(torch.int32) & 0xFFFF) - (d_dec.to(torch.int32) & 0xFFFF)).abs()
            results_max_err = int(diff.max().item())
        results[name] = {
            "compress_GBps": in_bytes / (c_wall * 1e-3) / 1e9,
            "decompress_GBps": in_bytes / (d_wall * 1e-3) / 1e9,
            "compress_ms": c_wall, "decompress_ms": d_wall,
            "compress_kernel_ms": c_ev, "decompress_kernel_ms": d_ev,
            "stream_bytes": s_bytes, "ratio": in_bytes / s_bytes,
            "compress_hbm_frac": (in_bytes + s_bytes) / (c_ev * 1e-3) / HBM_PEAK,
            "decompress_hbm_frac": (in_bytes + s_bytes) / (d_ev * 1e-3) / HBM_PEAK,
            "verified": ok,
        }
        if name != "lossless":
            results[name]["max_abs_err"] = results_max_err
        log(f"{name}: {json.dumps(results[name])}")

    # measured device-copy bandwidth for context (read + write of the input)
    cp = torch.empty_like(d_in)
    _, cp_ms, _ = timed_loop(torch, lambda: cp.copy_(d_in), 5)
    copy_gbps = 2 * in_bytes / (cp_ms * 1e-3) / 1e9
    del cp

    ll = results["lossless"]
    s_bytes = ll["stream_bytes"]
    achieved = (in_bytes + s_bytes) / (ll["compress_kernel_ms"] * 1e-3)
    cpu = None
    if not args.no_cpu:
        frames = min(args.cpu_frames, dims[0])
        gbps, threads, nbytes, dt = cpu_baseline(vol, frames, (1, 1, 1, 1), 1.0)
        cpu = {"value": gbps, "unit": "GB/s", "cores": threads, "kind": "port",
               "sample": f"lossless compression of the first {frames} frames ({nbytes / 1e9:.3f} GB), "
                         f"OpenMP two-pass oracle, {dt:.3f} s per pass"}
    rec = {
        "metric": "compress/decompress GB/s + % of HBM peak on 1 MI355X; p99 per-frame latency",
        "value": ll["compress_GBps"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ll["compress_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u16", "data": "synthetic",
        "config": {"workload": f"CSSI-shaped {dims[0]}x{dims[1]}x{dims[2]} uint16, Poisson 0.3 + 50 peaks/frame "
                               f"+ gap stripes, seed 2; lossless compression (bins 1 1 1 1, pooling 1)",
                   "dims": list(dims), "input_bytes": in_bytes},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None},
        "cpu_baseline": cpu,
        "modes": results,
        "device_copy_GBps": copy_gbps,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
