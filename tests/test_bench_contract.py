"""bench.py's contract on the host: one GPU only (more is refused with a
message and exit status 2, before any GPU work), and the CPU-baseline record
carries the fields its documentation names."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("argv, env", [(["--gpus", "2"], {}), (["--gpus", "8"], {}), ([], {"WORLD_SIZE": "2"})])
def test_more_than_one_gpu_is_refused(argv, env):
    e = dict(os.environ, **env)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + argv, env=e, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_cpu_baseline_record():
    sys.path.insert(0, os.path.join(ROOT, "sc25-lscomp_amd"))
    sys.path.insert(0, ROOT)
    import bench
    from lscomp import datagen
    vol = datagen.volume("cssi", 2, (6, 1813, 1558))
    rec = bench.cpu_baseline(vol, (1, 1, 1, 1), 1.0, frames_all=6, frames_one=2, reps=1)
    for k in ("value", "unit", "cores", "kind", "sample", "value_1_thread", "sample_1_thread", "configs",
              "cpu_model", "nproc", "omp_num_threads"):
        assert k in rec, k
    assert rec["unit"] == "GB/s" and rec["kind"] == "port" and rec["cores"] >= 1
    assert rec["value"] > 0 and rec["value_1_thread"] > 0
    assert set(rec["configs"]) == {"cssi_600x1813x1558_b3-5-10-15_p0.5", "small_16x512x512_lossless",
                                   "xpcs_512x1024x1024_lossless", "xpcs_512x1024x1024_b3x4_p0.5"}
    for c in rec["configs"].values():
        assert c["all_threads_GBps"] > 0 and c["one_thread_GBps"] > 0
        assert 1 <= c["frames_one_thread"] <= c["frames_all_threads"]
    assert np.isfinite(rec["value"])
