"""The lsCOMP_xpcs (uint16) and lsCOMP_cssi (uint32) CLIs' argument validation
(no GPU needed: they exit before any HIP call)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLIS = [os.path.join(ROOT, "sc25-lscomp_amd", "cli", name) for name in ("lsCOMP_xpcs", "lsCOMP_cssi")]


@pytest.mark.parametrize("cli", CLIS, ids=os.path.basename)
@pytest.mark.parametrize("args,msg", [
    (["-d", "0", "8", "8"], "-d values must be positive integers"),
    (["-b", "3", "2", "5", "6"], "-b values must satisfy x <= y <= z <= w"),
    (["-p", "1.5"], "-p value must be between 0 and 1"),
    (["-x", "1"], "Unknown argument -x"),
])
def test_cli_validation(tmp_path, cli, args, msg):
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    base = {"-i": [str(tmp_path / "none.bin")], "-d": ["1", "8", "8"], "-b": ["1", "1", "1", "1"], "-p": ["1"]}
    flag = args[0]
    if flag in base:
        base[flag] = args[1:]
        argv = [a for k, v in base.items() for a in [k, *v]]
    else:
        argv = [a for k, v in base.items() for a in [k, *v]] + args
    r = subprocess.run([cli, *argv], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and msg in r.stderr
