"""bench.py's command-line contract on a host without a GPU: more than one
GPU is refused with a message (the build measures one GPU), and the
options parse.  Nothing here touches a device."""

import subprocess
import sys

from conftest import ROOT


def _run(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_options_parse():
    r = _run("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--gpus", "--workload", "--fd-check", "--delta-q", "--nonlocal-only",
                "--subspace-only", "--hpsi-only"):
        assert opt in r.stdout
