"""bench.py's command-line contract, checked without a GPU: it measures one GPU, so a request
for several exits with a message before touching a device."""
import json
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=REPO)
    assert r.returncode != 0
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_bench_help_lists_the_legs():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120, cwd=REPO)
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--gpus", "--small-batches", "--otf-sps", "--uneven-jobs", "--serve-seconds",
                 "--variant-iters", "--sparse-products"):
        assert flag in r.stdout
