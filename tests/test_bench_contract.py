"""bench.py's command-line contract, the parts that need no GPU: more than one
GPU is refused with a message and exit status 2 before anything touches a
device, and the options the measurement is documented with exist with their
defaults (lanes 2, row order 3, the waves per workgroup chosen by precision,
the per-colour schedule, 5000 iterations).  No GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_documented_defaults():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--lanes", "--row-order", "--wpb", "--schedule", "--maxiter",
                "--snapshot", "--save-snapshot", "--workload", "--shape"):
        assert opt in r.stdout, opt
    sys.path.insert(0, ROOT)
    try:
        import bench
    finally:
        sys.path.pop(0)
    d = vars(bench.make_parser().parse_args([]))
    assert d["gpus"] == 1 and d["steps"] == 5 and d["warmup"] == 1
    assert d["lanes"] == 2 and d["row_order"] == 3 and d["wpb"] == 0 and d["schedule"] == 0
    assert d["maxiter"] == 5000 and d["workload"] == "packing" and d["shape"] == "sphere"
