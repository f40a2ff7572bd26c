"""bench.py's command line on the CPU: more than one GPU is refused with a
message and exit status 2 before anything touches a device, and the options
the measurement documents parse."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _bench(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True,
                          text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_documented_options_parse():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--gpus", "--no-serve", "--no-plain", "--no-batch",
                "--all-configs", "--serve-s22-contexts", "--tail-div", "--copy-calibration"):
        assert opt in r.stdout, opt
