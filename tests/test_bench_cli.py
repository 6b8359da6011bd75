"""bench.py's command line without a GPU: more than one GPU is refused with a message (there
is no multi-GPU path), and the options the profiles' commands use are accepted."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + list(args), capture_output=True,
                          text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_profile_options_are_accepted():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--profile-run", "--profile-variant", "--grouped-only", "--variants", "--no-cpu", "--steps",
                "--warmup"):
        assert opt in r.stdout, opt
