"""bench.py's command line on a host without a GPU: a request for more than
one GPU is refused with a message and exit status 2, before anything touches
the device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, env_extra=None):
    env = dict(os.environ, **(env_extra or {}))
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT, env=env,
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    p = _run(["--gpus", "2"])
    assert p.returncode == 2
    assert "one GPU only" in p.stderr
    assert p.stdout == ""


def test_world_size_above_one_is_refused():
    p = _run(["--gpus", "1"], {"WORLD_SIZE": "4"})
    assert p.returncode == 2
    assert "one GPU only" in p.stderr
