"""bench.py's command line on a host without a GPU: a run on more than one GPU
is refused with a message before anything touches a device, and the batch
legs that the bench runs in child processes are named options."""
import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr or "--gpus 1" in p.stderr


def test_legs_are_options():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0
    assert "--leg" in p.stdout and "--hw-queues" in p.stdout
    bad = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--leg", "nonsense"], capture_output=True,
                         text=True, timeout=120)
    assert bad.returncode != 0
