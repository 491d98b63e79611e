"""bench.py's command-line contract that holds without a GPU: more than one
GPU is refused with a message (the render path runs on one GPU), and the
options parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stdout


def test_options_parse():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--precision", "--cache-latents", "--no-latent-leg"):
        assert opt in p.stdout
