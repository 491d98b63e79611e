"""The cache heads divide the 8 trilinear latent sums by one weight sum with
fvc::div_recip (one IEEE reciprocal, then Markstein's correction per
numerator) instead of 8 IEEE divisions (cache_performance_renderer.cu:770-787
divides each sum), and the reprojection divides camera coordinates by their
depth and by the stepsize constants the same way.  The product header is
built for the host and checked against a / b bit for bit over every divisor
in +-[2^-8, 2] (the weight sums lie in [1/8, 1]) with random numerators and
signed zeros, and over 4 M random divisors of either sign in 2^-60 .. 2^60
with quotients in 2^-30 .. 2^30."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def test_div_recip_equals_ieee_division(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "div_recip_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950", "-x", "hip",
                    os.path.join(HERE, "div_recip_check.hip"), "-o", exe], check=True, capture_output=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(p.stdout)
    assert p.returncode == 0 and " 0 mismatches" in p.stdout, p.stdout
