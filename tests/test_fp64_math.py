"""The exact-range kernel's FP64 routines on the host (tests/fp64_math_check.cpp,
built with hipcc from srg_amd/csrc/tdbp_geom.h): sin and cos of 2*pi*f within
1e-11 of long-double sinl/cosl (the Taylor remainder bound is 7e-12), and the
Goldschmidt root correctly rounded from reciprocal-root estimates 1e-7 off."""
import os
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_fp64_math_routines():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "fp64_math_check")
        subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-x", "hip",
                        os.path.join(HERE, "fp64_math_check.cpp"), "-o", exe],
                       check=True, capture_output=True, timeout=300)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
