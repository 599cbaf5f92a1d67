"""The tracers' 1/det (pmkd::rcp_rn in pm_device.h: hardware reciprocal plus
one fma correction inside [2^-126, 2^126), the IEEE division outside) must
equal the IEEE quotient 1.0f / x bit for bit for every float input; the
check runs over all 2^32 bit patterns on the GPU (tests/cpp/rcp_exhaustive)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "rcp_exhaustive")


def test_rcp_rn_equals_ieee_division_everywhere():
    assert os.path.exists(EXE), "build tests/cpp first (__graft_entry__.build())"
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "checked 4294967296 mismatches 0" in out.stdout, out.stdout
