"""The C++ API with the reference's class names (include/pmkd/*.hpp) compiles
against the C ABI (CPU) and drives a build and all three tracers (GPU)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_shim")


def test_cpp_api_compiles():
    subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "cpp")])
    assert os.path.exists(BIN)


@pytest.mark.gpu
def test_cpp_api_on_gpu():
    assert os.path.exists(BIN), "build tests/cpp/test_shim first (__graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "C++ API OK" in r.stdout
