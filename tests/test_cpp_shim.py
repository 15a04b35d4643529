"""The header-only C++ plugin class (include/dmrgpp_amd_batched_gemm.hpp) compiles against the
library, and on the GPU passes the restated reference unit test test_BatchedGemm.cpp."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
BIN = os.path.join(CPP, "_build", "test_batched_gemm_shim")


def _build():
    subprocess.run(["make", "-s", "-C", CPP], check=True)


def test_shim_compiles_and_links():
    _build()
    assert os.path.exists(BIN)


@pytest.mark.gpu
def test_shim_batched_gemm_on_gpu():
    if not os.path.exists(BIN):
        _build()
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "passed" in r.stdout
