"""The header-only C++ plugin class (include/dmrgpp_amd_batched_gemm.hpp) and the drop-in header
with the reference's name (include/BatchedGemm.h) compile against the library; on the GPU they
pass the restated reference unit test test_BatchedGemm.cpp and the call pattern of
BatchedGemmPluginSc (constructor + apply_Htarget on a pair of shared-space views)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
BINS = {name: os.path.join(CPP, "_build", name)
        for name in ("test_batched_gemm_shim", "test_plugin_dropin")}


def _build():
    subprocess.run(["make", "-s", "-C", CPP], check=True)


def test_shim_compiles_and_links():
    _build()
    for path in BINS.values():
        assert os.path.exists(path)


def test_pool_cache_on_cpu():
    """The allocation cache behind the device and pinned pools (pool_cache.h) on malloc, under
    AddressSanitizer and UBSan: reuse within 2x, the cap never returns a buffer in use, a stale
    release is counted and changes nothing, and 8 threads keep the books balanced."""
    _build()
    exe = os.path.join(CPP, "_build", "test_pool_cache")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout


def test_pool_counters_export_without_gpu():
    """dmk_pool_counters is host-only bookkeeping: callable without a GPU, no stale releases."""
    import dmrgpp_amd as dm

    for host in (False, True):
        c = dm.pool_counters(host)
        assert c["bad_releases"] == 0
        assert set(c) == {"used_bytes", "cached_bytes", "allocs", "frees", "reuses", "bad_releases"}


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(BINS))
def test_shim_on_gpu(name):
    if not os.path.exists(BINS[name]):
        _build()
    r = subprocess.run([BINS[name]], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "passed" in r.stdout
