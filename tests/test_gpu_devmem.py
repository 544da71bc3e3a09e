"""The library's caching device allocator on the GPU (no reference
counterpart: the reference keeps its blocks in host memory).

- a freed block is handed out again without a new runtime allocation, and a
  block's contents survive until it is freed;
- a second solve on the same operator, and an operator built again on the
  same grid, take all their memory from idle blocks, and give bit-identical
  results (memory that held other data is never read before it is written);
- trimming gives wholly idle memory back to the runtime;
- with BDDA_DEVICE_CACHE=0 (plain hipMalloc / hipFree) a solve gives the same
  result as with the cache."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def small_problem(bdda, seed=3):
    import importlib
    T = importlib.import_module("block_dda_amd.targets")
    rng = np.random.default_rng(seed)
    n, f, lf, k = (14, 12, 13), 3, 0.5, 0.9
    C = int(np.prod(n))
    addr = np.sort(rng.choice(C, size=900, replace=False)).astype(np.int32)
    diag = np.repeat(np.full(addr.size, 1.0 / T.alpha_ldr(1.6 + 0.2j, lf, k)), f)
    B = rng.standard_normal((f * addr.size, 8)) + 1j * rng.standard_normal((f * addr.size, 8))
    return n, f, lf, k, addr, diag, B


def solve_once(bdda, prob, mode):
    n, f, lf, k, addr, diag, B = prob
    op = bdda.MBT_fft_init(n, f, lf, k)
    try:
        op.set_target(addr, diag)
        X, info, _ = op.solve(B, 1e-8, 500, method=bdda.COCG_RQ, jacobi=True, mode=mode)
    finally:
        op.close()
    assert info["converged"]
    return X


def test_freed_block_is_reused_with_contents_kept(bdda):
    L = bdda.lib()
    nbytes = 48 << 20
    p = ctypes.c_void_p()
    bdda._check(L.bdda_dev_alloc(nbytes, ctypes.byref(p)))
    h = np.arange(nbytes // 8, dtype=np.int64)
    bdda._check(L.bdda_memcpy_h2d(p, h.ctypes.data, nbytes))
    back = np.empty_like(h)
    bdda._check(L.bdda_memcpy_d2h(back.ctypes.data, p, nbytes))
    assert np.array_equal(back, h)
    s1 = bdda.dev_pool_stats()
    assert s1["in_use"] >= nbytes
    bdda._check(L.bdda_dev_free(p))
    q = ctypes.c_void_p()
    bdda._check(L.bdda_dev_alloc(nbytes, ctypes.byref(q)))
    s2 = bdda.dev_pool_stats()
    assert s2["runtime_allocs"] == s1["runtime_allocs"]
    assert s2["reuses"] > s1["reuses"]
    assert q.value % 512 == 0
    bdda._check(L.bdda_dev_free(q))


def test_solves_and_operators_reuse_idle_memory(bdda):
    prob = small_problem(bdda)
    X1 = solve_once(bdda, prob, bdda.MODE_MIXED)
    s1 = bdda.dev_pool_stats()
    # the operator (G^, twiddles, build transient, workspace), the target and
    # every Krylov block of the second round come from idle blocks
    X2 = solve_once(bdda, prob, bdda.MODE_MIXED)
    s2 = bdda.dev_pool_stats()
    assert s2["runtime_allocs"] == s1["runtime_allocs"], (s1, s2)
    assert s2["reuses"] > s1["reuses"]
    assert np.array_equal(X1, X2)
    X3 = solve_once(bdda, prob, bdda.MODE_FP64)
    X4 = solve_once(bdda, prob, bdda.MODE_FP64)
    assert np.array_equal(X3, X4)
    assert np.abs(X1 - X3).max() <= 1e-5 * np.abs(X3).max()


def test_trim_gives_idle_memory_back(bdda):
    L = bdda.lib()
    p = ctypes.c_void_p()
    bdda._check(L.bdda_dev_alloc(300 << 20, ctypes.byref(p)))
    bdda._check(L.bdda_dev_free(p))
    held = bdda.dev_pool_stats()["held"]
    released = bdda.dev_trim()
    after = bdda.dev_pool_stats()
    assert released >= 300 << 20
    assert after["held"] == held - released
    assert after["held"] >= after["in_use"]


SCRIPT = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
import conftest, test_gpu_devmem as t
bdda = conftest.load_pkg()
X = t.solve_once(bdda, t.small_problem(bdda), bdda.MODE_MIXED)
st = bdda.dev_pool_stats()
assert st["held"] == 0 and st["runtime_allocs"] == 0, st
np.save(sys.argv[2], X)
print("ok", flush=True)
"""


def test_cache_off_gives_the_same_result(bdda, tmp_path):
    X = solve_once(bdda, small_problem(bdda), bdda.MODE_MIXED)
    out = str(tmp_path / "x_nocache.npy")
    env = dict(os.environ, BDDA_DEVICE_CACHE="0")
    r = subprocess.run([sys.executable, "-c", SCRIPT, os.path.join(ROOT, "tests"), out], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr[-3000:]
    assert np.array_equal(np.load(out), X)
