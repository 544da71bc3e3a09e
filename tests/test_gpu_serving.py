"""Serving loop on the GPU: requests that arrive together are merged and run
concurrently, and every result equals the same request solved alone (one
operator, one block, nothing else on the device) within the solve tolerance:
cross sections within 1e-4 relative, dipoles within 1e-3."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_alone_vs_in_the_mix(bdda):
    S = importlib.import_module("block_dda_amd.serving")
    T = importlib.import_module("block_dda_amd.targets")
    scenes = S.synthetic_scenes(T, 4, seed=5, nmin=24, nmax=96, across=(4, 4), max_grid=64)
    reqs = S.synthetic_requests(T, scenes, 12, rate=1000.0, seed=6, s_choices=(2, 8), tol=1e-5)
    print("case serving", [sc.meta["nocc"] for sc in scenes], flush=True)
    be = S.GpuBackend(bdda)
    sv = S.Server(be, hbm_budget=32 << 30, workers=3).start()
    for r in reqs:
        sv.submit(r)
    assert sv.drain(timeout=300)
    st = sv.stats()
    sv.stop()
    print("stats", st, flush=True)
    assert st["failed"] == 0 and st["requests"] == 12
    assert st["merged_requests"] > 0
    assert st["max_concurrent_batches"] >= 2
    ops = {}
    for r in reqs:
        sc = r.scene
        if sc.key not in ops:
            ops[sc.key] = be.create(sc)
        op = ops[sc.key]
        X, info = be.solve(op, r.B, r.tol)
        assert info["converged"] and info["true_res"] <= r.tol
        assert r.info["converged"] and r.info["true_res"] <= r.tol
        xs = op.cross_sections(r.B, X)
        for q in ("Cext", "Cabs", "Csca"):
            mix = np.asarray(r.observables[q])
            assert np.abs(mix - xs[q]).max() <= 1e-4 * np.abs(xs["Cext"]).max(), (r.id, q)
        assert np.linalg.norm(r.X - X) / np.linalg.norm(X) <= 1e-3
    for op in ops.values():
        op.close()
