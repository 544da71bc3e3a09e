"""Serving loop on the GPU: requests that arrive together are merged and run
concurrently, and every result equals the same request solved alone (one
operator, one block, nothing else on the device) within the solve tolerance:
cross sections within 1e-4 relative, dipoles within 1e-3."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check_alone(be, reqs):
    """Each request solved alone on a fresh operator of its scene."""
    ops = {}
    try:
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
    finally:
        for op in ops.values():
            op.close()


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
    _check_alone(be, reqs)


# the stream workload's own distribution (N logUniform 64..8192 monomers,
# Df 1.8-2.8, kf 1.0-1.3, 4-8 dipoles across, s in {2, 8, 32, 64}, Poisson
# arrivals at 4 req/s replayed 8x faster); the test keeps the scenes whose
# 64-column batch needs at most TEST_CAP (the bench's cap is one GPU's
# 240 GiB) so that every request can also be solved alone within minutes
TEST_CAP = 40 << 30
WS_LIMIT = 4 << 30          # column chunking of the block matvecs beyond 4 GiB of workspace


def test_stream_workload_alone_vs_in_the_mix(bdda):
    S = importlib.import_module("block_dda_amd.serving")
    T = importlib.import_module("block_dda_amd.targets")
    be = S.GpuBackend(bdda, ws_limit=WS_LIMIT)
    draw = {}
    scenes = S.synthetic_scenes(T, 5, seed=2031, nmin=64, nmax=8192, across=(4, 8),
                                fits=lambda t: (be.fresh_need(t, 3, 64) or float("inf")) <= TEST_CAP, stats=draw)
    reqs = S.synthetic_requests(T, scenes, 16, rate=4.0, seed=2032, s_choices=(2, 8, 32, 64), tol=1e-5)
    print("case serving stream: scenes %s draws %s" % ([(sc.meta["N"], round(sc.meta["Df"], 2), sc.meta["n"],
                                                          sc.meta["nocc"]) for sc in scenes], draw), flush=True)
    sv = S.Server(be, hbm_budget=200 << 30, workers=3).start()
    S.replay(sv, reqs, speed=8.0)
    assert sv.drain(timeout=400)
    st = sv.stats()
    sv.stop()
    print("stats", st, flush=True)
    assert st["failed"] == 0 and st["requests"] == 16
    assert st["merged_requests"] >= 1
    assert st["max_concurrent_batches"] >= 2
    assert st["admitted_while_busy"] >= 1
    assert st["peak_bytes"] <= 200 << 30
    _check_alone(be, reqs)

    # tightened budget: room for the largest scene's 8-column batch but not
    # for every operator at once (evictions), and a 64-column request on the
    # largest scene that cannot fit at all (refused, the rest still served)
    need8 = [be.fresh_need(sc.target, 3, 8) for sc in scenes]
    big = int(np.argmax(need8))
    budget = int(need8[big] * 1.02)
    others = [sc for i, sc in enumerate(scenes) if i != big]
    held = sum(sum(be.estimate(sc, 8)[1:3]) for sc in others)      # operators + workspace kept cached
    # building the largest scene's operator beside the others' cached ones
    # does not fit: some must be evicted.  A sparse scene on a 500x500x300
    # lattice needs more than the whole budget for its operator: refused.
    huge_t = T.Target((500, 500, 300), 5.0, np.array([0, 500 * 500 * 300 - 1], dtype=np.int32))
    huge = S.Scene(huge_t, 0.01, np.full(6, 10.0 + 1.0j))
    assert held + need8[big] > budget and be.fresh_need(huge_t, 3, 2) > budget
    sv = S.Server(be, hbm_budget=budget, workers=2).start()
    rng = np.random.default_rng(7)
    small = []
    for sc in others + [scenes[big]] + others[:2]:
        w = T.plane_waves_sphere(sc.k, 4, int(rng.integers(1 << 30)))
        small.append(sv.submit(S.Request(sc, w.rhs(sc.target)[:, :8], tol=1e-5)))
        assert sv.drain(timeout=200)
    wide = sv.submit(S.Request(huge, T.plane_waves_xy_z(0.01).rhs(huge_t), tol=1e-5))
    assert sv.drain(timeout=200)
    w = T.plane_waves_sphere(scenes[big].k, 4, 98)                       # still served afterwards
    small.append(sv.submit(S.Request(scenes[big], w.rhs(scenes[big].target), tol=1e-5)))
    assert sv.drain(timeout=200)
    st2 = sv.stats()
    sv.stop()
    print("tightened budget %d: %s" % (budget, st2), flush=True)
    assert wide.error is not None and "budget" in wide.error
    assert all(r.error is None and r.info["converged"] for r in small)
    assert st2["evictions"] >= 1 and st2["refused"] == 1
    assert st2["peak_bytes"] <= budget
