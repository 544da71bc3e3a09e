"""Scheduler of the serving loop (block-dda_amd/serving.py) on a CPU stand-in
backend: every request completes with its own columns, same-scene requests
merge, batches of different scenes overlap, the HBM budget holds at every
admission, idle operators are evicted LRU-first, and a request too large
for the budget fails alone without stopping the server."""
import importlib
import threading
import time

import numpy as np
import pytest


@pytest.fixture(scope="module")
def S(bdda):
    return importlib.import_module("block_dda_amd.serving")


@pytest.fixture(scope="module")
def T(bdda):
    return importlib.import_module("block_dda_amd.targets")


class FakeBackend:
    """Sizes proportional to the scene; a solve sleeps and returns 2 B."""

    def __init__(self, base=100, ws_per_col=5, kry_per_col=3, solve_s=0.02):
        self.base, self.ws_per_col, self.kry_per_col, self.solve_s = base, ws_per_col, kry_per_col, solve_s
        self.created = []
        self.released = []
        self.lock = threading.Lock()
        self.live = 0
        self.max_live = 0
        # bytes this backend actually holds, as a device would: the build
        # transient while an operator is created, then its base, and the
        # workspace and Krylov blocks of running solves
        self.held = 0
        self.max_held = 0

    def _hold(self, d):
        with self.lock:
            self.held += d
            self.max_held = max(self.max_held, self.held)

    def estimate(self, scene, cols):
        b = self.base * scene.target.nocc // 10
        return 2 * b, b, self.ws_per_col * cols, self.kry_per_col * cols

    def create(self, scene):
        init_peak, base, _, _ = self.estimate(scene, 1)
        self._hold(init_peak)
        time.sleep(0.005)
        self._hold(base - init_peak)
        with self.lock:
            self.created.append(scene.key)
        return {"scene": scene, "base": base}

    def release(self, h):
        self._hold(-h["base"])
        with self.lock:
            self.released.append(h["scene"].key)

    def solve(self, h, B, tol):
        _, _, ws, kry = self.estimate(h["scene"], B.shape[1])
        self._hold(ws + kry)
        with self.lock:
            self.live += 1
            self.max_live = max(self.max_live, self.live)
        time.sleep(self.solve_s)
        with self.lock:
            self.live -= 1
        self._hold(-(ws + kry))
        return 2.0 * B, {"iters": 1, "converged": 1}

    def observables(self, h, B, X):
        return {"Cext": [float(np.abs(X).sum())]}


def scenes(S, T, n):
    out = []
    for i in range(n):
        t = T.Target((4, 4, 4), 0.5, np.arange(10 + i, dtype=np.int32))
        out.append(S.Scene(t, 1.0 + 0.1 * i, np.ones(3 * t.nocc) * (5 + i)))
    return out


def req(S, sc, cols, seed):
    rng = np.random.default_rng(seed)
    return S.Request(sc, rng.standard_normal((sc.rows, cols)) + 0j)


def test_all_complete_with_their_columns_and_merge(S, T):
    be = FakeBackend()
    sv = S.Server(be, hbm_budget=10 ** 6, workers=3).start()
    sc = scenes(S, T, 3)
    # hold the workers busy so that same-scene requests pile up and merge
    reqs = [req(S, sc[i % 3], 2 + (i % 2) * 6, i) for i in range(12)]
    for r in reqs:
        sv.submit(r)
    assert sv.drain(timeout=30)
    st = sv.stats()
    sv.stop()
    assert st["requests"] == 12 and st["failed"] == 0
    for r in reqs:
        assert np.array_equal(r.X, 2.0 * r.B)
        assert r.finish >= r.start >= r.arrival
    assert st["batches"] < 12 and st["merged_requests"] > 0
    assert st["max_concurrent_batches"] >= 2 and be.max_live >= 2
    assert sorted(set(be.created)) == sorted(s.key for s in sc)   # one operator per scene


def test_budget_and_lru_eviction(S, T):
    be = FakeBackend(solve_s=0.01)
    sc = scenes(S, T, 4)
    # each operator: base = 100 .. 130 (+ build transient 2x), budget fits about two
    sv = S.Server(be, hbm_budget=400, workers=2).start()
    reqs = [req(S, sc[i % 4], 2, i) for i in range(16)]
    for r in reqs:
        sv.submit(r)
        time.sleep(0.003)
    assert sv.drain(timeout=30)
    st = sv.stats()
    sv.stop()
    assert st["failed"] == 0
    assert st["peak_bytes"] <= 400
    assert be.max_held <= 400, be.max_held       # what the backend really held
    assert st["evictions"] > 0 and len(be.released) >= st["evictions"]


def test_cached_scene_with_a_larger_request(S, T):
    """A scene's operator is cached from a small request; a wider request of
    the same scene does not fit beside it.  Its operator must not be evicted
    under a need that assumed it stays: the server refuses the request.  A
    cached operator then serves requests within its own base plus the
    batch's workspace and Krylov blocks."""
    sc = scenes(S, T, 1)[0]               # base 100, build transient 200
    # budget 250: the 20-column request needs 100 + 100 + 60 = 260 fresh -> refused
    be = FakeBackend(solve_s=0.005)
    sv = S.Server(be, hbm_budget=250, workers=1).start()
    small = sv.submit(req(S, sc, 2, 0))
    assert sv.drain(timeout=10)
    wide = sv.submit(req(S, sc, 20, 1))
    assert sv.drain(timeout=10)
    sv.stop()
    assert small.error is None
    assert wide.error is not None and "budget" in wide.error
    assert be.max_held <= 250, be.max_held
    assert be.created.count(sc.key) == 1
    # a handle keeps no workspace between solves: after a 20-column request
    # (workspace 100), a budget shrunk to 120 still serves a 2-column request
    # beside the cached operator (100 + 10 + 6), without building it again
    be = FakeBackend(solve_s=0.005)
    sv = S.Server(be, hbm_budget=300, workers=1).start()
    wide = sv.submit(req(S, sc, 20, 0))
    assert sv.drain(timeout=10)
    with sv._cv:
        sv.budget = 120
    small = sv.submit(req(S, sc, 2, 1))
    assert sv.drain(timeout=10)
    st = sv.stats()
    sv.stop()
    assert small.error is None and wide.error is None
    assert np.array_equal(small.X, 2.0 * small.B)
    assert be.max_held <= 300, be.max_held
    assert be.created.count(sc.key) == 1 and st["evictions"] == 0
    assert be.held == 0


class OomError(RuntimeError):
    code = 2   # the library's ENOMEM status


class OomBackend(FakeBackend):
    """Runs out of device memory on the first `fails` solves of each scene;
    counts the trims the server asks for."""

    def __init__(self, fails=1, **kw):
        super().__init__(**kw)
        self.fails = fails
        self.seen = {}
        self.trims = 0

    def trim(self):
        with self.lock:
            self.trims += 1

    def solve(self, h, B, tol):
        key = h["scene"].key
        with self.lock:
            n = self.seen.get(key, 0)
            self.seen[key] = n + 1
        if n < self.fails:
            raise OomError("out of memory")
        return super().solve(h, B, tol)


def test_out_of_memory_batches_queue_again(S, T):
    """A batch that runs out of device memory gives idle memory back (trim)
    and its requests queue again, ahead of later ones; a request is retried
    once, so a second failure is reported."""
    be = OomBackend(fails=1)
    sv = S.Server(be, hbm_budget=10 ** 6, workers=2).start()
    sc = scenes(S, T, 2)
    reqs = [req(S, sc[i % 2], 2, i) for i in range(6)]
    for r in reqs:
        sv.submit(r)
    assert sv.drain(timeout=30)
    st = sv.stats()
    sv.stop()
    assert st["failed"] == 0 and st["requests"] == 6
    assert st["requeued_after_oom"] >= 2 and be.trims >= 2
    assert all(np.array_equal(r.X, 2.0 * r.B) for r in reqs)
    assert sum(r.retries for r in reqs) == st["requeued_after_oom"]
    assert all(r.retries <= 1 for r in reqs)
    # a scene that keeps running out of memory: reported after one retry
    be = OomBackend(fails=5)
    sv = S.Server(be, hbm_budget=10 ** 6, workers=1).start()
    r = sv.submit(req(S, sc[0], 2, 9))
    assert sv.drain(timeout=30)
    st = sv.stats()
    sv.stop()
    assert r.error is not None and "out of memory" in r.error and r.retries == 1
    assert st["requests"] == 1 and st["failed"] == 1


def test_too_large_request_fails_alone(S, T):
    be = FakeBackend()
    sc = scenes(S, T, 2)
    sv = S.Server(be, hbm_budget=250, workers=2).start()
    big = req(S, sc[1], 64, 0)            # 110 base + 64 * 8 > 250
    small = req(S, sc[0], 2, 1)
    sv.submit(big)
    sv.submit(small)
    assert sv.drain(timeout=30)
    sv.stop()
    assert big.error is not None and "budget" in big.error
    assert small.error is None and np.array_equal(small.X, 2.0 * small.B)


class RefusingBackend(FakeBackend):
    """Estimates fail for scenes of more than 12 dipoles (as the library
    refuses lattices beyond the passes' limit)."""

    def estimate(self, scene, cols):
        if scene.target.nocc > 12:
            raise RuntimeError("EUNSUPPORTED: padded grid beyond the passes' limit")
        return super().estimate(scene, cols)


def test_unsupported_scene_is_refused(S, T):
    be = RefusingBackend()
    sc = scenes(S, T, 4)                  # nocc 10 .. 13
    sv = S.Server(be, hbm_budget=10 ** 6, workers=2).start()
    bad = sv.submit(req(S, sc[3], 2, 0))
    good = sv.submit(req(S, sc[0], 2, 1))
    assert sv.drain(timeout=10)
    st = sv.stats()
    sv.stop()
    assert bad.error is not None and "EUNSUPPORTED" in bad.error
    assert good.error is None and st["failed"] == 1


def test_no_merge_mode_and_stream_generator(S, T):
    be = FakeBackend(solve_s=0.002)
    sc = scenes(S, T, 2)
    sv = S.Server(be, hbm_budget=10 ** 6, workers=1, merge=False).start()
    reqs = [req(S, sc[0], 2, i) for i in range(5)]
    for r in reqs:
        sv.submit(r)
    assert sv.drain(timeout=30)
    st = sv.stats()
    sv.stop()
    assert st["batches"] == 5 and st["merged_requests"] == 0
    # Poisson offsets increase, s from the choices, B = plane waves of the scene
    rq = S.synthetic_requests(T, sc, 20, rate=4.0, seed=3, s_choices=(2, 8))
    offs = [r.offset for r in rq]
    assert all(b > a for a, b in zip(offs, offs[1:]))
    assert {r.cols for r in rq} <= {2, 8}
    assert abs(np.mean(np.diff([0.0] + offs)) - 0.25) < 0.15


class BatchBackend(FakeBackend):
    """A backend whose solve_batch returns X = 3 B and, per request, the column
    sums of its slice of X: the worker must hand each request its own slice
    of both."""

    def solve_batch(self, h, B, tol, cols):
        X, info = self.solve(h, B, tol)
        X = 1.5 * X
        obs, c0 = [], 0
        for c in cols:
            obs.append({"Cext": np.abs(X[:, c0:c0 + c]).sum(axis=0).tolist()})
            c0 += c
        return X, info, obs

    def observables(self, h, B, X):  # not used when solve_batch exists
        raise AssertionError("observables called beside solve_batch")


def test_solve_batch_slices_per_request(S, T):
    be = BatchBackend(solve_s=0.05)
    sv = S.Server(be, hbm_budget=10 ** 6, workers=1).start()
    sc = scenes(S, T, 1)[0]
    # the first request occupies the worker; the next three merge into one batch
    reqs = [req(S, sc, c, 100 + i) for i, c in enumerate((2, 3, 1, 4))]
    for r in reqs:
        sv.submit(r)
        time.sleep(0.002)
    assert sv.drain(timeout=10)
    sv.stop()
    assert any(r.batch_size > 1 for r in reqs)
    for r in reqs:
        assert r.error is None
        np.testing.assert_allclose(r.X, 3.0 * r.B)
        np.testing.assert_allclose(r.observables["Cext"], np.abs(3.0 * r.B).sum(axis=0))
