"""Serving loop: continuous batching of DDA solve requests on one GPU.

The reference solves one target at a time (block_DDA_main.cpp, Solver.cpp);
this module has no counterpart there.  It turns the solve path into a
service:

* Requests (a scene -- target, wavenumber, polarizability -- and a block of
  incident fields B) arrive at any time and queue.
* A dispatcher admits work under an HBM budget: each scene's operator (the
  transformed interaction kernel, the occupancy map and DIAG_A) is built once
  and stays resident in an LRU cache; a batch is started only when its
  operator (if not cached), its solve workspace and its Krylov blocks fit the
  budget, evicting idle operators first.
* Requests for the same scene that are pending together are merged into one
  block solve (their B columns side by side, up to `max_cols`), so the block
  iteration shares matvecs across them.
* Batches of different scenes run concurrently, one worker thread each, on
  their own operator handle and HIP stream; a request that arrives while
  others run is admitted as soon as memory and a worker free up (continuous
  batching, not a queue in front of a single solver).
* Every request records arrival, start and finish; `stats()` reports p50/p99
  latency, solves/s and right-hand sides/s.

The execution backend is pluggable: `GpuBackend` drives the HIP library
through its C ABI; tests drive the scheduler with a CPU stand-in.
"""
import collections
import ctypes
import hashlib
import math
import queue
import threading
import time

import numpy as np


class Scene:
    """One system matrix: target occupancy, wavenumber and DIAG_A."""

    def __init__(self, target, k, diag, f=3, name=""):
        self.target = target
        self.k = float(k)
        self.diag = np.ascontiguousarray(diag, dtype=np.complex128)
        self.f = f
        self.name = name or target.name
        h = hashlib.sha1()
        h.update(np.asarray(target.n, dtype=np.int64).tobytes())
        h.update(np.float64(target.lf).tobytes())
        h.update(target.address.tobytes())
        h.update(np.float64(self.k).tobytes())
        h.update(self.diag.tobytes())
        h.update(bytes([f]))
        self.key = h.hexdigest()

    @property
    def rows(self):
        return self.f * self.target.nocc


class Request:
    _ids = iter(range(1 << 62))

    def __init__(self, scene, B, tol=1e-5, arrival=None, waves=None):
        self.id = next(Request._ids)
        self.scene = scene
        self.B = np.asfortranarray(B, dtype=np.complex128).reshape(scene.rows, -1, order="F")
        self.tol = float(tol)
        self.waves = waves
        self.arrival = arrival
        self.start = None
        self.finish = None
        self.X = None
        self.info = None
        self.observables = None
        self.error = None
        self.batch_size = 0
        self.retries = 0           # batches of this request that ran out of device memory and were queued again
        self.oom = False
        self.done = threading.Event()

    @property
    def cols(self):
        return self.B.shape[1]

    @property
    def latency(self):
        return None if self.finish is None else self.finish - self.arrival


class GpuBackend:
    """Operator handles and solves through the HIP library (one stream per handle)."""

    def __init__(self, pkg, mode=None, method=None, jacobi=True, itermax=3000, ws_limit=None):
        self.pkg = pkg
        self.mode = pkg.MODE_MIXED if mode is None else mode
        self.method = pkg.COCG_RQ if method is None else method
        self.jacobi = jacobi
        self.itermax = itermax
        self.ws_limit = ws_limit

    def estimate(self, scene, cols):
        """(init_peak, resident, workspace, krylov) bytes."""
        return self.estimate_target(scene.target.n, scene.f, scene.target.nocc, cols)

    def estimate_target(self, n, f, nocc, cols):
        e = self.pkg.estimate_bytes(n, f, nocc, cols, self.mode)
        ws = e["workspace"]
        if self.ws_limit:
            # a capped handle runs its block matvecs in column chunks, but
            # never fewer than one column, and the FP64 residual matvec's
            # column is the widest
            one64 = self.pkg.estimate_bytes(n, f, nocc, 1, self.pkg.MODE_FP64)["workspace"]
            ws = min(ws, max(int(self.ws_limit), one64))
        return e["init_peak"], e["resident"], ws, e["krylov"]

    def fresh_need(self, target, f, cols):
        """Bytes a batch of `cols` columns needs when it builds the operator
        (None when the lattice is beyond what the passes can transform)."""
        try:
            init_peak, base, ws, kry = self.estimate_target(target.n, f, target.nocc, cols)
        except self.pkg.BddaError:
            return None
        return max(init_peak, base + ws + kry)

    def create(self, scene):
        t = scene.target
        op = self.pkg.MBT_fft_init(t.n, scene.f, t.lf, scene.k)
        op.set_target(t.address, scene.diag)
        if self.ws_limit:
            op.set_workspace_limit(self.ws_limit)
        return op

    def release(self, op):
        op.close()

    def trim(self):
        """Give the device allocator's idle memory back to the runtime."""
        self.pkg.dev_trim()

    def solve(self, op, B, tol):
        X, info, _ = op.solve(B, tol, self.itermax, method=self.method, jacobi=self.jacobi, mode=self.mode)
        return X, info

    def observables(self, op, B, X):
        xs = op.cross_sections(B, X)
        return {q: v.tolist() for q, v in xs.items()}

    def solve_batch(self, op, B, tol, cols):
        """Solve a merged block and the cross sections of each request's
        columns (widths `cols`) with B and X kept in HBM in between: B goes up
        once, X comes down once, and each request's cross sections run on its
        column slice of the device blocks.  Returns X, info, [observables]."""
        pkg = self.pkg
        L = pkg.lib()
        B = np.asfortranarray(B, dtype=np.complex128)
        rows, s = B.shape
        nb = rows * s * 16
        dB, dX = ctypes.c_void_p(), ctypes.c_void_p()
        dp = ctypes.POINTER(ctypes.c_double)
        pkg._check(L.bdda_dev_alloc(nb, ctypes.byref(dB)))
        try:
            pkg._check(L.bdda_dev_alloc(nb, ctypes.byref(dX)))
            try:
                pkg._check(L.bdda_memcpy_h2d(dB, B.ctypes.data, nb))
                info = pkg.SolveInfo()
                rc = L.bdda_solve_dev(op.h, int(self.method), int(self.jacobi), int(self.mode), dB, s, float(tol),
                                      int(self.itermax), dX, ctypes.byref(info), None, 0)
                if rc not in (0, 5):  # a non-converged solve still returns X
                    pkg._check(rc)
                X = np.empty((rows, s), dtype=np.complex128, order="F")
                pkg._check(L.bdda_memcpy_d2h(X.ctypes.data, dX, nb))
                obs = []
                c0 = 0
                for c in cols:
                    out = [np.zeros(c) for _ in range(3)]
                    off = c0 * rows * 16
                    pkg._check(L.bdda_cross_sections_dev(op.h, ctypes.c_void_p(dB.value + off),
                                                         ctypes.c_void_p(dX.value + off), int(c), pkg.ALPHA_FIRST,
                                                         *[o.ctypes.data_as(dp) for o in out]))
                    obs.append({"Cext": out[0].tolist(), "Cabs": out[1].tolist(), "Csca": out[2].tolist()})
                    c0 += c
                return X, info.as_dict(), obs
            finally:
                L.bdda_dev_free(dX)
        finally:
            L.bdda_dev_free(dB)


def _out_of_memory(e):
    """The library's ENOMEM status (or a message that says so)."""
    return getattr(e, "code", None) == 2 or "out of memory" in str(e).lower()


class _Entry:
    def __init__(self, key, handle, base):
        self.key = key
        self.handle = handle
        self.base = base           # operator + target bytes
        self.busy = False
        self.last_used = time.perf_counter()


class Server:
    """Dispatcher + workers over a backend; see the module docstring.

    Memory model (bytes): every cached operator holds base; a running batch
    reserves its pass workspace and Krylov blocks (the library gives both
    back when the solve returns) and, while its operator is being built, the
    build transient.  A batch starts only when
    the cached operators plus all reservations stay within hbm_budget after
    evicting idle operators (least recently used first).  A scene that has
    waited longer than max_wait_s blocks admissions behind it until it fits.
    """

    def __init__(self, backend, hbm_budget, workers=4, max_cols=64, merge=True, max_wait_s=2.0, on_batch=None,
                 oom_retries=1):
        self.backend = backend
        self.oom_retries = int(oom_retries)   # times a request is queued again after running out of device memory
        self.requeued = 0
        self.on_batch = on_batch      # called by the worker with (scene key, finished requests, seconds) after each batch
        self.budget = int(hbm_budget)
        self.max_cols = int(max_cols)
        self.merge = merge
        self.nworkers = int(workers)
        self.max_wait_s = float(max_wait_s)
        self._cv = threading.Condition()
        self._pending = collections.OrderedDict()     # request id -> request, arrival order
        self._cache = collections.OrderedDict()       # scene key -> _Entry, LRU order
        self._inflight = set()                        # scene keys with a batch queued or running
        self._reserved = 0
        self._work = queue.Queue()
        self._threads = []
        self._running = False
        self.completed = []
        self.batches = []       # (scene key, request ids, t_start, t_end, worker)
        self.peak_bytes = 0
        self.evictions = 0

    # ---------------------------------------------------------------- API
    def start(self):
        self._running = True
        self._threads = [threading.Thread(target=self._dispatch, daemon=True)]
        self._threads += [threading.Thread(target=self._worker, args=(w,), daemon=True)
                          for w in range(self.nworkers)]
        for t in self._threads:
            t.start()
        return self

    def submit(self, req):
        if req.arrival is None:
            req.arrival = time.perf_counter()
        with self._cv:
            self._pending[req.id] = req
            self._cv.notify_all()
        return req

    def drain(self, timeout=None):
        t0 = time.perf_counter()
        with self._cv:
            while self._pending or self._inflight:
                left = None if timeout is None else timeout - (time.perf_counter() - t0)
                if left is not None and left <= 0:
                    return False
                self._cv.wait(0.05 if left is None else min(left, 0.05))
        return True

    def stop(self):
        self.drain()
        with self._cv:
            self._running = False
            self._cv.notify_all()
        for _ in range(self.nworkers):
            self._work.put(None)
        for t in self._threads:
            t.join()
        with self._cv:
            for e in self._cache.values():
                self.backend.release(e.handle)
            self._cache.clear()

    def used_bytes(self):
        return sum(e.base for e in self._cache.values()) + self._reserved

    def stats(self):
        ok = [r for r in self.completed if r.error is None]
        if not ok:
            return {"requests": len(self.completed), "failed": len(self.completed)}
        lat = np.array([r.latency for r in ok])
        # service time: from the start of a request's batch to its finish,
        # without the time it queued (latency = queueing + service)
        svc = np.array([r.finish - r.start for r in ok])
        t0 = min(r.arrival for r in self.completed)
        t1 = max(r.finish for r in self.completed)
        span = max(t1 - t0, 1e-9)
        return {
            "requests": len(self.completed), "failed": len(self.completed) - len(ok),
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max()),
            "service_p50_s": float(np.percentile(svc, 50)), "service_p99_s": float(np.percentile(svc, 99)),
            "solves_per_s": len(ok) / span,
            "rhs_per_s": sum(r.cols for r in ok) / span,
            "batches": len(self.batches),
            "merged_requests": sum(len(b[1]) for b in self.batches) - len(self.batches),
            "max_concurrent_batches": self._max_overlap(),
            "admitted_while_busy": self._admitted_while_busy(ok),
            "refused": sum(1 for r in self.completed if r.error is not None and "budget" in r.error),
            "peak_bytes": self.peak_bytes, "budget_bytes": self.budget, "evictions": self.evictions,
            "errors": sorted({r.error[:200] for r in self.completed if r.error is not None})[:3],
            "requeued_after_oom": self.requeued,
        }

    def _admitted_while_busy(self, reqs):
        """Requests that arrived while another batch was on the device and
        started before that batch finished (continuous batching)."""
        n = 0
        for r in reqs:
            for key, ids, t0, t1, _ in self.batches:
                if r.id not in ids and t0 <= r.arrival and r.start < t1:
                    n += 1
                    break
        return n

    # ---------------------------------------------------------------- internals
    def _max_overlap(self):
        ev = sorted([(b[2], 1) for b in self.batches] + [(b[3], -1) for b in self.batches])
        cur = best = 0
        for _, d in ev:
            cur += d
            best = max(best, cur)
        return best

    def _evict_for(self, need, keep=None):
        """Drop idle cached operators, least recently used first, until `need`
        fits; the operator of scene `keep` (the one being admitted, whose
        need assumes it stays) is never dropped."""
        for key in list(self._cache.keys()):
            if self.used_bytes() + need <= self.budget:
                return True
            e = self._cache[key]
            if key != keep and not e.busy and key not in self._inflight:
                self._drop(key)
        return self.used_bytes() + need <= self.budget

    def _drop(self, key):
        e = self._cache.pop(key)
        self.backend.release(e.handle)
        self.evictions += 1

    def _take_batch(self, now):
        """Oldest pending scene that is idle and fits (its requests merged), or None."""
        seen = set()
        for req in list(self._pending.values()):
            key = req.scene.key
            if key in seen:
                continue
            seen.add(key)
            if key in self._inflight:
                continue
            group, cols = [], 0
            for r in self._pending.values():
                if r.scene.key != key:
                    continue
                if group and (not self.merge or cols + r.cols > self.max_cols):
                    break
                group.append(r)
                cols += r.cols
            try:
                est = self.backend.estimate(req.scene, cols)
            except Exception as e:   # e.g. a lattice beyond what the passes transform
                for r in group:
                    r.error = "refused: %r" % (e,)
                return key, group, 0, None
            init_peak, base, ws, kry = est
            fresh = max(init_peak, base + ws + kry)   # need when the operator is built for this batch
            e = self._cache.get(key)
            if e is not None:
                # (the cached operator holds base; the batch needs its
                # workspace and Krylov blocks, never more than a fresh build)
                need = ws + kry
                self._cache.move_to_end(key)
                if self._evict_for(need, keep=key):
                    return key, group, need, est
            else:
                need = fresh
                if self._evict_for(need):
                    return key, group, need, est
            if self._reserved == 0 and fresh > self.budget and all(k == key for k in self._cache):
                for r in group:     # alone on an idle device and still too large
                    r.error = "needs %d bytes, budget %d" % (fresh, self.budget)
                return key, group, 0, None
            if now - req.arrival > self.max_wait_s:
                return None         # starving: admit nothing else until it fits
        return None

    def _dispatch(self):
        while True:
            with self._cv:
                item = None
                while self._running:
                    if self._pending and len(self._inflight) < self.nworkers:
                        item = self._take_batch(time.perf_counter())
                        if item is not None:
                            break
                    self._cv.wait(0.05)
                if not self._running:
                    return
                key, group, need, est = item
                for r in group:
                    del self._pending[r.id]
                if est is None:
                    t = time.perf_counter()
                    for r in group:
                        r.finish = t
                        self.completed.append(r)
                        r.done.set()
                    self._cv.notify_all()
                    continue
                self._reserved += need
                self._inflight.add(key)
                if key in self._cache:
                    self._cache[key].busy = True
                self.peak_bytes = max(self.peak_bytes, self.used_bytes())
            self._work.put((key, group, need, est))

    def _worker(self, wid):
        while True:
            item = self._work.get()
            if item is None:
                return
            key, group, need, (init_peak, base, ws, kry) = item
            t0 = time.perf_counter()
            for r in group:
                r.start = t0
                r.batch_size = len(group)
            scene = group[0].scene
            try:
                with self._cv:
                    entry = self._cache.get(key)
                if entry is None:
                    handle = self.backend.create(scene)
                    entry = _Entry(key, handle, base)
                    entry.busy = True
                    with self._cv:
                        # build transient over: the operator is cached, the
                        # batch keeps its workspace and Krylov reservation
                        self._cache[key] = entry
                        self._reserved -= need
                        need = ws + kry
                        self._reserved += need
                        self.peak_bytes = max(self.peak_bytes, self.used_bytes())
                B = np.concatenate([r.B for r in group], axis=1) if len(group) > 1 else group[0].B
                tol = min(r.tol for r in group)
                if hasattr(self.backend, "solve_batch"):
                    X, info, obs = self.backend.solve_batch(entry.handle, B, tol, [r.cols for r in group])
                else:
                    X, info = self.backend.solve(entry.handle, B, tol)
                    obs = None
                c0 = 0
                for i, r in enumerate(group):
                    r.X = X[:, c0:c0 + r.cols]
                    r.info = info
                    r.observables = obs[i] if obs is not None else self.backend.observables(entry.handle, r.B, r.X)
                    c0 += r.cols
            except Exception as e:  # reported per request; the server keeps running
                for r in group:
                    r.error = repr(e)
                    r.oom = _out_of_memory(e)
            t1 = time.perf_counter()
            # out of device memory (the estimates are estimates; idle memory
            # can be held by the allocator): idle memory goes back to the
            # runtime and the requests queue again, ahead of later arrivals
            retry = [r for r in group if r.error is not None and r.oom and r.retries < self.oom_retries]
            if retry and hasattr(self.backend, "trim"):
                try:
                    self.backend.trim()
                except Exception:
                    pass
            with self._cv:
                self._reserved -= need
                self._inflight.discard(key)
                e = self._cache.get(key)
                if e is not None:
                    e.busy = False
                    e.last_used = t1
                done = [r for r in group if r not in retry]
                for r in retry:
                    r.retries += 1
                    r.error = None
                    r.oom = False
                    r.start = None
                if retry:
                    self.requeued += len(retry)
                    self._pending = collections.OrderedDict(
                        [(r.id, r) for r in retry] + list(self._pending.items()))
                for r in done:
                    r.finish = t1
                    self.completed.append(r)
                self.batches.append((key, [r.id for r in group], t0, t1, wid))
                self._cv.notify_all()
            for r in done:
                r.done.set()
            if self.on_batch is not None and done:
                self.on_batch(key, done, t1 - t0)


# -------------------------------------------------------------------- streams

def log_uniform_int(rng, lo, hi):
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


def synthetic_scenes(pkg_targets, nscenes, seed, nmin=64, nmax=8192, across=(4, 8), f=3, max_grid=None, fits=None,
                     stats=None):
    """Seeded scenes: aggregates with N ~ logUniform[nmin, nmax] monomers,
    Df U[1.8, 2.8], kf U[1.0, 1.3], a0 = 20 nm, dipoles across U{across},
    wavelength U[300, 1000] nm, m = U[1.3, 2.5] + i U[0, 1]; LDR polarizability.
    Aggregation: cluster-cluster where its joins hold the law (Df <= 2.2),
    particle-cluster above (targets.aggregate_target).
    A draw is drawn again when `fits(target)` is false (a memory cap such as
    "a 64-column batch fits one GPU", see GpuBackend.fresh_need: an open
    Df = 1.8 aggregate of thousands of monomers at 8 dipoles across spans
    over a thousand cells per axis, beyond one GPU) or, if given, when its
    lattice exceeds max_grid cells on an axis.  `stats` (a dict) receives
    the number of draws and of rejected draws."""
    T = pkg_targets
    rng = np.random.default_rng(seed)
    scenes = []
    draws = rejected = 0
    for i in range(nscenes):
        while True:
            N = log_uniform_int(rng, nmin, nmax)
            Df = rng.uniform(1.8, 2.8)
            kf = rng.uniform(1.0, 1.3)
            ac = int(rng.integers(across[0], across[1] + 1))
            tgt = T.aggregate_target(N, Df, kf, 20.0, ac, int(rng.integers(1 << 30)))
            draws += 1
            if (max_grid is None or max(tgt.n) <= max_grid) and (fits is None or fits(tgt)):
                break
            rejected += 1
        wl = rng.uniform(300.0, 1000.0)
        m = complex(rng.uniform(1.3, 2.5), rng.uniform(0.0, 1.0))
        k = 2 * math.pi / wl
        alpha = T.alpha_ldr(m, tgt.lf, k)
        diag = T.diag_from_alpha(alpha, tgt.nocc, f)
        sc = Scene(tgt, k, diag, f, name="scene%d_N%d_wl%.0f" % (i, N, wl))
        sc.meta = {"N": N, "Df": Df, "kf": kf, "across": ac, "wl": wl, "m": [m.real, m.imag], "nocc": tgt.nocc,
                   "n": list(tgt.n), "aggregation": getattr(tgt, "aggregation", None)}
        scenes.append(sc)
    if stats is not None:
        stats.update({"draws": draws, "rejected": rejected, "rejected_share": rejected / max(draws, 1)})
    return scenes


def synthetic_requests(pkg_targets, scenes, nreq, rate, seed, s_choices=(2, 8, 32, 64), tol=1e-5):
    """Poisson arrivals (mean rate req/s) over the scenes; each request has s
    right-hand sides: s/2 seeded directions x 2 polarisations."""
    T = pkg_targets
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(nreq):
        t += rng.exponential(1.0 / rate)
        sc = scenes[int(rng.integers(len(scenes)))]
        s = int(s_choices[int(rng.integers(len(s_choices)))])
        waves = T.plane_waves_sphere(sc.k, max(1, s // 2), int(rng.integers(1 << 30)))
        B = waves.rhs(sc.target)[:, :s]
        r = Request(sc, B, tol=tol, waves=waves)
        r.offset = t
        reqs.append(r)
    return reqs


def replay(server, reqs, speed=1.0):
    """Submit each request at its arrival offset (seconds / speed) from now."""
    t0 = time.perf_counter()
    for r in reqs:
        due = t0 + r.offset / speed
        now = time.perf_counter()
        if due > now:
            time.sleep(due - now)
        r.arrival = time.perf_counter()
        server.submit(r)
    return t0
