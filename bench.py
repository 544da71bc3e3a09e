"""Benchmark of the MI355X block-DDA solve path.

A step is one time-to-solution: the block COCG-rq solve of the whole
right-hand-side block to a true FP64 relative residual ||B - A X||/||B|| of
1e-5, with B already resident in HBM and the transformed interaction kernel
cached (the reference also builds it before its solve timer,
Solver.cpp:55,59).  The default workload is the 16384-monomer aggregate
(a0 = 20 nm, Df = 2.8, kf = 1.0, seeded tunable aggregation, 4 dipoles across
each monomer, m = 2.3 + 0.6i, lambda = 400 nm, LDR polarizability, s = 16:
8 seeded directions x 2 polarisations), solved in mixed precision (complex
FP32 inner block iteration inside an FP64 defect-correction loop); the same
solve in pure FP64 is timed beside it.

Printed JSON line: metric/value (mixed time-to-solution, seconds), the FP64
time-to-solution and the speedup, a roofline object for the dominant kernel
(the fused x pass of the FP32 block matvec, HIP events on the library
stream), the sphere workload's cross sections and S11/S12 against Mie
theory, and a cpu_baseline measured on this host with the oracle (the FP64
C restatement of the reference algorithm: one 1-D FFT of length
f^2 prod(2n-1) per RHS per matvec, OpenMP over RHS columns).
"""
import argparse
import ctypes
import importlib
import importlib.util
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "time-to-solution & p99 latency at rel.res 1e-5; mixed-prec vs FP64 on MI355X"
HBM_PEAK_GBS = 8000.0
# Hardware counters cannot be read from inside this process: the FP32 x
# pass's HBM bytes per launch (FETCH_SIZE + WRITE_SIZE scaled by a 2 GiB
# device copy) and VALU wave instructions (SQ_INSTS_VALU) come from the
# committed profile of the workload, and only when that profile was taken on
# the kernel sources built now (their SHA-256 is recorded in it).
KERNEL_SOURCES = ["bdda_common.hpp", "fft_engine.hpp", "fft_plan.hpp", "mbt_plans.hpp"]


def kernel_source_stamp():
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "block-dda_amd", "csrc", name), "rb") as fh:
            h.update(name.encode() + b"\0" + fh.read())
    return h.hexdigest()


def profile_counters(config):
    """Counter figures of profiles/mbt_matvec_<config>_hbm_traffic.md (its
    "bench counters" block) and whether they describe the current kernel."""
    path = os.path.join(ROOT, "profiles", "mbt_matvec_%s_hbm_traffic.md" % config)
    if not os.path.exists(path):
        return None, "no committed counter profile for %s" % config
    vals, inside = {}, False
    for line in open(path):
        s = line.strip()
        if s == "bench counters:":
            inside = True
        elif inside and ":" in s:
            k, v = s.split(":", 1)
            vals[k.strip()] = v.strip()
        elif inside and not s:
            break
    if "source_sha256" not in vals:
        return None, "profile has no source stamp"
    if vals["source_sha256"] != kernel_source_stamp():
        return None, "profile stale: taken on other x-pass sources than the ones built now"
    return vals, "profile of the current x-pass sources"


def load_pkg():
    name = "block_dda_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = os.path.join(ROOT, "block-dda_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


METHOD_NAMES = {"cocg_rq": "COCG-rq", "bicgstab": "BiCGSTAB", "bicggr": "BiCGGR"}


def build_workload(name):
    W = importlib.import_module("block_dda_amd.workloads").build(name)
    log("workload: %s (%.1fs)" % (W.describe(), W.seconds))
    return W


def occupied_lines(n, address):
    """(ix, iy) lines that hold an occupied cell: the z passes and the y passes'
    short side move only their rows."""
    return int(np.unique(np.asarray(address, dtype=np.int64) // n[2]).size)


def pass_bytes(n, P, nocc, s, f, c, nc, nlines=None):
    nx, ny, nz = n
    Px, Py, Pz = P
    nxnyPz = (nx * ny if nlines is None else nlines) * Pz
    nxPyPz = nx * Py * Pz
    gh = nc * (Px // 2 + 1) * (Py // 2 + 1) * (Pz // 2 + 1) * c
    return {
        "z_fwd": s * f * c * (nocc + nxnyPz),
        "y_fwd": s * f * c * (nxnyPz + nxPyPz),
        "x_fused": s * f * c * 2 * nxPyPz + gh,
        "y_inv": s * f * c * (nxPyPz + nxnyPz),
        "z_inv": s * f * c * (nxnyPz + 2 * nocc),
    }


def time_matvec(pkg, op, rows, s, prec, reps):
    L = pkg.lib()
    c = 8 if prec == pkg.BDDA_PREC_FP32 else 16
    dX = ctypes.c_void_p()
    dY = ctypes.c_void_p()
    nbytes = rows * s * c
    pkg._check(L.bdda_dev_alloc(nbytes, ctypes.byref(dX)))
    pkg._check(L.bdda_dev_alloc(nbytes, ctypes.byref(dY)))
    rng = np.random.default_rng(0)
    if prec == pkg.BDDA_PREC_FP32:
        h = (rng.standard_normal((rows * s, 2)).astype(np.float32)).ravel()
    else:
        h = rng.standard_normal(rows * s * 2)
    pkg._check(L.bdda_memcpy_h2d(dX, h.ctypes.data, nbytes))
    pkg._check(L.bdda_mbt_set_timing(op.h, 1))
    names = ["z_fwd", "y_fwd", "x_fused", "y_inv", "z_inv"]
    acc = np.zeros(5)
    pt = (ctypes.c_float * 5)()
    pkg._check(L.bdda_mbt_apply_occupied_dev(op.h, prec, 1, dX, dY, s))  # warm-up
    t0 = time.perf_counter()
    for _ in range(reps):
        pkg._check(L.bdda_mbt_apply_occupied_dev(op.h, prec, 1, dX, dY, s))
        pkg._check(L.bdda_mbt_pass_times(op.h, pt))
        acc += np.array(pt[:])
    wall = (time.perf_counter() - t0) / reps
    pkg._check(L.bdda_mbt_set_timing(op.h, 0))
    L.bdda_dev_free(dX)
    L.bdda_dev_free(dY)
    ms = acc / reps
    return dict(zip(names, ms.tolist())), wall * 1e3


def solve_dev(pkg, op, dB, dX, s, tol, itermax, mode, method=0, jacobi=1):
    L = pkg.lib()
    info = pkg.SolveInfo()
    pkg._check(L.bdda_device_sync())
    t0 = time.perf_counter()
    rc = L.bdda_solve_dev(op.h, method, jacobi, mode, dB, s, tol, itermax, dX, ctypes.byref(info), None, 0)
    pkg._check(L.bdda_device_sync())
    wall = time.perf_counter() - t0
    if rc not in (0, 5):
        pkg._check(rc)
    d = info.as_dict()
    d["wall_s"] = wall
    return d


def residual_col_max(pkg, op, dB, dX, s):
    """Largest per-column FP64 true residual ||B_j - A X_j|| / ||B_j|| (outside the timed region)."""
    res = np.zeros(s)
    out = res.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    pkg._check(pkg.lib().bdda_residual_columns_dev(op.h, dB, dX, s, out))
    return float(res.max())


def host_memory():
    """(total, available) host bytes from /proc/meminfo."""
    vals = {}
    with open("/proc/meminfo") as fh:
        for line in fh:
            k, v = line.split(":", 1)
            vals[k] = int(v.split()[0]) * 1024
    return vals.get("MemTotal", 0), vals.get("MemAvailable", 0)


def cpu_baseline(tgt, s, n_matvecs, threads):
    """The oracle's MBT matvec (FP64 C restatement of MBT_fft_mvp: BT_pad, one
    1-D FFT of length F = 9 prod(2n - 1), the product, the inverse FFT, 1/F,
    BT_reconstruct) on ONE column at the workload's own F, timed on this
    host; scaled over the s columns and the FP64 GPU solve's matvec count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    F = orc.fft_length(tgt.n, 3)
    total, avail = host_memory()
    # Au_til, the two F-length work vectors and the four-step transpose buffer
    need = 4 * F * 16
    if need > 0.8 * avail:
        return {"value": None, "unit": "s", "cores": threads, "kind": "port",
                "sample": "not run: one real-shape column needs %.1f GB, %.1f GB available" % (need / 1e9,
                                                                                          avail / 1e9)}
    orc.fft_config(1 << 22, threads)
    # the timed product does not depend on Au_til's values, and the
    # reference builds Au_til before its solve timer (Solver.cpp:55,59)
    Au = np.full(F, 1e-3 + 0.0j)
    x = np.random.default_rng(1).standard_normal(3 * int(np.prod(tgt.n))) + 0j
    t0 = time.perf_counter()
    orc.mbt_fft_mvp(tgt.n, 3, Au, x)
    t_col = time.perf_counter() - t0
    del Au
    return {
        "value": t_col * s * n_matvecs, "unit": "s", "cores": threads, "kind": "port",
        "sample": ("oracle MBT matvec of one column at the workload's own shape (lattice %s, 1-D FFT of length "
                   "F=%d, four-step on %d OpenMP threads): %.1f s measured; x %d columns x %d block matvecs of "
                   "the FP64 GPU solve (extrapolated over columns and matvecs only; solve only)"
                   % (list(tgt.n), F, threads, t_col, s, n_matvecs)),
        "column_s": t_col, "F": F, "host_nproc": os.cpu_count(), "host_ram_bytes": total,
        "host_ram_available_bytes": avail,
    }


def observables_dev(pkg, op, dB, dX, waves, s, nang=181):
    """Cross sections (all columns) and S11/S12 (one plane per incident
    direction) of the device block dX, FP64 on the GPU."""
    L = pkg.lib()
    out = [np.zeros(s) for _ in range(3)]
    pkg._check(L.bdda_cross_sections_dev(op.h, dB, dX, s, 0, *[o.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                                                               for o in out]))
    u = np.ascontiguousarray(waves.dirs[0::2])
    e1 = np.ascontiguousarray(waves.pols[0::2].real)
    npairs = u.shape[0]
    S11 = np.zeros((npairs, nang))
    S12 = np.zeros((npairs, nang))
    dp = ctypes.POINTER(ctypes.c_double)
    pkg._check(L.bdda_scattering_matrix_dev(op.h, dX, npairs, u.ctypes.data_as(dp), e1.ctypes.data_as(dp), nang,
                                            S11.ctypes.data_as(dp), S12.ctypes.data_as(dp), None))
    return {"Cext": out[0], "Cabs": out[1], "Csca": out[2], "S11": S11, "S12": S12}


def sphere_leg(pkg):
    """The sphere workload (16 dipoles across, m = 1.33, ka = 1, x- and
    y-polarised waves along +z) solved in mixed precision on the GPU, its cross
    sections and S11/S12 (scattering plane xz) against Mie theory for the
    volume-equivalent sphere."""
    mie = importlib.import_module("block_dda_amd.mie")
    W = importlib.import_module("block_dda_amd.workloads").build("sphere16")
    t, k = W.target, W.k
    op = pkg.MBT_fft_init(t.n, 3, t.lf, k)
    op.set_target(t.address, W.diag)
    X, info, _ = op.solve(W.B, 1e-5, 500, method=pkg.COCG_RQ, jacobi=True, mode=pkg.MODE_MIXED)
    xs = op.cross_sections(W.B, X)
    S11, S12, _ = op.scattering_matrix(X, [[0, 0, 1.0]], [[1.0, 0, 0]])
    aeff = (3 * t.nocc * t.lf ** 3 / (4 * math.pi)) ** (1 / 3)
    M = mie.mie(W.cfg["m"], k * aeff)
    M11, M12 = mie.scattering_matrix(M["S1"], M["S2"])
    g = math.pi * aeff ** 2
    out = {"workload": W.describe(), "a_eff": aeff, "x_eff": k * aeff, "iters": info["iters"],
           "true_res": info["true_res"]}
    for q, c in (("Qext", "Cext"), ("Qabs", "Cabs"), ("Qsca", "Csca")):
        out[q] = (xs[c] / g).tolist()
        out[q + "_mie"] = M[q]
        out[q + "_err_rel_to_Qext"] = float(np.abs(xs[c] / g - M[q]).max() / M["Qext"])
    out["S11_max_err_rel_to_max"] = float(np.abs(S11[0] - M11).max() / M11.max())
    out["S12_max_err_rel_to_S11max"] = float(np.abs(S12[0] - M12).max() / M11.max())
    log("sphere vs Mie: %s" % out)
    return out


def configs_leg(pkg):
    """The other stated workloads, measured as one solve each after a warm-up
    solve (device buffers, operator cached): the 1024-monomer aggregate (8
    dipoles across, m = 1.95 + 0.79i, 550 nm, s = 32, block COCG-rq-Jacobi)
    and the 4096-monomer sweep aggregate (6 dipoles across, s = 8, block
    BiCGGR-Jacobi) at 300, 650 and 1000 nm, FP64 and mixed, true residual
    <= 1e-5."""
    WL = importlib.import_module("block_dda_amd.workloads")
    L = pkg.lib()
    out = []
    cases = [("agg1024", WL.build("agg1024"), pkg.COCG_RQ)]
    first = WL.sweep(300.0)
    cases += [("agg4096 sweep %g nm" % w, first if w == 300.0 else WL.sweep(w, target=first.target), pkg.BICGGR)
              for w in (300.0, 650.0, 1000.0)]
    for name, W, method in cases:
        t = W.target
        s = W.B.shape[1]
        rows = 3 * t.nocc
        op = pkg.MBT_fft_init(t.n, 3, t.lf, W.k)
        op.set_target(t.address, W.diag)
        dB, dX = ctypes.c_void_p(), ctypes.c_void_p()
        pkg._check(L.bdda_dev_alloc(rows * s * 16, ctypes.byref(dB)))
        pkg._check(L.bdda_dev_alloc(rows * s * 16, ctypes.byref(dX)))
        Bh = np.asfortranarray(W.B)
        pkg._check(L.bdda_memcpy_h2d(dB, Bh.ctypes.data, rows * s * 16))
        row = {"workload": name, "nocc": t.nocc, "n": list(t.n), "padded": list(op.grid), "s": s,
               "method": METHOD_NAMES[{pkg.COCG_RQ: "cocg_rq", pkg.BICGGR: "bicggr"}[method]]}
        for mname, mode in (("fp64", pkg.MODE_FP64), ("mixed", pkg.MODE_MIXED)):
            solve_dev(pkg, op, dB, dX, s, 1e-5, 3000, mode, method)
            r = solve_dev(pkg, op, dB, dX, s, 1e-5, 3000, mode, method)
            row[mname] = {"solve_s": r["wall_s"], "iters": r["iters"], "matvecs": r["matvecs"],
                          "true_res": r["true_res"], "converged": r["converged"]}
        row["speedup_mixed_vs_fp64"] = row["fp64"]["solve_s"] / row["mixed"]["solve_s"]
        log("config %s" % row)
        out.append(row)
        L.bdda_dev_free(dB)
        L.bdda_dev_free(dX)
        op.close()
    return out


def sweep_leg(pkg, nwl, method, workers=4, sequential=True):
    """The wavelength sweep on the 4096-monomer aggregate (6 dipoles across,
    s = 8, m(lambda) linear from 2.0 + 1.0i at 300 nm to 2.4 + 0.5i at
    1000 nm): nwl wavelengths uniform in 300-1000 nm, each with its own
    operator (G^ at its k) and DIAG_A.  Each wavelength is solved to true
    residual 1e-5 one after another, in mixed precision and in FP64 (device
    buffers, operator build included), and then all wavelengths at once
    through the serving loop, where up to four operators and their solves
    overlap on their own streams."""
    WL = importlib.import_module("block_dda_amd.workloads")
    S = importlib.import_module("block_dda_amd.serving")
    L = pkg.lib()
    wls = np.linspace(300.0, 1000.0, nwl)
    first = WL.sweep(float(wls[0]))
    tgt = first.target
    works = [first] + [WL.sweep(float(w), target=tgt) for w in wls[1:]]
    s = first.B.shape[1]
    rows = 3 * tgt.nocc
    log("sweep: %d wavelengths, %s" % (nwl, first.describe()))
    dB, dX = ctypes.c_void_p(), ctypes.c_void_p()
    pkg._check(L.bdda_dev_alloc(rows * s * 16, ctypes.byref(dB)))
    pkg._check(L.bdda_dev_alloc(rows * s * 16, ctypes.byref(dX)))
    per = {"mixed": [], "fp64": []}
    for W in (works if sequential else []):
        Bh = np.asfortranarray(W.B)
        for name, mode in (("mixed", pkg.MODE_MIXED), ("fp64", pkg.MODE_FP64)):
            pkg._check(L.bdda_device_sync())
            t0 = time.perf_counter()
            op = pkg.MBT_fft_init(tgt.n, 3, tgt.lf, W.k)
            op.set_target(tgt.address, W.diag)
            pkg._check(L.bdda_memcpy_h2d(dB, Bh.ctypes.data, rows * s * 16))
            r = solve_dev(pkg, op, dB, dX, s, 1e-5, 3000, mode, method)
            wall = time.perf_counter() - t0
            op.close()
            per[name].append({"wl": W.cfg["wl"], "s": wall, "iters": r["iters"], "matvecs": r["matvecs"],
                              "true_res": r["true_res"], "converged": r["converged"]})
        log("sweep %.1f nm: mixed %.3f s (%d its), fp64 %.3f s (%d its)" % (
            W.cfg["wl"], per["mixed"][-1]["s"], per["mixed"][-1]["iters"], per["fp64"][-1]["s"],
            per["fp64"][-1]["iters"]))
    L.bdda_dev_free(dB)
    L.bdda_dev_free(dX)
    out = {"wavelengths": nwl, "workload": first.describe(), "method": method}
    for name in (per if sequential else []):
        v = per[name]
        it = [x["iters"] for x in v]
        out[name] = {"total_s": float(sum(x["s"] for x in v)), "mean_s": float(np.mean([x["s"] for x in v])),
                     "iters_min": int(min(it)), "iters_max": int(max(it)),
                     "true_res_max": float(max(x["true_res"] for x in v)),
                     "all_converged": bool(all(x["converged"] for x in v))}
    if sequential:
        out["speedup_mixed_vs_fp64"] = out["fp64"]["total_s"] / out["mixed"]["total_s"]
    # all wavelengths at once through the serving loop (mixed precision)
    sv = S.Server(S.GpuBackend(pkg, method=method), hbm_budget=200 << 30, workers=workers, max_cols=64).start()
    t0 = time.perf_counter()
    for W in works:
        sv.submit(S.Request(S.Scene(tgt, W.k, W.diag, name="sweep%.0f" % W.cfg["wl"]), W.B, tol=1e-5))
    sv.drain()
    span = time.perf_counter() - t0
    st = sv.stats()
    sv.stop()
    out["served"] = {"workers": workers, "makespan_s": span, "p50_s": st.get("p50_s"), "p99_s": st.get("p99_s"),
                     "failed": st.get("failed"), "max_concurrent_batches": st.get("max_concurrent_batches")}
    log("sweep served: %s" % out["served"])
    return out


SERVING_BUDGET = 240 << 30      # of the MI355X's 288 GB HBM
SERVING_WS_LIMIT = 32 << 30     # block matvecs beyond this workspace run in column chunks


def serving_leg(pkg, nreq, rate, nscenes=24, cap=SERVING_BUDGET, workers=4):
    """Poisson stream of solve requests through the serving loop (mixed
    precision, block COCG-rq-Jacobi, tol 1e-5): p50/p99 latency, solves/s.
    The scenes follow the stream workload's distribution (N logUniform
    64..8192 monomers, Df 1.8-2.8, kf 1.0-1.3, 4-8 dipoles across, lambda
    300-1000 nm); a draw is kept when a 64-column batch of it fits the HBM
    per-scene cap `cap` by bdda_estimate_bytes (operator build transient, or
    operator + workspace under the chunking cap + Krylov blocks) and its
    padded grid is within the passes' limit; the share of draws this rejects
    is reported.  The server's own budget is the GPU's SERVING_BUDGET."""
    S = importlib.import_module("block_dda_amd.serving")
    T = importlib.import_module("block_dda_amd.targets")
    be = S.GpuBackend(pkg, ws_limit=SERVING_WS_LIMIT)
    t0 = time.time()
    draw = {}
    scenes = S.synthetic_scenes(T, nscenes, seed=2026, nmin=64, nmax=8192, across=(4, 8),
                                fits=lambda t: (be.fresh_need(t, 3, 64) or float("inf")) <= cap,
                                stats=draw)
    reqs = S.synthetic_requests(T, scenes, nreq, rate=rate, seed=2027, s_choices=(2, 8, 32, 64), tol=1e-5)
    log("serving: %d scenes (nocc %s, lattices %s) built in %.1fs, %s" % (
        len(scenes), [sc.meta["nocc"] for sc in scenes], [sc.meta["n"] for sc in scenes], time.time() - t0, draw))
    done = [0]

    def on_batch(key, group, secs):
        done[0] += len(group)
        log("served %d/%d: batch of %d request(s), %d columns, %d dipoles, %.2f s" % (
            done[0], nreq, len(group), sum(r.cols for r in group), group[0].scene.target.nocc, secs))

    sv = S.Server(be, hbm_budget=SERVING_BUDGET, workers=workers, max_cols=64, on_batch=on_batch).start()
    t1 = time.time()
    S.replay(sv, reqs)
    sv.drain()
    st = sv.stats()
    sv.stop()
    st["stream_s"] = time.time() - t1
    st["rate_req_per_s"] = rate
    st["scene_cap"] = ("64-column batch within %d GiB by bdda_estimate_bytes (block matvecs chunked beyond %d GiB "
                       "of workspace) and padded grid within the passes' limit; server budget %d GiB" % (
                           cap >> 30, SERVING_WS_LIMIT >> 30, SERVING_BUDGET >> 30))
    st["scene_draws"] = draw
    st["scenes"] = [sc.meta for sc in scenes]
    st["s_choices"] = [2, 8, 32, 64]
    log("serving: %s" % {k: v for k, v in st.items() if k != "scenes"})
    return st


def reference_leg(pkg, sizes, wl0=0.3):
    """The reference's own published workloads: the BC aggregates
    DDA_target_0_agg_N{size} at wl0, f = 6, four seeded incident waves,
    Mie-a_1/b_1 polarizabilities, block BiCGSTAB-Jacobi to the reference's
    tol 1e-2 (Solver.cpp:69-73, Solver.hpp:27-28), built from the fixture
    copies of the reference's target files and material table.  Reports the
    GPU solve time in FP64 (the reference's arithmetic and stopping rule)
    and in mixed precision (true FP64 residual <= 1e-2), the reference's
    published solve time on its CPU (context only: other hardware, whole
    seconds) and the relative errors of the FP64 result against the
    published mean Csca, MAC and SSA."""
    R = importlib.import_module("block_dda_amd.results")
    M = importlib.import_module("block_dda_amd.materials")
    gold = os.path.join(ROOT, "tests", "golden")
    data = os.path.join(gold, "reference_data")
    secs = json.load(open(os.path.join(gold, "reference_results.json")))["sections"]
    mats = [M.Material.named("BC", data), M.Material.constant(2.25, 1.0, 1.0)]
    out = []
    for N in sizes:
        fname = "DDA_target_0_agg_N%d_kf1.0_Df2.8.out" % N
        s = [x for x in secs if x["file"] == "DDA_result_BC_rpp20nm.dat" and x["target_fname"] == fname
             and abs(x["wl0"] - wl0) < 1e-12][0]
        t0 = time.perf_counter()
        inp = R.SectionInputs(os.path.join(data, fname), mats, 20e-3, M.Material.vacuum(), wl0, s["N_inc"], 6)
        t_inputs = time.perf_counter() - t0
        t = inp.target
        t0 = time.perf_counter()
        op = pkg.MBT_fft_init(t.n, 6, t.lf, inp.k)
        op.set_target(t.address, inp.diag)
        pkg._check(pkg.lib().bdda_device_sync())
        t_init = time.perf_counter() - t0
        row = {"target": fname, "wl0": wl0, "n": list(t.n), "nocc": t.nocc, "f": 6, "L": s["N_inc"],
               "padded": list(op.grid), "inputs_s": t_inputs, "init_s": t_init,
               "published_solver_time_s": s["DDA_solver_time"]}
        for name, mode in (("fp64", pkg.MODE_FP64), ("mixed", pkg.MODE_MIXED)):
            op.solve(inp.B, 1e-2, 300, method=pkg.BICGSTAB, jacobi=True, mode=mode)     # warm-up
            pkg._check(pkg.lib().bdda_device_sync())
            t0 = time.perf_counter()
            X, info, _ = op.solve(inp.B, 1e-2, 300, method=pkg.BICGSTAB, jacobi=True, mode=mode)
            wall = time.perf_counter() - t0
            xs = op.cross_sections(inp.B, X, alpha_mode=pkg.ALPHA_FIRST)
            r = inp.result(xs["Cext"], xs["Cabs"])
            row[name] = {"solve_s": wall, "iters": info["iters"], "matvecs": info["matvecs"],
                         "true_res": info["true_res"], "converged": info["converged"],
                         "err_vs_published": {q: abs(r[q] - s[q]) / abs(s[q])
                                              for q in ("mean_Csca", "mean_MAC", "mean_SSA", "MAC_mie")}}
        op.close()
        log("reference section: %s" % row)
        out.append(row)
    return out


def compare_observables(mx, ref, geom_cs):
    """Errors of the mixed-precision observables against the FP64 solve's."""
    res = {}
    for q in ("Cext", "Cabs", "Csca"):
        res["Q" + q[1:] + "_fp64_mean"] = float(np.mean(ref[q]) / geom_cs)
        res["Q" + q[1:] + "_mixed_mean"] = float(np.mean(mx[q]) / geom_cs)
        res[q + "_max_rel_err"] = float(np.max(np.abs(mx[q] - ref[q]) / np.abs(ref[q])))
    scale = np.max(ref["S11"])
    res["S11_max_err_rel_to_max"] = float(np.max(np.abs(mx["S11"] - ref["S11"])) / scale)
    res["S12_max_err_rel_to_S11max"] = float(np.max(np.abs(mx["S12"] - ref["S12"])) / scale)
    res["angles"] = int(ref["S11"].shape[1])
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=3, help="timed FP64 solves (at least 3, median reported)")
    ap.add_argument("--config", default="agg16384", choices=["agg16384", "agg1024", "agg4096", "agg256", "sphere16"])
    ap.add_argument("--method", default=None, choices=["cocg_rq", "bicgstab", "bicggr"],
                    help="block solver (Jacobi variant); the wavelength-sweep workload agg4096 is solved "
                         "with bicggr in its parity tests")
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--itermax", type=int, default=3000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-sphere", action="store_true", help="skip the sphere-against-Mie leg")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other workloads (1024-monomer s = 32, sweep aggregate at three wavelengths)")
    ap.add_argument("--sweep-workers", type=int, default=4, help="serving-loop workers for the sweep")
    ap.add_argument("--sweep-served-only", action="store_true", help="skip the sweep's one-by-one solves")
    ap.add_argument("--sweep", type=int, default=0,
                    help="run the wavelength sweep over this many wavelengths (300-1000 nm) instead of the solve "
                         "benchmark, with --method (default bicggr for the sweep), and print its summary")
    ap.add_argument("--matvec-reps", type=int, default=5)
    ap.add_argument("--matvec-only", action="store_true",
                    help="time the block matvecs and exit (for counter runs)")
    ap.add_argument("--serving-requests", type=int, default=200,
                    help="requests in the serving stream leg (0 skips it)")
    ap.add_argument("--serving-workers", type=int, default=4,
                    help="serving-loop workers (concurrent batches on the GPU)")
    ap.add_argument("--serving-rate", type=float, default=4.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serving-cap-gib", type=int, default=48,
                    help="per-scene cap of the stream's scene draws (GiB a 64-column batch may need); 240 admits "
                         "everything one GPU holds, but at 4 req/s that mix is more work than one GPU serves")
    ap.add_argument("--reference-sizes", default="16384",
                    help="comma-separated monomer counts of the reference's published BC aggregates to solve "
                         "(f = 6, BiCGSTAB-Jacobi, tol 1e-2); empty skips the leg")
    ap.add_argument("--reference-only", action="store_true", help="run only the published-workload leg")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="also run one device-to-device copy of this many bytes: the known-size "
                         "reference for HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the solve path is measured on a single MI355X "
              "(the path would shard by right-hand-side columns with an s x s all-reduce per iteration; "
              "that is not run here).", file=sys.stderr)
        sys.exit(2)
    pkg = load_pkg()
    if args.method is None:
        args.method = "bicggr" if args.sweep > 0 else "cocg_rq"
    method = {"cocg_rq": pkg.COCG_RQ, "bicgstab": pkg.BICGSTAB, "bicggr": pkg.BICGGR}[args.method]
    if args.sweep > 0:
        out = sweep_leg(pkg, args.sweep, method, args.sweep_workers, not args.sweep_served_only)
        out["method"] = METHOD_NAMES[args.method]
        print(json.dumps(out), flush=True)
        return
    ref_sizes = [int(v) for v in args.reference_sizes.split(",") if v.strip()]
    if args.reference_only:
        print(json.dumps({"reference_published": reference_leg(pkg, ref_sizes)}), flush=True)
        return
    W = build_workload(args.config)
    tgt, k, diag, B = W.target, W.k, W.diag, W.B
    s = B.shape[1]
    # the process's first operator also pays the HIP runtime's start-up (code
    # objects, first allocations, about 0.15 s); a small operator takes it
    # first, so that init_s is the build of this workload's G^ and target
    warm = pkg.MBT_fft_init((8, 8, 8), 3, 1.0, 1.0)
    warm.close()
    pkg._check(pkg.lib().bdda_device_sync())
    t0 = time.time()
    op = pkg.MBT_fft_init(tgt.n, 3, tgt.lf, k)
    op.set_target(tgt.address, diag)
    pkg._check(pkg.lib().bdda_device_sync())
    init_s = time.time() - t0
    P = op.grid
    log("operator: padded grid %s, init %.2fs" % (P, init_s))
    rows = 3 * tgt.nocc
    L = pkg.lib()
    dB = ctypes.c_void_p()
    dX = ctypes.c_void_p()
    pkg._check(L.bdda_dev_alloc(rows * s * 16, ctypes.byref(dB)))
    pkg._check(L.bdda_dev_alloc(rows * s * 16, ctypes.byref(dX)))
    Bh = np.asfortranarray(B)
    pkg._check(L.bdda_memcpy_h2d(dB, Bh.ctypes.data, rows * s * 16))

    # per-pass kernel times of the block matvec (HIP events, library stream)
    mv = {}
    for name, prec, c in (("fp32", pkg.BDDA_PREC_FP32, 8), ("fp64", pkg.BDDA_PREC_FP64, 16)):
        ms, wall = time_matvec(pkg, op, rows, s, prec, args.matvec_reps)
        by = pass_bytes(tgt.n, P, tgt.nocc, s, 3, c, 6, occupied_lines(tgt.n, tgt.address))
        mv[name] = {"pass_ms": ms, "wall_ms": wall,
                    "pass_GBs": {kk: by[kk] / (ms[kk] * 1e-3) / 1e9 for kk in ms},
                    "matvec_GBs": sum(by.values()) / (sum(ms.values()) * 1e-3) / 1e9,
                    "bytes": by}
        log("matvec %s: %s  total %.2f ms (%.0f GB/s algorithmic)" % (
            name, {kk: round(v, 3) for kk, v in ms.items()}, sum(ms.values()), mv[name]["matvec_GBs"]))

    # what a plain streaming kernel reaches on this GPU (2 x 2 GiB buffers)
    probe = None
    pa, pb = ctypes.c_void_p(), ctypes.c_void_p()
    pbytes = 2 << 30
    if L.bdda_dev_alloc(pbytes, ctypes.byref(pa)) == 0:
        if L.bdda_dev_alloc(pbytes, ctypes.byref(pb)) == 0:
            cg, rg, wg = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
            pkg._check(L.bdda_hbm_probe(pa, pb, pbytes, ctypes.byref(cg), ctypes.byref(rg), ctypes.byref(wg)))
            probe = {"copy_GBs": cg.value, "read_GBs": rg.value, "write_GBs": wg.value, "buffer_bytes": pbytes,
                     "kernel": "bdda_hbm_probe: 16-byte vector streaming copy/read/write, best of 5"}
            L.bdda_dev_free(pb)
        L.bdda_dev_free(pa)
    log("hbm probe: %s" % probe)
    if args.copy_bytes > 0:
        cA, cB = ctypes.c_void_p(), ctypes.c_void_p()
        pkg._check(L.bdda_dev_alloc(args.copy_bytes, ctypes.byref(cA)))
        pkg._check(L.bdda_dev_alloc(args.copy_bytes, ctypes.byref(cB)))
        pkg._check(L.bdda_memcpy_d2d(cB, cA, args.copy_bytes))
        pkg._check(L.bdda_device_sync())
        L.bdda_dev_free(cA)
        L.bdda_dev_free(cB)
        log("copy: %d bytes device to device" % args.copy_bytes)
    if args.matvec_only:
        print(json.dumps({"matvec": {kk: {"pass_ms": v["pass_ms"], "matvec_GBs": v["matvec_GBs"],
                                          "bytes": v["bytes"]} for kk, v in mv.items()},
                          "copy_bytes": args.copy_bytes}), flush=True)
        return
    # FP64 reference-semantics solve (recurrence residual < tol)
    fp64 = None
    obs64 = None
    if not args.no_fp64:
        # the same warm-up as the mixed solve, then at least three timed
        # solves; the comparison uses the medians of both
        for _ in range(args.warmup):
            log("fp64 warmup: %s" % solve_dev(pkg, op, dB, dX, s, args.tol, args.itermax, pkg.MODE_FP64, method))
        runs64 = []
        for _ in range(max(3, args.fp64_steps)):
            runs64.append(solve_dev(pkg, op, dB, dX, s, args.tol, args.itermax, pkg.MODE_FP64, method))
            log("fp64 step: %s" % runs64[-1])
        fp64 = dict(runs64[-1])
        fp64["wall_s_runs"] = [r["wall_s"] for r in runs64]
        fp64["wall_s_median"] = float(np.median(fp64["wall_s_runs"]))
        fp64["true_res_col_max"] = residual_col_max(pkg, op, dB, dX, s)
        obs64 = observables_dev(pkg, op, dB, dX, W.waves, s)
    # mixed precision (true FP64 residual <= tol): warmup + timed steps
    for _ in range(args.warmup):
        r = solve_dev(pkg, op, dB, dX, s, args.tol, args.itermax, pkg.MODE_MIXED, method)
        log("mixed warmup: %s" % r)
    runs = []
    pkg._check(L.bdda_device_sync())
    t_all = time.perf_counter()
    for _ in range(args.steps):
        runs.append(solve_dev(pkg, op, dB, dX, s, args.tol, args.itermax, pkg.MODE_MIXED, method))
        log("mixed step: %s" % runs[-1])
    pkg._check(L.bdda_device_sync())
    total = time.perf_counter() - t_all
    runs[-1]["true_res_col_max"] = residual_col_max(pkg, op, dB, dX, s)
    # the same solve stopping on every column's true residual (no reference
    # counterpart: the reference's rule is the block norm)
    cols = solve_dev(pkg, op, dB, dX, s, args.tol, args.itermax, pkg.MODE_MIXED_COLUMNS, method)
    cols["true_res_col_max"] = residual_col_max(pkg, op, dB, dX, s)
    log("mixed, per-column stop: %s" % cols)
    # X of the block-rule solve again for the observables below
    solve_dev(pkg, op, dB, dX, s, args.tol, args.itermax, pkg.MODE_MIXED, method)
    obs = None
    if obs64 is not None:
        aeff = (3.0 * tgt.nocc * tgt.lf ** 3 / (4.0 * math.pi)) ** (1.0 / 3.0)
        obs = compare_observables(observables_dev(pkg, op, dB, dX, W.waves, s), obs64, math.pi * aeff ** 2)
        obs["a_eff"] = aeff
        log("observables (mixed vs fp64): %s" % obs)
    ms_per_step = total / max(args.steps, 1) * 1e3
    tts = ms_per_step / 1e3
    mixed_median = float(np.median([r["wall_s"] for r in runs]))

    xf = mv["fp32"]
    dom_bytes = xf["bytes"]["x_fused"]
    dom_ms = xf["pass_ms"]["x_fused"]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    counters, counters_note = profile_counters(args.config)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": float(counters["x_traffic_bytes"]) if counters else None,
                "traffic_source": counters_note,
                "frac_of_streaming_copy": (achieved / probe["copy_GBs"]) if probe else None,
                "kernel": "k_xfused_plan<float> at Px=%d (x FFT fwd + G^ multiply + x FFT inv)" % P[0],
                "bytes_per_launch": dom_bytes, "avg_launch_ms": dom_ms}
    if counters:
        # 4 cycles per wave64 VALU instruction on each of the 1024 SIMDs
        valu = float(counters["x_valu_wave_insts"])
        valu_ms = valu * 4 / (1024 * 2.1e9) * 1e3
        roofline["limiter"] = ("VALU issue (packed FP32 disabled): %.3g wave instructions per launch take "
                               "%.2f ms at 2.1 GHz, %.0f %% of the launch" % (valu, valu_ms, 100 * valu_ms / dom_ms))
    # the main workload's operator (with its pass workspace) and blocks are
    # released before the other legs, which size their memory against the HBM
    L.bdda_dev_free(dB)
    L.bdda_dev_free(dX)
    op.close()
    serving = (serving_leg(pkg, args.serving_requests, args.serving_rate, cap=args.serving_cap_gib << 30,
                           workers=args.serving_workers)
               if args.serving_requests > 0 else None)
    sphere = None if args.no_sphere else sphere_leg(pkg)
    other = None if args.no_configs else configs_leg(pkg)
    refpub = reference_leg(pkg, ref_sizes) if ref_sizes else None
    cpu = None
    if not args.no_cpu:
        try:
            nmv = fp64["matvecs"] if fp64 else runs[-1]["matvecs"]
            # the GPU box's CPU share is 16 cores
            cpu = cpu_baseline(tgt, s, nmv, threads=min(16, os.cpu_count() or 1))
            log("cpu baseline: %s" % cpu)
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline failed: %r" % e)
    out = {
        "metric": METRIC, "value": tts, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "complex64 inner + complex128 outer", "data": "synthetic (seeded aggregate, plane waves)",
        "config": {"workload": "%s -> padded %s, block %s (Jacobi), tol %g" % (
            W.describe(), list(P), METHOD_NAMES[args.method], args.tol),
                   "parallelism": "single GPU"},
        "time_to_solution_with_init_s": tts + init_s,
        "mixed": {"iters": runs[-1]["iters"], "outer": runs[-1]["outer_iters"], "matvecs": runs[-1]["matvecs"],
                  "true_res": runs[-1]["true_res"], "true_res_col_max": runs[-1]["true_res_col_max"],
                  "converged": runs[-1]["converged"], "wall_s_median": mixed_median,
                  "wall_s_runs": [r["wall_s"] for r in runs]},
        "mixed_per_column_stop": {k2: cols[k2] for k2 in ("iters", "outer_iters", "matvecs", "true_res",
                                                           "true_res_col_max", "converged", "wall_s")},
        "fp64_time_s": fp64["wall_s_median"] if fp64 else None,
        "fp64": {k2: fp64[k2] for k2 in ("iters", "matvecs", "true_res", "true_res_col_max", "recurrence_res",
                                          "converged", "wall_s_median", "wall_s_runs")}
        if fp64 else None,
        # medians of both, each after the same warm-up
        "speedup_mixed_vs_fp64": (fp64["wall_s_median"] / mixed_median) if fp64 else None,
        "observables": obs,
        "serving": serving,
        "init_s": init_s,
        "hbm_probe": probe,
        "matvec": {kk: {"pass_ms": v["pass_ms"], "matvec_GBs": v["matvec_GBs"],
                        "pass_GBs": v["pass_GBs"]} for kk, v in mv.items()},
        "roofline": roofline,
        "sphere_vs_mie": sphere,
        "other_configs": other,
        "reference_published": refpub,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
