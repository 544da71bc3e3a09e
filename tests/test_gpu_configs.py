"""The measured configurations beyond the 16384-monomer one, as parity cases:

  * the 1024-monomer aggregate (8 dipoles across, m = 1.95 + 0.79i, 550 nm,
    s = 32): block COCG-rq-Jacobi in FP64 with the reference's stopping rule
    and in mixed precision; both reach a true residual of 1e-5 and their cross
    sections agree within 1e-4;
  * the wavelength sweep on the 4096-monomer aggregate (6 dipoles across, m
    linear in wavelength, s = 8) at its two ends and middle, solved with
    mixed-precision block BiCGGR-Jacobi against FP64 BiCGGR-Jacobi, and at
    all 64 of its wavelengths in mixed precision.
"""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel_max(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _check_pair(bdda, W, method, tol=1e-5):
    t = W.target
    op = bdda.MBT_fft_init(t.n, 3, t.lf, W.k)
    op.set_target(t.address, W.diag)
    X64, i64, _ = op.solve(W.B, tol, 3000, method=method, jacobi=True, mode=bdda.MODE_FP64)
    Xm, im, _ = op.solve(W.B, tol, 3000, method=method, jacobi=True, mode=bdda.MODE_MIXED)
    print(W.name, W.cfg["wl"], "fp64", i64["iters"], i64["true_res"], "mixed", im["iters"], im["true_res"],
          flush=True)
    assert i64["converged"] and im["converged"] and im["true_res"] <= tol
    assert i64["true_res"] <= 3 * tol        # the recurrence residual tracks the true one
    x64 = op.cross_sections(W.B, X64)
    xm = op.cross_sections(W.B, Xm)
    for q in ("Cext", "Cabs", "Csca"):
        assert _rel_max(xm[q], x64[q]) <= 1e-4, q
    op.close()


def test_agg1024_cocg_rq(bdda):
    W = importlib.import_module("block_dda_amd.workloads").build("agg1024")
    print("case", W.describe(), flush=True)
    _check_pair(bdda, W, bdda.COCG_RQ)


def test_wavelength_sweep_bicggr(bdda):
    wl = importlib.import_module("block_dda_amd.workloads")
    first = wl.sweep(300.0)
    print("case", first.describe(), flush=True)
    for lam in (300.0, 650.0, 1000.0):
        W = first if lam == 300.0 else wl.sweep(lam, target=first.target)
        _check_pair(bdda, W, bdda.BICGGR)


def test_wavelength_sweep_all_64(bdda):
    """All 64 wavelengths of the sweep (uniform in 300-1000 nm, m(lambda)
    linear from 2.0 + 1.0i to 2.4 + 0.5i, s = 8), each with its own operator
    and DIAG_A, solved by mixed-precision block BiCGGR-Jacobi to a true
    residual of 1e-5; the cross sections are those of an absorbing particle
    (Cext > Cabs > 0) and vary smoothly with the wavelength."""
    wl = importlib.import_module("block_dda_amd.workloads")
    first = wl.sweep(300.0)
    t = first.target
    print("case sweep 64", first.describe(), flush=True)
    qext = []
    for lam in np.linspace(300.0, 1000.0, 64):
        W = first if lam == 300.0 else wl.sweep(lam, target=t)
        op = bdda.MBT_fft_init(t.n, 3, t.lf, W.k)
        op.set_target(t.address, W.diag)
        X, info, _ = op.solve(W.B, 1e-5, 3000, method=bdda.BICGGR, jacobi=True, mode=bdda.MODE_MIXED)
        assert info["converged"] and info["true_res"] <= 1e-5, (lam, info)
        xs = op.cross_sections(W.B, X)
        assert np.all(xs["Cabs"] > 0) and np.all(xs["Cext"] > xs["Cabs"]), lam
        qext.append(float(np.mean(xs["Cext"])))
        op.close()
    q = np.array(qext)
    # neighbouring wavelengths (11 nm apart) differ by a few per cent at most
    assert np.max(np.abs(np.diff(q)) / q[1:]) < 0.1
