"""The 16384-monomer aggregate at full size (grid 170x171x171 padded to 343^3,
549,239 dipoles, s = 16), where the oracle would take hours: size-independent
properties of the block matvec, and the mixed-precision solve checked by its
own FP64 residual and against the FP64 solve."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def big(bdda):
    W = importlib.import_module("block_dda_amd.workloads").build("agg16384")
    t = W.target
    print("case agg16384 grid", t.n, "nocc", t.nocc, flush=True)
    op = bdda.MBT_fft_init(t.n, 3, t.lf, W.k)
    op.set_target(t.address, W.diag)
    yield W, op
    op.close()


def test_fullsize_matvec_properties(big):
    W, op = big
    rows = 3 * W.target.nocc
    rng = np.random.default_rng(20261203)
    X = rng.uniform(-1, 1, (rows, 3)) + 1j * rng.uniform(-1, 1, (rows, 3))
    Y = rng.uniform(-1, 1, (rows, 3)) + 1j * rng.uniform(-1, 1, (rows, 3))
    a, b = 0.7 - 0.2j, -1.3 + 0.4j
    AXY = op.apply_occupied(np.concatenate([X, Y, a * X + b * Y], axis=1))
    AX, AY, AZ = AXY[:, :3], AXY[:, 3:6], AXY[:, 6:]
    # linearity
    assert rel(AZ, a * AX + b * AY) <= 1e-12
    # A = DIAG + G is complex symmetric for f = 3: X^T (A Y) = (Y^T (A X))^T
    G1 = X.T @ AY
    G2 = (Y.T @ AX).T
    assert np.abs(G1 - G2).max() <= 1e-11 * np.abs(G1).max()
    # the Jacobi variant is jpre = 1/DIAG applied to the same product
    AXj = op.apply_occupied(X, jacobi=True)
    assert rel(AXj, AX / W.diag[:, None]) <= 1e-13
    # FP32 product within the FP32 tolerance of the FP64 one
    AX32 = op.apply_occupied(X, precision="fp32")
    assert rel(AX32, AX) <= 1e-5


def test_fullsize_mixed_solve_matches_fp64(bdda, big):
    W, op = big
    B = W.B
    tol = 1e-5
    Xm, im, _ = op.solve(B, tol, 3000, method=bdda.COCG_RQ, jacobi=True, mode=bdda.MODE_MIXED)
    print("mixed", im, flush=True)
    assert im["converged"] == 1 and im["true_res"] <= tol
    # the residual, recomputed here through the FP64 matvec, per column too
    R = B - op.apply_occupied(Xm)
    assert np.linalg.norm(R) / np.linalg.norm(B) <= tol * 1.001
    col = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
    assert col.max() <= 5 * tol
    X64, i64, _ = op.solve(B, tol, 3000, method=bdda.COCG_RQ, jacobi=True, mode=bdda.MODE_FP64)
    print("fp64", i64, flush=True)
    assert i64["converged"] == 1
    # two solutions at residual 1e-5 differ by at most cond(A) x 1e-5
    assert rel(Xm, X64) <= 1e-3
    # cross sections of the mixed solve within 1e-4 of the FP64 solve's
    xm = op.cross_sections(B, Xm)
    x64 = op.cross_sections(B, X64)
    print("Cext mixed", xm["Cext"][:4], "fp64", x64["Cext"][:4], flush=True)
    for q in ("Cext", "Cabs", "Csca"):
        assert np.abs(xm[q] - x64[q]).max() <= 1e-4 * np.abs(x64[q]).max(), q
    w = W.waves
    u, e1 = w.dirs[0::2], w.pols[0::2].real
    S11m, S12m, _ = op.scattering_matrix(Xm, u, e1)
    S11d, S12d, _ = op.scattering_matrix(X64, u, e1)
    assert np.abs(S11m - S11d).max() <= 1e-4 * S11d.max()
    assert np.abs(S12m - S12d).max() <= 1e-4 * S11d.max()


def test_fullsize_bicgstab_mixed_matches_fp64(bdda, big):
    """The workload's second solver, block BiCGSTAB with Jacobi
    (bl_bicgstab_jacobi_mvp_fft.cpp:12-97), in mixed precision against the same
    solver in FP64: true residuals at tol, cross sections within 1e-4."""
    W, op = big
    B = W.B
    tol = 1e-5
    Xm, im, _ = op.solve(B, tol, 3000, method=bdda.BICGSTAB, jacobi=True, mode=bdda.MODE_MIXED)
    print("bicgstab mixed", im, flush=True)
    assert im["converged"] == 1 and im["true_res"] <= tol
    X64, i64, _ = op.solve(B, tol, 3000, method=bdda.BICGSTAB, jacobi=True, mode=bdda.MODE_FP64)
    print("bicgstab fp64", i64, flush=True)
    assert i64["converged"] == 1
    R = B - op.apply_occupied(X64)
    assert np.linalg.norm(R) / np.linalg.norm(B) <= 1.001 * tol
    assert rel(Xm, X64) <= 1e-3
    xm = op.cross_sections(B, Xm)
    x64 = op.cross_sections(B, X64)
    for q in ("Cext", "Cabs", "Csca"):
        assert np.abs(xm[q] - x64[q]).max() <= 1e-4 * np.abs(x64[q]).max(), q
