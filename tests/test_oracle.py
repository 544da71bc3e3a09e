"""CPU tests pinning the oracle (oracle/oracle.c) before it is trusted.

Each check is chosen so that a plausible mistake in the restatement fails:
  - FFT: against numpy.fft on lengths with 2/3/5/7, mid primes and
    Bluestein primes;
  - BT_fft ordering: the reference's own f=1 label map
    (module_tests/mvp_fft_tests/mvp_fft_example.cpp:37-101) must give
    Au = [1..27] for n=(2,2,2);
  - BT_pad / BT_reconstruct: the literal recursions against the closed-form
    index maps, exhaustively on small grids with f in {1, 3, 6};
  - application_function: against blocks derived by calculus
    (tests/golden/green_kat.json), plus symmetry / parity / static-limit
    invariants;
  - MBT_fft_mvp: against the dense block-Toeplitz product (numpy brute force);
  - solvers: against numpy.linalg.solve, and for L = 1 against textbook
    single-RHS COCG / BiCGSTAB / BiCGGR written out here;
  - qr_reduced: orthonormality, triangularity, positive real diagonal,
    reconstruction, and LAPACK QR after phase normalisation.
"""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5, 7, 8, 12, 31, 49, 61, 62, 67, 127, 331, 993, 2 * 3 * 5 * 7 * 11,
                               9 * 31 * 31 * 31])
def test_fft_against_numpy(orc, L):
    rng = np.random.default_rng(L)
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    ref = np.fft.fft(x)
    assert np.abs(orc.fft(x, -1) - ref).max() <= 1e-13 * np.abs(ref).max()
    refi = np.fft.ifft(x) * L
    assert np.abs(orc.fft(x, +1) - refi).max() <= 1e-13 * np.abs(refi).max()


@pytest.mark.parametrize("L", [96, 3072, 3729, 30030, 38533, 127, 1 << 16])
def test_fourstep_fft_against_numpy(orc, L):
    """The four-step path (forced at small lengths here) against numpy,
    with one thread and with four; primes fall back to the direct plan."""
    rng = np.random.default_rng(L)
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    ref, refi = np.fft.fft(x), np.fft.ifft(x) * L
    try:
        for threads in (1, 4):
            orc.fft_config(64, threads)
            assert np.abs(orc.fft(x, -1) - ref).max() <= 1e-13 * np.abs(ref).max()
            assert np.abs(orc.fft(x, +1) - refi).max() <= 1e-13 * np.abs(refi).max()
    finally:
        orc.fft_config()


def test_fourstep_matvec_against_dense(orc):
    """The MBT matvec through four-step transforms equals the dense product."""
    n, f, lf, k = (4, 3, 5), 3, 0.7, 1.1
    rng = np.random.default_rng(4)
    x = rng.standard_normal(f * 60) + 1j * rng.standard_normal(f * 60)
    A = orc.dense_offdiag(n, f, lf, k)
    Au = orc.mbt_fft_init(n, f, lf, k)
    try:
        orc.fft_config(64, 3)
        y = orc.mbt_fft_mvp(n, f, Au, x)
    finally:
        orc.fft_config()
    ref = A @ x
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


def test_bt_fft_label_kat(orc):
    kat = json.load(open(os.path.join(GOLD, "kat_label_map.json")))
    table = {(tuple(a), tuple(b)): lab for a, b, lab in kat["labels"]}

    def elem(n_ind):
        return np.array([[table[(tuple(n_ind[0]), tuple(n_ind[1]))]]], dtype=complex)

    Au = orc.bt_fft_custom(kat["n"], 1, elem)
    assert np.array_equal(Au.real, np.arange(1, 28)) and not Au.imag.any()
    # MBT_fft_init reverses it before the FFT (MBT_fft_init.cpp:34)
    assert np.array_equal(Au[::-1].real, np.arange(27, 0, -1))


def _pad_closed(n, f, x):
    nx, ny, nz = n
    My, Mz = 2 * ny - 1, 2 * nz - 1
    xz = np.zeros(f * f * nx * My * Mz, dtype=complex)
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                for q in range(f):
                    xz[((ix * My + iy) * Mz + iz) * f * f + q] = x[f * ((ix * ny + iy) * nz + iz) + q]
    return xz


def _rec_closed(n, f, bz):
    nx, ny, nz = n
    My, Mz = 2 * ny - 1, 2 * nz - 1
    b = np.zeros(f * nx * ny * nz, dtype=complex)
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                for r in range(f):
                    b[f * ((ix * ny + iy) * nz + iz) + r] = bz[
                        (((ix + nx - 1) * My + iy + ny - 1) * Mz + iz + nz - 1) * f * f + r * f + f - 1]
    return b


@pytest.mark.parametrize("n", [(1, 1, 1), (2, 3, 4), (3, 2, 2), (1, 4, 2), (4, 1, 3), (2, 2, 5)])
@pytest.mark.parametrize("f", [1, 3, 6])
def test_bt_pad_reconstruct_closed_form(orc, n, f):
    rng = np.random.default_rng(7)
    C = int(np.prod(n))
    x = rng.standard_normal(f * C) + 1j * rng.standard_normal(f * C)
    assert np.array_equal(orc.bt_pad(n, f, x), _pad_closed(n, f, x))
    F = orc.fft_length(n, f)
    bz = rng.standard_normal(F) + 1j * rng.standard_normal(F)
    assert np.array_equal(orc.bt_reconstruct(n, f, bz), _rec_closed(n, f, bz))


def test_application_function_calculus_kat(orc):
    kat = json.load(open(os.path.join(GOLD, "green_kat.json")))
    for c in kat["cases"]:
        k = complex(*c["k"])
        for f in (3, 6):
            ref = np.array([[complex(*v) for v in row] for row in c["f%d" % f]])
            got = orc.application_function(list(c["d"]) + [0, 0, 0], f, c["lf"], k)
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (c["d"], f)


def test_application_function_invariants(orc):
    rng = np.random.default_rng(3)
    for _ in range(20):
        d = rng.integers(-4, 5, size=3)
        if not d.any():
            continue
        lf, k = 0.8, 1.1 + 0.05j
        ind = list(d) + [0, 0, 0]
        ind_m = list(-d) + [0, 0, 0]
        M = orc.application_function(ind, 6, lf, k)
        Mm = orc.application_function(ind_m, 6, lf, k)
        G, Fm = M[:3, :3], M[:3, 3:]
        assert np.allclose(G, G.T, rtol=0, atol=1e-14 * np.abs(G).max())          # symmetric
        assert np.allclose(Fm, -Fm.T, rtol=0, atol=1e-14 * np.abs(Fm).max())       # antisymmetric
        assert np.allclose(Mm[:3, :3], G, rtol=1e-13)                              # even in d
        assert np.allclose(Mm[:3, 3:], -Fm, rtol=1e-13)                            # odd in d
        assert np.allclose(M[3:, 3:], G) and np.allclose(M[3:, :3], -Fm)           # block structure
        assert np.allclose(orc.application_function(ind, 3, lf, k), G, rtol=1e-15)  # f=3 is the E block
    # static limit: k -> 0 gives -G = (I - 3uu)/r^3
    d = np.array([2, -1, 3])
    X = d * 0.5
    r = np.linalg.norm(X)
    u = X / r
    ref = (np.eye(3) - 3 * np.outer(u, u)) / r ** 3
    got = orc.application_function(list(d) + [0, 0, 0], 3, 0.5, 1e-9)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    assert not orc.application_function([1, 1, 1, 1, 1, 1], 3, 0.5, 1.0).any()    # zero self term
    with pytest.raises(ValueError):
        orc.application_function([1, 0, 0, 0, 0, 0], 4, 1.0, 1.0)


@pytest.mark.parametrize("f", [3, 6])
@pytest.mark.parametrize("n", [(2, 3, 4), (3, 2, 2), (4, 4, 3), (1, 3, 2), (5, 4, 5)])
def test_mbt_mvp_dense(orc, n, f):
    rng = np.random.default_rng(11)
    lf, k = 0.7, 1.3 + 0.05j
    Au = orc.mbt_fft_init(n, f, lf, k)
    C = int(np.prod(n))
    x = rng.standard_normal(f * C) + 1j * rng.standard_normal(f * C)
    A = orc.dense_offdiag(n, f, lf, k)
    ref = A @ x
    assert np.abs(orc.mbt_fft_mvp(n, f, Au, x) - ref).max() <= 1e-12 * np.abs(ref).max()


def _problem(orc, f, n=(4, 3, 5), nocc=40, L=3, seed=1, alpha=0.3 + 0.05j, k=0.6):
    rng = np.random.default_rng(seed)
    C = int(np.prod(n))
    addr = np.sort(rng.choice(C, size=nocc, replace=False)).astype(np.int32)
    diag = np.full(f * nocc, 1 / alpha, dtype=complex)
    Au = orc.mbt_fft_init(n, f, 1.0, k)
    B = rng.standard_normal((f * nocc, L)) + 1j * rng.standard_normal((f * nocc, L))
    A = orc.dense_system(n, f, 1.0, k, addr, diag)
    return n, addr, diag, Au, B, A


@pytest.mark.parametrize("f", [3, 6])
def test_block_matvec_dense(orc, f):
    n, addr, diag, Au, B, A = _problem(orc, f)
    for jac in (False, True):
        Y = orc.block_matvec(n, f, addr, Au, diag, B, jacobi=jac)
        ref = A @ B
        if jac:
            ref = ref / diag[:, None]
        assert np.abs(Y - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("method,f", [("cocg_rq", 3), ("bicgstab", 3), ("bicggr", 3), ("bicgstab", 6),
                                      ("bicggr", 6)])
@pytest.mark.parametrize("jacobi", [False, True])
def test_solvers_direct(orc, method, f, jacobi):
    n, addr, diag, Au, B, A = _problem(orc, f)
    X, it, hist = orc.solve(method, n, f, addr, Au, diag, B, 1e-10, 400, jacobi=jacobi)
    Xd = np.linalg.solve(A, B)
    # block COCG terminates after ceil(n/L) steps in exact arithmetic; the
    # FP64 loss of bilinear orthogonality costs about 20 % more here.  The
    # BiCG products take up to about twice that on these non-normal systems.
    nl = -(-A.shape[0] // B.shape[1])
    assert hist[-1] < 1e-10 and it <= (1.3 if method == "cocg_rq" else 2.2) * nl, (it, nl)
    assert np.linalg.norm(A @ X - B) / np.linalg.norm(B) < 1e-9
    assert np.linalg.norm(X - Xd) / np.linalg.norm(Xd) < 1e-8


def _cocg(A, b, tol, itmax):
    x = np.zeros_like(b); r = b.copy(); p = r.copy(); rr = r @ r; hist = []
    for _ in range(itmax):
        q = A @ p; a = rr / (p @ q); x = x + a * p; r = r - a * q
        hist.append(np.linalg.norm(r) / np.linalg.norm(b))
        if hist[-1] < tol:
            break
        rn = r @ r; p = r + (rn / rr) * p; rr = rn
    return x, hist


def _bicgstab(A, b, tol, itmax):
    x = np.zeros_like(b); r = b.copy(); p = r.copy(); r0 = r.copy(); hist = []
    for _ in range(itmax):
        v = A @ p; a = (r0.conj() @ r) / (r0.conj() @ v); t = r - a * v; z = A @ t
        w = (z.conj() @ t) / (z.conj() @ z); x = x + a * p + w * t; rn = t - w * z
        hist.append(np.linalg.norm(rn) / np.linalg.norm(b))
        if hist[-1] < tol:
            break
        beta = -(r0.conj() @ z) / (r0.conj() @ v)
        p = rn + (p - w * v) * beta; r = rn
    return x, hist


def _bicggr(A, b, tol, itmax):
    x = np.zeros_like(b); r = b.copy(); p = r.copy(); r0 = r.copy(); v = A @ r; w = v.copy(); hist = []
    for _ in range(itmax):
        a = (r0.conj() @ r) / (r0.conj() @ v); q = (w.conj() @ r) / (w.conj() @ w)
        u = (p - q * v) * a; y = A @ u; x = x + q * r + u; rn = r - q * w - y
        hist.append(np.linalg.norm(rn) / np.linalg.norm(b))
        if hist[-1] < tol:
            break
        w = A @ rn; g = (r0.conj() @ rn / q) / (r0.conj() @ r); r = rn; p = r + u * g; v = w + y * g
    return x, hist


def _krylov_basis(Aop, Bop, k):
    """Orthonormal basis of span{B, A B, ..., A^(k-1) B}."""
    blocks, Y = [], Bop
    for _ in range(k):
        blocks.append(Y)
        Y = Aop @ Y
    V, _ = np.linalg.qr(np.concatenate(blocks, axis=1))
    return V


@pytest.mark.parametrize("jacobi", [False, True])
def test_cocg_rq_iterates_are_block_galerkin(orc, jacobi):
    """Block COCG (the reference's bl_cocg_rq recurrences,
    bl_cocg_rq_jacobi_mvp_fft.cpp:41-75) on a complex-symmetric A: the k-th
    iterate is fixed by the bilinear Galerkin condition V^T (B - A X_k) = 0
    on V = span{B, AB, ..., A^(k-1) B}, i.e. X_k = V (V^T A V)^(-1) V^T B,
    and its residual blocks are mutually bilinearly orthogonal,
    R_i^T R_j = 0 (i != j).  A wrong alpha, beta, tau or xi term breaks both.
    (DIAG is a scalar here, so the Jacobi operator is complex symmetric too.)"""
    n, addr, diag, Au, B, A = _problem(orc, 3)
    Aop, Bop = (A / diag[:, None], B / diag[:, None]) if jacobi else (A, B)
    assert np.allclose(Aop, Aop.T, atol=1e-12 * np.abs(Aop).max())
    R = [Bop]
    for k in range(1, 5):
        X, it, _ = orc.solve("cocg_rq", n, 3, addr, Au, diag, B, 1e-30, k, jacobi=jacobi)
        assert it == k
        V = _krylov_basis(Aop, Bop, k)
        Xg = V @ np.linalg.solve(V.T @ Aop @ V, V.T @ Bop)
        assert np.linalg.norm(X - Xg) / np.linalg.norm(Xg) <= 1e-9, k
        R.append(Bop - Aop @ X)
    for i in range(len(R)):
        for j in range(i + 1, len(R)):
            G = R[i].T @ R[j]
            assert np.abs(G).max() <= 1e-8 * np.linalg.norm(R[i]) * np.linalg.norm(R[j]), (i, j)


@pytest.mark.parametrize("method", ["bicgstab", "bicggr"])
@pytest.mark.parametrize("f", [3, 6])
def test_bicg_iterates_in_block_krylov_space(orc, method, f):
    """Block BiCGSTAB and BiCGGR use two matvecs per iteration
    (bl_bicgstab_mvp_fft.cpp:43-84, bl_bicggr_mvp_fft.cpp:57-106), so the
    k-th iterate lies in span{B, AB, ..., A^(2k-1) B} and, in general, not in
    the space one power smaller."""
    n, addr, diag, Au, B, A = _problem(orc, f)
    for k in (1, 2):
        X, it, _ = orc.solve(method, n, f, addr, Au, diag, B, 1e-30, k)
        assert it == k
        V = _krylov_basis(A, B, 2 * k)
        assert np.linalg.norm(X - V @ (V.conj().T @ X)) <= 1e-10 * np.linalg.norm(X), k
        W = _krylov_basis(A, B, 2 * k - 1)
        assert np.linalg.norm(X - W @ (W.conj().T @ X)) >= 1e-4 * np.linalg.norm(X), k


@pytest.mark.parametrize("method,ref", [("cocg_rq", _cocg), ("bicgstab", _bicgstab), ("bicggr", _bicggr)])
def test_single_rhs_reduces_to_textbook(orc, method, ref):
    n, addr, diag, Au, B, A = _problem(orc, 3, L=1, seed=5)
    b = B[:, 0]
    X, it, hist = orc.solve(method, n, 3, addr, Au, diag, B, 1e-9, 300)
    xr, hr = ref(A, b, 1e-9, 300)
    # BiCG-type recurrences amplify rounding ~10x per step on this non-normal
    # system, so the histories are compared where rounding is still small
    m = 6
    assert np.allclose(hist[:m], hr[:m], rtol=1e-10, atol=0)
    # the late iterations of BiCGSTAB here move with the FFT's rounding by
    # up to about 15 % of the ~130 steps; both must still converge
    assert abs(len(hist) - len(hr)) <= 0.2 * len(hr)
    assert hist[-1] < 1e-9 and hr[-1] < 1e-9
    assert np.linalg.norm(X[:, 0] - xr) / np.linalg.norm(xr) < 1e-7


def test_qr_reduced_invariants(orc):
    rng = np.random.default_rng(2)
    A = rng.standard_normal((60, 5)) + 1j * rng.standard_normal((60, 5))
    Q, R = orc.qr_reduced(A)
    assert np.abs(Q.conj().T @ Q - np.eye(5)).max() < 1e-14
    assert np.abs(np.tril(R, -1)).max() == 0
    d = np.diag(R)
    assert (d.real > 0).all() and not d.imag.any()
    assert np.abs(Q @ R - A).max() < 1e-13 * np.abs(A).max()
    Qn, Rn = np.linalg.qr(A)
    ph = np.diag(Rn) / np.abs(np.diag(Rn))
    assert np.allclose(Qn * ph[None, :], Q, atol=1e-13)
    assert np.allclose(Rn / ph[:, None], R, atol=1e-12)


def test_qr_reduced_nearly_dependent_columns(orc):
    """qr_reduced keeps only the first-pass projection coefficients in R
    (the comma expressions at qr_reduced.cpp:18,22 make the `tt == 0` test
    true only in pass 0).  With modified Gram-Schmidt the second-pass
    coefficients it drops are O(u |a_k|) whatever the conditioning, so the
    quirk moves results at rounding level only: on columns equal up to 1e-7
    (condition number ~1e7) Q stays orthonormal after the second pass, Q R
    reproduces A to rounding and R matches LAPACK's R (phases normalised)."""
    rng = np.random.default_rng(12)
    m, n = 80, 5
    base = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    A = np.stack([base + 1e-7 * (rng.standard_normal(m) + 1j * rng.standard_normal(m)) for _ in range(n)], axis=1)
    Q, R = orc.qr_reduced(A)
    assert np.abs(Q.conj().T @ Q - np.eye(n)).max() < 1e-12
    assert np.linalg.norm(A - Q @ R) <= 1e-14 * np.linalg.norm(A)
    d = np.diag(R)
    assert (d.real > 0).all() and not d.imag.any() and np.abs(np.tril(R, -1)).max() == 0
    Qn, Rn = np.linalg.qr(A)
    ph = np.diag(Rn) / np.abs(np.diag(Rn))
    assert np.abs(Rn / ph[:, None] - R).max() <= 1e-12 * np.abs(R).max()


@pytest.mark.parametrize("full", [False, True])
@pytest.mark.parametrize("s", [1, 2, 7, 16, 64])
def test_lu_solve_against_lapack(orc, s, full):
    """The solvers' s x s solves (Eigen PartialPivLU / FullPivLU .solve(),
    bl_cocg_rq_mvp_fft.cpp:56,66) against LAPACK's solve: a random complex
    matrix, and one whose leading entries are zero so that the elimination
    must pivot (rows, and for FullPivLU columns) to proceed."""
    rng = np.random.default_rng(30 + s)
    A = rng.standard_normal((s, s)) + 1j * rng.standard_normal((s, s))
    B = rng.standard_normal((s, s)) + 1j * rng.standard_normal((s, s))
    ref = np.linalg.solve(A, B)
    X = orc.lu_solve(A, B, full)
    assert np.linalg.norm(X - ref) <= 1e-12 * np.linalg.norm(ref) * max(1.0, np.linalg.cond(A) / 1e3)
    if s >= 2:
        P = np.roll(np.eye(s), 1, axis=0) + 1e-3 * (rng.standard_normal((s, s)) + 1j * rng.standard_normal((s, s)))
        P[0, 0] = 0.0
        X2 = orc.lu_solve(P, B, full)
        ref2 = np.linalg.solve(P, B)
        assert np.linalg.norm(X2 - ref2) <= 1e-12 * np.linalg.norm(ref2)
