"""Python access to the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module; the product path never does.

Two independent restatements live here:

* ``lib()`` -- ctypes binding of ``oracle/_build/liboracle.so`` (the C
  restatement of the reference recursions, FFT matvec and solvers, see
  oracle.h for the file:line map).
* numpy brute force -- ``green_block`` / ``dense_offdiag`` build the dense
  block-Toeplitz interaction matrix directly from the closed form of
  application_function.cpp:36-39 (f=3) and :50-67 (f=6), without any FFT or
  index recursion, so that the C matvec and the HIP matvec can both be
  pinned against plain matrix algebra on tiny grids.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "liboracle.so")
REF_GMIE_SO = os.path.join(HERE, "_ref", "libgmie_ref.so")

_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
ELEM_FN = ctypes.CFUNCTYPE(None, _ip, ctypes.c_int, ctypes.c_void_p, _dp)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        L = ctypes.CDLL(SO)
        L.orc_fft_length.restype = ctypes.c_int64
        L.orc_bt_pad_length.restype = ctypes.c_int64
        L.orc_bt_pad.restype = ctypes.c_int64
        L.orc_bt_reconstruct.restype = ctypes.c_int64
        L.orc_fft.argtypes = [ctypes.c_int64, _dp, ctypes.c_int]
        L.orc_fft_config.argtypes = [ctypes.c_int64, ctypes.c_int]
        L.orc_fft_config.restype = None
        L.orc_qr_reduced.argtypes = [_dp, ctypes.c_int64, ctypes.c_int, _dp, _dp]
        L.orc_lu_solve.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_lu_solve.restype = None
        L.orc_bt_fft_custom.argtypes = [_ip, ctypes.c_int, ELEM_FN, ctypes.c_void_p, _dp]
        L.orc_block_matvec.argtypes = [_ip, ctypes.c_int, _ip, ctypes.c_int64, _dp, _dp,
                                       ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_solve.argtypes = [ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int, _ip,
                                ctypes.c_int64, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                ctypes.c_int, _dp, _ip, _dp, ctypes.c_int]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _n(n):
    return np.ascontiguousarray(n, dtype=np.int32)


def fft_length(n, f):
    return int(lib().orc_fft_length(_i(_n(n)), f))


def application_function(n_ind, f, lf, k):
    n_ind = np.ascontiguousarray(np.asarray(n_ind, dtype=np.int32).reshape(6))
    out = np.zeros(f * f, dtype=np.complex128)
    k = complex(k)
    rc = lib().orc_application_function(_i(n_ind), f, ctypes.c_double(lf), ctypes.c_double(k.real),
                                        ctypes.c_double(k.imag), _d(out))
    if rc:
        raise ValueError("application_function: unknown f")
    return out.reshape(f, f, order="F")


def bt_fft(n, f, lf, k):
    nn = _n(n)
    out = np.zeros(fft_length(n, f), dtype=np.complex128)
    k = complex(k)
    rc = lib().orc_bt_fft(_i(nn), f, ctypes.c_double(lf), ctypes.c_double(k.real),
                          ctypes.c_double(k.imag), _d(out))
    if rc:
        raise ValueError("bt_fft: unknown f")
    return out


def bt_fft_custom(n, f, elem):
    """BT_fft recursion with a Python element function elem(n_ind 2x3) -> f x f."""
    nn = _n(n)
    out = np.zeros(fft_length(n, f), dtype=np.complex128)

    def cb(n_ind_p, ff, ctx, out_p):
        ind = np.ctypeslib.as_array(n_ind_p, shape=(6,)).reshape(2, 3).copy()
        blk = np.asarray(elem(ind), dtype=np.complex128).reshape(ff, ff)
        dst = np.ctypeslib.as_array(out_p, shape=(2 * ff * ff,))
        dst[:] = np.asfortranarray(blk).ravel(order="F").view(np.float64)

    fn = ELEM_FN(cb)
    lib().orc_bt_fft_custom(_i(nn), f, fn, None, _d(out))
    return out


def bt_pad(n, f, x):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros(int(lib().orc_bt_pad_length(_i(_n(n)), f)), dtype=np.complex128)
    ln = lib().orc_bt_pad(_i(_n(n)), f, _d(x), _d(out))
    assert ln == out.size
    return out


def bt_reconstruct(n, f, bz):
    bz = np.ascontiguousarray(bz, dtype=np.complex128)
    out = np.zeros(f * int(np.prod(n)), dtype=np.complex128)
    ln = lib().orc_bt_reconstruct(_i(_n(n)), f, _d(bz), _d(out))
    assert ln == out.size
    return out


def fft_config(fourstep_min=1 << 24, threads=1):
    """Transforms of at least fourstep_min points run as four-step FFTs on
    `threads` OpenMP threads (when not inside the block matvec's own loop)."""
    lib().orc_fft_config(int(fourstep_min), int(threads))


def fft(x, sign=-1):
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft(y.size, _d(y), sign)
    return y


def mbt_fft_init(n, f, lf, k):
    out = np.zeros(fft_length(n, f), dtype=np.complex128)
    k = complex(k)
    rc = lib().orc_mbt_fft_init(_i(_n(n)), f, ctypes.c_double(lf), ctypes.c_double(k.real),
                                ctypes.c_double(k.imag), _d(out))
    if rc:
        raise ValueError("mbt_fft_init failed")
    return out


def mbt_fft_mvp(n, f, Au_til, p_hat):
    p_hat = np.ascontiguousarray(p_hat, dtype=np.complex128)
    out = np.zeros(f * int(np.prod(n)), dtype=np.complex128)
    rc = lib().orc_mbt_fft_mvp(_i(_n(n)), f, _d(Au_til), _d(p_hat), _d(out))
    if rc:
        raise RuntimeError("mbt_fft_mvp failed")
    return out


def qr_reduced(A):
    A = np.asfortranarray(A, dtype=np.complex128)
    m, nc = A.shape
    Q = np.zeros((m, nc), dtype=np.complex128, order="F")
    R = np.zeros((nc, nc), dtype=np.complex128, order="F")
    lib().orc_qr_reduced(_d(A), m, nc, _d(Q), _d(R))
    return Q, R


def lu_solve(A, B, full=False):
    """X = A^{-1} B by LU with partial or full pivoting (the solvers' s x s solves)."""
    A = np.asfortranarray(A, dtype=np.complex128)
    B = np.asfortranarray(B, dtype=np.complex128)
    s = A.shape[0]
    B2 = B.reshape(s, -1, order="F")
    X = np.zeros_like(B2, order="F")
    lib().orc_lu_solve(_d(A), _d(B2), s, B2.shape[1], int(bool(full)), _d(X))
    return X.reshape(B.shape, order="F")


def block_matvec(n, f, address, Au_til, diag, X, jacobi=False, nthreads=1):
    X = np.asfortranarray(X, dtype=np.complex128)
    Y = np.zeros_like(X, order="F")
    addr = np.ascontiguousarray(address, dtype=np.int32)
    diag = np.ascontiguousarray(diag, dtype=np.complex128)
    rc = lib().orc_block_matvec(_i(_n(n)), f, _i(addr), addr.size, _d(Au_til), _d(diag),
                                int(jacobi), _d(X), _d(Y), X.shape[1], nthreads)
    if rc:
        raise RuntimeError("block_matvec failed")
    return Y


METHODS = {"cocg_rq": 0, "bicgstab": 1, "bicggr": 2}


def solve(method, n, f, address, Au_til, diag, B, tol, itermax, jacobi=False, nthreads=1):
    """Runs bl_<method>[_jacobi]_mvp_fft; returns (X, iters, residual history)."""
    B = np.asfortranarray(B, dtype=np.complex128)
    X = np.zeros_like(B, order="F")
    addr = np.ascontiguousarray(address, dtype=np.int32)
    diag = np.ascontiguousarray(diag, dtype=np.complex128)
    hist = np.zeros(max(itermax, 1))
    iters = ctypes.c_int(0)
    rc = lib().orc_solve(METHODS[method], int(jacobi), _i(_n(n)), f, _i(addr), addr.size,
                         _d(Au_til), _d(diag), _d(B), B.shape[1], tol, itermax, _d(X),
                         ctypes.byref(iters), _d(hist), nthreads)
    if rc:
        raise RuntimeError("orc_solve failed: %d" % rc)
    return X, iters.value, hist[: iters.value]


# ---------------------------------------------------------------------------
# numpy brute force (independent of the C code above)
# ---------------------------------------------------------------------------

def green_block(d, f, lf, k):
    """f x f block of application_function for integer displacement d (3,)."""
    X = np.asarray(d, dtype=np.float64) * lf
    r = float(np.sqrt(X @ X))
    if r == 0.0:
        return np.zeros((f, f), dtype=np.complex128)
    u = X / r
    uu = np.outer(u, u)
    Id = np.eye(3)
    ikr = 1j * k * r
    G = np.exp(ikr) / r * (k * k * (Id - uu) - (1.0 - ikr) / (r * r) * (Id - 3.0 * uu))
    if f == 3:
        return -G
    ux = np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]])
    Fm = -k * k * np.exp(ikr) / r * (1.0 - 1.0 / ikr) * ux
    out = np.zeros((6, 6), dtype=np.complex128)
    out[:3, :3] = -G
    out[:3, 3:] = -Fm
    out[3:, :3] = Fm
    out[3:, 3:] = -G
    return out


def cell_index(n):
    nx, ny, nz = n
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)  # address order


def dense_offdiag(n, f, lf, k):
    """Dense (f C) x (f C) block matrix with blocks M(lf (i - j)) over the cuboid."""
    cells = cell_index(n)
    C = cells.shape[0]
    A = np.zeros((f * C, f * C), dtype=np.complex128)
    cache = {}
    for i in range(C):
        for j in range(C):
            d = tuple(cells[i] - cells[j])
            if d not in cache:
                cache[d] = green_block(d, f, lf, k)
            A[f * i:f * i + f, f * j:f * j + f] = cache[d]
    return A


def dense_system(n, f, lf, k, address, diag):
    """Occupied-dipole system matrix A = diag(DIAG) + offdiag restricted to address."""
    full = dense_offdiag(n, f, lf, k)
    idx = (f * np.asarray(address)[:, None] + np.arange(f)[None, :]).ravel()
    A = full[np.ix_(idx, idx)]
    A = A + np.diag(np.asarray(diag, dtype=np.complex128))
    return A


# ---------------------------------------------------------------------------
# observables (numpy restatements)
# ---------------------------------------------------------------------------

def cross_sections(B, X, diag, f, k, alpha_each=False):
    """Cext, Cabs, Csca per column, restating Result_manager.cpp:65-122:
    Cext = 4 pi k Im(EHinc.col(l).dot(PM.col(l))) (Eigen's dot conjugates its
    left operand), Cabs = 4 pi k sum (Im(conj(1/alpha)) - 2k^3/3) |P|^2 with
    alpha of dipole 0 for every dipole (alpha_each=False, as the reference)
    or each dipole's own; 1/alpha = DIAG_A (Solver.cpp:31-47)."""
    B = np.asarray(B, dtype=np.complex128).reshape(B.shape[0], -1)
    X = np.asarray(X, dtype=np.complex128).reshape(X.shape[0], -1)
    diag = np.asarray(diag, dtype=np.complex128)
    rows = X.shape[0]
    d = diag if alpha_each else np.tile(diag[:f], rows // f)
    w = np.imag(np.conj(d)) - 2.0 * k ** 3 / 3.0
    Cext = 4 * np.pi * k * np.imag(np.sum(np.conj(B) * X, axis=0))
    Cabs = 4 * np.pi * k * np.sum(w[:, None] * np.abs(X) ** 2, axis=0)
    return Cext, Cabs, Cext - Cabs


def far_field(X, positions, f, k, ndirs):
    """Far-field vector E' = k^2 sum_j [(I - n n) P_j - n x M_j] e^{-ik n.r_j}
    (E_far = E' e^{ikr}/r) for each column of X at unit directions ndirs (m, 3).
    Returns (m, 3, L)."""
    X = np.asarray(X, dtype=np.complex128).reshape(X.shape[0], -1)
    N = positions.shape[0]
    Xr = X.reshape(N, f, -1)
    ph = np.exp(-1j * k * (ndirs @ positions.T))               # (m, N)
    AP = np.einsum("mj,jcl->mcl", ph, Xr[:, :3, :])            # (m, 3, L)
    nn = ndirs[:, :, None]
    E = AP - nn * np.sum(nn * AP, axis=1, keepdims=True)
    if f == 6:
        AM = np.einsum("mj,jcl->mcl", ph, Xr[:, 3:, :])
        E = E - np.cross(np.broadcast_to(nn, AM.shape), AM, axis=1)
    return k * k * E


def amplitude_matrix(X, positions, f, k, u, e1, nang):
    """S1..S4 (npairs, nang, 4) of column pairs (2p: pol e1_p, 2p+1: u_p x e1_p)
    in the plane of u_p and e1_p: S2 = -ik theta.E'(e1), S4 = -ik phi.E'(e1),
    S3 = -ik theta.E'(e2), S1 = -ik phi.E'(e2), theta_a = pi a/(nang-1)."""
    X = np.asarray(X, dtype=np.complex128).reshape(X.shape[0], -1)
    u = np.asarray(u, dtype=np.float64).reshape(-1, 3)
    e1 = np.asarray(e1, dtype=np.float64).reshape(-1, 3)
    th = np.pi * np.arange(nang) / (nang - 1)
    out = np.zeros((u.shape[0], nang, 4), dtype=np.complex128)
    for p in range(u.shape[0]):
        n = np.cos(th)[:, None] * u[p] + np.sin(th)[:, None] * e1[p]
        tv = -np.sin(th)[:, None] * u[p] + np.cos(th)[:, None] * e1[p]
        pv = np.cross(u[p], e1[p])
        E = far_field(X[:, 2 * p:2 * p + 2], positions, f, k, n)   # (nang, 3, 2)
        eth = np.einsum("mc,mcl->ml", tv, E)
        eph = np.einsum("c,mcl->ml", pv, E)
        out[p, :, 0] = -1j * k * eph[:, 1]
        out[p, :, 1] = -1j * k * eth[:, 0]
        out[p, :, 2] = -1j * k * eth[:, 1]
        out[p, :, 3] = -1j * k * eph[:, 0]
    return out


def scattering_matrix(S):
    """S11, S12 from S1..S4 (Bohren & Huffman 3.16)."""
    m2 = np.abs(S) ** 2
    return 0.5 * m2.sum(axis=-1), 0.5 * (m2[..., 1] - m2[..., 0] + m2[..., 3] - m2[..., 2])
