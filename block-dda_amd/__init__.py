"""block-DDA solve path on AMD Instinct MI355X (gfx950).

Python host mirror of the reference's solve-path interface, calling the
in-tree HIP library ``libbdda.so`` through its C ABI (``include/bdda.h``).
Function names, argument order and meaning follow the reference:

* ``MBT_fft_init(n, f, lf, k)``      -- mvp_fft/MBT_fft_init.cpp:10-42
  (returns a device operator standing in for ``Au_til``; Solver.cpp:55)
* ``MBT_fft_mvp(n, f, Au_til, p_hat)`` -- mvp_fft/MBT_fft_mvp.cpp:8-43
* ``BT_pad`` / ``BT_reconstruct`` / ``BT_fft`` / ``application_function``
  -- mvp_fft/*.cpp (CPU layout helpers)
* ``bl_cocg_rq_mvp_fft`` ... ``bl_bicggr_jacobi_mvp_fft`` (six solvers,
  DDA_linear_solvers.hpp:1-6) with the reference's parameter list
  ``(n, f, address, Au_til, DIAG_A, B, tol, itermax)`` plus keyword options
  for precision mode; they return X like the reference.
* ``qr_reduced(A)`` -- bl_krylov_mvp_fft/qr_reduced.cpp:9-31

The reference exits the process on bad input; this mirror raises
``BddaError`` carrying the library's status code instead.  There is no CPU
fallback: if the HIP library cannot be loaded every call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbdda.so")

BDDA_PREC_FP64 = 1
BDDA_PREC_FP32 = 2
MODE_FP64, MODE_FP32, MODE_MIXED, MODE_MIXED_COLUMNS = 0, 1, 2, 3
COCG_RQ, BICGSTAB, BICGGR = 0, 1, 2
STATUS = {0: "OK", 1: "EINVAL", 2: "ENOMEM", 3: "EHIP", 4: "EBREAKDOWN", 5: "ENOCONV", 6: "ENAN",
          7: "EUNSUPPORTED"}


class BddaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (STATUS.get(code, "?"), code, msg))
        self.code = code


class c64(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("outer_iters", ctypes.c_int), ("matvecs", ctypes.c_int),
                ("converged", ctypes.c_int), ("status", ctypes.c_int), ("recurrence_res", ctypes.c_double),
                ("true_res", ctypes.c_double), ("time_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64

# every symbol include/bdda.h declares, with (restype, argtypes)
SIGNATURES = {
    "bdda_last_error": (ctypes.c_char_p, []),
    "bdda_version": (ctypes.c_int, []),
    "bdda_fft_length": (ctypes.c_int64, [_ip, ctypes.c_int]),
    "bdda_application_function": (ctypes.c_int, [_ip, ctypes.c_int, ctypes.c_double, c64, _vp]),
    "bdda_bt_fft": (ctypes.c_int, [_ip, ctypes.c_int, ctypes.c_double, c64, _vp]),
    "bdda_bt_pad": (ctypes.c_int, [_ip, ctypes.c_int, _vp, _vp]),
    "bdda_bt_reconstruct": (ctypes.c_int, [_ip, ctypes.c_int, _vp, _vp]),
    "bdda_mbt_create": (ctypes.c_int, [_ip, ctypes.c_int, ctypes.c_double, c64, ctypes.c_int,
                                       ctypes.POINTER(_vp)]),
    "bdda_mbt_destroy": (None, [_vp]),
    "bdda_mbt_grid": (ctypes.c_int, [_vp, _ip]),
    "bdda_mbt_apply": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _i64, _vp, _i64, ctypes.c_int]),
    "bdda_mbt_set_target": (ctypes.c_int, [_vp, _ip, _i64, _vp]),
    "bdda_mbt_apply_occupied": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _i64, _vp, _i64,
                                               ctypes.c_int]),
    "bdda_solve": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _i64, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_int, _vp, _i64, ctypes.POINTER(SolveInfo), _dp,
                                  ctypes.c_int]),
    "bdda_qr_reduced": (ctypes.c_int, [_vp, _i64, ctypes.c_int, _vp, _vp]),
    "bdda_lu_solve": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "bdda_hbm_probe": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _dp, _dp, _dp]),
    "bdda_dev_alloc": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "bdda_dev_free": (ctypes.c_int, [_vp]),
    "bdda_dev_trim": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t)]),
    "bdda_dev_pool_stats": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t),
                                           ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "bdda_memcpy_h2d": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "bdda_memcpy_d2h": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "bdda_memcpy_d2d": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "bdda_device_sync": (ctypes.c_int, []),
    "bdda_solve_dev": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_int, _vp, ctypes.POINTER(SolveInfo), _dp,
                                      ctypes.c_int]),
    "bdda_mbt_apply_occupied_dev": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int]),
    "bdda_mbt_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "bdda_mbt_pass_times": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "bdda_mbt_algorithmic_bytes": (ctypes.c_double, [_vp, ctypes.c_int, ctypes.c_int]),
    "bdda_cross_sections": (ctypes.c_int, [_vp, _vp, _i64, _vp, _i64, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]),
    "bdda_cross_sections_dev": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]),
    "bdda_cross_sections_dev_noalloc": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                                       _dp]),
    "bdda_residual_columns": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, _vp, ctypes.c_int64, ctypes.c_int, _dp]),
    "bdda_residual_columns_dev": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _dp]),
    "bdda_scattering_matrix": (ctypes.c_int, [_vp, _vp, _i64, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp,
                                              _vp]),
    "bdda_scattering_matrix_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, _vp]),
    "bdda_estimate_bytes": (ctypes.c_int, [_ip, ctypes.c_int, _i64, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                           ctypes.POINTER(_i64)]),
    "bdda_mbt_set_workspace_limit": (ctypes.c_int, [_vp, ctypes.c_size_t]),
}

ALPHA_FIRST, ALPHA_EACH = 0, 1


def estimate_bytes(n, f, nocc, L, mode=MODE_MIXED):
    """HBM estimate in bytes: dict(init_peak, resident, workspace, krylov)."""
    vals = [_i64() for _ in range(4)]
    keep, p = _n3(n)
    _check(lib().bdda_estimate_bytes(p, int(f), int(nocc), int(L), int(mode), *[ctypes.byref(v) for v in vals]))
    return dict(zip(("init_peak", "resident", "workspace", "krylov"), (v.value for v in vals)))


def lib():
    """Load libbdda.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libbdda.so not built: run __graft_entry__.build() or make -C block-dda_amd")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise BddaError(rc, lib().bdda_last_error().decode())


def _n3(n):
    a = np.ascontiguousarray(np.asarray(n, dtype=np.int32).reshape(3))
    return a, a.ctypes.data_as(_ip)


def _cz(a):
    return np.asfortranarray(a, dtype=np.complex128)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _k(k):
    k = complex(k)
    return c64(k.real, k.imag)


# ---------------------------------------------------------------- layout

def application_function(n, f, n_ind, level, lf, k):
    """f x f block for n_ind (2 x 3): application_function.cpp:7-78."""
    ind = np.ascontiguousarray(np.asarray(n_ind, dtype=np.int32).reshape(6))
    out = np.zeros(f * f, dtype=np.complex128)
    _check(lib().bdda_application_function(ind.ctypes.data_as(_ip), f, lf, _k(k), _ptr(out)))
    return out.reshape(f, f, order="F")


def BT_fft(n, f, lf, k):
    """Au of length f^2 prod(2n-1) (BT_fft.cpp:8-53, called with n_ind = ones, level 1)."""
    a, p = _n3(n)
    out = np.zeros(int(lib().bdda_fft_length(p, f)), dtype=np.complex128)
    _check(lib().bdda_bt_fft(p, f, lf, _k(k), _ptr(out)))
    return out


def BT_pad(n, f, x):
    a, p = _n3(n)
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros(f * f * a[0] * (2 * a[1] - 1) * (2 * a[2] - 1), dtype=np.complex128)
    _check(lib().bdda_bt_pad(p, f, _ptr(x), _ptr(out)))
    return out


def BT_reconstruct(n, f, bz):
    a, p = _n3(n)
    bz = np.ascontiguousarray(bz, dtype=np.complex128)
    out = np.zeros(f * int(np.prod(a)), dtype=np.complex128)
    _check(lib().bdda_bt_reconstruct(p, f, _ptr(bz), _ptr(out)))
    return out


# ---------------------------------------------------------------- operator

class MbtOperator:
    """Device operator built by MBT_fft_init (stands in for Au_til)."""

    def __init__(self, n, f, lf, k, precisions=("fp64", "fp32")):
        mask = (BDDA_PREC_FP64 if "fp64" in precisions else 0) | (BDDA_PREC_FP32 if "fp32" in precisions else 0)
        a, p = _n3(n)
        h = _vp()
        _check(lib().bdda_mbt_create(p, int(f), float(lf), _k(k), mask, ctypes.byref(h)))
        self.h = h
        self.n = tuple(int(v) for v in a)
        self.f = int(f)
        self.lf = float(lf)
        self.k = complex(k)
        self._target = None

    def close(self):
        if getattr(self, "h", None):
            lib().bdda_mbt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def grid(self):
        P = (ctypes.c_int * 3)()
        _check(lib().bdda_mbt_grid(self.h, ctypes.cast(P, _ip)))
        return tuple(P)

    def apply(self, x, precision="fp64"):
        """MBT_fft_mvp on cuboid vectors (f*C,) or (f*C, ncols)."""
        x = _cz(x)
        vec = x.ndim == 1
        X = x.reshape(x.shape[0], -1, order="F")
        Y = np.zeros_like(X, order="F")
        prec = BDDA_PREC_FP64 if precision == "fp64" else BDDA_PREC_FP32
        _check(lib().bdda_mbt_apply(self.h, prec, _ptr(X), X.shape[0], _ptr(Y), Y.shape[0], X.shape[1]))
        return Y[:, 0] if vec else Y

    def set_target(self, address, DIAG_A):
        addr = np.ascontiguousarray(address, dtype=np.int32)
        diag = np.ascontiguousarray(DIAG_A, dtype=np.complex128)
        if diag.size != self.f * addr.size:
            raise BddaError(1, "DIAG_A must have f * len(address) entries")
        key = (addr.tobytes(), diag.tobytes())
        if self._target != key:
            _check(lib().bdda_mbt_set_target(self.h, addr.ctypes.data_as(_ip), addr.size, _ptr(diag)))
            self._target = key

    def apply_occupied(self, X, precision="fp64", jacobi=False):
        X = _cz(X)
        vec = X.ndim == 1
        X2 = X.reshape(X.shape[0], -1, order="F")
        Y = np.zeros_like(X2, order="F")
        prec = BDDA_PREC_FP64 if precision == "fp64" else BDDA_PREC_FP32
        _check(lib().bdda_mbt_apply_occupied(self.h, prec, int(jacobi), _ptr(X2), X2.shape[0], _ptr(Y),
                                             Y.shape[0], X2.shape[1]))
        return Y[:, 0] if vec else Y

    def solve(self, B, tol, itermax, method=COCG_RQ, jacobi=False, mode=MODE_FP64, hist_len=0):
        B = _cz(B)
        X = np.empty_like(B, order="F")  # every entry is written (also when itermax is reached)
        info = SolveInfo()
        hist = np.zeros(max(hist_len, 1))
        rc = lib().bdda_solve(self.h, int(method), int(jacobi), int(mode), _ptr(B), B.shape[0], B.shape[1],
                              float(tol), int(itermax), _ptr(X), X.shape[0], ctypes.byref(info),
                              hist.ctypes.data_as(_dp) if hist_len else None, int(hist_len))
        if rc not in (0, 5):
            _check(rc)
        return X, info.as_dict(), hist[:hist_len]


    def residual_columns(self, B, X):
        """Per-column FP64 true residuals ||B_j - A X_j|| / ||B_j|| (one FP64 block matvec)."""
        B = _cz(B).reshape(B.shape[0], -1, order="F")
        X = _cz(X).reshape(X.shape[0], -1, order="F")
        res = np.zeros(B.shape[1])
        _check(lib().bdda_residual_columns(self.h, _ptr(B), B.shape[0], _ptr(X), X.shape[0], B.shape[1],
                                           res.ctypes.data_as(_dp)))
        return res

    def set_workspace_limit(self, nbytes):
        _check(lib().bdda_mbt_set_workspace_limit(self.h, int(nbytes)))

    def cross_sections(self, B, X, alpha_mode=ALPHA_FIRST):
        """Cext, Cabs, Csca per column (Result_manager.cpp:65-122), FP64 on the GPU."""
        B = _cz(B).reshape(B.shape[0], -1, order="F")
        X = _cz(X).reshape(X.shape[0], -1, order="F")
        L = B.shape[1]
        out = [np.zeros(L) for _ in range(3)]
        _check(lib().bdda_cross_sections(self.h, _ptr(B), B.shape[0], _ptr(X), X.shape[0], L, int(alpha_mode),
                                         *[o.ctypes.data_as(_dp) for o in out]))
        return {"Cext": out[0], "Cabs": out[1], "Csca": out[2]}

    def scattering_matrix(self, X, u, e1, nang=181):
        """S11, S12 (npairs x nang) and S1..S4 (npairs x nang x 4) of column pairs
        (2p: polarisation e1_p, 2p+1: u_p x e1_p) in the plane of u_p and e1_p."""
        X = _cz(X).reshape(X.shape[0], -1, order="F")
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 3)
        e1 = np.ascontiguousarray(e1, dtype=np.float64).reshape(-1, 3)
        npairs = u.shape[0]
        if X.shape[1] < 2 * npairs or e1.shape[0] != npairs:
            raise BddaError(1, "need two columns of X and one e1 per direction")
        S11 = np.zeros((npairs, nang))
        S12 = np.zeros((npairs, nang))
        S = np.zeros((npairs, nang, 4), dtype=np.complex128)
        _check(lib().bdda_scattering_matrix(self.h, _ptr(X), X.shape[0], npairs, u.ctypes.data_as(_dp),
                                            e1.ctypes.data_as(_dp), int(nang), S11.ctypes.data_as(_dp),
                                            S12.ctypes.data_as(_dp), _ptr(S)))
        return S11, S12, S


def MBT_fft_init(n, f, lf, k, plan_fwd=None):
    """Solver::fft_prep + MBT_fft_init (Solver.cpp:17-25,55): builds the device operator."""
    return MbtOperator(n, f, lf, k)


def MBT_fft_mvp(n, f, Au_til, p_hat, plan_fwd=None, plan_inv=None, precision="fp64"):
    """Off-diagonal MBT product on one cuboid vector (MBT_fft_mvp.cpp:8-43)."""
    if tuple(int(v) for v in n) != Au_til.n or int(f) != Au_til.f:
        raise BddaError(1, "n / f do not match the operator")
    return Au_til.apply(p_hat, precision)


def _solver(method, jacobi):
    def run(n, f, address, Au_til, DIAG_A, B, tol, itermax, plan_fwd=None, plan_inv=None, mode=MODE_FP64,
            return_info=False):
        if tuple(int(v) for v in n) != Au_til.n or int(f) != Au_til.f:
            raise BddaError(1, "n / f do not match the operator")
        Au_til.set_target(address, DIAG_A)
        X, info, _ = Au_til.solve(B, tol, itermax, method=method, jacobi=jacobi, mode=mode)
        return (X, info) if return_info else X
    return run


bl_cocg_rq_mvp_fft = _solver(COCG_RQ, False)
bl_cocg_rq_jacobi_mvp_fft = _solver(COCG_RQ, True)
bl_bicgstab_mvp_fft = _solver(BICGSTAB, False)
bl_bicgstab_jacobi_mvp_fft = _solver(BICGSTAB, True)
bl_bicggr_mvp_fft = _solver(BICGGR, False)
bl_bicggr_jacobi_mvp_fft = _solver(BICGGR, True)


def lu_solve(A, B, full=False):
    """X = A^{-1} B on the GPU (the block solvers' s x s solves: Eigen
    PartialPivLU / FullPivLU .solve(), bl_cocg_rq_mvp_fft.cpp:56,66)."""
    A = _cz(A)
    B = _cz(B)
    s = A.shape[0]
    B2 = B.reshape(s, -1, order="F")
    X = np.zeros_like(B2, order="F")
    _check(lib().bdda_lu_solve(_ptr(A), _ptr(B2), _ptr(X), s, B2.shape[1], int(bool(full))))
    return X.reshape(B.shape, order="F")


def dev_pool_stats():
    """The device allocator's bytes held from the runtime and in live blocks,
    runtime allocations made and requests served from idle blocks (current
    device)."""
    held, used = ctypes.c_size_t(), ctypes.c_size_t()
    nalloc, reuse = ctypes.c_int64(), ctypes.c_int64()
    _check(lib().bdda_dev_pool_stats(ctypes.byref(held), ctypes.byref(used), ctypes.byref(nalloc),
                                     ctypes.byref(reuse)))
    return {"held": held.value, "in_use": used.value, "runtime_allocs": nalloc.value, "reuses": reuse.value}


def dev_trim():
    """Give the device allocator's wholly idle memory back to the runtime;
    returns the bytes released."""
    n = ctypes.c_size_t()
    _check(lib().bdda_dev_trim(ctypes.byref(n)))
    return n.value


def qr_reduced(A):
    """(Q, R) of qr_reduced.cpp:9-31 computed on the GPU."""
    A = _cz(A)
    m, n = A.shape
    Q = np.zeros((m, n), dtype=np.complex128, order="F")
    R = np.zeros((n, n), dtype=np.complex128, order="F")
    _check(lib().bdda_qr_reduced(_ptr(A), m, n, _ptr(Q), _ptr(R)))
    return Q, R
