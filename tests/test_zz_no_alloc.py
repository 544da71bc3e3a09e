"""bdda_cross_sections_dev_noalloc (partial sums in the library's file-scope
device array, no device allocation) against bdda_cross_sections_dev (partial
sums in a per-call hipMalloc buffer), through the C ABI on device blocks.

Two calls in a row on different operators, the second with fewer rows and
fewer columns, so stale partial sums of the first call would show; both
alpha modes; f = 3 and f = 6.  The reduction order is the same in both
entry points, so the results must be bit-identical.  Also: more than the
stated maximum of 256 columns is refused with BDDA_EINVAL.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


def _dev_block(bdda, A):
    L = bdda.lib()
    A = np.asfortranarray(A, dtype=np.complex128)
    p = ctypes.c_void_p()
    bdda._check(L.bdda_dev_alloc(A.nbytes, ctypes.byref(p)))
    bdda._check(L.bdda_memcpy_h2d(p, A.ctypes.data, A.nbytes))
    return p


def _xs(bdda, fn, op, dB, dX, L, mode):
    out = [np.full(L, np.nan) for _ in range(3)]
    bdda._check(fn(op.h, dB, dX, L, mode, *[o.ctypes.data_as(_dp) for o in out]))
    return out


@pytest.mark.parametrize("f", [3, 6])
def test_noalloc_cross_sections_match_allocating_entry(bdda, f):
    L = bdda.lib()
    rng = np.random.default_rng(700 + f)
    cases = [((9, 8, 7), 300, 24), ((6, 5, 4), 40, 3)]   # second call: smaller input
    for call, (n, nocc, cols) in enumerate(cases):
        print("case noalloc f=%d call=%d n=%s nocc=%d cols=%d" % (f, call, n, nocc, cols), flush=True)
        C = int(np.prod(n))
        addr = np.sort(rng.choice(C, size=nocc, replace=False)).astype(np.int32)
        diag = (rng.standard_normal(f * nocc) + 1j * rng.standard_normal(f * nocc)) * 2 + 8
        op = bdda.MBT_fft_init(n, f, 0.3, 1.1)
        op.set_target(addr, diag)
        rows = f * nocc
        B = rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))
        X = rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))
        dB, dX = _dev_block(bdda, B), _dev_block(bdda, X)
        try:
            for mode in (bdda.ALPHA_FIRST, bdda.ALPHA_EACH):
                ref = _xs(bdda, L.bdda_cross_sections_dev, op, dB, dX, cols, mode)
                got = _xs(bdda, L.bdda_cross_sections_dev_noalloc, op, dB, dX, cols, mode)
                for q, r, g in zip(("Cext", "Cabs", "Csca"), ref, got):
                    assert np.all(np.isfinite(g)), q
                    assert np.array_equal(g, r), (q, mode, np.abs(g - r).max())
        finally:
            L.bdda_dev_free(dB)
            L.bdda_dev_free(dX)
            op.close()


def test_noalloc_cross_sections_rejects_too_many_columns(bdda):
    L = bdda.lib()
    rng = np.random.default_rng(9)
    n, nocc, cols = (4, 4, 4), 5, 257
    addr = np.arange(nocc, dtype=np.int32)
    op = bdda.MBT_fft_init(n, 3, 0.5, 1.0)
    op.set_target(addr, np.full(3 * nocc, 4.0 + 1.0j))
    B = rng.standard_normal((3 * nocc, cols)) + 0j
    dB = _dev_block(bdda, B)
    try:
        out = [np.zeros(cols) for _ in range(3)]
        rc = L.bdda_cross_sections_dev_noalloc(op.h, dB, dB, cols, 0, *[o.ctypes.data_as(_dp) for o in out])
        assert rc == 1, rc
        assert b"256" in L.bdda_last_error()
    finally:
        L.bdda_dev_free(dB)
        op.close()
