"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/bdda.h declares, and its CPU layout helpers (application_function,
BT_fft, BT_pad, BT_reconstruct) agree with the oracle's reference recursions.
No GPU compute is called here."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "bdda.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bdda_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header(bdda):
    L = bdda.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(bdda.SIGNATURES), "Python bindings must cover the header exactly"
    assert L.bdda_version() >= 1


@pytest.mark.parametrize("n", [(2, 2, 2), (2, 3, 4), (3, 1, 2)])
@pytest.mark.parametrize("f", [3, 6])
def test_layout_helpers_match_oracle(bdda, orc, n, f):
    rng = np.random.default_rng(0)
    C = int(np.prod(n))
    x = rng.standard_normal(f * C) + 1j * rng.standard_normal(f * C)
    assert np.array_equal(bdda.BT_pad(n, f, x), orc.bt_pad(n, f, x))
    F = orc.fft_length(n, f)
    bz = rng.standard_normal(F) + 1j * rng.standard_normal(F)
    assert np.array_equal(bdda.BT_reconstruct(n, f, bz), orc.bt_reconstruct(n, f, bz))
    k = 1.2 + 0.1j
    Au = bdda.BT_fft(n, f, 0.6, k)
    ref = orc.bt_fft(n, f, 0.6, k)
    assert np.abs(Au - ref).max() <= 1e-13 * np.abs(ref).max()
    blk = bdda.application_function(n, f, [[2, 1, 3], [1, 2, 1]], 3, 0.6, k)
    assert np.abs(blk - orc.application_function([2, 1, 3, 1, 2, 1], f, 0.6, k)).max() < 1e-13


def test_layout_helpers_reject_bad_f(bdda):
    with pytest.raises(bdda.BddaError) as e:
        bdda.application_function((2, 2, 2), 4, [[1, 1, 1], [1, 1, 2]], 3, 1.0, 1.0)
    assert e.value.code == 1
