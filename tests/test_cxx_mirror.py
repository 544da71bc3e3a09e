"""The reference-signature C++ interface (include/bdda_block_dda.hpp).

tests/cxx/mirror_check.cpp calls application_function, BT_fft, BT_pad,
BT_reconstruct, MBT_fft_init, MBT_fft_mvp, the six block solvers and
qr_reduced with the reference's argument lists (MBT_utils.hpp:1-6,
DDA_linear_solvers.hpp:1-6, linear_algebra_addon.hpp:1) and records inputs
and outputs; these tests check the records against the oracle.  The CPU
tests cover the layout helpers and the reference's error behaviour (message
on stderr, exit status EXIT_FAILURE); the GPU test covers the device path.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CXX_DIR = os.path.join(ROOT, "tests", "cxx")
EXE = os.path.join(CXX_DIR, "_build", "mirror_check")


@pytest.fixture(scope="module")
def exe(bdda):
    bdda.lib()  # libbdda.so must exist for the link
    if not os.path.exists(EXE):
        subprocess.check_call(["make", "-s", "-C", CXX_DIR])
    return EXE


def read_records(path):
    out = {}
    with open(path, "rb") as fh:
        buf = fh.read()
    pos = 0
    while pos < len(buf):
        nl = int(np.frombuffer(buf, np.int32, 1, pos)[0]); pos += 4
        name = buf[pos:pos + nl].decode(); pos += nl
        kind = int(np.frombuffer(buf, np.int32, 1, pos)[0]); pos += 4
        cnt = int(np.frombuffer(buf, np.int64, 1, pos)[0]); pos += 8
        dt = {0: np.complex128, 1: np.int32, 2: np.float64}[kind]
        out[name] = np.frombuffer(buf, dt, cnt, pos).copy()
        pos += cnt * np.dtype(dt).itemsize
    return out


def test_layout_helpers(exe, orc, tmp_path):
    path = str(tmp_path / "layout.bin")
    subprocess.check_call([exe, "layout", path])
    R = read_records(path)
    n, lf, k = (2, 3, 4), 0.6, 1.2 + 0.1j
    for f in (3, 6):
        t = "f%d_" % f
        app = R[t + "app"].reshape(f, f, order="F")
        ref = orc.application_function([2, 1, 3, 1, 2, 1], f, lf, k)
        assert np.abs(app - ref).max() <= 1e-13 * np.abs(ref).max()
        full = orc.bt_fft(n, f, lf, k)
        assert np.abs(R[t + "bt_fft_l1"] - full).max() <= 1e-13 * np.abs(full).max()
        # a sub-projection is one block of the projection one level up:
        # level-1 pair (2, 1) is block j = n0 - 2 = 0 of the full vector
        # (BT_fft.cpp:39-44) ...
        b1 = f * f * (2 * n[1] - 1) * (2 * n[2] - 1)
        l2 = R[t + "bt_fft_l2"]
        assert l2.size == b1
        assert np.abs(l2 - full[(n[0] - 2) * b1:(n[0] - 1) * b1]).max() <= 1e-13 * np.abs(full).max()
        # ... and level-1 pair (1, 2), level-2 pair (3, 1) is block
        # 2 + n0 - 2 = n0 (BT_fft.cpp:46-49), then block n1 - 3 below it
        b2 = f * f * (2 * n[2] - 1)
        l3 = R[t + "bt_fft_l3"]
        off = n[0] * b1 + (n[1] - 3) * b2
        assert l3.size == b2
        assert np.abs(l3 - full[off:off + b2]).max() <= 1e-13 * np.abs(full).max()
        assert np.array_equal(R[t + "bt_pad"], orc.bt_pad(n, f, R[t + "x"]))
        assert np.array_equal(R[t + "bt_reconstruct"], orc.bt_reconstruct(n, f, R[t + "bz"]))


@pytest.mark.parametrize("mode,msg", [("badf", "unknown f value"), ("badtoken", "Au_til was not made")])
def test_reference_error_behaviour(exe, mode, msg):
    p = subprocess.run([exe, mode], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1          # EXIT_FAILURE, as the reference's exit paths
    assert msg in p.stderr
    assert "not reached" not in p.stdout


@pytest.mark.gpu
def test_device_path(exe, orc, tmp_path):
    path = str(tmp_path / "gpu.bin")
    print("case mirror_check gpu", flush=True)
    p = subprocess.run([exe, "gpu", path], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    R = read_records(path)
    n, lf, k = (6, 5, 7), 0.5, 0.9
    for f in (3, 6):
        t = "f%d_" % f
        Au = orc.mbt_fft_init(n, f, lf, k)
        ref = orc.mbt_fft_mvp(n, f, Au, R[t + "p_hat"])
        assert np.abs(R[t + "mvp"] - ref).max() <= 1e-12 * np.abs(ref).max()
        addr, diag = R[t + "address"], R[t + "diag"]
        B = R[t + "B"].reshape(-1, 2, order="F")
        methods = [("bicgstab", False), ("bicgstab", True), ("bicggr", False), ("bicggr", True)]
        if f == 3:
            methods = [("cocg_rq", False), ("cocg_rq", True)] + methods
        for method, jac in methods:
            name = method + ("_jacobi" if jac else "")
            X = R[t + name].reshape(-1, 2, order="F")
            Xo, _, _ = orc.solve(method, n, f, addr, Au, diag, B, 1e-9, 300, jacobi=jac)
            assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-7, (f, name)
    # the reference's per-iteration line, e.g. bl_cocg_rq_jacobi_mvp_fft.cpp:71
    assert "bl_cocg_rq_jacobi_mvp_fft: iter= 0 relative err= " in p.stdout
    assert "bl_bicggr: iter= 0 relative err= " in p.stdout
    A = R["qr_A"].reshape(60, 5, order="F")
    Qo, Ro = orc.qr_reduced(A)
    assert np.abs(R["qr_Q"].reshape(60, 5, order="F") - Qo).max() < 1e-12
    assert np.abs(R["qr_R"].reshape(5, 5, order="F") - Ro).max() < 1e-12
