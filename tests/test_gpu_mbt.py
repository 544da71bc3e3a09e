"""GPU parity of the MBT operator (HIP library through the C ABI) against
the oracle: the dense block-Toeplitz product on tiny grids, the oracle's
1-D FFT restatement of MBT_fft_mvp on medium grids, and size-independent
properties (linearity, complex symmetry, embedding independence) on larger
grids.  Tolerances: 1e-12 relative in FP64, 1e-5 in FP32."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64, TOL32 = 1e-12, 1e-5


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def rnd(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("f", [3, 6])
@pytest.mark.parametrize("n", [(2, 3, 4), (1, 3, 2), (5, 4, 5), (4, 4, 3), (6, 5, 7), (1, 1, 1)])
def test_mvp_dense(bdda, orc, n, f):
    print("case", n, f, flush=True)
    rng = np.random.default_rng(1)
    lf, k = 0.7, 1.3 + 0.05j
    op = bdda.MBT_fft_init(n, f, lf, k)
    x = rnd(rng, f * int(np.prod(n)))
    ref = orc.dense_offdiag(n, f, lf, k) @ x
    if not np.abs(ref).max():
        assert not np.abs(bdda.MBT_fft_mvp(n, f, op, x)).max()
        return
    assert rel(bdda.MBT_fft_mvp(n, f, op, x), ref) <= TOL64
    assert rel(op.apply(x, "fp32"), ref) <= TOL32


@pytest.mark.parametrize("f", [3, 6])
@pytest.mark.parametrize("n", [(12, 10, 9), (17, 9, 8), (9, 16, 15)])
def test_mvp_oracle_medium(bdda, orc, n, f):
    print("case", n, f, flush=True)
    rng = np.random.default_rng(2)
    lf, k = 0.05, 2.1 + 0.0j
    op = bdda.MBT_fft_init(n, f, lf, k)
    Au = orc.mbt_fft_init(n, f, lf, k)
    X = rnd(rng, f * int(np.prod(n)), 2)
    Y = op.apply(X)
    for j in range(2):
        ref = orc.mbt_fft_mvp(n, f, Au, X[:, j])
        assert rel(Y[:, j], ref) <= TOL64
    assert rel(op.apply(X, "fp32"), Y) <= TOL32


@pytest.mark.parametrize("f", [3, 6])
def test_occupied_matvec_vs_oracle(bdda, orc, f):
    rng = np.random.default_rng(3)
    n = (7, 6, 5)
    C = int(np.prod(n))
    addr = np.sort(rng.choice(C, size=120, replace=False)).astype(np.int32)
    alpha = rng.standard_normal(120) * 0.1 + 0.3 + 0.05j
    diag = np.repeat(1.0 / alpha, f)
    lf, k = 0.4, 1.7
    op = bdda.MBT_fft_init(n, f, lf, k)
    Au = orc.mbt_fft_init(n, f, lf, k)
    op.set_target(addr, diag)
    X = rnd(rng, f * addr.size, 3)
    for jac in (False, True):
        ref = orc.block_matvec(n, f, addr, Au, diag, X, jacobi=jac)
        assert rel(op.apply_occupied(X, "fp64", jac), ref) <= TOL64
        assert rel(op.apply_occupied(X, "fp32", jac), ref) <= TOL32
    # unsorted address order must give the same products
    perm = rng.permutation(addr.size)
    op2 = bdda.MBT_fft_init(n, f, lf, k)
    op2.set_target(addr[perm], np.repeat(1.0 / alpha[perm], f))
    Xp = X.reshape(addr.size, f, 3)[perm].reshape(-1, 3)
    Yp = op2.apply_occupied(Xp)
    ref = orc.block_matvec(n, f, addr, Au, diag, X)
    assert rel(Yp, ref.reshape(addr.size, f, 3)[perm].reshape(-1, 3)) <= TOL64


@pytest.mark.parametrize("f", [3, 6])
@pytest.mark.parametrize("jac", [False, True])
def test_occupied_matvec_sparse_lines(bdda, orc, jac, f):
    """Occupied matvec on the planned passes (14x16x16 -> 32^3) when most
    (ix, iy) lines are empty: the passes then transform only the occupied
    lines.  Occupied lines include the first and the last line, a line whose
    only cell is its last z, and a full line; the rest are random.  f = 6
    runs the planned z and y passes with the all-in-LDS x pass."""
    print("case sparse lines jacobi", jac, "f", f, flush=True)
    rng = np.random.default_rng(31)
    n = (14, 16, 16)
    nxny, nz = n[0] * n[1], n[2]
    lines = set(rng.choice(nxny, size=40, replace=False).tolist()) | {0, nxny - 1, 77, 150}
    cells = []
    for l in sorted(lines):
        if l == 77:
            cells.append(l * nz + nz - 1)
        elif l == 150:
            cells.extend(l * nz + np.arange(nz))
        else:
            cells.extend(l * nz + rng.choice(nz, size=int(rng.integers(1, nz)), replace=False))
    addr = np.array(sorted(cells), dtype=np.int32)
    alpha = rng.standard_normal(addr.size) * 0.1 + 0.3 + 0.05j
    diag = np.repeat(1.0 / alpha, f)
    lf, k = 0.4, 1.7
    op = bdda.MBT_fft_init(n, f, lf, k)
    assert op.grid == (32, 32, 32)
    Au = orc.mbt_fft_init(n, f, lf, k)
    op.set_target(addr, diag)
    X = rnd(rng, f * addr.size, 3)
    ref = orc.block_matvec(n, f, addr, Au, diag, X, jacobi=jac)
    assert rel(op.apply_occupied(X, "fp64", jac), ref) <= TOL64
    assert rel(op.apply_occupied(X, "fp32", jac), ref) <= TOL32
    # the same operator after a denser target, then back: the line lists follow the target
    addr2 = np.sort(rng.choice(int(np.prod(n)), size=900, replace=False)).astype(np.int32)
    diag2 = np.repeat(1.0 / (rng.standard_normal(addr2.size) * 0.1 + 0.3 + 0.05j), f)
    op.set_target(addr2, diag2)
    X2 = rnd(rng, f * addr2.size, 2)
    ref2 = orc.block_matvec(n, f, addr2, Au, diag2, X2, jacobi=jac)
    assert rel(op.apply_occupied(X2, "fp64", jac), ref2) <= TOL64
    op.set_target(addr, diag)
    assert rel(op.apply_occupied(X, "fp64", jac), ref) <= TOL64


@pytest.mark.parametrize("f", [3, 6])
@pytest.mark.parametrize("n", [(14, 16, 10), (14, 10, 16)])
def test_occupied_lines_mixed_plans(bdda, orc, n, f):
    """Occupied matvec where only one of the y and z axes has a planned size
    (14x16x10 -> 32x32x20: planned y passes with line flags, all-in-LDS z
    passes over every line; 14x10x16 -> 32x20x32: planned z passes over the
    occupied lines, all-in-LDS y passes) against the oracle."""
    print("case mixed plans", n, "f", f, flush=True)
    rng = np.random.default_rng(43)
    nxny, nz = n[0] * n[1], n[2]
    lines = np.sort(rng.choice(nxny, size=nxny // 3, replace=False))
    addr = np.concatenate([l * nz + np.flatnonzero(rng.random(nz) < 0.5) for l in lines]).astype(np.int32)
    diag = np.repeat(1.0 / (rng.standard_normal(addr.size) * 0.1 + 0.3 + 0.05j), f)
    lf, k = 0.4, 1.7
    op = bdda.MBT_fft_init(n, f, lf, k)
    assert op.grid == ((32, 32, 20) if n[2] == 10 else (32, 20, 32))
    op.set_target(addr, diag)
    Au = orc.mbt_fft_init(n, f, lf, k)
    X = rnd(rng, f * addr.size, 2)
    for jac in (False, True):
        ref = orc.block_matvec(n, f, addr, Au, diag, X, jacobi=jac)
        assert rel(op.apply_occupied(X, "fp64", jac), ref) <= TOL64
        assert rel(op.apply_occupied(X, "fp32", jac), ref) <= TOL32


@pytest.mark.parametrize("n", [(96, 6, 7), (97, 5, 6)])
def test_x_plan_192(bdda, orc, n):
    """Occupied matvec on grids whose x axis pads to 192 (the FP32 fused x
    pass runs it as 8 x 8 x 3, the FP64 one as 8 x 6 x 4; 97 pads to 200 or
    more and stays off that plan) against the oracle, with and without Jacobi
    scaling, over 3 right-hand sides."""
    print("case x plan 192", n, flush=True)
    rng = np.random.default_rng(47)
    f = 3
    C = int(np.prod(n))
    addr = np.sort(rng.choice(C, size=C * 2 // 3, replace=False)).astype(np.int32)
    diag = np.repeat(1.0 / (rng.standard_normal(addr.size) * 0.1 + 0.3 + 0.05j), f)
    lf, k = 0.45, 1.3
    op = bdda.MBT_fft_init(n, f, lf, k)
    assert (op.grid[0] == 192) == (n[0] == 96)
    op.set_target(addr, diag)
    Au = orc.mbt_fft_init(n, f, lf, k)
    X = rnd(rng, f * addr.size, 3)
    for jac in (False, True):
        ref = orc.block_matvec(n, f, addr, Au, diag, X, jacobi=jac)
        assert rel(op.apply_occupied(X, "fp64", jac), ref) <= TOL64
        assert rel(op.apply_occupied(X, "fp32", jac), ref) <= TOL32


# padded sizes with a compile-time plan: every size of the table below 512
# and the four-stage ones above, each landed on one axis (extent P / 2)
PLAN_SIZES = [40, 56, 72, 80, 112, 120, 168, 200, 240, 288, 336, 392, 576, 640, 720, 784, 896, 960, 1024]
PLAN_CASES = [(P, ax, 3) for P in PLAN_SIZES for ax in (0, 1, 2)]
PLAN_CASES += [(P, ax, 6) for P in (40, 56, 72, 80, 112, 120, 168, 200, 240, 288, 336, 392) for ax in (0, 2)]


@pytest.mark.parametrize("P,axis,f", PLAN_CASES)
def test_plan_sizes(bdda, orc, P, axis, f):
    """Occupied matvec on a grid one of whose axes pads to the plan size P
    (the others stay small and run the all-in-LDS kernels) against the
    oracle, FP64 and FP32, with Jacobi scaling, over 2 right-hand sides."""
    print("case plan size", P, "axis", axis, "f", f, flush=True)
    rng = np.random.default_rng(P * 7 + axis * 3 + f)
    n = [5, 4, 6]
    n[axis] = P // 2
    n = tuple(n)
    C = int(np.prod(n))
    addr = np.sort(rng.choice(C, size=C // 3, replace=False)).astype(np.int32)
    diag = np.repeat(1.0 / (rng.standard_normal(addr.size) * 0.1 + 0.3 + 0.05j), f)
    lf, k = 0.21, 1.4 + 0.01j
    op = bdda.MBT_fft_init(n, f, lf, k)
    assert op.grid[axis] == P
    op.set_target(addr, diag)
    Au = orc.mbt_fft_init(n, f, lf, k)
    X = rnd(rng, f * addr.size, 2)
    ref = orc.block_matvec(n, f, addr, Au, diag, X, jacobi=True)
    assert rel(op.apply_occupied(X, "fp64", True), ref) <= TOL64
    assert rel(op.apply_occupied(X, "fp32", True), ref) <= TOL32


# sizes whose passes are tuned apart from the table's defaults: FP32 x plans
# on 256 threads (280, 336) or as 6 4 4 4 (384), x groups over two
# right-hand sides (FP32 512, FP64 448 and 512), row pitches padded off
# multiples of 64 elements (384, 448, 512 along z), two LDS exchange buffers
# in the FP32 x groups (280, 320, 384, 448, 512; at 512 with a group's two
# right-hand sides taking the buffers in turn), and 343 = 7^3, the
# 16384-monomer workload's padded size (seven-wave x groups of 392
# butterflies, z inverse groups of five lines, y groups on 448 threads)
TUNED_CASES = [(P, ax, 3) for P in (280, 320, 336, 343, 384, 448, 512) for ax in (0, 1, 2)]
TUNED_CASES += [(P, 0, 6) for P in (384, 448, 512)]


@pytest.mark.parametrize("P,axis,f", TUNED_CASES)
def test_tuned_sizes(bdda, orc, P, axis, f):
    """Occupied matvec on a grid one of whose axes pads to P against the
    oracle, FP64 and FP32, with Jacobi scaling, over 3 right-hand sides (an
    odd count, so a group of two right-hand sides is left with one)."""
    print("case tuned size", P, "axis", axis, "f", f, flush=True)
    rng = np.random.default_rng(P * 11 + axis * 5 + f)
    n = [5, 4, 6]
    n[axis] = P // 2
    n = tuple(n)
    C = int(np.prod(n))
    addr = np.sort(rng.choice(C, size=C // 3, replace=False)).astype(np.int32)
    diag = np.repeat(1.0 / (rng.standard_normal(addr.size) * 0.1 + 0.3 + 0.05j), f)
    lf, k = 0.19, 1.3 + 0.02j
    op = bdda.MBT_fft_init(n, f, lf, k)
    assert op.grid[axis] == P
    op.set_target(addr, diag)
    Au = orc.mbt_fft_init(n, f, lf, k)
    X = rnd(rng, f * addr.size, 3)
    ref = orc.block_matvec(n, f, addr, Au, diag, X, jacobi=True)
    assert rel(op.apply_occupied(X, "fp64", True), ref) <= TOL64
    assert rel(op.apply_occupied(X, "fp32", True), ref) <= TOL32


@pytest.mark.parametrize("jac", [False, True])
def test_occupied_lines_against_cuboid(bdda, jac):
    """Occupied matvec (only occupied lines transformed along z) against the
    cuboid-layout matvec (every line transformed) of the same operator, on
    100x128x120 (y padded to 256, where the y passes hold two butterflies per
    thread in their first and last stages).  30 % of the (ix, iy) lines are
    occupied, each by a random subset of its cells."""
    print("case occupied lines vs cuboid jacobi", jac, flush=True)
    rng = np.random.default_rng(41)
    n = (100, 128, 120)
    f = 3
    nxny, nz = n[0] * n[1], n[2]
    lines = np.sort(rng.choice(nxny, size=int(0.3 * nxny), replace=False))
    cells = [l * nz + np.flatnonzero(rng.random(nz) < 0.5) for l in lines]
    addr = np.concatenate(cells).astype(np.int32)
    m = addr.size
    diag = np.repeat(1.0 / (rng.standard_normal(m) * 0.1 + 0.3 + 0.05j), f)
    op = bdda.MBT_fft_init(n, f, 0.02, 2.5)
    assert op.grid[1] == 256
    op.set_target(addr, diag)
    X = rnd(rng, f * m, 2)
    C = int(np.prod(n))
    xc = np.zeros((f * C, 2), dtype=complex)
    idx = (f * addr[:, None] + np.arange(f)[None, :]).ravel()
    xc[idx] = X
    yc = op.apply(xc)
    ref = diag[:, None] * X + yc[idx]
    if jac:
        ref = ref / diag[:, None]
    assert rel(op.apply_occupied(X, "fp64", jac), ref) <= TOL64
    assert rel(op.apply_occupied(X, "fp32", jac), ref) <= TOL32


def test_properties_large(bdda):
    """Linearity and complex symmetry (f=3: x^T A y = y^T A x) on 48x40x36."""
    rng = np.random.default_rng(4)
    n = (48, 40, 36)
    op = bdda.MBT_fft_init(n, 3, 0.02, 3.0 + 0.01j)
    N = 3 * int(np.prod(n))
    x, y = rnd(rng, N), rnd(rng, N)
    a, b = 0.3 - 1.1j, 2.0 + 0.5j
    Ax, Ay, Axy = op.apply(np.stack([x, y, a * x + b * y], axis=1)).T
    assert rel(Axy, a * Ax + b * Ay) <= TOL64
    assert abs(x @ Ay - y @ Ax) <= TOL64 * abs(x @ Ay) * 10
    Ax32 = op.apply(x, "fp32")
    assert rel(Ax32, Ax) <= TOL32


@pytest.mark.parametrize("f", [3, 6])
def test_embedding_independence(bdda, f):
    """A zero layer added to the cuboid changes the padded grid but not the product."""
    rng = np.random.default_rng(5)
    n = (9, 7, 6)
    n2 = (11, 7, 8)
    x = rnd(rng, *n, f)
    x2 = np.zeros(n2 + (f,), dtype=complex)
    x2[: n[0], :, : n[2]] = x
    op = bdda.MBT_fft_init(n, f, 0.3, 1.4)
    op2 = bdda.MBT_fft_init(n2, f, 0.3, 1.4)
    assert op.grid != op2.grid
    y = op.apply(x.ravel()).reshape(n + (f,))
    y2 = op2.apply(x2.ravel()).reshape(n2 + (f,))
    assert rel(y2[: n[0], :, : n[2]], y) <= TOL64


def test_operator_errors(bdda):
    with pytest.raises(bdda.BddaError) as e:
        bdda.MBT_fft_init((4, 4, 4), 4, 1.0, 1.0)
    assert e.value.code == 1
    with pytest.raises(bdda.BddaError):
        bdda.MBT_fft_init((0, 4, 4), 3, 1.0, 1.0)
    op = bdda.MBT_fft_init((4, 4, 4), 3, 1.0, 1.0)
    with pytest.raises(bdda.BddaError):
        op.set_target(np.array([0, 64], dtype=np.int32), np.ones(6))   # outside the cuboid
    with pytest.raises(bdda.BddaError):
        op.set_target(np.array([3, 3], dtype=np.int32), np.ones(6))    # duplicate address
    with pytest.raises(bdda.BddaError):
        op.apply_occupied(np.ones(6))                                   # no target yet
