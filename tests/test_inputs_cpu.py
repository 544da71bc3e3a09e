"""The reference's input path (targets, materials, polarizability, incident
waves) and result post-processing, restated in block-dda_amd, pinned on
the CPU:

  * the seeded rotations' random numbers against libstdc++ itself
    (tests/golden/random_golden.json from make_random_golden.cpp);
  * a_1/b_1 of magnetic and non-magnetic spheres against the reference's
    gmie_coeff (tests/golden/mie_gmie.json, cases_em);
  * the target-file reader, material interpolation, the layouts of B and
    DIAG_A for f = 3 and f = 6 against literal loop restatements of
    Target_manager.cpp:262-299 and Solver.cpp:31-47, the sample statistics
    of Basic_Stats;
  * end to end: the oracle solving what these builders produce reproduces
    the reference's own published result sections (mean/std of Csca, MAC
    and SSA, MAC_mie) -- f = 6 block BiCGSTAB-Jacobi sections of
    DDA_result_BC_rpp20nm.dat and DDA_result_MAGNETITE_rpp40nm.dat, a coated
    target, and the f = 3 block COCG-rq-Jacobi section of result.dat.
"""
import importlib
import json
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DATA = os.path.join(GOLD, "reference_data")


@pytest.fixture(scope="module")
def mods(bdda):
    return {m: importlib.import_module("block_dda_amd." + m)
            for m in ("targets", "materials", "incident", "results", "mie")}


def sections():
    return json.load(open(os.path.join(GOLD, "reference_results.json")))["sections"]


def find_section(file, fname, wl0):
    for s in sections():
        if s["file"] == file and s["target_fname"] == fname and abs(s["wl0"] - wl0) < 1e-12:
            return s
    raise KeyError((file, fname, wl0))


# ------------------------------------------------------------- random numbers

def test_random_engine_matches_libstdcxx(mods):
    I = mods["incident"]
    g = json.load(open(os.path.join(GOLD, "random_golden.json")))
    e = I.MinstdRand0()
    assert [e() for _ in range(len(g["engine_outputs"]))] == g["engine_outputs"]
    e, u = I.MinstdRand0(), I.UniformReal(0.0, 1.0)
    assert [u(e) for _ in range(len(g["draws"]))] == g["draws"]       # bit-exact


def test_incident_waves(mods):
    I, M = mods["incident"], mods["materials"]
    inc = I.IncidentWave(M.Material.vacuum(), 0.3).set_random_incident_directions(6)
    assert inc.N_inc == 6
    assert np.allclose(inc.uinc[0], [0, 0, 1]) and np.allclose(inc.einc[0], [1, 0, 0])
    assert np.allclose(inc.hinc[0], [0, 1, 0])
    assert abs(inc.k - 2 * math.pi / 0.3) < 1e-12
    g = json.load(open(os.path.join(GOLD, "random_golden.json")))["draws"]
    for l in range(1, 6):
        R = inc.rot[l]
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14) and abs(np.linalg.det(R) - 1) < 1e-14
        # the axis and angle are the (l-1)-th triple of draws
        t, p, s = g[3 * (l - 1):3 * l]
        th, ph, psi = math.acos(2 * t - 1), 2 * math.pi * p, 2 * math.pi * s
        n = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
        assert np.allclose(R @ n, n, atol=1e-14)                       # the axis is fixed
        assert abs(np.trace(R) - (1 + 2 * math.cos(psi))) < 1e-13      # by angle psi
        w = np.cross(n, [1.0, 0, 0]); w /= np.linalg.norm(w)
        assert abs(np.dot(np.cross(w, R @ w), n) - math.sin(psi)) < 1e-13   # right-handed
        u, e, h = inc.uinc[l], inc.einc[l].real, inc.hinc[l].real
        assert np.allclose(u, R @ [0, 0, 1.0]) and abs(u @ e) < 1e-14
        assert np.allclose(np.cross(u, e), h, atol=1e-14)                 # H = u x E in vacuum
    # a medium: k = k0 sqrt(eper/mper), |H| = sqrt(eper/mper)
    med = M.Material.constant(1.77, 1.0)
    inc = I.IncidentWave(med, 0.5)
    assert abs(inc.k - 2 * math.pi / 0.5 * math.sqrt(1.77)) < 1e-12
    assert abs(inc.hinc[0][1] - math.sqrt(1.77)) < 1e-15


# ------------------------------------------------------------- polarizability

def test_mie_em_coefficients_match_gmie(mods):
    mie = mods["mie"]
    g = json.load(open(os.path.join(GOLD, "mie_gmie.json")))
    for c in g["cases_em"]:
        a, b = mie.coefficients_em(c["wl0"], c["r_p"], complex(*c["eper_p"]), complex(*c["mper_p"]))
        ga = np.array([complex(*v) for v in c["a"]])
        gb = np.array([complex(*v) for v in c["b"]])
        n = min(ga.size, a.size)
        assert n >= 2
        assert np.abs(a[:n] - ga[:n]).max() <= 1e-10 * np.abs(ga).max(), c["name"]
        assert np.abs(b[:n] - gb[:n]).max() <= 1e-10 * max(np.abs(gb).max(), 1e-300), c["name"]


def test_alpha_mie_small_sphere_limit(mods):
    """alpha_E = 3 i a_1 / (2 k^3) tends to the Clausius-Mossotti value
    r^3 (eps - 1)/(eps + 2) for a small sphere, and for a non-absorbing
    sphere obeys the optical theorem Im(1/alpha_E) = -2 k^3 / 3."""
    T, M = mods["targets"], mods["materials"]
    t = T.Target((2, 2, 2), 1.0, np.arange(2, dtype=np.int32), vol=[4 / 3 * math.pi * 1e-9] * 2,
                 material_index=[0, 1])
    mats = [M.Material.constant(2.5 + 0.4j), M.Material.constant(1.7)]
    wl0 = 1.0
    k = 2 * math.pi / wl0
    aE, aH = T.alpha_mie(t, mats, wl0, M.Material.vacuum(), k, f=6)
    r = 1e-3
    for a, eps in zip(aE, (2.5 + 0.4j, 1.7)):
        cm = r ** 3 * (eps - 1) / (eps + 2)
        assert abs(a - cm) <= 1e-5 * abs(cm)
    assert abs((1 / aE[1]).imag + 2 * k ** 3 / 3) <= 1e-9 * abs(1 / aE[1])
    assert abs((1 / aH[1]).imag + 2 * k ** 3 / 3) <= 1e-6 * abs(1 / aH[1])


def test_ldr_polarizability(mods):
    """LDR (Draine & Goodman 1993): alpha_LDR -> alpha_CM as kd -> 0, its
    published coefficients, and Im(1/alpha) = -2 k^3/3 for a real index."""
    T = mods["targets"]
    assert (T.LDR_B1, T.LDR_B2, T.LDR_B3) == (-1.8915316, 0.1648469, -1.7700004)
    d, m = 0.1, 1.5
    assert abs(T.alpha_ldr(m, d, 1e-6) - T.alpha_cm(m, d)) <= 1e-10 * abs(T.alpha_cm(m, d))
    for k in (0.5, 2.0, 6.0):
        a = T.alpha_ldr(m, d, k)
        assert abs((1 / a).imag + 2 * k ** 3 / 3) <= 1e-10 * abs(1 / a)
        # the second-order correction term, as published
        kd = k * d
        corr = (1 / a - 1 / T.alpha_cm(m, d)) * d ** 3
        assert abs(corr.real - (T.LDR_B1 + m * m * T.LDR_B2) * kd ** 2) <= 1e-12


def test_sphere_reference_rule(mods):
    t = mods["targets"].sphere_reference(16, 2.0)
    assert t.n == (16, 16, 16) and t.nocc == 2106 and abs(t.lf - 0.125) < 1e-15


# ------------------------------------------------------------- aggregates

def _connected_no_overlap(X, a0):
    from scipy.spatial import cKDTree
    tree = cKDTree(X)
    d, _ = tree.query(X, k=2)
    assert d[:, 1].min() >= 2 * a0 * (1 - 1e-9)                  # no two monomers overlap
    pairs = tree.query_pairs(2 * a0 * (1 + 1e-9), output_type="ndarray")
    parent = list(range(len(X)))

    def root(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i
    for i, j in pairs:
        parent[root(i)] = root(j)
    assert len({root(i) for i in range(len(X))}) == 1             # one touching cluster


def _law_n(X, kf, Df, a0):
    c = X.mean(axis=0)
    rg = np.sqrt(((X - c) ** 2).sum(axis=1).mean())
    return kf * (rg / a0) ** Df


def test_cca_law_at_every_join(mods):
    """Cluster-cluster aggregation (Df 1.8, kf 1.3): every join sets the
    centre distance of the law exactly, so each joined cluster and the whole
    aggregate obey N = kf (Rg/a0)^Df; monomers touch and never overlap; the
    same seed gives the same aggregate."""
    T = mods["targets"]
    N, Df, kf, a0 = 512, 1.8, 1.3, 20.0
    X, log = T.aggregate_cca(N, Df, kf, a0, seed=11)
    assert X.shape == (N, 3) and len(log) == 63                   # 64 subclusters of 8
    for nA, nB, gw, gu in log:
        assert abs(gu - gw) <= 1e-12 * gw
    assert abs(_law_n(X, kf, Df, a0) - N) <= 1e-9 * N
    _connected_no_overlap(X, a0)
    X2, _ = T.aggregate_cca(N, Df, kf, a0, seed=11)
    assert np.array_equal(X, X2)
    X3, _ = T.aggregate_cca(N, Df, kf, a0, seed=12)
    assert not np.allclose(X, X3)


def test_cca_merge_sets_gamma(mods):
    """One join: the joined cluster's Rg follows from the two clusters' Rg
    and the centre distance (parallel-axis theorem), and equals the law."""
    T = mods["targets"]
    rng = np.random.default_rng(3)
    A = T.aggregate(64, 1.8, 1.3, 1.0, 1)
    B = T.aggregate(64, 1.8, 1.3, 1.0, 2)
    X, gw, gu = T.cca_merge(A, B, 1.8, 1.3, 1.0, rng)
    assert abs(gu - gw) <= 1e-12 * gw
    cA, cB = X[:64].mean(axis=0), X[64:].mean(axis=0)
    assert abs(np.linalg.norm(cB - cA) - gw) <= 1e-9 * gw
    assert abs(_law_n(X, 1.3, 1.8, 1.0) - 128) <= 1e-9 * 128
    _connected_no_overlap(X, 1.0)


def test_cca_compact_law_widens_and_reports(mods):
    """Df 2.2: joins that the law would interpenetrate are widened (logged),
    the aggregate still touches everywhere and never overlaps."""
    T = mods["targets"]
    X, log = T.aggregate_cca(512, 2.2, 1.0, 20.0, seed=5)
    assert any(gu > gw * (1 + 1e-9) for _, _, gw, gu in log)
    _connected_no_overlap(X, 20.0)


def test_pc_law_at_every_step(mods):
    """Particle-cluster aggregation (the configs' Df 2.8, kf 1.0): from the
    fourth monomer on every prefix obeys the law to rounding."""
    T = mods["targets"]
    N, Df, kf, a0 = 300, 2.8, 1.0, 20.0
    X = T.aggregate(N, Df, kf, a0, 7)
    for m in range(4, N + 1):
        assert abs(_law_n(X[:m], kf, Df, a0) - m) <= 1e-9 * m, m
    _connected_no_overlap(X, a0)
    assert np.array_equal(X, T.aggregate(N, Df, kf, a0, 7))


def test_aggregate_target_method(mods):
    T = mods["targets"]
    t1 = T.aggregate_target(64, 1.8, 1.3, 20.0, 4, seed=1)
    t2 = T.aggregate_target(64, 2.8, 1.0, 20.0, 4, seed=1)
    assert t1.aggregation == "cca" and t2.aggregation == "pc"
    assert T.aggregate_target(64, 2.8, 1.0, 20.0, 4, seed=1, method="cca").aggregation == "cca"
    with pytest.raises(ValueError):
        T.aggregate_target(64, 2.8, 1.0, 20.0, 4, seed=1, method="dla")
    # voxels: spacing 2 a0 / across, each occupied cell centre inside a monomer
    d = 2 * 20.0 / 4
    assert abs(t2.lf - d) < 1e-12 and t2.nocc > 0


# ------------------------------------------------------------- targets, layouts

def test_read_target_file(mods):
    T = mods["targets"]
    t = T.read_target_file(os.path.join(DATA, "DDA_target_0_agg_N8_kf1.0_Df2.8.out"), 20e-3)
    assert t.n == (5, 4, 5) and t.nocc == 8 and abs(t.lf - 0.02) < 1e-15
    assert t.address.tolist() == [31, 40, 65, 67, 7, 52, 16, 19]          # file order
    assert np.allclose(t.vol, 4.18879 * 0.02 ** 3) and t.material_index.tolist() == [0] * 8
    pos = t.positions()
    assert np.allclose(pos[0], 0.02 * np.array([1, 2, 1]))                # 31 = 20*1 + 5*2 + 1
    assert t.file_header["aggregate_volume"] == 33.5103
    c = T.read_target_file(os.path.join(DATA, "DDA_target_2_agg_N8_kf1.0_Df2.8.out"), 20e-3)
    assert c.nocc == 49 and set(c.material_index.tolist()) == {0, 1}
    with pytest.raises(ValueError):
        T.read_target_file(os.path.join(DATA, "DDA_target_2_agg_N8_kf1.0_Df2.8.out"), 20e-3, nmaterials=1)


def test_material_table(mods):
    M = mods["materials"]
    bc = M.Material.named("BC", DATA)
    assert bc.density == 1.8
    assert bc.eper(0.3) == 2.895 + 2.576j and bc.mper(0.4) == 1.0
    assert abs(bc.eper(0.35) - (2.895 + 2.576j + 3.058 + 2.594j) / 2) < 1e-14
    with pytest.raises(ValueError):
        bc.eper(0.2)
    # the reference's interpolated magnetite value at 0.3 um, as it printed it
    mag = M.Material.named("MAGNETITE", DATA)
    s = find_section("DDA_result_MAGNETITE_rpp40nm.dat", "DDA_target_0_agg_N8_kf1.0_Df2.8.out", 0.3)
    e = mag.eper(0.3)
    assert abs(e.real - s["eper_particle"][0]) < 1e-5 and abs(e.imag - s["eper_particle"][1]) < 1e-4


@pytest.mark.parametrize("f", [3, 6])
def test_rhs_and_diag_layouts(mods, f):
    """B = EHinc and DIAG_A against literal loops over (m, l) as in
    Target_manager.cpp:262-299 and Solver.cpp:31-47."""
    T, I, M = mods["targets"], mods["incident"], mods["materials"]
    t = T.read_target_file(os.path.join(DATA, "DDA_target_2_agg_N16_kf1.0_Df2.8.out"), 20e-3)
    inc = I.IncidentWave(M.Material.vacuum(), 0.4).set_random_incident_directions(4)
    B = inc.waves().rhs(t, f)
    aE = np.arange(1, t.nocc + 1) * (1 + 0.5j)
    aH = np.arange(1, t.nocc + 1) * (2 - 0.25j)
    D = T.diag_from_alpha(aE, t.nocc, f, aH if f == 6 else None)
    assert B.shape == (f * t.nocc, 4) and D.shape == (f * t.nocc,)
    nx, ny, nz = t.n
    for m in range(t.nocc):
        a = int(t.address[m])
        ix, iy, iz = a // (ny * nz), (a - ny * nz * (a // (ny * nz))) // nz, a % nz
        pos = t.lf * np.array([ix, iy, iz], dtype=float)
        for l in range(4):
            ph = np.exp(1j * inc.k * np.dot(inc.uinc[l], pos))
            assert np.allclose(B[f * m:f * m + 3, l], inc.einc[l] * ph, atol=1e-15)
            if f == 6:
                assert np.allclose(B[6 * m + 3:6 * m + 6, l], inc.hinc[l] * ph, atol=1e-15)
        assert np.allclose(D[f * m:f * m + 3], 1 / aE[m])
        if f == 6:
            assert np.allclose(D[6 * m + 3:6 * m + 6], 1 / aH[m])


def test_basic_stats(mods):
    R = mods["results"]
    v = np.array([0.3, 0.35, 0.31, 0.42])
    st = R.basic_stats(v)
    assert abs(st["mean"] - v.mean()) < 1e-15
    assert abs(st["std"] - v.std(ddof=1)) < 1e-12


# ------------------------------------------------------------- end to end

CASES = [
    # (result file, target file, wl0, primary-particle radius in um used by the run)
    ("DDA_result_BC_rpp20nm.dat", "DDA_target_0_agg_N8_kf1.0_Df2.8.out", 0.3, 20e-3),
    ("DDA_result_BC_rpp20nm.dat", "DDA_target_0_agg_N16_kf1.0_Df2.8.out", 1.0, 20e-3),
    ("DDA_result_BC_rpp20nm.dat", "DDA_target_0_agg_N32_kf1.0_Df2.8.out", 0.5, 20e-3),
    ("DDA_result_BC_rpp20nm.dat", "DDA_target_2_agg_N8_kf1.0_Df2.8.out", 0.3, 20e-3),
    ("DDA_result_MAGNETITE_rpp40nm.dat", "DDA_target_0_agg_N16_kf1.0_Df2.8.out", 0.3, 40e-3),
    ("result.dat", "DDA_target_0_agg_N512_kf1.0_Df2.8.out", 0.3, 30e-3),
]


def reference_inputs(mods, s, scale):
    R, M = mods["results"], mods["materials"]
    core = "MAGNETITE" if "MAGNETITE" in s["file"] else "BC"
    mats = [M.Material.named(core, DATA), M.Material.constant(2.25, 1.0, 1.0)]   # block_DDA_main.cpp:66-72
    return R.SectionInputs(os.path.join(DATA, s["target_fname"]), mats, scale, M.Material.vacuum(), s["wl0"],
                           s["N_inc"], s["f"])


def check_section(r, s, mean_tol, std_tol):
    vol = s.get("volume_particle", s.get("volume_core"))
    assert abs(r["props"]["volume_0"] - vol) <= 2e-6 * vol
    for q in ("mean_Csca", "mean_MAC", "mean_SSA", "MAC_mie"):
        assert abs(r[q] - s[q]) <= mean_tol * abs(s[q]), (q, r[q], s[q])
    for q in ("std_Csca", "std_MAC", "std_SSA"):
        assert abs(r[q] - s[q]) <= std_tol * abs(s[q]), (q, r[q], s[q])


@pytest.mark.parametrize("case", CASES, ids=[c[1].split("_kf")[0] + "@%g" % c[2] for c in CASES])
def test_oracle_reproduces_published_section(mods, orc, case):
    """The reference's solver (Solver.cpp:61-73: f = 3 COCG-rq-Jacobi, f = 6
    BiCGSTAB-Jacobi; tol 1e-2, itermax 300) on the oracle, fed by the input
    builders: the published numbers carry six significant digits."""
    file, fname, wl0, scale = case
    s = find_section(file, fname, wl0)
    assert s["file"] != "result.dat" or s["f"] == 3
    inp = reference_inputs(mods, s, scale)
    t = inp.target
    assert t.n == (s["nx"], s["ny"], s["nz"]) and t.nocc == s["num_element_occupy"]
    Au = orc.mbt_fft_init(t.n, inp.f, t.lf, inp.k)
    X, it, _ = orc.solve(inp.method_name, t.n, inp.f, t.address, Au, inp.diag, inp.B, 1e-2, 300, jacobi=True,
                         nthreads=min(4, os.cpu_count() or 1))
    Cext, Cabs, _ = orc.cross_sections(inp.B, X, inp.diag, inp.f, inp.k.real)
    check_section(inp.result(Cext, Cabs), s, mean_tol=2e-5, std_tol=5e-4)
