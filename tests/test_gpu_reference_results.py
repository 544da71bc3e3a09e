"""The reference's own published results reproduced on the GPU.

Each case is one result section of the reference's result files (the f = 6
block BiCGSTAB-Jacobi sections of DDA_result_BC_rpp20nm.dat and
DDA_result_MAGNETITE_rpp40nm.dat, and the f = 3 block COCG-rq-Jacobi section
of result.dat): the target file, materials, seeded incident waves and
Mie-a_1/b_1 polarizabilities built as the reference builds them, solved by
the HIP path in FP64 with the reference's solver, tol 1e-2 and itermax 300
(Solver.cpp:61-73, Solver.hpp:27-28), and post-processed as Result_manager
does.  The published numbers carry six significant digits, and most cases
agree to about 1e-6.  But a solve stopped at a recurrence residual of 1e-2
fixes its cross sections only to about that tolerance: where the block
BiCGSTAB residual is erratic near the stop (the N128 aggregate at 0.3 um:
0.0116, 0.0129, 0.0166, 0.0158, 0.0032 on the oracle), rounding decides the
stopping iteration and the result moves by a few 1e-3.  The bars are
therefore the solve's tolerance, 1e-2, relative for the means and MAC_mie,
and 1e-2 of the mean for the standard deviations over the four waves (each
wave's value is fixed to that accuracy; a relative bar on the spread of
nearly equal numbers would amplify it).
"""
import importlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DATA = os.path.join(GOLD, "reference_data")

BC, MAG, F3 = "DDA_result_BC_rpp20nm.dat", "DDA_result_MAGNETITE_rpp40nm.dat", "result.dat"
CASES = [(F3, 0, 512, 0.3, 30e-3)]
CASES += [(BC, 0, n, 0.3, 20e-3) for n in (8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096)]
CASES += [(BC, 0, n, 1.0, 20e-3) for n in (64, 256, 1024)]
CASES += [(BC, 2, n, 0.3, 20e-3) for n in (64, 256, 512)]
CASES += [(MAG, 0, n, wl, 40e-3) for n, wl in ((256, 0.3), (1024, 0.5))]


def _section(file, fname, wl0):
    for s in json.load(open(os.path.join(GOLD, "reference_results.json")))["sections"]:
        if s["file"] == file and s["target_fname"] == fname and abs(s["wl0"] - wl0) < 1e-12:
            return s
    raise KeyError((file, fname, wl0))


@pytest.mark.parametrize("case", CASES, ids=["%s-t%d-N%d@%g" % (c[0].split("_")[-1].split(".")[0], c[1], c[2], c[3])
                                             for c in CASES])
def test_gpu_reproduces_published_section(bdda, case):
    file, kind, N, wl0, scale = case
    fname = "DDA_target_%d_agg_N%d_kf1.0_Df2.8.out" % (kind, N)
    print("case reference section %s %s wl0=%g" % (file, fname, wl0), flush=True)
    R = importlib.import_module("block_dda_amd.results")
    M = importlib.import_module("block_dda_amd.materials")
    s = _section(file, fname, wl0)
    core = "MAGNETITE" if file == MAG else "BC"
    mats = [M.Material.named(core, DATA), M.Material.constant(2.25, 1.0, 1.0)]   # block_DDA_main.cpp:66-72
    inp = R.SectionInputs(os.path.join(DATA, fname), mats, scale, M.Material.vacuum(), wl0, s["N_inc"], s["f"])
    t = inp.target
    assert t.n == (s["nx"], s["ny"], s["nz"]) and t.nocc == s["num_element_occupy"]
    r, X = R.solve_section(bdda, inp, tol=1e-2, itermax=300)
    errs = {q: abs(r[q] - s[q]) / abs(s[q]) for q in ("mean_Csca", "mean_MAC", "mean_SSA", "MAC_mie", "std_Csca",
                                                      "std_MAC", "std_SSA")}
    print("  f=%d iters=%d recurrence_res=%.3g errors %s" % (inp.f, r["info"]["iters"], r["info"]["recurrence_res"],
                                                            {q: "%.1e" % e for q, e in errs.items()}), flush=True)
    assert r["info"]["converged"]
    vol = s.get("volume_particle", s.get("volume_core"))
    assert abs(r["props"]["volume_0"] - vol) <= 2e-6 * vol
    for q in ("mean_Csca", "mean_MAC", "mean_SSA", "MAC_mie"):
        assert errs[q] <= 1e-2, (q, r[q], s[q])
    for q in ("Csca", "MAC", "SSA"):
        # each wave's value is fixed to about tol times the mean, so is their spread
        assert abs(r["std_" + q] - s["std_" + q]) <= 1e-2 * abs(s["mean_" + q]), (q, r["std_" + q], s["std_" + q])
