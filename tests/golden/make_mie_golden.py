"""Writes mie_gmie.json: Mie efficiencies, S1/S2 at 181 angles and the first
Mie coefficients from the reference's own gmie (gmie/gmie.cpp,
gmie/gmie_coeff.cpp), compiled from its sources by `make -C oracle ref` into
oracle/_ref/libgmie_ref.so.  Each case is a sphere of radius r_p in vacuum
at wavelength wl0 with relative permittivity eper (mu = 1), i.e. refractive
index m = sqrt(eper) and size parameter x = 2 pi r_p / wl0.

Run (in a tree where oracle/_ref is built):
    python tests/golden/make_mie_golden.py
"""
import ctypes
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SO = os.path.join(ROOT, "oracle", "_ref", "libgmie_ref.so")

# (name, m, x): the sphere16 workload (m = 1.33, ka = 1), the aggregate
# materials at small size parameters, and a larger absorbing sphere
CASES = [
    ("sphere16", 1.33 + 0.0j, 1.0),
    ("m1.5_x2", 1.5 + 0.1j, 2.0),
    ("agg_mat_x0.5", 2.3 + 0.6j, 0.5),
    ("agg1024_mat_x3", 1.95 + 0.79j, 3.0),
    ("m1.33_x10", 1.33 + 0.01j, 10.0),
]
NANG = 181
NCOEF = 8
# (name, wl0 [um], r_p [um], eper_p, mper_p)
EM_CASES = [
    ("bc_0.3um_rp20nm", 0.3, 0.02, 2.895 + 2.576j, 1.0 + 0.0j),
    ("bc_1.0um_rp30nm", 1.0, 0.03, 4.146 + 2.501j, 1.0 + 0.0j),
    ("magnetite_0.3um_rp40nm", 0.3, 0.04, 3.88492 + 4.0195j, 1.0 + 0.0j),
    ("magnetic_x1", 0.5, 0.08, 2.0 + 0.3j, 1.6 + 0.2j),
    ("magnetic_x3", 0.6, 0.3, 3.1 + 1.0j, 0.8 + 0.05j),
]


def main():
    lib = ctypes.CDLL(SO)
    D = ctypes.POINTER(ctypes.c_double)
    lib.ref_gmie.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, D, D, D, D, D, D, D]
    lib.ref_gmie_coeff.argtypes = [ctypes.c_double, ctypes.c_double, D, D, D, D, ctypes.c_int, D, D]
    out = {"source": "gmie/gmie.cpp via oracle/_ref/libgmie_ref.so", "nang": NANG, "cases": []}
    for name, m, x in CASES:
        wl0 = 2 * math.pi
        rp = x  # k0 = 1
        eper = m * m
        ep = (ctypes.c_double * 2)(eper.real, eper.imag)
        one = (ctypes.c_double * 2)(1.0, 0.0)
        q = (ctypes.c_double * 3)()
        S1 = (ctypes.c_double * (2 * NANG))()
        S2 = (ctypes.c_double * (2 * NANG))()
        lib.ref_gmie(wl0, rp, NANG, ep, one, one, one, q, S1, S2)
        a = (ctypes.c_double * (2 * NCOEF))()
        b = (ctypes.c_double * (2 * NCOEF))()
        nc = lib.ref_gmie_coeff(wl0, rp, ep, one, one, one, NCOEF, a, b)
        out["cases"].append({
            "name": name, "m": [m.real, m.imag], "x": x,
            "Qsca": q[0], "Qabs": q[1], "Qext": q[2],
            "S1": [[S1[2 * j], S1[2 * j + 1]] for j in range(NANG)],
            "S2": [[S2[2 * j], S2[2 * j + 1]] for j in range(NANG)],
            "a": [[a[2 * j], a[2 * j + 1]] for j in range(nc)],
            "b": [[b[2 * j], b[2 * j + 1]] for j in range(nc)],
        })
    # gmie_coeff with the arguments Target_manager::set_polarizability passes
    # (wl0 [um], element radius r_p, eper_p, mper_p, vacuum medium): the
    # reference's BC and magnetite primary particles, and magnetic spheres
    # (mper_p != 1) that exercise the mper factors of a_n and b_n
    out["cases_em"] = []
    for name, wl0, rp, eper, mper in EM_CASES:
        ep = (ctypes.c_double * 2)(eper.real, eper.imag)
        mp = (ctypes.c_double * 2)(mper.real, mper.imag)
        one = (ctypes.c_double * 2)(1.0, 0.0)
        a = (ctypes.c_double * (2 * NCOEF))()
        b = (ctypes.c_double * (2 * NCOEF))()
        nc = lib.ref_gmie_coeff(wl0, rp, ep, mp, one, one, NCOEF, a, b)
        out["cases_em"].append({
            "name": name, "wl0": wl0, "r_p": rp, "eper_p": [eper.real, eper.imag], "mper_p": [mper.real, mper.imag],
            "a": [[a[2 * j], a[2 * j + 1]] for j in range(nc)],
            "b": [[b[2 * j], b[2 * j + 1]] for j in range(nc)],
        })
    with open(os.path.join(HERE, "mie_gmie.json"), "w") as fh:
        json.dump(out, fh, indent=0)
    print("wrote %d cases, %d em cases" % (len(out["cases"]), len(out["cases_em"])))


if __name__ == "__main__":
    main()
