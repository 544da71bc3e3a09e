"""Optical properties of a solved target, as the reference's Result_manager
reports them (Result_manager.cpp:10-175, Basic_Stats.hpp:11-63), and the
driver of one result section (block_DDA_main.cpp:138-163).

Per incident wave l, from the cross sections Cext_l and Cabs_l of the solve
(Cabs with the polarizability of element 0 for every element, as the
reference computes it, Result_manager.cpp:73,97-98):
  Q = C / (pi r_ve^2), SSA = (Cext - Cabs) / Cext,
  MAC = Cabs [um^2] * 1e-12 / mass [g]  (m^2/g),
with volume V = sum of element volumes [um^3], mass = V * 1e-12 * density
[g/cm^3], r_ve = (3 V / 4 pi)^(1/3).  A target whose elements are not all of
one permittivity is split at its first element of another permittivity into
a core (before it) and a coat (Result_manager.cpp:24-51); MAC then uses the
core's mass, and Q is left undefined (NaN) because the reference never sets
the particle cross section of such targets.

The section's statistics over the L waves are the sample mean and standard
deviation of Basic_Stats: var = (mean(v^2) - mean(v)^2) N / (N - 1).
MAC_mie is the volume-equivalent sphere's (Result_manager.cpp:137-175): Mie
Qabs of a sphere of radius r_ve (core radius for coated targets) of the first
element's material.
"""
import math

import numpy as np

from . import incident as I
from . import mie
from . import targets as T

UM3_TO_CM3 = 1.0e-12
UM2_TO_M2 = 1.0e-12


def basic_stats(v):
    v = np.asarray(v, dtype=np.float64)
    N = v.size
    s = float(np.sum(v))
    mean = s / N
    sq_mean = float(np.dot(v, v)) / N
    var = (sq_mean - mean * mean) * N * 1.0 / (N - 1.0) if N > 1 else float("nan")
    return {"mean": mean, "std": math.sqrt(var) if var >= 0 else float("nan"), "var": var}


def target_properties(target, materials, wl0):
    """Volumes, masses and volume-equivalent radii (Result_manager.cpp:24-51)."""
    eper = [materials[i].eper(wl0) for i in target.material_index]
    mper = [materials[i].mper(wl0) for i in target.material_index]
    split = next((i for i, e in enumerate(eper) if e != eper[0]), len(eper))
    vol = np.asarray(target.vol, dtype=np.float64)
    m0 = materials[int(target.material_index[0])]
    p = {"is_homogeneous": split == len(eper), "eper_0": eper[0], "mper_0": mper[0], "density_0": m0.density}
    V0 = float(np.sum(vol[:split]))
    p["volume_0"] = V0
    p["mass_0"] = V0 * UM3_TO_CM3 * m0.density
    p["ve_radius_0"] = np.cbrt(V0 * 3.0 / (4.0 * math.pi))
    p["ve_cross_section_0"] = math.pi * p["ve_radius_0"] ** 2
    if not p["is_homogeneous"]:
        p["volume_coat"] = float(np.sum(vol[split:]))
        p["eper_coat"] = eper[split]
        p["coat_to_core_vratio"] = p["volume_coat"] / V0
    return p


def mie_mac(props, wl0, medium):
    """MAC of the volume-equivalent sphere (Result_manager.cpp:137-175)."""
    r = props["ve_radius_0"]
    a, b = mie.coefficients_em(wl0, r, props["eper_0"], props["mper_0"], medium.eper(wl0), medium.mper(wl0))
    x = (np.sqrt(complex(medium.eper(wl0)) * complex(medium.mper(wl0))).real * 2.0 * math.pi / wl0) * r
    n = np.arange(1, a.size + 1)
    qsca = 2.0 / x ** 2 * float(np.sum((2 * n + 1) * (np.abs(a) ** 2 + np.abs(b) ** 2)))
    qext = 2.0 / x ** 2 * float(np.sum((2 * n + 1) * (a + b).real))
    cabs = (qext - qsca) * math.pi * r * r
    return cabs * UM2_TO_M2 / props["mass_0"]


def section(target, materials, wl0, medium, Cext, Cabs):
    """One result section from the per-wave cross sections (um^2)."""
    props = target_properties(target, materials, wl0)
    Cext = np.asarray(Cext, dtype=np.float64)
    Cabs = np.asarray(Cabs, dtype=np.float64)
    Csca = Cext - Cabs
    g = props["ve_cross_section_0"] if props["is_homogeneous"] else float("nan")
    MAC = Cabs * UM2_TO_M2 / props["mass_0"]
    SSA = (Cext - Cabs) / Cext
    out = {"props": props, "Cext": Cext.tolist(), "Cabs": Cabs.tolist(), "Csca": Csca.tolist(),
           "Qext": (Cext / g).tolist(), "Qabs": (Cabs / g).tolist(), "Qsca": (Csca / g).tolist(),
           "MAC": MAC.tolist(), "SSA": SSA.tolist(), "MAC_mie": mie_mac(props, wl0, medium)}
    for name, v in (("Csca", Csca), ("MAC", MAC), ("SSA", SSA)):
        st = basic_stats(v)
        out["mean_" + name] = st["mean"]
        out["std_" + name] = st["std"]
    return out


class SectionInputs:
    """What Solver receives for one (wavelength, target) pair of the
    reference's main loop (block_DDA_main.cpp:145-156): the target read from
    its file, the incident waves with the seeded rotations, the Mie-a_1
    (and b_1) polarizabilities, DIAG_A and B = EHinc."""

    def __init__(self, target_path, materials, scale_factor, medium, wl0, N_inc, f):
        self.target = T.read_target_file(target_path, scale_factor, nmaterials=len(materials))
        self.materials, self.medium, self.wl0, self.f = materials, medium, float(wl0), int(f)
        self.inc = I.IncidentWave(medium, wl0).set_random_incident_directions(N_inc)
        self.k = complex(self.inc.k)
        self.alpha_E, self.alpha_H = T.alpha_mie(self.target, materials, wl0, medium, self.k, f)
        self.diag = T.diag_from_alpha(self.alpha_E, self.target.nocc, f, self.alpha_H)
        self.B = self.inc.waves().rhs(self.target, f)

    @property
    def method_name(self):
        """Solver::DDA_solve's choice (Solver.cpp:61-73)."""
        return "cocg_rq" if self.f == 3 else "bicgstab"

    def result(self, Cext, Cabs):
        return section(self.target, self.materials, self.wl0, self.medium, Cext, Cabs)


def solve_section(pkg, inp, tol=1e-2, itermax=300, mode=None):
    """Solver::prepare_DDA_matrix + DDA_solve + Result_manager on the GPU:
    the reference's solver for f (block COCG-rq-Jacobi for f = 3, block
    BiCGSTAB-Jacobi for f = 6), its tol 1e-2 and itermax 300
    (Solver.hpp:27-28), FP64 unless `mode` says otherwise."""
    t = inp.target
    op = pkg.MBT_fft_init(t.n, inp.f, t.lf, inp.k)
    try:
        op.set_target(t.address, inp.diag)
        method = pkg.COCG_RQ if inp.f == 3 else pkg.BICGSTAB
        X, info, _ = op.solve(inp.B, tol, itermax, method=method, jacobi=True,
                              mode=pkg.MODE_FP64 if mode is None else mode)
        xs = op.cross_sections(inp.B, X, alpha_mode=pkg.ALPHA_FIRST)
    finally:
        op.close()
    out = inp.result(xs["Cext"], xs["Cabs"])
    out["info"] = info
    return out, X
