"""Named workloads of the solve path: target, wavenumber, DIAG_A and B.

Each workload is the input set the reference's Solver would receive
(Solver.cpp:28-56: n, lf, address, DIAG_A = 1/alpha, EHinc) for one of the
measured configurations:

  sphere16   homogeneous sphere, 16 dipoles across its diameter (SPHERE rule,
             Target_manager.cpp:84-112), m = 1.33, ka = 1.0, x- and
             y-polarised plane waves along +z (s = 2)
  agg1024    1024-monomer aggregate (a0 = 20 nm, Df = 2.8, kf = 1.0), 8 dipoles
             across each monomer, m = 1.95 + 0.79i, lambda = 550 nm,
             16 directions x 2 polarisations (s = 32)
  agg16384   16384-monomer aggregate, 4 dipoles across each monomer,
             m = 2.3 + 0.6i, lambda = 400 nm, 8 directions x 2 polarisations
             (s = 16)
  agg256     256-monomer aggregate with the agg16384 material and waves (a
             small case for tests)
  agg4096    4096-monomer aggregate, 6 dipoles across each monomer, for the
             wavelength sweep: `sweep(wl)` gives the workload at wavelength wl
             (300-1000 nm) with m linear in wl from 2.0 + 1.0i to 2.4 + 0.5i and
             4 directions x 2 polarisations (s = 8)

Lengths are in nm for the aggregates and in units of the sphere radius for
sphere16.  The polarizability is the lattice dispersion relation one.
"""
import math
import time

import numpy as np

from . import targets as T

CONFIGS = {
    "sphere16": dict(kind="sphere", nmax=16, m=1.33 + 0.0j, ka=1.0),
    "agg256": dict(kind="agg", N=256, Df=2.8, kf=1.0, a0=20.0, across=4, m=2.3 + 0.6j, wl=400.0, ndir=8,
                   seed=256),
    "agg1024": dict(kind="agg", N=1024, Df=2.8, kf=1.0, a0=20.0, across=8, m=1.95 + 0.79j, wl=550.0, ndir=16,
                    seed=1024),
    "agg4096": dict(kind="agg", N=4096, Df=2.8, kf=1.0, a0=20.0, across=6, m=2.2 + 0.75j, wl=650.0, ndir=4,
                    seed=4096),
    "agg16384": dict(kind="agg", N=16384, Df=2.8, kf=1.0, a0=20.0, across=4, m=2.3 + 0.6j, wl=400.0, ndir=8,
                     seed=16384),
}


class Workload:
    def __init__(self, name, cfg, target, k, alpha, diag, waves, B, seconds):
        self.name, self.cfg, self.target, self.k = name, cfg, target, k
        self.alpha, self.diag, self.waves, self.B = alpha, diag, waves, B
        self.seconds = seconds

    @property
    def s(self):
        return self.B.shape[1]

    def describe(self):
        c = self.cfg
        t = self.target
        if c["kind"] == "sphere":
            return ("%s: sphere, %d dipoles across, grid %s, N=%d dipoles, m=%s, ka=%g, s=%d"
                    % (self.name, c["nmax"], list(t.n), t.nocc, c["m"], c["ka"], self.s))
        return ("%s: %d monomers a0=%gnm Df=%g kf=%g, %d dipoles across, grid %s, N=%d dipoles, m=%s, "
                "lambda=%gnm, s=%d" % (self.name, c["N"], c["a0"], c["Df"], c["kf"], c["across"], list(t.n),
                                       t.nocc, c["m"], c["wl"], self.s))


def sweep_material(wl):
    """m(lambda) of the sweep: linear from 2.0 + 1.0i at 300 nm to 2.4 + 0.5i at 1000 nm."""
    t = (wl - 300.0) / 700.0
    return complex(2.0 + 0.4 * t, 1.0 - 0.5 * t)


def sweep(wl, target=None):
    """The agg4096 workload at wavelength wl (nm); pass `target` to reuse a built lattice."""
    cfg = dict(CONFIGS["agg4096"], wl=float(wl), m=sweep_material(wl))
    return build("agg4096", cfg=cfg, target=target)


def build(name, cfg=None, target=None):
    cfg = CONFIGS[name] if cfg is None else cfg
    t0 = time.time()
    if cfg["kind"] == "sphere":
        radius = 1.0
        tgt = T.sphere_reference(cfg["nmax"], 2.0 * radius)
        k = cfg["ka"] / radius
        waves = T.plane_waves_xy_z(k)
    else:
        tgt = target if target is not None else T.aggregate_target(cfg["N"], cfg["Df"], cfg["kf"], cfg["a0"],
                                                                   cfg["across"], cfg["seed"])
        k = 2.0 * math.pi / cfg["wl"]
        waves = T.plane_waves_sphere(k, cfg["ndir"], cfg["seed"])
    alpha = T.alpha_ldr(cfg["m"], tgt.lf, k, S=0.0)
    diag = T.diag_from_alpha(alpha, tgt.nocc, 3)
    B = waves.rhs(tgt)
    return Workload(name, cfg, tgt, k, alpha, diag, waves, np.asfortranarray(B), time.time() - t0)
