"""Inputs of the solve path: targets, polarizabilities and incident fields.

Produces exactly what the reference's Solver consumes (Target.hpp:17-31):
the cuboid ``n``, the lattice spacing ``lf``, the 0-based ``address`` of each
occupied cell (ny*nz*ix + nz*iy + iz, Target_manager.cpp:98,167-169), the
diagonal ``DIAG_A = 1/alpha`` per component (Solver.cpp:31-47) and the
incident-field block ``B`` (Target_manager.cpp:262-299:
B[f m + q, l] = e_l[q] exp(i k u_l . pos_m), pos_m = lf * (ix, iy, iz)).

Targets
  * ``sphere_reference(nmax, diameter)`` -- the reference SPHERE rule
    (Target_manager.cpp:8-18, 84-112): lf = diameter/nmax, n = nmax per axis,
    occupied when |lf*(ix,iy,iz) - (r,r,r)| <= r.
  * ``aggregate(N, Df, kf, a0, seed)`` -- seeded tunable aggregation of equal
    spherical monomers obeying N = kf (Rg/a0)^Df at every step (the
    particle-cluster variant of the tunable algorithm of Filippov, Zurita &
    Rosner, J. Colloid Interface Sci. 229, 261 (2000)); every new monomer
    touches one monomer of the cluster and overlaps none.
  * ``voxelize(centers, a0, dipoles_across)`` -- cells of spacing
    d = 2 a0 / dipoles_across whose centres lie inside a monomer.
Polarizabilities (Gaussian units, matching application_function's G)
  * ``alpha_cm`` -- Clausius-Mossotti, (3 d^3 / 4 pi)(m^2 - 1)/(m^2 + 2)
  * ``alpha_ldr`` -- lattice dispersion relation (Draine & Goodman, ApJ 405,
    685 (1993)) with S = sum_j (a_j e_j)^2 supplied by the caller.
Incident waves
  * ``plane_waves_xy_z`` -- x- and y-polarised waves along +z (config 1)
  * ``plane_waves_sphere(ndir, seed)`` -- ndir seeded directions uniform on
    the sphere, two orthogonal polarisations each.
"""
import os

import numpy as np

LDR_B1, LDR_B2, LDR_B3 = -1.8915316, 0.1648469, -1.7700004


class Target:
    def __init__(self, n, lf, address, name="", monomers=None, vol=None, material_index=None):
        self.n = tuple(int(v) for v in n)
        self.lf = float(lf)
        self.address = np.ascontiguousarray(address, dtype=np.int32)
        self.name = name
        self.monomers = monomers
        # per element: physical volume and material (index into the list of
        # materials the target was read with); None for generated lattices
        self.vol = None if vol is None else np.asarray(vol, dtype=np.float64)
        self.material_index = None if material_index is None else np.asarray(material_index, dtype=np.int32)

    @property
    def nocc(self):
        return int(self.address.size)

    def positions(self):
        nx, ny, nz = self.n
        a = self.address.astype(np.int64)
        ix = a // (ny * nz)
        iy = (a - ix * ny * nz) // nz
        iz = a - ix * ny * nz - iy * nz
        return self.lf * np.stack([ix, iy, iz], axis=1).astype(np.float64)

    def summary(self):
        return {"name": self.name, "n": list(self.n), "lf": self.lf, "nocc": self.nocc}


# ---------------------------------------------------------------- targets

def read_target_file(path, scale_factor=1.0, nmaterials=2):
    """Target_manager::add_to_target_queue(file, material_0, material_1,
    scale_factor) (Target_manager.cpp:114-202).  The file holds, one per
    line: the primary-particle radius, the aggregate volume, the coating
    volume, "nx ny nz", lf, then one "address vol material_index" line per
    element (address 0-based, ny nz ix + nz iy + iz).  Lengths are scaled by
    scale_factor (lf) and volumes by its cube; elements keep the file's order
    (the rows of B and X follow it).  A material index outside
    0 .. nmaterials-1 raises ValueError (the reference exits)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    try:
        pp_radius = float(lines[0])
        agg_vol = float(lines[1])
        coat_vol = float(lines[2])
        n = tuple(int(v) for v in lines[3].split()[:3])
        lf = float(lines[4]) * scale_factor
    except (IndexError, ValueError) as e:
        raise ValueError("wrong data format in %s: %s" % (path, e))
    if len(n) != 3:
        raise ValueError("wrong data format in %s: lattice size" % path)
    addr, vol, mat = [], [], []
    for ln in lines[5:]:
        tok = ln.split()
        if not tok:
            continue
        if len(tok) < 3:
            raise ValueError("wrong data format in %s: %r" % (path, ln))
        addr.append(int(tok[0]))
        vol.append(float(tok[1]) * scale_factor ** 3)
        mat.append(int(tok[2]))
    mat = np.asarray(mat, dtype=np.int32)
    if mat.size and (mat.min() < 0 or mat.max() >= nmaterials):
        raise ValueError("unknown material in %s" % path)
    t = Target(n, lf, np.asarray(addr, dtype=np.int32), name=os.path.basename(path), vol=vol, material_index=mat)
    t.file_header = {"pp_radius": pp_radius, "aggregate_volume": agg_vol, "coating_volume": coat_vol}
    return t


def sphere_reference(nmax, diameter):
    """Target_manager::add_to_target_queue(SPHERE, nmax, {D, D, D}, ...)."""
    lf = diameter / nmax
    n = (nmax, nmax, nmax)
    r = diameter / 2.0
    ix, iy, iz = np.meshgrid(np.arange(nmax), np.arange(nmax), np.arange(nmax), indexing="ij")
    pos = np.stack([ix, iy, iz], axis=-1).reshape(-1, 3) * lf
    inside = np.linalg.norm(pos - r, axis=1) <= r
    address = (ix * nmax * nmax + iy * nmax + iz).reshape(-1)[inside]
    return Target(n, lf, np.sort(address), name="sphere_ref_n%d" % nmax)


def aggregate(N, Df, kf, a0, seed, max_tries=400):
    """Monomer centres of a seeded tunable aggregate (see module docstring).

    Monomer m (m >= 2) is placed at distance Gamma from the cluster's centre
    of mass, where Gamma follows from requiring the (m+1)-monomer cluster to
    obey Rg^2 = a0^2 ((m+1)/kf)^(2/Df):
        Gamma^2 = (m+1)/m * ((m+1) Rg_{m+1}^2 - m Rg_m^2),
    touching a randomly chosen monomer whose distance from the centre allows
    it, on a random point of the circle where both constraints hold; overlap
    with every other monomer is rejected (spatial hash of cell 2 a0).
    """
    rng = np.random.default_rng(seed)
    X = np.zeros((N, 3))
    if N == 1:
        return X
    u = rng.normal(size=3)
    X[1] = 2 * a0 * u / np.linalg.norm(u)
    two_a = 2.0 * a0
    lim2 = (two_a * (1 - 1e-9)) ** 2
    grid = {}

    def key(p):
        return (int(np.floor(p[0] / two_a)), int(np.floor(p[1] / two_a)), int(np.floor(p[2] / two_a)))

    def insert(i):
        grid.setdefault(key(X[i]), []).append(i)

    def overlaps(p):
        kx, ky, kz = key(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for i in grid.get((kx + dx, ky + dy, kz + dz), ()):
                        d = X[i] - p
                        if d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < lim2:
                            return True
        return False

    insert(0)
    insert(1)
    s1 = X[0] + X[1]
    s2 = float((X[0] ** 2).sum() + (X[1] ** 2).sum())
    for m in range(2, N):
        P = X[:m]
        c = s1 / m
        rg2 = s2 / m - float(c @ c)
        rg2_new = a0 * a0 * ((m + 1) / kf) ** (2.0 / Df)
        gam2 = (m + 1) / m * ((m + 1) * rg2_new - m * rg2)
        gam = np.sqrt(max(gam2, 0.0))
        D = np.sqrt(((P - c) ** 2).sum(axis=1))
        cand = np.nonzero(np.abs(D - gam) <= two_a)[0]
        placed = False
        for _ in range(max_tries):
            if cand.size:
                j = cand[rng.integers(cand.size)]
                Dj = D[j]
                if Dj < 1e-12:
                    continue
                # circle where |p - c| = gam and |p - X_j| = 2 a0
                h = (gam * gam - two_a * two_a + Dj * Dj) / (2 * Dj)
                rho2 = gam * gam - h * h
                if rho2 < 0:
                    continue
                e = (P[j] - c) / Dj
                t = np.cross(e, rng.normal(size=3))
                t /= np.linalg.norm(t)
                p = c + h * e + np.sqrt(rho2) * t
            else:  # no monomer at the required distance: attach anywhere on the surface
                j = rng.integers(m)
                v = rng.normal(size=3)
                p = P[j] + two_a * v / np.linalg.norm(v)
            if not overlaps(p):
                X[m] = p
                insert(m)
                s1 = s1 + p
                s2 += float(p @ p)
                placed = True
                break
        if not placed:
            raise RuntimeError("aggregate: could not place monomer %d" % m)
    return X - X.mean(axis=0)


def voxelize(centers, a0, dipoles_across, name="aggregate"):
    """Cells (spacing d = 2 a0 / dipoles_across) whose centres lie inside a monomer."""
    d = 2.0 * a0 / dipoles_across
    lo = centers.min(axis=0) - a0
    hi = centers.max(axis=0) + a0
    n = np.maximum(np.ceil((hi - lo) / d - 1e-9).astype(int), 1)
    occ = np.zeros(tuple(n), dtype=bool)
    k = int(np.ceil(a0 / d)) + 1
    offs = np.arange(-k, k + 1)
    ox, oy, oz = np.meshgrid(offs, offs, offs, indexing="ij")
    offs3 = np.stack([ox.ravel(), oy.ravel(), oz.ravel()], axis=1)
    base = np.floor((centers - lo) / d).astype(int)
    for ci, b in zip(centers, base):
        cells = b + offs3
        ok = np.all((cells >= 0) & (cells < n), axis=1)
        cells = cells[ok]
        cc = lo + (cells + 0.5) * d
        inside = ((cc - ci) ** 2).sum(axis=1) <= a0 * a0
        cells = cells[inside]
        occ[cells[:, 0], cells[:, 1], cells[:, 2]] = True
    ix, iy, iz = np.nonzero(occ)
    address = (ix * n[1] * n[2] + iy * n[2] + iz).astype(np.int32)
    return Target(tuple(n), d, np.sort(address), name=name, monomers=centers)


def aggregate_target(N, Df, kf, a0, dipoles_across, seed):
    c = aggregate(N, Df, kf, a0, seed)
    t = voxelize(c, a0, dipoles_across, name="agg_N%d_Df%.1f_kf%.1f_seed%d" % (N, Df, kf, seed))
    return t


# ------------------------------------------------------- polarizability

def alpha_cm(m, d):
    m2 = complex(m) ** 2
    return 3.0 * d ** 3 / (4.0 * np.pi) * (m2 - 1.0) / (m2 + 2.0)


def alpha_ldr(m, d, k, S=0.0):
    m2 = complex(m) ** 2
    acm = alpha_cm(m, d)
    kd = k * d
    corr = (LDR_B1 + m2 * LDR_B2 + m2 * LDR_B3 * S) * kd ** 2 - 2.0j / 3.0 * kd ** 3
    return acm / (1.0 + acm / d ** 3 * corr)


def diag_from_alpha(alpha, nocc, f=3, alpha_h=None):
    """DIAG_A of Solver::prepare_DDA_matrix: f = 3, 1/alpha_E repeated over
    the three components of each dipole (Solver.cpp:31-38); f = 6, 1/alpha_E
    for rows 6 m .. 6 m + 2 and 1/alpha_H for rows 6 m + 3 .. 6 m + 5
    (Solver.cpp:39-47).  alpha and alpha_h: one value or one per dipole."""
    a = np.broadcast_to(np.asarray(alpha, dtype=np.complex128), (nocc,))
    if f == 3:
        return np.repeat(1.0 / a, 3)
    if f != 6:
        raise ValueError("unknown f value (3 or 6)")
    if alpha_h is None:
        raise ValueError("f = 6 needs the magnetic polarizability alpha_h")
    h = np.broadcast_to(np.asarray(alpha_h, dtype=np.complex128), (nocc,))
    return np.concatenate([np.repeat((1.0 / a)[:, None], 3, axis=1), np.repeat((1.0 / h)[:, None], 3, axis=1)],
                          axis=1).reshape(6 * nocc)


def alpha_mie(target, materials, wl0, medium, k, f=3):
    """Electric (and for f = 6 magnetic) polarizability of every element from
    its sphere's first Mie coefficients, as Target_manager::set_polarizability
    (Target_manager.cpp:215-260): r_p = (3 vol / 4 pi)^(1/3),
    alpha_E = 3 i a_1 / (2 k^3), alpha_H = 3 i b_1 / (2 k^3), with a_1, b_1 of
    gmie_coeff(wl0, r_p, eper_p, mper_p, eper_m, mper_m) for the element's
    material.  Returns (alpha_E, alpha_H or None), one value per element."""
    from . import mie
    if target.vol is None or target.material_index is None:
        raise ValueError("target has no element volumes / materials (read it with read_target_file)")
    k = complex(k)
    eper_m, mper_m = medium.eper(wl0), medium.mper(wl0)
    aE = np.empty(target.nocc, dtype=np.complex128)
    aH = np.empty(target.nocc, dtype=np.complex128)
    cache = {}
    for i, (v, mi) in enumerate(zip(target.vol, target.material_index)):
        key = (float(v), int(mi))
        if key not in cache:
            mat = materials[int(mi)]
            r_p = np.cbrt(v * 3.0 / (4.0 * np.pi))
            a, b = mie.coefficients_em(wl0, r_p, mat.eper(wl0), mat.mper(wl0), eper_m, mper_m)
            cache[key] = (3.0j * a[0] / (2.0 * k ** 3), 3.0j * b[0] / (2.0 * k ** 3))
        aE[i], aH[i] = cache[key]
    return aE, (aH if f == 6 else None)


# ------------------------------------------------------- incident waves

class Waves:
    def __init__(self, k, dirs, pols, hfields=None):
        k = complex(k)
        self.k = k.real if k.imag == 0.0 else k
        self.dirs = np.asarray(dirs, dtype=np.float64)   # (L, 3) unit propagation vectors
        self.pols = np.asarray(pols, dtype=np.complex128)  # (L, 3) unit polarisation vectors
        # (L, 3) magnetic amplitudes: u x e times the medium's impedance
        # factor sqrt(eper/mper) when not given (1 in vacuum)
        self.hfields = (np.cross(self.dirs, self.pols) if hfields is None
                        else np.asarray(hfields, dtype=np.complex128))

    @property
    def L(self):
        return self.dirs.shape[0]

    def rhs(self, target, f=3):
        """f = 3: B[3 m + q, l] = e_l[q] exp(i k u_l . pos_m)  (Target_manager.cpp:267-275);
        f = 6: rows 6 m .. 6 m + 2 hold e_l, rows 6 m + 3 .. 6 m + 5 hold h_l, times the
        same phase (Target_manager.cpp:277-292)."""
        if f not in (3, 6):
            raise ValueError("unknown f value (3 or 6)")
        pos = target.positions()
        phase = np.exp(1j * self.k * (pos @ self.dirs.T))               # (N, L)
        amp = self.pols.T if f == 3 else np.concatenate([self.pols.T, self.hfields.T], axis=0)
        B = phase[:, None, :] * amp[None, :, :]                          # (N, f, L)
        return np.asfortranarray(B.reshape(f * target.nocc, self.L))


def plane_waves_xy_z(k):
    return Waves(k, [[0, 0, 1], [0, 0, 1]], [[1, 0, 0], [0, 1, 0]])


def plane_waves_sphere(k, ndir, seed):
    rng = np.random.default_rng(seed)
    dirs, pols = [], []
    for _ in range(ndir):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        a = np.array([1.0, 0, 0]) if abs(u[0]) < 0.9 else np.array([0, 1.0, 0])
        e1 = np.cross(u, a)
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(u, e1)
        dirs += [u, u]
        pols += [e1, e2]
    return Waves(k, dirs, pols)
