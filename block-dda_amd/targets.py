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
  * ``aggregate_cca(N, Df, kf, a0, seed)`` -- seeded cluster-cluster tunable
    aggregation: particle-cluster subclusters joined in equal pairs at the
    law's centre distance (``cca_merge``); ``aggregate_target`` chooses it
    for Df <= CCA_MAX_DF, where the joins can hold the law.
  * ``read_target_file(path, scale_factor)`` -- the reference's target files
    (Target_manager.cpp:114-202).
  * ``voxelize(centers, a0, dipoles_across)`` -- cells of spacing
    d = 2 a0 / dipoles_across whose centres lie inside a monomer.
Polarizabilities (Gaussian units, matching application_function's G)
  * ``alpha_cm`` -- Clausius-Mossotti, (3 d^3 / 4 pi)(m^2 - 1)/(m^2 + 2)
  * ``alpha_ldr`` -- lattice dispersion relation (Draine & Goodman, ApJ 405,
    685 (1993)) with S = sum_j (a_j e_j)^2 supplied by the caller.
  * ``alpha_mie`` -- the reference's 3 i a_1 / (2 k^3) and 3 i b_1 / (2 k^3)
    (Target_manager.cpp:215-260).
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


def _unit(v):
    return v / np.linalg.norm(v)


def _perp_basis(e, rng):
    """Two unit vectors orthogonal to unit e and to each other, at a random angle."""
    t = _unit(np.cross(e, rng.normal(size=3)))
    return t, np.cross(e, t)


def _circle_point(p, q, rp, rq, rng):
    """A random point x with |x - p| = rp and |x - q| = rq (None if none exists)."""
    d = np.linalg.norm(q - p)
    if d < 1e-300:
        return None
    h = (rp * rp - rq * rq + d * d) / (2.0 * d)
    rho2 = rp * rp - h * h
    if rho2 < -1e-12 * rp * rp:
        return None
    e = (q - p) / d
    t, _ = _perp_basis(e, rng)
    return p + h * e + np.sqrt(max(rho2, 0.0)) * t


def _rotation_taking(u, v, spin):
    """Rotation taking unit u to unit v, then turning by `spin` about v."""
    c = float(np.clip(u @ v, -1.0, 1.0))
    w = np.cross(u, v)
    s = np.linalg.norm(w)
    if s < 1e-12:
        if c > 0:
            R = np.eye(3)
        else:                    # half turn about any axis orthogonal to u
            a = _unit(np.cross(u, [1.0, 0, 0] if abs(u[0]) < 0.9 else [0, 1.0, 0]))
            R = 2.0 * np.outer(a, a) - np.eye(3)
    else:
        k = w / s
        K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        R = np.eye(3) + s * K + (1 - c) * (K @ K)
    K = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    S = np.eye(3) + np.sin(spin) * K + (1 - np.cos(spin)) * (K @ K)
    return S @ R


def _rg2(X):
    c = X.mean(axis=0)
    return float(((X - c) ** 2).sum(axis=1).mean()), c


def cca_merge(A, B, Df, kf, a0, rng, max_tries=4000, widen_every=40):
    """Join cluster B to cluster A (monomer centres, radius a0) so that the
    joined cluster obeys N = kf (Rg/a0)^Df, with Rg the point-mass radius of
    gyration: the centres of mass are set Gamma apart with
        Gamma^2 = N^2/(nA nB) a0^2 (N/kf)^(2/Df) - N/nB RgA^2 - N/nA RgB^2,
    N = nA + nB (Filippov, Zurita & Rosner, J. Colloid Interface Sci. 229,
    261 (2000)); one monomer i of A and one monomer j of B touch, and no
    other pair overlaps.  The contact pair is drawn among the pairs whose
    distances dA, dB from their centres admit a contact at that Gamma; the
    contact point, the position of B's centre on its circle and B's spin
    about the contact are drawn at random, and rejected while any pair
    overlaps.  When no pair admits a contact at Gamma (small clusters with
    kf, Df that ask for an Rg the touching pair cannot have), Gamma is
    clamped to the nearest distance that does; when `widen_every` draws in a
    row overlap (compact laws), Gamma grows by 2 %.  Returns (joined
    centres, Gamma wanted, Gamma used)."""
    from scipy.spatial import cKDTree
    nA, nB = A.shape[0], B.shape[0]
    N = nA + nB
    rgA2, cA = _rg2(A)
    rgB2, cB0 = _rg2(B)
    two_a = 2.0 * a0
    g2 = N * N / (nA * nB) * a0 * a0 * (N / kf) ** (2.0 / Df) - N / nB * rgA2 - N / nA * rgB2
    gam_want = float(np.sqrt(max(g2, 0.0)))
    dA = np.linalg.norm(A - cA, axis=1)
    dB = np.linalg.norm(B - cB0, axis=1)
    # a pair (i, j) can touch at centre distance gam when
    #   gam - dA_i - dB_j <= 2a  and  |dA_i - dB_j| - gam <= 2a;
    # clamp gam into the range where some pair can: at most the two farthest
    # monomers' reach, at least the closest |dA_i - dB_j| less 2a
    sB = np.sort(dB)
    pos = np.clip(np.searchsorted(sB, dA), 1, sB.size - 1) if sB.size > 1 else np.zeros(dA.size, dtype=int)
    near = np.minimum(np.abs(dA - sB[pos]), np.abs(dA - sB[np.maximum(pos - 1, 0)]))
    lo = max(0.0, float(near.min()) - two_a)
    hi = float(dA.max() + dB.max() + two_a)
    gam = min(max(gam_want, lo), hi)
    slack = two_a * (1 + 1e-12)
    lim = two_a * (1.0 - 1e-9)
    treeA = cKDTree(A)
    for attempt in range(max_tries):
        if attempt and attempt % widen_every == 0:
            # compact laws (Df near 3) can ask for clusters closer than any
            # overlap-free contact allows: widen Gamma by 2 % per round
            gam = min(gam * 1.02, hi)
        ok_i = np.nonzero(gam - dA - dB.max() <= slack)[0]  # i reaches out far enough for some j
        i = int(ok_i[rng.integers(ok_i.size)])
        cand = np.nonzero((gam - dA[i] - dB <= slack) & (np.abs(dA[i] - dB) - gam <= slack))[0]
        if cand.size == 0:
            continue
        j = int(cand[rng.integers(cand.size)])
        ri, di, dj = A[i] - cA, dA[i], dB[j]
        # D = |cB - r_i| closes both triangles (0, r_i, cB) and (r_i, x_j, cB)
        Dlo = max(abs(gam - di), abs(dj - two_a))
        Dhi = min(gam + di, dj + two_a)
        if Dlo > Dhi * (1 + 1e-12):
            continue
        D = rng.uniform(Dlo, max(Dlo, Dhi))
        if di < 1e-12 * a0:
            cB = _unit(rng.normal(size=3)) * gam
        else:
            cB = _circle_point(np.zeros(3), ri, gam, D, rng)
            if cB is None:
                continue
        if dj < 1e-12 * a0:
            xj = cB
        else:
            xj = _circle_point(ri, cB, two_a, dj, rng)
            if xj is None:
                continue
        if dj < 1e-12 * a0:
            R = _rotation_taking(np.array([0, 0, 1.0]), _unit(rng.normal(size=3)), rng.uniform(0, 2 * np.pi))
        else:
            R = _rotation_taking(_unit(B[j] - cB0), _unit(xj - cB), rng.uniform(0, 2 * np.pi))
        Bn = cA + cB + (B - cB0) @ R.T
        d, _ = treeA.query(Bn, k=1, distance_upper_bound=lim)
        if np.all(np.isinf(d)):
            return np.concatenate([A, Bn]), gam_want, gam
    raise RuntimeError("cca_merge: no contact without overlap after %d tries" % max_tries)


def aggregate_cca(N, Df, kf, a0, seed, subcluster=8):
    """Monomer centres of a seeded cluster-cluster tunable aggregate.

    N monomers are dealt into 2^L subclusters of near-equal size (at most
    `subcluster` monomers each), each grown by the particle-cluster variant
    (`aggregate`) under the same law; the subclusters are then joined in
    pairs, level by level (cluster-cluster aggregation of equal-sized
    clusters, as in the tunable algorithm of Filippov et al. 2000), each
    join by `cca_merge`.  Deterministic for a given seed.  Returns
    (centres, merge log), the log holding per join (nA, nB, Gamma wanted,
    Gamma used)."""
    rng = np.random.default_rng(seed)
    if N <= subcluster:
        return aggregate(N, Df, kf, a0, int(rng.integers(1 << 62))), []
    levels = int(np.ceil(np.log2(N / subcluster)))
    m = 1 << levels
    sizes = [N // m + (1 if q < N % m else 0) for q in range(m)]
    clusters = [aggregate(sz, Df, kf, a0, int(rng.integers(1 << 62))) for sz in sizes]
    log = []
    while len(clusters) > 1:
        nxt = []
        for q in range(0, len(clusters), 2):
            A, B = clusters[q], clusters[q + 1]
            if B.shape[0] > A.shape[0]:
                A, B = B, A
            X, gw, gu = cca_merge(A, B, Df, kf, a0, rng)
            log.append((A.shape[0], B.shape[0], gw, gu))
            nxt.append(X)
        clusters = nxt
    X = clusters[0]
    return X - X.mean(axis=0), log


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


# Cluster-cluster joins of equal clusters can meet N = kf (Rg/a0)^Df exactly
# only while the law lets the two clusters sit far enough apart: the
# centre distance it asks for is Gamma = 2 Rg sqrt(2^(2/Df) - 1), 2.15 Rg at
# Df = 1.8 but 1.60 Rg at Df = 2.8, where two such clusters admit an
# overlap-free contact only beyond about 2.6 Rg.  Measured (cca_merge's
# log): every join exact at Df 1.8 / kf 1.3; a few widened joins up to
# Df 2.0; Df 2.2 ends 8-35 % above the law's N at 4096 monomers; Df 2.5
# cannot be joined.  Above CCA_MAX_DF the particle-cluster variant is used.
CCA_MAX_DF = 2.2


def aggregate_target(N, Df, kf, a0, dipoles_across, seed, method="auto"):
    """Voxelized seeded aggregate: method "cca" (cluster-cluster,
    `aggregate_cca`), "pc" (particle-cluster, `aggregate`) or "auto" (CCA
    when Df <= CCA_MAX_DF, where its joins hold the law, else PC)."""
    if method == "auto":
        method = "cca" if Df <= CCA_MAX_DF else "pc"
    if method == "cca":
        c, _ = aggregate_cca(N, Df, kf, a0, seed)
    elif method == "pc":
        c = aggregate(N, Df, kf, a0, seed)
    else:
        raise ValueError("unknown aggregation method %r" % method)
    t = voxelize(c, a0, dipoles_across, name="agg_%s_N%d_Df%.1f_kf%.1f_seed%d" % (method, N, Df, kf, seed))
    t.aggregation = method
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
