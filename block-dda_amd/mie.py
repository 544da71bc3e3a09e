"""Mie theory for a homogeneous sphere: the exact reference for sphere targets.

Efficiencies Q_ext, Q_sca, Q_abs, the coefficients a_n, b_n and the amplitude
functions S1(theta), S2(theta) of a sphere of size parameter x = k a and
relative refractive index m in a non-absorbing medium, with the conventions
of Bohren & Huffman, "Absorption and Scattering of Light by Small Particles"
(1983), chapter 4:

  a_n = [(D_n(mx)/m + n/x) psi_n(x) - psi_{n-1}(x)] /
        [(D_n(mx)/m + n/x) xi_n(x)  - xi_{n-1}(x)]
  b_n = [(m D_n(mx) + n/x) psi_n(x) - psi_{n-1}(x)] /
        [(m D_n(mx) + n/x) xi_n(x)  - xi_{n-1}(x)]
  S1 = sum (2n+1)/(n(n+1)) (a_n pi_n + b_n tau_n),
  S2 = sum (2n+1)/(n(n+1)) (a_n tau_n + b_n pi_n),

psi_n, xi_n = psi_n - i chi_n the Riccati-Bessel functions (upward
recurrence), D_n the logarithmic derivative of psi_n(mx) (downward
recurrence from well above the truncation order, which is stable for
absorbing spheres), n_stop = x + 4 x^(1/3) + 2 (Wiscombe's criterion).
The reference carries the same quantities in gmie/gmie.cpp and
gmie/gmie_coeff.cpp (also for magnetic spheres and absorbing media; this
module covers magnetic spheres (``coefficients_em``, the polarizabilities
of the reference's targets) in a non-absorbing medium, the case of every
target here).  Angles: theta_j = pi j / (nang - 1), j = 0 .. nang - 1, as gmie
(gmie.cpp:112-115).
"""
import numpy as np


def _nstop(x):
    return int(np.floor(x + 4.0 * x ** (1.0 / 3.0) + 2.0))


def coefficients_em(wl0, r_p, eper_p, mper_p=1.0, eper_m=1.0, mper_m=1.0, nmax=None):
    """a_n, b_n of a sphere of radius r_p with relative permittivity eper_p
    and permeability mper_p in a medium (eper_m, mper_m) at free-space
    wavelength wl0 -- the arguments of gmie_coeff (gmie/gmie_coeff.cpp:8-103):
    m_p = sqrt(eper_p mper_p), m_m = sqrt(eper_m mper_m), relative index
    m = m_p / m_m, x = m_m (2 pi / wl0) r_p, and the magnetic sphere's
    coefficients (Kerker et al., JOSA 73, 765 (1983)), in which 1/m of the
    a_n numerator and denominator becomes mper_p / (mper_m m) and m of b_n
    becomes mper_m m / mper_p.  The medium must be non-absorbing."""
    m_m = np.sqrt(complex(eper_m) * complex(mper_m))
    if abs(m_m.imag) > 1e-14 * abs(m_m):
        raise ValueError("absorbing media are not supported")
    m_p = np.sqrt(complex(eper_p) * complex(mper_p))
    x = m_m.real * 2.0 * np.pi / float(wl0) * float(r_p)
    return coefficients(m_p / m_m.real, x, nmax, mu_p=complex(mper_p), mu_m=complex(mper_m))


def coefficients(m, x, nmax=None, mu_p=1.0, mu_m=1.0):
    """a_n, b_n for n = 1 .. nmax (default n_stop); mu_p / mu_m is the
    particle's relative permeability (1: the non-magnetic sphere)."""
    m = complex(m)
    x = float(x)
    mu = complex(mu_p) / complex(mu_m)
    ns = _nstop(x) if nmax is None else int(nmax)
    mx = m * x
    nmx = int(max(ns, abs(mx))) + 16
    D = np.zeros(nmx + 1, dtype=complex)
    for n in range(nmx, 0, -1):
        D[n - 1] = n / mx - 1.0 / (D[n] + n / mx)
    # Riccati-Bessel psi_n(x), chi_n(x): psi_{-1} = cos x, psi_0 = sin x,
    # chi_{-1} = -sin x, chi_0 = cos x, f_n = (2n-1)/x f_{n-1} - f_{n-2}
    psi = np.zeros(ns + 1)
    chi = np.zeros(ns + 1)
    psi_m1, psi[0] = np.cos(x), np.sin(x)
    chi_m1, chi[0] = -np.sin(x), np.cos(x)
    for n in range(1, ns + 1):
        prev_psi = psi[n - 2] if n >= 2 else psi_m1
        prev_chi = chi[n - 2] if n >= 2 else chi_m1
        psi[n] = (2 * n - 1) / x * psi[n - 1] - prev_psi
        chi[n] = (2 * n - 1) / x * chi[n - 1] - prev_chi
    xi = psi - 1j * chi
    n = np.arange(1, ns + 1)
    Dn = D[1:ns + 1]
    ta = mu / m * Dn + n / x
    tb = m / mu * Dn + n / x
    a = (ta * psi[1:] - psi[:-1]) / (ta * xi[1:] - xi[:-1])
    b = (tb * psi[1:] - psi[:-1]) / (tb * xi[1:] - xi[:-1])
    return a, b


def mie(m, x, nang=181):
    """Returns dict(Qext, Qsca, Qabs, S1, S2, theta, a, b)."""
    a, b = coefficients(m, x)
    ns = a.size
    n = np.arange(1, ns + 1)
    Qsca = 2.0 / x ** 2 * np.sum((2 * n + 1) * (np.abs(a) ** 2 + np.abs(b) ** 2))
    Qext = 2.0 / x ** 2 * np.sum((2 * n + 1) * (a + b).real)
    theta = np.pi * np.arange(nang) / (nang - 1)
    mu = np.cos(theta)
    # angular functions pi_n, tau_n: pi_0 = 0, pi_1 = 1,
    # pi_n = (2n-1)/(n-1) mu pi_{n-1} - n/(n-1) pi_{n-2}, tau_n = n mu pi_n - (n+1) pi_{n-1}
    pi_prev = np.zeros(nang)
    pi_cur = np.ones(nang)
    fn = (2 * n + 1) / (n * (n + 1.0))
    S1 = np.zeros(nang, dtype=complex)
    S2 = np.zeros(nang, dtype=complex)
    for i in range(ns):
        nn = i + 1
        tau = nn * mu * pi_cur - (nn + 1) * pi_prev
        S1 += fn[i] * (a[i] * pi_cur + b[i] * tau)
        S2 += fn[i] * (a[i] * tau + b[i] * pi_cur)
        pi_next = ((2 * nn + 1) * mu * pi_cur - (nn + 1) * pi_prev) / nn
        pi_prev, pi_cur = pi_cur, pi_next
    return {"Qext": float(Qext), "Qsca": float(Qsca), "Qabs": float(Qext - Qsca), "S1": S1, "S2": S2,
            "theta": theta, "a": a, "b": b}


def scattering_matrix(S1, S2):
    """S11 = (|S1|^2 + |S2|^2)/2, S12 = (|S2|^2 - |S1|^2)/2 (Bohren & Huffman 4.74)."""
    return 0.5 * (np.abs(S1) ** 2 + np.abs(S2) ** 2), 0.5 * (np.abs(S2) ** 2 - np.abs(S1) ** 2)
