"""Incident plane waves of the reference (Incident_Wave.cpp:8-49) and the
right-hand-side block B they give (Target_manager.cpp:262-299).

Wave 0 propagates along +z with E along +x and H = sqrt(eper_m/mper_m) +y
(Incident_Wave.cpp:14-25).  ``set_random_incident_directions(L)`` adds
L - 1 waves, each the default wave turned by a rotation drawn from
``std::default_random_engine`` with its default seed and
``std::uniform_real_distribution<double>(0, 1)`` (Incident_Wave.cpp:29-49):
per wave, theta = acos(2 u1 - 1), phi = 2 pi u2 (the rotation axis
n = (sin theta cos phi, sin theta sin phi, cos theta)) and psi = 2 pi u3 (the
angle), in that order.  The rotation is R = cos psi I + (1 - cos psi) n n^T
+ sin psi [n]x (Rodrigues' formula; Eigen's AngleAxisd), and
u_l = R (0,0,1), e_l = R (1,0,0), h_l = R (0, sqrt(eper_m/mper_m), 0).

The generator is restated here rather than called, so that the directions
are the reference's on any platform:
  * libstdc++'s default_random_engine is minstd_rand0, x <- 16807 x mod
    (2^31 - 1), seed 1;
  * generate_canonical<double, 53> takes k = max(1, ceil(53 / floor(log2 R)))
    = 2 draws, R = max - min + 1 = 2^31 - 2, and returns
    ((g1 - 1) + (g2 - 1) R) / R^2 in double arithmetic (the next double
    below 1 if that rounds to 1).
tests/golden/make_random_golden.cpp writes the same numbers from libstdc++
itself for the test that pins this restatement.
"""
import math

import numpy as np

from .targets import Waves


class MinstdRand0:
    """std::minstd_rand0: x <- 16807 x mod 2147483647 (default seed 1)."""
    A = 16807
    M = 2147483647
    MIN, MAX = 1, M - 1

    def __init__(self, seed=1):
        s = int(seed) % self.M
        self.x = s if s != 0 else 1

    def __call__(self):
        self.x = (self.A * self.x) % self.M
        return self.x


def generate_canonical_53(g):
    """libstdc++ generate_canonical<double, 53>(g) for a 31-bit engine."""
    R = float(g.MAX - g.MIN + 1)                  # 2147483646, exact in double
    log2r = int(math.log(R) / math.log(2.0))      # 30 (truncated, as size_t)
    k = max(1, (53 + log2r - 1) // log2r)         # 2
    s, t = 0.0, 1.0
    for _ in range(k):
        s += float(g() - g.MIN) * t
        t *= R
    r = s / t
    return math.nextafter(1.0, 0.0) if r >= 1.0 else r


class UniformReal:
    """std::uniform_real_distribution<double>(a, b)."""

    def __init__(self, a=0.0, b=1.0):
        self.a, self.b = float(a), float(b)

    def __call__(self, g):
        return generate_canonical_53(g) * (self.b - self.a) + self.a


def rotation(psi, n):
    """R = cos psi I + (1 - cos psi) n n^T + sin psi [n]x."""
    n = np.asarray(n, dtype=np.float64)
    c, s = math.cos(psi), math.sin(psi)
    nx = np.array([[0.0, -n[2], n[1]], [n[2], 0.0, -n[0]], [-n[1], n[0], 0.0]])
    return c * np.eye(3) + (1.0 - c) * np.outer(n, n) + s * nx


class IncidentWave:
    """Incident_Wave(medium, wl0) (Incident_Wave.cpp:8-27)."""

    def __init__(self, medium, wl0):
        self.wl0 = float(wl0)
        self.k0 = 2.0 * math.pi / self.wl0
        self.eper_m = complex(medium.eper(self.wl0))
        self.mper_m = complex(medium.mper(self.wl0))
        self.k = self.k0 * np.sqrt(self.eper_m / self.mper_m)
        self.u0 = np.array([0.0, 0.0, 1.0])
        self.e0 = np.array([1.0, 0.0, 0.0], dtype=np.complex128)
        self.h0 = np.array([0.0, np.sqrt(self.eper_m / self.mper_m), 0.0], dtype=np.complex128)
        self.rot = [np.eye(3)]
        self.uinc = [self.u0.copy()]
        self.einc = [self.e0.copy()]
        self.hinc = [self.h0.copy()]

    @property
    def N_inc(self):
        return len(self.uinc)

    def set_random_incident_directions(self, L, seed=1):
        """Waves 1 .. L-1 from the default-seeded engine (Incident_Wave.cpp:29-49)."""
        g = MinstdRand0(seed)
        urf = UniformReal(0.0, 1.0)
        del self.rot[1:], self.uinc[1:], self.einc[1:], self.hinc[1:]
        for _ in range(1, int(L)):
            theta = math.acos(2.0 * urf(g) - 1.0)
            phi = 2.0 * math.pi * urf(g)
            n = np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])
            psi = 2.0 * math.pi * urf(g)
            R = rotation(psi, n)
            self.rot.append(R)
            self.uinc.append(R @ self.u0)
            self.einc.append(R @ self.e0)
            self.hinc.append(R @ self.h0)
        return self

    def waves(self):
        """The waves as a targets.Waves (k, directions, E and H amplitudes)."""
        return Waves(self.k, np.array(self.uinc), np.array(self.einc), hfields=np.array(self.hinc))
