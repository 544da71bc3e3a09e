"""Writes green_kat.json: interaction blocks from calculus, not from the closed form.

application_function.cpp:36-39 states -G with G = e^{ikr}/r [k^2 (I - uu) -
(1 - ikr)/r^2 (I - 3uu)], and :59-67 the 6x6 block [[-G, -F], [F, -G]] with
F = -k^2 e^{ikr}/r (1 - 1/(ikr)) [u]_x.  Those are the dyadic Green tensor
G = (k^2 I + grad grad) phi and F = i k [grad phi]_x of phi = e^{ik|x|}/|x|.
This script evaluates the second form with mpmath numerical differentiation
at 40 digits, so the fixture pins the closed form (and its signs) from the
definition rather than from a retyped formula.

Run: python tests/golden/make_green_kat.py   (writes tests/golden/green_kat.json)
"""
import json
import os

import mpmath as mp

mp.mp.dps = 40

PROBES = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, -1, 0), (2, 1, -1), (-3, 2, 1),
          (0, -2, 3), (4, 0, -1), (1, 1, 1), (-1, -2, -2), (5, -3, 2), (2, 2, 0), (0, 3, -4)]
CASES = [  # (lf, k)
    (1.0, mp.mpc(1.3, 0.0)),
    (0.25, mp.mpc(2.0, 0.0)),
    (0.7, mp.mpc(1.1, 0.15)),
    (10.0, mp.mpc(0.05, 0.0)),
]


def blocks(d, lf, k):
    x0 = [mp.mpf(v) * lf for v in d]

    def phi(x, y, z):
        r = mp.sqrt(x * x + y * y + z * z)
        return mp.exp(1j * k * r) / r

    p0 = phi(*x0)
    G = [[None] * 3 for _ in range(3)]
    for a in range(3):
        for b in range(3):
            order = [0, 0, 0]
            order[a] += 1
            order[b] += 1
            h = mp.diff(phi, x0, tuple(order))
            G[a][b] = h + (k * k * p0 if a == b else 0)
    grad = []
    for a in range(3):
        order = [0, 0, 0]
        order[a] = 1
        grad.append(mp.diff(phi, x0, tuple(order)))
    gx = [[0, -grad[2], grad[1]], [grad[2], 0, -grad[0]], [-grad[1], grad[0], 0]]
    F = [[1j * k * gx[a][b] for b in range(3)] for a in range(3)]
    M3 = [[-G[a][b] for b in range(3)] for a in range(3)]
    M6 = [[0] * 6 for _ in range(6)]
    for a in range(3):
        for b in range(3):
            M6[a][b] = -G[a][b]
            M6[a][b + 3] = -F[a][b]
            M6[a + 3][b] = F[a][b]
            M6[a + 3][b + 3] = -G[a][b]
    enc = lambda M: [[[float(mp.re(v)), float(mp.im(v))] for v in row] for row in M]
    return enc(M3), enc(M6)


def main():
    out = {"doc": __doc__.strip().splitlines()[0], "cases": []}
    for lf, k in CASES:
        for d in PROBES:
            m3, m6 = blocks(d, lf, k)
            out["cases"].append({"d": list(d), "lf": lf, "k": [float(k.real), float(k.imag)],
                                 "f3": m3, "f6": m6})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "green_kat.json")
    with open(path, "w") as fh:
        json.dump(out, fh)
    print("wrote", path, len(out["cases"]), "cases")


if __name__ == "__main__":
    main()
