"""Optical constants of the target materials, as the reference's Material
(Material.cpp:11-124) provides them to the solve path's inputs.

  * ``Material.vacuum()`` -- eper = mper = 1, density 0 (Material.cpp:14-21).
  * ``Material.constant(eper, mper, density)`` -- fixed values
    (Material.cpp:22-29).
  * ``Material.from_table(path, density, name)`` -- a table of free-space
    wavelength (um), relative permittivity and permeability, one line per
    wavelength as ``wl0 (re,im) (re,im)``; lines whose first token is ``##``
    are comments (Material.cpp:60-86).  Values between rows are linear in
    wl0 on the first bracketing row pair (Material.cpp:88-124); outside the
    table the reference exits, this raises ValueError.
  * ``Material.named(name, table_dir)`` -- the reference's named materials
    with their densities in g/cm^3: BC 1.8 (material_bc.dat), MAGNETITE
    5.17, WATER 1.0, ICE 0.917 (Material.cpp:30-57).
"""
import os
import re

_PAIR = re.compile(r"\(\s*([-+0-9.eE]+)\s*,\s*([-+0-9.eE]+)\s*\)")

NAMED = {
    "BC": ("material_bc.dat", 1.8),
    "MAGNETITE": ("material_magnetite.dat", 5.17),
    "WATER": ("material_water.dat", 1.0),
    "ICE": ("material_ice.dat", 0.917),
}


class Material:
    def __init__(self, name, density, eper=1.0, mper=1.0, table=None):
        self.name = name
        self.density = float(density)
        self.const_eper = complex(eper)
        self.const_mper = complex(mper)
        self.table = table          # (wl0 list, eper list, mper list) or None

    @classmethod
    def vacuum(cls):
        return cls("VACUUM", 0.0)

    @classmethod
    def constant(cls, eper, mper=1.0, density=1.0):
        return cls("CONSTANT", density, eper, mper)

    @classmethod
    def from_table(cls, path, density, name=None):
        wl, ep, mp = [], [], []
        with open(path) as fh:
            for line in fh:
                tok = line.split()
                if not tok or tok[0] == "##":
                    continue
                pairs = _PAIR.findall(line)
                if len(pairs) < 2:
                    raise ValueError("wrong data format in %s: %r" % (path, line))
                wl.append(float(tok[0]))
                ep.append(complex(float(pairs[0][0]), float(pairs[0][1])))
                mp.append(complex(float(pairs[1][0]), float(pairs[1][1])))
        return cls(name or os.path.basename(path), density, table=(wl, ep, mp))

    @classmethod
    def named(cls, name, table_dir):
        fname, density = NAMED[name]
        return cls.from_table(os.path.join(table_dir, fname), density, name)

    def _interp(self, wl0, col):
        wl = self.table[0]
        vals = self.table[col]
        if not wl or wl0 < wl[0] or wl0 > wl[-1]:
            raise ValueError("%s is unknown at wavelength %g" % (self.name, wl0))
        for i in range(len(wl) - 1):
            if wl[i] <= wl0 <= wl[i + 1]:
                return vals[i] + (vals[i + 1] - vals[i]) / (wl[i + 1] - wl[i]) * (wl0 - wl[i])
        return vals[-1]     # a one-row table at exactly its wavelength

    def eper(self, wl0):
        return self.const_eper if self.table is None else self._interp(float(wl0), 1)

    def mper(self, wl0):
        return self.const_mper if self.table is None else self._interp(float(wl0), 2)
