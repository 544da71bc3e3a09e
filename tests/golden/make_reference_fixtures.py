"""Writes the reference's own input and output data as test fixtures.

Inputs (copied unchanged into tests/golden/reference_data/):
  * aggregate target files DDA_target_0_agg_N{8..16384}_kf1.0_Df2.8.out and
    DDA_target_2_agg_N{8..512}_kf1.0_Df2.8.out (format
    Target_manager.hpp:33-47, reader Target_manager.cpp:114-202);
  * the material tables material_bc.dat and material_magnetite.dat.
Outputs (tests/golden/reference_results.json): every result section of the
reference's published result files DDA_result_BC_rpp20nm.dat,
DDA_result_MAGNETITE_rpp40nm.dat, DDA_result.dat and result.dat, as written
by Result_manager's operator<< (Result_manager.cpp:178-245) and the MAC_mie
line after it (block_DDA_main.cpp:174-181): key/value pairs, complex values
as [re, im], with the file and the line where each section starts.

Run in a container where the reference tree is mounted:
    python tests/golden/make_reference_fixtures.py [/root/reference]
"""
import json
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "reference_data")
RESULT_FILES = ["DDA_result_BC_rpp20nm.dat", "DDA_result_MAGNETITE_rpp40nm.dat", "DDA_result.dat", "result.dat"]
TARGETS = (["DDA_target_0_agg_N%d_kf1.0_Df2.8.out" % n for n in (8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096,
                                                                8192, 16384)]
           + ["DDA_target_2_agg_N%d_kf1.0_Df2.8.out" % n for n in (8, 16, 32, 64, 128, 256, 512)])
MATERIALS = ["material_bc.dat", "material_magnetite.dat"]
_CPLX = re.compile(r"^\(\s*([-+0-9.eE]+)\s*,\s*([-+0-9.eE]+)\s*\)$")


def _value(s):
    m = _CPLX.match(s)
    if m:
        return [float(m.group(1)), float(m.group(2))]
    try:
        return float(s) if any(c in s for c in ".eE") else int(s)
    except ValueError:
        return s


def parse_results(path):
    sections, cur = [], None
    with open(path) as fh:
        for lineno, line in enumerate(fh, 1):
            line = line.rstrip("\n")
            if "Start of result section" in line:
                cur = {"file": os.path.basename(path), "line": lineno}
                continue
            if cur is None:
                continue
            if "End of result section" in line:
                cur["_closed"] = True
                continue
            parts = line.split("\t")
            if len(parts) == 2 and parts[0] and not parts[0].startswith("-"):
                cur[parts[0]] = _value(parts[1].strip())
                if parts[0] == "MAC_mie" and cur.get("_closed"):
                    del cur["_closed"]
                    sections.append(cur)
                    cur = None
    return sections


def main(ref):
    os.makedirs(OUT_DIR, exist_ok=True)
    for name in TARGETS + MATERIALS:
        shutil.copyfile(os.path.join(ref, name), os.path.join(OUT_DIR, name))
    out = {"source": "result sections of the reference's published result files", "sections": []}
    for name in RESULT_FILES:
        out["sections"] += parse_results(os.path.join(ref, name))
    with open(os.path.join(HERE, "reference_results.json"), "w") as fh:
        json.dump(out, fh, indent=0)
    print("wrote %d sections, %d data files" % (len(out["sections"]), len(TARGETS) + len(MATERIALS)))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
