"""Extract the NASA-9 (CEA format) thermodynamic and transport entries of the
reference's cold-flow sample no28 (species H2, N2, O2) from the reference's
data files store/therm_lib/NASA/core/thermo.inp and trans.inp, plus the
sample's chem.inp and control_chem.inp, into small fixtures in the same
formats.  The data are NASA Glenn's published CEA database (McBride, Zehe &
Gordon, NASA/TP-2002-211556); the files are read as text only.

    python tests/golden/make_nasa9_fixture.py [/root/reference]
"""
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
SPECIES = ("H2", "N2", "O2")


def thermo():
    lines = open(os.path.join(REF, "store/therm_lib/NASA/core/thermo.inp")).read().splitlines()
    k = 0
    while lines[k].startswith("!"):
        k += 1
    out = ["! NASA-9 entries of H2, N2, O2 (CEA thermo.inp format)", lines[k], lines[k + 1]]
    k += 2
    while not lines[k].startswith("END"):
        name = lines[k][:18].strip()
        nsec = int(lines[k + 1][:2])
        entry = lines[k:k + 2 + 3 * nsec]
        if name in SPECIES:
            out += entry
        k += 2 + 3 * nsec
    out.append("END PRODUCTS")
    out.append("END REACTANTS")
    return out


def trans():
    lines = open(os.path.join(REF, "store/therm_lib/NASA/core/trans.inp")).read().splitlines()
    out = [lines[0]]
    k = 1
    while not lines[k][:3].upper() == "END":
        a, b = lines[k][:15].strip(), lines[k][16:31].strip()
        iv, ic = int(lines[k][35]), int(lines[k][37])
        entry = lines[k:k + 1 + iv + ic]
        if a in SPECIES and (b == "" or b in SPECIES):
            out += entry  # pure species and their binary pairs (which the reader skips)
        k += 1 + iv + ic
    out.append("end")
    return out


def main():
    with open(os.path.join(HERE, "nasa9_thermo.inp"), "w") as f:
        f.write("\n".join(thermo()) + "\n")
    with open(os.path.join(HERE, "nasa9_trans.inp"), "w") as f:
        f.write("\n".join(trans()) + "\n")
    s = os.path.join(REF, "sample/no28.cold_flow_NASA")
    shutil.copyfile(os.path.join(s, "chem.inp"), os.path.join(HERE, "no28_chem.inp"))
    shutil.copyfile(os.path.join(s, "control_chem.inp"), os.path.join(HERE, "no28_control_chem.inp"))


if __name__ == "__main__":
    main()
