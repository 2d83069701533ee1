"""Mechanisms and states shared by the chemistry tests."""
import os

import numpy as np

from shared_cfd_amd import chem

HERE = os.path.dirname(os.path.abspath(__file__))
NO27 = os.path.join(HERE, "golden", "no27_chem.inp")

# NASA-7 record of H2O2 (GRI-Mech 3.0 thermo data) in the 80-column format
_H2O2 = [
    "H2O2                    H   2O   2          G    200.00   3500.00 1000.00      1",
    " 4.16500285E+00 4.90831694E-03-1.90139225E-06 3.71185986E-10-2.87908305E-14    2",
    "-1.78617877E+04 2.91615662E+00 4.27611269E+00-5.42822417E-04 1.67335701E-05    3",
    "-2.15770813E-08 8.62454363E-12-1.77025821E+04 3.43505074E+00                   4",
]
NO27_9 = ["N2", "O2", "H2", "OH", "O", "H", "HO2", "H2O", "H2O2"]


def no27_plus_h2o2(outdir):
    """The sample no27 H2/air mechanism (8 species, 9 reactions) with an
    inert ninth species H2O2, so it runs in the 9-species flow kernels."""
    lines = open(NO27).read().split("\n")
    out = []
    in_thermo = False
    for ln in lines:
        if ln.strip() == "N2 O2 H2 OH O H HO2 H2O":
            ln = ln + " H2O2"
        if ln.strip() == "thermo all":
            in_thermo = True
        if in_thermo and ln.strip() == "end":
            out.extend(_H2O2)
            in_thermo = False
        out.append(ln)
    os.makedirs(outdir, exist_ok=True)
    p = os.path.join(outdir, "chem.inp")
    with open(p, "w") as f:
        f.write("\n".join(out))
    return p


def random_states(m, n, seed, tmin=700.0, tmax=2600.0):
    """n states (concentrations mol/cm^3 then T) with every species present."""
    rng = np.random.default_rng(seed)
    y = 10.0 ** rng.uniform(-10, -5, (n, m.ns))
    T = rng.uniform(tmin, tmax, (n, 1))
    return np.hstack([y, T])


def mixture_states(m, n, seed, p=1.01325e5, tmin=700.0, tmax=2200.0):
    """(vrho kg/m^3, rho, E J/kg, T) of random mixtures at pressure p."""
    from oracle import oracle as O

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        rel = rng.random(m.ns) ** 3 + 1e-6
        T = rng.uniform(tmin, tmax)
        vrho, rho, E = O.chem_rho_rel2abs(m, p, T, rel)
        out.append((vrho, rho, E, T))
    return out


def ignition_states(m, n, seed, p=1.01325e5, tmin=1000.0, tmax=2200.0):
    """(vrho kg/m^3, T) of H2/air mixtures seeded with radicals, for the
    no27 species order (N2, O2, H2, OH, O, H, HO2, H2O) and any species
    after those at small amounts: states whose ignition over 1e-7..1e-5 s
    needs more than 10 CHEMEQ2 attempts."""
    from oracle import oracle as O

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        rel = np.empty(m.ns)
        rel[:8] = [0.70, 0.20, rng.uniform(0.02, 0.3), *(10.0 ** rng.uniform(-6, -3, 4)), 10.0 ** rng.uniform(-4, -1)]
        rel[8:] = 1e-3 * rng.random(m.ns - 8)
        T = rng.uniform(tmin, tmax)
        vrho, rho, E = O.chem_rho_rel2abs(m, p, T, rel)
        out.append((vrho, T))
    return out


def no27_plus_inert(outdir, k):
    """The sample no27 mechanism with k inert species appended, N2-like
    copies named N2A, N2B, ... (same NASA-7 record and element counts), for
    flows and 0-D states with 8 + k species."""
    names = ["N2" + chr(ord("A") + i) for i in range(k)]
    lines = open(NO27).read().split("\n")
    out = []
    in_thermo = False
    n2 = None
    for idx, ln in enumerate(lines):
        if ln.strip() == "N2 O2 H2 OH O H HO2 H2O":
            ln = ln + " " + " ".join(names)
        if ln.strip() == "thermo all":
            in_thermo = True
        if in_thermo and ln.startswith("N2 "):
            n2 = lines[idx:idx + 4]
        if in_thermo and ln.strip() == "end":
            for nm in names:
                out.append(nm.ljust(18) + n2[0][18:])
                out.extend(n2[1:])
            in_thermo = False
        out.append(ln)
    os.makedirs(outdir, exist_ok=True)
    p = os.path.join(outdir, "chem.inp")
    with open(p, "w") as f:
        f.write("\n".join(out))
    return p, ["N2", "O2", "H2", "OH", "O", "H", "HO2", "H2O"] + names
