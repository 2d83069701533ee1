"""Chemkin mechanisms for the reacting solver: the ``shcfd_mech`` table
(include/shcfd_types.h), parsing through the library, host-side boundary
states, the batched 0-D GPU harness, and the synthetic 9-species /
21-reaction H2-O2-N2 mechanism of the reacting mixing-layer workload.

The reference reads ``chem.inp``, ``MW.inp`` and ``trans.inp`` at start-up
(store/therm_lib/chemkin/core/sub_chem.f90:191-598, :788-889); the same
files are parsed here by ``shcfd_mech_parse``.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(HERE, "data")
MW_INP = os.path.join(DATA, "MW.inp")

MAX_NE, MAX_NS, MAX_NR, MAX_TSEC = 8, 16, 48, 5
RXN_NONE, RXN_REV, RXN_LOW, RXN_TROE, RXN_M, RXN_M_REV = range(6)
THERMAL_IDEAL, THERMAL_CHEMKIN = 0, 2
REACTION_NONE, REACTION_CHEMEQ2, REACTION_POINT_IMPLICIT = 0, 1, 2
FEX, FEXVAR, JEX, CALCT, CHEMEQ2, POINT_IMPLICIT, DVODE = range(7)

_d = ctypes.c_double
_i = ctypes.c_int


class Mech(ctypes.Structure):
    """Mirror of ``shcfd_mech``."""

    _fields_ = [
        ("ne", _i), ("ns", _i), ("nr", _i), ("nt", _i),
        ("mw", _d * MAX_NS), ("inv_mw", _d * MAX_NS),
        ("tthre", (_d * 3) * MAX_NS),
        ("coeff", ((_d * 7) * 2) * MAX_NS),
        ("es", (_d * MAX_NE) * MAX_NS),
        ("reverse_off", _i * MAX_NR), ("kind", _i * MAX_NR),
        ("num_nu", (_i * 2) * MAX_NR),
        ("ind_nu", ((_i * 3) * 2) * MAX_NR),
        ("snu", _i * MAX_NR), ("exist_m", _i * MAX_NR), ("num_m", _i * MAX_NR),
        ("ind_m", (_i * MAX_NS) * MAX_NR),
        ("men", (_d * MAX_NS) * MAX_NR),
        ("abe", (_d * 3) * MAX_NR), ("cabe", (_d * 3) * MAX_NR), ("troe", (_d * 4) * MAX_NR),
        ("tr2th", _i * MAX_NS), ("nsec_trans", _i * MAX_NS),
        ("trange", ((_d * 2) * MAX_TSEC) * MAX_NS),
        ("trans", ((_d * 4) * MAX_TSEC) * MAX_NS),
        ("rtol", _d), ("atol", _d),
    ]

    def array(self, name):
        return np.ctypeslib.as_array(getattr(self, name))


def _lib():
    from .native import lib

    L = lib()
    if not getattr(L, "_chem_bound", False):
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.shcfd_mech_parse.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p,
                                       ctypes.c_char_p, ctypes.c_int]
        L.shcfd_set_mechanism.argtypes = [vp, ctypes.c_void_p]
        L.shcfd_proceed_reaction.argtypes = [vp]
        L.shcfd_chem_boundary_state.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, dp, dp, dp, dp]
        L.shcfd_chem_eval.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, dp, dp,
                                      ctypes.POINTER(ctypes.c_int), ctypes.c_double]
        L._chem_bound = True
    return L


def parse(chem_inp: str, trans_inp: str | None = None, mw_inp: str = MW_INP, rtol: float = 1e-6,
          atol: float = 1e-6) -> Mech:
    """read_cheminp + set_trans_data through the library's parser."""
    L = _lib()
    m = Mech()
    err = ctypes.create_string_buffer(256)
    rc = L.shcfd_mech_parse(chem_inp.encode(), mw_inp.encode(), trans_inp.encode() if trans_inp else None,
                            ctypes.byref(m), err, 256)
    if rc != 0:
        raise ValueError(f"shcfd_mech_parse({chem_inp}): {err.value.decode()}")
    m.rtol, m.atol = rtol, atol
    return m


def boundary_state(m: Mech, p: float, T: float, vrho_rel):
    """calc_boundary (thermal_model.f90:326-356): w(dimw), vhi(ns), dhi(ns)."""
    L = _lib()
    ns = m.ns
    vr = np.ascontiguousarray(vrho_rel, dtype=np.float64)
    w = np.zeros(ns + 8)
    vhi = np.zeros(ns)
    dhi = np.zeros(ns)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.shcfd_chem_boundary_state(ctypes.byref(m), p, T, vr.ctypes.data_as(dp), w.ctypes.data_as(dp),
                                     vhi.ctypes.data_as(dp), dhi.ctypes.data_as(dp))
    if rc != 0:
        raise ValueError(f"shcfd_chem_boundary_state failed ({rc})")
    return w, vhi, dhi


def row_sizes(op: int, ns: int):
    n1 = ns + 1
    return {FEX: (n1, n1), FEXVAR: (n1, 2 * n1), JEX: (n1, n1 * n1), CALCT: (ns + 3, 4 + 2 * ns),
            CHEMEQ2: (n1, ns), POINT_IMPLICIT: (n1, n1), DVODE: (n1, ns)}[op]


def chem_eval(m: Mech, op: int, states, dt: float = 0.0):
    """Batched 0-D evaluation on the GPU; returns (out, status)."""
    L = _lib()
    ldi, ldo = row_sizes(op, m.ns)
    x = np.ascontiguousarray(states, dtype=np.float64).reshape(-1, ldi)
    n = x.shape[0]
    out = np.zeros((n, ldo))
    st = np.zeros(n, dtype=np.int32)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.shcfd_chem_eval(ctypes.byref(m), op, n, x.ctypes.data_as(dp), out.ctypes.data_as(dp),
                           st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), dt)
    if rc != 0:
        raise RuntimeError(f"shcfd_chem_eval failed ({rc})")
    return out, st


# ---------------------------------------------------------------------------
# synthetic H2-O2-N2 mechanism (9 species, 21 reversible reactions)
# ---------------------------------------------------------------------------
# Base NASA-7 fits (GRI-Mech 3.0 thermo data; low range 200/300-1000 K, high
# range 1000-3500/5000 K), as (elements, high a1..a7, low a1..a7).
_BASE_THERMO = {
    "N2": ({"N": 2}, [2.92664e0, 1.4879768e-3, -5.68476e-7, 1.0097038e-10, -6.753351e-15, -9.227977e2, 5.980528e0],
           [3.298677e0, 1.4082404e-3, -3.963222e-6, 5.641515e-9, -2.444854e-12, -1.0208999e3, 3.950372e0]),
    "O2": ({"O": 2}, [3.28253784e0, 1.48308754e-3, -7.57966669e-7, 2.09470555e-10, -2.16717794e-14, -1.08845772e3,
                      5.45323129e0],
           [3.78245636e0, -2.99673416e-3, 9.84730201e-6, -9.68129509e-9, 3.24372837e-12, -1.06394356e3,
            3.65767573e0]),
    "H2": ({"H": 2}, [3.3372792e0, -4.94024731e-5, 4.99456778e-7, -1.79566394e-10, 2.00255376e-14, -9.50158922e2,
                      -3.20502331e0],
           [2.34433112e0, 7.98052075e-3, -1.9478151e-5, 2.01572094e-8, -7.37611761e-12, -9.17935173e2,
            6.83010238e-1]),
    "OH": ({"O": 1, "H": 1}, [3.09288767e0, 5.48429716e-4, 1.26505228e-7, -8.79461556e-11, 1.17412376e-14,
                              3.858657e3, 4.4766961e0],
           [3.99201543e0, -2.40131752e-3, 4.61793841e-6, -3.88113333e-9, 1.3641147e-12, 3.61508056e3,
            -1.03925458e-1]),
    "O": ({"O": 1}, [2.56942078e0, -8.59741137e-5, 4.19484589e-8, -1.00177799e-11, 1.22833691e-15, 2.92175791e4,
                     4.78433864e0],
          [3.1682671e0, -3.27931884e-3, 6.64306396e-6, -6.12806624e-9, 2.11265971e-12, 2.91222592e4,
           2.05193346e0]),
    "H": ({"H": 1}, [2.5e0, 0.0, 0.0, 0.0, 0.0, 2.54736599e4, -4.46682914e-1],
          [2.5e0, 0.0, 0.0, 0.0, 0.0, 2.54736599e4, -4.46682853e-1]),
    "HO2": ({"H": 1, "O": 2}, [4.0172109e0, 2.23982013e-3, -6.3365815e-7, 1.1424637e-10, -1.07908535e-14,
                               1.11856713e2, 3.78510215e0],
            [4.30179801e0, -4.74912051e-3, 2.11582891e-5, -2.42763894e-8, 9.29225124e-12, 2.9480804e2,
             3.71666245e0]),
    "H2O": ({"H": 2, "O": 1}, [3.03399249e0, 2.17691804e-3, -1.64072518e-7, -9.7041987e-11, 1.68200992e-14,
                               -3.00042971e4, 4.9667701e0],
            [4.19864056e0, -2.0364341e-3, 6.52040211e-6, -5.48797062e-9, 1.77197817e-12, -3.02937267e4,
             -8.49032208e-1]),
    "H2O2": ({"H": 2, "O": 2}, [4.16500285e0, 4.90831694e-3, -1.90139225e-6, 3.71185986e-10, -2.87908305e-14,
                                -1.78617877e4, 2.91615662e0],
             [4.27611269e0, -5.42822417e-4, 1.67335701e-5, -2.15770813e-8, 8.62454363e-12, -1.77025821e4,
              3.43505074e0]),
}
SPECIES9 = ["N2", "O2", "H2", "OH", "O", "H", "HO2", "H2O", "H2O2"]

# reaction templates: (equation, kind, third-body efficiencies); all
# element-balanced and reversible
_TEMPLATES = [
    ("H+O2=O+OH", RXN_NONE, None),
    ("O+H2=H+OH", RXN_NONE, None),
    ("H2+OH=H2O+H", RXN_NONE, None),
    ("O+H2O=OH+OH", RXN_NONE, None),
    ("H2+M=H+H+M", RXN_M, {"H2": 2.5, "H2O": 12.0}),
    ("O+O+M=O2+M", RXN_M_REV, {"H2": 2.5, "H2O": 12.0}),
    ("O+H+M=OH+M", RXN_M, {"H2": 2.5, "H2O": 12.0}),
    ("H+OH+M=H2O+M", RXN_M, {"H2": 0.73, "H2O": 3.65}),
    ("H+O2(+M)=HO2(+M)", RXN_TROE, {"H2": 1.3, "H2O": 14.0}),
    ("HO2+H=H2+O2", RXN_NONE, None),
    ("HO2+H=OH+OH", RXN_NONE, None),
    ("HO2+O=O2+OH", RXN_REV, None),
    ("HO2+OH=H2O+O2", RXN_NONE, None),
    ("HO2+HO2=H2O2+O2", RXN_NONE, None),
    ("H2O2(+M)=OH+OH(+M)", RXN_TROE, {"H2O": 12.0, "H2": 2.5}),
    ("H2O2+H=H2O+OH", RXN_NONE, None),
    ("H2O2+H=HO2+H2", RXN_NONE, None),
    ("H2O2+O=OH+HO2", RXN_NONE, None),
    ("H2O2+OH=H2O+HO2", RXN_NONE, None),
    ("OH+OH(+M)=H2O2(+M)", RXN_LOW, {"H2O": 6.0}),
    ("H+HO2=O+H2O", RXN_NONE, None),
]

_CAL = 4.184


def _nasa_row(a, idx):
    s = ""
    for v in a:
        s += f"{v:15.8E}"
    return f"{s:<75}{idx:5d}"


def _perturbed_thermo(rng):
    """NASA-7 fits scaled coefficient-wise by U[0.95, 1.05]; the high-range
    a6/a7 are then re-fitted so that h and s stay continuous at 1000 K."""
    out = {}
    tc = 1000.0
    for sp in SPECIES9:
        el, hi, lo = _BASE_THERMO[sp]
        lo = [v * rng.uniform(0.95, 1.05) for v in lo]
        hi = [v * rng.uniform(0.95, 1.05) for v in hi]
        ph_lo = sum(lo[i] * tc ** i / (i + 1) for i in range(5))
        ph_hi = sum(hi[i] * tc ** i / (i + 1) for i in range(5))
        hi[5] = lo[5] + tc * (ph_lo - ph_hi)
        s_lo = lo[0] * math.log(tc) + sum(lo[i] * tc ** i / i for i in range(1, 5))
        s_hi = hi[0] * math.log(tc) + sum(hi[i] * tc ** i / i for i in range(1, 5))
        hi[6] = lo[6] + (s_lo - s_hi)
        out[sp] = (el, hi, lo)
    return out


def _fmt_num(v):
    return f"{v:.4E}"


def write_synthetic_mechanism(seed: int, outdir: str):
    """Write chem.inp and trans.inp of the synthetic 9-species / 21-reaction
    mechanism; returns their paths.  Arrhenius parameters from ``seed``:
    A log-uniform in [1e8, 1e16] (cgs), beta uniform in [-1, 1], Ea uniform
    in [0, 2e5] J/mol written in cal/mol; LOW/REV parameters likewise."""
    rng = np.random.default_rng(seed)
    os.makedirs(outdir, exist_ok=True)
    thermo = _perturbed_thermo(rng)
    lines = ["! synthetic H2/O2/N2 mechanism (9 species, 21 reactions), seed %d" % seed,
             "elements", " H O N", "end", "species", " " + " ".join(SPECIES9), "end", "thermo all",
             "   300.000  1000.000  5000.000"]
    for sp in SPECIES9:
        el, hi, lo = thermo[sp]
        head = f"{sp:<24}"
        for e, cnt in el.items():
            head += f"{e:<2}{cnt:3d}"
        head = f"{head:<44}G{200.0:10.2f}{5000.0:10.2f}{1000.0:8.2f}"
        lines.append(f"{head:<79}1")
        lines.append(_nasa_row(hi[0:5], 2))
        lines.append(_nasa_row(hi[5:7] + lo[0:3], 3))
        lines.append(f"{''.join(f'{v:15.8E}' for v in lo[3:7]):<79}4")
    lines += ["end", "reactions      cal/mole  moles"]

    def arr():
        A = 10.0 ** rng.uniform(8.0, 16.0)
        b = rng.uniform(-1.0, 1.0)
        E = rng.uniform(0.0, 2.0e5) / _CAL
        return A, b, E

    def g_rt(sp, T):
        _, hi, lo = thermo[sp]
        a = lo if T <= 1000.0 else hi
        h = a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T
        s = a[0] * math.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + a[6]
        return h - s

    def rev_fit(eq, A, b, E):
        """Explicit reverse parameters fitted to k_f / K_c at 600, 1200 and
        2400 K, so REV reactions keep (nearly) the mechanism's equilibrium."""
        lhs, rhs = eq.replace("(+M)", "").replace("+M", "").split("=")
        reac, prod = lhs.split("+"), rhs.split("+")
        Ts = [600.0, 1200.0, 2400.0]
        rows, vals = [], []
        for T in Ts:
            lnkc = -(sum(g_rt(s, T) for s in prod) - sum(g_rt(s, T) for s in reac)) + \
                (len(prod) - len(reac)) * math.log(1.01325e6 / (8.3144621e7 * T))
            rows.append([1.0, math.log(T), -1.0 / T])
            vals.append(math.log(A) + b * math.log(T) - E * _CAL / 8.314462 / T - lnkc)
        lnA, br, er = np.linalg.solve(np.array(rows), np.array(vals))
        return math.exp(lnA), br, er * 8.314462 / _CAL

    def h_rt(sp, T):
        _, hi, lo = thermo[sp]
        a = lo if T <= 1000.0 else hi
        return a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T

    def exothermic(eq):
        """Write the template in its exothermic direction (at 1000 K), so the
        random activation energy belongs to the direction that may have
        Ea = 0 and the K_c-derived reverse rate stays bounded."""
        lhs, rhs = eq.split("=")
        strip = lambda s: s.replace("(+M)", "").replace("+M", "").split("+")
        dh = sum(h_rt(s, 1000.0) for s in strip(rhs)) - sum(h_rt(s, 1000.0) for s in strip(lhs))
        return eq if dh <= 0.0 else rhs + "=" + lhs

    for eq, kind, eff in _TEMPLATES:
        eq = exothermic(eq)
        A, b, E = arr()
        lines.append(f" {eq:<30} {A:12.4E} {b:8.4f} {E:12.3f}")
        if kind in (RXN_REV, RXN_M_REV):
            A2, b2, E2 = rev_fit(eq, A, b, E)
            lines.append(f"     REV / {A2:12.4E} {b2:8.4f} {E2:12.3f} /")
        if kind in (RXN_LOW, RXN_TROE):
            A2, b2, E2 = arr()
            lines.append(f"     LOW / {A2:12.4E} {b2:8.4f} {E2:12.3f} /")
        if kind == RXN_TROE:
            a = rng.uniform(0.4, 0.8)
            t3, t1, t2 = rng.uniform(50.0, 500.0), rng.uniform(1000.0, 3000.0), rng.uniform(3000.0, 8000.0)
            lines.append(f"     TROE / {a:.4f} {t3:.2f} {t1:.2f} {t2:.2f} /")
        if eff:
            lines.append("     " + " ".join(f"{k}/{v:g}/" for k, v in eff.items()))
    lines.append("end")
    chem = os.path.join(outdir, "chem.inp")
    with open(chem, "w") as f:
        f.write("\n".join(lines) + "\n")

    # viscosity fits ln(mu[micropoise]) = A ln T + B/T + C/T^2 + D, two ranges
    # continuous at 1000 K, from mu(300 K) and a power law, scaled by seed
    mu300 = {"N2": 178.0, "O2": 207.0, "H2": 89.0, "OH": 180.0, "O": 190.0, "H": 90.0, "HO2": 200.0,
             "H2O": 100.0, "H2O2": 180.0}
    tl = ["transport property coefficients (synthetic)"]
    for sp in SPECIES9:
        e1 = 0.69 * rng.uniform(0.95, 1.05)
        e2 = 0.62 * rng.uniform(0.95, 1.05)
        B1 = rng.uniform(-60.0, 60.0)
        C1 = rng.uniform(-2e3, 2e3)
        m300 = mu300[sp] * rng.uniform(0.95, 1.05)
        D1 = math.log(m300) - (e1 * math.log(300.0) + B1 / 300.0 + C1 / 300.0 ** 2)
        B2, C2 = rng.uniform(-60.0, 60.0), rng.uniform(-2e4, 2e4)
        ln1000 = e1 * math.log(1000.0) + B1 / 1000.0 + C1 / 1e6 + D1
        D2 = ln1000 - (e2 * math.log(1000.0) + B2 / 1000.0 + C2 / 1e6)
        tl.append(f"{sp:<15} {'':15}   V2C0  synthetic")
        tl.append(f" V{200.0:7.1f}  {1000.0:7.1f}  {e1:15.8E}{B1:15.8E}{C1:15.8E}{D1:15.8E}")
        tl.append(f" V{1000.0:7.1f}  {5000.0:7.1f}  {e2:15.8E}{B2:15.8E}{C2:15.8E}{D2:15.8E}")
    tl.append("end")
    trans = os.path.join(outdir, "trans.inp")
    with open(trans, "w") as f:
        f.write("\n".join(tl) + "\n")
    return chem, trans
