"""Cold flow with NASA-9 thermodynamics (the reference's thermal model 5,
store/therm_lib/NASA/cold): the ``shcfd_nasa9`` table (include/shcfd_types.h),
parsing of chem.inp / CEA thermo.inp / CEA trans.inp / control_chem.inp
through the library, the static states of the fuel and oxidizer streams
(initialize_cold_flow, store/therm_lib/NASA/core/sub_chem.f90:388-488), and a
fuel-jet mixing case on a flat plate.

The flow carries nY = 2 "species": the mass fractions of the fuel stream and
of the oxidizer stream, whose compositions are frozen (no reaction)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import cases as C

MAX_NS, MAX_SEC, MAX_TSEC = 8, 6, 3
THERMAL_COLD_NASA9 = 5
FUEL, OXIDIZER = 0, 1

_d = ctypes.c_double
_i = ctypes.c_int


class Nasa9(ctypes.Structure):
    """Mirror of ``shcfd_nasa9``."""

    _fields_ = [
        ("ns", _i), ("nt", _i),
        ("nsec", _i * MAX_NS), ("nsec_trans", _i * MAX_NS), ("tr2th", _i * MAX_NS),
        ("mw", _d * MAX_NS),
        ("trange", ((_d * 2) * MAX_SEC) * MAX_NS),
        ("co", ((_d * 9) * MAX_SEC) * MAX_NS),
        ("trange_trans", ((_d * 2) * MAX_TSEC) * MAX_NS),
        ("trans", ((_d * 4) * MAX_TSEC) * MAX_NS),
        ("nf", _d * MAX_NS), ("no", _d * MAX_NS),
    ]

    def array(self, name):
        return np.ctypeslib.as_array(getattr(self, name))


def _lib():
    from .native import lib

    L = lib()
    if not getattr(L, "_nasa9_bound", False):
        dp = ctypes.POINTER(ctypes.c_double)
        L.shcfd_nasa9_parse.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                        ctypes.c_void_p, dp, ctypes.c_char_p, ctypes.c_int]
        L.shcfd_set_nasa9.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.shcfd_nasa9_stream.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp, dp]
        L._nasa9_bound = True
    return L


def parse(chem_inp: str, thermo_inp: str, trans_inp: str | None, control_chem_inp: str):
    """(tables, (po, To, pf, Tf)) from the reference's input files."""
    L = _lib()
    t = Nasa9()
    pt = (ctypes.c_double * 4)()
    err = ctypes.create_string_buffer(512)
    rc = L.shcfd_nasa9_parse(chem_inp.encode(), thermo_inp.encode(), trans_inp.encode() if trans_inp else None,
                             control_chem_inp.encode(), ctypes.byref(t), pt, err, 512)
    if rc != 0:
        raise ValueError(f"shcfd_nasa9_parse: {err.value.decode()}")
    return t, tuple(pt)


def stream_state(t: Nasa9, which: int, p: float, T: float):
    """(q(5), w(10), vhi(2)) of the fuel (0) or oxidizer (1) stream at rest."""
    L = _lib()
    q, w, v = np.zeros(5), np.zeros(10), np.zeros(2)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = L.shcfd_nasa9_stream(ctypes.byref(t), which, p, T, dp(q), dp(w), dp(v))
    if rc != 0:
        raise ValueError("shcfd_nasa9_stream: bad arguments")
    return q, w, v


def mixing_case(t: Nasa9, pt, ni=96, nj=48, u_ox=60.0, u_fuel=30.0, seed=0, scheme=C.SCHEME_LUSGS,
                viscous=1, cfl=0.5, length=0.1, height=0.05) -> C.Case:
    """A fuel stream entering under an oxidizer stream along a flat plate
    (the layout of the reference's sample no28): the lower third of the
    inflow is fuel at (pf, Tf) moving at u_fuel, the rest oxidizer at
    (po, To) moving at u_ox; the interior starts as a smoothed mixture with a
    small random perturbation; outflow and top extrapolated, no-slip plate.
    Every state is a frozen mixture of the two streams (nY = 2)."""
    po, To, pf, Tf = pt
    rng = np.random.default_rng(seed)
    qf, wf, vf = stream_state(t, FUEL, pf, Tf)
    qo, wo, vo = stream_state(t, OXIDIZER, po, To)
    xh, rh = C.rect_nodes(ni, nj, 0.0, length, 0.0, height)
    yc = (np.arange(nj) + 0.5) / nj
    Z = 0.5 * (1.0 - np.tanh((yc - 1.0 / 3.0) * 12.0))          # fuel fraction across the layer
    Z = np.clip(Z[:, None] * np.ones((1, ni)) + 0.02 * rng.standard_normal((nj, ni)), 0.0, 1.0)
    # mix at equal pressure by mass: rho from the ideal mixture of the two streams' specific volumes
    vol = Z / wf[0] + (1.0 - Z) / wo[0]
    rho = 1.0 / vol
    Yf = Z
    u = Z * u_fuel + (1.0 - Z) * u_ox
    e = Z * (qf[4] / wf[0]) + (1.0 - Z) * (qo[4] / wo[0])       # frozen mixing conserves e per kg
    q = np.stack([rho * Yf, rho * (1.0 - Yf), rho * u, 0.0 * rho, rho * (e + 0.5 * u * u)], axis=-1)
    q0 = C._fill_q(ni, nj, q)
    # initial w: p and R of a perfect mixture at the stream temperature (the start of the T iteration)
    w0 = C.ref_zeros(10, ni, nj)
    Rm = Z * wf[8] + (1.0 - Z) * wo[8]
    w0[2:nj + 2, 2:ni + 2, 0] = rho
    w0[2:nj + 2, 2:ni + 2, 3] = rho * Rm * (Z * Tf + (1.0 - Z) * To)
    w0[2:nj + 2, 2:ni + 2, 8] = Rm
    w_f = list(wf); w_f[1] = u_fuel; w_f[7] = wf[7] + 0.5 * u_fuel ** 2
    w_o = list(wo); w_o[1] = u_ox; w_o[7] = wo[7] + 0.5 * u_ox ** 2
    pool = np.array(w_f + list(vf) + w_o + list(vo), dtype=np.float64)
    jf = max(1, nj // 3)
    rows = [[1, C.SIDE_IMIN, 1, jf, C.BC_FIXED, 0], [1, C.SIDE_IMIN, jf + 1, nj, C.BC_FIXED, 12],
            [1, C.SIDE_IMAX, 1, nj, C.BC_EXTRAP, 0], [1, C.SIDE_JMIN, 1, ni, C.BC_NOSLIP if viscous else C.BC_REFLECT, 0],
            [1, C.SIDE_JMAX, 1, ni, C.BC_EXTRAP, 0]]
    params = C.make_params(scheme=scheme, time_step=1 if scheme == C.SCHEME_PRECON_LUSGS else 0, viscous=viscous,
                           thermal=THERMAL_COLD_NASA9, cfl=cfl, k_mus=-1.0, kappa=1.4, rgas=287.0, nu=0.0)
    case = C.Case("cold_mixing", [ni], [nj], 2, params, [xh], [rh], [q0], np.array(rows, dtype=np.int32), pool,
                  np.zeros((0, 11), np.int32), description=f"cold-flow fuel/oxidizer layer {ni}x{nj}",
                  w0=[w0], nasa9=t)
    return case
