"""Airfoil shape-sensitivity case: a NACA 4-digit symmetric section on an
algebraic C-mesh, with Hicks-Henne wall bumps as design variables.

The reference ships the NACA0012 sample (sample/no03.n0012: ``n0012.x``,
``naca0012.dat``) but no adjoint; this module builds the synthetic inviscid
case the shape-gradient workload is defined on:

* C-mesh of ``ni x nj`` cells.  i runs along the C: lower wake (from the far
  field to the trailing edge), lower surface, leading edge, upper surface,
  upper wake; j runs from the wall/wake line outwards.  The wake line is a
  cut of the block onto itself (j-min, order -1, the ``cut_copro`` j-j row).
* BCs: curved slip wall on the airfoil (j-min), free stream on the outer
  boundary (j-max), zero gradient on the two downstream ends (i-min, i-max).
* Design variables: ``nb`` Hicks-Henne bumps per surface,
  b_k(x) = sin(pi x^m_k)^t with m_k = ln(1/2)/ln(x_k), added to the wall
  ordinate (outwards) and carried into the field with a decay along j, so
  the node coordinates are an affine function X(p) = X0 + Phi p and dX/dp =
  Phi exactly.
* Objectives: lift and drag coefficients from the pressure (momentum) flux
  through the wall faces, C = sum_faces ds (TG_u a_x + TG_v a_y) with
  a = -e / (q_inf c) for e the lift or drag direction.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import cases as C


def naca4_thickness(x, t=0.12):
    """Half thickness of a NACA 4-digit section (closed trailing edge)."""
    return 5.0 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 - 0.1036 * x ** 4)


def hicks_henne(x, nb=32, width=3.0):
    """(nb, len(x)) bump functions with peaks spread over (0, 1)."""
    xk = 0.5 * (1.0 - np.cos(np.pi * (np.arange(nb) + 1.0) / (nb + 1.0)))
    m = np.log(0.5) / np.log(xk)
    xc = np.clip(x, 0.0, 1.0)
    return np.sin(np.pi * xc[None, :] ** m[:, None]) ** width


@dataclasses.dataclass
class AirfoilCase:
    case: C.Case
    ni: int
    nj: int
    nwake: int
    mach: float
    alpha: float            # radians
    rho_inf: float
    p_inf: float
    kappa: float
    rgas: float
    phi: np.ndarray         # (npar, 2, ni+1, nj+1): dX/dp (x and y of every node)
    x0: np.ndarray          # (2, ni+1, nj+1) baseline nodes
    npar: int

    @property
    def v_inf(self):
        return self.mach * np.sqrt(self.kappa * self.p_inf / self.rho_inf)

    def directions(self):
        """Lift and drag unit vectors for the current inflow angle."""
        a = self.alpha
        return np.array([-np.sin(a), np.cos(a)]), np.array([np.cos(a), np.sin(a)])

    def wall_range(self):
        """First and last i of the airfoil wall faces (1-based cells)."""
        return self.nwake + 1, self.ni - self.nwake


def _wall_line(ni, nwake, wake_len, t):
    """Nodes of the j=0 line: lower wake, lower surface, upper surface, upper wake."""
    nsurf = (ni - 2 * nwake) // 2
    # cosine clustering at both ends of each surface
    s = 0.5 * (1.0 - np.cos(np.pi * np.arange(nsurf + 1) / nsurf))
    xs_lower = 1.0 - s                 # TE -> LE
    xs_upper = s                       # LE -> TE
    # wake: geometric growth away from the trailing edge
    dx0 = xs_upper[-1] - xs_upper[-2]
    r = 1.0
    for _ in range(200):               # growth factor for the wake length
        total = dx0 * (r ** np.arange(1, nwake + 1)).sum()
        r *= (wake_len / total) ** (1.0 / (2 * nwake))
    xw = 1.0 + np.concatenate([[0.0], np.cumsum(dx0 * r ** np.arange(1, nwake + 1))])
    x = np.concatenate([xw[::-1], xs_lower[1:], xs_upper[1:], xw[1:]])
    sgn = np.concatenate([np.zeros(nwake), -np.ones(nsurf), np.ones(nsurf + 1), np.zeros(nwake)])
    # nodes on the surface: sign -1 lower, +1 upper (the LE node, counted in
    # the upper list, has zero thickness)
    y = sgn * naca4_thickness(np.clip(x, 0.0, 1.0), t)
    y[: nwake + 1] = 0.0
    y[-(nwake + 1):] = 0.0
    return x, y, sgn


def naca_cmesh(ni=512, nj=128, nwake=64, mach=0.8, alpha_deg=1.25, t=0.12, radius=15.0, wake_len=15.0,
               first_height=2e-3, nb=32, amplitudes=None, cfl=5.0, order=2, rho_inf=1.4, p_inf=1.0,
               kappa=1.4, params=None, reynolds=None) -> AirfoilCase:
    """The NACA 4-digit (thickness ``t``) airfoil case: inviscid with a
    curved slip wall, or laminar with a no-slip wall at the chord Reynolds
    number ``reynolds`` (nu = V_inf * chord / reynolds).

    ``amplitudes`` (2*nb values, upper then lower surface bumps, in chords)
    deform the wall; the case starts from the free stream.  LU-SGS with local
    time steps unless ``params`` (a dict of shcfd_params fields) says
    otherwise; ``order`` 1: first-order face states, 2: MUSCL with the
    reference limiter and kappa=-1, 3: kappa=1/3."""
    assert (ni - 2 * nwake) % 2 == 0
    x, y, sgn = _wall_line(ni, nwake, wake_len, t)
    npt = ni + 1
    # outer boundary: lines y = -/+R along the wake, a half circle of
    # radius R around the trailing edge for the airfoil part
    xo = np.empty(npt)
    yo = np.empty(npt)
    xo[: nwake + 1] = x[: nwake + 1]
    yo[: nwake + 1] = -radius
    xo[-(nwake + 1):] = x[-(nwake + 1):]
    yo[-(nwake + 1):] = radius
    i0, i1 = nwake, ni - nwake
    # arc parameter by arclength along the surface
    seg = np.hypot(np.diff(x[i0:i1 + 1]), np.diff(y[i0:i1 + 1]))
    sarc = np.concatenate([[0.0], np.cumsum(seg)]) / seg.sum()
    th = -0.5 * np.pi - np.pi * sarc
    xo[i0:i1 + 1] = 1.0 + radius * np.cos(th)
    yo[i0:i1 + 1] = radius * np.sin(th)
    # geometric stretching in j with the given first height
    dist = radius
    r = 1.1
    for _ in range(200):
        total = first_height * (r ** nj - 1.0) / (r - 1.0)
        r *= (dist / total) ** (1.0 / (4 * nj))
    f = np.concatenate([[0.0], np.cumsum(first_height * r ** np.arange(nj))])
    f = f / f[-1]
    X0 = np.empty((2, npt, nj + 1))
    X0[0] = x[:, None] + f[None, :] * (xo - x)[:, None]
    X0[1] = y[:, None] + f[None, :] * (yo - y)[:, None]
    # Hicks-Henne modes: wall ordinate moves outwards, decaying into the field
    hh = hicks_henne(x, nb)
    decay = (1.0 - f) ** 2
    npar = 2 * nb
    phi = np.zeros((npar, 2, npt, nj + 1))
    upper = sgn > 0
    lower = sgn < 0
    for k in range(nb):
        phi[k, 1] = (hh[k] * upper)[:, None] * decay[None, :]
        phi[nb + k, 1] = (-hh[k] * lower)[:, None] * decay[None, :]
    p = np.zeros(npar) if amplitudes is None else np.asarray(amplitudes, dtype=np.float64)
    Xn = X0 + np.tensordot(p, phi, axes=(0, 0))
    alpha = np.deg2rad(alpha_deg)
    rgas = p_inf / rho_inf  # T_inf = 1
    c_inf = np.sqrt(kappa * p_inf / rho_inf)
    u_inf, v_inf = mach * c_inf * np.cos(alpha), mach * c_inf * np.sin(alpha)
    qc = C.cons_from_prim(np.full((nj, ni), rho_inf), np.full((nj, ni), u_inf), np.full((nj, ni), v_inf),
                          np.full((nj, ni), p_inf), kappa)
    q = [C._fill_q(ni, nj, qc)]
    nu = 0.0 if reynolds is None else mach * c_inf / float(reynolds)
    pool = np.array(C.fixed_state_pool(rho_inf, u_inf, v_inf, p_inf, kappa, rgas, mu=rho_inf * nu), dtype=np.float64)
    rows = [
        [1, C.SIDE_IMIN, 1, nj, C.BC_EXTRAP, 0],
        [1, C.SIDE_IMAX, 1, nj, C.BC_EXTRAP, 0],
        [1, C.SIDE_JMIN, nwake + 1, ni - nwake, C.BC_SLIP_NORMAL if reynolds is None else C.BC_NOSLIP, 0],
        [1, C.SIDE_JMAX, 1, ni, C.BC_FIXED, 0],
    ]
    # wake: j-min cells i=1..nwake exchange with i=ni..ni-nwake+1
    cut = [[4, nwake, -1, 1, 1, 1, -1, 1, 1, ni, -1]]
    kw = dict(scheme=C.SCHEME_LUSGS, time_step=1, high_order=0 if order == 1 else 1, limiter=C.LIMITER_REFERENCE,
              cfl=cfl, k_mus=-1.0 if order == 2 else 1.0 / 3.0, kappa=kappa, rgas=rgas, nu=nu,
              viscous=0 if reynolds is None else 1,
              res_rate_warning=1e300, res_rate_error=1e300)
    kw.update(params or {})
    params = C.make_params(**kw)
    case = C.Case("naca_cmesh", [ni], [nj], 1, params, [Xn[0].copy()], [Xn[1].copy()], q,
                  np.array(rows, dtype=np.int32), pool, np.array(cut, dtype=np.int32),
                  description=f"NACA 4-digit t={t:.2f} C-mesh {ni}x{nj}, M={mach}, alpha={alpha_deg} deg, "
                              + ("inviscid" if reynolds is None else f"laminar Re={reynolds:g}"))
    return AirfoilCase(case, ni, nj, nwake, mach, alpha, rho_inf, p_inf, kappa, rgas, phi, X0, npar)
