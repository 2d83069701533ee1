"""Case set-up for the shcfd solver: grids, boundary/cut tables, initial states.

The reference builds a case from a Plot3D grid, a generated cut table
(``cut_copro000.inp``, store/cond/core/mod_cond.py:86-91) and a user
``condition.f90`` holding ``set_IC``/``set_BC`` (e.g.
sample/no06.step_flow/condition.f90).  Here the same information is data:

* node coordinates ``xh``/``rh`` per block, shape ``(ni+1, nj+1)`` indexed
  ``[i, j]`` (the reference's ``xh(0:ni,0:nj)``),
* BC rows (6 ints: block, side, start, end, kind, pool offset) plus a
  double pool, replacing the user ``set_BC`` loops,
* cut rows (11 ints, the ``cut_copro`` format),
* initial conservative state ``q`` per block in the reference layout.

Arrays in the *reference layout* are numpy arrays of shape
``(nj+4, ni+4, K)`` in C order, i.e. Fortran ``f(1:K,-1:ni+2,-1:nj+2)``;
element ``(k, i, j)`` is ``a[j+1, i+1, k]``.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
from typing import List, Optional

import numpy as np

# ---- mirrors of include/shcfd_types.h ----------------------------------------
SCHEME_EULER, SCHEME_RK2, SCHEME_LUSGS, SCHEME_NR, SCHEME_PRECON, SCHEME_DUAL, SCHEME_PRECON_LUSGS = 0, 1, 2, 3, 4, 5, 6
LIMITER_REFERENCE, LIMITER_MINMOD = 0, 1
(BC_EXTRAP, BC_REFLECT, BC_NOSLIP, BC_FIXED, BC_INFLOW_TUV, BC_OUTFLOW_P, BC_SLIP_NORMAL, BC_INFLOW_MIX_P,
 BC_INJECT_WALL) = range(9)
SIDE_IMIN, SIDE_IMAX, SIDE_JMIN, SIDE_JMAX = range(4)
DIM_PLANAR, DIM_AXISYMMETRIC = 0, 1

FIELDS = [
    "Q", "QP", "QPP", "W", "VHI", "DHI", "WHLI", "WHRI", "WHLJ", "WHRJ",
    "TGI", "TGJ", "TGVI", "TGVJ", "VOL", "DSI", "DSJ", "VNI", "VNJ", "DSCI", "DSCJ",
    "VNCI", "VNCJ", "GEOJACI", "GEOJACJ", "X", "R", "XH", "RH", "DT", "ALPHA", "PHI",
    "PRESCA", "PRE1", "PRE2", "DQDT", "AP", "AM", "BP", "BM", "DQ", "AREA", "GEOJACC",
]
F = {name: idx for idx, name in enumerate(FIELDS)}


def field_ncomp(name: str, ny: int) -> int:
    dimq = ny + 3
    dimw = dimq + 5
    if name in ("Q", "QP", "QPP", "TGI", "TGJ", "TGVI", "TGVJ", "PRE1", "PRE2", "DQDT", "DQ"):
        return dimq
    if name in ("W", "WHLI", "WHRI", "WHLJ", "WHRJ"):
        return dimw
    if name in ("VHI", "DHI"):
        return ny
    if name in ("VNI", "VNJ", "VNCI", "VNCJ"):
        return 2
    if name in ("GEOJACI", "GEOJACJ", "GEOJACC", "PRESCA"):
        return 4
    if name in ("AP", "AM", "BP", "BM"):
        return dimq * dimq
    return 1


class Params(ctypes.Structure):
    """ctypes mirror of ``shcfd_params`` (include/shcfd_types.h)."""

    _fields_ = [
        ("scheme", ctypes.c_int), ("time_step", ctypes.c_int), ("high_order", ctypes.c_int),
        ("limiter", ctypes.c_int), ("viscous", ctypes.c_int), ("thermal", ctypes.c_int),
        ("num_internal_loop", ctypes.c_int), ("reaction", ctypes.c_int),
        ("phi_mus", ctypes.c_double), ("k_mus", ctypes.c_double), ("cfl", ctypes.c_double),
        ("cfl_tau", ctypes.c_double), ("omega_max", ctypes.c_double), ("omega_min", ctypes.c_double),
        ("dqmax", ctypes.c_double), ("res_rate_warning", ctypes.c_double),
        ("res_rate_error", ctypes.c_double), ("kappa", ctypes.c_double), ("rgas", ctypes.c_double),
        ("nu", ctypes.c_double), ("dimension", ctypes.c_int),
    ]


def make_params(**kw) -> Params:
    d = dict(scheme=SCHEME_LUSGS, time_step=0, high_order=1, limiter=LIMITER_REFERENCE, viscous=0,
             thermal=0, num_internal_loop=21, reaction=0, phi_mus=1.0, k_mus=-1.0, cfl=0.5, cfl_tau=1e-2,
             omega_max=0.1, omega_min=1e-3, dqmax=1e-2, res_rate_warning=1e-5, res_rate_error=1e-3,
             kappa=1.4, rgas=287.0, nu=1.6e-5, dimension=DIM_PLANAR)
    d.update(kw)
    p = Params()
    for k, v in d.items():
        setattr(p, k, v)
    return p


def ref_zeros(k: int, ni: int, nj: int) -> np.ndarray:
    return np.zeros((nj + 4, ni + 4, k), dtype=np.float64)


@dataclasses.dataclass
class Case:
    name: str
    ni: List[int]
    nj: List[int]
    ny: int
    params: Params
    xh: List[np.ndarray]
    rh: List[np.ndarray]
    q0: List[np.ndarray]
    bc_rows: np.ndarray
    bc_pool: np.ndarray
    cut_rows: np.ndarray
    description: str = ""
    meta: dict = dataclasses.field(default_factory=dict)
    mech: Optional[object] = None           # chem.Mech for the Chemkin thermal model
    w0: Optional[List[np.ndarray]] = None   # initial w per block (T start of calc_T)
    block_nu: Optional[List[float]] = None  # per-block kinematic viscosity (batches)
    nasa9: Optional[object] = None          # nasa9.Nasa9 tables of the cold-flow thermal model

    @property
    def nblocks(self) -> int:
        return len(self.ni)

    @property
    def ncells(self) -> int:
        return int(sum(a * b for a, b in zip(self.ni, self.nj)))

    @property
    def dimq(self) -> int:
        return self.ny + 3

    @property
    def dimw(self) -> int:
        return self.ny + 8

    def nodes_ref(self, b: int):
        """xh/rh of block b as reference-layout arrays (K=1)."""
        out = []
        for arr in (self.xh[b], self.rh[b]):
            a = ref_zeros(1, self.ni[b], self.nj[b])
            a[1:self.nj[b] + 2, 1:self.ni[b] + 2, 0] = arr.T
            out.append(a)
        return out


# ---- helpers --------------------------------------------------------------------
def cons_from_prim(rho, u, v, p, kappa):
    """q = (rho, rho u, rho v, rho E) for a single-species ideal gas."""
    return np.stack([rho, rho * u, rho * v, p / (kappa - 1.0) + 0.5 * rho * (u * u + v * v)], axis=-1)


def _fill_q(ni, nj, qcell):
    """qcell: (nj, ni, dimq) interior values -> reference-layout q."""
    q = ref_zeros(qcell.shape[-1], ni, nj)
    q[2:nj + 2, 2:ni + 2, :] = qcell
    return q


def rect_nodes(ni, nj, x0, x1, y0, y1):
    xs = np.linspace(x0, x1, ni + 1)
    ys = np.linspace(y0, y1, nj + 1)
    xh = np.repeat(xs[:, None], nj + 1, axis=1)
    rh = np.repeat(ys[None, :], ni + 1, axis=0)
    return xh.copy(), rh.copy()


def fixed_state_pool(rho, u, v, p, kappa, rgas, mu=0.0, ny=1):
    """w(1:dimw) + vhi(1:nY) for a single-species ideal-gas FIXED BC."""
    T = p / (rho * rgas)
    H = kappa / (kappa - 1.0) * rgas * T + 0.5 * (u * u + v * v)
    w = [rho, u, v, p] + [1.0] * ny + [kappa, H, rgas, mu]
    vhi = [kappa / (kappa - 1.0) * rgas * T] * ny
    return w + vhi


# ---- C1: Sod shock tube ---------------------------------------------------------
def sod(ni=400, nj=4, limiter=LIMITER_REFERENCE, scheme=SCHEME_LUSGS, cfl=0.5, order=2,
        high_order=1) -> Case:
    """Sod tube on x in [0,1], uniform dx, nj cells of height dx; zero-gradient
    BCs on all sides as in sample/no00.shock_tube/condition.f90."""
    kappa, rgas = 1.4, 287.0
    dx = 1.0 / ni
    xh, rh = rect_nodes(ni, nj, 0.0, 1.0, 0.0, nj * dx)
    xc = (np.arange(ni) + 0.5) * dx
    left = xc < 0.5
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1e5, 1e4)
    z = np.zeros_like(rho)
    qrow = cons_from_prim(rho, z, z, p, kappa)
    qcell = np.repeat(qrow[None, :, :], nj, axis=0)
    rows = []
    for side, n in ((SIDE_IMIN, nj), (SIDE_IMAX, nj), (SIDE_JMIN, ni), (SIDE_JMAX, ni)):
        rows.append([1, side, 1, n, BC_EXTRAP, 0])
    params = make_params(scheme=scheme, time_step=0, high_order=high_order, limiter=limiter, cfl=cfl,
                         k_mus=-1.0 if order == 2 else 1.0 / 3.0, kappa=kappa, rgas=rgas, nu=1.6e-5,
                         num_internal_loop=21, cfl_tau=5.0, omega_max=1.0, dqmax=0.5,
                         res_rate_error=1e300, res_rate_warning=1e300)
    return Case("sod", [ni], [nj], 1, params, [xh], [rh], [_fill_q(ni, nj, qcell)],
                np.array(rows, dtype=np.int32), np.zeros(1), np.zeros((0, 11), dtype=np.int32),
                description=f"Sod shock tube {ni}x{nj}")


# ---- C2: Mach-3 forward-facing step (2 blocks) ------------------------------------
def forward_step(n_per_unit=320, scheme=SCHEME_DUAL, num_internal_loop=21, cfl=0.5, cfl_tau=5.0,
                 omega_max=1.0, dqmax=0.5, k_mus=1.0 / 3.0, res_rate_warning=1e-5, res_rate_error=1e-3) -> Case:
    """Woodward-Colella step: domain [0,3]x[0,1], step x>=0.6, y<0.2.

    Block 1: x in [0,0.6], y in [0,1]  -> (0.6 h^-1) x (h^-1) cells.
    Block 2: x in [0.6,3], y in [0.2,1] -> (2.4 h^-1) x (0.8 h^-1) cells.
    One i-i cut joins block 1's i-max side above the step to block 2's i-min
    side (the topology of sample/no06.step_flow/forward_step.x).  Inflow
    rho=1.4, p=1, u=3 (gamma=1.4, so c=1 and M=3); slip walls; zero-gradient
    outflow; the BC kinds of sample/no06.step_flow/condition.f90.

    Inner-loop controls: cfl_tau 5, omega_max 1, Dqmax 0.5 converge each
    physical step's inner loop (res/res1 ~ 1e-13 after 20 inner iterations);
    with the template values of store/time_scheme/dual/2d/control.part.inp
    (1e-2, 1e-1, 1e-2) res/res1 stays near 0.9 and the reference's own
    ResRateErr 1e-3 check, kept here, would stop the run at step 2."""
    h = 1.0 / n_per_unit
    n1i, n1j = int(round(0.6 * n_per_unit)), n_per_unit
    n2i, n2j = int(round(2.4 * n_per_unit)), int(round(0.8 * n_per_unit))
    nstep = n1j - n2j
    kappa, rgas = 1.4, 1.0
    rho0, u0, p0 = 1.4, 3.0, 1.0
    xh1, rh1 = rect_nodes(n1i, n1j, 0.0, 0.6, 0.0, 1.0)
    xh2, rh2 = rect_nodes(n2i, n2j, 0.6, 3.0, 0.2, 1.0)
    q = []
    for ni, nj in ((n1i, n1j), (n2i, n2j)):
        qc = cons_from_prim(np.full((nj, ni), rho0), np.full((nj, ni), u0), np.zeros((nj, ni)),
                            np.full((nj, ni), p0), kappa)
        q.append(_fill_q(ni, nj, qc))
    pool = np.array(fixed_state_pool(rho0, u0, 0.0, p0, kappa, rgas, mu=0.0), dtype=np.float64)
    rows = [
        [1, SIDE_IMIN, 1, n1j, BC_FIXED, 0],
        [1, SIDE_IMAX, 1, nstep, BC_REFLECT, 0],
        [1, SIDE_JMIN, 1, n1i, BC_REFLECT, 0],
        [1, SIDE_JMAX, 1, n1i, BC_REFLECT, 0],
        [2, SIDE_IMAX, 1, n2j, BC_EXTRAP, 0],
        [2, SIDE_JMIN, 1, n2i, BC_REFLECT, 0],
        [2, SIDE_JMAX, 1, n2i, BC_REFLECT, 0],
    ]
    cut = [[1, n2j, 1, 1, n1i, nstep + 1, 1, 2, 1, 1, -1]]
    params = make_params(scheme=scheme, time_step=0, high_order=1, limiter=LIMITER_REFERENCE, cfl=cfl,
                         k_mus=k_mus, kappa=kappa, rgas=rgas, nu=1.6e-5, num_internal_loop=num_internal_loop,
                         cfl_tau=cfl_tau, omega_max=omega_max, omega_min=1e-3, dqmax=dqmax,
                         res_rate_warning=res_rate_warning, res_rate_error=res_rate_error)
    return Case("forward_step", [n1i, n2i], [n1j, n2j], 1, params, [xh1, xh2], [rh1, rh2], q,
                np.array(rows, dtype=np.int32), pool, np.array(cut, dtype=np.int32),
                description=f"Mach-3 forward step, h=1/{n_per_unit}, {n1i}x{n1j}+{n2i}x{n2j} cells")


# ---- files of the Fortran host driver ---------------------------------------------
def write_case(case: "Case", path: str, max_step: int, out_period: int = 1, min_rms: float = 1e-14,
               fast: bool = True, mech_files=None) -> None:
    """case.inp for shared_cfd_amd/main_shcfd (the Fortran main program).
    ``mech_files`` = (chem.inp, MW.inp, trans.inp or None) for the Chemkin
    thermal model, (chem.inp, thermo.inp, trans.inp or None,
    control_chem.inp) for the cold-flow model."""
    p = case.params
    with open(path, "w") as f:
        f.write("SHCFD-CASE\n")
        f.write(f"{case.nblocks} {case.ny}\n")
        for a, b in zip(case.ni, case.nj):
            f.write(f"{a} {b}\n")
        f.write(f"{p.scheme} {p.time_step} {p.high_order} {p.limiter} {p.viscous} {p.thermal} "
                f"{p.num_internal_loop} {int(fast)}\n")
        vals = [p.phi_mus, p.k_mus, p.cfl, p.cfl_tau, p.omega_max, p.omega_min, p.dqmax,
                min(p.res_rate_warning, 1e300), min(p.res_rate_error, 1e300), p.kappa, p.rgas, p.nu]
        f.write(" ".join(repr(float(v)) for v in vals) + "\n")
        f.write(f"{p.reaction} {int(case.w0 is not None)} {p.dimension}\n")
        if p.thermal == 2:
            if mech_files is None:
                raise ValueError("write_case: the Chemkin thermal model needs mech_files")
            chem_inp, mw_inp, trans_inp = mech_files
            for fn in (chem_inp, mw_inp, trans_inp or "-"):
                f.write(os.path.abspath(fn) if fn != "-" else "-")
                f.write("\n")
            f.write(f"{case.mech.rtol!r} {case.mech.atol!r}\n")
        elif p.thermal == 5:
            if mech_files is None:
                raise ValueError("write_case: the cold-flow model needs mech_files")
            for fn in mech_files:  # chem.inp, thermo.inp, trans.inp or None, control_chem.inp
                f.write((os.path.abspath(fn) if fn else "-") + "\n")
        f.write(f"{float(max_step)!r} {min_rms!r} {out_period}\n")
        f.write(f"{case.bc_rows.shape[0]} {case.bc_pool.size}\n")
        for r in case.bc_rows:
            f.write(" ".join(str(int(x)) for x in r) + "\n")
        if case.bc_pool.size:
            f.write(" ".join(repr(float(x)) for x in case.bc_pool) + "\n")
        f.write(f"{case.cut_rows.shape[0]}\n")
        for r in case.cut_rows:
            f.write(" ".join(str(int(x)) for x in r) + "\n")
        for b in range(case.nblocks):
            ni, nj = case.ni[b], case.nj[b]
            for arr in (case.xh[b], case.rh[b]):
                f.write(" ".join(repr(float(x)) for x in arr.T.ravel()) + "\n")  # i fastest
            qi = case.q0[b][2:nj + 2, 2:ni + 2, :]
            f.write(" ".join(repr(float(x)) for x in qi.ravel()) + "\n")  # component fastest
            if case.w0 is not None:
                wi = case.w0[b][2:nj + 2, 2:ni + 2, :]
                f.write(" ".join(repr(float(x)) for x in wi.ravel()) + "\n")


def read_result_bin(path: str, case: "Case"):
    """Read result.bin / restart.bin in the out_bin layout of
    store/core/inout.f90:44-56 (big-endian Fortran sequential records):
    step, tt, then (q(1:dimq), R, gamma, p) for i=0..ni+1, j=0..nj+1 of
    every plane.  Returns (step, tt, [array (nj+2, ni+2, dimq+3)])."""
    import struct

    with open(path, "rb") as f:
        data = f.read()
    recs, pos = [], 0
    while pos < len(data):
        (n,) = struct.unpack(">i", data[pos:pos + 4])
        recs.append(data[pos + 4:pos + 4 + n])
        pos += 8 + n
    step = struct.unpack(">i", recs[0])[0]
    tt = struct.unpack(">d", recs[1])[0]
    vals = np.frombuffer(recs[2], dtype=">f8").astype(np.float64)
    out, k = [], 0
    for ni, nj in zip(case.ni, case.nj):
        n = (nj + 2) * (ni + 2) * (case.dimq + 3)
        out.append(vals[k:k + n].reshape(nj + 2, ni + 2, case.dimq + 3))
        k += n
    return step, tt, out


# ---- the same grid cut into two blocks (block-interface tests) -----------------------
def _rotate_block(xh, rh, q):
    """Re-index a block by a quarter turn: new (i', j') = (j, n_i+1-i).
    Physical coordinates and velocity components are unchanged, the index
    space stays right-handed."""
    xr = xh[::-1, :].T.copy()
    rr = rh[::-1, :].T.copy()
    # interior cells: new[j'-1, i'-1] = old[j-1, i-1] with j = i', i = n+1-j'
    qi = q[2:-2, 2:-2, :]
    qn = np.transpose(qi, (1, 0, 2))[::-1, :, :].copy()
    qq = ref_zeros(q.shape[-1], qn.shape[1], qn.shape[0])
    qq[2:-2, 2:-2, :] = qn
    return xr, rr, qq


def to_original(case: "Case", b: int, interior_arr: np.ndarray) -> np.ndarray:
    """Interior array of a split block re-indexed back to the unsplit grid's
    orientation (identity unless the block was rotated)."""
    if case.meta.get("rotated", [False] * case.nblocks)[b]:
        return interior_arr[::-1].transpose(1, 0, 2)
    return interior_arr


def split_case_i(case: "Case", k: int, orient: str = "ii") -> "Case":
    """Cut block 1 of `case` after column k into two blocks joined by one cut.
    orient: 'ii' (geo 1), 'ij' (block 2 rotated, geo 2), 'ji' (block 1 rotated,
    geo 3) or 'jj' (both rotated, geo 4).  All physical sides become
    zero-gradient except the inflow (i-min), kept as in `case`."""
    assert case.nblocks == 1
    ni, nj = case.ni[0], case.nj[0]
    xh, rh, q = case.xh[0], case.rh[0], case.q0[0]
    n1, n2 = k, ni - k
    x1, r1 = xh[: k + 1].copy(), rh[: k + 1].copy()
    x2, r2 = xh[k:].copy(), rh[k:].copy()
    q1 = ref_zeros(q.shape[-1], n1, nj)
    q1[2:-2, 2:-2] = q[2:-2, 2:2 + n1]
    q2 = ref_zeros(q.shape[-1], n2, nj)
    q2[2:-2, 2:-2] = q[2:-2, 2 + n1:2 + ni]
    rot1, rot2 = orient[0] == "j", orient[1] == "j"
    if rot1:
        x1, r1, q1 = _rotate_block(x1, r1, q1)
    if rot2:
        x2, r2, q2 = _rotate_block(x2, r2, q2)
    # sides of the original orientation -> sides of the (possibly rotated) block
    rot_side = {SIDE_IMIN: SIDE_JMAX, SIDE_IMAX: SIDE_JMIN, SIDE_JMIN: SIDE_IMIN, SIDE_JMAX: SIDE_IMAX}
    inflow = [r for r in case.bc_rows if r[1] == SIDE_IMIN]
    rows = []
    for r in inflow:
        rows.append([1, rot_side[SIDE_IMIN] if rot1 else SIDE_IMIN, 1, nj, int(r[4]), int(r[5])])
    for side in (SIDE_JMIN, SIDE_JMAX):
        rows.append([1, rot_side[side] if rot1 else side, 1, n1, BC_EXTRAP, 0])
        rows.append([2, rot_side[side] if rot2 else side, 1, n2, BC_EXTRAP, 0])
    rows.append([2, rot_side[SIDE_IMAX] if rot2 else SIDE_IMAX, 1, nj, BC_EXTRAP, 0])
    # cut: block 1 column n1 <-> block 2 column 1
    if not rot1 and not rot2:
        cut = [1, nj, 1, 1, n1, 1, 1, 2, 1, 1, -1]
    elif not rot1 and rot2:   # block 2 side is j-max (j' = n2)
        cut = [2, nj, 1, 1, n1, 1, 1, 2, n2, 1, 1]
    elif rot1 and not rot2:   # block 1 side is j-min (j' = 1)
        cut = [3, nj, 1, 1, 1, 1, -1, 2, 1, 1, -1]
    else:
        cut = [4, nj, 1, 1, 1, 1, -1, 2, n2, 1, 1]
    nis = [nj if rot1 else n1, nj if rot2 else n2]
    njs = [n1 if rot1 else nj, n2 if rot2 else nj]
    out = Case(case.name + "_split" + orient, nis, njs, case.ny, case.params, [x1, x2], [r1, r2], [q1, q2],
               np.array(rows, dtype=np.int32), case.bc_pool.copy(), np.array([cut], dtype=np.int32),
               description=case.description + f" split at i={k} ({orient})")
    out.meta["rotated"] = [rot1, rot2]
    return out


# ---- random smooth states on skewed grids (parity tests) -----------------------------
def random_case(seed=0, ni=(13,), nj=(9,), ny=1, scheme=SCHEME_LUSGS, viscous=0, limiter=0,
                time_step=0, k_mus=1.0 / 3.0, high_order=1, mach=0.5, with_cut=False,
                dimension=DIM_PLANAR) -> Case:
    """Smooth random perturbations of a uniform flow on skewed curvilinear
    blocks; all sides zero-gradient except a FIXED inflow and a slip wall.
    Axisymmetric (``dimension=DIM_AXISYMMETRIC``): the grid is lifted off the
    axis to r >= 0.3 (y is the radius)."""
    rng = np.random.default_rng(seed)
    kappa, rgas = 1.4, 287.0
    xhs, rhs, qs = [], [], []
    x_off = 0.0
    for b, (n_i, n_j) in enumerate(zip(ni, nj)):
        s = np.linspace(0, 1, n_i + 1)[:, None]
        t = np.linspace(0, 1, n_j + 1)[None, :]
        xh = x_off + s * 1.0 + 0.05 * np.sin(2 * np.pi * t) * (1 - s) * 0.5 + 0.02 * rng.standard_normal((n_i + 1, n_j + 1)) / max(n_i, n_j)
        rh = t * 0.8 + 0.04 * np.sin(np.pi * s) + 0.02 * rng.standard_normal((n_i + 1, n_j + 1)) / max(n_i, n_j)
        if dimension == DIM_AXISYMMETRIC:
            rh = rh + 0.3
        if with_cut and b > 0:
            xh[0, :] = xhs[-1][-1, :]
            rh[0, :] = rhs[-1][-1, :]
        xhs.append(xh)
        rhs.append(rh)
        x_off = float(xh[-1, :].mean())
        T = 300.0 * (1 + 0.1 * rng.random((n_j, n_i)))
        rho = 1.2 * (1 + 0.2 * rng.random((n_j, n_i)))
        c0 = np.sqrt(kappa * rgas * 300.0)
        u = mach * c0 * (1 + 0.3 * rng.standard_normal((n_j, n_i)))
        v = 0.3 * mach * c0 * rng.standard_normal((n_j, n_i))
        p = rho * rgas * T
        if ny == 1:
            qc = cons_from_prim(rho, u, v, p, kappa)
        else:
            Y = rng.random((n_j, n_i, ny)) + 0.1
            Y /= Y.sum(axis=-1, keepdims=True)
            qc = np.concatenate([rho[..., None] * Y, (rho * u)[..., None], (rho * v)[..., None],
                                 (p / (kappa - 1) + 0.5 * rho * (u * u + v * v))[..., None]], axis=-1)
        qs.append(_fill_q(n_i, n_j, qc))
    pool = np.array(fixed_state_pool(1.2, 0.5 * np.sqrt(kappa * rgas * 300.0), 0.0, 1.2 * rgas * 300.0,
                                     kappa, rgas, mu=1.2 * 1.6e-5, ny=ny) + [300.0, 100.0, 0.0, 1.1e5],
                    dtype=np.float64)
    rows = []
    cut = []
    nb = len(ni)
    for b in range(nb):
        first, last = b == 0, b == nb - 1
        if first:
            rows.append([b + 1, SIDE_IMIN, 1, nj[b], BC_FIXED, 0])
        if last:
            rows.append([b + 1, SIDE_IMAX, 1, nj[b], BC_EXTRAP, 0])
        rows.append([b + 1, SIDE_JMIN, 1, ni[b], BC_NOSLIP if viscous else BC_REFLECT, 0])
        rows.append([b + 1, SIDE_JMAX, 1, ni[b], BC_EXTRAP, 0])
        if not last:
            if with_cut:
                cut.append([1, nj[b], 1, b + 1, ni[b], 1, 1, b + 2, 1, 1, -1])
            else:
                rows.append([b + 1, SIDE_IMAX, 1, nj[b], BC_EXTRAP, 0])
                rows.append([b + 2, SIDE_IMIN, 1, nj[b + 1], BC_EXTRAP, 0])
    params = make_params(scheme=scheme, time_step=time_step, high_order=high_order, limiter=limiter,
                         viscous=viscous, cfl=0.8, k_mus=k_mus, kappa=kappa, rgas=rgas, nu=1.6e-5 * 1e3,
                         num_internal_loop=4, cfl_tau=2.0, omega_max=1.0, omega_min=1e-6, dqmax=0.5,
                         res_rate_warning=1e300, res_rate_error=1e300, dimension=dimension)
    return Case(f"random{seed}", list(ni), list(nj), ny, params, xhs, rhs, qs,
                np.array(rows, dtype=np.int32), pool,
                np.array(cut, dtype=np.int32).reshape(-1, 11), description="random smooth test state")


# ---- reacting flame-sheet mixing layer (Chemkin thermo, nY = ns) ---------------
def mixing_layer(mech, ni=1024, nj=512, seed=0, scheme=SCHEME_LUSGS, reaction=2, viscous=1, cfl=0.5,
                 length=0.05, height=0.025, T_fuel=300.0, T_ox=1500.0, u_fuel=60.0, u_ox=120.0,
                 p0=1.01325e5, order=3, fuel=None, ox=None, species=None) -> Case:
    """Two parallel streams meeting at a flame sheet: fuel (y < height/2,
    T_fuel) and oxidiser (y > height/2, T_ox) at pressure p0, uniform
    Cartesian grid.  Mass fractions are drawn from ``seed`` and normalised
    unless given: fuel H2 in [0.05, 0.3] with N2, oxidiser O2 in [0.2, 0.3],
    H2O in [0, 0.1] with N2.  Inflow (i-min) holds each stream fixed, as
    calc_boundary does in sample/no27.chemkin_reaction/condition.f90
    (static enthalpy plus the kinetic term); the outflow is zero-gradient
    and the j sides are slip walls.  The initial state is the two streams
    with the sheet at mid-height.  ``mech`` is a chem.Mech with ns = nY;
    ``species`` names its species in order (needed for drawn compositions)."""
    from . import chem

    ns = mech.ns
    rng = np.random.default_rng(seed)
    species = list(species) if species is not None else None

    def rel(ys):
        r = np.zeros(ns)
        for name, y in ys.items():
            r[species.index(name)] = y
        return r / r.sum()

    if fuel is None or ox is None:
        if species is None:
            raise ValueError("mixing_layer: give the species names or explicit stream compositions")
        yH2 = rng.uniform(0.05, 0.3)
        fuel = rel({"H2": yH2, "N2": 1.0 - yH2})
        yO2, yH2O = rng.uniform(0.2, 0.3), rng.uniform(0.0, 0.1)
        ox = rel({"O2": yO2, "H2O": yH2O, "N2": 1.0 - yO2 - yH2O})
    states = []
    for T, y, u in ((T_fuel, fuel, u_fuel), (T_ox, ox, u_ox)):
        w, vhi, dhi = chem.boundary_state(mech, p0, T, y)
        w = w.copy()
        w[1] = u
        w[ns + 3 + 2] = w[ns + 3 + 2] + 0.5 * u * u   # H += kinetic energy
        states.append((w, vhi))
    dimq, dimw = ns + 3, ns + 8
    jm = nj // 2
    qcell = np.zeros((nj, ni, dimq))
    wcell = np.zeros((nj, ni, dimw))
    for s, (w, vhi) in enumerate(states):
        rows = slice(0, jm) if s == 0 else slice(jm, nj)
        rho, u, p = w[0], w[1], w[3]
        e_int = w[ns + 5] - 0.5 * u * u - p / rho       # H - u^2/2 - p/rho
        qcell[rows, :, :ns] = rho * w[4:4 + ns]
        qcell[rows, :, ns] = rho * u
        qcell[rows, :, ns + 1] = 0.0
        qcell[rows, :, ns + 2] = rho * (e_int + 0.5 * u * u)
        wcell[rows, :, :] = w
    w0 = ref_zeros(dimw, ni, nj)
    w0[2:nj + 2, 2:ni + 2, :] = wcell
    pool = np.concatenate([np.concatenate(st) for st in states])
    off = dimw + ns
    rows = [[1, SIDE_IMIN, 1, jm, BC_FIXED, 0], [1, SIDE_IMIN, jm + 1, nj, BC_FIXED, off],
            [1, SIDE_IMAX, 1, nj, BC_EXTRAP, 0], [1, SIDE_JMIN, 1, ni, BC_REFLECT, 0],
            [1, SIDE_JMAX, 1, ni, BC_REFLECT, 0]]
    xh, rh = rect_nodes(ni, nj, 0.0, length, 0.0, height)
    k_mus = -1.0 if order == 2 else 1.0 / 3.0
    params = make_params(scheme=scheme, time_step=0, high_order=1, limiter=LIMITER_REFERENCE, viscous=viscous,
                         thermal=2, reaction=reaction, cfl=cfl, k_mus=k_mus, kappa=1.4, rgas=287.0, nu=0.0,
                         num_internal_loop=2)
    return Case("mixing_layer", [ni], [nj], ns, params, [xh], [rh], [_fill_q(ni, nj, qcell)],
                np.array(rows, dtype=np.int32), pool, np.zeros((0, 11), dtype=np.int32),
                description=f"reacting mixing layer {ni}x{nj}, {ns} species / {mech.nr} reactions, "
                            f"fuel {T_fuel:g} K / oxidiser {T_ox:g} K, p {p0:g} Pa",
                mech=mech, w0=[w0])


# ---- the reference's sample no27 (H2 injected into an air boundary layer) ------------
R_UNI = 8.3144621e3   # J/(kmol K), chem_var of the reference's Chemkin model


def no27_cold(mech, state, species=("N2", "O2", "H2", "OH", "O", "H", "HO2", "H2O"), u_in=0.603, v_wall=0.02,
              p_ox=1e5, T_ox=300.0, p_f=1e5, T_f=300.0, cfl=0.5) -> Case:
    """sample/no27.chemkin_reaction as its cold-flow initialisation runs it
    (checkout.ini.inp, control.ini.inp): one 223 x 99 block (ideal_bl.x),
    preconditioned LU-SGS with local time steps, viscous, Chemkin thermo
    without reactions, MUSCL off, CFL 0.5.  ``state`` holds the grid and the
    restart state (tests/golden/no27_state.npz); ``mech`` is the parsed
    no27 mechanism with its transport data.

    Boundaries of the sample's condition.f90 at the steps of that run (the
    ignition spot it ramps in after step 40000 is off): i-min oxidizer
    inflow (N2:O2 = 4:1 by moles at T_ox, u = u_in, p from the interior,
    :53-66), i-max and j-max outflow at 1e5 Pa (:69-79, :135-145), j-min a
    slip wall for i <= 54 (:93-99) and a wall injecting fuel (H2 at T_f,
    v = v_wall) beyond (:104-131)."""
    from . import chem

    ns = mech.ns
    xh, rh = np.asarray(state["x"], dtype=np.float64), np.asarray(state["y"], dtype=np.float64)
    ni, nj = xh.shape[0] - 1, xh.shape[1] - 1
    q = np.asarray(state["q"], dtype=np.float64)
    assert q.shape == (nj, ni, ns + 3)
    rho = q[..., :ns].sum(-1)
    dimw = ns + 8
    w0 = ref_zeros(dimw, ni, nj)
    w0[2:nj + 2, 2:ni + 2, 0] = rho
    w0[2:nj + 2, 2:ni + 2, 3] = state["p"]
    w0[2:nj + 2, 2:ni + 2, ns + 6] = state["rgas"]

    def stream(p, T, moles):
        rel = np.zeros(ns)
        for name, n in moles.items():
            k = species.index(name)
            rel[k] = n * mech.mw[k]
        w, vhi, _ = chem.boundary_state(mech, p, T, rel)
        mw = R_UNI / w[ns + 6]
        return w, vhi, mw

    wo, vhio, mwo = stream(p_ox, T_ox, {"N2": 4.0, "O2": 1.0})
    wf, vhif, mwf = stream(p_f, T_f, {"H2": 2.0})
    wo = wo.copy()
    wo[1], wo[2] = u_in, 0.0
    pool_o = list(wo) + list(vhio) + [mwo, R_UNI * T_ox]
    pool_f = list(wf) + list(vhif) + [mwf, R_UNI * T_f, v_wall]
    off_f = len(pool_o)
    off_p = off_f + len(pool_f)
    pool = np.array(pool_o + pool_f + [1e5], dtype=np.float64)
    rows = [[1, SIDE_IMIN, 1, nj, BC_INFLOW_MIX_P, 0], [1, SIDE_IMAX, 1, nj, BC_OUTFLOW_P, off_p],
            [1, SIDE_JMIN, 1, 54, BC_REFLECT, 0], [1, SIDE_JMIN, 55, ni, BC_INJECT_WALL, off_f],
            [1, SIDE_JMAX, 1, ni, BC_OUTFLOW_P, off_p]]
    params = make_params(scheme=SCHEME_PRECON_LUSGS, time_step=1, high_order=1, phi_mus=0.0, k_mus=1.0 / 3.0,
                         limiter=LIMITER_REFERENCE, viscous=1, thermal=2, reaction=0, cfl=cfl, kappa=1.4,
                         rgas=287.0, nu=0.0, num_internal_loop=2)
    return Case("no27_cold", [ni], [nj], ns, params, [xh], [rh], [_fill_q(ni, nj, q)],
                np.array(rows, dtype=np.int32), pool, np.zeros((0, 11), dtype=np.int32),
                description=f"sample no27 (H2 injection into air), {ni}x{nj} cells, {ns} species, cold flow, "
                            "preconditioned LU-SGS, local dt, viscous, first order",
                mech=mech, w0=[w0], meta={"step": int(state["step"])})


# ---- batch of independent laminar plate / channel blocks -------------------------
def batched_plates(n=256, seed=0, ni_range=(64, 512), nj_range=(32, 256), mach_range=(0.05, 5.0),
                   re_range=(1e3, 1e6), cfl=1.0, order=2, length=1.0, height=0.25) -> Case:
    """``n`` independent laminar flat-plate or channel blocks (one block each,
    no cuts), sizes ni x nj drawn uniformly from the ranges, Mach uniform and
    Reynolds number log-uniform in the ranges (Re = U L / nu, so the
    kinematic viscosity is set per block).  Air at rho = 1.2, p = 1e5, ideal
    gas (gamma 1.4, R 287).  i-min: fixed free stream; i-max: zero gradient;
    j-min: no-slip wall; j-max: no-slip wall (channel) or zero gradient
    (plate), chosen by seed.  Uniform free stream as the initial state.
    Preconditioned LU-SGS with local time steps, viscous."""
    rng = np.random.default_rng(seed)
    kappa, rgas = 1.4, 287.0
    rho0, p0 = 1.2, 1e5
    c0 = float(np.sqrt(kappa * p0 / rho0))
    nis, njs, xhs, rhs, qs, nus = [], [], [], [], [], []
    rows, pool = [], []
    for b in range(n):
        ni = int(rng.integers(ni_range[0], ni_range[1] + 1))
        nj = int(rng.integers(nj_range[0], nj_range[1] + 1))
        mach = float(rng.uniform(*mach_range))
        re = float(10.0 ** rng.uniform(np.log10(re_range[0]), np.log10(re_range[1])))
        channel = bool(rng.integers(0, 2))
        u0 = mach * c0
        nu = u0 * length / re
        xh, rh = rect_nodes(ni, nj, 0.0, length, 0.0, height)
        qcell = np.broadcast_to(cons_from_prim(np.array(rho0), np.array(u0), np.array(0.0), np.array(p0), kappa),
                                (nj, ni, 4)).copy()
        off = len(pool)
        pool.extend(fixed_state_pool(rho0, u0, 0.0, p0, kappa, rgas, mu=rho0 * nu))
        rows += [[b + 1, SIDE_IMIN, 1, nj, BC_FIXED, off], [b + 1, SIDE_IMAX, 1, nj, BC_EXTRAP, 0],
                 [b + 1, SIDE_JMIN, 1, ni, BC_NOSLIP, 0],
                 [b + 1, SIDE_JMAX, 1, ni, BC_NOSLIP if channel else BC_EXTRAP, 0]]
        nis.append(ni)
        njs.append(nj)
        xhs.append(xh)
        rhs.append(rh)
        qs.append(_fill_q(ni, nj, qcell))
        nus.append(nu)
    k_mus = -1.0 if order == 2 else 1.0 / 3.0
    params = make_params(scheme=SCHEME_PRECON_LUSGS, time_step=1, high_order=1, limiter=LIMITER_REFERENCE,
                         viscous=1, cfl=cfl, k_mus=k_mus, kappa=kappa, rgas=rgas, nu=nus[0])
    cells = int(sum(a * b for a, b in zip(nis, njs)))
    return Case("batched_plates", nis, njs, 1, params, xhs, rhs, qs, np.array(rows, dtype=np.int32),
                np.array(pool, dtype=np.float64), np.zeros((0, 11), dtype=np.int32),
                description=f"{n} independent laminar plate/channel blocks, {cells} cells, "
                            f"Mach {mach_range[0]:g}-{mach_range[1]:g}, Re {re_range[0]:g}-{re_range[1]:g}, "
                            f"preconditioned LU-SGS, local dt, viscous",
                block_nu=nus, meta={"seed": seed})


def sub_case(case: Case, blocks) -> Case:
    """The blocks ``blocks`` of a batch of independent blocks as a case of
    their own (BC rows renumbered, pool shared)."""
    blocks = list(blocks)
    rows = []
    for r in case.bc_rows:
        if int(r[0]) - 1 in blocks:
            rr = [int(x) for x in r]
            rr[0] = blocks.index(int(r[0]) - 1) + 1
            rows.append(rr)
    return Case(case.name + "_sub", [case.ni[b] for b in blocks], [case.nj[b] for b in blocks], case.ny,
                case.params, [case.xh[b] for b in blocks], [case.rh[b] for b in blocks],
                [case.q0[b] for b in blocks], np.array(rows, dtype=np.int32), case.bc_pool,
                np.zeros((0, 11), dtype=np.int32), description=case.description,
                block_nu=None if case.block_nu is None else [case.block_nu[b] for b in blocks])
