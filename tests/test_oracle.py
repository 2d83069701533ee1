"""Pins of the CPU oracle (oracle/shcfd_oracle.c) that do not come from the
oracle's own formulas: closed forms, invariants and brute-force solves.

The reference ships no tests or golden vectors for this path and cannot be
built here (its kernels need the Python-2 code generator's output), so each
oracle routine is checked against what the mathematics fixes, chosen so that
a dropped term, a wrong sign or a transposed index fails one of them."""
import numpy as np
import pytest

from shared_cfd_amd import cases as C
from parity import interior, iface, jface


def _oracle(case, geom=True):
    from oracle.oracle import Oracle

    return Oracle(case, init_geometry=geom)


def _int(o, name, b=0):
    return interior(o.field(name, b), o.case.ni[b], o.case.nj[b])


# --------------------------------------------------------------------------- geometry
def test_geometry_closed_form_uniform():
    case = C.sod(ni=8, nj=3)
    o = _oracle(case)
    dx = 1.0 / 8
    np.testing.assert_allclose(_int(o, "VOL"), dx * dx, rtol=1e-14)
    np.testing.assert_allclose(iface(o.field("DSI", 0), 8, 3), dx, rtol=1e-14)
    np.testing.assert_allclose(jface(o.field("DSJ", 0), 8, 3), dx, rtol=1e-14)
    np.testing.assert_allclose(iface(o.field("VNI", 0), 8, 3)[..., 0], 1.0)
    np.testing.assert_allclose(jface(o.field("VNJ", 0), 8, 3)[..., 1], 1.0)
    np.testing.assert_allclose(_int(o, "DSCI"), dx, rtol=1e-14)
    np.testing.assert_allclose(_int(o, "VNCI")[..., 0], 1.0)
    np.testing.assert_allclose(_int(o, "VNCJ")[..., 1], 1.0)
    # inverse metric of a uniform grid: dxi/dx = 1/dx
    g = iface(o.field("GEOJACI", 0), 8, 3)
    np.testing.assert_allclose(g[..., 0], 1 / dx, rtol=1e-12)
    np.testing.assert_allclose(g[..., 3], 1 / dx, rtol=1e-12)
    np.testing.assert_allclose(g[..., 1], 0.0, atol=1e-9)


def test_geometry_closed_cells_and_shoelace():
    """Sum of ds*n over the four faces of every cell is zero (discrete
    divergence theorem), and Vol equals the shoelace area of the node quad."""
    case = C.random_case(seed=4, ni=(11,), nj=(7,))
    o = _oracle(case)
    ni, nj = 11, 7
    dsi, dsj = o.field("DSI", 0)[..., 0], o.field("DSJ", 0)[..., 0]
    vni, vnj = o.field("VNI", 0), o.field("VNJ", 0)
    xh, rh = case.xh[0], case.rh[0]
    for j in range(1, nj + 1):
        for i in range(1, ni + 1):
            J, I = j + 1, i + 1
            s = dsi[J, I] * vni[J, I] - dsi[J, I - 1] * vni[J, I - 1] + dsj[J, I] * vnj[J, I] - dsj[J - 1, I] * vnj[J - 1, I]
            assert np.max(np.abs(s)) < 1e-14
            px = [xh[i - 1, j - 1], xh[i, j - 1], xh[i, j], xh[i - 1, j]]
            py = [rh[i - 1, j - 1], rh[i, j - 1], rh[i, j], rh[i - 1, j]]
            area = 0.5 * sum(px[k] * py[(k + 1) % 4] - px[(k + 1) % 4] * py[k] for k in range(4))
            assert abs(o.field("VOL", 0)[J, I, 0] - area) < 1e-14 * abs(area) + 1e-17


# --------------------------------------------------------------------------- primitive / thermo
def test_primitive_round_trip():
    case = C.random_case(seed=2, ni=(9,), nj=(6,))
    o = _oracle(case)
    o.set_w()
    o.set_thermo_prop()
    q = _int(o, "Q")
    w = _int(o, "W")
    kap, R = case.params.kappa, case.params.rgas
    rho = q[..., 0]
    u, v = q[..., 1] / rho, q[..., 2] / rho
    p = (kap - 1) * (q[..., 3] - 0.5 * rho * (u * u + v * v))
    np.testing.assert_allclose(w[..., 0], rho, rtol=1e-15)
    np.testing.assert_allclose(w[..., 1], u, rtol=1e-14)
    np.testing.assert_allclose(w[..., 3], p, rtol=1e-12)
    np.testing.assert_allclose(w[..., 5], kap)
    np.testing.assert_allclose(w[..., 6], kap / (kap - 1) * p / rho + 0.5 * (u * u + v * v), rtol=1e-13)
    np.testing.assert_allclose(_int(o, "VHI")[..., 0], kap / (kap - 1) * p / rho, rtol=1e-12)


# --------------------------------------------------------------------------- MUSCL
def _muscl_case(field, k_mus, limiter=0):
    case = C.sod(ni=12, nj=3, limiter=limiter)
    case.params.k_mus = k_mus
    o = _oracle(case)
    w = o.field("W", 0)
    w[...] = field[..., None]
    o.set_hv()
    return o


@pytest.mark.parametrize("k_mus", [-1.0, 1.0 / 3.0])
def test_muscl_linear_exact(k_mus):
    """Du = Dl (linear data) gives s = 1 and the exact face value."""
    nj, ni = 3, 12
    ii = np.arange(-1, ni + 3, dtype=float)
    field = np.broadcast_to(2.0 + 0.5 * ii[None, :], (nj + 4, ni + 4)).copy()
    o = _muscl_case(field, k_mus)
    L = iface(o.field("WHLI", 0), ni, nj)[..., 0]
    Rr = iface(o.field("WHRI", 0), ni, nj)[..., 0]
    faces = 2.0 + 0.5 * (np.arange(0, ni + 1) + 0.5)
    np.testing.assert_allclose(L, np.broadcast_to(faces, L.shape), rtol=1e-15)
    np.testing.assert_allclose(Rr, np.broadcast_to(faces, Rr.shape), rtol=1e-15)


def test_muscl_extremum_not_clipped_and_constant():
    """Du*Dl < 0 gives s < 0 (the reference's limiter is not clipped at an
    extremum: muscl.f90:32); Du = Dl = 0 gives the cell value."""
    nj, ni = 3, 12
    ii = np.arange(-1, ni + 3)
    field = np.broadcast_to(np.where(ii % 2 == 0, 1.0, 2.0)[None, :], (nj + 4, ni + 4)).copy()
    o = _muscl_case(field, -1.0)
    L = iface(o.field("WHLI", 0), ni, nj)[..., 0]
    # cell i has Du = -Dl = d, s = -1, delta = 0.5*(-1)*((1+s*k)Du + (1-s*k)Dl) with k=-1: = -(Du... )
    w0 = field[2, 1:ni + 2]
    du = field[2, 2:ni + 3] - w0
    s = -1.0
    delta = 0.5 * s * ((1 - s) * du + (1 + s) * (-du))
    np.testing.assert_allclose(L[0], w0 + 0.5 * delta, rtol=1e-15)
    assert np.any(np.abs(L[0] - w0) > 0)  # really not clipped
    const = np.full((nj + 4, ni + 4), 3.0)
    o = _muscl_case(const, 1.0 / 3.0)
    np.testing.assert_array_equal(iface(o.field("WHLI", 0), ni, nj)[..., 0], 3.0)


def test_minmod_clips():
    nj, ni = 3, 12
    ii = np.arange(-1, ni + 3)
    field = np.broadcast_to(np.where(ii % 2 == 0, 1.0, 2.0)[None, :], (nj + 4, ni + 4)).copy()
    o = _muscl_case(field, -1.0, limiter=1)
    L = iface(o.field("WHLI", 0), ni, nj)[..., 0]
    np.testing.assert_array_equal(L[0], field[2, 1:ni + 2])


# --------------------------------------------------------------------------- SLAU
def _flux_case(wl, wr, n):
    """One-face evaluation through the oracle's set_tg on a 2x2 block."""
    case = C.sod(ni=2, nj=2)
    o = _oracle(case)
    for name, w in (("WHLI", wl), ("WHRI", wr)):
        f = o.field(name, 0)
        f[...] = np.asarray(w)[None, None, :]
    vni = o.field("VNI", 0)
    vni[...] = np.asarray(n)[None, None, :]
    o.set_tg()
    return iface(o.field("TGI", 0), 2, 2)[0, 0].copy()


def _w(rho, u, v, p, kap=1.4, R=287.0):
    H = kap / (kap - 1) * p / rho + 0.5 * (u * u + v * v)
    return [rho, u, v, p, 1.0, kap, H, R, 0.0]


def _phys_flux(w, n):
    rho, u, v, p, _, _, H = w[:7]
    un = n[0] * u + n[1] * v
    return np.array([rho * un, rho * u * un + p * n[0], rho * v * un + p * n[1], rho * H * un])


@pytest.mark.parametrize("u,v", [(30.0, 10.0), (-400.0, 50.0), (700.0, -30.0), (0.0, 0.0)])
def test_slau_consistency(u, v):
    """Equal left/right states give the physical flux F(w).n (consistency;
    uses beta+ + beta- = 1)."""
    n = np.array([0.6, 0.8])
    w = _w(1.2, u, v, 1.1e5)
    np.testing.assert_allclose(_flux_case(w, w, n), _phys_flux(w, n), rtol=1e-13, atol=1e-9)


def test_slau_antisymmetry_and_contact():
    rng = np.random.default_rng(0)
    for _ in range(20):
        wl = _w(rng.uniform(0.5, 2), rng.uniform(-500, 500), rng.uniform(-500, 500), rng.uniform(5e4, 2e5))
        wr = _w(rng.uniform(0.5, 2), rng.uniform(-500, 500), rng.uniform(-500, 500), rng.uniform(5e4, 2e5))
        th = rng.uniform(0, 2 * np.pi)
        n = np.array([np.cos(th), np.sin(th)])
        f1 = _flux_case(wl, wr, n)
        f2 = _flux_case(wr, wl, -n)
        np.testing.assert_allclose(f1, -f2, rtol=1e-13, atol=1e-8)
    # stationary contact: no mass flux, momentum flux = p n
    f = _flux_case(_w(1.0, 0, 0, 1e5), _w(0.2, 0, 0, 1e5), [1.0, 0.0])
    assert abs(f[0]) < 1e-12 and abs(f[3]) < 1e-6
    np.testing.assert_allclose(f[1], 1e5, rtol=1e-14)


def test_slau_supersonic_upwind():
    """Both sides supersonic to the right: beta_L = 1, beta_R = 0, so the
    pressure term is p_L exactly, and every convected quantity is taken from
    the left state (momentum = mdot*u_L + p_L n, energy = mdot*H_L)."""
    n = np.array([1.0, 0.0])
    wl = _w(1.0, 900.0, 40.0, 1e5)
    wr = _w(0.8, 850.0, 20.0, 0.9e5)
    f = _flux_case(wl, wr, n)
    mdot = f[0]
    assert mdot > 0
    np.testing.assert_allclose(f[1], mdot * wl[1] + wl[3], rtol=1e-14)
    np.testing.assert_allclose(f[2], mdot * wl[2], rtol=1e-14)
    np.testing.assert_allclose(f[3], mdot * wl[6], rtol=1e-14)
    # mirrored: supersonic to the left takes everything from the right state
    f = _flux_case(_w(1.0, -900.0, 40.0, 1e5), _w(0.8, -850.0, 20.0, 0.9e5), n)
    assert f[0] < 0
    np.testing.assert_allclose(f[1], f[0] * -850.0 + 0.9e5, rtol=1e-14)


# --------------------------------------------------------------------------- viscous
def test_viscous_couette_closed_form():
    """Linear shear u = U y and linear temperature: central differences are
    exact, so TGvj = (0, mu U, 0, u mu U + k dT/dy) on horizontal faces."""
    case = C.sod(ni=6, nj=5)
    case.params.viscous = 1
    o = _oracle(case)
    ni, nj = 6, 5
    kap, R = 1.4, 287.0
    X, Y = o.field("X", 0)[..., 0], o.field("R", 0)[..., 0]
    U, dTdy, mu0 = 3000.0, 5000.0, 2e-3
    w = o.field("W", 0)
    rho = 1.2
    T = 300.0 + dTdy * Y
    w[..., 0] = rho
    w[..., 1] = U * Y
    w[..., 2] = 0.0
    w[..., 3] = rho * R * T
    w[..., 4] = 1.0
    w[..., 5] = kap
    w[..., 7] = R
    w[..., 8] = mu0
    o.field("VHI", 0)[..., 0] = kap / (kap - 1) * R * T
    o.set_tgv()
    tj = jface(o.field("TGVJ", 0), ni, nj)
    yf = 0.5 * (Y[1:nj + 2, 2:ni + 2] + Y[2:nj + 3, 2:ni + 2])
    k = mu0 * kap / (kap - 1) * R
    np.testing.assert_allclose(tj[..., 1], mu0 * U, rtol=1e-10)
    np.testing.assert_allclose(tj[..., 2], 0.0, atol=1e-8)
    np.testing.assert_allclose(tj[..., 3], U * yf * mu0 * U + k * dTdy, rtol=1e-10)
    np.testing.assert_allclose(tj[..., 0], 0.0, atol=1e-12)


# --------------------------------------------------------------------------- time step
def test_dt_closed_form():
    case = C.sod(ni=10, nj=4)
    o = _oracle(case)
    w = o.field("W", 0)
    u, v, p, rho, kap = 100.0, -40.0, 1e5, 1.2, 1.4
    w[..., 0], w[..., 1], w[..., 2], w[..., 3], w[..., 5], w[..., 8] = rho, u, v, p, kap, 0.0
    dtmin = o.set_dt()
    a = np.sqrt(kap * p / rho)
    dx = 0.1
    expect = 0.5 * dx * dx / (2 * (abs(u) + a) * dx + 2 * (abs(v) + a) * dx)
    np.testing.assert_allclose(_int(o, "DT"), expect, rtol=1e-14)
    assert abs(dtmin - expect) < 1e-14 * expect


# --------------------------------------------------------------------------- Jacobians
def _prep(case):
    o = _oracle(case)
    o.prepare()
    o.set_dt()
    o.set_hv(); o.set_tg(); o.set_tgv()
    o.set_abpm()
    return o


def test_jacobian_homogeneity_and_split():
    """Euler homogeneity: (n1 A + n2 B) q = F(q).n; Ap + Am = A~ and
    Ap - Am = nu I."""
    case = C.random_case(seed=7, ni=(7,), nj=(5,))
    o = _prep(case)
    Ap, Am = _int(o, "AP"), _int(o, "AM")
    q, w = _int(o, "Q"), _int(o, "W")
    n = _int(o, "VNCI")
    for j in range(5):
        for i in range(7):
            At = (Ap[j, i] + Am[j, i]).reshape(4, 4, order="F")
            nu = (Ap[j, i] - Am[j, i]).reshape(4, 4, order="F")
            np.testing.assert_allclose(nu, nu[0, 0] * np.eye(4), atol=1e-9 * abs(nu[0, 0]))
            F = _phys_flux(w[j, i], n[j, i])
            np.testing.assert_allclose(At @ q[j, i], F, rtol=1e-12, atol=1e-9 * np.max(np.abs(F)))


def test_jacobian_finite_difference():
    """A~ against central differences of F(q).n (the reference's own check
    routine, store/debug/debug.f90:1-55, perturbs by 1 %)."""
    case = C.random_case(seed=8, ni=(4,), nj=(3,))
    o = _prep(case)
    Ap, Am = _int(o, "AP"), _int(o, "AM")
    q, n = _int(o, "Q"), _int(o, "VNCI")
    kap = case.params.kappa

    def flux(qc, nn):
        rho = qc[0]
        u, v = qc[1] / rho, qc[2] / rho
        p = (kap - 1) * (qc[3] - 0.5 * rho * (u * u + v * v))
        return _phys_flux([rho, u, v, p, 1.0, kap, (qc[3] + p) / rho], nn)

    j, i = 1, 2
    At = (Ap[j, i] + Am[j, i]).reshape(4, 4, order="F")
    J = np.zeros((4, 4))
    for k in range(4):
        h = 1e-6 * max(abs(q[j, i, k]), 1e-3)
        e = np.zeros(4); e[k] = h
        J[:, k] = (flux(q[j, i] + e, n[j, i]) - flux(q[j, i] - e, n[j, i])) / (2 * h)
    np.testing.assert_allclose(At, J, rtol=1e-6, atol=1e-6 * np.max(np.abs(J)))


def test_dual_preconditioner_algebra():
    """pre2.pre1 = c^2, Gamma * Gamma^-1 = I and S * S^-1 = I for the
    dual-time scalars (sch_dual.raw.f90:62-69, DualTimeStep.tex)."""
    case = C.random_case(seed=3, ni=(6,), nj=(5,), scheme=C.SCHEME_DUAL, viscous=1)
    o = _prep(case)
    ps = _int(o, "PRESCA")
    p1, p2, phi = _int(o, "PRE1"), _int(o, "PRE2"), _int(o, "PHI")[..., 0]
    w = _int(o, "W")
    c2 = 1.4 * w[..., 3] / w[..., 0]
    d = np.sum(p1 * p2, axis=-1)
    np.testing.assert_allclose(d, c2, rtol=1e-12)
    np.testing.assert_allclose(phi + ps[..., 1] + phi * ps[..., 1] * d, 0.0, atol=1e-12 * np.max(np.abs(phi)))
    np.testing.assert_allclose(ps[..., 2] + ps[..., 3] + ps[..., 2] * ps[..., 3] * d, 0.0,
                               atol=1e-12 * np.max(np.abs(ps[..., 3])))


# --------------------------------------------------------------------------- LU-SGS: brute force
def _cell_rhs(F, I, J, kind):
    """Right-hand side of one cell from the stored fluxes: the flux balance,
    minus Vol dq/dt for the inner-loop schemes, with the preconditioner
    (I + pre1 pre2^T) on the flux balance (precon) or on the whole (dual)."""
    tgi, tgj = F("TGI") - F("TGVI"), F("TGJ") - F("TGVJ")
    dsi, dsj = F("DSI")[..., 0], F("DSJ")[..., 0]
    r = dsi[J, I - 1] * tgi[J, I - 1] - dsi[J, I] * tgi[J, I] + dsj[J - 1, I] * tgj[J - 1, I] - dsj[J, I] * tgj[J, I]
    p1, p2 = F("PRE1")[J, I], F("PRE2")[J, I]
    if kind == "precon":
        r = r + np.dot(p2, r) * p1
    if kind in ("dual", "nr", "precon"):
        r = r - F("VOL")[J, I, 0] * F("DQDT")[J, I]
    if kind == "dual":
        r = r + F("PRESCA")[J, I][1] * np.dot(p2, r) * p1
    return r


def _brute_force_check(o, kind):
    """The LU-SGS update satisfies (D+L) D^-1 (D+U) Dq = R' exactly, with
    the block-lower/upper couplings dsc*A+/B+ and dsc*A-/B- of the
    neighbouring cells, D = 1/alpha (LU-SGS, NR, precon) or S/alpha (dual),
    and Dq = 0 outside the block (time_lusgs_viscous.f90, time_NR_viscous.f90,
    time_precon_viscous.f90, time_dual_viscous.f90)."""
    dual = kind == "dual"
    case = o.case
    ni, nj = case.ni[0], case.nj[0]
    N = ni * nj * 4
    idx = lambda i, j: ((j - 1) * ni + (i - 1)) * 4
    F = lambda name: o.field(name, 0)
    Dm = np.zeros((N, N)); Lm = np.zeros((N, N)); Um = np.zeros((N, N))
    Rv = np.zeros(N)
    for j in range(1, nj + 1):
        for i in range(1, ni + 1):
            J, I = j + 1, i + 1
            al = F("ALPHA")[J, I, 0]
            if dual:
                ps = F("PRESCA")[J, I]
                p1, p2 = F("PRE1")[J, I], F("PRE2")[J, I]
                S = ps[0] * (np.eye(4) + ps[2] * np.outer(p1, p2))
                D = S / al
            else:
                D = np.eye(4) / al
            a = idx(i, j)
            Dm[a:a + 4, a:a + 4] = D
            if i > 1:
                b = idx(i - 1, j)
                Lm[a:a + 4, b:b + 4] = -F("DSCI")[J, I - 1, 0] * F("AP")[J, I - 1].reshape(4, 4, order="F")
            if j > 1:
                b = idx(i, j - 1)
                Lm[a:a + 4, b:b + 4] = -F("DSCJ")[J - 1, I, 0] * F("BP")[J - 1, I].reshape(4, 4, order="F")
            if i < ni:
                b = idx(i + 1, j)
                Um[a:a + 4, b:b + 4] = F("DSCI")[J, I + 1, 0] * F("AM")[J, I + 1].reshape(4, 4, order="F")
            if j < nj:
                b = idx(i, j + 1)
                Um[a:a + 4, b:b + 4] = F("DSCJ")[J + 1, I, 0] * F("BM")[J + 1, I].reshape(4, 4, order="F")
            Rv[a:a + 4] = _cell_rhs(F, I, J, kind)
    M = (Dm + Lm) @ np.linalg.solve(Dm, Dm + Um)
    dq = np.concatenate([F("DQ")[j + 1, i + 1] for j in range(1, nj + 1) for i in range(1, ni + 1)])
    res = M @ dq - Rv
    assert np.linalg.norm(res) <= 1e-11 * np.linalg.norm(Rv), np.linalg.norm(res) / np.linalg.norm(Rv)


def test_lusgs_brute_force():
    case = C.random_case(seed=12, ni=(6,), nj=(5,), scheme=C.SCHEME_LUSGS)
    o = _prep(case)
    o.implicit(0)
    _brute_force_check(o, "lusgs")


def test_dual_lusgs_brute_force():
    case = C.random_case(seed=13, ni=(5,), nj=(6,), scheme=C.SCHEME_DUAL, viscous=1)
    o = _oracle(case)
    o.prepare()
    o.set_dt()
    o.copy_field(C.F["QPP"], C.F["QP"]); o.copy_field(C.F["QP"], C.F["Q"])
    q = o.field("Q", 0)
    q[2:-2, 2:-2] *= 1.001
    o.set_w(); o.set_thermo_prop(); o.set_bc(); o.set_hv(); o.set_tg(); o.set_tgv()
    o.set_abpm(); o.set_dqdt()
    q_before = _int(o, "Q").copy()
    err, res, om = o.implicit(1)
    _brute_force_check(o, "dual")
    # the update is q + min(omega_max, omega) dq, and res = sum dq_E^2
    dq = _int(o, "DQ")
    np.testing.assert_allclose(res, np.sum(dq[..., 3] ** 2), rtol=1e-12)
    np.testing.assert_allclose(_int(o, "Q"), q_before + min(case.params.omega_max, om) * dq, rtol=1e-14)


def _inner_state(o, kind):
    """Primitives, fluxes, set_ABpm and set_dqdt at the current q; returns the
    unsteady right-hand side of every interior cell."""
    o.set_w(); o.set_thermo_prop(); o.set_bc(); o.set_hv(); o.set_tg(); o.set_tgv()
    o.set_abpm(); o.set_dqdt()
    F = lambda name: o.field(name, 0)
    ni, nj = o.case.ni[0], o.case.nj[0]
    return np.array([_cell_rhs(F, i + 1, j + 1, kind) for j in range(1, nj + 1) for i in range(1, ni + 1)])


def _inner_case(scheme, seed, cfl=None):
    case = C.random_case(seed=seed, ni=(6,), nj=(5,), scheme=scheme, viscous=1,
                         time_step=1 if scheme == C.SCHEME_PRECON else 0, mach=0.3)
    if cfl is not None:
        case.params.cfl = cfl
    case.params.omega_max = 1.0
    o = _oracle(case)
    o.prepare()
    if scheme == C.SCHEME_NR:
        o.set_dt()
    o.copy_field(C.F["QPP"], C.F["QP"]); o.copy_field(C.F["QP"], C.F["Q"])
    return case, o


@pytest.mark.parametrize("scheme,kind", [(C.SCHEME_NR, "nr"), (C.SCHEME_PRECON, "precon")])
def test_inner_loop_brute_force(scheme, kind):
    """NR and precon inner iterations: the sweeps solve the same
    approximate factorisation as LU-SGS with the BDF2 right-hand side, and
    the update is q + min(omega_max, omega) dq with res = sum dq_E^2
    (NR/2d/time_NR_viscous.f90, precon/2d/time_precon_viscous.f90)."""
    case, o = _inner_case(scheme, 41)
    q = o.field("Q", 0)
    q[2:-2, 2:-2] *= 1.0 + 1e-3 * np.cos(np.arange(q.shape[1] - 4))[None, :, None]
    _inner_state(o, kind)
    q_before = _int(o, "Q").copy()
    err, res, om = o.implicit(1)
    assert err == 0
    _brute_force_check(o, kind)
    dq = _int(o, "DQ")
    np.testing.assert_allclose(res, np.sum(dq[..., 3] ** 2), rtol=1e-12)
    np.testing.assert_allclose(_int(o, "Q"), q_before + min(case.params.omega_max, om) * dq, rtol=1e-14)


@pytest.mark.parametrize("scheme,kind", [(C.SCHEME_NR, "nr"), (C.SCHEME_PRECON, "precon")])
def test_inner_loop_diagonal_small_dt(scheme, kind):
    """At a tiny time step the BDF2 term 3 Vol/(2 dt) dominates the
    operator, so one inner iteration must nearly zero the unsteady
    residual R(q) - Vol (3q - 4qp + qpp)/(2 dt): a diagonal Vol/dt (the
    steady LU-SGS one) would leave half of it, a wrong sign of the dq/dt
    term would grow it (sch_NR.raw.f90:134-136, sch_precon.raw.f90:194-196)."""
    case, o = _inner_case(scheme, 43, cfl=1e-5)
    o.set_first_dual(0)
    r0 = _inner_state(o, kind)
    err, _, _ = o.implicit(1)
    assert err == 0
    r1 = _inner_state(o, kind)
    assert np.linalg.norm(r1) < 1e-3 * np.linalg.norm(r0), np.linalg.norm(r1) / np.linalg.norm(r0)


def test_set_dqdt_bdf2_exact():
    """set_dqdt is exact for data quadratic in time (BDF2) and, on the first
    physical step, for data linear in time (sch_NR.raw.f90:143-172)."""
    case, o = _inner_case(C.SCHEME_NR, 45)
    dt = o.dt_grbl()
    assert dt > 0.0
    q0 = o.field("Q", 0).copy()
    a = 1e-2 * np.sin(np.arange(q0.size)).reshape(q0.shape)
    b = 3e-3 * np.cos(np.arange(q0.size)).reshape(q0.shape)
    qt = lambda n: q0 * (1.0 + a * n + b * n ** 2)   # q at t = n dt
    o.field("QPP", 0)[...] = qt(0); o.field("QP", 0)[...] = qt(1); o.field("Q", 0)[...] = qt(2)
    o.set_first_dual(0)
    o.set_dqdt()
    exact = interior(q0 * (a + 4.0 * b) / dt, 6, 5)
    np.testing.assert_allclose(_int(o, "DQDT"), exact, rtol=1e-9, atol=1e-12 * np.abs(exact).max())
    o.field("QP", 0)[...] = q0; o.field("Q", 0)[...] = q0 * (1.0 + a)
    o.set_first_dual(1)
    o.set_dqdt()
    exact = interior(q0 * a / dt, 6, 5)
    np.testing.assert_allclose(_int(o, "DQDT"), exact, rtol=1e-9, atol=1e-12 * np.abs(exact).max())


def test_precon_scaling_eigenvalue():
    """The precon scheme's (I + pre1 pre2^T) scales the acoustic direction
    (Y, u, v, H) by Ur^2/c^2 with Ur = min(max(nu/dmin, |u|), c) and leaves
    the directions orthogonal to pre2 alone (sch_precon.raw.f90:44-68)."""
    case, o = _inner_case(C.SCHEME_PRECON, 47)
    o.set_abpm()
    w = _int(o, "W")
    p1, p2 = _int(o, "PRE1"), _int(o, "PRE2")
    rho, u, v, p = w[..., 0], w[..., 1], w[..., 2], w[..., 3]
    c = np.sqrt(w[..., 5] * p / rho)
    dmin = np.minimum(_int(o, "DSCI")[..., 0], _int(o, "DSCJ")[..., 0])
    ur = np.minimum(np.maximum(w[..., 8] / rho / dmin, np.hypot(u, v)), c)
    lam = 1.0 + np.sum(p2 * p1, axis=-1)
    np.testing.assert_allclose(lam, ur ** 2 / c ** 2, rtol=1e-12)
    # pre1 is parallel to (Y, u, v, H)
    h = w[..., 6]
    base = np.stack([np.ones_like(u), u, v, h], axis=-1)
    np.testing.assert_allclose(p1 / base, (p1 / base)[..., :1] * np.ones(4), rtol=1e-12)


# --------------------------------------------------------------------------- invariants of whole steps
@pytest.mark.parametrize("scheme", [C.SCHEME_EULER, C.SCHEME_RK2, C.SCHEME_LUSGS, C.SCHEME_PRECON_LUSGS,
                                    C.SCHEME_NR, C.SCHEME_PRECON, C.SCHEME_DUAL])
def test_free_stream_preservation(scheme):
    """A uniform flow on a skewed grid with consistent boundary states stays
    uniform (closed cells + SLAU consistency)."""
    local = scheme in (C.SCHEME_PRECON_LUSGS, C.SCHEME_PRECON)
    case = C.random_case(seed=21, ni=(9,), nj=(7,), scheme=scheme, time_step=1 if local else 0)
    kap, R = 1.4, 287.0
    q0 = C.cons_from_prim(np.array(1.2), np.array(150.0), np.array(20.0), np.array(1.0e5), kap)
    case.q0[0][2:-2, 2:-2, :] = q0
    case.bc_rows = np.array([[1, s, 1, 99, C.BC_FIXED, 0] for s in range(4)], dtype=np.int32)
    case.bc_pool = np.array(C.fixed_state_pool(1.2, 150.0, 20.0, 1.0e5, kap, R, mu=1.2 * case.params.nu))
    o = _oracle(case)
    o.prepare()
    for _ in range(2):
        err, _, _ = o.step()
        assert err == 0
    q = _int(o, "Q")
    np.testing.assert_allclose(q, np.broadcast_to(q0, q.shape), rtol=1e-11)


def test_block_split_invariance():
    """Explicit Euler on one block equals the same grid cut into two blocks
    joined by an i-i cut (section_exchange): identical to rounding."""
    one = C.random_case(seed=5, ni=(16,), nj=(6,), scheme=C.SCHEME_EULER)
    one.bc_rows = np.array([r if r[1] == C.SIDE_IMIN else [r[0], r[1], r[2], r[3], C.BC_EXTRAP, 0]
                            for r in one.bc_rows], dtype=np.int32)
    two = C.split_case_i(one, 7)
    o1, o2 = _oracle(one), _oracle(two)
    o1.prepare(); o2.prepare()
    for _ in range(3):
        assert o1.step()[0] == 0 and o2.step()[0] == 0
    q1 = _int(o1, "Q")
    q2 = np.concatenate([_int(o2, "Q", 0), _int(o2, "Q", 1)], axis=1)
    np.testing.assert_array_equal(q1, q2)


@pytest.mark.parametrize("orient", ["ij", "ji", "jj"])
def test_cut_orientations(orient):
    """The transposed-block cut forms (geo 2, 3, 4) give the same ghost
    values as the i-i form on the same physical grid."""
    one = C.random_case(seed=6, ni=(12,), nj=(8,), scheme=C.SCHEME_EULER)
    one.bc_rows = np.array([r if r[1] == C.SIDE_IMIN else [r[0], r[1], r[2], r[3], C.BC_EXTRAP, 0]
                            for r in one.bc_rows], dtype=np.int32)
    ref = C.split_case_i(one, 5)
    alt = C.split_case_i(one, 5, orient=orient)
    o1, o2 = _oracle(ref), _oracle(alt)
    o1.prepare(); o2.prepare()
    for _ in range(2):
        assert o1.step()[0] == 0 and o2.step()[0] == 0
    for b in range(2):
        a = _int(o1, "Q", b)
        c = C.to_original(alt, b, _int(o2, "Q", b))
        np.testing.assert_allclose(c, a, rtol=1e-13, atol=1e-13 * np.max(np.abs(a)))


# --------------------------------------------------------------------------- exact Riemann solution
from riemann import exact_riemann  # noqa: E402


def test_exact_riemann_textbook_values():
    """Toro's test 1 (table 4.2): p* = 0.30313, u* = 0.92745."""
    x = np.array([0.2, 0.6, 0.75, 0.9])
    rho = exact_riemann(x, 0.25, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1))
    # contact at 0.5 + 0.92745*0.25 = 0.732, shock at 0.5 + 1.75216*0.25 = 0.938
    np.testing.assert_allclose(rho[0], 1.0)
    np.testing.assert_allclose(rho[1], 0.42632, rtol=2e-4)  # rho*L, left of the contact
    np.testing.assert_allclose(rho[2:], 0.26557, rtol=2e-4)  # rho*R, between contact and shock


@pytest.mark.parametrize("limiter,tol", [(C.LIMITER_MINMOD, 0.012), (C.LIMITER_REFERENCE, 0.012)])
def test_sod_against_exact_riemann(limiter, tol):
    """C1: 400x4 cells, MUSCL order 2, LU-SGS, CFL 0.5, 200 steps."""
    case = C.sod(400, 4, limiter=limiter)
    o = _oracle(case)
    o.prepare()
    for _ in range(200):
        assert o.step()[0] == 0
    t = o.tt()
    xc = (np.arange(400) + 0.5) / 400
    rho = _int(o, "W")[1, :, 0]
    ex = exact_riemann(xc, t, (1.0, 0.0, 1e5), (0.125, 0.0, 1e4))
    l1 = np.mean(np.abs(rho - ex))
    assert l1 < tol, l1
    # shock position within 2.5 cells of the exact one
    jump = np.where(np.abs(np.diff(ex)) > 0.05)[0]
    x_shock = xc[jump[-1]]
    i_half = np.where(rho > 0.5 * (ex[jump[-1]] + 0.125))[0][-1]
    assert abs(xc[i_half] - x_shock) <= 2.5 / 400
