"""Pins of the oracle where the reference's formula is only restated:
boundary walls, species diffusion in the viscous flux, fall-off rate
constants.  Each check compares the oracle with what geometry or a textbook
fixes, and each has a planted-mistake counterpart (a wrong wall kind, a
dropped factor) that the same check rejects, so the pin is not vacuous.

- Slip wall (REFLECT, sample/no06.step_flow/condition.f90:100-154): with
  first-order face states a block with a slip wall on a straight side
  evolves exactly as the upper half of the mirror-doubled domain.
- Curved-wall slip (SLIP_NORMAL): the same mirror test on a grid rotated by
  an arbitrary angle.
- No-slip wall (sample/no05.boundary_layer_ideal/condition.f90:110-118):
  Couette flow u = U y gives the wall shear mu U exactly.
- Species diffusion of set_TGv
  (store/viscosity/viscous/no-nV/2d/sch_viscous.f90:146-170): linear mass
  fractions on a skewed grid give fluxes rho D grad(Y_k).n, summing to zero,
  and the energy flux sum_k h_k rho D grad(Y_k).n.
- Lindemann and Troe fall-off (store/therm_lib/chemkin/uvhp/uv/FJ.f90:
  224-245): k_f of each fall-off reaction of the synthetic mechanism equals
  the textbook form (Gilbert, Luther and Troe 1983, as in the Chemkin
  manual) evaluated here in numpy from the mechanism tables.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle.chem_inp import read_cheminp
from oracle.oracle import Oracle
from shared_cfd_amd import cases as C
from shared_cfd_amd import chem

from parity import interior, jface

KAPPA, RGAS = 1.4, 287.0


# --------------------------------------------------------------------------- walls
def _skewed_nodes(ni, nj, seed, wall_straight=True):
    """Nodes on [0,1] x [0,0.6], jittered, with the j = 0 row on y = 0."""
    rng = np.random.default_rng(seed)
    s = np.linspace(0, 1, ni + 1)[:, None]
    t = np.linspace(0, 1, nj + 1)[None, :]
    xh = s + 0.03 * np.sin(3 * t) + 0.15 * rng.standard_normal((ni + 1, nj + 1)) / ni
    rh = 0.6 * t + 0.04 * np.sin(np.pi * s) * t + 0.1 * rng.standard_normal((ni + 1, nj + 1)) / nj
    if wall_straight:
        rh[:, 0] = 0.0
    return xh, rh


def _smooth_state(ni, nj, seed):
    rng = np.random.default_rng(seed)
    c0 = np.sqrt(KAPPA * RGAS * 300.0)
    rho = 1.2 * (1 + 0.2 * rng.random((nj, ni)))
    u = 0.4 * c0 * (1 + 0.3 * rng.standard_normal((nj, ni)))
    v = 0.3 * c0 * rng.standard_normal((nj, ni))
    p = rho * RGAS * 300.0 * (1 + 0.1 * rng.random((nj, ni)))
    return rho, u, v, p


def _case(ni, nj, xh, rh, rho, u, v, p, wall_kind, viscous=0):
    q = C._fill_q(ni, nj, C.cons_from_prim(rho, u, v, p, KAPPA))
    rows = [[1, C.SIDE_IMIN, 1, nj, C.BC_EXTRAP, 0], [1, C.SIDE_IMAX, 1, nj, C.BC_EXTRAP, 0],
            [1, C.SIDE_JMIN, 1, ni, wall_kind, 0], [1, C.SIDE_JMAX, 1, ni, C.BC_EXTRAP, 0]]
    params = C.make_params(scheme=C.SCHEME_EULER, time_step=0, high_order=0, viscous=viscous, cfl=0.4,
                           kappa=KAPPA, rgas=RGAS, nu=1e-3)
    return C.Case("pin", [ni], [nj], 1, params, [xh], [rh], [q], np.array(rows, dtype=np.int32), np.zeros(1),
                  np.zeros((0, 11), dtype=np.int32))


def _rotate(x, y, th):
    c, s = np.cos(th), np.sin(th)
    return c * x - s * y, s * x + c * y


def _mirror_pair(ni, nj, seed, theta, wall_kind):
    """(one block with `wall_kind` on its j-min side, the doubled block with
    the mirror image below the wall), both rotated by theta."""
    xh, rh = _skewed_nodes(ni, nj, seed)
    rho, u, v, p = _smooth_state(ni, nj, seed + 1)
    # doubled: rows nj-1 .. 0 of the image (y -> -y, v -> -v), then the block
    xd = np.concatenate([xh[:, :0:-1], xh], axis=1)
    yd = np.concatenate([-rh[:, :0:-1], rh], axis=1)
    fl = lambda a: np.concatenate([a[::-1], a], axis=0)  # noqa: E731
    rd, ud, vd, pd = fl(rho), fl(u), np.concatenate([-v[::-1], v], axis=0), fl(p)
    xr, yr = _rotate(xh, rh, theta)
    ur, vr = _rotate(u, v, theta)
    xdr, ydr = _rotate(xd, yd, theta)
    udr, vdr = _rotate(ud, vd, theta)
    one = _case(ni, nj, xr, yr, rho, ur, vr, p, wall_kind)
    two = _case(ni, 2 * nj, xdr, ydr, rd, udr, vdr, pd, C.BC_EXTRAP)
    return one, two


def _mirror_gap(one, two, steps=2):
    o1, o2 = Oracle(one), Oracle(two)
    o1.prepare()
    o2.prepare()
    for _ in range(steps):
        assert o1.step()[0] == 0 and o2.step()[0] == 0
    ni, nj = one.ni[0], one.nj[0]
    a = interior(o1.field("Q", 0), ni, nj)
    b = interior(o2.field("Q", 0), ni, 2 * nj)[nj:]
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_slip_wall_is_a_mirror():
    one, two = _mirror_pair(14, 6, seed=1, theta=0.0, wall_kind=C.BC_REFLECT)
    assert _mirror_gap(one, two) <= 1e-13
    # planted mistake: the normal velocity not negated (a zero-gradient side)
    one_bad, _ = _mirror_pair(14, 6, seed=1, theta=0.0, wall_kind=C.BC_EXTRAP)
    assert _mirror_gap(one_bad, two) > 1e-4


@pytest.mark.parametrize("theta", [0.37, -1.1])
def test_slip_normal_wall_is_a_mirror_on_a_rotated_grid(theta):
    one, two = _mirror_pair(14, 6, seed=2, theta=theta, wall_kind=C.BC_SLIP_NORMAL)
    assert _mirror_gap(one, two) <= 1e-12
    # planted mistake: only the y velocity negated, as for an axis-aligned wall
    one_bad, _ = _mirror_pair(14, 6, seed=2, theta=theta, wall_kind=C.BC_REFLECT)
    assert _mirror_gap(one_bad, two) > 1e-4


def _couette(wall_kind, ni=8, nj=5, U=300.0, mu=2e-3):
    xh, rh = C.rect_nodes(ni, nj, 0.0, 1.0, 0.0, 0.5)
    yc = (np.arange(nj) + 0.5) * 0.5 / nj
    rho = np.full((nj, ni), 1.2)
    u = np.repeat((U * yc)[:, None], ni, axis=1)
    p = np.full((nj, ni), 1e5)
    case = _case(ni, nj, xh, rh, rho, u, np.zeros_like(u), p, wall_kind, viscous=1)
    case.params.nu = mu / 1.2
    o = Oracle(case)
    o.set_w()
    o.set_thermo_prop()
    o.set_bc()
    o.set_tgv()
    return jface(o.field("TGVJ", 0), ni, nj)[0, 2:ni - 2].copy()   # wall faces away from the corners


def test_noslip_wall_shear_of_couette_flow():
    U, mu = 300.0, 2e-3
    tw = _couette(C.BC_NOSLIP, U=U, mu=mu)
    # (mass, x momentum, y momentum, energy) through the wall face, normal +y:
    # tau_xy = mu dU/dy exactly; no normal stress, wall velocity zero, adiabatic
    np.testing.assert_allclose(tw[:, 1], mu * U, rtol=1e-12)
    np.testing.assert_allclose(tw[:, 0], 0.0, atol=1e-15)
    np.testing.assert_allclose(tw[:, 2], 0.0, atol=1e-12 * mu * U)
    np.testing.assert_allclose(tw[:, 3], 0.0, atol=1e-9 * mu * U * U)
    # planted mistake: a slip wall has no shear
    assert np.max(np.abs(_couette(C.BC_REFLECT, U=U, mu=mu)[:, 1])) < 1e-3 * mu * U


# --------------------------------------------------------------------------- species diffusion
def test_species_diffusion_of_linear_mass_fractions():
    case = C.random_case(seed=3, ni=(12,), nj=(9,), ny=2, viscous=1)
    o = Oracle(case)
    ni, nj = 12, 9
    X, Y = o.field("X", 0)[..., 0], o.field("R", 0)[..., 0]
    w = o.field("W", 0)
    vhi = o.field("VHI", 0)
    b = np.array([0.7, -0.4])     # grad Y_1; Y_2 = 1 - Y_1
    mu, rho, h = 3e-3, 1.1, np.array([2.5e6, -4.0e5])
    Y1 = 0.4 + b[0] * X + b[1] * Y
    w[..., 0] = rho
    w[..., 1] = 0.0
    w[..., 2] = 0.0
    w[..., 3] = rho * RGAS * 300.0
    w[..., 4] = Y1
    w[..., 5] = 1.0 - Y1
    w[..., 6] = KAPPA
    w[..., 8] = RGAS
    w[..., 9] = mu
    vhi[..., 0] = h[0]
    vhi[..., 1] = h[1]
    o.set_tgv()
    for f, vn, sl in (("TGVI", "VNI", np.s_[2:nj + 2, 2:ni + 1]), ("TGVJ", "VNJ", np.s_[2:nj + 1, 2:ni + 2])):
        t = o.field(f, 0)[sl]
        n = o.field(vn, 0)[sl]
        dY1 = b[0] * n[..., 0] + b[1] * n[..., 1]
        scale = mu * np.max(np.abs(dY1))
        # rho D = mu / Sc with Sc = 1
        np.testing.assert_allclose(t[..., 0], mu * dY1, rtol=0, atol=1e-12 * scale)
        np.testing.assert_allclose(t[..., 1], -mu * dY1, rtol=0, atol=1e-12 * scale)
        np.testing.assert_allclose(t[..., 0] + t[..., 1], 0.0, atol=1e-13 * scale)
        np.testing.assert_allclose(t[..., 2:4], 0.0, atol=1e-9 * scale)
        e = mu * (h[0] - h[1]) * dY1      # sum_k h_k rho D grad(Y_k).n
        assert np.max(np.abs(e)) > 1e3 * scale
        np.testing.assert_allclose(t[..., 4], e, rtol=0, atol=1e-10 * np.max(np.abs(e)))


# --------------------------------------------------------------------------- fall-off
_RXN_FIELDS = ["reverse_off", "kind", "num_nu", "ind_nu", "snu", "exist_m", "num_m", "ind_m", "men", "abe", "cabe",
               "troe"]


def _one_reaction(m, i, irreversible=False):
    """A copy of mechanism m holding only reaction i."""
    m1 = type(m)()
    ctypes.memmove(ctypes.byref(m1), ctypes.byref(m), ctypes.sizeof(m))
    for f in _RXN_FIELDS:
        a = getattr(m1, f)
        src = getattr(m, f)[i]
        if isinstance(src, (int, float)):
            a[0] = src
        else:
            ctypes.memmove(ctypes.byref(a, 0), ctypes.byref(src), ctypes.sizeof(src))
    if irreversible:
        m1.reverse_off[0] = 1
    m1.nr = 1
    return m1


def _textbook_kf(m, i, n, T, troe=True):
    """k_f of fall-off reaction i: Lindemann k_inf Pr/(1+Pr), times Troe's F."""
    lnA, beta, ER = m.abe[i][0], m.abe[i][1], m.abe[i][2]
    kinf = np.exp(lnA + beta * np.log(T) - ER / T)
    k0 = np.exp(m.cabe[i][0] + m.cabe[i][1] * np.log(T) - m.cabe[i][2] / T)
    eff = np.ones(m.ns)
    for q in range(m.num_m[i]):          # men = efficiency - 1 of the q-th listed species
        eff[m.ind_m[i][q]] += m.men[i][q]
    M = float(np.dot(eff, n[:m.ns]))
    Pr = k0 * M / kinf
    k = kinf * Pr / (1.0 + Pr)
    if troe and m.kind[i] == chem.RXN_TROE:
        a, t3, t1, t2 = m.troe[i][0], 1.0 / m.troe[i][1], 1.0 / m.troe[i][2], m.troe[i][3]
        Fc = (1 - a) * np.exp(-T / t3) + a * np.exp(-T / t1) + np.exp(-t2 / T)
        c = -0.4 - 0.67 * np.log10(Fc)
        nn = 0.75 - 1.27 * np.log10(Fc)
        f1 = (np.log10(Pr) + c) / (nn - 0.14 * (np.log10(Pr) + c))
        k = k * 10.0 ** (np.log10(Fc) / (1.0 + f1 * f1))
    return k


def _oracle_kf(m1, n, T):
    """k_f from the oracle's Fex_var of the one-reaction mechanism: the
    production rate of a species made only on the product side (the forward
    direction alone) divided by the reactant concentration product."""
    p, _ = O.chem_fex_var(m1, np.append(n[:m1.ns], T))
    rs = [m1.ind_nu[0][0][k] for k in range(m1.num_nu[0][0])]
    ps = [m1.ind_nu[0][1][k] for k in range(m1.num_nu[0][1])]
    sp = next(s for s in ps if s not in rs)
    return p[sp] / (ps.count(sp) * np.prod([n[s] for s in rs]))


@pytest.fixture(scope="module")
def synthetic(tmp_path_factory):
    d = tmp_path_factory.mktemp("fo")
    c, t = chem.write_synthetic_mechanism(11, str(d / "syn"))
    return read_cheminp(c, chem.MW_INP, t)


def test_falloff_rate_constants_match_the_textbook_form(synthetic):
    m = synthetic
    fo = [i for i in range(m.nr) if m.kind[i] in (chem.RXN_TROE, chem.RXN_LOW)]
    assert sorted(m.kind[i] for i in fo) == [chem.RXN_LOW, chem.RXN_TROE, chem.RXN_TROE]
    rng = np.random.default_rng(7)
    worst, troe_effect = 0.0, 0.0
    for i in fo:
        m1 = _one_reaction(m, i)
        for _ in range(40):
            T = rng.uniform(700.0, 2800.0)
            n = 10.0 ** rng.uniform(-9, -4, m.ns)        # mol/cm^3: Pr spans the fall-off region
            k_or = _oracle_kf(m1, n, T)
            k_tb = _textbook_kf(m, i, n, T)
            worst = max(worst, abs(k_or / k_tb - 1.0))
            if m.kind[i] == chem.RXN_TROE:
                troe_effect = max(troe_effect, abs(_textbook_kf(m, i, n, T, troe=False) / k_tb - 1.0))
    assert worst <= 1e-12, worst
    # planted mistake: Lindemann's form for a Troe reaction misses by far more
    assert troe_effect > 1e-2


def test_irreversible_falloff_reaction_keeps_the_high_pressure_limit(synthetic):
    """A quirk of the reference that the oracle keeps: the fall-off factor is
    evaluated only for reversible reactions (FJ.f90:229-234 sits inside the
    bi-directional branch), so an irreversible (+M) fall-off reaction runs at
    k_inf."""
    m = synthetic
    i = next(i for i in range(m.nr) if m.kind[i] == chem.RXN_TROE)
    m1 = _one_reaction(m, i, irreversible=True)
    T, n = 1400.0, np.full(m.ns, 1e-6)
    kinf = np.exp(m.abe[i][0] + m.abe[i][1] * np.log(T) - m.abe[i][2] / T)
    assert abs(_oracle_kf(m1, n, T) / kinf - 1.0) <= 1e-13


# --------------------------------------------------------------------------- inflow / outflow
def test_inflow_tuv_and_outflow_p_ghost_states():
    """The subsonic inflow of sample/no05.boundary_layer_ideal/condition.f90
    (T, u, v given, p from the interior) and a fixed-pressure outflow: the
    ghost states are the ideal-gas states those data fix, H = cp T + |V|^2/2
    with cp = gamma R/(gamma - 1), in both ghost layers."""
    ni, nj = 6, 4
    xh, rh = C.rect_nodes(ni, nj, 0.0, 1.0, 0.0, 0.5)
    rho, u, v, p = _smooth_state(ni, nj, 5)
    case = _case(ni, nj, xh, rh, rho, u, v, p, C.BC_EXTRAP)
    Tin, uin, vin, pout = 310.0, 120.0, -7.0, 0.93e5
    case.bc_rows = np.array([[1, C.SIDE_IMIN, 1, nj, C.BC_INFLOW_TUV, 0], [1, C.SIDE_IMAX, 1, nj, C.BC_OUTFLOW_P, 3],
                             [1, C.SIDE_JMIN, 1, ni, C.BC_EXTRAP, 0], [1, C.SIDE_JMAX, 1, ni, C.BC_EXTRAP, 0]],
                            dtype=np.int32)
    case.bc_pool = np.array([Tin, uin, vin, pout])
    case.params.nu = 1.5e-5
    o = Oracle(case)
    o.set_w()
    o.set_thermo_prop()
    o.set_bc()
    w = o.field("W", 0)
    for gi in (1, 0):                          # ghost columns i = 0 and i = -1
        g, it = w[2:nj + 2, gi], w[2:nj + 2, 2]
        np.testing.assert_allclose(g[:, 3], it[:, 3], rtol=1e-15)            # p from the interior
        np.testing.assert_allclose(g[:, 0], it[:, 3] / (RGAS * Tin), rtol=1e-14)
        np.testing.assert_allclose(g[:, 1:3], np.broadcast_to([uin, vin], (nj, 2)), rtol=1e-15)
        cp = KAPPA * RGAS / (KAPPA - 1.0)
        np.testing.assert_allclose(g[:, 6], cp * Tin + 0.5 * (uin ** 2 + vin ** 2), rtol=1e-14)
        np.testing.assert_allclose(g[:, 8], 1.5e-5 * g[:, 0], rtol=1e-14)      # mu = rho nu
    for gi in (ni + 2, ni + 3):                # ghost columns i = ni+1, ni+2
        g, it = w[2:nj + 2, gi], w[2:nj + 2, ni + 1]
        np.testing.assert_allclose(g[:, 3], pout, rtol=1e-15)
        np.testing.assert_allclose(g[:, [0, 1, 2]], it[:, [0, 1, 2]], rtol=1e-15)


# --------------------------------------------------------------------------- the reference's own solution
def _no27(**kw):
    st = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "no27_state.npz"))
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    m = chem.parse(os.path.join(golden, "no27_chem.inp"), os.path.join(golden, "no27_trans.inp"))
    return C.no27_cold(m, st, **kw)


def _rms_steps(case, n):
    o = Oracle(case)
    o.prepare()
    out = []
    for _ in range(n):
        assert o.step()[0] == 0
        out.append(o.energy_rms())
    return out


def test_reference_no27_state_is_a_near_fixed_point():
    """The state the reference wrote after 40000 steps of its no27
    cold-flow initialisation (8 species, viscous, preconditioned LU-SGS,
    local time steps, first order; sample/no27.chemkin_reaction/restart.bin,
    checkout.ini.inp, control.ini.inp), continued by the oracle with the
    sample's boundary conditions (condition.f90): the energy RMS of
    check_convergence (store/core/check_convergence.f90:22-28) is small and
    keeps falling, as a continuation of a converging run.  A planted mistake
    -- the fuel stream 10 K too warm -- moves the state off that fixed point
    and raises the RMS by more than four orders of magnitude."""
    r = _rms_steps(_no27(), 10)
    assert r[0] < 3e-5
    assert all(b < a for a, b in zip(r, r[1:]))
    assert _rms_steps(_no27(T_f=310.0), 1)[0] > 1e4 * r[0]
