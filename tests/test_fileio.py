"""On-disk formats and block-interface detection (shared_cfd_amd/fileio.py)
against the reference's own data: its sample grids (Plot3D), its no27
restart.bin (through a committed fixture, and the file itself when the
reference is checked out), and the cut tables the library's cases use."""
import json
import os

import numpy as np
import pytest

from shared_cfd_amd import airfoil as A
from shared_cfd_amd import cases as C
from shared_cfd_amd import fileio as F

HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLES = "/root/reference/sample"
need_ref = pytest.mark.skipif(not os.path.isdir(SAMPLES), reason="reference samples not present")


def _sample_grid(name):
    d = os.path.join(SAMPLES, name)
    return F.read_plot3d(os.path.join(d, [f for f in os.listdir(d) if f.endswith(".x")][0]))


# --------------------------------------------------------------------------- round trips
def test_plot3d_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    xs = [rng.standard_normal((7, 5)), rng.standard_normal((3, 9)) * 1e-7]
    ys = [rng.standard_normal((7, 5)), rng.standard_normal((3, 9)) * 1e5]
    p = str(tmp_path / "g.x")
    F.write_plot3d(p, xs, ys)
    xr, yr = F.read_plot3d(p)
    for a, b in zip(xs + ys, xr + yr):
        assert np.array_equal(a, b)


def test_restart_and_geometry_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    ni, nj, dimq = [5, 3], [4, 6], 4
    blocks = [rng.standard_normal((b + 2, a + 2, dimq + 3)) for a, b in zip(ni, nj)]
    p = str(tmp_path / "restart.bin")
    F.write_restart(p, 123, 0.25, blocks)
    step, tt, out = F.read_restart(p, ni, nj, dimq)
    assert step == 123 and tt == 0.25
    for a, b in zip(blocks, out):
        assert np.array_equal(a, b)
    # the library's own reader of the out_bin layout agrees
    case = C.sod(5, 4)
    case.ni, case.nj = ni, nj
    s2, t2, o2 = C.read_result_bin(p, case)
    assert s2 == 123 and all(np.array_equal(a, b) for a, b in zip(blocks, o2))
    xs = [rng.standard_normal((a + 1, b + 1)) for a, b in zip(ni, nj)]
    ys = [rng.standard_normal((a + 1, b + 1)) for a, b in zip(ni, nj)]
    g = str(tmp_path / "geometry.bin")
    F.write_geometry(g, xs, ys)
    for (a, b, x, y), xe, ye, ae, be in zip(F.read_geometry(g), xs, ys, ni, nj):
        assert (a, b) == (ae, be) and np.array_equal(x, xe) and np.array_equal(y, ye)


def test_split_records(tmp_path):
    """gfortran continuation sub-records (negative length markers) join."""
    import struct
    p = str(tmp_path / "r.bin")
    with open(p, "wb") as f:
        for n, body in ((-4, b"abcd"), (2, b"ef")):
            f.write(struct.pack(">i", n) + body + struct.pack(">i", n))
    assert F.read_records(p) == [b"abcdef"]


# --------------------------------------------------------------------------- interface detection
def test_cmesh_wake_is_a_self_cut():
    ac = A.naca_cmesh(ni=96, nj=24, nwake=12)
    it = F.detect_interfaces([ac.case.xh[0]], [ac.case.rh[0]])
    assert it.cuts.tolist() == ac.case.cut_rows.tolist() == [[4, 12, -1, 1, 1, 1, -1, 1, 1, 96, -1]]
    assert sorted(it.uncovered) == [(1, 0, 1, 24), (1, 1, 1, 24), (1, 2, 13, 84), (1, 3, 1, 96)]


def test_forward_step_case_interface():
    """The two-block forward-step case's hand-written cut row is what the
    detection finds on its own grid."""
    case = C.forward_step(n_per_unit=40, num_internal_loop=3)
    it = F.detect_interfaces(case.xh, case.rh)
    assert it.cuts.tolist() == case.cut_rows.tolist()


@need_ref
def test_reference_forward_step_grid_interface():
    """The reference's sample no06 (forward_step.x, h = 0.01) has the
    topology of the library's forward-step case at the same spacing."""
    xs, ys = _sample_grid("no06.step_flow")
    assert [x.shape for x in xs] == [(61, 101), (241, 81)]
    it = F.detect_interfaces(xs, ys)
    assert it.cuts.tolist() == C.forward_step(n_per_unit=100, num_internal_loop=3).cut_rows.tolist()


@need_ref
def test_reference_n0012_wake():
    xs, ys = _sample_grid("no03.n0012")
    ni = xs[0].shape[0] - 1
    it = F.detect_interfaces(xs, ys)
    assert it.cuts.shape == (1, 11)
    geo, width, order, p1, c1, s1, pm1, p2, c2, s2, pm2 = it.cuts[0]
    assert (geo, order, p1, c1, s1, pm1, p2, c2, s2, pm2) == (4, -1, 1, 1, 1, -1, 1, 1, ni, -1)
    assert 0 < width < ni // 2


def _cell_centres(x, y):
    xc = 0.25 * (x[:-1, :-1] + x[1:, :-1] + x[:-1, 1:] + x[1:, 1:])
    yc = 0.25 * (y[:-1, :-1] + y[1:, :-1] + y[:-1, 1:] + y[1:, 1:])
    return xc.T, yc.T  # (nj, ni)


def _step_case(xs, ys):
    """Inviscid explicit-Euler case on a grid: all uncovered boundary
    stretches FIXED at the free stream, a smooth initial perturbation that
    depends on the cell centre only (so every decomposition starts alike)."""
    kap, R = 1.4, 287.0
    it = F.detect_interfaces(xs, ys)
    q = []
    for x, y in zip(xs, ys):
        xc, yc = _cell_centres(x, y)
        rho = 1.2 * (1.0 + 0.05 * np.sin(2.0 * xc) * np.cos(3.0 * yc))
        u = 200.0 + 20.0 * np.cos(xc + yc)
        v = 10.0 * np.sin(4.0 * yc)
        p = 1.0e5 * (1.0 + 0.02 * np.cos(2.0 * xc - yc))
        q.append(C._fill_q(x.shape[0] - 1, x.shape[1] - 1, C.cons_from_prim(rho, u, v, p, kap)))
    rows = [[p, side, s, e, C.BC_FIXED, 0] for (p, side, s, e) in it.uncovered]
    pool = C.fixed_state_pool(1.2, 200.0, 0.0, 1.0e5, kap, R)
    params = C.make_params(scheme=C.SCHEME_EULER, time_step=0, high_order=1, limiter=C.LIMITER_REFERENCE,
                           cfl=0.4, kappa=kap, rgas=R, nu=0.0)
    ni = [x.shape[0] - 1 for x in xs]
    nj = [x.shape[1] - 1 for x in xs]
    return C.Case("step_grid", ni, nj, 1, params, [x.copy() for x in xs], [y.copy() for y in ys], q,
                  np.array(rows, dtype=np.int32), np.array(pool), it.cuts)


@need_ref
def test_reference_step_decompositions_agree():
    """The reference's forward-step samples no06-no09 cover the same cells
    with different block decompositions and index orientations (geo codes of
    the detected cuts differ).  Explicit steps with the detected cut tables
    give the same state in every cell, matched by cell centre."""
    from oracle.oracle import Oracle

    states = {}
    geos = set()
    for name in ("no06.step_flow", "no07.step_flow2", "no08.step_flow3", "no09.step_flow4"):
        xs, ys = _sample_grid(name)
        case = _step_case(xs, ys)
        geos.update(int(g) for g in case.cut_rows[:, 0])
        o = Oracle(case)
        o.prepare()
        for _ in range(3):
            assert o.step()[0] == 0
        d = {}
        for b in range(case.nblocks):
            xc, yc = _cell_centres(xs[b], ys[b])
            qb = o.field("Q", b)[2:-2, 2:-2]
            for key, val in zip(zip(np.round(xc, 6).ravel(), np.round(yc, 6).ravel()), qb.reshape(-1, 4)):
                d[key] = val
        states[name] = d
    assert len(geos) >= 3  # several cut orientations exercised
    ref = states["no07.step_flow2"]
    for name, d in states.items():
        assert d.keys() == ref.keys(), name
        a = np.array([d[k] for k in ref])
        b = np.array([ref[k] for k in ref])
        assert np.max(np.abs(a - b) / np.abs(b).max(axis=0)) < 1e-11, name


# --------------------------------------------------------------------------- the reference's no27 restart state
def _no27_fixture():
    with open(os.path.join(HERE, "golden", "no27_restart_cells.json")) as f:
        return json.load(f)


def test_no27_restart_state_matches_our_thermo():
    """Cells of the reference's no27 restart.bin (H2 jet into air at 300 K,
    1 bar): the stored mixture gas constant is R_uni sum(Y_k/MW_k) with the
    molecular weights of the sample's chem.inp, and calc_T of (rho_k, e)
    reproduces the stored temperature p/(rho R) and ratio of specific heats.
    The file holds q after the last update but R, gamma, p of the state
    before it (out_bin writes w as the last set_thermo_prop left it), so the
    tight checks use the cells of pure air, where one step changes little;
    the reacting cells agree to what one step of chemistry moves."""
    from oracle import oracle as O
    from shared_cfd_amd import chem

    fx = _no27_fixture()
    m = chem.parse(os.path.join(HERE, "golden", "no27_chem.inp"))
    mw = np.array([m.mw[k] for k in range(m.ns)])
    R_UNI = 8.3144621e3
    nair = 0
    for j, i, c in fx["cells"]:
        c = np.array(c)
        vr, ru, rv, E, R, g, p = c[:8], c[8], c[9], c[10], c[11], c[12], c[13]
        rho = vr.sum()
        e = (E - 0.5 * (ru * ru + rv * rv) / rho) / rho
        T, out, it = O.chem_calc_t(m, vr, rho, e, 300.0)
        assert it > 0
        Tp = p / (rho * R)
        assert abs(T / Tp - 1.0) < 1e-4
        if np.sum(np.abs(vr[2:])) / rho < 1e-12:  # pure N2/O2
            nair += 1
            assert abs(R_UNI * np.sum(vr / rho / mw) / R - 1.0) < 1e-11
            assert abs(T / Tp - 1.0) < 1e-6
            assert abs(out[0] / g - 1.0) < 1e-8
    assert nair > 100


@need_ref
def test_reads_reference_restart():
    fx = _no27_fixture()
    xs, _ = _sample_grid("no27.chemkin_reaction")
    ni, nj = xs[0].shape[0] - 1, xs[0].shape[1] - 1
    assert (ni, nj) == (fx["ni"], fx["nj"])
    step, tt, blocks = F.read_restart(os.path.join(SAMPLES, "no27.chemkin_reaction", "restart.bin"), [ni], [nj], 11)
    assert (step, tt) == (fx["step"], fx["tt"])
    for j, i, c in fx["cells"]:
        assert np.array_equal(blocks[0][j, i], np.array(c))


def test_subrecord_marker_signs(tmp_path):
    """Records split into subrecords carry gfortran's markers: head negative
    when a subrecord follows, tail negative when one precedes."""
    import struct

    path = str(tmp_path / "sub.bin")
    rec = bytes(range(250)) * 4          # 1000 bytes -> 400 + 400 + 200
    F.write_records(path, [rec, b"xyz"], limit=400)
    data = open(path, "rb").read()
    marks, pos = [], 0
    while pos < len(data):
        (h,) = struct.unpack(">i", data[pos:pos + 4])
        (t,) = struct.unpack(">i", data[pos + 4 + abs(h):pos + 8 + abs(h)])
        marks.append((h, t))
        pos += 8 + abs(h)
    assert marks == [(-400, 400), (-400, -400), (200, -200), (3, 3)]
    assert F.read_records(path) == [rec, b"xyz"]
