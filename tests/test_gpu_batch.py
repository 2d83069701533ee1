"""Batches of independent blocks (uneven laminar plate/channel cases with a
kinematic viscosity per block): parity with the CPU oracle, identity with
one-block contexts, HIP-graph replays, block freezing and the autotuner."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from shared_cfd_amd import cases as C
from shared_cfd_amd.native import OPT_FAST_IMPLICIT, OPT_TILE_ROWS, Solver

pytestmark = pytest.mark.gpu


def _batch(n=6, seed=3):
    return C.batched_plates(n=n, seed=seed, ni_range=(20, 90), nj_range=(12, 70))


def _q(s, case):
    return [s.download("Q", b)[2:case.nj[b] + 2, 2:case.ni[b] + 2].copy() for b in range(case.nblocks)]


def test_batch_matches_oracle():
    case = _batch()
    o = Oracle(case)
    s = Solver(case)
    o.prepare()
    s.prepare()
    for k in range(4):
        assert o.step()[0] == 0
        s.step(k + 1, 0)
    for b, qs in enumerate(_q(s, case)):
        qo = o.field("Q", b)[2:case.nj[b] + 2, 2:case.ni[b] + 2]
        err = np.max(np.abs(qs - qo)) / np.max(np.abs(qo))
        assert err < 1e-11, (b, err)
    # per-block energy RMS of check_convergence
    r = s.energy_rms_blocks()
    for b in range(case.nblocks):
        ni, nj = case.ni[b], case.nj[b]
        d = o.field("Q", b)[2:nj + 2, 2:ni + 2, 3] - o.field("QP", b)[2:nj + 2, 2:ni + 2, 3]
        ref = np.sqrt(np.sum(d * d)) / (ni * nj)
        assert abs(r[b] - ref) <= 1e-9 * ref + 1e-300


@pytest.mark.parametrize("fast", [1, 0])
def test_batch_equals_one_block_contexts(fast):
    case = _batch(n=5, seed=4)
    s = Solver(case, fast=bool(fast))
    s.prepare()
    s.iterate(1, 3, 0)
    qb = _q(s, case)
    for b in range(case.nblocks):
        sub = C.sub_case(case, [b])
        t = Solver(sub, fast=bool(fast))
        t.prepare()
        t.iterate(1, 3, 0)
        assert np.array_equal(_q(t, sub)[0], qb[b]), b
        t.close()


def test_graph_replays_match_eager_launches():
    case = _batch(n=4, seed=5)
    a = Solver(case)
    a.prepare()
    for k in range(9):
        a.step(k + 1, 0)
    ref = _q(a, case)
    g = Solver(case)
    g.prepare()
    g.iterate(1, 4, 2)        # capture 2 steps, replay twice
    g.step(5, 0)              # eager work between replays
    g.iterate(6, 4, 2)        # replay the same graph twice more
    for b, (x, y) in enumerate(zip(_q(g, case), ref)):
        assert np.array_equal(x, y), b


def test_frozen_blocks_keep_their_state():
    case = _batch(n=4, seed=6)
    s = Solver(case)
    s.prepare()
    s.iterate(1, 2, 0)
    before = _q(s, case)
    s.set_block_active([1, 0, 1, 0])
    s.iterate(3, 3, 0)
    after = _q(s, case)
    assert np.array_equal(after[1], before[1]) and np.array_equal(after[3], before[3])
    assert not np.array_equal(after[0], before[0])


def test_autotune_restores_state_and_picks_a_schedule():
    case = _batch(n=4, seed=7)
    s = Solver(case)
    s.prepare()
    s.iterate(1, 2, 0)
    before = _q(s, case)
    choice = s.autotune(3)
    assert choice["tile_rows"] in (1, 2, 4) and choice["graph_steps"] in (0, 8) and choice["ms_per_step"] > 0
    for x, y in zip(_q(s, case), before):
        assert np.array_equal(x, y)
    # the tuned schedule gives the same answer as the default one
    s.iterate(3, 8, -1)
    ref = Solver(case)
    ref.prepare()
    ref.iterate(1, 10, 0)
    for x, y in zip(_q(s, case), _q(ref, case)):
        assert np.max(np.abs(x - y)) <= 1e-10 * np.max(np.abs(y))


def test_tile_rows_do_not_change_cell_results():
    case = _batch(n=3, seed=8)
    out = []
    for rows in (4, 1, 2):
        s = Solver(case)
        s.set_option(OPT_TILE_ROWS, rows)
        s.prepare()
        s.iterate(1, 3, 0)
        out.append(_q(s, case))
        s.close()
    for o in out[1:]:
        for x, y in zip(o, out[0]):
            assert np.array_equal(x, y)


def test_dual_batch_with_more_strips_than_cus_matches_per_routine_sweeps():
    """Dual time on a batch with more quad strips than the GPU has compute
    units: the optimised update then forms the backward operators in its
    right-hand-side pass instead of on idle CUs, and numbers the strips of
    all blocks interleaved (strip k of every block before strip k+1 of any).
    Two physical steps agree with the per-routine sweeps: the increments to
    1e-11 of their size (the optimised path rounds differently)."""
    case = C.batched_plates(n=48, seed=5)
    nstrips = sum((nj + 15) // 16 for nj in case.nj)
    assert nstrips > 256, nstrips  # an MI355X has 256 CUs
    p = case.params
    p.scheme = C.SCHEME_DUAL
    p.time_step = 0  # dual time steps every cell with the global dt
    p.num_internal_loop = 4
    p.cfl_tau, p.omega_max, p.dqmax = 5.0, 1.0, 0.5
    p.res_rate_warning = p.res_rate_error = 1e300
    f, s = Solver(case, fast=True), Solver(case, fast=False)
    f.prepare()
    s.prepare()
    q0 = _q(s, case)
    for k in range(2):
        rf, rs = f.step(k + 1, 0), s.step(k + 1, 0)
        assert np.isfinite(rf).all() and np.isfinite(rs).all()
    # the increments of the two steps, relative to their size per block
    worst = 0.0
    for qf, qs, qi in zip(_q(f, case), _q(s, case), q0):
        dq = np.max(np.abs(qs - qi))
        assert dq > 0.0
        worst = max(worst, float(np.max(np.abs(qf - qs)) / dq))
    print(f"dual batch, {nstrips} strips: max rel {worst:.3e} of the increments")
    assert worst <= 1e-11
