"""DVODE on the GPU: CHEMEQ2's hand-over of cells that need more than 10
attempts (sub_chem.f90:1751-1770) and the per-cell DVODE reaction
(:1236-1315), in the 0-D harness and in a reacting flow.

Checked against the host build of the same code (tests/host/chem_host.cpp:
the same bits) and against the oracle, which calls the reference's own
DVODE (oracle/_ref).  The library evaluates DVODE's ** with the chemistry's
exp and log, the reference with its runtime's pow; host tests show that with
the same pow the restatement takes the reference's steps one for one, so
here the two differ by step-size roundings amplified by the error
estimate's cancellation: relative 1e-7 at rtol = 1e-10 on species above
1e-12 of the largest (measured: a few 1e-9)."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle.chem_inp import read_cheminp
from oracle.oracle import Oracle
from shared_cfd_amd import cases as C
from shared_cfd_amd import chem
from shared_cfd_amd.native import Solver

from chem_cases import NO27, ignition_states, no27_plus_inert
from parity import interior

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(O.DVODE_REF), reason="oracle/_ref (reference DVODE) not built")]
HOST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host", "libchem_host.so")
TOL = 1e-7


def _host_eval(m, op, rows, dt):
    L = ctypes.CDLL(HOST)
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    out = np.zeros((rows.shape[0], m.ns))
    st = np.zeros(rows.shape[0], dtype=np.int32)
    rc = L.host_chem_eval(ctypes.byref(m), op, rows.shape[0], rows.ctypes.data_as(dp), out.ctypes.data_as(dp),
                          st.ctypes.data_as(ip), ctypes.c_double(dt))
    assert rc == 0
    return out, st


def _mechs(kind, tmp_path, rtol, atol):
    if kind == "no27":
        lm, om = chem.parse(NO27), read_cheminp(NO27, chem.MW_INP)
    else:
        p, _ = no27_plus_inert(str(tmp_path / kind), int(kind.split("+")[1]))
        lm, om = chem.parse(p), read_cheminp(p, chem.MW_INP)
    lm.rtol = om.rtol = rtol
    lm.atol = om.atol = atol
    return lm, om


def _rel(a, b):
    big = b > 1e-12 * b.max()
    return float(np.max(np.abs(a - b)[big] / b[big]))


@pytest.mark.parametrize("kind", ["no27", "no27+4", "no27+8"])
@pytest.mark.parametrize("op", [chem.DVODE, chem.CHEMEQ2])
def test_zero_d_dvode_matches_host_build_and_reference(tmp_path, kind, op):
    lm, om = _mechs(kind, tmp_path, 1e-10, 1e-22)
    states = ignition_states(om, 48, seed=11)
    rows = np.array([np.append(v, T) for v, T in states])
    for dt in (1e-7, 1e-6, 1e-5):
        print(f"0-D {kind} op {op} dt {dt}", flush=True)
        out, st = chem.chem_eval(lm, op, rows, dt)
        hout, hst = _host_eval(lm, op, rows, dt)
        assert np.array_equal(st, hst)
        assert np.array_equal(out, hout)
        if op == chem.CHEMEQ2:
            assert np.all(st == 11)   # every cell is handed to DVODE at this rtol
        else:
            assert np.all(st > 10)
        for q in range(0, rows.shape[0], 4):
            fn = O.chem_reaction_dvode if op == chem.DVODE else O.chem_reaction_chemeq2
            e, v, it = fn(om, rows[q, -1], rows[q, :-1], dt)
            assert e == 0
            assert _rel(out[q], v) <= TOL, (q, dt)


def test_zero_d_chemeq2_with_and_without_hand_over(tmp_path):
    """At rtol = 1e-6 some cells finish within 10 CHEMEQ2 attempts and some
    are handed over; both kinds match the host build bit for bit and the
    reference to the integration tolerance."""
    lm, om = _mechs("no27", tmp_path, 1e-6, 1e-12)
    states = ignition_states(om, 96, seed=12, tmin=900.0)
    rows = np.array([np.append(v, T) for v, T in states])
    out, st = chem.chem_eval(lm, chem.CHEMEQ2, rows, 2e-7)
    hout, hst = _host_eval(lm, chem.CHEMEQ2, rows, 2e-7)
    assert np.array_equal(st, hst) and np.array_equal(out, hout)
    assert np.any(st == 11) and np.any((st >= 1) & (st <= 10))
    for q in range(rows.shape[0]):
        e, v, it = O.chem_reaction_chemeq2(om, rows[q, -1], rows[q, :-1], 2e-7)
        assert e == 0 and it == st[q]
        assert _rel(out[q], v) <= (1e-11 if it <= 10 else 1e-3), (q, it)


def test_flow_chemeq2_hands_cells_to_dvode(tmp_path):
    """A reacting mixing layer under CHEMEQ2 whose hot cells need DVODE: the
    GPU's list-and-drain hand-over against the oracle's per-cell one."""
    lm, om = _mechs("no27", tmp_path, 1e-10, 1e-22)
    names = ["N2", "O2", "H2", "OH", "O", "H", "HO2", "H2O"]

    def rel(d):
        r = np.zeros(len(names))
        for k, v in d.items():
            r[names.index(k)] = v
        return r / r.sum()

    # radical-seeded streams: about half the cells need DVODE each step
    fuel = rel({"H2": 0.2, "N2": 0.8, "H": 1e-4})
    ox = rel({"N2": 0.7, "O2": 0.22, "H2O": 0.05, "OH": 0.01, "O": 0.01, "H": 1e-3, "HO2": 1e-3})
    case = C.mixing_layer(lm, ni=40, nj=24, seed=5, species=names, reaction=1, T_fuel=1000.0, T_ox=1500.0,
                          fuel=fuel, ox=ox)
    o, s = Oracle(case), Solver(case)
    o.prepare()
    s.prepare()
    calls = O.dvode_calls()
    s.chem_iterations(reset=True)
    for k in range(2):
        print(f"flow step {k + 1}", flush=True)
        assert o.step()[0] == 0
        s.step(k + 1, 0)
    handed = O.dvode_calls() - calls
    assert handed > 0.2 * 2 * case.ncells
    iters = s.chem_iterations()
    assert iters > 11 * handed   # the DVODE steps are counted with the attempts
    for b in range(case.nblocks):
        ni, nj = case.ni[b], case.nj[b]
        qs, qo = interior(s.download("Q", b), ni, nj), interior(o.field("Q", b), ni, nj)
        e = float(np.max(np.abs(qs - qo)) / np.max(np.abs(qo)))
        assert e <= TOL, (b, e)
