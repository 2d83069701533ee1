"""shcfd_chem_eval_noalloc (device scratch, no allocation) against
shcfd_chem_eval (per-call allocations) through the C ABI: the same kernels
on the same rows, so outputs and status words must be identical.  Two calls
in a row, the second with fewer states, so that the shared scratch holds the
first call's contents when the second one runs."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from shared_cfd_amd import chem

from chem_cases import NO27, random_states

pytestmark = pytest.mark.gpu

E_TOO_LARGE = 11


@pytest.fixture(scope="module")
def mechs(tmp_path_factory):
    d = tmp_path_factory.mktemp("na")
    c, t = chem.write_synthetic_mechanism(11, str(d / "syn"))
    return {"no27": chem.parse(NO27), "syn": chem.parse(c, t)}


def _call(fn, m, op, x, dt):
    ldi, ldo = chem.row_sizes(op, m.ns)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, ldi)
    n = x.shape[0]
    out = np.full((n, ldo), np.nan)
    st = np.full(n, -99, dtype=np.int32)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = fn(ctypes.byref(m), op, n, x.ctypes.data_as(dp), out.ctypes.data_as(dp),
            st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.c_double(dt))
    return rc, out, st


def _states(m, op, n, seed):
    if op in (chem.FEX, chem.FEXVAR, chem.JEX):
        return random_states(m, n, seed)
    # H2/air-like mixtures with radicals at 1e-4 of the mixture, 900-1500 K
    rng = np.random.default_rng(seed)
    base = np.full(m.ns, 1e-4)
    base[:3] = (0.7, 0.2, 0.03)
    base[-1] = 0.01
    rows = []
    for _ in range(n):
        rel = base * rng.uniform(0.5, 1.5, m.ns)
        T = rng.uniform(900.0, 1500.0)
        vrho, rho, E = O.chem_rho_rel2abs(m, 1.01325e5, T, rel)
        rows.append(np.concatenate([vrho, [rho, E, 0.9 * T + 50.0]]) if op == chem.CALCT else np.append(vrho, T))
    return np.array(rows)


@pytest.mark.parametrize("which", ["no27", "syn"])
@pytest.mark.parametrize("op", [chem.JEX, chem.CALCT, chem.POINT_IMPLICIT, chem.CHEMEQ2])
def test_noalloc_equals_allocating_entry_point(mechs, which, op):
    L = chem._lib()
    m = mechs[which]
    m.rtol = 1e-6
    for call, n in enumerate((1500, 300)):
        x = _states(m, op, n, seed=10 * op + call)
        print(f"noalloc {which} op={op} call={call} n={n}", flush=True)
        rc_a, out_a, st_a = _call(L.shcfd_chem_eval, m, op, x, 1e-7)
        rc_b, out_b, st_b = _call(L.shcfd_chem_eval_noalloc, m, op, x, 1e-7)
        assert rc_a == 0 and rc_b == 0
        assert np.array_equal(st_a, st_b)
        assert np.array_equal(out_a, out_b, equal_nan=True)
        assert np.all(st_b != -99)


def test_noalloc_rejects_requests_beyond_its_scratch(mechs):
    L = chem._lib()
    m = mechs["syn"]
    ldi, ldo = chem.row_sizes(chem.JEX, m.ns)
    n = (16 << 20) // (8 * (ldi + ldo)) + 1
    x = np.zeros((n, ldi))
    rc, out, st = _call(L.shcfd_chem_eval_noalloc, m, chem.JEX, x, 0.0)
    assert rc == E_TOO_LARGE
    assert np.all(st == -99)          # nothing was computed or copied back
