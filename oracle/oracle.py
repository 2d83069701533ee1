"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py to check and time the HIP library.  The
product path (shared_cfd_amd) never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_DVODE = None
# the reference's own DVODE, compiled from store/therm_lib/chemkin/core/dvode.f
# by `make ref` wherever the reference's sources are present
DVODE_REF = os.path.join(_HERE, "_ref", "libdvode_ref.so")
REFERENCE = "/root/reference"


def build(force: bool = False) -> str:
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    if os.path.isdir(REFERENCE) and (force or not os.path.exists(DVODE_REF)):
        subprocess.run(["make", "-s", "-C", _HERE, "ref", "REF=" + REFERENCE] + (["-B"] if force else []),
                       check=True)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = ctypes.CDLL(so)
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.or_create.restype = vp
        L.or_create.argtypes = [ctypes.c_int, ip, ip, ctypes.c_int, ctypes.c_void_p]
        L.or_destroy.argtypes = [vp]
        L.or_set_params.argtypes = [vp, ctypes.c_void_p]
        L.or_field.restype = dp
        L.or_field.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.or_field_ncomp.argtypes = [vp, ctypes.c_int]
        L.or_set_cut_table.argtypes = [vp, ip, ctypes.c_int]
        L.or_set_bc_table.argtypes = [vp, ip, ctypes.c_int, dp, ctypes.c_int]
        for name in ("or_init_geometry", "or_set_vnc_dsc", "or_set_geojac", "or_set_w", "or_set_thermo_prop",
                     "or_set_bc", "or_set_hv", "or_set_tg", "or_set_tgv", "or_set_abpm", "or_set_dqdt",
                     "or_calc_next_step"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = None
        L.or_calc_next_step_rk2.argtypes = [vp, ctypes.c_int]
        L.or_set_dt.argtypes = [vp]
        L.or_set_dt.restype = ctypes.c_double
        L.or_energy_rms.argtypes = [vp]
        L.or_energy_rms.restype = ctypes.c_double
        L.or_dt_grbl.argtypes = [vp]
        L.or_dt_grbl.restype = ctypes.c_double
        L.or_tt.argtypes = [vp]
        L.or_tt.restype = ctypes.c_double
        L.or_set_first_dual.argtypes = [vp, ctypes.c_int]
        L.or_set_fixed_dt.argtypes = [vp, ctypes.c_double]
        L.or_copy_field.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.or_calc_next_step_implicit.argtypes = [vp, ctypes.c_int, dp, dp]
        L.or_prepare.argtypes = [vp]
        L.or_step.argtypes = [vp, dp, dp]
        L.or_omp_threads.restype = ctypes.c_int
        mp = ctypes.c_void_p
        L.or_set_mechanism.argtypes = [vp, mp]
        L.or_set_block_nu.argtypes = [vp, dp]
        L.or_chem_error.argtypes = [vp]
        L.or_proceed_reaction.argtypes = [vp]
        L.or_chem_fex.argtypes = [mp, dp, dp]
        L.or_chem_fex.restype = None
        L.or_chem_fex_var.argtypes = [mp, dp, dp, dp]
        L.or_chem_fex_var.restype = None
        L.or_chem_jex.argtypes = [mp, dp, dp]
        L.or_chem_jex.restype = None
        L.or_chem_rho_rel2abs.argtypes = [mp, ctypes.c_double, ctypes.c_double, dp, dp, dp]
        L.or_chem_rho_rel2abs.restype = None
        L.or_chem_calc_t.argtypes = [mp, dp, ctypes.c_double, ctypes.c_double, dp, dp]
        L.or_cold_flow.argtypes = [mp, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp, dp]
        L.or_set_nasa9.argtypes = [vp, mp]
        L.or_chem_chemeq2.argtypes = [mp, dp, ctypes.c_double, ctypes.c_double, ip]
        L.or_chem_reaction_chemeq2.argtypes = [mp, ctypes.c_double, dp, ctypes.c_double, ip]
        L.or_chem_point_implicit.argtypes = [mp, ctypes.c_double, dp, ctypes.c_double, ip, dp]
        L.or_dgefa.argtypes = [dp, ctypes.c_int, ip]
        L.or_dgesl.argtypes = [dp, ctypes.c_int, ip, dp]
        L.or_dgesl.restype = None
        L.or_exp.argtypes = [ctypes.c_double]
        L.or_exp.restype = ctypes.c_double
        L.or_log.argtypes = [ctypes.c_double]
        L.or_log.restype = ctypes.c_double
        L.or_chem_reaction_dvode.argtypes = [mp, ctypes.c_double, dp, ctypes.c_double, ip]
        L.or_set_dvode.argtypes = [vp]
        L.or_set_dvode.restype = None
        L.or_dvode_calls.restype = ctypes.c_long
        if os.path.exists(DVODE_REF):
            global _DVODE
            _DVODE = ctypes.CDLL(DVODE_REF)
            L.or_set_dvode(ctypes.cast(_DVODE.dvode_, ctypes.c_void_p))
        _LIB = L
    return _LIB


def have_dvode() -> bool:
    """True when the reference's DVODE (oracle/_ref) is loaded: CHEMEQ2's
    hand-over and the DVODE reaction need it."""
    return bool(lib().or_have_dvode())


def dvode_calls() -> int:
    """Number of DVODE integrations the oracle has run in this process."""
    return int(lib().or_dvode_calls())


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """One oracle context holding a whole case (all blocks)."""

    def __init__(self, case, init_geometry=True, load_q=True):
        from shared_cfd_amd import cases as C  # case data only (no device code)

        self.C = C
        self.case = case
        L = lib()
        self.L = L
        self.ni = np.array(case.ni, dtype=np.int32)
        self.nj = np.array(case.nj, dtype=np.int32)
        self.params = case.params
        self.ctx = L.or_create(len(case.ni), _ip(self.ni), _ip(self.nj), case.ny, ctypes.byref(self.params))
        if not self.ctx:
            raise RuntimeError("or_create failed")
        for b in range(case.nblocks):
            xh, rh = case.nodes_ref(b)
            self.field("XH", b)[...] = xh
            self.field("RH", b)[...] = rh
            if load_q:
                self.field("Q", b)[...] = case.q0[b]
        cut = np.ascontiguousarray(case.cut_rows, dtype=np.int32)
        L.or_set_cut_table(self.ctx, _ip(cut), int(cut.shape[0]))
        bc = np.ascontiguousarray(case.bc_rows, dtype=np.int32)
        pool = np.ascontiguousarray(case.bc_pool, dtype=np.float64)
        L.or_set_bc_table(self.ctx, _ip(bc), int(bc.shape[0]), _dp(pool), int(pool.size))
        if init_geometry:
            L.or_init_geometry(self.ctx)
        self.mech = None
        if getattr(case, "mech", None) is not None:
            from oracle.chem_inp import OrMech

            # the oracle keeps its own copy of the tables (same C layout)
            self.mech = OrMech.from_buffer_copy(bytes(case.mech))
            if L.or_set_mechanism(self.ctx, ctypes.byref(self.mech)) != 0:
                raise RuntimeError("or_set_mechanism failed")
        self.nasa9 = None
        if getattr(case, "nasa9", None) is not None:
            self.nasa9 = type(case.nasa9).from_buffer_copy(bytes(case.nasa9))  # the oracle's own copy
            if L.or_set_nasa9(self.ctx, ctypes.byref(self.nasa9)) != 0:
                raise RuntimeError("or_set_nasa9 failed")
        if getattr(case, "w0", None) is not None:
            for b in range(case.nblocks):
                self.field("W", b)[...] = case.w0[b]
        if getattr(case, "block_nu", None) is not None:
            self._bnu = np.ascontiguousarray(case.block_nu, dtype=np.float64)
            L.or_set_block_nu(self.ctx, _dp(self._bnu))

    def __del__(self):
        try:
            if self.ctx:
                self.L.or_destroy(self.ctx)
                self.ctx = None
        except Exception:
            pass

    def set_params(self, params):
        self.params = params
        self.L.or_set_params(self.ctx, ctypes.byref(params))

    def field(self, name, b):
        """numpy view (nj+4, ni+4, K) of field `name`, block b (writable)."""
        fid = self.C.F[name]
        k = self.L.or_field_ncomp(self.ctx, fid)
        ni, nj = int(self.ni[b]), int(self.nj[b])
        ptr = self.L.or_field(self.ctx, fid, b)
        return np.ctypeslib.as_array(ptr, shape=(nj + 4, ni + 4, k))

    def __getattr__(self, name):
        # or_<routine> passthroughs: o.set_w(), o.set_tg(), ...
        L = self.__dict__.get("L")
        if L is not None and hasattr(L, "or_" + name):
            fn = getattr(L, "or_" + name)
            return lambda *a: fn(self.ctx, *a)
        raise AttributeError(name)

    def step(self):
        r1, r2 = ctypes.c_double(0.0), ctypes.c_double(0.0)
        err = self.L.or_step(self.ctx, ctypes.byref(r1), ctypes.byref(r2))
        return err, r1.value, r2.value

    def implicit(self, step_internal=1):
        res, om = ctypes.c_double(0.0), ctypes.c_double(0.0)
        err = self.L.or_calc_next_step_implicit(self.ctx, step_internal, ctypes.byref(res), ctypes.byref(om))
        return err, res.value, om.value


# ---- 0-D chemistry (per state; n = ns concentrations then T) --------------
def chem_fex(m, n):
    n = np.ascontiguousarray(n, dtype=np.float64)
    out = np.zeros(m.ns + 1)
    lib().or_chem_fex(ctypes.byref(m), _dp(n), _dp(out))
    return out


def chem_fex_var(m, n):
    n = np.ascontiguousarray(n, dtype=np.float64)
    p, q = np.zeros(m.ns + 1), np.zeros(m.ns + 1)
    lib().or_chem_fex_var(ctypes.byref(m), _dp(n), _dp(p), _dp(q))
    return p, q


def chem_jex(m, n):
    """Jacobian d(dndt_a)/d(n_b) as a (ns+1, ns+1) array indexed [a, b]."""
    n = np.ascontiguousarray(n, dtype=np.float64)
    N = m.ns + 1
    J = np.zeros(N * N)
    lib().or_chem_jex(ctypes.byref(m), _dp(n), _dp(J))
    return J.reshape(N, N, order="F")


def chem_rho_rel2abs(m, p, T, vrho_rel):
    v = np.array(vrho_rel, dtype=np.float64)
    rho, E = ctypes.c_double(), ctypes.c_double()
    lib().or_chem_rho_rel2abs(ctypes.byref(m), p, T, _dp(v), ctypes.byref(rho), ctypes.byref(E))
    return v, rho.value, E.value


def cold_flow(t, y0, y1, E, T0):
    """(T, out(7) = MWave, kappa, mu, DHi(2), vhi(2), iterations) of the
    oracle's cold_flow restatement."""
    T = ctypes.c_double(T0)
    out = np.zeros(7)
    it = lib().or_cold_flow(ctypes.byref(t), y0, y1, E, ctypes.byref(T), _dp(out))
    return T.value, out, it


def chem_calc_t(m, vrho, rho, E, T0):
    v = np.ascontiguousarray(vrho, dtype=np.float64)
    T = ctypes.c_double(T0)
    out = np.zeros(3 + 2 * m.ns)
    it = lib().or_chem_calc_t(ctypes.byref(m), _dp(v), rho, E, ctypes.byref(T), _dp(out))
    return T.value, out, it


def chem_reaction_chemeq2(m, T, vrho, dt):
    v = np.array(vrho, dtype=np.float64)
    att = ctypes.c_int(0)
    e = lib().or_chem_reaction_chemeq2(ctypes.byref(m), T, _dp(v), dt, ctypes.byref(att))
    return e, v, att.value


def chem_reaction_dvode(m, T, vrho, dt):
    """reaction (sub_chem.f90:1236-1315): (error, vrho, DVODE steps)."""
    v = np.array(vrho, dtype=np.float64)
    st = ctypes.c_int(0)
    e = lib().or_chem_reaction_dvode(ctypes.byref(m), T, _dp(v), dt, ctypes.byref(st))
    return e, v, st.value


def chem_point_implicit(m, T, vrho, dt):
    v = np.array(vrho, dtype=np.float64)
    it = ctypes.c_int(0)
    T1 = ctypes.c_double(0.0)
    e = lib().or_chem_point_implicit(ctypes.byref(m), T, _dp(v), dt, ctypes.byref(it), ctypes.byref(T1))
    return e, v, it.value, T1.value


def dgefa_dgesl(a, b):
    a = np.array(a, dtype=np.float64, order="F")
    n = a.shape[0]
    ipvt = np.zeros(n, dtype=np.int32)
    x = np.array(b, dtype=np.float64)
    info = lib().or_dgefa(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, _ip(ipvt))
    if info == 0:
        lib().or_dgesl(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, _ip(ipvt), _dp(x))
    return info, x
