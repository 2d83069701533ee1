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


def build(force: bool = False) -> str:
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
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
        L.or_copy_field.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.or_calc_next_step_implicit.argtypes = [vp, ctypes.c_int, dp, dp]
        L.or_prepare.argtypes = [vp]
        L.or_step.argtypes = [vp, dp, dp]
        L.or_omp_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


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
