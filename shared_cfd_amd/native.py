"""ctypes binding of libshcfd.so (the HIP library) and a host-side solver
object that drives it the way the reference's generated main program drives
its Fortran subroutines (store/core/main.body.raw.f90,
store/time_scheme/*/main.main.raw.f90).

There is no CPU fallback: if the library is missing, or no GPU is present,
every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import cases as C

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libshcfd.so")
_LIB = None

# every symbol declared in include/shcfd.h
EXPORTS = [
    "shcfd_create", "shcfd_destroy", "shcfd_set_params", "shcfd_strerror", "shcfd_version",
    "shcfd_upload_field", "shcfd_download_field", "shcfd_set_cut_table", "shcfd_set_bc_table",
    "shcfd_init_geometry", "shcfd_set_time", "shcfd_get_time", "shcfd_synchronize", "shcfd_set_w",
    "shcfd_set_thermo_prop", "shcfd_set_bc", "shcfd_set_hv", "shcfd_set_tg", "shcfd_set_tgv", "shcfd_set_dt",
    "shcfd_set_abpm", "shcfd_set_dqdt", "shcfd_calc_next_step_implicit", "shcfd_calc_next_step",
    "shcfd_calc_next_step_rk2", "shcfd_copy_state", "shcfd_energy_rms", "shcfd_prepare", "shcfd_step",
    "shcfd_check_error", "shcfd_profile_enable", "shcfd_profile_read", "shcfd_set_option",
    "shcfd_mech_parse", "shcfd_set_mechanism", "shcfd_load_mechanism", "shcfd_proceed_reaction",
    "shcfd_chem_boundary_state", "shcfd_chem_eval", "shcfd_chem_eval_noalloc", "shcfd_chem_iterations",
    "shcfd_set_block_nu", "shcfd_set_block_active", "shcfd_energy_rms_blocks", "shcfd_iterate", "shcfd_autotune",
    "shcfd_adjoint_objective", "shcfd_objective", "shcfd_adjoint_upload", "shcfd_adjoint_download",
    "shcfd_adjoint_poolbar", "shcfd_residual_vjp", "shcfd_adjoint_solve", "shcfd_adjoint_solve_gmres",
    "shcfd_set_fixed_dt", "shcfd_snapshot_save", "shcfd_snapshot_load", "shcfd_snapshot_free",
    "shcfd_adjoint_unsteady_reset", "shcfd_adjoint_unsteady_step", "shcfd_adjoint_unsteady_poolbar",
    "shcfd_nasa9_parse", "shcfd_set_nasa9", "shcfd_nasa9_stream", "shcfd_load_nasa9",
]
ADJ_LAMBDA, ADJ_QBAR, ADJ_XBAR, ADJ_XBAR_SUM, ADJ_LAMBDA_LATER = 0, 1, 2, 3, 4
OPT_FAST_IMPLICIT, OPT_SPIN_LIMIT, OPT_SKEW_SWEEP, OPT_TILE_ROWS, OPT_JVP_SWEEP = 1, 2, 3, 4, 5

# kernel classes of shcfd_profile_read (include/shcfd.h)
KCLASS = ["PRIMITIVE", "BC", "HV", "TG", "TGV", "DT", "ABPM", "DQDT", "RHS", "SWEEP_FWD", "SWEEP_BWD",
          "REDUCE", "UPDATE", "EXPLICIT", "RESIDUAL", "PREP", "THERMO_CHEM", "REACTION"]


class ShcfdError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        msg = lib().shcfd_strerror(code).decode() if _LIB is not None else str(code)
        super().__init__(f"{where}: {msg} (code {code})")


def lib():
    """Load libshcfd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run `python -m shared_cfd_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.shcfd_strerror.restype = ctypes.c_char_p
        L.shcfd_strerror.argtypes = [ctypes.c_int]
        L.shcfd_create.argtypes = [ctypes.c_int, ip, ip, ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(vp)]
        L.shcfd_destroy.argtypes = [vp]
        L.shcfd_set_params.argtypes = [vp, ctypes.c_void_p]
        L.shcfd_upload_field.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.shcfd_download_field.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.shcfd_set_cut_table.argtypes = [vp, ip, ctypes.c_int]
        L.shcfd_set_bc_table.argtypes = [vp, ip, ctypes.c_int, dp, ctypes.c_int]
        for n in ("shcfd_init_geometry", "shcfd_synchronize", "shcfd_set_w", "shcfd_set_hv", "shcfd_set_tg",
                  "shcfd_set_tgv", "shcfd_set_abpm", "shcfd_set_dqdt", "shcfd_calc_next_step", "shcfd_prepare",
                  "shcfd_check_error"):
            getattr(L, n).argtypes = [vp]
        L.shcfd_set_thermo_prop.argtypes = [vp, ctypes.c_int]
        L.shcfd_set_bc.argtypes = [vp, ctypes.c_int]
        L.shcfd_set_dt.argtypes = [vp, ctypes.c_int, dp]
        L.shcfd_set_time.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.shcfd_get_time.argtypes = [vp, dp, dp]
        L.shcfd_calc_next_step_implicit.argtypes = [vp, ctypes.c_int, dp, dp]
        L.shcfd_calc_next_step_rk2.argtypes = [vp, ctypes.c_int]
        L.shcfd_copy_state.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.shcfd_energy_rms.argtypes = [vp, dp]
        L.shcfd_step.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.shcfd_profile_enable.argtypes = [vp, ctypes.c_int]
        L.shcfd_set_option.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.shcfd_profile_read.argtypes = [vp, ctypes.c_int, dp, ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
        L.shcfd_adjoint_objective.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                              ctypes.c_double]
        L.shcfd_objective.argtypes = [vp, dp]
        L.shcfd_adjoint_upload.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.shcfd_adjoint_download.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.shcfd_adjoint_poolbar.argtypes = [vp, dp]
        L.shcfd_residual_vjp.argtypes = [vp]
        L.shcfd_adjoint_solve.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp,
                                          ctypes.POINTER(ctypes.c_int)]
        L.shcfd_adjoint_solve_gmres.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp,
                                                ctypes.POINTER(ctypes.c_int)]
        L.shcfd_set_fixed_dt.argtypes = [vp, ctypes.c_double]
        L.shcfd_snapshot_save.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.shcfd_snapshot_load.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.shcfd_snapshot_free.argtypes = [vp]
        L.shcfd_adjoint_unsteady_reset.argtypes = [vp]
        L.shcfd_adjoint_unsteady_step.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp,
                                                  ctypes.POINTER(ctypes.c_int)]
        L.shcfd_adjoint_unsteady_poolbar.argtypes = [vp, dp]
        L.shcfd_set_block_nu.argtypes = [vp, dp]
        L.shcfd_set_block_active.argtypes = [vp, ip]
        L.shcfd_energy_rms_blocks.argtypes = [vp, dp]
        L.shcfd_iterate.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.shcfd_autotune.argtypes = [vp, ctypes.c_int, dp]
        _LIB = L
    return _LIB


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def check(rc, where):
    if rc != 0:
        raise ShcfdError(rc, where)


class Solver:
    """A device-resident case: all blocks' state in HBM, driven through the
    C ABI.  Method names follow the reference's subroutine names."""

    def __init__(self, case: C.Case, load_q: bool = True, fast: bool = True):
        L = lib()
        self.L = L
        self.case = case
        self.params = case.params
        self.ni = np.array(case.ni, dtype=np.int32)
        self.nj = np.array(case.nj, dtype=np.int32)
        h = ctypes.c_void_p()
        check(L.shcfd_create(case.nblocks, _ip(self.ni), _ip(self.nj), case.ny, ctypes.byref(self.params),
                             ctypes.byref(h)), "shcfd_create")
        self.ctx = h
        for b in range(case.nblocks):
            xh, rh = case.nodes_ref(b)
            self.upload("XH", b, xh)
            self.upload("RH", b, rh)
            if load_q:
                self.upload("Q", b, case.q0[b])
        cut = np.ascontiguousarray(case.cut_rows, dtype=np.int32)
        check(L.shcfd_set_cut_table(self.ctx, _ip(cut), int(cut.shape[0])), "shcfd_set_cut_table")
        bc = np.ascontiguousarray(case.bc_rows, dtype=np.int32)
        pool = np.ascontiguousarray(case.bc_pool, dtype=np.float64)
        check(L.shcfd_set_bc_table(self.ctx, _ip(bc), int(bc.shape[0]), _dp(pool), int(pool.size)),
              "shcfd_set_bc_table")
        if case.mech is not None:
            from . import chem

            chem._lib()
            check(L.shcfd_set_mechanism(self.ctx, ctypes.byref(case.mech)), "shcfd_set_mechanism")
        if case.nasa9 is not None:
            from . import nasa9 as _n9

            check(_n9._lib().shcfd_set_nasa9(self.ctx, ctypes.byref(case.nasa9)), "shcfd_set_nasa9")
        if case.w0 is not None and load_q:
            for b in range(case.nblocks):
                self.upload("W", b, case.w0[b])
        if case.block_nu is not None:
            bnu = np.ascontiguousarray(case.block_nu, dtype=np.float64)
            check(L.shcfd_set_block_nu(self.ctx, _dp(bnu)), "shcfd_set_block_nu")
        check(L.shcfd_init_geometry(self.ctx), "shcfd_init_geometry")
        self.set_option(OPT_FAST_IMPLICIT, int(fast))

    def set_option(self, option, value):
        check(self.L.shcfd_set_option(self.ctx, option, int(value)), "shcfd_set_option")

    def close(self):
        if getattr(self, "ctx", None):
            self.L.shcfd_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        self.params = params
        check(self.L.shcfd_set_params(self.ctx, ctypes.byref(params)), "shcfd_set_params")

    def upload(self, name, b, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        k = C.field_ncomp(name, self.case.ny)
        assert a.shape == (self.case.nj[b] + 4, self.case.ni[b] + 4, k), (name, a.shape)
        check(self.L.shcfd_upload_field(self.ctx, C.F[name], b, _dp(a)), "shcfd_upload_field")

    def download(self, name, b):
        k = C.field_ncomp(name, self.case.ny)
        a = np.zeros((self.case.nj[b] + 4, self.case.ni[b] + 4, k), dtype=np.float64)
        check(self.L.shcfd_download_field(self.ctx, C.F[name], b, _dp(a)), "shcfd_download_field")
        return a

    # ---- per-routine drop-ins (reference names) ----
    def set_w(self):
        check(self.L.shcfd_set_w(self.ctx), "set_w")

    def set_thermo_prop(self, step=0):
        check(self.L.shcfd_set_thermo_prop(self.ctx, step), "set_thermo_prop")

    def set_BC(self, step=0):
        check(self.L.shcfd_set_bc(self.ctx, step), "set_BC")

    def proceed_reaction(self):
        check(self.L.shcfd_proceed_reaction(self.ctx), "proceed_reaction")
        check(self.L.shcfd_check_error(self.ctx), "proceed_reaction")

    # ---- batches of independent blocks ----
    def set_block_active(self, active):
        a = np.ascontiguousarray(active, dtype=np.int32)
        check(self.L.shcfd_set_block_active(self.ctx, _ip(a)), "set_block_active")

    def energy_rms_blocks(self):
        r = np.zeros(self.case.nblocks)
        check(self.L.shcfd_energy_rms_blocks(self.ctx, _dp(r)), "energy_rms_blocks")
        return r

    def iterate(self, first_step, nsteps, graph_steps=-1):
        check(self.L.shcfd_iterate(self.ctx, int(first_step), int(nsteps), int(graph_steps)), "iterate")

    def autotune(self, trial_steps=4):
        c = np.zeros(4)
        check(self.L.shcfd_autotune(self.ctx, int(trial_steps), _dp(c)), "autotune")
        return {"fast": int(c[0]), "tile_rows": int(c[1]), "graph_steps": int(c[2]), "ms_per_step": float(c[3])}

    def chem_iterations(self, reset=True):
        """Newton iterations / CHEMEQ2 attempts summed over cells since the last reset."""
        v = ctypes.c_double(0.0)
        self.L.shcfd_chem_iterations.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        check(self.L.shcfd_chem_iterations(self.ctx, ctypes.byref(v), int(reset)), "chem_iterations")
        return v.value

    def set_HV(self):
        check(self.L.shcfd_set_hv(self.ctx), "set_HV")

    def set_TG(self):
        check(self.L.shcfd_set_tg(self.ctx), "set_TG")

    def set_TGv(self):
        check(self.L.shcfd_set_tgv(self.ctx), "set_TGv")

    def set_dt(self, step=0):
        v = ctypes.c_double(0.0)
        check(self.L.shcfd_set_dt(self.ctx, step, ctypes.byref(v)), "set_dt")
        return v.value

    def set_ABpm(self):
        check(self.L.shcfd_set_abpm(self.ctx), "set_ABpm")

    def set_dqdt(self):
        check(self.L.shcfd_set_dqdt(self.ctx), "set_dqdt")

    def calc_next_step_implicit(self, step_internal=1):
        r, o = ctypes.c_double(0.0), ctypes.c_double(0.0)
        check(self.L.shcfd_calc_next_step_implicit(self.ctx, step_internal, ctypes.byref(r), ctypes.byref(o)),
              "calc_next_step_implicit")
        return r.value, o.value

    def calc_next_step(self):
        check(self.L.shcfd_calc_next_step(self.ctx), "calc_next_step")

    def calc_next_step_rk2(self, stage):
        check(self.L.shcfd_calc_next_step_rk2(self.ctx, stage), "calc_next_step_rk2")

    def copy_state(self, dst, src):
        check(self.L.shcfd_copy_state(self.ctx, C.F[dst], C.F[src]), "copy_state")

    def energy_rms(self):
        v = ctypes.c_double(0.0)
        check(self.L.shcfd_energy_rms(self.ctx, ctypes.byref(v)), "energy_rms")
        return v.value

    def set_time(self, tt, first_dual):
        check(self.L.shcfd_set_time(self.ctx, tt, int(first_dual)), "set_time")

    def get_time(self):
        a, b = ctypes.c_double(0.0), ctypes.c_double(0.0)
        check(self.L.shcfd_get_time(self.ctx, ctypes.byref(a), ctypes.byref(b)), "get_time")
        return a.value, b.value

    def synchronize(self):
        check(self.L.shcfd_synchronize(self.ctx), "synchronize")

    # ---- whole steps ----
    def prepare(self):
        check(self.L.shcfd_prepare(self.ctx), "prepare")

    def step(self, step=1, step_res=0):
        r1, r2 = ctypes.c_double(0.0), ctypes.c_double(0.0)
        check(self.L.shcfd_step(self.ctx, step, step_res, ctypes.byref(r1), ctypes.byref(r2)), "step")
        return r1.value, r2.value

    # ---- discrete adjoint (include/shcfd.h) ----
    def adjoint_objective(self, block, i0, i1, ax, ay):
        """J = sum over j-min faces i0..i1 of `block` of dsj (TG_u ax + TG_v ay)."""
        check(self.L.shcfd_adjoint_objective(self.ctx, block, i0, i1, ax, ay), "adjoint_objective")

    def objective(self):
        v = ctypes.c_double(0.0)
        check(self.L.shcfd_objective(self.ctx, ctypes.byref(v)), "objective")
        return v.value

    def adjoint_upload_lambda(self, b, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == (self.case.nj[b] + 4, self.case.ni[b] + 4, self.case.dimq)
        check(self.L.shcfd_adjoint_upload(self.ctx, ADJ_LAMBDA, b, _dp(a)), "adjoint_upload")

    def adjoint_download(self, what, b):
        k = 2 if what in (ADJ_XBAR, ADJ_XBAR_SUM) else self.case.dimq
        out = np.zeros((self.case.nj[b] + 4, self.case.ni[b] + 4, k), dtype=np.float64)
        check(self.L.shcfd_adjoint_download(self.ctx, what, b, _dp(out)), "adjoint_download")
        return out

    def adjoint_poolbar(self):
        out = np.zeros(max(1, self.case.bc_pool.size), dtype=np.float64)
        check(self.L.shcfd_adjoint_poolbar(self.ctx, _dp(out)), "adjoint_poolbar")
        return out[: self.case.bc_pool.size]

    def residual_vjp(self):
        check(self.L.shcfd_residual_vjp(self.ctx), "residual_vjp")

    def adjoint_solve(self, max_iter=20000, rtol=1e-10, check_every=100):
        r0, r, it = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_int(0)
        check(self.L.shcfd_adjoint_solve(self.ctx, max_iter, rtol, check_every, ctypes.byref(r0), ctypes.byref(r),
                                         ctypes.byref(it)), "adjoint_solve")
        return r0.value, r.value, it.value

    def adjoint_solve_gmres(self, restart=40, max_iter=2000, rtol=1e-10):
        """Restarted GMRES, right-preconditioned by the transposed LU-SGS
        sweeps: (initial RMS, final RMS, Krylov steps)."""
        r0, r, it = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_int(0)
        check(self.L.shcfd_adjoint_solve_gmres(self.ctx, restart, max_iter, rtol, ctypes.byref(r0), ctypes.byref(r),
                                               ctypes.byref(it)), "adjoint_solve_gmres")
        return r0.value, r.value, it.value

    # ---- fixed time step, snapshots, unsteady adjoint ----
    def set_fixed_dt(self, dt):
        check(self.L.shcfd_set_fixed_dt(self.ctx, float(dt)), "set_fixed_dt")

    def snapshot_save(self, slot, field="Q"):
        check(self.L.shcfd_snapshot_save(self.ctx, slot, C.F[field]), "snapshot_save")

    def snapshot_load(self, slot, field="Q"):
        check(self.L.shcfd_snapshot_load(self.ctx, slot, C.F[field]), "snapshot_load")

    def snapshot_free(self):
        check(self.L.shcfd_snapshot_free(self.ctx), "snapshot_free")

    def adjoint_unsteady_reset(self):
        check(self.L.shcfd_adjoint_unsteady_reset(self.ctx), "adjoint_unsteady_reset")

    def adjoint_unsteady_step(self, max_iter=20000, rtol=1e-10, check_every=20):
        r0, r, it = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_int(0)
        check(self.L.shcfd_adjoint_unsteady_step(self.ctx, max_iter, rtol, check_every, ctypes.byref(r0),
                                                 ctypes.byref(r), ctypes.byref(it)), "adjoint_unsteady_step")
        return r0.value, r.value, it.value

    def adjoint_unsteady_poolbar(self):
        out = np.zeros(max(1, self.case.bc_pool.size), dtype=np.float64)
        check(self.L.shcfd_adjoint_unsteady_poolbar(self.ctx, _dp(out)), "adjoint_unsteady_poolbar")
        return out[: self.case.bc_pool.size]

    # ---- per-kernel timing ----
    def profile(self, classes=None):
        """Time the given kernel classes (names of KCLASS; None = all, [] = off)."""
        mask = -1 if classes is None else sum(1 << KCLASS.index(c) for c in classes)
        check(self.L.shcfd_profile_enable(self.ctx, mask), "profile_enable")

    def profile_read(self, reset=True):
        """{class: (total_ms, launches)} for every kernel class with launches."""
        out = {}
        for k, name in enumerate(KCLASS):
            ms, n = ctypes.c_double(0.0), ctypes.c_longlong(0)
            check(self.L.shcfd_profile_read(self.ctx, k, ctypes.byref(ms), ctypes.byref(n), int(reset)), "profile_read")
            if n.value:
                out[name] = (ms.value, n.value)
        return out
