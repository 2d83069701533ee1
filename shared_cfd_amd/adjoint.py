"""Shape and inflow-angle gradients of airfoil force coefficients through the
discrete adjoint of the GPU solver, behind a torch.autograd.Function.

The primal is the steady LU-SGS solution, inviscid or laminar, on the C-mesh
of shared_cfd_amd.airfoil; the adjoint (include/shcfd.h, shcfd_adjoint_*) solves
(dR/dq)^T lambda = -(dJ/dq)^T with the transposed LU-SGS iteration and returns
dJ/dX per node and dJ/d(free-stream state).  The chain to the design
variables is exact and linear:

* Hicks-Henne amplitudes p: X = X0 + Phi p, so dJ/dp = Phi^T dJ/dX;
* inflow angle alpha: the free-stream state (u, v) = V (cos a, sin a) enters
  the far-field ghosts, and the lift/drag directions rotate with alpha.

The reference has no adjoint; PyTorch is used only for this differentiable
interface, every derivative is computed by the HIP kernels.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np
import torch

from . import airfoil as A
from . import cases as C
from .native import ADJ_LAMBDA, ADJ_LAMBDA_LATER, ADJ_XBAR, ADJ_XBAR_SUM, Solver


@dataclasses.dataclass
class PrimalResult:
    steps: int
    rms0: float
    rms: float
    seconds: float


@dataclasses.dataclass
class AdjointResult:
    iterations: int
    res0: float
    res: float
    seconds: float
    dJ_dp: np.ndarray
    dJ_dalpha: float


class AirfoilFlow:
    """One airfoil configuration (amplitudes p, inflow angle) on the GPU."""

    def __init__(self, amplitudes=None, alpha_deg=1.25, **mesh):
        self.mesh = dict(mesh)
        self.ac = A.naca_cmesh(amplitudes=amplitudes, alpha_deg=alpha_deg, **mesh)
        self.solver = Solver(self.ac.case)
        self.solver.prepare()
        self.step = 0
        self.q_inf = 0.5 * self.ac.rho_inf * self.ac.v_inf ** 2
        self.i0, self.i1 = self.ac.wall_range()

    def close(self):
        self.solver.close()

    # ---- primal --------------------------------------------------------------
    def converge(self, rtol=1e-10, max_steps=200000, chunk=200) -> PrimalResult:
        """Steps until the energy RMS change per step falls below rtol times
        its value after the first chunk (or max_steps)."""
        s = self.solver
        t0 = time.perf_counter()
        s.iterate(self.step + 1, chunk, 0)
        self.step += chunk
        rms0 = rms = s.energy_rms()
        while self.step < max_steps and rms > rtol * rms0:
            s.iterate(self.step + 1, chunk, 0)
            self.step += chunk
            rms = s.energy_rms()
            if not np.isfinite(rms):
                raise FloatingPointError("primal diverged")
        s.synchronize()
        return PrimalResult(self.step, rms0, rms, time.perf_counter() - t0)

    # ---- objectives ------------------------------------------------------------
    def directions(self):
        """Weights a of J = sum ds (TG_u a_x + TG_v a_y) for C_L and C_D (the
        force on the body is minus the momentum flux into the fluid)."""
        e_l, e_d = self.ac.directions()
        return -e_l / self.q_inf, -e_d / self.q_inf

    def force_flux(self):
        """(sum ds TG_u, sum ds TG_v) over the wall faces."""
        s = self.solver
        s.adjoint_objective(0, self.i0, self.i1, 1.0, 0.0)
        fx = s.objective()
        s.adjoint_objective(0, self.i0, self.i1, 0.0, 1.0)
        fy = s.objective()
        return np.array([fx, fy])

    def coefficients(self):
        """(C_L, C_D) of the current state."""
        F = self.force_flux()
        a_l, a_d = self.directions()
        return float(F @ a_l), float(F @ a_d)

    # ---- adjoint ---------------------------------------------------------------
    def reset_adjoint(self):
        """lambda = 0 (the next solve starts from zero instead of the last
        solution)."""
        self.solver.adjoint_upload_lambda(0, np.zeros((self.ac.nj + 4, self.ac.ni + 4, 4)))

    def gradient(self, w_l, w_d, rtol=1e-10, max_iter=200000, check_every=100, gmres=0) -> AdjointResult:
        """d(w_l C_L + w_d C_D)/d(p, alpha) at the current (converged) state.

        gmres = m > 0 solves the adjoint system by GMRES(m) (max_iter Krylov
        steps) instead of the fixed-point transposed LU-SGS iteration."""
        s = self.solver
        a_l, a_d = self.directions()
        a = w_l * a_l + w_d * a_d
        s.adjoint_objective(0, self.i0, self.i1, float(a[0]), float(a[1]))
        t0 = time.perf_counter()
        if gmres:
            res0, res, its = s.adjoint_solve_gmres(gmres, max_iter, rtol)
        else:
            res0, res, its = s.adjoint_solve(max_iter, rtol, check_every)
        s.synchronize()
        secs = time.perf_counter() - t0
        ni, nj = self.ac.ni, self.ac.nj
        xb = s.adjoint_download(ADJ_XBAR, 0)[1:nj + 2, 1:ni + 2]    # node (i,j) at [j+1, i+1]
        dJdX = np.stack([xb[..., 0].T, xb[..., 1].T])              # (2, ni+1, nj+1)
        dJ_dp = np.tensordot(self.ac.phi, dJdX, axes=([1, 2, 3], [0, 1, 2]))
        # inflow angle: far-field state and the rotation of the force directions
        pbar = s.adjoint_poolbar()
        al, V = self.ac.alpha, self.ac.v_inf
        dpool = pbar[1] * (-V * np.sin(al)) + pbar[2] * (V * np.cos(al))
        F = self.force_flux()
        da_l = np.array([np.cos(al), np.sin(al)]) / self.q_inf
        da_d = np.array([np.sin(al), -np.cos(al)]) / self.q_inf
        dJ_da = dpool + F @ (w_l * da_l + w_d * da_d)
        return AdjointResult(its, res0, res, secs, dJ_dp, float(dJ_da))


class AirfoilCoefficients(torch.autograd.Function):
    """(C_L, C_D) = f(amplitudes, alpha [rad]) with adjoint gradients.

    forward converges a fresh primal; backward runs one adjoint solve for the
    incoming weights.  Extra mesh options go in `opts` (a dict)."""

    @staticmethod
    def forward(ctx, amplitudes, alpha, opts):
        p = amplitudes.detach().cpu().numpy().astype(np.float64)
        flow = AirfoilFlow(amplitudes=p, alpha_deg=float(np.rad2deg(float(alpha))), **opts.get("mesh", {}))
        flow.converge(rtol=opts.get("rtol", 1e-10), max_steps=opts.get("max_steps", 200000))
        cl, cd = flow.coefficients()
        ctx.flow = flow
        ctx.opts = opts
        return torch.tensor([cl, cd], dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        g = grad.detach().cpu().numpy()
        flow = ctx.flow
        try:
            r = flow.gradient(float(g[0]), float(g[1]), rtol=ctx.opts.get("rtol", 1e-10),
                              gmres=ctx.opts.get("gmres", 0), max_iter=ctx.opts.get("adj_max_iter", 200000))
        finally:
            flow.close()
        return torch.from_numpy(r.dJ_dp), torch.tensor(r.dJ_dalpha, dtype=torch.float64), None


# ---------------------------------------------------------------------------
# unsteady: impulsively started airfoil, BDF2 dual-time steps
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class UnsteadyPrimal:
    steps: int
    J: float
    res_ratio_max: float     # largest last/first inner residual over the steps
    seconds: float


@dataclasses.dataclass
class UnsteadyAdjoint:
    steps: int
    iterations: int          # transposed LU-SGS iterations over all steps
    res_ratio_max: float
    seconds: float           # reverse sweep, recomputation included
    recomputed_steps: int
    dJ_dp: np.ndarray
    dJ_dalpha: float


class UnsteadyAirfoilFlow:
    """The airfoil started impulsively from the free stream and advanced by
    ``nsteps`` BDF2 dual-time steps of the NR scheme (or, with ``scheme`` =
    SCHEME_DUAL, the low-Mach preconditioned dual scheme) with a fixed time
    step ``dt`` (``inner - 1`` inner iterations each).  J = mean over the
    last ``window`` steps of w_l C_L + w_d C_D.

    Checkpointing: with ``checkpoint_every`` = k > 1 the forward run keeps
    the states (q^m, q^(m-1)) at every k-th step on the device; the reverse
    sweep recomputes each segment of k steps from its checkpoint (the solver
    is deterministic, so the recomputed states are bit-identical) and runs
    the adjoint steps of the segment backwards.  k = 1 keeps every state."""

    def __init__(self, amplitudes=None, alpha_deg=1.25, dt=0.01, inner=41, scheme=C.SCHEME_NR, cfl_tau=5.0, **mesh):
        prm = dict(scheme=scheme, time_step=0, num_internal_loop=inner, omega_max=1.0, dqmax=1.0, cfl_tau=cfl_tau)
        prm.update(mesh.pop("params", {}) or {})
        self.ac = A.naca_cmesh(amplitudes=amplitudes, alpha_deg=alpha_deg, params=prm, **mesh)
        self.solver = Solver(self.ac.case)
        self.solver.prepare()
        self.solver.set_fixed_dt(dt)
        self.dt = dt
        self.q_inf = 0.5 * self.ac.rho_inf * self.ac.v_inf ** 2
        self.i0, self.i1 = self.ac.wall_range()
        self.nsteps = 0

    def close(self):
        self.solver.close()

    def directions(self):
        e_l, e_d = self.ac.directions()
        return -e_l / self.q_inf, -e_d / self.q_inf

    def _force_flux(self):
        s = self.solver
        s.adjoint_objective(0, self.i0, self.i1, 1.0, 0.0)
        fx = s.objective()
        s.adjoint_objective(0, self.i0, self.i1, 0.0, 1.0)
        fy = s.objective()
        return np.array([fx, fy])

    # ---- forward -----------------------------------------------------------
    def run(self, nsteps, window, w_l=1.0, w_d=0.0, checkpoint_every=1) -> UnsteadyPrimal:
        s = self.solver
        k = int(checkpoint_every)
        assert nsteps >= 1 and 1 <= window <= nsteps and k >= 1
        self.nsteps, self.window, self.k = nsteps, window, k
        self.w = (w_l, w_d)
        self.forces = {}
        a_l, a_d = self.directions()
        t0 = time.perf_counter()
        s.snapshot_save(0, "Q")
        s.snapshot_save(1, "Q")      # q^(-1) is never read; keep the slot valid
        J, worst = 0.0, 0.0
        for n in range(1, nsteps + 1):
            r1, r2 = s.step(n, 0)
            worst = max(worst, r2 / r1 if r1 > 0 else 0.0)
            if n > nsteps - window:
                F = self._force_flux()
                self.forces[n] = F
                J += float(F @ (w_l * a_l + w_d * a_d)) / window
            if k == 1:
                s.snapshot_save(n, "Q")
            elif n % k == 0 and n < nsteps:
                s.snapshot_save(2 * (n // k), "Q")
                s.snapshot_save(2 * (n // k) + 1, "QP")
        s.synchronize()
        self.J = J
        return UnsteadyPrimal(nsteps, J, worst, time.perf_counter() - t0)

    # ---- reverse -----------------------------------------------------------
    def gradient(self, rtol=1e-10, max_iter=20000, check_every=10) -> UnsteadyAdjoint:
        s = self.solver
        N, k, window, dt = self.nsteps, self.k, self.window, self.dt
        w_l, w_d = self.w
        a_l, a_d = self.directions()
        a = (w_l * a_l + w_d * a_d) / window
        t0 = time.perf_counter()
        s.adjoint_unsteady_reset()
        its, worst, recomputed = 0, 0.0, 0
        ncp = (N + k - 1) // k
        seg0 = 2 * ncp                      # first segment slot (k > 1)

        def slot_of(m, k0):
            """(slot, field) holding q^m while segment (k0, ...] is live."""
            if k == 1:
                return m, "Q"
            if m > k0:
                return seg0 + (m - k0 - 1), "Q"
            if m == k0:
                return 2 * (k0 // k), "Q"
            return 2 * (k0 // k) + 1, "Q"   # m = k0 - 1: the checkpoint's QP copy

        starts = [0] if k == 1 else list(range(0, N, k))
        for k0 in reversed(starts):
            k1 = N if k == 1 else min(k0 + k, N)
            if k > 1:
                s.snapshot_load(2 * (k0 // k), "Q")
                s.snapshot_load(2 * (k0 // k) + 1, "QP")
                s.set_time(k0 * dt, 1 if k0 == 0 else 0)
                for n in range(k0 + 1, k1 + 1):
                    s.step(n, 0)
                    s.snapshot_save(seg0 + (n - k0 - 1), "Q")
                    recomputed += 1
            for n in range(k1, k0, -1):
                for m, fld in ((n, "Q"), (max(n - 1, 0), "QP"), (max(n - 2, 0), "QPP")):
                    if k > 1 and m < k0 - 1:
                        m = k0 - 1
                    sl, _ = slot_of(m, k0)
                    s.snapshot_load(sl, fld)
                s.set_time(n * dt, 1 if n == 1 else 0)
                wt = a if n > N - window else np.zeros(2)
                s.adjoint_objective(0, self.i0, self.i1, float(wt[0]), float(wt[1]))
                r0, r, it = s.adjoint_unsteady_step(max_iter, rtol, check_every)
                its += it
                worst = max(worst, r / r0 if r0 > 0 else 0.0)
        s.synchronize()
        secs = time.perf_counter() - t0
        ni, nj = self.ac.ni, self.ac.nj
        xb = s.adjoint_download(ADJ_XBAR_SUM, 0)[1:nj + 2, 1:ni + 2]
        dJdX = np.stack([xb[..., 0].T, xb[..., 1].T])
        dJ_dp = np.tensordot(self.ac.phi, dJdX, axes=([1, 2, 3], [0, 1, 2]))
        # inflow angle: far-field state over all steps, the rotation of the
        # force directions, and the initial (free-stream) state q^0 through
        # U^1 = ... + Vol q^0/dt and U^2 = ... - Vol q^0/(2 dt)
        al, V = self.ac.alpha, self.ac.v_inf
        pbar = s.adjoint_unsteady_poolbar()
        dpool = pbar[1] * (-V * np.sin(al)) + pbar[2] * (V * np.cos(al))
        da_l = np.array([np.cos(al), np.sin(al)]) / self.q_inf
        da_d = np.array([np.sin(al), -np.cos(al)]) / self.q_inf
        drot = sum(float(F @ (w_l * da_l + w_d * da_d)) for F in self.forces.values()) / window
        lam1 = s.adjoint_download(ADJ_LAMBDA, 0)[2:nj + 2, 2:ni + 2]
        lam2 = s.adjoint_download(ADJ_LAMBDA_LATER, 0)[2:nj + 2, 2:ni + 2]
        vol = s.download("VOL", 0)[2:nj + 2, 2:ni + 2, 0]
        q0bar = vol[..., None] * (lam1 - 0.5 * lam2) / dt
        rho = self.ac.rho_inf
        dq0 = np.array([0.0, -rho * V * np.sin(al), rho * V * np.cos(al), 0.0])
        dic = float(np.sum(q0bar * dq0))
        return UnsteadyAdjoint(N, its, worst, secs, recomputed, dJ_dp, float(dpool + drot + dic))
