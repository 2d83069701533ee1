#!/usr/bin/env python3
"""Throughput of the shared_cfd solver hot path on one MI355X.

Workload: the Mach-3 forward-facing step, 2 structured blocks
(192x320 + 768x256 = 258,048 cells in a 960x320 index space, h = 1/320),
ideal gas, MUSCL (order 3, reference limiter) + SLAU, preconditioned
LU-SGS in dual time with "InternalLoop Max Number : 21" (20 inner
iterations per physical step), FP64.  One bench "step" = one physical time
step = 20 inner iterations over all cells; a cell-update is one inner
iteration of one cell.  Inputs are resident in HBM before timing starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line (metric, value = Mcell-updates/s, roofline of the
dominant kernel from HIP events over the timed region, cpu_baseline of the
CPU oracle on the host cores over a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 Mcell-updates/s vs plain GPU & CPU; adjoint/primal cost at 1e-10 tol"
HBM_PEAK_GBS = 8000.0

# algorithmic HBM bytes per cell for one launch of each kernel class
# (minimum-pass schedule: residual reads q, qp, qpp and 13 inviscid geometry
# doubles and writes R; each sweep/update pass reads two dimq vectors and
# writes one; neighbour reads are counted once)
def alg_bytes_per_cell(kclass: str, dimq: int = 4, geo: int = 13) -> float:
    d = 8.0
    return {
        "SWEEP_FWD": 3 * dimq * d,      # read R, q; write dq*
        "SWEEP_BWD": 3 * dimq * d,      # read dq*, q; write dq
        "UPDATE": 3 * dimq * d,         # read dq, q; write q
        "RHS": (3 * dimq + geo + dimq) * d,       # read q, qp, qpp, G; write R
        "RESIDUAL": (3 * dimq + geo + dimq) * d,
    }.get(kclass, 0.0)


# Algorithmic FP64 operations of one point-implicit Newton iteration for one
# cell (add/mul = 1, FMA = 2; exp/log counted apart), for ns species and nr
# reactions: NASA-7 sums with derivatives 5 x 7 x 2 x ns; rates ~17 per
# reaction (once: F(y) is the dn/dt Jex forms beside the Jacobian); Jacobian
# scatter nr x (ns+1) x 2 x 3; LU 2/3 (ns+1)^3 and solve 2 (ns+1)^2.
def chem_flops_per_iteration(ns: int, nr: int) -> float:
    n1 = ns + 1
    return 5 * 7 * 2 * ns + 17 * nr + nr * n1 * 2 * 3 + (2.0 / 3.0) * n1 ** 3 + 2 * n1 ** 2


FP64_PEAK_TFLOPS = 78.6


def bench_reacting(a):
    """C4: reacting flame-sheet mixing layer, 1024x512 cells, synthetic 9
    species / 21 reactions, point-implicit LU-SGS (global dt, viscous)."""
    import tempfile

    from shared_cfd_amd import cases as C
    from shared_cfd_amd import chem
    from shared_cfd_amd.native import Solver

    d = tempfile.mkdtemp(prefix="shcfd_mech_")
    c, t = chem.write_synthetic_mechanism(0, d)
    m = chem.parse(c, t)
    case = C.mixing_layer(m, ni=a.ni, nj=a.nj, seed=0, species=chem.SPECIES9, reaction=2)
    cells = case.ncells
    s = Solver(case)
    s.prepare()
    step = 0
    for _ in range(a.warmup):
        step += 1
        s.step(step, 0)
    s.profile(None)
    step += 1
    s.step(step, 0)
    prof_all = s.profile_read()
    s.chem_iterations(reset=True)
    s.profile(["REACTION"])
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step += 1
        s.step(step, 0)
    s.synchronize()
    t1 = time.perf_counter()
    prof = s.profile_read()
    iters = s.chem_iterations(reset=True)
    ms, n = prof["REACTION"]
    avg_s = ms / n / 1e3
    flops_launch = chem_flops_per_iteration(m.ns, m.nr) * iters / n
    achieved = flops_launch / avg_s / 1e12
    tot = sum(x[0] for x in prof_all.values())
    out = {
        "metric": "FP64 Mcell-updates/s (reacting, point-implicit chemistry)",
        "value": round(cells * a.steps / (t1 - t0) / 1e6, 3), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round((t1 - t0) / a.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": case.description + ", LU-SGS + point-implicit backward Euler per cell",
                   "cells": cells, "species": m.ns, "reactions": m.nr},
        "roofline": {"bound": "fp64-valu", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 5), "traffic": None,
                     "kernel": "REACTION", "avg_launch_us": round(avg_s * 1e6, 3),
                     "newton_iterations_per_cell": round(iters / n / cells, 3),
                     "alg_flops_per_iteration": chem_flops_per_iteration(m.ns, m.nr),
                     "share_of_step": {k: round(v[0] / tot, 4) for k, v in prof_all.items() if v[0] > 0}},
    }
    s.close()
    if not a.no_cpu:
        from oracle.oracle import Oracle

        o = Oracle(case)
        o.prepare()
        tc0 = time.perf_counter()
        for _ in range(a.cpu_steps):
            o.step()
        tc1 = time.perf_counter()
        out["cpu_baseline"] = {"value": round(cells * a.cpu_steps / (tc1 - tc0) / 1e6, 4), "unit": "Mcell-updates/s",
                               "cores": o.L.or_omp_threads(), "kind": "port",
                               "sample": f"{a.cpu_steps} step(s) of the same {cells}-cell reacting workload, "
                                         f"C oracle (OpenMP), {tc1 - tc0:.1f} s"}
    return out


def bench_sod(a):
    """C1: Sod shock tube, 400x4 cells, MUSCL order 2 with minmod, LU-SGS,
    CFL 0.5, 200 steps: a graph of 20 steps is captured and replayed for
    steps 1-20 (untimed), then replayed 9 times for steps 21-200 (timed)."""
    from shared_cfd_amd import cases as C
    from shared_cfd_amd.native import Solver

    case = C.sod(400, 4, limiter=C.LIMITER_MINMOD, scheme=C.SCHEME_LUSGS, cfl=0.5)
    s = Solver(case)
    s.prepare()
    s.iterate(1, 20, 20)
    s.synchronize()
    t0 = time.perf_counter()
    s.iterate(21, 180, 20)
    s.synchronize()
    t1 = time.perf_counter()
    _, tt = s.get_time()
    s.close()
    return {"workload": case.description + ", MUSCL 2 + minmod, LU-SGS, CFL 0.5, 200 steps (steps 21-200 timed)",
            "cells": case.ncells, "steps_timed": 180, "seconds": round(t1 - t0, 5),
            "value": round(case.ncells * 180 / (t1 - t0) / 1e6, 3), "unit": "Mcell-updates/s", "t_end": tt}


def bench_batched_key(a):
    """C3 as an extra key of the default line: the 256-block batch (graph +
    autotuned schedule) for a few steps, against the same blocks one by one
    with the optimised kernels (graph of steps per block) and with the
    per-routine kernels (eager)."""
    from shared_cfd_amd import cases as C
    from shared_cfd_amd.native import Solver

    case = C.batched_plates(n=256, seed=0)
    cells = case.ncells
    s = Solver(case)
    s.prepare()
    choice = s.autotune(2)
    s.iterate(1, 8, -1)
    s.synchronize()
    t0 = time.perf_counter()
    s.iterate(9, a.batch_steps, -1)
    s.synchronize()
    t1 = time.perf_counter()
    s.close()
    batched = cells * a.batch_steps / (t1 - t0) / 1e6
    seq = {}
    for name, fast, gs in (("sequential_optimised", True, 4), ("sequential_plain", False, 0)):
        tt, nc = 0.0, 0
        for b in range(case.nblocks):
            sub = C.sub_case(case, [b])
            p = Solver(sub, fast=fast)
            p.prepare()
            p.iterate(1, max(gs, 1), gs)          # warm-up (and capture)
            p.synchronize()
            tb0 = time.perf_counter()
            p.iterate(2, a.seq_steps, gs)
            p.synchronize()
            tt += time.perf_counter() - tb0
            nc += sub.ncells * a.seq_steps
            p.close()
        seq[name] = {"value": round(nc / tt / 1e6, 3), "unit": "Mcell-updates/s", "instances": case.nblocks,
                     "steps": a.seq_steps}
    return {"workload": case.description, "cells": cells, "blocks": case.nblocks, "steps": a.batch_steps,
            "value": round(batched, 3), "unit": "Mcell-updates/s", "autotuned": choice, **seq,
            "speedup_vs_sequential_plain": round(batched / seq["sequential_plain"]["value"], 2),
            "speedup_vs_sequential_optimised": round(batched / seq["sequential_optimised"]["value"], 2)}


def bench_batched(a):
    """C3: 256 independent laminar plate/channel blocks of uneven sizes in one
    context (one launch per kernel over all blocks, autotuned schedule), against
    the same blocks solved one by one with the per-routine kernels."""
    from shared_cfd_amd import cases as C
    from shared_cfd_amd.native import Solver

    case = C.batched_plates(n=a.instances, seed=0)
    cells = case.ncells
    s = Solver(case)
    s.prepare()
    # --no-autotune keeps the default schedule (fast sweeps, graph of steps):
    # a kernel trace then holds the measured steps alone
    choice = None if a.no_autotune else s.autotune(4)
    s.iterate(1, 8, -1)  # warm-up with the tuned schedule (captures the graph)
    s.synchronize()
    t0 = time.perf_counter()
    s.iterate(9, a.steps, -1)
    s.synchronize()
    t1 = time.perf_counter()
    batched = cells * a.steps / (t1 - t0) / 1e6
    s.close()
    # the same instances one by one, per-routine kernels, eager launches
    seq_time, seq_cells = 0.0, 0
    for b in range(case.nblocks):
        if b >= a.seq_instances:
            break
        sub = C.sub_case(case, [b])
        p = Solver(sub, fast=False)
        p.prepare()
        p.iterate(1, 1, 0)
        p.synchronize()
        tb0 = time.perf_counter()
        p.iterate(2, a.seq_steps, 0)
        p.synchronize()
        seq_time += time.perf_counter() - tb0
        seq_cells += sub.ncells * a.seq_steps
        p.close()
    seq = seq_cells / seq_time / 1e6 if seq_cells else None
    return {
        "metric": "FP64 Mcell-updates/s (batched independent blocks)",
        "value": round(batched, 3), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": 8,
        "ms_per_step": round((t1 - t0) / a.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": case.description, "cells": cells, "blocks": case.nblocks, "autotuned": choice},
        "sequential_plain": {"value": round(seq, 3), "unit": "Mcell-updates/s",
                             "instances": min(a.seq_instances, case.nblocks), "steps": a.seq_steps} if seq else None,
        "speedup_vs_sequential": round(batched / seq, 2) if seq else None,
    }


def bench_adjoint(a, fd=True):
    """Discrete adjoint on the NACA 4-digit (12 %) C-mesh, 512x128 cells,
    Mach 0.8, alpha 1.25 deg, 64 Hicks-Henne amplitudes drawn from seed 0 in
    [-1e-3, 1e-3] chord: the steady primal and the adjoint of C_L (then C_D)
    each iterated to 1e-10 relative; cost = wall time of the adjoint solve
    over that of the primal solve.  First-order face states: with the
    reference limiter the second-order scheme stalls at Mach 0.8, and its
    exact linearisation does not converge under the LU-SGS iteration."""
    import numpy as np
    from shared_cfd_amd.adjoint import AirfoilFlow

    mesh = dict(ni=a.adj_ni, nj=a.adj_nj, nwake=a.adj_ni // 8, mach=0.8, cfl=2.0, order=1)
    rng = np.random.default_rng(0)
    p0 = rng.uniform(-1e-3, 1e-3, 64)
    f = AirfoilFlow(amplitudes=p0, alpha_deg=1.25, **mesh)
    pr = f.converge(rtol=1e-10, max_steps=400000)
    cl, cd = f.coefficients()
    gl = f.gradient(1.0, 0.0, rtol=1e-10, max_iter=400000)
    gd = f.gradient(0.0, 1.0, rtol=1e-10, max_iter=400000)
    f.close()
    out = {
        "workload": f"NACA 4-digit t=0.12 C-mesh {mesh['ni']}x{mesh['nj']}, M=0.8, alpha=1.25 deg, inviscid, "
                    "LU-SGS local dt (first-order face states), 64 Hicks-Henne amplitudes (seed 0)",
        "cells": mesh["ni"] * mesh["nj"], "tol": 1e-10,
        "primal": {"steps": pr.steps, "seconds": round(pr.seconds, 3), "rms_ratio": pr.rms / pr.rms0},
        "adjoint_cl": {"iterations": gl.iterations, "seconds": round(gl.seconds, 3), "res_ratio": gl.res / gl.res0},
        "adjoint_cd": {"iterations": gd.iterations, "seconds": round(gd.seconds, 3), "res_ratio": gd.res / gd.res0},
        "cost_ratio": round(gl.seconds / pr.seconds, 3),
        "cost_ratio_per_iteration": round((gl.seconds / gl.iterations) / (pr.seconds / pr.steps), 3),
        "CL": cl, "CD": cd, "dCL_dalpha": gl.dJ_dalpha, "dCD_dalpha": gd.dJ_dalpha,
        "dCL_dp_norm": float(np.linalg.norm(gl.dJ_dp)),
    }
    if fd:
        h, k = 1e-7, 20
        vals = []
        for sgn in (1.0, -1.0):
            p = p0.copy()
            p[k] += sgn * h
            ff = AirfoilFlow(amplitudes=p, alpha_deg=1.25, **mesh)
            ff.converge(rtol=1e-12, max_steps=600000)
            vals.append(ff.coefficients()[0])
            ff.close()
        fdv = (vals[0] - vals[1]) / (2 * h)
        out["fd_check"] = {"param": k, "h": h, "adjoint": float(gl.dJ_dp[k]), "central_fd": fdv,
                           "rel_err": abs(fdv - gl.dJ_dp[k]) / abs(fdv)}
    if not a.no_unsteady:
        out["unsteady"] = bench_adjoint_unsteady(a, p0, mesh)
    return out


def bench_adjoint_unsteady(a, p0, mesh, nsteps=200, window=100, dt=0.002, inner=61, every=15):
    """Unsteady variant: the same airfoil started impulsively from the free
    stream, 200 BDF2 dual-time steps (NR scheme, fixed dt, 60 inner
    iterations each); J = mean C_L over the last 100 steps.  The reverse
    sweep of unsteady adjoint steps (each to 1e-10 of its right-hand side,
    starting from the later step's lambda) runs once with
    every state kept on the device and once with checkpoints every 15 steps
    (segments recomputed); cost = reverse sweep time over forward run time."""
    import numpy as np
    from shared_cfd_amd.adjoint import UnsteadyAirfoilFlow

    res = {"workload": f"impulsive start, {nsteps} BDF2 dual-time steps (NR scheme), dt={dt}, "
                       f"{inner - 1} inner iterations/step, J = mean C_L over the last {window} steps",
           "tol": 1e-10}
    for k in (1, every):
        f = UnsteadyAirfoilFlow(amplitudes=p0, alpha_deg=1.25, dt=dt, inner=inner, **mesh)
        pr = f.run(nsteps, window, checkpoint_every=k)
        g = f.gradient(rtol=1e-10, max_iter=20000, check_every=20)
        f.close()
        key = "store_all" if k == 1 else f"checkpoint_every_{k}"
        res[key] = {"primal_seconds": round(pr.seconds, 3), "primal_inner_res_ratio_max": pr.res_ratio_max,
                    "adjoint_seconds": round(g.seconds, 3), "adjoint_iterations": g.iterations,
                    "adjoint_res_ratio_max": g.res_ratio_max, "recomputed_steps": g.recomputed_steps,
                    "cost_ratio": round(g.seconds / pr.seconds, 3), "J": pr.J, "dJ_dalpha": g.dJ_dalpha,
                    "dJ_dp_norm": float(np.linalg.norm(g.dJ_dp))}
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--seq-instances", type=int, default=256)
    ap.add_argument("--seq-steps", type=int, default=5)
    ap.add_argument("--no-autotune", action="store_true", help="batched: default schedule, no autotuner trials")
    ap.add_argument("--workload", choices=["step", "reacting", "batched", "adjoint"], default="step",
                    help="step: Mach-3 forward step (the headline metric); reacting: mixing layer with "
                         "chemistry; batched: 256 independent plate/channel blocks; adjoint: airfoil shape "
                         "gradients with a finite-difference spot check")
    ap.add_argument("--adj-ni", type=int, default=512)
    ap.add_argument("--adj-nj", type=int, default=128)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-unsteady", action="store_true", help="skip the unsteady adjoint variant")
    ap.add_argument("--ni", type=int, default=1024)
    ap.add_argument("--nj", type=int, default=512)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-per-unit", type=int, default=320)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="CPU sample length (default: 15 physical steps; reacting: 30 steps)")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--prof-steps", type=int, default=2, help="eager profiled steps for kernel durations")
    ap.add_argument("--no-extra", action="store_true", help="skip the Sod, batched and reacting keys")
    ap.add_argument("--batch-steps", type=int, default=16)
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "whole blocks would shard across GPUs, but no multi-GPU path is built.",
              file=sys.stderr)
        return 2
    if a.workload == "reacting":
        if a.cpu_steps is None:
            a.cpu_steps = 30
        print(json.dumps(bench_reacting(a)))
        return 0
    if a.workload == "batched":
        print(json.dumps(bench_batched(a)))
        return 0
    if a.workload == "adjoint":
        r = bench_adjoint(a, fd=True)
        print(json.dumps({"metric": "adjoint/primal cost at 1e-10 tol", "value": r["cost_ratio"], "unit": "ratio",
                          "n_gpus": 1, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
                          "vs_baseline": None, "config": {"workload": r["workload"]}, "adjoint": r}))
        return 0

    from shared_cfd_amd import cases as C
    from shared_cfd_amd.native import Solver

    if a.cpu_steps is None:
        a.cpu_steps = 15
    case = C.forward_step(n_per_unit=a.n_per_unit, num_internal_loop=21)
    inner = case.params.num_internal_loop - 1
    cells = case.ncells
    s = Solver(case)
    s.prepare()
    step = 1
    # warm-up: physical steps replayed from a graph of one step (captured here)
    s.iterate(step, max(a.warmup, 1), 1)
    step += max(a.warmup, 1)
    # kernel durations: HIP events on the library's stream around every
    # launch, over eager steps (the launches a graph replay makes are the same
    # kernels with the same arguments); first all classes, to find the
    # dominant one, then that class alone
    s.profile(None)
    s.step(step, 0)
    step += 1
    prof_all = s.profile_read()
    dominant = max((k for k in prof_all if alg_bytes_per_cell(k) > 0), key=lambda k: prof_all[k][0])
    s.profile([dominant])
    for _ in range(a.prof_steps):
        s.step(step, 0)
        step += 1
    prof = s.profile_read()
    s.profile([])
    # timed region: K physical steps as graph replays (one physical step of
    # 20 inner iterations per graph; the host checks each step's inner
    # residual ratio after its replay, as check_internal_loop_convergence does)
    s.synchronize()
    t0 = time.perf_counter()
    s.iterate(step, a.steps, 1)
    s.synchronize()
    t1 = time.perf_counter()
    step += a.steps
    dt = t1 - t0
    value = cells * inner * a.steps / dt / 1e6
    ms_dom, n_dom = prof[dominant]
    avg_s = ms_dom / n_dom / 1e3
    bytes_launch = alg_bytes_per_cell(dominant) * cells
    achieved = bytes_launch / avg_s / 1e9
    # HBM bytes per launch of the dominant kernel from rocprofv3 --pmc
    # FETCH_SIZE / WRITE_SIZE runs on this workload (copy-calibrated); a
    # profile-sourced figure, not measured in this run
    traffic = None
    tr_path = os.path.join(ROOT, "profiles", "forward_step_258k_traffic.json")
    if a.n_per_unit == 320 and os.path.exists(tr_path):
        with open(tr_path) as f:
            traffic = json.load(f).get(dominant)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "traffic_source": "rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE of this workload "
                                  "(profiles/forward_step_258k_traffic.json), not measured in this run",
                "traffic_note": ("the forward sweep launch also forms the backward operators (HA, HB: 256 B per "
                                 "cell written) on the CUs its strips leave idle; traffic counts both, the "
                                 "algorithmic bytes only the sweep") if dominant == "SWEEP_FWD" else None,
                "kernel": dominant, "avg_launch_us": round(avg_s * 1e6, 3), "launches_timed": n_dom,
                "alg_bytes_per_launch": bytes_launch,
                "share_of_step": {k: round(v[0] / sum(x[0] for x in prof_all.values()), 4) for k, v in prof_all.items()}}

    # the per-routine drop-in ("plain GPU": one kernel per reference routine,
    # stored Jacobians) on the same workload, for comparison
    plain = None
    if a.plain_steps > 0:
        s.close()
        sp = Solver(case, fast=False)
        sp.prepare()
        sp.step(1, 0)
        sp.synchronize()
        tp0 = time.perf_counter()
        for k in range(a.plain_steps):
            sp.step(2 + k, 0)
        sp.synchronize()
        tp1 = time.perf_counter()
        plain = {"value": round(cells * inner * a.plain_steps / (tp1 - tp0) / 1e6, 3), "unit": "Mcell-updates/s",
                 "steps": a.plain_steps}
        sp.close()

    cpu = None
    if not a.no_cpu:
        from oracle.oracle import Oracle

        o = Oracle(case)
        o.prepare()
        tc0 = time.perf_counter()
        for _ in range(a.cpu_steps):
            err, _, _ = o.step()
        tc1 = time.perf_counter()
        cpu_val = cells * inner * a.cpu_steps / (tc1 - tc0) / 1e6
        cpu = {"value": round(cpu_val, 3), "unit": "Mcell-updates/s", "cores": o.L.or_omp_threads(), "kind": "port",
               "sample": f"{a.cpu_steps} physical step(s) x {inner} inner iterations of the same "
                         f"{cells}-cell forward-step workload, C oracle (OpenMP), {tc1 - tc0:.1f} s"}

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(dt / a.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": case.description + ", dual time 20 inner iterations/step, MUSCL3 + SLAU, FP64",
                   "cells": cells, "inner_iterations_per_step": inner, "blocks": case.nblocks},
        "roofline": roofline, "cpu_baseline": cpu, "plain_gpu": plain,
    }
    # the other workloads, measured in the same run: Sod (400x4, 200 steps),
    # the 256-block batch with its sequential controls, and the reacting
    # mixing layer (1024x512, 9 species / 21 reactions)
    if not a.no_extra:
        s.close()
        out["sod"] = bench_sod(a)
        out["batched"] = bench_batched_key(a)
        ra = argparse.Namespace(**vars(a))
        ra.steps, ra.warmup, ra.ni, ra.nj, ra.cpu_steps = 3, 1, 1024, 512, 3
        out["reacting"] = bench_reacting(ra)
    # the metric's second figure: adjoint/primal cost at 1e-10 (airfoil
    # shape-gradient workload; the finite-difference check runs only with
    # --workload adjoint)
    if not a.no_adjoint:
        s.close()
        out["adjoint"] = bench_adjoint(a, fd=False)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
