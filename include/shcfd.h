/*
 * shcfd.h -- C ABI of libshcfd, the MI355X (gfx950) implementation of the
 * shared_cfd per-iteration solver hot path.
 *
 * Every entry point returns an int status (SHCFD_OK or an SHCFD_E_* code,
 * see shcfd_types.h; shcfd_strerror gives the text).  Arguments are plain
 * pointers and sizes.  Host arrays are borrowed only for the duration of a
 * call.  All device work of one context is queued on one HIP stream; calls
 * that return a scalar synchronise that stream, all others are asynchronous.
 *
 * The reference exposes no library API: its main program calls external
 * Fortran subroutines by name (generated main.f90, e.g.
 * store/time_scheme/dual/2d/main.main.raw.f90).  Each per-routine entry
 * point below replaces one of those subroutines; the citation gives the
 * reference routine it stands in for.  The Fortran host module
 * shared_cfd_amd/fortran/shcfd_f.f90 binds them with ISO_C_BINDING under the
 * reference's names.
 */
#ifndef SHCFD_H
#define SHCFD_H

#include "shcfd_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct shcfd_ctx shcfd_ctx;

/* ---- context ------------------------------------------------------------ */
/* Replaces the compile-time sizes of store/core/n_grid.raw.f90:1-20 (Nplane,
 * ni(:), nj(:), nY) and the module state of store/core/variable.f90:1-49:
 * allocates all blocks' state in device memory (packed, no nimax padding). */
int shcfd_create(int nblocks, const int *ni, const int *nj, int ny, const shcfd_params *params,
                 shcfd_ctx **out);
int shcfd_destroy(shcfd_ctx *ctx);
/* run-time control.inp keys: store/core/read_control.f90:61-92,
 * store/high_order/muscl/muscl.f90:76-116, store/time_scheme/dual/2d/sch_dual.raw.f90:301-340 */
int shcfd_set_params(shcfd_ctx *ctx, const shcfd_params *params);
const char *shcfd_strerror(int code);
/* options (no counterpart in the reference) */
enum {
    SHCFD_OPT_FAST_IMPLICIT = 1, /* 1 (default): optimised LU-SGS sweeps for nY = 1;
                                    0: the per-routine path with stored Jacobians */
    SHCFD_OPT_SPIN_LIMIT = 2,    /* bound on polls of a sweep hand-off before
                                    SHCFD_E_SPIN is raised                          */
    SHCFD_OPT_SKEW_SWEEP = 3,    /* 1 (default): whole steps with nY > 1 keep the
                                    split Jacobians in the strip-skewed layout that
                                    the row-per-lane sweeps read coalesced; 0: the
                                    Ap/Am/Bp/Bm fields                              */
    SHCFD_OPT_TILE_ROWS = 4,     /* rows of the 64-wide cell tiles of the per-cell
                                    kernels: 1, 2 or 4 (default)                    */
    SHCFD_OPT_JVP_SWEEP = 5      /* 1 (default): LU-SGS and NR at nY > 1 apply the
                                    split Jacobians in closed form from the cell
                                    state inside the 16-row sweeps (no operator
                                    copy); 0: the sweeps read a copy of the
                                    matrices (bit-identical to the field sweeps)   */
};
int shcfd_set_option(shcfd_ctx *ctx, int option, int value);
int shcfd_version(void);

/* ---- data movement (reference layout, see shcfd_types.h) ----------------- */
int shcfd_upload_field(shcfd_ctx *ctx, int field, int block, const double *host);
int shcfd_download_field(shcfd_ctx *ctx, int field, int block, double *host);
/* block-interface table, the cut_copro000.inp rows read by set_cut_copro
 * (store/architecture/PC/mod_mpi_dummy.f90:59-79) */
int shcfd_set_cut_table(shcfd_ctx *ctx, const int *rows, int nrows);
/* physical boundary segments replacing the user set_BC loops
 * (e.g. sample/no06.step_flow/condition.f90:31-157) */
int shcfd_set_bc_table(shcfd_ctx *ctx, const int *rows, int nrows, const double *pool, int npool);
/* geometry from the uploaded nodes (fields XH, RH):
 * init_geometry, store/dim/2d/geometry.f90:6-25; set_vnc_dsc,
 * store/time_scheme/LU-SGS/2d/sch_lusgs.raw.f90:143-181; set_geojac,
 * store/viscosity/viscous/no-nV/2d/sch_viscous.f90:1-54 */
int shcfd_init_geometry(shcfd_ctx *ctx);
/* time and the dual-time BDF1 flag (tt, first_dual of the reference) */
int shcfd_set_time(shcfd_ctx *ctx, double tt, int first_dual);
int shcfd_get_time(shcfd_ctx *ctx, double *dt_grbl, double *tt);
int shcfd_synchronize(shcfd_ctx *ctx);

/* ---- per-routine drop-ins ------------------------------------------------ */
/* set_w: store/core/scheme.f90:1-29 */
int shcfd_set_w(shcfd_ctx *ctx);
/* set_thermo_prop (ideal gas): store/therm_lib/ideal/thermal_model.raw.f90:7-33 */
int shcfd_set_thermo_prop(shcfd_ctx *ctx, int step);
/* set_BC: section_exchange + BCs + set_corners,
 * store/cond/no-nV/condition.tail.f90:2-70, :331-356 */
int shcfd_set_bc(shcfd_ctx *ctx, int step);
/* set_HV: store/high_order/muscl/muscl.f90:7-73 (upwind: upwind.f90:6-36) */
int shcfd_set_hv(shcfd_ctx *ctx);
/* set_TG (SLAU): store/core/scheme.f90:31-259 */
int shcfd_set_tg(shcfd_ctx *ctx);
/* set_TGv: store/viscosity/viscous/no-nV/2d/sch_viscous.f90:56-272
 * (inviscid: store/viscosity/non-viscous/sch_viscous.f90, TGv = 0) */
int shcfd_set_tgv(shcfd_ctx *ctx);
/* set_dt: store/time_step/global/set_dt.f90:1-43 / local/set_dt.f90:1-36;
 * dt_grbl_out (may be NULL) receives the global step (synchronises) */
int shcfd_set_dt(shcfd_ctx *ctx, int step, double *dt_grbl_out);
/* set_ABpm: store/time_scheme/LU-SGS/2d/sch_lusgs.raw.f90:1-141,
 * NR/2d/sch_NR.raw.f90:1-141, precon/2d/sch_precon.raw.f90:1-201,
 * preconLU-SGS/2d/sch_precon.raw.f90:1-202, dual/2d/sch_dual.raw.f90:1-209 */
int shcfd_set_abpm(shcfd_ctx *ctx);
/* set_dqdt: store/time_scheme/dual/2d/sch_dual.raw.f90:211-240 (the same in
 * NR/2d/sch_NR.raw.f90:143-172 and precon/2d/sch_precon.raw.f90:203-232) */
int shcfd_set_dqdt(shcfd_ctx *ctx);
/* calc_next_step_implicit: LU-SGS/2d/time_lusgs_viscous.f90:1-65,
 * preconLU-SGS/2d/time_precon_viscous.f90:1-66, NR/2d/time_NR_viscous.f90:1-102,
 * precon/2d/time_precon_viscous.f90:1-103, dual/2d/time_dual_viscous.f90:1-125.
 * res_out/omega_out (inner-loop schemes only, may be NULL) synchronise. */
int shcfd_calc_next_step_implicit(shcfd_ctx *ctx, int step_internal, double *res_out, double *omega_out);
/* calc_next_step (explicit Euler): store/time_scheme/euler/time_euler_viscous.raw.f90:1-23 */
int shcfd_calc_next_step(shcfd_ctx *ctx);
/* calc_next_step_RK2_first/second: store/time_scheme/RK2/time_RK2_viscous.raw.f90:1-59 */
int shcfd_calc_next_step_rk2(shcfd_ctx *ctx, int stage);
/* qp = q / qpp = qp copies of the main loops (interior cells) */
int shcfd_copy_state(shcfd_ctx *ctx, int dst_field, int src_field);
/* energy-change RMS of check_convergence: store/core/check_convergence.f90:17-28 */
int shcfd_energy_rms(shcfd_ctx *ctx, double *rms_out);

/* ---- whole steps --------------------------------------------------------- */
/* main.body.raw.f90:14-15: part_primitive + set_BC before the main loop */
int shcfd_prepare(shcfd_ctx *ctx);
/* one physical step of store/time_scheme/<scheme>/main.main.raw.f90 with
 * check_internal_loop_convergence (sch_dual.raw.f90:284-299) for the
 * inner-loop schemes (NR, precon, dual).  res_first/res_last: inner
 * residuals of the first and last inner iteration (may be NULL). */
int shcfd_step(shcfd_ctx *ctx, int step, int step_res, double *res_first, double *res_last);
/* pending asynchronous error (omega, NaN, sweep flag timeout); clears it */
int shcfd_check_error(shcfd_ctx *ctx);

/* ---- Chemkin chemistry ---------------------------------------------------- */
/* read_cheminp (store/therm_lib/chemkin/core/sub_chem.f90:191-598) on
 * chem_inp with the element weights of mw_inp, and set_trans_data
 * (:788-889) on trans_inp (NULL: no transport data, mu = 0).  Host only.
 * out->rtol/atol (control_chem.inp, read_fo_composition :934-935) are set
 * to 0 and left to the caller.  On error a message goes to errbuf. */
int shcfd_mech_parse(const char *chem_inp, const char *mw_inp, const char *trans_inp, shcfd_mech *out, char *errbuf,
                     int errlen);
/* make mech the context's mechanism (the reference's module chem state, set
 * up by init_therm); mech->ns must equal nY */
int shcfd_set_mechanism(shcfd_ctx *ctx, const shcfd_mech *mech);
/* init_therm of the reactive model (store/therm_lib/chemkin/flow/reactive/
 * main.part_init.f90 -> read_cheminp, set_trans_data, control_chem.inp
 * tolerances): shcfd_mech_parse of the three files, then
 * shcfd_set_mechanism.  For hosts that do not hold the table themselves
 * (the Fortran driver). */
int shcfd_load_mechanism(shcfd_ctx *ctx, const char *chem_inp, const char *mw_inp, const char *trans_inp, double rtol,
                         double atol);

/* ---- cold flow with NASA-9 thermodynamics (thermal model 5) ---------------
 * The reference's init_pack_cold_flow (store/therm_lib/NASA/core/
 * sub_chem.f90:380-387): read_cheminp (:2-67) of chem.inp, set_therm_data
 * (:68-166) of the CEA thermo.inp, set_trans_data (:167-245) of the CEA
 * trans.inp (optional: NULL gives mu = 0) and read_fo_composition (:247-300)
 * of control_chem.inp, with the stream mole ratios normalised to mol/kg
 * (calc_ini, :420-441).  pt (may be NULL) receives the stream conditions
 * (po, To, pf, Tf).  Host code; on error a message goes to errbuf. */
int shcfd_nasa9_parse(const char *chem_inp, const char *thermo_inp, const char *trans_inp,
                      const char *control_chem_inp, shcfd_nasa9 *out, double *pt, char *errbuf, int errlen);
/* make tables the context's cold-flow data; the context must have nY = 2
 * and thermal = SHCFD_THERMAL_COLD_NASA9.  set_thermo_prop then runs
 * store/therm_lib/NASA/cold/thermal_model.f90:1-26 per cell. */
int shcfd_set_nasa9(shcfd_ctx *ctx, const shcfd_nasa9 *tables);
/* the static state of the fuel (which = 0) or oxidizer (which = 1) stream at
 * (p, T): initialize_cold_flow / set_static_qw (sub_chem.f90:388-488):
 * q(1:5), w(1:10), vhi(1:2).  Host code. */
int shcfd_nasa9_stream(const shcfd_nasa9 *tables, int which, double p, double T, double *q, double *w, double *vhi);
/* init_pack_cold_flow for hosts that do not hold the tables themselves (the
 * Fortran driver): shcfd_nasa9_parse of the four files (trans_inp may be
 * NULL or ""), then shcfd_set_nasa9. */
int shcfd_load_nasa9(shcfd_ctx *ctx, const char *chem_inp, const char *thermo_inp, const char *trans_inp,
                     const char *control_chem_inp);
/* proceed_reaction: store/therm_lib/chemkin/flow/reactive/thermal_model.f90:45-73
 * (reaction_chemeq2 per cell over dt_grbl; or the point-implicit update,
 * params.reaction) */
int shcfd_proceed_reaction(shcfd_ctx *ctx);
/* instrumentation, no counterpart in the reference: Newton iterations
 * (point-implicit) or CHEMEQ2 attempts summed over all cells and calls since
 * the last reset (synchronises) */
int shcfd_chem_iterations(shcfd_ctx *ctx, double *total, int reset);
/* calc_boundary: thermal_model.f90:326-356 -- rho_rel2abs (sub_chem.f90:1036-1085)
 * and calc_T (:1507-1645) for a fixed state of pressure p [Pa], temperature
 * T [K] and relative composition vrho_rel(ns).  Evaluated on the host.
 * w(dimw) = rho, 0, 0, p, Y(ns), gamma, H, R, mu; vhi(ns); dhi(ns) (may be NULL). */
int shcfd_chem_boundary_state(const shcfd_mech *mech, double p, double T, const double *vrho_rel, double *w,
                              double *vhi, double *dhi);
/* Batched 0-D evaluation on the GPU of independent states (the per-cell
 * routines the reference's DVODE callbacks and chem_drive use).  Row
 * layouts per state (N = ns+1):
 *   SHCFD_CHEM_FEX      in n(ns),T          out dndt(N)          Fex,  FJ.f90:152-300
 *   SHCFD_CHEM_FEXVAR   in n(ns),T          out p(N), m(N)       Fex_var, FJ.f90:301-453
 *   SHCFD_CHEM_JEX      in n(ns),T          out dndn(N,N) col-major  Jex, FJ.f90:454-708
 *   SHCFD_CHEM_CALCT    in vrho(ns),rho,E,T0  out T,kappa,MWave,mu,DHi(ns),vhi(ns)  calc_T
 *   SHCFD_CHEM_CHEMEQ2  in vrho(ns),T       out vrho(ns)   reaction_chemeq2 over dt, with its DVODE
 *                       hand-over after 10 attempts (sub_chem.f90:1684-1819)
 *   SHCFD_CHEM_POINT_IMPLICIT  in vrho(ns),T  out vrho(ns),T1  point-implicit (backward-Euler) update over dt
 *   SHCFD_CHEM_DVODE    in vrho(ns),T       out vrho(ns)   reaction (DVODE over dt), sub_chem.f90:1236-1315
 * status[i] (may be NULL): Newton iterations / CHEMEQ2 attempts (11: DVODE
 * finished the interval) / DVODE steps, or -code.
 * Synchronous; allocates and frees its own device buffers. */
int shcfd_chem_eval(const shcfd_mech *mech, int op, int nstates, const double *in, double *out, int *status,
                    double dt);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, code between graph captures, harnesses that
 * forbid hipMalloc): shcfd_chem_eval with the same arguments, rows, results
 * and error behaviour, but it calls no hipMalloc or hipFree.  It stages the
 * mechanism table, the input and output rows and the status words in one
 * fixed device array per GPU of SHCFD_CHEM_SCRATCH_BYTES, held until the
 * process ends: a call needs
 *   256-aligned sizeof(shcfd_mech) + 256-aligned 8*n*row_in
 *   + 256-aligned 8*n*row_out + 256-aligned 4*n
 * bytes and returns SHCFD_E_TOO_LARGE (nothing computed) when that exceeds
 * the array.  Synchronous, on the null stream of the current device.  All
 * calls on one GPU share its array: calls that overlap on one GPU (from
 * several host threads) race and must be serialised by the caller. */
#define SHCFD_CHEM_SCRATCH_BYTES (16u << 20)
int shcfd_chem_eval_noalloc(const shcfd_mech *mech, int op, int nstates, const double *in, double *out,
                            int *status, double dt);

/* ---- batches of independent blocks (no counterpart in the reference, whose
 * kinematic viscosity is one compile-time constant, thermal_model.raw.f90:4,
 * and whose run stops for all blocks at once) ----------------------------- */
/* per-block kinematic viscosity of the ideal gas (mu = rho nu); shcfd_create
 * sets every block to params->nu */
int shcfd_set_block_nu(shcfd_ctx *ctx, const double *nu);
/* per-block flag: 0 freezes a block: the implicit and explicit updates and
 * the chemistry source leave its q unchanged */
int shcfd_set_block_active(shcfd_ctx *ctx, const int *active);
/* the energy-change RMS of check_convergence (check_convergence.f90:17-28)
 * for each block separately */
int shcfd_energy_rms_blocks(shcfd_ctx *ctx, double *rms);
/* nsteps physical steps (numbered from first_step) with no host
 * synchronisation between steps; graph_steps > 0 captures that many steps
 * into one HIP graph and replays it, 0 launches eagerly, < 0 uses the
 * schedule chosen by shcfd_autotune.  For the schemes with a dual-time inner
 * loop (NR, precon, dual) a graph holds whole physical steps (every inner
 * iteration, its reductions and updates); each step stores its first and
 * last inner residual on the device, and after each replay the host applies
 * check_internal_loop_convergence (sch_dual.raw.f90:284-299) to the steps in
 * order.  A failing step is then reported after the whole replay, i.e. with
 * the state of the replay's last step; eager launches (graph_steps = 0) stop
 * at the failing step, as the reference does. */
int shcfd_iterate(shcfd_ctx *ctx, int first_step, int nsteps, int graph_steps);
/* times candidate schedules (optimised or per-routine sweeps, tile rows of
 * the per-cell kernels, eager or graph launches) over trial_steps steps each
 * and keeps the fastest; the flow state is restored afterwards.
 * choice[4] (may be NULL) = fast sweeps, tile rows, graph steps, ms/step. */
int shcfd_autotune(shcfd_ctx *ctx, int trial_steps, double *choice);

/* ---- fixed physical time step and device snapshots (no counterpart in
 * the reference, whose global time step always follows the CFL rule of
 * set_dt; a time-accurate run that is differentiated or compared across
 * perturbed inputs needs the same dt in every run, and checkpointing needs
 * states kept on the device) ---------------------------------------------- */
/* dt > 0 holds dt_grbl at dt in set_dt (global time step only); 0 restores
 * the CFL rule */
int shcfd_set_fixed_dt(shcfd_ctx *ctx, double dt);
/* copy field (SHCFD_F_Q, _QP or _QPP, all cells) to / from snapshot slot
 * (0 <= slot < SHCFD_MAX_SNAPSHOTS; device memory of a slot is allocated on
 * its first save and kept until shcfd_snapshot_free or shcfd_destroy) */
int shcfd_snapshot_save(shcfd_ctx *ctx, int slot, int field);
int shcfd_snapshot_load(shcfd_ctx *ctx, int slot, int field);
int shcfd_snapshot_free(shcfd_ctx *ctx);

/* ---- discrete adjoint (no counterpart in the reference, which has no
 * adjoint; for shape and inflow-angle gradients of ideal-gas solutions,
 * inviscid or laminar (viscous): nY = 1, planar, optimised implicit path;
 * steady: LU-SGS, unsteady: the NR or dual scheme with a fixed time step;
 * boundary kinds up to the curved slip wall) ------------------------------ */
/* what of shcfd_adjoint_upload / shcfd_adjoint_download */
enum {
    SHCFD_ADJ_LAMBDA = 0, /* adjoint state, dimq per cell (reference cell layout) */
    SHCFD_ADJ_QBAR = 1,   /* lambda^T dR/dq + dJ/dq per cell                      */
    SHCFD_ADJ_XBAR = 2,   /* lambda^T dR/dX + dJ/dX per node: (x, y) in the cell
                             layout at the node's index, as node uploads use     */
    SHCFD_ADJ_XBAR_SUM = 3,   /* unsteady: node bars summed over the steps done   */
    SHCFD_ADJ_LAMBDA_LATER = 4 /* unsteady: lambda of the step after the last one
                                  done (lambda^2 once step 1 is done)             */
};
/* objective J = sum over the j-min faces i in [i0, i1] of block (0-based
 * block, 1-based i) of dsj * ((TGj - TGvj)_u * ax + (TGj - TGvj)_v * ay):
 * the momentum flux through a wall (pressure, plus the viscous stress when
 * the flow is viscous), so ax, ay = -e/(q_inf c) give the force
 * coefficient along e */
int shcfd_adjoint_objective(shcfd_ctx *ctx, int block, int i0, int i1, double ax, double ay);
/* evaluates the residual pipeline at the current state and J */
int shcfd_objective(shcfd_ctx *ctx, double *J);
int shcfd_adjoint_upload(shcfd_ctx *ctx, int what, int block, const double *host);
int shcfd_adjoint_download(shcfd_ctx *ctx, int what, int block, double *host);
/* bars of the fixed boundary states (npool doubles, the bc pool layout) */
int shcfd_adjoint_poolbar(shcfd_ctx *ctx, double *host);
/* one reverse pass at the current state with the current lambda: fills
 * QBAR, XBAR and the pool bars */
int shcfd_residual_vjp(shcfd_ctx *ctx);
/* solves (dR/dq)^T lambda = -(dJ/dq)^T by the transposed LU-SGS iteration
 * lambda += M^-T (dR/dq^T lambda + dJ/dq^T), starting from the current
 * lambda, until the RMS of the bracket falls below rtol times its first
 * value or max_iter iterations; the RMS is checked every check_every
 * iterations.  On return QBAR, XBAR and the pool bars belong to the final
 * lambda, so XBAR is dJ/dX and the pool bars dJ/d(fixed states). */
int shcfd_adjoint_solve(shcfd_ctx *ctx, int max_iter, double rtol, int check_every, double *res0, double *res,
                        int *iters);
/* The same adjoint system solved by restarted GMRES(restart),
 * right-preconditioned by the transposed LU-SGS sweeps, until the RMS of
 * dR/dq^T lambda + dJ/dq falls below rtol times its value at the starting
 * lambda (the current one) or max_iter Krylov steps were taken.  For the
 * second-order (MUSCL) residual, whose exact linearisation the fixed-point
 * iteration above need not converge on.  Allocates (restart + 5) adjoint
 * fields for the call. */
int shcfd_adjoint_solve_gmres(shcfd_ctx *ctx, int restart, int max_iter, double rtol, double *res0, double *res,
                              int *iters);
/* Unsteady adjoint of the dual-time (BDF2) steps of the NR or dual scheme
 * with a fixed time step: U^n = R(q^n) - Vol (a0 q^n + a1 q^(n-1) + a2 q^(n-2))/dt
 * = 0, with (a0, a1, a2) = (3/2, -2, 1/2) and (1, -1, 0) on the first step.
 * For J = sum_n J^n(q^n), the adjoint steps run from the last physical step
 * to the first; each solves
 *   (dU^n/dq^n)^T lambda^n = -(dJ^n/dq^n + Vol/dt (2 lambda^(n+1) - lambda^(n+2)/2))^T
 * by the transposed LU-SGS iteration (the NR operator as preconditioner)
 * and adds lambda^nT dU^n/dX + dJ^n/dX to XBAR_SUM and lambda^nT dU^n/d(fixed
 * states) to the pool-bar sum.  reset zeroes lambda, the later lambda and
 * the sums.  Before each step the caller restores Q = q^n, QP = q^(n-1),
 * QPP = q^(n-2) (shcfd_snapshot_load), sets first_dual with shcfd_set_time
 * and the weights of J^n with shcfd_adjoint_objective (0, 0 for none).
 * Each step starts from lambda^(n+1) and stops once the RMS of its residual
 * falls below rtol times the RMS of its right-hand side (the residual at
 * lambda = 0), which step returns in res0. */
int shcfd_adjoint_unsteady_reset(shcfd_ctx *ctx);
int shcfd_adjoint_unsteady_step(shcfd_ctx *ctx, int max_iter, double rtol, int check_every, double *res0,
                                double *res, int *iters);
/* the pool-bar sum (npool doubles) */
int shcfd_adjoint_unsteady_poolbar(shcfd_ctx *ctx, double *host);

/* ---- per-kernel timing (no counterpart in the reference, whose timers
 * are commented-out stubs: store/debug/time_barrier.f90:1-37) ------------- */
/* kernel classes */
enum {
    SHCFD_K_PRIMITIVE = 0, SHCFD_K_BC, SHCFD_K_HV, SHCFD_K_TG, SHCFD_K_TGV, SHCFD_K_DT, SHCFD_K_ABPM,
    SHCFD_K_DQDT, SHCFD_K_RHS, SHCFD_K_SWEEP_FWD, SHCFD_K_SWEEP_BWD, SHCFD_K_REDUCE, SHCFD_K_UPDATE,
    SHCFD_K_EXPLICIT, SHCFD_K_RESIDUAL, SHCFD_K_PREP, SHCFD_K_THERMO_CHEM, SHCFD_K_REACTION, SHCFD_K_COUNT
};
/* mask: bit k set = record HIP events (on the context stream) around every
 * launch of kernel class k; -1 = all classes, 0 = off */
int shcfd_profile_enable(shcfd_ctx *ctx, int mask);
/* synchronise, then total milliseconds and launch count of one class since
 * the last reset; reset != 0 clears the class afterwards */
int shcfd_profile_read(shcfd_ctx *ctx, int kernel_class, double *total_ms, long long *launches, int reset);

#ifdef __cplusplus
}
#endif
#endif /* SHCFD_H */
