/*
 * shcfd_oracle.h -- CPU restatement of the reference solver's per-iteration
 * hot path (ISASShimadaLab/shared_cfd, Fortran 90), written in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in shared_cfd_amd/ and the timed CPU baseline of bench.py.
 * It is never linked into, loaded by or called from the product path.
 *
 * Data live in the reference layout: one AoS array per field and block,
 * component fastest, ghosted index space (K,-1:ni+2,-1:nj+2) -- see
 * include/shcfd_types.h.  Each routine cites the reference file:line it
 * restates; arithmetic is written in the reference's evaluation order and the
 * library is compiled with -ffp-contract=off.
 */
#ifndef SHCFD_ORACLE_H
#define SHCFD_ORACLE_H

#include <stddef.h>
#include "../include/shcfd_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

or_ctx *or_create(int nb, const int *ni, const int *nj, int ny, const shcfd_params *p);
void or_destroy(or_ctx *c);
int or_set_params(or_ctx *c, const shcfd_params *p);
int or_set_block_nu(or_ctx *c, const double *nu);
/* pointer to the AoS array of one field/block (size or_field_size doubles) */
double *or_field(or_ctx *c, int field, int block);
int or_field_ncomp(or_ctx *c, int field);
size_t or_field_size(or_ctx *c, int field, int block);
int or_set_cut_table(or_ctx *c, const int *rows, int nrows);
int or_set_bc_table(or_ctx *c, const int *rows, int nrows, const double *pool, int npool);

/* geometry: nodes xh/rh must be filled (fields SHCFD_F_XH/RH) */
void or_init_geometry(or_ctx *c);     /* dim/2d/geometry.f90:6-169           */
void or_set_vnc_dsc(or_ctx *c);       /* LU-SGS/2d/sch_lusgs.raw.f90:143-181 */
void or_set_geojac(or_ctx *c);        /* viscous/no-nV/2d/sch_viscous.f90:1-54 */

/* per-routine restatements */
void or_set_w(or_ctx *c);             /* core/scheme.f90:1-29                */
void or_set_thermo_prop(or_ctx *c);   /* therm_lib/ideal/thermal_model.raw.f90:7-33 */
void or_set_bc(or_ctx *c);            /* cond/no-nV/condition.tail.f90 + user set_BC */
void or_set_hv(or_ctx *c);            /* high_order/muscl/muscl.f90:7-73     */
void or_set_tg(or_ctx *c);            /* core/scheme.f90:31-259              */
void or_set_tgv(or_ctx *c);           /* viscous/no-nV/2d/sch_viscous.f90:56-272 */
double or_set_dt(or_ctx *c);          /* time_step/{global,local}/set_dt.f90 */
void or_set_abpm(or_ctx *c);          /* LU-SGS / preconLU-SGS / dual set_ABpm */
void or_set_dqdt(or_ctx *c);          /* dual/2d/sch_dual.raw.f90:211-240    */
int or_calc_next_step_implicit(or_ctx *c, int step_internal, double *res, double *omega);
void or_calc_next_step(or_ctx *c);    /* euler/time_euler_viscous.raw.f90    */
void or_calc_next_step_rk2(or_ctx *c, int stage); /* RK2/time_RK2_viscous.raw.f90 */
double or_energy_rms(or_ctx *c);      /* core/check_convergence.f90:17-28    */
void or_copy_field(or_ctx *c, int dst, int src);
double or_dt_grbl(or_ctx *c);
void or_set_fixed_dt(or_ctx *c, double dt);
double or_tt(or_ctx *c);
void or_set_first_dual(or_ctx *c, int first);

/* whole steps, mirroring store/time_scheme/<scheme>/main.main.raw.f90.
 * Return 0 or an SHCFD_E_* code.  res_last/res_first: dual only. */
int or_prepare(or_ctx *c);            /* main.body.raw.f90: part_primitive + set_BC */
int or_step(or_ctx *c, double *res_first, double *res_last);
int or_omp_threads(void);

/* ---- Chemkin chemistry (shcfd_chem_oracle.c) ----------------------------
 * n/y vectors: ns concentrations (mol/cm^3) then T (K). */
int or_set_mechanism(or_ctx *c, const shcfd_mech *m);
int or_chem_error(or_ctx *c);
int or_proceed_reaction(or_ctx *c);   /* chemkin/flow/reactive/thermal_model.f90:45-73 */
void or_chem_fex(const shcfd_mech *m, const double *n, double *dndt);                    /* FJ.f90:152-300 */
void or_chem_fex_var(const shcfd_mech *m, const double *n, double *dndtp, double *dndtm); /* FJ.f90:301-453 */
void or_chem_jex(const shcfd_mech *m, const double *n, double *J);                       /* FJ.f90:454-708 */
void or_chem_rho_rel2abs(const shcfd_mech *m, double p, double T, double *vrho, double *rho, double *E);
int or_chem_calc_t(const shcfd_mech *m, const double *vrho, double rho, double E, double *T, double *out);
int or_chem_chemeq2(const shcfd_mech *m, double *y, double tout, double rtol, int *attempts);
int or_chem_reaction_chemeq2(const shcfd_mech *m, double T, double *vrho, double tout, int *attempts);
int or_chem_reaction_dvode(const shcfd_mech *m, double T, double *vrho, double tout, int *steps);
void or_set_dvode(void *fn);
int or_have_dvode(void);
long or_dvode_calls(void);
int or_chem_point_implicit(const shcfd_mech *m, double T, double *vrho, double dt, int *iters, double *T_out);
int or_dgefa(double *a, int n, int *ipvt);
double or_exp(double x);
double or_log(double x);
void or_dgesl(const double *a, int n, const int *ipvt, double *b);

/* ---- NASA-9 cold flow (shcfd_chem_oracle.c) ------------------------------
 * out = (MWave, kappa, mu, DHi(1:2), vhi(1:2)); *T: start value in, solution
 * out; returns the Newton iterations (NASA/core/sub_chem.f90:489-610) */
int or_cold_flow(const shcfd_nasa9 *m, double y0, double y1, double E, double *T, double *out);
int or_set_nasa9(or_ctx *c, const shcfd_nasa9 *m);

#ifdef __cplusplus
}
#endif
#endif
