/*
 * shcfd_types.h -- plain-C vocabulary shared by the shcfd C-ABI (shcfd.h):
 * field identifiers, boundary-condition kinds, scheme codes and the
 * per-case parameter block.  No functions are declared here.
 *
 * Layout convention for every field that crosses the ABI ("reference
 * layout"): one block at a time, Fortran order with the component index
 * fastest and the reference's two ghost layers,
 *     f(1:K, -1:ni+2, -1:nj+2)
 * i.e. element (k,i,j) (k 0-based, i/j Fortran cell indices) lives at
 *     k + K*((i+1) + (ni+4)*(j+1)).
 * This is the layout of the reference's module arrays
 * (store/core/variable.f90:3-49) restricted to one plane.  Face arrays
 * (TGi(0:ni,1:nj), wHlj(1:ni,0:nj), ...), node arrays (xh(0:ni,0:nj)) and
 * cell-centre arrays use the same ghosted index space; entries outside the
 * reference's declared range are zero.
 */
#ifndef SHCFD_TYPES_H
#define SHCFD_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
enum {
    SHCFD_OK = 0,
    SHCFD_E_ARG = 1,           /* bad argument / size / table row           */
    SHCFD_E_HIP = 2,           /* HIP runtime error                          */
    SHCFD_E_NOMEM = 3,         /* device allocation failed                   */
    SHCFD_E_NAN = 4,           /* NaN in RMS or inner residual               */
    SHCFD_E_OMEGA = 5,         /* dual time: omega < omega_min               */
    SHCFD_E_RESRATE = 6,       /* dual time: res/res1 > ResRateErr           */
    SHCFD_E_TNEWTON = 7,       /* T-from-e Newton iteration limit            */
    SHCFD_E_CHEM = 8,          /* chemistry integrator failure               */
    SHCFD_E_SPIN = 9,          /* sweep wavefront flag wait exceeded bound   */
    SHCFD_E_STATE = 10,        /* call made in the wrong state / scheme      */
    SHCFD_E_TOO_LARGE = 11     /* request exceeds the fixed device scratch of
                                  a no-allocation entry point              */
};

/* ---- time schemes (checkout.inp "time scheme", checkout_flow.py:134-185) -- */
enum {
    SHCFD_SCHEME_EULER = 0,
    SHCFD_SCHEME_RK2 = 1,
    SHCFD_SCHEME_LUSGS = 2,
    SHCFD_SCHEME_NR = 3,         /* dual-time inner loop, plain LU-SGS operator: NR/2d        */
    SHCFD_SCHEME_PRECON = 4,     /* dual-time inner loop, preconditioned operator: precon/2d  */
    SHCFD_SCHEME_DUAL = 5,
    SHCFD_SCHEME_PRECON_LUSGS = 6
};

/* ---- geometry ("dimension" of checkout.inp, checkout_flow.py) ----------- */
enum {
    SHCFD_DIM_PLANAR = 0,        /* store/dim/2d                                              */
    SHCFD_DIM_AXISYMMETRIC = 1   /* store/dim/q2d: x axial, r radial (r >= 0), per radian     */
};

/* ---- MUSCL limiter ------------------------------------------------------- */
enum {
    SHCFD_LIMITER_REFERENCE = 0,  /* s=2DuDl/(Du^2+Dl^2+1e-300), muscl.f90:32 */
    SHCFD_LIMITER_MINMOD = 1      /* addition: minmod slopes                 */
};

/* ---- boundary-condition segment kinds ------------------------------------ */
/* A segment covers cells start..end (Fortran, 1-based, inclusive) along one
 * side of one block and fills the two ghost layers of w and vhi, as the
 * user set_BC loops of the reference samples do
 * (sample/no06.step_flow/condition.f90:31-157,
 *  sample/no05.boundary_layer_ideal/condition.f90). */
enum {
    SHCFD_BC_EXTRAP = 0,     /* g1 = interior, g2 = g1                          */
    SHCFD_BC_REFLECT = 1,    /* as EXTRAP, normal velocity (u on i-sides, v on
                                j-sides) negated in g1; g2 = g1               */
    SHCFD_BC_NOSLIP = 2,     /* as EXTRAP, u and v negated in g1; g2 = g1      */
    SHCFD_BC_FIXED = 3,      /* g1 = g2 = given w(1:dimw) and vhi(1:nY)        */
    SHCFD_BC_INFLOW_TUV = 4, /* T,u,v given, p from interior (ideal gas)        */
    SHCFD_BC_OUTFLOW_P = 5,  /* as EXTRAP, p replaced by a given value in g1    */
    SHCFD_BC_SLIP_NORMAL = 6, /* slip wall for a curved side: velocity mirrored
                                about the face normal (addition)             */
    SHCFD_BC_INFLOW_MIX_P = 7, /* inflow of a given mixture at a given T, p from
                                the interior: g1 = pool w(1:dimw) and
                                vhi(1:nY), rho = (p MW)/(R_uni T), H = pool H
                                (static) + |V|^2/2; pool: w, vhi, MW,
                                R_uni T (sample/no27.chemkin_reaction/
                                condition.f90:53-66)                        */
    SHCFD_BC_INJECT_WALL = 8 /* wall injecting a given mixture at a given T:
                                as INFLOW_MIX_P, with the tangential velocity
                                negated and the normal one 2 vw - interior
                                (u on i-sides, v on j-sides); pool: w, vhi,
                                MW, R_uni T, vw (condition.f90:120-127)     */
};
/* sides */
enum { SHCFD_SIDE_IMIN = 0, SHCFD_SIDE_IMAX = 1, SHCFD_SIDE_JMIN = 2, SHCFD_SIDE_JMAX = 3 };
/* one BC row = 6 ints: block(1-based), side, start, end, kind, pool offset   */
#define SHCFD_BC_ROW 6
/* one cut row = 11 ints, the cut_copro000.inp format
 * (geo,width,order,p1,c1,s1,pm1,p2,c2,s2,pm2), store/cond/core/mod_cond.py:86-91 */
#define SHCFD_CUT_ROW 11

/* device snapshot slots per context (shcfd_snapshot_save) */
#define SHCFD_MAX_SNAPSHOTS 4096

/* ---- field identifiers (upload/download in the reference layout) --------- */
enum {
    SHCFD_F_Q = 0,      /* q(1:dimq)        conservative                     */
    SHCFD_F_QP,         /* qp(1:dimq)                                        */
    SHCFD_F_QPP,        /* qpp(1:dimq)                                       */
    SHCFD_F_W,          /* w(1:dimw)        rho,u,v,p,Y(nY),gamma,H,R,mu     */
    SHCFD_F_VHI,        /* vhi(1:nY)                                         */
    SHCFD_F_DHI,        /* DHi(1:nY)                                         */
    SHCFD_F_WHLI,       /* wHli(1:dimw)  i-face left states                  */
    SHCFD_F_WHRI,
    SHCFD_F_WHLJ,
    SHCFD_F_WHRJ,
    SHCFD_F_TGI,        /* TGi(1:dimq)   inviscid i-face flux                */
    SHCFD_F_TGJ,
    SHCFD_F_TGVI,       /* TGvi(1:dimq)  viscous i-face flux                 */
    SHCFD_F_TGVJ,
    SHCFD_F_VOL,        /* Vol                                               */
    SHCFD_F_DSI,        /* dsi(0:ni,1:nj)                                    */
    SHCFD_F_DSJ,
    SHCFD_F_VNI,        /* vni(1:2)                                          */
    SHCFD_F_VNJ,
    SHCFD_F_DSCI,       /* dsci (cell centre, 0 on ghosts)                   */
    SHCFD_F_DSCJ,
    SHCFD_F_VNCI,       /* vnci(1:2)                                         */
    SHCFD_F_VNCJ,
    SHCFD_F_GEOJACI,    /* geojaci(1:2,1:2)                                  */
    SHCFD_F_GEOJACJ,
    SHCFD_F_X,          /* cell centre x(0:ni+1,0:nj+1)                      */
    SHCFD_F_R,
    SHCFD_F_XH,         /* node xh(0:ni,0:nj)                                */
    SHCFD_F_RH,
    SHCFD_F_DT,         /* dt_mat                                            */
    SHCFD_F_ALPHA,      /* alpha                                             */
    SHCFD_F_PHI,        /* phi (precon / dual)                               */
    SHCFD_F_PRESCA,     /* pre_sca(1:4) (dual)                               */
    SHCFD_F_PRE1,       /* pre1(1:dimq)                                      */
    SHCFD_F_PRE2,       /* pre2(1:dimq)                                      */
    SHCFD_F_DQDT,       /* dqdt(1:dimq) (dual)                               */
    SHCFD_F_AP,         /* Ap(1:dimq,1:dimq)                                 */
    SHCFD_F_AM,
    SHCFD_F_BP,
    SHCFD_F_BM,
    SHCFD_F_DQ,         /* Dq(1:dimq) of the last implicit update            */
    SHCFD_F_AREA,       /* Area: planar cell area (axisymmetric; 0 in 2-D)   */
    SHCFD_F_GEOJACC,    /* geojacc(1:2,1:2) cell-centre metrics (axisym.)    */
    SHCFD_F_COUNT
};

/* ---- thermal models ("thermal model" of checkout.inp) -------------------- */
enum {
    SHCFD_THERMAL_IDEAL = 0,    /* store/therm_lib/ideal                        */
    SHCFD_THERMAL_CHEMKIN = 2,  /* store/therm_lib/chemkin (NASA-7, reactive)   */
    SHCFD_THERMAL_COLD_NASA9 = 5 /* store/therm_lib/NASA/cold: frozen mixing of a
                                   fuel and an oxidizer stream (nY = 2), NASA-9 */
};

/* ---- NASA-9 tables of the cold-flow model ----------------------------------
 * What read_cheminp / set_therm_data / set_trans_data / read_fo_composition
 * produce for it (store/therm_lib/NASA/core/sub_chem.f90:1-300, module chem
 * of mod_chem.raw.f90:17-40), 0-based: per species up to 6 temperature
 * sections of the CEA thermo.inp polynomials (co[0..6] = a1..a7, co[7] = b1,
 * co[8] = b2), the CEA trans.inp viscosity fits of the species that have one
 * (ln mu[micropoise] = A ln T + B/T + C/T^2 + D, in trans.inp order, tr2th =
 * species index) and the two streams in mol/kg. */
#define SHCFD_NASA9_MAX_NS 8
#define SHCFD_NASA9_MAX_SEC 6
#define SHCFD_NASA9_MAX_TSEC 3
typedef struct shcfd_nasa9 {
    int ns, nt;                                               /* species, transport entries          */
    int nsec[SHCFD_NASA9_MAX_NS];                             /* num_sctn                            */
    int nsec_trans[SHCFD_NASA9_MAX_NS];                       /* num_sctn_trans per transport entry   */
    int tr2th[SHCFD_NASA9_MAX_NS];                            /* transport entry -> species           */
    double mw[SHCFD_NASA9_MAX_NS];                            /* g/mol                               */
    double trange[SHCFD_NASA9_MAX_NS][SHCFD_NASA9_MAX_SEC][2];
    double co[SHCFD_NASA9_MAX_NS][SHCFD_NASA9_MAX_SEC][9];
    double trange_trans[SHCFD_NASA9_MAX_NS][SHCFD_NASA9_MAX_TSEC][2];
    double trans[SHCFD_NASA9_MAX_NS][SHCFD_NASA9_MAX_TSEC][4];
    double nf[SHCFD_NASA9_MAX_NS];                            /* fuel stream, mol/kg                 */
    double no[SHCFD_NASA9_MAX_NS];                            /* oxidizer stream, mol/kg             */
} shcfd_nasa9;
/* ---- chemistry source per physical step (the point-implicit slot of
 * store/time_scheme/{LU-SGS/2d,RK2}/main.main.point_implicit.f90) --------- */
enum {
    SHCFD_REACTION_NONE = 0,
    SHCFD_REACTION_CHEMEQ2 = 1, /* reaction_chemeq2 per cell (sub_chem.f90:1684-1819) */
    SHCFD_REACTION_POINT_IMPLICIT = 2 /* one linearised backward-Euler step per
                                  cell: (I - dt J) dn = dt F with the analytic
                                  Jacobian Jex and LINPACK-style LU (dvode.f DGEFA/DGESL) */
};

/* ---- Chemkin mechanism tables --------------------------------------------
 * What read_cheminp / set_trans_data produce
 * (store/therm_lib/chemkin/core/sub_chem.f90:191-598, :788-889; module chem,
 * mod_chem.raw.f90:30-70), with 0-based indices.  Units are the reference's:
 * cgs (mol/cm^3, erg), abe = (ln A, beta, Ea/Rc [K]) with Rc = 1.987 cal/mol/K
 * as a single-precision literal (mod_chem.raw.f90:9), troe = (a, 1/T***,
 * 1/T*, T**), men = efficiency - 1. */
#define SHCFD_CHEM_MAX_NE 8
#define SHCFD_CHEM_MAX_NS 16
#define SHCFD_CHEM_MAX_NR 48
#define SHCFD_CHEM_MAX_TSEC 5
enum {
    SHCFD_RXN_NONE = 0,  /* reverse rate from K_c                          */
    SHCFD_RXN_REV = 1,   /* explicit REV /A b E/                           */
    SHCFD_RXN_LOW = 2,   /* Lindemann fall-off (+M) with LOW /A b E/       */
    SHCFD_RXN_TROE = 3,  /* Troe fall-off                                   */
    SHCFD_RXN_M = 4,     /* +M third body                                   */
    SHCFD_RXN_M_REV = 5  /* +M with REV                                     */
};
/* 0-D chemistry operations of shcfd_chem_eval */
enum {
    SHCFD_CHEM_FEX = 0, SHCFD_CHEM_FEXVAR = 1, SHCFD_CHEM_JEX = 2, SHCFD_CHEM_CALCT = 3, SHCFD_CHEM_CHEMEQ2 = 4,
    SHCFD_CHEM_POINT_IMPLICIT = 5, SHCFD_CHEM_DVODE = 6
};
typedef struct shcfd_mech {
    int ne, ns, nr, nt;                                  /* elements, species, reactions, transport species */
    double mw[SHCFD_CHEM_MAX_NS];                        /* g/mol                                     */
    double inv_mw[SHCFD_CHEM_MAX_NS];
    double tthre[SHCFD_CHEM_MAX_NS][3];                  /* Tlow, Thigh, Tcommon                      */
    double coeff[SHCFD_CHEM_MAX_NS][2][7];               /* NASA-7: [0] T <= Tcommon, [1] above       */
    double es[SHCFD_CHEM_MAX_NS][SHCFD_CHEM_MAX_NE];     /* element counts                            */
    int reverse_off[SHCFD_CHEM_MAX_NR];                  /* Rstate(:,1): 1 for "=>"                   */
    int kind[SHCFD_CHEM_MAX_NR];                         /* Rstate(:,2): SHCFD_RXN_*                  */
    int num_nu[SHCFD_CHEM_MAX_NR][2];                    /* reactant / product count (<= 3 each)      */
    int ind_nu[SHCFD_CHEM_MAX_NR][2][3];                 /* species index per occurrence              */
    int snu[SHCFD_CHEM_MAX_NR];                          /* num_nu[1] - num_nu[0]                     */
    int exist_m[SHCFD_CHEM_MAX_NR];
    int num_m[SHCFD_CHEM_MAX_NR];
    int ind_m[SHCFD_CHEM_MAX_NR][SHCFD_CHEM_MAX_NS];
    double men[SHCFD_CHEM_MAX_NR][SHCFD_CHEM_MAX_NS];
    double abe[SHCFD_CHEM_MAX_NR][3];
    double cabe[SHCFD_CHEM_MAX_NR][3];                   /* REV or LOW parameters                     */
    double troe[SHCFD_CHEM_MAX_NR][4];
    int tr2th[SHCFD_CHEM_MAX_NS];                        /* transport entry -> species index          */
    int nsec_trans[SHCFD_CHEM_MAX_NS];
    double trange[SHCFD_CHEM_MAX_NS][SHCFD_CHEM_MAX_TSEC][2];
    double trans[SHCFD_CHEM_MAX_NS][SHCFD_CHEM_MAX_TSEC][4]; /* ln mu = A lnT + B/T + C/T^2 + D (micropoise) */
    double rtol, atol;                                   /* control_chem.inp tolerances (CHEMEQ2)     */
} shcfd_mech;

/* ---- per-case parameters --------------------------------------------------
 * Build-time choices of the reference's checkout.inp (checkout_flow.py:66-318)
 * and the run-time keys of control.inp (read_control.f90:61-92,
 * muscl.f90:76-116, sch_dual.raw.f90:301-340), with the ideal-gas constants
 * (compile-time KAPPA/RGAS/NU in the reference, checkout_flow.py:217-221)
 * made run-time values. */
typedef struct shcfd_params {
    int scheme;            /* SHCFD_SCHEME_*                                  */
    int time_step;         /* 0 global (dt_grbl), 1 local (dt_mat)            */
    int high_order;        /* 0 upwind, 1 MUSCL                               */
    int limiter;           /* SHCFD_LIMITER_*                                 */
    int viscous;           /* 0 inviscid, 1 viscous no-nV/2d                  */
    int thermal;           /* SHCFD_THERMAL_*                                 */
    int num_internal_loop; /* "InternalLoop Max Number" (N-1 inner iterations)*/
    int reaction;          /* SHCFD_REACTION_*: chemistry source per step     */
    double phi_mus;        /* "MUSCL ON(1)/OFF(0)"                            */
    double k_mus;          /* -1 (order 2) or 1/3 (order 3)                   */
    double cfl;            /* "CFL Number"                                    */
    double cfl_tau;        /* "InternalLoop CFL tau"                          */
    double omega_max;      /* "InternalLoop Omega Max"                        */
    double omega_min;      /* "InternalLoop Omega Min"                        */
    double dqmax;          /* "InternalLoop Dqrate Max"                       */
    double res_rate_warning;
    double res_rate_error;
    double kappa;          /* ratio of specific heats                         */
    double rgas;           /* gas constant                                    */
    double nu;             /* kinematic viscosity (mu = rho*nu)               */
    int dimension;         /* SHCFD_DIM_*                                     */
} shcfd_params;

#ifdef __cplusplus
}
#endif
#endif /* SHCFD_TYPES_H */
