/*
 * shcfd_oracle.c -- CPU restatement of the shared_cfd flow path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and the
 * CPU baseline timed by bench.py.  The product never links or calls this.
 *
 * Every routine follows one reference subroutine (file:line given at each
 * function, paths relative to the reference's store/ directory).  Loop
 * order, operand order and the evaluation order of each expression follow
 * the Fortran text; the file is compiled with -ffp-contract=off so that no
 * multiply-add is fused.  OpenMP is used only where the reference itself
 * parallelises, and every reduction is done in the reference's single-thread
 * order so results do not depend on the thread count.
 */
#include "shcfd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct or_ctx {
    int nb, ny, dimq, dimw;
    int *ni, *nj;
    shcfd_params p;
    int ncomp[SHCFD_F_COUNT];
    double **f[SHCFD_F_COUNT];
    int *cut;
    int ncut;
    int *bc;
    int nbc;
    double *pool;
    int npool;
    double dt_grbl, tt;
    double dt_fixed; /* > 0: the global time step is held at this value */
    int first_dual;
    double res1;
    double *scratch; /* per-cell res contributions, max block size */
    double *bnu;     /* kinematic viscosity per block (default: params nu) */
    shcfd_mech mech; /* Chemkin tables (thermal = SHCFD_THERMAL_CHEMKIN) */
    shcfd_nasa9 nasa9; /* cold-flow tables (thermal = SHCFD_THERMAL_COLD_NASA9) */
    int has_nasa9;
    int has_mech;
    int chem_err;    /* first chemistry failure since the step began (or_step clears it) */
};

/* element (k,i,j) of field fld, block b; i,j are Fortran indices in
 * -1..n+2; k is 0-based */
#define CELL(c, b, i, j) ((size_t)((i) + 1) + (size_t)((c)->ni[b] + 4) * (size_t)((j) + 1))
#define A_(c, fld, b, k, i, j) ((c)->f[fld][b][(size_t)(k) + (size_t)(c)->ncomp[fld] * CELL(c, b, i, j)])
#define Q(k, i, j) A_(c, SHCFD_F_Q, b, k, i, j)
#define QP(k, i, j) A_(c, SHCFD_F_QP, b, k, i, j)
#define QPP(k, i, j) A_(c, SHCFD_F_QPP, b, k, i, j)
#define W(k, i, j) A_(c, SHCFD_F_W, b, k, i, j)
#define VHI(k, i, j) A_(c, SHCFD_F_VHI, b, k, i, j)
#define DHI(k, i, j) A_(c, SHCFD_F_DHI, b, k, i, j)
#define SC(fld, i, j) A_(c, fld, b, 0, i, j)
#define VEC(fld, k, i, j) A_(c, fld, b, k, i, j)

/* 0-based component indices of w (reference 1-based: 1..4, 4+k, indxg..indxMu) */
#define IW_RHO 0
#define IW_U 1
#define IW_V 2
#define IW_P 3
#define IW_Y(k) (4 + (k))
#define IW_G(c) ((c)->dimq + 1)
#define IW_H(c) ((c)->dimq + 2)
#define IW_R(c) ((c)->dimq + 3)
#define IW_MU(c) ((c)->dimq + 4)

static int field_ncomp(int fld, int ny) {
    int dimq = ny + 3, dimw = dimq + 5;
    switch (fld) {
    case SHCFD_F_Q: case SHCFD_F_QP: case SHCFD_F_QPP:
    case SHCFD_F_TGI: case SHCFD_F_TGJ: case SHCFD_F_TGVI: case SHCFD_F_TGVJ:
    case SHCFD_F_PRE1: case SHCFD_F_PRE2: case SHCFD_F_DQDT: case SHCFD_F_DQ:
        return dimq;
    case SHCFD_F_W: case SHCFD_F_WHLI: case SHCFD_F_WHRI: case SHCFD_F_WHLJ: case SHCFD_F_WHRJ:
        return dimw;
    case SHCFD_F_VHI: case SHCFD_F_DHI:
        return ny;
    case SHCFD_F_VNI: case SHCFD_F_VNJ: case SHCFD_F_VNCI: case SHCFD_F_VNCJ:
        return 2;
    case SHCFD_F_GEOJACI: case SHCFD_F_GEOJACJ: case SHCFD_F_GEOJACC: case SHCFD_F_PRESCA:
        return 4;
    case SHCFD_F_AP: case SHCFD_F_AM: case SHCFD_F_BP: case SHCFD_F_BM:
        return dimq * dimq;
    default:
        return 1;
    }
}

or_ctx *or_create(int nb, const int *ni, const int *nj, int ny, const shcfd_params *p) {
    if (nb <= 0 || ny <= 0) return NULL;
    or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
    c->nb = nb;
    c->ny = ny;
    c->dimq = ny + 3;
    c->dimw = c->dimq + 5;
    c->ni = (int *)malloc(sizeof(int) * nb);
    c->nj = (int *)malloc(sizeof(int) * nb);
    size_t maxcells = 0;
    for (int b = 0; b < nb; b++) {
        if (ni[b] < 2 || nj[b] < 2) { free(c->ni); free(c->nj); free(c); return NULL; }
        c->ni[b] = ni[b];
        c->nj[b] = nj[b];
        size_t n = (size_t)(ni[b] + 4) * (size_t)(nj[b] + 4);
        if (n > maxcells) maxcells = n;
    }
    for (int fld = 0; fld < SHCFD_F_COUNT; fld++) {
        c->ncomp[fld] = field_ncomp(fld, ny);
        c->f[fld] = (double **)malloc(sizeof(double *) * nb);
        for (int b = 0; b < nb; b++)
            c->f[fld][b] = (double *)calloc((size_t)c->ncomp[fld] * (size_t)(ni[b] + 4) * (size_t)(nj[b] + 4), sizeof(double));
    }
    c->scratch = (double *)calloc(maxcells, sizeof(double));
    if (p) c->p = *p;
    c->bnu = (double *)malloc(sizeof(double) * nb);
    for (int b = 0; b < nb; b++) c->bnu[b] = c->p.nu;
    c->first_dual = 1;
    return c;
}

void or_destroy(or_ctx *c) {
    if (!c) return;
    for (int fld = 0; fld < SHCFD_F_COUNT; fld++) {
        for (int b = 0; b < c->nb; b++) free(c->f[fld][b]);
        free(c->f[fld]);
    }
    free(c->ni); free(c->nj); free(c->cut); free(c->bc); free(c->pool); free(c->scratch); free(c->bnu);
    free(c);
}

int or_set_params(or_ctx *c, const shcfd_params *p) { c->p = *p; return 0; }
/* per-block kinematic viscosity: an extension for batches of independent
 * cases (the reference has one compile-time NU, thermal_model.raw.f90:4) */
int or_set_block_nu(or_ctx *c, const double *nu) {
    for (int b = 0; b < c->nb; b++) c->bnu[b] = nu[b];
    return 0;
}
double *or_field(or_ctx *c, int fld, int b) {
    if (fld < 0 || fld >= SHCFD_F_COUNT || b < 0 || b >= c->nb) return NULL;
    return c->f[fld][b];
}
int or_field_ncomp(or_ctx *c, int fld) { return (fld < 0 || fld >= SHCFD_F_COUNT) ? 0 : c->ncomp[fld]; }
size_t or_field_size(or_ctx *c, int fld, int b) {
    return (size_t)c->ncomp[fld] * (size_t)(c->ni[b] + 4) * (size_t)(c->nj[b] + 4);
}
double or_dt_grbl(or_ctx *c) { return c->dt_grbl; }
double or_tt(or_ctx *c) { return c->tt; }
void or_set_first_dual(or_ctx *c, int first) { c->first_dual = first; }
int or_omp_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int or_set_cut_table(or_ctx *c, const int *rows, int nrows) {
    free(c->cut);
    c->cut = NULL;
    c->ncut = 0;
    if (nrows <= 0) return 0;
    c->cut = (int *)malloc(sizeof(int) * SHCFD_CUT_ROW * nrows);
    memcpy(c->cut, rows, sizeof(int) * SHCFD_CUT_ROW * nrows);
    c->ncut = nrows;
    return 0;
}

int or_set_bc_table(or_ctx *c, const int *rows, int nrows, const double *pool, int npool) {
    free(c->bc);
    free(c->pool);
    c->bc = NULL; c->pool = NULL; c->nbc = 0; c->npool = 0;
    if (nrows > 0) {
        c->bc = (int *)malloc(sizeof(int) * SHCFD_BC_ROW * nrows);
        memcpy(c->bc, rows, sizeof(int) * SHCFD_BC_ROW * nrows);
        c->nbc = nrows;
    }
    if (npool > 0) {
        c->pool = (double *)malloc(sizeof(double) * npool);
        memcpy(c->pool, pool, sizeof(double) * npool);
        c->npool = npool;
    }
    return 0;
}

void or_copy_field(or_ctx *c, int dst, int src) {
    for (int b = 0; b < c->nb; b++)
        memcpy(c->f[dst][b], c->f[src][b], sizeof(double) * or_field_size(c, src, b));
}

/* ======================================================================== */
/* geometry: dim/2d/geometry.f90                                            */
/* ======================================================================== */
static double area_triangle(double a1, double a2, double b1, double b2, double c1, double c2) {
    /* dim/2d/geometry.f90:170-175 */
    return 0.5 * (((a1 * (b2 - c2)) + (b1 * (c2 - a2))) + (c1 * (a2 - b2)));
}
static double dis_point(double x1, double y1, double x2, double y2) {
    /* dim/2d/geometry.f90:176-181 */
    return sqrt(((x1 - x2) * (x1 - x2)) + ((y1 - y2) * (y1 - y2)));
}

void or_init_geometry(or_ctx *c) {
    const int ax = c->p.dimension == SHCFD_DIM_AXISYMMETRIC;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
#define XH(i, j) SC(SHCFD_F_XH, i, j)
#define RH(i, j) SC(SHCFD_F_RH, i, j)
#define XC(i, j) SC(SHCFD_F_X, i, j)
#define RC(i, j) SC(SHCFD_F_R, i, j)
        /* set_xr, geometry.f90:79-106 (per-plane extents; the reference
         * indexes the ghost column with the whole ni(:) array, which equals
         * this only for a single plane) */
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                XC(i, j) = 0.25 * (((XH(i - 1, j - 1) + XH(i, j - 1)) + XH(i - 1, j)) + XH(i, j));
                RC(i, j) = 0.25 * (((RH(i - 1, j - 1) + RH(i, j - 1)) + RH(i - 1, j)) + RH(i, j));
            }
        for (int j = 1; j <= nj; j++) {
            XC(0, j) = 2.0 * XC(1, j) - XC(2, j);
            RC(0, j) = 2.0 * RC(1, j) - RC(2, j);
            XC(ni + 1, j) = 2.0 * XC(ni, j) - XC(ni - 1, j);
            RC(ni + 1, j) = 2.0 * RC(ni, j) - RC(ni - 1, j);
        }
        for (int i = 0; i <= ni + 1; i++) {
            XC(i, 0) = 2.0 * XC(i, 1) - XC(i, 2);
            RC(i, 0) = 2.0 * RC(i, 1) - RC(i, 2);
            XC(i, nj + 1) = 2.0 * XC(i, nj) - XC(i, nj - 1);
            RC(i, nj + 1) = 2.0 * RC(i, nj) - RC(i, nj - 1);
        }
        /* set_Area_Vol, dim/2d/geometry.f90:107-122 (Area = 0 in 2-D);
         * dim/q2d/geometry.f90:107-122 (Area planar, Vol per radian) */
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double aabd = area_triangle(XH(i - 1, j - 1), RH(i - 1, j - 1), XH(i, j - 1), RH(i, j - 1), XH(i - 1, j), RH(i - 1, j));
                double abcd = area_triangle(XH(i, j - 1), RH(i, j - 1), XH(i, j), RH(i, j), XH(i - 1, j), RH(i - 1, j));
                if (ax) {
                    SC(SHCFD_F_AREA, i, j) = aabd + abcd;
                    SC(SHCFD_F_VOL, i, j) = ((((RH(i - 1, j - 1) + RH(i, j - 1)) + RH(i - 1, j)) * aabd) / 3.0) +
                                            ((((RH(i, j - 1) + RH(i, j)) + RH(i - 1, j)) * abcd) / 3.0);
                } else {
                    SC(SHCFD_F_AREA, i, j) = 0.0;
                    SC(SHCFD_F_VOL, i, j) = aabd + abcd;
                }
            }
        /* set_dsij, dim/2d/geometry.f90:123-140; dim/q2d/geometry.f90:123-141
         * (times the mean node r) */
        for (int j = 0; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double ds = dis_point(XH(i - 1, j), RH(i - 1, j), XH(i, j), RH(i, j));
                SC(SHCFD_F_DSJ, i, j) = ax ? (ds * (RH(i - 1, j) + RH(i, j))) / 2.0 : ds;
            }
        for (int j = 1; j <= nj; j++)
            for (int i = 0; i <= ni; i++) {
                double ds = dis_point(XH(i, j - 1), RH(i, j - 1), XH(i, j), RH(i, j));
                SC(SHCFD_F_DSI, i, j) = ax ? (ds * (RH(i, j - 1) + RH(i, j))) / 2.0 : ds;
            }
        /* set_vn, geometry.f90:141-169 */
        for (int j = 1; j <= nj; j++)
            for (int i = 0; i <= ni; i++) {
                double nx = RH(i, j) - RH(i, j - 1);
                double nr = -(XH(i, j) - XH(i, j - 1));
                double sd = sqrt((nx * nx) + (nr * nr));
                VEC(SHCFD_F_VNI, 0, i, j) = nx / sd;
                VEC(SHCFD_F_VNI, 1, i, j) = nr / sd;
            }
        for (int j = 0; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double nx = -(RH(i, j) - RH(i - 1, j));
                double nr = XH(i, j) - XH(i - 1, j);
                double sd = sqrt((nx * nx) + (nr * nr));
                VEC(SHCFD_F_VNJ, 0, i, j) = nx / sd;
                VEC(SHCFD_F_VNJ, 1, i, j) = nr / sd;
            }
    }
    or_set_vnc_dsc(c);
    or_set_geojac(c);
}

void or_set_vnc_dsc(or_ctx *c) {
    /* time_scheme/LU-SGS/2d/sch_lusgs.raw.f90:143-181; LU-SGS/q2d/
     * sch_lusgs.raw.f90:151-186 (dsci, dscj times the mean r of the two
     * mid-face points, after the normals) */
    const int ax = c->p.dimension == SHCFD_DIM_AXISYMMETRIC;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double a1 = (XH(i - 1, j - 1) + XH(i, j - 1)) * 0.5, a2 = (RH(i - 1, j - 1) + RH(i, j - 1)) * 0.5;
                double b1 = (XH(i, j - 1) + XH(i, j)) * 0.5, b2 = (RH(i, j - 1) + RH(i, j)) * 0.5;
                double c1 = (XH(i - 1, j) + XH(i, j)) * 0.5, c2 = (RH(i - 1, j) + RH(i, j)) * 0.5;
                double d1 = (XH(i - 1, j - 1) + XH(i - 1, j)) * 0.5, d2 = (RH(i - 1, j - 1) + RH(i - 1, j)) * 0.5;
                double dsci = sqrt(((c1 - a1) * (c1 - a1)) + ((c2 - a2) * (c2 - a2)));
                double dscj = sqrt(((b1 - d1) * (b1 - d1)) + ((b2 - d2) * (b2 - d2)));
                VEC(SHCFD_F_VNCI, 0, i, j) = (b1 - d1) / dscj;
                VEC(SHCFD_F_VNCI, 1, i, j) = (b2 - d2) / dscj;
                VEC(SHCFD_F_VNCJ, 0, i, j) = (c1 - a1) / dsci;
                VEC(SHCFD_F_VNCJ, 1, i, j) = (c2 - a2) / dsci;
                if (ax) {
                    dsci = (dsci * (c2 + a2)) * 0.5;
                    dscj = (dscj * (b2 + d2)) * 0.5;
                }
                SC(SHCFD_F_DSCI, i, j) = dsci;
                SC(SHCFD_F_DSCJ, i, j) = dscj;
            }
    }
}

void or_set_geojac(or_ctx *c) {
    /* viscosity/viscous/no-nV/2d/sch_viscous.f90:1-54 */
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
        for (int j = 1; j <= nj; j++)
            for (int i = 0; i <= ni; i++) {
                double dxdxi = XC(i + 1, j) - XC(i, j);
                double dydxi = RC(i + 1, j) - RC(i, j);
                double dxdeta = (((XC(i + 1, j + 1) + XC(i, j + 1)) - XC(i + 1, j - 1)) - XC(i, j - 1)) * 0.25;
                double dydeta = (((RC(i + 1, j + 1) + RC(i, j + 1)) - RC(i + 1, j - 1)) - RC(i, j - 1)) * 0.25;
                double det = (dxdxi * dydeta) - (dydxi * dxdeta);
                VEC(SHCFD_F_GEOJACI, 0, i, j) = dydeta / det;
                VEC(SHCFD_F_GEOJACI, 2, i, j) = -dxdeta / det;
                VEC(SHCFD_F_GEOJACI, 1, i, j) = -dydxi / det;
                VEC(SHCFD_F_GEOJACI, 3, i, j) = dxdxi / det;
            }
        for (int j = 0; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double dxdxi = (((XC(i + 1, j + 1) + XC(i + 1, j)) - XC(i - 1, j + 1)) - XC(i - 1, j)) * 0.25;
                double dydxi = (((RC(i + 1, j + 1) + RC(i + 1, j)) - RC(i - 1, j + 1)) - RC(i - 1, j)) * 0.25;
                double dxdeta = XC(i, j + 1) - XC(i, j);
                double dydeta = RC(i, j + 1) - RC(i, j);
                double det = (dxdxi * dydeta) - (dydxi * dxdeta);
                VEC(SHCFD_F_GEOJACJ, 0, i, j) = dydeta / det;
                VEC(SHCFD_F_GEOJACJ, 2, i, j) = -dxdeta / det;
                VEC(SHCFD_F_GEOJACJ, 1, i, j) = -dydxi / det;
                VEC(SHCFD_F_GEOJACJ, 3, i, j) = dxdxi / det;
            }
        if (c->p.dimension == SHCFD_DIM_AXISYMMETRIC)
            /* cell-centred metrics: viscosity/viscous/no-nV/q2d/sch_viscous.f90:10-28 */
            for (int j = 1; j <= nj; j++)
                for (int i = 1; i <= ni; i++) {
                    double dxdxi = (XC(i + 1, j) - XC(i - 1, j)) * 0.5;
                    double dydxi = (RC(i + 1, j) - RC(i - 1, j)) * 0.5;
                    double dxdeta = (XC(i, j + 1) - XC(i, j - 1)) * 0.5;
                    double dydeta = (RC(i, j + 1) - RC(i, j - 1)) * 0.5;
                    double det = (dxdxi * dydeta) - (dydxi * dxdeta);
                    VEC(SHCFD_F_GEOJACC, 0, i, j) = dydeta / det;
                    VEC(SHCFD_F_GEOJACC, 2, i, j) = -dxdeta / det;
                    VEC(SHCFD_F_GEOJACC, 1, i, j) = -dydxi / det;
                    VEC(SHCFD_F_GEOJACC, 3, i, j) = dxdxi / det;
                }
    }
}
#undef XH
#undef RH
#undef XC
#undef RC

/* ======================================================================== */
/* primitive variables and ideal-gas thermo                                 */
/* ======================================================================== */
void or_set_w(or_ctx *c) {
    /* core/scheme.f90:1-29 */
    int ny = c->ny;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
#pragma omp parallel for
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double tmp = 0.0;
                for (int k = 0; k < ny; k++) tmp = tmp + Q(k, i, j);
                W(IW_RHO, i, j) = tmp;
                tmp = 1.0 / tmp;
                W(IW_U, i, j) = Q(ny, i, j) * tmp;
                W(IW_V, i, j) = Q(ny + 1, i, j) * tmp;
                for (int k = 0; k < ny; k++) W(IW_Y(k), i, j) = Q(k, i, j) * tmp;
            }
    }
}

int or_set_mechanism(or_ctx *c, const shcfd_mech *m) {
    if (!m || m->ns != c->ny) return SHCFD_E_ARG;
    c->mech = *m;
    c->has_mech = 1;
    return 0;
}
int or_chem_error(or_ctx *c) { return c->chem_err; }

/* therm_lib/chemkin/flow/reactive/thermal_model.f90:13-42: per cell, T from
 * e by calc_T starting at the previous p/(rho R), then R, p, gamma, H, mu,
 * DHi, vhi.  R_uni = 8.3144621e3 J/(kmol K) (chem_var). */
static void set_thermo_prop_chemkin(or_ctx *c) {
    int ny = c->ny, err = 0;
    const shcfd_mech *m = &c->mech;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
#pragma omp parallel for reduction(max : err)
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double rho = W(IW_RHO, i, j), u = W(IW_U, i, j), v = W(IW_V, i, j);
                double ei = Q(ny + 2, i, j) / rho - 0.5 * (u * u + v * v);
                double T = W(IW_P, i, j) / (rho * W(IW_R(c), i, j));
                double vrho[SHCFD_CHEM_MAX_NS], out[3 + 2 * SHCFD_CHEM_MAX_NS];
                for (int k = 0; k < ny; k++) vrho[k] = Q(k, i, j);
                if (or_chem_calc_t(m, vrho, rho, ei, &T, out) < 0) err = SHCFD_E_TNEWTON;
                double rg = 8.3144621e3 / out[1];
                W(IW_R(c), i, j) = rg;
                W(IW_P, i, j) = rho * rg * T;
                W(IW_G(c), i, j) = out[0];
                W(IW_H(c), i, j) = (Q(ny + 2, i, j) + W(IW_P, i, j)) / rho;
                W(IW_MU(c), i, j) = out[2];
                for (int k = 0; k < ny; k++) {
                    DHI(k, i, j) = out[3 + k];
                    VHI(k, i, j) = out[3 + ny + k];
                }
            }
    }
    if (err && !c->chem_err) c->chem_err = err;
}

/* proceed_reaction: thermal_model.f90:45-73 -- per interior cell,
 * T = p/(rho R) and reaction_chemeq2(T, q(1:nY), dt_grbl); or, for
 * SHCFD_REACTION_POINT_IMPLICIT, the linearised backward-Euler update. */
int or_proceed_reaction(or_ctx *c) {
    int ny = c->ny, err = 0;
    const shcfd_mech *m = &c->mech;
    double dt = c->dt_grbl;
    if (!c->has_mech) return SHCFD_E_STATE;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
#pragma omp parallel for reduction(max : err)
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double T = W(IW_P, i, j) / W(IW_RHO, i, j) / W(IW_R(c), i, j);
                double vrho[SHCFD_CHEM_MAX_NS];
                for (int k = 0; k < ny; k++) vrho[k] = Q(k, i, j);
                int e = (c->p.reaction == SHCFD_REACTION_POINT_IMPLICIT)
                            ? or_chem_point_implicit(m, T, vrho, dt, NULL, NULL)
                            : or_chem_reaction_chemeq2(m, T, vrho, dt, NULL);
                if (e) err = e > err ? e : err;
                for (int k = 0; k < ny; k++) Q(k, i, j) = vrho[k];
            }
    }
    if (err && !c->chem_err) c->chem_err = err;
    return err;
}

int or_set_nasa9(or_ctx *c, const shcfd_nasa9 *m) {
    if (c->ny != 2) return SHCFD_E_ARG;
    c->nasa9 = *m;
    c->has_nasa9 = 1;
    return 0;
}

/* store/therm_lib/NASA/cold/thermal_model.f90:1-26 */
static void set_thermo_prop_cold(or_ctx *c) {
    const double R_uni = 8.3144621e3; /* core/prmtr.f90:4 */
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
#pragma omp parallel for
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double out[7];
                double ei = Q(4, i, j) / W(IW_RHO, i, j) - 0.5 * (W(IW_U, i, j) * W(IW_U, i, j) + W(IW_V, i, j) * W(IW_V, i, j));
                double T = W(IW_P, i, j) / (W(IW_RHO, i, j) * W(IW_R(c), i, j));
                or_cold_flow(&c->nasa9, W(IW_Y(0), i, j), W(IW_Y(1), i, j), ei, &T, out);
                W(IW_P, i, j) = W(IW_RHO, i, j) * (R_uni / out[0]) * T;
                W(IW_G(c), i, j) = out[1];
                W(IW_H(c), i, j) = (Q(4, i, j) + W(IW_P, i, j)) / W(IW_RHO, i, j);
                W(IW_MU(c), i, j) = out[2];
                W(IW_R(c), i, j) = R_uni / out[0];
                DHI(0, i, j) = out[3];
                DHI(1, i, j) = out[4];
                VHI(0, i, j) = out[5];
                VHI(1, i, j) = out[6];
            }
    }
}

void or_set_thermo_prop(or_ctx *c) {
    if (c->p.thermal == SHCFD_THERMAL_CHEMKIN) {
        set_thermo_prop_chemkin(c);
        return;
    }
    if (c->p.thermal == SHCFD_THERMAL_COLD_NASA9) {
        set_thermo_prop_cold(c);
        return;
    }
    /* therm_lib/ideal/thermal_model.raw.f90:7-33 */
    int ny = c->ny;
    double kap = c->p.kappa, rg = c->p.rgas;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
        double nu = c->bnu[b];
#pragma omp parallel for
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double rho = W(IW_RHO, i, j), u = W(IW_U, i, j), v = W(IW_V, i, j);
                double ei = Q(ny + 2, i, j) / rho - 0.5 * ((u * u) + (v * v));
                double T = (ei * (kap - 1.0)) / rg;
                W(IW_P, i, j) = (rho * rg) * T;
                W(IW_G(c), i, j) = kap;
                W(IW_H(c), i, j) = (Q(ny + 2, i, j) + W(IW_P, i, j)) / rho;
                W(IW_MU(c), i, j) = rho * nu;
                W(IW_R(c), i, j) = rg;
                DHI(0, i, j) = 0.0;
                VHI(0, i, j) = ((kap / (kap - 1.0)) * rg) * T;
            }
    }
}

/* ======================================================================== */
/* boundary conditions                                                       */
/* ======================================================================== */
static void copy_wv(or_ctx *c, int bd, int id, int jd, int bs, int is, int js) {
    int b;
    for (int k = 0; k < c->dimw; k++) {
        b = bs; double v = W(k, is, js);
        b = bd; W(k, id, jd) = v;
    }
    for (int k = 0; k < c->ny; k++) {
        b = bs; double v = VHI(k, is, js);
        b = bd; VHI(k, id, jd) = v;
    }
}

static void section_exchange(or_ctx *c) {
    /* cond/no-nV/condition.tail.f90:2-70; rows use 1-based plane ids */
    for (int r = 0; r < c->ncut; r++) {
        const int *t = c->cut + SHCFD_CUT_ROW * r;
        int geo = t[0], width = t[1], order = t[2];
        int p1 = t[3] - 1, c1 = t[4], s1 = t[5], pm1 = t[6];
        int p2 = t[7] - 1, c2 = t[8], s2 = t[9], pm2 = t[10];
        for (int inc = 0; inc < width; inc++) {
            int a1 = s1 + inc, a2 = s2 + inc * order;
            switch (geo) {
            case 1: /* i-i */
                copy_wv(c, p1, c1 + pm1, a1, p2, c2, a2);
                copy_wv(c, p1, c1 + 2 * pm1, a1, p2, c2 - pm2, a2);
                copy_wv(c, p2, c2 + pm2, a2, p1, c1, a1);
                copy_wv(c, p2, c2 + 2 * pm2, a2, p1, c1 - pm1, a1);
                break;
            case 2: /* i-j */
                copy_wv(c, p1, c1 + pm1, a1, p2, a2, c2);
                copy_wv(c, p1, c1 + 2 * pm1, a1, p2, a2, c2 - pm2);
                copy_wv(c, p2, a2, c2 + pm2, p1, c1, a1);
                copy_wv(c, p2, a2, c2 + 2 * pm2, p1, c1 - pm1, a1);
                break;
            case 3: /* j-i */
                copy_wv(c, p1, a1, c1 + pm1, p2, c2, a2);
                copy_wv(c, p1, a1, c1 + 2 * pm1, p2, c2 - pm2, a2);
                copy_wv(c, p2, c2 + pm2, a2, p1, a1, c1);
                copy_wv(c, p2, c2 + 2 * pm2, a2, p1, a1, c1 - pm1);
                break;
            case 4: /* j-j */
                copy_wv(c, p1, a1, c1 + pm1, p2, a2, c2);
                copy_wv(c, p1, a1, c1 + 2 * pm1, p2, a2, c2 - pm2);
                copy_wv(c, p2, a2, c2 + pm2, p1, a1, c1);
                copy_wv(c, p2, a2, c2 + 2 * pm2, p1, a1, c1 - pm1);
                break;
            default:
                break;
            }
        }
    }
}

/* One BC segment, following the user set_BC loops of the samples
 * (sample/no06.step_flow/condition.f90:52-154,
 *  sample/no05.boundary_layer_ideal/condition.f90, i=1/2 inflow and
 *  p-fixed outflow; SLIP_NORMAL is an addition for curved walls). */
static void bc_segment(or_ctx *c, const int *row) {
    int b = row[0] - 1, side = row[1], s = row[2], e = row[3], kind = row[4], off = row[5];
    int ni = c->ni[b], nj = c->nj[b];
    int n = (side <= SHCFD_SIDE_IMAX) ? nj : ni;
    if (s < 1) s = 1;
    if (e > n) e = n;
    const double *pv = c->pool ? c->pool + off : NULL;
    for (int t = s; t <= e; t++) {
        int ii, jj, ig1, jg1, ig2, jg2;
        switch (side) {
        case SHCFD_SIDE_IMIN: ii = 1; jj = t; ig1 = 0; jg1 = t; ig2 = -1; jg2 = t; break;
        case SHCFD_SIDE_IMAX: ii = ni; jj = t; ig1 = ni + 1; jg1 = t; ig2 = ni + 2; jg2 = t; break;
        case SHCFD_SIDE_JMIN: ii = t; jj = 1; ig1 = t; jg1 = 0; ig2 = t; jg2 = -1; break;
        default: ii = t; jj = nj; ig1 = t; jg1 = nj + 1; ig2 = t; jg2 = nj + 2; break;
        }
        switch (kind) {
        case SHCFD_BC_EXTRAP:
        case SHCFD_BC_REFLECT:
        case SHCFD_BC_NOSLIP:
        case SHCFD_BC_OUTFLOW_P:
        case SHCFD_BC_SLIP_NORMAL:
            for (int k = 0; k < c->dimw; k++) W(k, ig1, jg1) = W(k, ii, jj);
            if (kind == SHCFD_BC_REFLECT) {
                int kv = (side <= SHCFD_SIDE_IMAX) ? IW_U : IW_V;
                W(kv, ig1, jg1) = -W(kv, ii, jj);
            } else if (kind == SHCFD_BC_NOSLIP) {
                W(IW_U, ig1, jg1) = -W(IW_U, ii, jj);
                W(IW_V, ig1, jg1) = -W(IW_V, ii, jj);
            } else if (kind == SHCFD_BC_OUTFLOW_P) {
                W(IW_P, ig1, jg1) = pv[0];
            } else if (kind == SHCFD_BC_SLIP_NORMAL) {
                /* face normal of the boundary face */
                double nx, ny;
                if (side == SHCFD_SIDE_IMIN) { nx = VEC(SHCFD_F_VNI, 0, 0, jj); ny = VEC(SHCFD_F_VNI, 1, 0, jj); }
                else if (side == SHCFD_SIDE_IMAX) { nx = VEC(SHCFD_F_VNI, 0, ni, jj); ny = VEC(SHCFD_F_VNI, 1, ni, jj); }
                else if (side == SHCFD_SIDE_JMIN) { nx = VEC(SHCFD_F_VNJ, 0, ii, 0); ny = VEC(SHCFD_F_VNJ, 1, ii, 0); }
                else { nx = VEC(SHCFD_F_VNJ, 0, ii, nj); ny = VEC(SHCFD_F_VNJ, 1, ii, nj); }
                double u = W(IW_U, ii, jj), v = W(IW_V, ii, jj);
                double un = (nx * u) + (ny * v);
                W(IW_U, ig1, jg1) = u - (2.0 * un) * nx;
                W(IW_V, ig1, jg1) = v - (2.0 * un) * ny;
            }
            for (int k = 0; k < c->ny; k++) VHI(k, ig1, jg1) = VHI(k, ii, jj);
            break;
        case SHCFD_BC_FIXED:
            for (int k = 0; k < c->dimw; k++) W(k, ig1, jg1) = pv[k];
            for (int k = 0; k < c->ny; k++) VHI(k, ig1, jg1) = pv[c->dimw + k];
            break;
        case SHCFD_BC_INFLOW_MIX_P:
        case SHCFD_BC_INJECT_WALL: {
            /* sample/no27.chemkin_reaction/condition.f90:53-66 (oxidizer
             * inflow: w = wo, rho = p*MWo/(R_uni*To), p from the interior,
             * H = wo(H) + |V|^2/2) and :120-127 (fuel-injecting wall:
             * w = wf, u = -u(1), v = 0.02*2 - v(1)) */
            int dw = c->dimw, ny = c->ny;
            for (int k = 0; k < dw; k++) W(k, ig1, jg1) = pv[k];
            for (int k = 0; k < ny; k++) VHI(k, ig1, jg1) = pv[dw + k];
            double pin = W(IW_P, ii, jj);
            W(IW_RHO, ig1, jg1) = (pin * pv[dw + ny]) / pv[dw + ny + 1];
            if (kind == SHCFD_BC_INJECT_WALL) {
                int kn = (side <= SHCFD_SIDE_IMAX) ? IW_U : IW_V, kt = (side <= SHCFD_SIDE_IMAX) ? IW_V : IW_U;
                W(kt, ig1, jg1) = -W(kt, ii, jj);
                W(kn, ig1, jg1) = (pv[dw + ny + 2] * 2.0) - W(kn, ii, jj);
            }
            W(IW_P, ig1, jg1) = pin;
            double u = W(IW_U, ig1, jg1), v = W(IW_V, ig1, jg1);
            W(IW_H(c), ig1, jg1) = pv[IW_H(c)] + 0.5 * ((u * u) + (v * v));
            break;
        }
        case SHCFD_BC_INFLOW_TUV: {
            /* sample/no05.boundary_layer_ideal/condition.f90, i=1/2 block */
            double kap = c->p.kappa, rg = c->p.rgas;
            double T = pv[0], u = pv[1], v = pv[2];
            double rho = W(IW_P, ii, jj) / (rg * T);
            W(IW_RHO, ig1, jg1) = rho;
            W(IW_U, ig1, jg1) = u;
            W(IW_V, ig1, jg1) = v;
            W(IW_P, ig1, jg1) = W(IW_P, ii, jj);
            for (int k = 0; k < c->ny; k++) W(IW_Y(k), ig1, jg1) = 1.0;
            W(IW_G(c), ig1, jg1) = kap;
            W(IW_H(c), ig1, jg1) = ((kap / (kap - 1.0)) * rg) * T + 0.5 * ((u * u) + (v * v));
            W(IW_R(c), ig1, jg1) = rg;
            W(IW_MU(c), ig1, jg1) = c->bnu[b] * rho;
            for (int k = 0; k < c->ny; k++) VHI(k, ig1, jg1) = ((kap / (kap - 1.0)) * rg) * T;
            break;
        }
        default:
            break;
        }
        for (int k = 0; k < c->dimw; k++) W(k, ig2, jg2) = W(k, ig1, jg1);
        for (int k = 0; k < c->ny; k++) VHI(k, ig2, jg2) = VHI(k, ig1, jg1);
    }
}

static void set_corners(or_ctx *c) {
    /* cond/no-nV/condition.tail.f90:331-356 */
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
        for (int pi = -1; pi <= 1; pi += 2) {
            int i = (pi < 0) ? 1 : ni;
            for (int pj = -1; pj <= 1; pj += 2) {
                int j = (pj < 0) ? 1 : nj;
                for (int k = 0; k < c->dimw; k++)
                    W(k, i + pi, j + pj) = (1.0 / 3.0) * ((W(k, i, j) + W(k, i, j + pj)) + W(k, i + pi, j));
                for (int k = 0; k < c->ny; k++)
                    VHI(k, i + pi, j + pj) = (1.0 / 3.0) * ((VHI(k, i, j) + VHI(k, i, j + pj)) + VHI(k, i + pi, j));
            }
        }
    }
}

void or_set_bc(or_ctx *c) {
    /* user set_BC skeleton (sample/no06.step_flow/condition.f90:31-157):
     * section_exchange -> i-side BCs -> j-side BCs -> set_corners */
    section_exchange(c);
    for (int r = 0; r < c->nbc; r++)
        if (c->bc[SHCFD_BC_ROW * r + 1] <= SHCFD_SIDE_IMAX) bc_segment(c, c->bc + SHCFD_BC_ROW * r);
    for (int r = 0; r < c->nbc; r++)
        if (c->bc[SHCFD_BC_ROW * r + 1] > SHCFD_SIDE_IMAX) bc_segment(c, c->bc + SHCFD_BC_ROW * r);
    set_corners(c);
}

/* ======================================================================== */
/* MUSCL / upwind face states: high_order/muscl/muscl.f90:7-73             */
/* ======================================================================== */
static inline void muscl_delta(double Du, double Dl, double k_mus, int limiter, double *du, double *dl) {
    if (limiter == SHCFD_LIMITER_MINMOD) {
        double m;
        if (Du * Dl <= 0.0) m = 0.0;
        else m = (fabs(Du) < fabs(Dl)) ? Du : Dl;
        *du = m;
        *dl = m;
        return;
    }
    double s = ((2.0 * Du) * Dl) / (((Du * Du) + (Dl * Dl)) + 1e-300);
    *du = (0.5 * s) * (((1.0 + k_mus * s) * Du) + ((1.0 - k_mus * s) * Dl));
    *dl = (0.5 * s) * (((1.0 + k_mus * s) * Dl) + ((1.0 - k_mus * s) * Du));
}

void or_set_hv(or_ctx *c) {
    int dimw = c->dimw;
    double phi = c->p.phi_mus, kk = c->p.k_mus;
    int lim = c->p.limiter;
    int muscl = c->p.high_order;
    double hp = phi * 0.5;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
        if (!muscl) {
            /* high_order/upwind/upwind.f90:6-36 */
            for (int j = 1; j <= nj; j++)
                for (int i = 0; i <= ni; i++)
                    for (int k = 0; k < dimw; k++) {
                        VEC(SHCFD_F_WHLI, k, i, j) = W(k, i, j);
                        VEC(SHCFD_F_WHRI, k, i, j) = W(k, i + 1, j);
                    }
            for (int j = 0; j <= nj; j++)
                for (int i = 1; i <= ni; i++)
                    for (int k = 0; k < dimw; k++) {
                        VEC(SHCFD_F_WHLJ, k, i, j) = W(k, i, j);
                        VEC(SHCFD_F_WHRJ, k, i, j) = W(k, i, j + 1);
                    }
            continue;
        }
#pragma omp parallel for
        for (int j = 1; j <= nj; j++)
            for (int i = 0; i <= ni; i++)
                for (int k = 0; k < dimw; k++) {
                    double du0, dl0, du1, dl1;
                    /* delta_ui at cell i, delta_li at cell i+1 */
                    muscl_delta(W(k, i + 1, j) - W(k, i, j), W(k, i, j) - W(k, i - 1, j), kk, lim, &du0, &dl0);
                    muscl_delta(W(k, i + 2, j) - W(k, i + 1, j), W(k, i + 1, j) - W(k, i, j), kk, lim, &du1, &dl1);
                    VEC(SHCFD_F_WHLI, k, i, j) = W(k, i, j) + hp * du0;
                    VEC(SHCFD_F_WHRI, k, i, j) = W(k, i + 1, j) - hp * dl1;
                }
#pragma omp parallel for
        for (int j = 0; j <= nj; j++)
            for (int i = 1; i <= ni; i++)
                for (int k = 0; k < dimw; k++) {
                    double du0, dl0, du1, dl1;
                    muscl_delta(W(k, i, j + 1) - W(k, i, j), W(k, i, j) - W(k, i, j - 1), kk, lim, &du0, &dl0);
                    muscl_delta(W(k, i, j + 2) - W(k, i, j + 1), W(k, i, j + 1) - W(k, i, j), kk, lim, &du1, &dl1);
                    VEC(SHCFD_F_WHLJ, k, i, j) = W(k, i, j) + hp * du0;
                    VEC(SHCFD_F_WHRJ, k, i, j) = W(k, i, j + 1) - hp * dl1;
                }
    }
}

/* ======================================================================== */
/* SLAU inviscid flux: core/scheme.f90:31-259                               */
/* ======================================================================== */
static inline double fsign1(double x) { return (x >= 0.0) ? 1.0 : -1.0; }

static void slau_face(const or_ctx *c, const double *wl, const double *wr, double n1, double n2, double *tg) {
    int ny = c->ny, dimq = c->dimq;
    double rhol = wl[0], ul = wl[1], vl = wl[2], pl = wl[3], gl = wl[dimq + 1];
    double rhor = wr[0], ur = wr[1], vr = wr[2], pr = wr[3], gr = wr[dimq + 1];
    double al = sqrt((gl * pl) / rhol);
    double unl = (n1 * ul) + (n2 * vl);
    double ar = sqrt((gr * pr) / rhor);
    double unr = (n1 * ur) + (n2 * vr);
    double dp = pr - pl;
    double a_bar = 0.5 * (ar + al);
    double p_bar = 0.5 * (pr + pl);
    double ml = unl / a_bar;
    double mr = unr / a_bar;
    double mt = (1.0 / a_bar) * sqrt(0.5 * ((((ul * ul) + (vl * vl)) + (ur * ur)) + (vr * vr)));
    double m_tilde = (1.0 < mt) ? 1.0 : mt;
    double chi = (1.0 - m_tilde) * (1.0 - m_tilde);
    double t = (ml < 0.0) ? ml : 0.0;
    double gp = -((t > -1.0) ? t : -1.0);
    t = (mr > 0.0) ? mr : 0.0;
    double gm = (t < 1.0) ? t : 1.0;
    double g = gp * gm;
    double un_bar_abs = ((fabs(unr) * rhor) + (fabs(unl) * rhol)) / (rhor + rhol);
    double unp = ((1.0 - g) * un_bar_abs) + (g * fabs(unl));
    double unm = ((1.0 - g) * un_bar_abs) + (g * fabs(unr));
    double dm = 0.5 * (((rhol * (unl + unp)) + (rhor * (unr - unm))) - ((chi / a_bar) * dp));
    double betar, betal;
    if (fabs(mr) < 1.0) betar = (0.25 * (2.0 + mr)) * ((mr - 1.0) * (mr - 1.0));
    else betar = 0.5 * (1.0 - fsign1(mr));
    if (fabs(ml) < 1.0) betal = (0.25 * (2.0 - ml)) * ((ml + 1.0) * (ml + 1.0));
    else betal = 0.5 * (1.0 + fsign1(ml));
    double p_tilde = (p_bar + ((0.5 * (betal - betar)) * (pl - pr))) + (((1.0 - chi) * ((betar + betal) - 1.0)) * p_bar);
    double cp = 0.5 * (dm + fabs(dm)), cm = 0.5 * (dm - fabs(dm));
    for (int k = 0; k < dimq; k++) {
        double phil, phir, nk;
        if (k < ny) { phil = wl[4 + k]; phir = wr[4 + k]; nk = 0.0; }
        else if (k == ny) { phil = ul; phir = ur; nk = n1; }
        else if (k == ny + 1) { phil = vl; phir = vr; nk = n2; }
        else { phil = wl[dimq + 2]; phir = wr[dimq + 2]; nk = 0.0; }
        tg[k] = ((cp * phil) + (cm * phir)) + (p_tilde * nk);
    }
}

void or_set_tg(or_ctx *c) {
    int dimw = c->dimw, dimq = c->dimq;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
#pragma omp parallel for
        for (int j = 1; j <= nj; j++) {
            double wl[64], wr[64], tg[32];
            for (int i = 0; i <= ni; i++) {
                for (int k = 0; k < dimw; k++) { wl[k] = VEC(SHCFD_F_WHLI, k, i, j); wr[k] = VEC(SHCFD_F_WHRI, k, i, j); }
                slau_face(c, wl, wr, VEC(SHCFD_F_VNI, 0, i, j), VEC(SHCFD_F_VNI, 1, i, j), tg);
                for (int k = 0; k < dimq; k++) VEC(SHCFD_F_TGI, k, i, j) = tg[k];
            }
        }
#pragma omp parallel for
        for (int j = 0; j <= nj; j++) {
            double wl[64], wr[64], tg[32];
            for (int i = 1; i <= ni; i++) {
                for (int k = 0; k < dimw; k++) { wl[k] = VEC(SHCFD_F_WHLJ, k, i, j); wr[k] = VEC(SHCFD_F_WHRJ, k, i, j); }
                slau_face(c, wl, wr, VEC(SHCFD_F_VNJ, 0, i, j), VEC(SHCFD_F_VNJ, 1, i, j), tg);
                for (int k = 0; k < dimq; k++) VEC(SHCFD_F_TGJ, k, i, j) = tg[k];
            }
        }
    }
}

/* ======================================================================== */
/* viscous face flux: viscosity/viscous/no-nV/2d/sch_viscous.f90:56-272     */
/* ======================================================================== */
void or_set_tgv(or_ctx *c) {
    int ny = c->ny, dimq = c->dimq;
    if (!c->p.viscous) {
        /* viscosity/non-viscous/sch_viscous.f90: TGv = 0 */
        for (int b = 0; b < c->nb; b++) {
            memset(c->f[SHCFD_F_TGVI][b], 0, sizeof(double) * or_field_size(c, SHCFD_F_TGVI, b));
            memset(c->f[SHCFD_F_TGVJ][b], 0, sizeof(double) * or_field_size(c, SHCFD_F_TGVJ, b));
        }
        return;
    }
    const double Sc = 1.0, Pr = 1.0;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
        double *Tdeg = c->scratch;
        size_t P = (size_t)(ni + 4);
#define TD(i, j) Tdeg[(size_t)((i) + 1) + P * (size_t)((j) + 1)]
        for (int j = 0; j <= nj + 1; j++)
            for (int i = 0; i <= ni + 1; i++)
                TD(i, j) = (W(IW_P, i, j) / W(IW_RHO, i, j)) / W(IW_R(c), i, j);
        for (int pass = 0; pass < 2; pass++) {
            int jlo = pass == 0 ? 1 : 0, ilo = pass == 0 ? 0 : 1;
#pragma omp parallel for
            for (int j = jlo; j <= nj; j++)
                for (int i = ilo; i <= ni; i++) {
                    int i1 = pass == 0 ? i + 1 : i, j1 = pass == 0 ? j : j + 1;
                    double rho = (W(IW_RHO, i, j) + W(IW_RHO, i1, j1)) * 0.5;
                    double u = (W(IW_U, i, j) + W(IW_U, i1, j1)) * 0.5;
                    double v = (W(IW_V, i, j) + W(IW_V, i1, j1)) * 0.5;
                    double gamm = (W(IW_G(c), i, j) + W(IW_G(c), i1, j1)) * 0.5;
                    double Rg = (W(IW_R(c), i, j) + W(IW_R(c), i1, j1)) * 0.5;
                    double mu = (W(IW_MU(c), i, j) + W(IW_MU(c), i1, j1)) * 0.5;
                    double lhi[32], dwdxi[32], dwdeta[32], dwdx[32], dwdr[32];
                    for (int k = 0; k < ny; k++) lhi[k] = (VHI(k, i, j) + VHI(k, i1, j1)) * 0.5;
                    double D = (mu / rho) / Sc;
                    double k_heat = (((mu * gamm) / (gamm - 1.0)) * Rg) / Pr;
                    double dudxi, dvdxi, dTdxi, dudeta, dvdeta, dTdeta;
                    const double *gj;
                    if (pass == 0) {
                        dudxi = W(IW_U, i + 1, j) - W(IW_U, i, j);
                        dvdxi = W(IW_V, i + 1, j) - W(IW_V, i, j);
                        dTdxi = TD(i + 1, j) - TD(i, j);
                        for (int k = 0; k < ny; k++) dwdxi[k] = W(IW_Y(k), i + 1, j) - W(IW_Y(k), i, j);
                        dudeta = (((W(IW_U, i + 1, j + 1) + W(IW_U, i, j + 1)) - W(IW_U, i + 1, j - 1)) - W(IW_U, i, j - 1)) * 0.25;
                        dvdeta = (((W(IW_V, i + 1, j + 1) + W(IW_V, i, j + 1)) - W(IW_V, i + 1, j - 1)) - W(IW_V, i, j - 1)) * 0.25;
                        dTdeta = (((TD(i + 1, j + 1) + TD(i, j + 1)) - TD(i + 1, j - 1)) - TD(i, j - 1)) * 0.25;
                        for (int k = 0; k < ny; k++)
                            dwdeta[k] = (((W(IW_Y(k), i + 1, j + 1) + W(IW_Y(k), i, j + 1)) - W(IW_Y(k), i + 1, j - 1)) - W(IW_Y(k), i, j - 1)) * 0.25;
                        gj = &VEC(SHCFD_F_GEOJACI, 0, i, j);
                    } else {
                        dudxi = (((W(IW_U, i + 1, j + 1) + W(IW_U, i + 1, j)) - W(IW_U, i - 1, j + 1)) - W(IW_U, i - 1, j)) * 0.25;
                        dvdxi = (((W(IW_V, i + 1, j + 1) + W(IW_V, i + 1, j)) - W(IW_V, i - 1, j + 1)) - W(IW_V, i - 1, j)) * 0.25;
                        dTdxi = (((TD(i + 1, j + 1) + TD(i + 1, j)) - TD(i - 1, j + 1)) - TD(i - 1, j)) * 0.25;
                        for (int k = 0; k < ny; k++)
                            dwdxi[k] = (((W(IW_Y(k), i + 1, j + 1) + W(IW_Y(k), i + 1, j)) - W(IW_Y(k), i - 1, j + 1)) - W(IW_Y(k), i - 1, j)) * 0.25;
                        dudeta = W(IW_U, i, j + 1) - W(IW_U, i, j);
                        dvdeta = W(IW_V, i, j + 1) - W(IW_V, i, j);
                        dTdeta = TD(i, j + 1) - TD(i, j);
                        for (int k = 0; k < ny; k++) dwdeta[k] = W(IW_Y(k), i, j + 1) - W(IW_Y(k), i, j);
                        gj = &VEC(SHCFD_F_GEOJACJ, 0, i, j);
                    }
                    /* geojac(a,b) stored at a + 2b (0-based) */
                    double g11 = gj[0], g21 = gj[1], g12 = gj[2], g22 = gj[3];
                    double dudx = (dudxi * g11) + (dudeta * g21);
                    double dvdx = (dvdxi * g11) + (dvdeta * g21);
                    double dTdx = (dTdxi * g11) + (dTdeta * g21);
                    double dudr = (dudxi * g12) + (dudeta * g22);
                    double dvdr = (dvdxi * g12) + (dvdeta * g22);
                    double dTdr = (dTdxi * g12) + (dTdeta * g22);
                    double hidwdx = 0.0, hidwdr = 0.0;
                    for (int k = 0; k < ny; k++) {
                        dwdx[k] = (dwdxi[k] * g11) + (dwdeta[k] * g21);
                        hidwdx = hidwdx + (dwdx[k] * lhi[k]);
                        dwdr[k] = (dwdxi[k] * g12) + (dwdeta[k] * g22);
                        hidwdr = hidwdr + (dwdr[k] * lhi[k]);
                    }
                    double divu = dudx + dvdr;
                    /* q2d: + v / r(i,j), the centre radius of the face's lower-index
                     * cell (viscous/no-nV/q2d/sch_viscous.f90:173-174, :262-263) */
                    if (c->p.dimension == SHCFD_DIM_AXISYMMETRIC) divu = divu + (v / SC(SHCFD_F_R, i, j));
                    double t_xx = -(((2.0 / 3.0) * mu) * divu) + ((2.0 * mu) * dudx);
                    double t_rr = -(((2.0 / 3.0) * mu) * divu) + ((2.0 * mu) * dvdr);
                    double t_xr = mu * (dvdx + dudr);
                    double qx = -(k_heat * dTdx) - ((rho * D) * hidwdx);
                    double qr = -(k_heat * dTdr) - ((rho * D) * hidwdr);
                    double Ev[32], Fv[32];
                    for (int k = 0; k < ny; k++) Ev[k] = (rho * D) * dwdx[k];
                    Ev[ny] = t_xx;
                    Ev[ny + 1] = t_xr;
                    Ev[ny + 2] = ((u * t_xx) + (v * t_xr)) - qx;
                    for (int k = 0; k < ny; k++) Fv[k] = (rho * D) * dwdr[k];
                    Fv[ny] = t_xr;
                    Fv[ny + 1] = t_rr;
                    Fv[ny + 2] = ((u * t_xr) + (v * t_rr)) - qr;
                    int fvn = pass == 0 ? SHCFD_F_VNI : SHCFD_F_VNJ;
                    int fout = pass == 0 ? SHCFD_F_TGVI : SHCFD_F_TGVJ;
                    double n1 = VEC(fvn, 0, i, j), n2 = VEC(fvn, 1, i, j);
                    for (int k = 0; k < dimq; k++) VEC(fout, k, i, j) = (Ev[k] * n1) + (Fv[k] * n2);
                }
        }
#undef TD
    }
}

/* ======================================================================== */
/* local / global time step: time_step/{local,global}/set_dt.f90            */
/* ======================================================================== */
double or_set_dt(or_ctx *c) {
    double cfl = c->p.cfl;
    double tmp = 1e300;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double rho = W(IW_RHO, i, j), u = W(IW_U, i, j), v = W(IW_V, i, j), p = W(IW_P, i, j);
                double a = sqrt((W(IW_G(c), i, j) * p) / rho);
                double uds = 0.0, un;
                un = (VEC(SHCFD_F_VNI, 0, i, j) * u) + (VEC(SHCFD_F_VNI, 1, i, j) * v);
                uds = uds + (fabs(un) + a) * SC(SHCFD_F_DSI, i, j);
                un = (VEC(SHCFD_F_VNI, 0, i - 1, j) * u) + (VEC(SHCFD_F_VNI, 1, i - 1, j) * v);
                uds = uds + (fabs(un) + a) * SC(SHCFD_F_DSI, i - 1, j);
                un = (VEC(SHCFD_F_VNJ, 0, i, j) * u) + (VEC(SHCFD_F_VNJ, 1, i, j) * v);
                uds = uds + (fabs(un) + a) * SC(SHCFD_F_DSJ, i, j);
                un = (VEC(SHCFD_F_VNJ, 0, i, j - 1) * u) + (VEC(SHCFD_F_VNJ, 1, i, j - 1) * v);
                uds = uds + (fabs(un) + a) * SC(SHCFD_F_DSJ, i, j - 1);
                double nr = W(IW_MU(c), i, j) / rho;
                double dt = (cfl * SC(SHCFD_F_VOL, i, j)) / ((uds > nr) ? uds : nr);
                SC(SHCFD_F_DT, i, j) = dt;
                if (tmp > dt) tmp = dt;
            }
    }
    if (c->p.time_step == 0) {
        c->dt_grbl = (c->dt_fixed > 0.0) ? c->dt_fixed : tmp;
        c->tt = c->tt + c->dt_grbl;
    }
    return tmp;
}

void or_set_fixed_dt(or_ctx *c, double dt) { c->dt_fixed = dt; }

/* ======================================================================== */
/* flux Jacobians: LU-SGS/2d/sch_lusgs.raw.f90:1-141,                       */
/* preconLU-SGS/2d/sch_precon.raw.f90:1-202, dual/2d/sch_dual.raw.f90:1-209 */
/* ======================================================================== */
/* A and B (column-major, A[k + l*dimq] = A(k,l)) at one cell */
static void euler_jacobians(const or_ctx *c, const double *wc, const double *dhi, double *A, double *B) {
    int ny = c->ny, dq = c->dimq;
    double u = wc[1], v = wc[2], gamm = wc[dq + 1], H = wc[dq + 2];
    double gmmad = gamm - 1.0;
    double Vs = (u * u) + (v * v);
    double tA[32], tB[32];
#define AM_(k, l) A[(k) + dq * (l)]
#define BM_(k, l) B[(k) + dq * (l)]
    int a1 = ny, a2 = ny + 1, a3 = ny + 2;
    AM_(a1, a1) = u * (3.0 - gamm);
    AM_(a2, a1) = v;
    AM_(a3, a1) = H - (u * u) * gmmad;
    AM_(a1, a2) = -(v * gmmad);
    AM_(a2, a2) = u;
    AM_(a3, a2) = -((u * v) * gmmad);
    AM_(a1, a3) = gmmad;
    AM_(a2, a3) = 0.0;
    AM_(a3, a3) = u * gamm;
    for (int k = 0; k < ny; k++) {
        tA[k] = -(wc[4 + k] * u);
        AM_(k, a1) = wc[4 + k];
        AM_(k, a2) = 0.0;
        AM_(k, a3) = 0.0;
    }
    tA[a1] = -(u * u) + (0.5 * gmmad) * Vs;
    tA[a2] = -(u * v);
    tA[a3] = u * (-H + (0.5 * gmmad) * Vs);
    BM_(a1, a1) = v;
    BM_(a2, a1) = -(u * gmmad);
    BM_(a3, a1) = -((u * v) * gmmad);
    BM_(a1, a2) = u;
    BM_(a2, a2) = v * (3.0 - gamm);
    BM_(a3, a2) = H - (v * v) * gmmad;
    BM_(a1, a3) = 0.0;
    BM_(a2, a3) = gmmad;
    BM_(a3, a3) = v * gamm;
    for (int k = 0; k < ny; k++) {
        tB[k] = -(wc[4 + k] * v);
        BM_(k, a1) = 0.0;
        BM_(k, a2) = wc[4 + k];
        BM_(k, a3) = 0.0;
    }
    tB[a1] = -(v * u);
    tB[a2] = -(v * v) + (0.5 * gmmad) * Vs;
    tB[a3] = v * (-H + (0.5 * gmmad) * Vs);
    for (int k = 0; k < ny; k++) {
        for (int l = 0; l < ny; l++) {
            AM_(l, k) = tA[l];
            BM_(l, k) = tB[l];
        }
        AM_(k, k) = tA[k] + u;
        AM_(a1, k) = tA[a1] + dhi[k];
        AM_(a2, k) = tA[a2];
        AM_(a3, k) = tA[a3] + dhi[k] * u;
        BM_(k, k) = tB[k] + v;
        BM_(a1, k) = tB[a1];
        BM_(a2, k) = tB[a2] + dhi[k];
        BM_(a3, k) = tB[a3] + dhi[k] * v;
    }
#undef AM_
#undef BM_
}

void or_set_abpm(or_ctx *c) {
    int ny = c->ny, dq = c->dimq, dw = c->dimw;
    int scheme = c->p.scheme;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
#pragma omp parallel for
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double wc[64], dhi[32], A[1024], B[1024];
                for (int k = 0; k < dw; k++) wc[k] = W(k, i, j);
                for (int k = 0; k < ny; k++) dhi[k] = DHI(k, i, j);
                double rho = wc[0], u = wc[1], v = wc[2], p = wc[3], gamm = wc[dq + 1], H = wc[dq + 2];
                double cs = sqrt((gamm * p) / rho);
                double gmmad = gamm - 1.0;
                double Vs = (u * u) + (v * v);
                double n1a = VEC(SHCFD_F_VNCI, 0, i, j), n2a = VEC(SHCFD_F_VNCI, 1, i, j);
                double n1b = VEC(SHCFD_F_VNCJ, 0, i, j), n2b = VEC(SHCFD_F_VNCJ, 1, i, j);
                double unA = (n1a * u) + (n2a * v);
                double unB = (n1b * u) + (n2b * v);
                double dsci = SC(SHCFD_F_DSCI, i, j), dscj = SC(SHCFD_F_DSCJ, i, j);
                double vol = SC(SHCFD_F_VOL, i, j);
                double nuA, nuB;
                double dtv = (c->p.time_step == 0) ? c->dt_grbl : SC(SHCFD_F_DT, i, j);
                double pre1[32], pre2[32], pscale = 1.0;
                int precond = (scheme == SHCFD_SCHEME_DUAL || scheme == SHCFD_SCHEME_PRECON_LUSGS ||
                               scheme == SHCFD_SCHEME_PRECON);
                if (!precond) {
                    /* sch_lusgs.raw.f90:38-41, sch_NR.raw.f90:40-41 */
                    nuA = fabs(unA) + cs;
                    nuB = fabs(unB) + cs;
                    if (c->p.dimension == SHCFD_DIM_AXISYMMETRIC) {
                        /* dpdq, kept in pre2: LU-SGS/q2d/sch_lusgs.raw.f90:43-49,
                         * NR/q2d/sch_NR.raw.f90:43-49 */
                        for (int k = 0; k < ny; k++) VEC(SHCFD_F_PRE2, k, i, j) = ((0.5 * gmmad) * Vs) + dhi[k];
                        VEC(SHCFD_F_PRE2, ny, i, j) = -(gmmad * u);
                        VEC(SHCFD_F_PRE2, ny + 1, i, j) = -(gmmad * v);
                        VEC(SHCFD_F_PRE2, ny + 2, i, j) = gmmad;
                    }
                } else {
                    /* sch_precon.raw.f90:44-59 / sch_dual.raw.f90:44-59 */
                    double dmin = (dsci < dscj) ? dsci : dscj;
                    double Vv = (wc[dq + 4] / rho) / dmin;
                    double sv = sqrt(Vs);
                    double Ur = (Vv > sv) ? Vv : sv;
                    Ur = (Ur < cs) ? Ur : cs;
                    double phi = (1.0 / (Ur * Ur)) - (1.0 / (cs * cs));
                    SC(SHCFD_F_PHI, i, j) = phi;
                    double rr = Ur / cs;
                    double alp = 0.5 * (1.0 - rr * rr);
                    double ud = unA * (1.0 - alp);
                    double cd = sqrt(((alp * alp) * (unA * unA)) + (Ur * Ur));
                    nuA = fabs(ud) + cd;
                    ud = unB * (1.0 - alp);
                    cd = sqrt(((alp * alp) * (unB * unB)) + (Ur * Ur));
                    nuB = fabs(ud) + cd;
                    double uds = (nuA * (SC(SHCFD_F_DSI, i, j) + SC(SHCFD_F_DSI, i - 1, j))) +
                                 (nuB * (SC(SHCFD_F_DSJ, i, j) + SC(SHCFD_F_DSJ, i, j - 1)));
                    double nr = wc[dq + 4] / rho;
                    double mx = (uds > nr) ? uds : nr;
                    if (scheme == SHCFD_SCHEME_PRECON_LUSGS || scheme == SHCFD_SCHEME_PRECON) {
                        /* precon/2d/sch_precon.raw.f90:58-68 (dt_mat and the phh-scaled pre1) */
                        double dtm = (c->p.cfl * vol) / mx;
                        SC(SHCFD_F_DT, i, j) = dtm;
                        dtv = dtm;
                        double phh = -1.0 / ((cs * cs) + (1.0 / phi));
                        for (int k = 0; k < ny; k++) pre1[k] = phh * wc[4 + k];
                        pre1[ny] = phh * u;
                        pre1[ny + 1] = phh * v;
                        pre1[ny + 2] = phh * H;
                    } else {
                        double dtau = (c->p.cfl_tau * vol) / mx;
                        double alpha = 1.0 / (((vol / dtau) + (dsci * nuA)) + (dscj * nuB));
                        SC(SHCFD_F_ALPHA, i, j) = alpha;
                        double beta = ((3.0 * vol) / (2.0 * dtv)) * alpha;
                        double ip = 1.0 / phi;
                        VEC(SHCFD_F_PRESCA, 0, i, j) = 1.0 + beta;
                        VEC(SHCFD_F_PRESCA, 1, i, j) = -1.0 / (ip + (cs * cs));
                        VEC(SHCFD_F_PRESCA, 2, i, j) = (-1.0 / (ip + (cs * cs))) / ((1.0 / beta) + 1.0);
                        VEC(SHCFD_F_PRESCA, 3, i, j) = beta / (((1.0 + beta) / phi) + (cs * cs));
                        pscale = VEC(SHCFD_F_PRESCA, 1, i, j);
                        for (int k = 0; k < ny; k++) pre1[k] = wc[4 + k];
                        pre1[ny] = u;
                        pre1[ny + 1] = v;
                        pre1[ny + 2] = H;
                    }
                    for (int k = 0; k < ny; k++) pre2[k] = (0.5 * gmmad) * Vs + dhi[k];
                    pre2[ny] = -(gmmad * u);
                    pre2[ny + 1] = -(gmmad * v);
                    pre2[ny + 2] = gmmad;
                    for (int k = 0; k < dq; k++) {
                        VEC(SHCFD_F_PRE1, k, i, j) = pre1[k];
                        VEC(SHCFD_F_PRE2, k, i, j) = pre2[k];
                    }
                }
                euler_jacobians(c, wc, dhi, A, B);
                if (precond) {
                    /* Gamma^-1 rank-1 update: sch_precon.raw.f90:149-170,
                     * sch_dual.raw.f90:156-181 */
                    for (int pass = 0; pass < 2; pass++) {
                        double *M = pass == 0 ? A : B;
                        for (int k = 0; k < dq; k++) {
                            double tmp = 0.0;
                            for (int l = 0; l < dq; l++) tmp = tmp + pre2[l] * M[l + dq * k];
                            if (scheme == SHCFD_SCHEME_DUAL) tmp = pscale * tmp;
                            for (int l = 0; l < dq; l++) M[l + dq * k] = M[l + dq * k] + pre1[l] * tmp;
                        }
                    }
                }
                double *Ap = &VEC(SHCFD_F_AP, 0, i, j), *Am = &VEC(SHCFD_F_AM, 0, i, j);
                double *Bp = &VEC(SHCFD_F_BP, 0, i, j), *Bm = &VEC(SHCFD_F_BM, 0, i, j);
                for (int l = 0; l < dq; l++)
                    for (int k = 0; k < dq; k++) {
                        int e = k + dq * l;
                        double at = (n1a * A[e]) + (n2a * B[e]);
                        double bt = (n1b * A[e]) + (n2b * B[e]);
                        Ap[e] = at * 0.5;
                        Am[e] = at * 0.5;
                        Bp[e] = bt * 0.5;
                        Bm[e] = bt * 0.5;
                    }
                for (int k = 0; k < dq; k++) {
                    int e = k + dq * k;
                    Ap[e] = Ap[e] + nuA * 0.5;
                    Am[e] = Am[e] - nuA * 0.5;
                    Bp[e] = Bp[e] + nuB * 0.5;
                    Bm[e] = Bm[e] - nuB * 0.5;
                }
                if (scheme == SHCFD_SCHEME_NR || scheme == SHCFD_SCHEME_PRECON)
                    /* BDF2 diagonal: NR/2d/sch_NR.raw.f90:134-136, precon/2d/sch_precon.raw.f90:194-196 */
                    SC(SHCFD_F_ALPHA, i, j) = 1.0 / ((((3.0 * vol) / (2.0 * dtv)) + (dsci * nuA)) + (dscj * nuB));
                else if (scheme != SHCFD_SCHEME_DUAL)
                    SC(SHCFD_F_ALPHA, i, j) = 1.0 / (((vol / dtv) + (dsci * nuA)) + (dscj * nuB));
            }
    }
}

void or_set_dqdt(or_ctx *c) {
    /* dual/2d/sch_dual.raw.f90:211-240 */
    int dq = c->dimq;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double dtv = (c->p.time_step == 0) ? c->dt_grbl : SC(SHCFD_F_DT, i, j);
                for (int k = 0; k < dq; k++) {
                    if (c->first_dual)
                        VEC(SHCFD_F_DQDT, k, i, j) = (Q(k, i, j) - QP(k, i, j)) / dtv;
                    else
                        VEC(SHCFD_F_DQDT, k, i, j) = (((3.0 * Q(k, i, j)) - (4.0 * QP(k, i, j))) + QPP(k, i, j)) / (2.0 * dtv);
                }
            }
    }
}

/* ======================================================================== */
/* implicit update: LU-SGS/2d/time_lusgs_viscous.f90:1-65,                  */
/* preconLU-SGS/2d/time_precon_viscous.f90:1-66,                            */
/* NR/2d/time_NR_viscous.f90:1-102, precon/2d/time_precon_viscous.f90:1-103, */
/* dual/2d/time_dual_viscous.f90:1-125                                      */
/* ======================================================================== */
/* Sq + Svq of the radial momentum of an axisymmetric cell: the pressure
 * (dim/q2d/Sq.f90:1-18) minus the hoop stress of viscous flow
 * (viscosity/viscous/no-nV/q2d/sch_viscous.f90:290-319) */
static double radial_source(or_ctx *c, int b, int i, int j) {
    double svq = 0.0;
    if (c->p.viscous) {
        double v = W(IW_V, i, j), mu = W(IW_MU(c), i, j), r = SC(SHCFD_F_R, i, j);
        double dudxi = (W(IW_U, i + 1, j) - W(IW_U, i - 1, j)) * 0.5;
        double dvdxi = (W(IW_V, i + 1, j) - W(IW_V, i - 1, j)) * 0.5;
        double dudeta = (W(IW_U, i, j + 1) - W(IW_U, i, j - 1)) * 0.5;
        double dvdeta = (W(IW_V, i, j + 1) - W(IW_V, i, j - 1)) * 0.5;
        double dudx = (dudxi * VEC(SHCFD_F_GEOJACC, 0, i, j)) + (dudeta * VEC(SHCFD_F_GEOJACC, 1, i, j));
        double dvdr = (dvdxi * VEC(SHCFD_F_GEOJACC, 2, i, j)) + (dvdeta * VEC(SHCFD_F_GEOJACC, 3, i, j));
        double divu = (dudx + dvdr) + (v / r);
        double t_tt = -(((2.0 / 3.0) * mu) * divu) + (((2.0 * mu) * v) / r);
        svq = -t_tt;
    }
    return W(IW_P, i, j) + svq;
}

/* flux balance of one cell: time_lusgs_viscous.f90:12-22 (Area = 0 in 2-D);
 * LU-SGS/q2d/time_lusgs_viscous.f90:15-20: + Area*(Sq+Svq), nonzero on the
 * radial momentum (component ny+1, 0-based) only */
static inline void flux_balance(or_ctx *c, int b, int i, int j, double *out) {
    int dq = c->dimq;
    const int ax = c->p.dimension == SHCFD_DIM_AXISYMMETRIC;
    const double area = ax ? SC(SHCFD_F_AREA, i, j) : 0.0;
    const double src = ax ? radial_source(c, b, i, j) : 0.0;
    for (int k = 0; k < dq; k++) {
        double s = ((((SC(SHCFD_F_DSI, i - 1, j) * (VEC(SHCFD_F_TGI, k, i - 1, j) - VEC(SHCFD_F_TGVI, k, i - 1, j))) -
                      (SC(SHCFD_F_DSI, i, j) * (VEC(SHCFD_F_TGI, k, i, j) - VEC(SHCFD_F_TGVI, k, i, j)))) +
                     (SC(SHCFD_F_DSJ, i, j - 1) * (VEC(SHCFD_F_TGJ, k, i, j - 1) - VEC(SHCFD_F_TGVJ, k, i, j - 1)))) -
                    (SC(SHCFD_F_DSJ, i, j) * (VEC(SHCFD_F_TGJ, k, i, j) - VEC(SHCFD_F_TGVJ, k, i, j))));
        out[k] = ax ? s + (area * (k == c->ny + 1 ? src : 0.0)) : s + 0.0;
    }
}

static inline void matvec(int dq, const double *M, const double *x, double *y) {
    for (int k = 0; k < dq; k++) {
        double s = 0.0;
        for (int l = 0; l < dq; l++) s = s + M[k + dq * l] * x[l];
        y[k] = s;
    }
}

static inline double dotp(int dq, const double *a, const double *x) {
    double s = 0.0;
    for (int l = 0; l < dq; l++) s = s + a[l] * x[l];
    return s;
}

int or_calc_next_step_implicit(or_ctx *c, int step_internal, double *res_out, double *omega_out) {
    int dq = c->dimq, ny = c->ny;
    int scheme = c->p.scheme;
    int dual = scheme == SHCFD_SCHEME_DUAL; /* S-scaled three-stage sweeps */
    /* inner loop of dual time: dq/dt term, res and the omega-limited update */
    int inner = dual || scheme == SHCFD_SCHEME_NR || scheme == SHCFD_SCHEME_PRECON;
    int prec = scheme == SHCFD_SCHEME_PRECON_LUSGS || scheme == SHCFD_SCHEME_PRECON;
    const int ax = c->p.dimension == SHCFD_DIM_AXISYMMETRIC;
    double omega = 1e300, res = 0.0;
    int err = 0;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
#define DQ(k, i, j) VEC(SHCFD_F_DQ, k, i, j)
        memset(c->f[SHCFD_F_DQ][b], 0, sizeof(double) * or_field_size(c, SHCFD_F_DQ, b));
        /* RHS */
#pragma omp parallel for
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double s[32];
                flux_balance(c, b, i, j, s);
                if (dual) {
                    double d[32];
                    for (int k = 0; k < dq; k++) d[k] = s[k] - SC(SHCFD_F_VOL, i, j) * VEC(SHCFD_F_DQDT, k, i, j);
                    double dt = dotp(dq, &VEC(SHCFD_F_PRE2, 0, i, j), d);
                    double ps2 = VEC(SHCFD_F_PRESCA, 1, i, j);
                    for (int k = 0; k < dq; k++) DQ(k, i, j) = d[k] + (ps2 * dt) * VEC(SHCFD_F_PRE1, k, i, j);
                } else if (prec) {
                    /* time_precon_viscous.f90:24-26 (precon/2d), :24 (preconLU-SGS/2d) */
                    double dt = dotp(dq, &VEC(SHCFD_F_PRE2, 0, i, j), s);
                    for (int k = 0; k < dq; k++) {
                        double x = s[k] + VEC(SHCFD_F_PRE1, k, i, j) * dt;
                        DQ(k, i, j) = inner ? x - SC(SHCFD_F_VOL, i, j) * VEC(SHCFD_F_DQDT, k, i, j) : x;
                    }
                } else {
                    /* time_lusgs_viscous.f90:21, time_NR_viscous.f90:24 */
                    for (int k = 0; k < dq; k++)
                        DQ(k, i, j) = inner ? s[k] - SC(SHCFD_F_VOL, i, j) * VEC(SHCFD_F_DQDT, k, i, j) : s[k];
                }
            }
        /* forward sweep on hyperplanes sm=i+j ascending (ghost Dq = 0) */
        for (int sm = 2; sm <= ni + nj; sm++) {
            int ilo = (1 > sm - nj) ? 1 : sm - nj, ihi = (sm - 1 < ni) ? sm - 1 : ni;
#pragma omp parallel for
            for (int i = ilo; i <= ihi; i++) {
                int j = sm - i;
                double ya[32], yb[32];
                matvec(dq, &VEC(SHCFD_F_AP, 0, i - 1, j), &DQ(0, i - 1, j), ya);
                matvec(dq, &VEC(SHCFD_F_BP, 0, i, j - 1), &DQ(0, i, j - 1), yb);
                double dci = SC(SHCFD_F_DSCI, i - 1, j), dcj = SC(SHCFD_F_DSCJ, i, j - 1);
                double al = SC(SHCFD_F_ALPHA, i, j);
                for (int k = 0; k < dq; k++) {
                    double sm_k = (dci * ya[k]) + (dcj * yb[k]);
                    DQ(k, i, j) = al * (DQ(k, i, j) + sm_k);
                }
                if (dual) {
                    double d = dotp(dq, &VEC(SHCFD_F_PRE2, 0, i, j), &DQ(0, i, j));
                    double ps4 = VEC(SHCFD_F_PRESCA, 3, i, j), ips1 = 1.0 / VEC(SHCFD_F_PRESCA, 0, i, j);
                    for (int k = 0; k < dq; k++) DQ(k, i, j) = DQ(k, i, j) + (ps4 * d) * VEC(SHCFD_F_PRE1, k, i, j);
                    for (int k = 0; k < dq; k++) DQ(k, i, j) = ips1 * DQ(k, i, j);
                }
                if (ax) {
                    /* implicit pressure source: LU-SGS/q2d/time_lusgs_viscous.f90:38-44,
                     * NR/q2d/time_NR_viscous.f90:43-49 (dpdq kept in pre2) */
                    const double *dpdq = &VEC(SHCFD_F_PRE2, 0, i, j);
                    double tmp = 0.0;
                    for (int l = 0; l <= ny; l++) tmp = tmp + (DQ(l, i, j) * dpdq[l]);
                    tmp = tmp + (DQ(ny + 2, i, j) * dpdq[ny + 2]);
                    tmp = DQ(ny + 1, i, j) + ((al * SC(SHCFD_F_AREA, i, j)) * tmp);
                    DQ(ny + 1, i, j) = tmp / (1.0 - ((al * SC(SHCFD_F_AREA, i, j)) * dpdq[ny + 1]));
                }
            }
        }
        if (dual) {
            /* second step: time_dual_viscous.f90:53-60 */
            for (int j = 1; j <= nj; j++)
                for (int i = 1; i <= ni; i++) {
                    double d = dotp(dq, &VEC(SHCFD_F_PRE2, 0, i, j), &DQ(0, i, j));
                    double ps3 = VEC(SHCFD_F_PRESCA, 2, i, j), ps1 = VEC(SHCFD_F_PRESCA, 0, i, j);
                    for (int k = 0; k < dq; k++) DQ(k, i, j) = DQ(k, i, j) + (ps3 * d) * VEC(SHCFD_F_PRE1, k, i, j);
                    for (int k = 0; k < dq; k++) DQ(k, i, j) = ps1 * DQ(k, i, j);
                }
        }
        /* backward sweep, sm descending (ghost Dq = 0 above/right) */
        double *cres = c->scratch;
        for (int sm = ni + nj; sm >= 2; sm--) {
            int ilo = (1 > sm - nj) ? 1 : sm - nj, ihi = (sm - 1 < ni) ? sm - 1 : ni;
#pragma omp parallel for
            for (int i = ilo; i <= ihi; i++) {
                int j = sm - i;
                double ya[32], yb[32];
                matvec(dq, &VEC(SHCFD_F_AM, 0, i + 1, j), &DQ(0, i + 1, j), ya);
                matvec(dq, &VEC(SHCFD_F_BM, 0, i, j + 1), &DQ(0, i, j + 1), yb);
                double dci = SC(SHCFD_F_DSCI, i + 1, j), dcj = SC(SHCFD_F_DSCJ, i, j + 1);
                double al = SC(SHCFD_F_ALPHA, i, j);
                for (int k = 0; k < dq; k++) {
                    double sm_k = (dci * ya[k]) + (dcj * yb[k]);
                    DQ(k, i, j) = DQ(k, i, j) - al * sm_k;
                }
                if (dual) {
                    double d = dotp(dq, &VEC(SHCFD_F_PRE2, 0, i, j), &DQ(0, i, j));
                    double ps4 = VEC(SHCFD_F_PRESCA, 3, i, j), ips1 = 1.0 / VEC(SHCFD_F_PRESCA, 0, i, j);
                    for (int k = 0; k < dq; k++) DQ(k, i, j) = DQ(k, i, j) + (ps4 * d) * VEC(SHCFD_F_PRE1, k, i, j);
                    for (int k = 0; k < dq; k++) DQ(k, i, j) = ips1 * DQ(k, i, j);
                }
                if (inner) {
                    double e = DQ(ny + 2, i, j);
                    cres[(size_t)(i - 1) + (size_t)ni * (size_t)(j - 1)] = e * e;
                }
            }
        }
        if (inner) {
            /* res and omega in the reference's single-thread visiting order:
             * time_dual_viscous.f90:68-93, time_NR_viscous.f90:55-69 */
            for (int sm = ni + nj; sm >= 2; sm--) {
                int ilo = (1 > sm - nj) ? 1 : sm - nj, ihi = (sm - 1 < ni) ? sm - 1 : ni;
                for (int i = ilo; i <= ihi; i++) {
                    int j = sm - i;
                    res = res + cres[(size_t)(i - 1) + (size_t)ni * (size_t)(j - 1)];
                    for (int k = 0; k < ny; k++) {
                        if (DQ(k, i, j) < 0.0) {
                            double om = (-Q(k, i, j) / DQ(k, i, j)) * c->p.dqmax;
                            omega = (omega < om) ? omega : om;
                            if (omega < c->p.omega_min) err = SHCFD_E_OMEGA;
                        }
                    }
                }
            }
        } else {
            /* add Dq: time_lusgs_viscous.f90:57-63 */
            for (int j = 1; j <= nj; j++)
                for (int i = 1; i <= ni; i++)
                    for (int k = 0; k < dq; k++) Q(k, i, j) = Q(k, i, j) + DQ(k, i, j);
        }
    }
    if (inner) {
        /* time_dual_viscous.f90:98-124, time_NR_viscous.f90:74-101 (no NaN stop) */
        if (dual && res != res) err = err ? err : SHCFD_E_NAN;
        if (step_internal == 1) c->res1 = res;
        if (omega_out) *omega_out = omega;
        double om = (c->p.omega_max < omega) ? c->p.omega_max : omega;
        for (int b = 0; b < c->nb; b++) {
            int ni = c->ni[b], nj = c->nj[b];
            for (int j = 1; j <= nj; j++)
                for (int i = 1; i <= ni; i++)
                    for (int k = 0; k < dq; k++) Q(k, i, j) = Q(k, i, j) + om * DQ(k, i, j);
        }
    }
    if (res_out) *res_out = res;
#undef DQ
    return err;
}

void or_calc_next_step(or_ctx *c) {
    /* time_scheme/euler/time_euler_viscous.raw.f90:1-23 */
    int dq = c->dimq;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double s[32];
                flux_balance(c, b, i, j, s);
                double dtv = (c->p.time_step == 0) ? c->dt_grbl : SC(SHCFD_F_DT, i, j);
                double f = dtv / SC(SHCFD_F_VOL, i, j);
                for (int k = 0; k < dq; k++) Q(k, i, j) = QP(k, i, j) + f * s[k];
            }
    }
}

void or_calc_next_step_rk2(or_ctx *c, int stage) {
    /* time_scheme/RK2/time_RK2_viscous.raw.f90:1-59 */
    int dq = c->dimq;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double s[32];
                flux_balance(c, b, i, j, s);
                double dtv = (c->p.time_step == 0) ? c->dt_grbl : SC(SHCFD_F_DT, i, j);
                double f = stage == 1 ? (0.5 * dtv) / SC(SHCFD_F_VOL, i, j) : dtv / SC(SHCFD_F_VOL, i, j);
                for (int k = 0; k < dq; k++) Q(k, i, j) = QP(k, i, j) + f * s[k];
                if (stage == 2)
                    for (int k = 0; k < c->ny; k++) Q(k, i, j) = (Q(k, i, j) > 1e-40) ? Q(k, i, j) : 1e-40;
            }
    }
}

double or_energy_rms(or_ctx *c) {
    /* core/check_convergence.f90:17-28 */
    int ny = c->ny;
    double tmp = 0.0, sm = 0.0;
    for (int b = 0; b < c->nb; b++) {
        int ni = c->ni[b], nj = c->nj[b];
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double d = Q(ny + 2, i, j) - QP(ny + 2, i, j);
                tmp = tmp + d * d;
            }
        sm = sm + (double)(ni * nj);
    }
    return sqrt(tmp) / sm;
}

/* ======================================================================== */
/* main loops                                                               */
/* ======================================================================== */
static void part_primitive(or_ctx *c) { or_set_w(c); or_set_thermo_prop(c); }
static void part_main(or_ctx *c) { or_set_bc(c); or_set_hv(c); or_set_tg(c); or_set_tgv(c); }

int or_prepare(or_ctx *c) {
    /* core/main.body.raw.f90:14-15 */
    part_primitive(c);
    or_set_bc(c);
    return 0;
}

int or_step(or_ctx *c, double *res_first, double *res_last) {
    int err = 0;
    c->chem_err = 0;
    switch (c->p.scheme) {
    case SHCFD_SCHEME_LUSGS: /* LU-SGS/2d/main.main.raw.f90:1-6 */
        if (c->p.reaction) { /* LU-SGS/2d/main.main.point_implicit.f90:1-9 */
            part_primitive(c);
            or_set_dt(c);
            err = or_proceed_reaction(c);
            or_copy_field(c, SHCFD_F_QP, SHCFD_F_Q);
            part_primitive(c);
            part_main(c);
            or_set_abpm(c);
            int e = or_calc_next_step_implicit(c, 0, NULL, NULL);
            if (!err) err = e;
            break;
        }
        or_copy_field(c, SHCFD_F_QP, SHCFD_F_Q);
        part_primitive(c);
        or_set_dt(c);
        part_main(c);
        or_set_abpm(c);
        err = or_calc_next_step_implicit(c, 0, NULL, NULL);
        break;
    case SHCFD_SCHEME_PRECON_LUSGS: /* preconLU-SGS/2d/main.main.raw.f90:1-5 */
        or_copy_field(c, SHCFD_F_QP, SHCFD_F_Q);
        part_primitive(c);
        part_main(c);
        or_set_abpm(c);
        err = or_calc_next_step_implicit(c, 0, NULL, NULL);
        break;
    case SHCFD_SCHEME_EULER: /* euler/main.main.raw.f90 */
        or_copy_field(c, SHCFD_F_QP, SHCFD_F_Q);
        part_primitive(c);
        or_set_dt(c);
        part_main(c);
        or_calc_next_step(c);
        break;
    case SHCFD_SCHEME_RK2: /* RK2/main.main.raw.f90 */
        if (c->p.reaction) { /* RK2/main.main.point_implicit.f90:1-11 */
            part_primitive(c);
            or_set_dt(c);
            err = or_proceed_reaction(c);
            or_copy_field(c, SHCFD_F_QP, SHCFD_F_Q);
            part_primitive(c);
            part_main(c);
            or_calc_next_step_rk2(c, 1);
            part_primitive(c);
            part_main(c);
            or_calc_next_step_rk2(c, 2);
            break;
        }
        or_copy_field(c, SHCFD_F_QP, SHCFD_F_Q);
        part_primitive(c);
        or_set_dt(c);
        part_main(c);
        or_calc_next_step_rk2(c, 1);
        part_primitive(c);
        part_main(c);
        or_calc_next_step_rk2(c, 2);
        break;
    case SHCFD_SCHEME_NR:     /* NR/2d/main.main.raw.f90:1-20 (the same loop as dual) */
    case SHCFD_SCHEME_PRECON: /* precon/2d/main.main.raw.f90:1-19 (no set_dt: set_ABpm sets dt_mat) */
    case SHCFD_SCHEME_DUAL: { /* dual/2d/main.main.raw.f90:1-20 */
        part_primitive(c);
        if (c->p.scheme != SHCFD_SCHEME_PRECON) or_set_dt(c);
        or_copy_field(c, SHCFD_F_QPP, SHCFD_F_QP);
        or_copy_field(c, SHCFD_F_QP, SHCFD_F_Q);
        double res = 0.0;
        for (int j = 1; j < c->p.num_internal_loop; j++) {
            part_primitive(c);
            part_main(c);
            or_set_abpm(c);
            or_set_dqdt(c);
            double om;
            int e = or_calc_next_step_implicit(c, j, &res, &om);
            if (e && !err) err = e;
            if (j == 1 && res_first) *res_first = res;
        }
        if (res_last) *res_last = res;
        /* check_internal_loop_convergence: sch_dual.raw.f90:284-299 (the
         * caller supplies the step / step_res context; only the flag here) */
        c->first_dual = 0;
        break;
    }
    default:
        return SHCFD_E_STATE;
    }
    if (!err && c->chem_err) err = c->chem_err;
    return err;
}
