// shcfd_kernels.h -- per-cell / per-face device functions shared by the plain
// per-routine kernels and the fused kernels, plus kernel declarations.
#pragma once
#include "shcfd_device.h"
#include "shcfd_ad.h"

// Species counts 1..SHCFD_MAX_NY have compiled kernels (1 = ideal gas), the
// same counts for the flow, the Chemkin chemistry and the 0-D harness.  A
// source whose kernels take the species count as a template parameter is
// compiled once per instantiation group (SHCFD_INST_GROUP = 0 ..
// SHCFD_INST_GROUPS-1, given by the build); each object instantiates its
// group's counts, and group 0 also holds the source's kernels that have no
// species parameter.
#define SHCFD_MAX_NY 16
#define SHCFD_INST_GROUPS 5
#ifndef SHCFD_INST_GROUP
#define SHCFD_INST_GROUP 0
#endif
#if SHCFD_INST_GROUP == 0
#define SHCFD_FOR_EACH_NY(X) X(1) X(2)
#elif SHCFD_INST_GROUP == 1
#define SHCFD_FOR_EACH_NY(X) X(3) X(4) X(5) X(6)
#elif SHCFD_INST_GROUP == 2
#define SHCFD_FOR_EACH_NY(X) X(7) X(8) X(9) X(10)
#elif SHCFD_INST_GROUP == 3
#define SHCFD_FOR_EACH_NY(X) X(11) X(12) X(13)
#else
#define SHCFD_FOR_EACH_NY(X) X(14) X(15) X(16)
#endif
#define SHCFD_FOR_EACH_NS(X) SHCFD_FOR_EACH_NY(X)
#define SHCFD_FOR_EACH_FLOW_NS(X) SHCFD_FOR_EACH_NY(X)
// the point-implicit flow kernel keeps each lane's (NS+1)^2 Newton matrix in
// LDS up to this species count (64 lanes x 121 doubles = 62 KB); above it the
// register kernel runs
#define SHCFD_LDS_NEWTON_MAX_NS 10

namespace shcfd {

// MUSCL slope pair of one cell: high_order/muscl/muscl.f90:28-35
// (limiter 1 = minmod, an addition).  Generic in the scalar type (see
// shcfd_ad.h); the double instantiation is the primal.
template <typename T>
__device__ inline void muscl_delta_t(const T& Du, const T& Dl, double k_mus, int limiter, T& du, T& dl) {
    if (limiter == SHCFD_LIMITER_MINMOD) {
        if (val(Du) * val(Dl) <= 0.0) {
            du = T(0.0);
            dl = T(0.0);
            return;
        }
        const T m = (tfabs(Du) < tfabs(Dl)) ? Du : Dl;
        du = m;
        dl = m;
        return;
    }
    T s = ((2.0 * Du) * Dl) / (((Du * Du) + (Dl * Dl)) + 1e-300);
    du = (0.5 * s) * (((1.0 + k_mus * s) * Du) + ((1.0 - k_mus * s) * Dl));
    dl = (0.5 * s) * (((1.0 + k_mus * s) * Dl) + ((1.0 - k_mus * s) * Du));
}
__device__ inline void muscl_delta(double Du, double Dl, double k_mus, int limiter, double& du, double& dl) {
    muscl_delta_t<double>(Du, Dl, k_mus, limiter, du, dl);
}

// SLAU flux of one face: core/scheme.f90:71-153.  wl/wr are full w
// vectors (rho,u,v,p,Y(NY),gamma,H,...), n the unit face normal.  Generic in
// the scalar type; slau_face below is the primal (double) form.
template <int NY, typename T>
__device__ inline void slau_face_t(const T* wl, const T* wr, const T& n1, const T& n2, T* tg) {
    constexpr int dq = NY + 3;
    T rhol = wl[0], ul = wl[1], vl = wl[2], pl = wl[3], gl = wl[dq + 1];
    T rhor = wr[0], ur = wr[1], vr = wr[2], pr = wr[3], gr = wr[dq + 1];
    T al = sqrt((gl * pl) / rhol);
    T unl = (n1 * ul) + (n2 * vl);
    T ar = sqrt((gr * pr) / rhor);
    T unr = (n1 * ur) + (n2 * vr);
    T dp = pr - pl;
    T a_bar = 0.5 * (ar + al);
    T p_bar = 0.5 * (pr + pl);
    T ml = unl / a_bar;
    T mr = unr / a_bar;
    T mt = (1.0 / a_bar) * sqrt(0.5 * ((((ul * ul) + (vl * vl)) + (ur * ur)) + (vr * vr)));
    T m_tilde = (1.0 < val(mt)) ? T(1.0) : mt;
    T chi = (1.0 - m_tilde) * (1.0 - m_tilde);
    T tt = (val(ml) < 0.0) ? ml : T(0.0);
    T gp = -((val(tt) > -1.0) ? tt : T(-1.0));
    tt = (val(mr) > 0.0) ? mr : T(0.0);
    T gm = (val(tt) < 1.0) ? tt : T(1.0);
    T g = gp * gm;
    T un_bar_abs = ((tfabs(unr) * rhor) + (tfabs(unl) * rhol)) / (rhor + rhol);
    T unp = ((1.0 - g) * un_bar_abs) + (g * tfabs(unl));
    T unm = ((1.0 - g) * un_bar_abs) + (g * tfabs(unr));
    T dm = 0.5 * (((rhol * (unl + unp)) + (rhor * (unr - unm))) - ((chi / a_bar) * dp));
    T betar, betal;
    if (::fabs(val(mr)) < 1.0) betar = (0.25 * (2.0 + mr)) * ((mr - 1.0) * (mr - 1.0));
    else betar = T(0.5 * (1.0 - sgn1(mr)));
    if (::fabs(val(ml)) < 1.0) betal = (0.25 * (2.0 - ml)) * ((ml + 1.0) * (ml + 1.0));
    else betal = T(0.5 * (1.0 + sgn1(ml)));
    T p_tilde = (p_bar + ((0.5 * (betal - betar)) * (pl - pr))) + (((1.0 - chi) * ((betar + betal) - 1.0)) * p_bar);
    T cp = 0.5 * (dm + tfabs(dm)), cm = 0.5 * (dm - tfabs(dm));
#pragma unroll
    for (int k = 0; k < NY; k++) tg[k] = ((cp * wl[4 + k]) + (cm * wr[4 + k])) + (p_tilde * 0.0);
    tg[NY] = ((cp * ul) + (cm * ur)) + (p_tilde * n1);
    tg[NY + 1] = ((cp * vl) + (cm * vr)) + (p_tilde * n2);
    tg[NY + 2] = ((cp * wl[dq + 2]) + (cm * wr[dq + 2])) + (p_tilde * 0.0);
}
template <int NY>
__device__ inline void slau_face(const double* wl, const double* wr, double n1, double n2, double* tg) {
    slau_face_t<NY, double>(wl, wr, n1, n2, tg);
}

// viscous flux of one face: viscosity/viscous/no-nV/2d/sch_viscous.f90:106-178
// (DIR 0: i-face (i,j) between i and i+1; DIR 1: j-face between j and j+1).
// Axisymmetric: the dilatation gains v/r with the face-averaged v over the
// centre radius of cell (i,j), the lower-index cell of the face, as
// viscous/no-nV/q2d/sch_viscous.f90:173-174 and :262-263 write it.
// Generic in the scalar type, in two parts: viscous_in_t forms the face
// averages and the xi/eta differences of (u, v, T, Y) from the w vectors of
// the stencil cells w[0..5] = c, c1 (the face's two cells), a, b, e, f --
//   DIR 0: a = (i+1,j+1), b = (i,j+1), e = (i+1,j-1), f = (i,j-1);
//   DIR 1: a = (i+1,j+1), b = (i+1,j), e = (i-1,j+1), f = (i-1,j) --
// (xi differences c1 - c and eta differences (a + b - e - f)/4 for DIR 0,
// the other way round for DIR 1); viscous_out_t forms the flux from them,
// the face-averaged species enthalpies lhi, the face metrics g (geojac), the
// unit normal (n1, n2) and, axisymmetric, the centre radius rc of c.
template <int NY, typename T>
struct ViscIn {
    T rho, u, v, gamm, Rg, mu;  // face averages
    T dxi[3 + NY], deta[3 + NY];  // differences of (u, v, T, Y_k) along xi and eta
};
template <int NY, int DIR, typename T>
__device__ inline void viscous_in_t(const T* const* w, ViscIn<NY, T>& z) {
    constexpr int dq = NY + 3, ig = dq + 1, ir = dq + 3, imu = dq + 4;
    const T *wc = w[0], *w1 = w[1], *wa = w[2], *wb = w[3], *we = w[4], *wf = w[5];
    auto Td = [&](const T* x) { return (x[IW_P] / x[IW_RHO]) / x[ir]; };
    z.rho = (wc[IW_RHO] + w1[IW_RHO]) * 0.5;
    z.u = (wc[IW_U] + w1[IW_U]) * 0.5;
    z.v = (wc[IW_V] + w1[IW_V]) * 0.5;
    z.gamm = (wc[ig] + w1[ig]) * 0.5;
    z.Rg = (wc[ir] + w1[ir]) * 0.5;
    z.mu = (wc[imu] + w1[imu]) * 0.5;
    T* dn = DIR == 0 ? z.dxi : z.deta;   // c1 - c
    T* dc = DIR == 0 ? z.deta : z.dxi;   // (a + b - e - f) / 4
    dn[0] = w1[IW_U] - wc[IW_U];
    dn[1] = w1[IW_V] - wc[IW_V];
    dn[2] = Td(w1) - Td(wc);
#pragma unroll
    for (int k = 0; k < NY; k++) dn[3 + k] = w1[IW_Y(k)] - wc[IW_Y(k)];
    dc[0] = (((wa[IW_U] + wb[IW_U]) - we[IW_U]) - wf[IW_U]) * 0.25;
    dc[1] = (((wa[IW_V] + wb[IW_V]) - we[IW_V]) - wf[IW_V]) * 0.25;
    dc[2] = (((Td(wa) + Td(wb)) - Td(we)) - Td(wf)) * 0.25;
#pragma unroll
    for (int k = 0; k < NY; k++) dc[3 + k] = (((wa[IW_Y(k)] + wb[IW_Y(k)]) - we[IW_Y(k)]) - wf[IW_Y(k)]) * 0.25;
}
template <int NY, typename T>
__device__ inline void viscous_out_t(const ViscIn<NY, T>& z, const T* lhi, const T* g, const T& n1, const T& n2,
                                     bool axi, const T& rc, T* out) {
    const double Sc = 1.0, Pr = 1.0;
    const T rho = z.rho, u = z.u, v = z.v, gamm = z.gamm, Rg = z.Rg, mu = z.mu;
    T D = (mu / rho) / Sc;
    T k_heat = (((mu * gamm) / (gamm - 1.0)) * Rg) / Pr;
    const T g11 = g[0], g21 = g[1], g12 = g[2], g22 = g[3];
    T dudx = (z.dxi[0] * g11) + (z.deta[0] * g21);
    T dvdx = (z.dxi[1] * g11) + (z.deta[1] * g21);
    T dTdx = (z.dxi[2] * g11) + (z.deta[2] * g21);
    T dudr = (z.dxi[0] * g12) + (z.deta[0] * g22);
    T dvdr = (z.dxi[1] * g12) + (z.deta[1] * g22);
    T dTdr = (z.dxi[2] * g12) + (z.deta[2] * g22);
    T dwdx[NY], dwdr[NY];
    T hidwdx = T(0.0), hidwdr = T(0.0);
#pragma unroll
    for (int k = 0; k < NY; k++) {
        dwdx[k] = (z.dxi[3 + k] * g11) + (z.deta[3 + k] * g21);
        hidwdx = hidwdx + (dwdx[k] * lhi[k]);
        dwdr[k] = (z.dxi[3 + k] * g12) + (z.deta[3 + k] * g22);
        hidwdr = hidwdr + (dwdr[k] * lhi[k]);
    }
    T divu = dudx + dvdr;
    if (axi) divu = divu + (v / rc);
    T t_xx = -(((2.0 / 3.0) * mu) * divu) + ((2.0 * mu) * dudx);
    T t_rr = -(((2.0 / 3.0) * mu) * divu) + ((2.0 * mu) * dvdr);
    T t_xr = mu * (dvdx + dudr);
    T qx = -(k_heat * dTdx) - ((rho * D) * hidwdx);
    T qr = -(k_heat * dTdr) - ((rho * D) * hidwdr);
#pragma unroll
    for (int k = 0; k < NY; k++) out[k] = (((rho * D) * dwdx[k]) * n1) + (((rho * D) * dwdr[k]) * n2);
    out[NY] = (t_xx * n1) + (t_xr * n2);
    out[NY + 1] = (t_xr * n1) + (t_rr * n2);
    out[NY + 2] = ((((u * t_xx) + (v * t_xr)) - qx) * n1) + ((((u * t_xr) + (v * t_rr)) - qr) * n2);
}

// face metrics (geojaci / geojacj) from the centre coordinates of the same
// stencil (c, c1, a, b, e, f): viscosity/viscous/no-nV/2d/sch_viscous.f90:1-54
template <int DIR, typename T>
__device__ inline void geojac_t(const T* xc, const T* rc, T* g) {
    T dxdxi, dydxi, dxdeta, dydeta;
    if (DIR == 0) {
        dxdxi = xc[1] - xc[0];
        dydxi = rc[1] - rc[0];
        dxdeta = (((xc[2] + xc[3]) - xc[4]) - xc[5]) * 0.25;
        dydeta = (((rc[2] + rc[3]) - rc[4]) - rc[5]) * 0.25;
    } else {
        dxdxi = (((xc[2] + xc[3]) - xc[4]) - xc[5]) * 0.25;
        dydxi = (((rc[2] + rc[3]) - rc[4]) - rc[5]) * 0.25;
        dxdeta = xc[1] - xc[0];
        dydeta = rc[1] - rc[0];
    }
    T det = (dxdxi * dydeta) - (dydxi * dxdeta);
    g[0] = dydeta / det;
    g[2] = -dxdeta / det;
    g[1] = -dydxi / det;
    g[3] = dxdxi / det;
}

// stencil cells of a viscous face in viscous_in_t's order (c, c1, a, b, e, f)
template <int DIR>
__device__ inline void viscous_stencil(long long c, long long P, long long* s) {
    s[0] = c;
    s[1] = DIR == 0 ? c + 1 : c + P;
    s[2] = c + 1 + P;
    s[3] = DIR == 0 ? c + P : c + 1;
    s[4] = DIR == 0 ? c + 1 - P : c - 1 + P;
    s[5] = DIR == 0 ? c - P : c - 1;
}

// the primal (double) instance, reading the fields
template <int NY, int DIR>
__device__ inline void viscous_face(const DevCtx& d, const DevBlock& B, int i, int j, double* out) {
    constexpr int dw = NY + 8;
    const long long c = cidx(B, i, j);
    long long st[6];
    viscous_stencil<DIR>(c, B.pitch, st);
    double ws[6][dw];
    const double* wp[6];
#pragma unroll
    for (int s = 0; s < 6; s++) {
#pragma unroll
        for (int k = 0; k < dw; k++) ws[s][k] = FLD(d, SHCFD_F_W, k, st[s]);
        wp[s] = ws[s];
    }
    double lhi[NY];
#pragma unroll
    for (int k = 0; k < NY; k++) lhi[k] = (FLD(d, SHCFD_F_VHI, k, c) + FLD(d, SHCFD_F_VHI, k, st[1])) * 0.5;
    const int fg = DIR == 0 ? SHCFD_F_GEOJACI : SHCFD_F_GEOJACJ;
    const double g[4] = {FLD(d, fg, 0, c), FLD(d, fg, 1, c), FLD(d, fg, 2, c), FLD(d, fg, 3, c)};
    const int fvn = DIR == 0 ? SHCFD_F_VNI : SHCFD_F_VNJ;
    const bool axi = d.p.dimension == SHCFD_DIM_AXISYMMETRIC;
    ViscIn<NY, double> z;
    viscous_in_t<NY, DIR, double>(wp, z);
    viscous_out_t<NY, double>(z, lhi, g, FLD(d, fvn, 0, c), FLD(d, fvn, 1, c), axi,
                              axi ? FLD(d, SHCFD_F_R, 0, c) : 1.0, out);
}

// radial-momentum source Sq + Svq of an axisymmetric cell: the pressure
// (dim/q2d/Sq.f90:1-18) minus the hoop stress t_tt of viscous flow
// (viscosity/viscous/no-nV/q2d/sch_viscous.f90:290-319; 0 when inviscid)
template <int NY>
__device__ inline double radial_source(const DevCtx& d, const DevBlock& B, long long c) {
    constexpr int imu = NY + 3 + 4;
    const long long P = B.pitch;
    auto Wf = [&](int k, long long cc) { return FLD(d, SHCFD_F_W, k, cc); };
    double svq = 0.0;
    if (d.p.viscous) {
        const double v = Wf(IW_V, c), mu = Wf(imu, c), r = FLD(d, SHCFD_F_R, 0, c);
        double dudxi = (Wf(IW_U, c + 1) - Wf(IW_U, c - 1)) * 0.5;
        double dvdxi = (Wf(IW_V, c + 1) - Wf(IW_V, c - 1)) * 0.5;
        double dudeta = (Wf(IW_U, c + P) - Wf(IW_U, c - P)) * 0.5;
        double dvdeta = (Wf(IW_V, c + P) - Wf(IW_V, c - P)) * 0.5;
        double dudx = (dudxi * FLD(d, SHCFD_F_GEOJACC, 0, c)) + (dudeta * FLD(d, SHCFD_F_GEOJACC, 1, c));
        double dvdr = (dvdxi * FLD(d, SHCFD_F_GEOJACC, 2, c)) + (dvdeta * FLD(d, SHCFD_F_GEOJACC, 3, c));
        double divu = (dudx + dvdr) + (v / r);
        double t_tt = -(((2.0 / 3.0) * mu) * divu) + (((2.0 * mu) * v) / r);
        svq = -t_tt;
    }
    return Wf(IW_P, c) + svq;
}

// flux balance of one cell: LU-SGS/2d/time_lusgs_viscous.f90:12-22 (Area = 0
// in 2-D), LU-SGS/q2d/time_lusgs_viscous.f90:15-20 (Area*(Sq+Svq), nonzero on
// the radial momentum only)
template <int NY>
__device__ inline void flux_balance(const DevCtx& d, const DevBlock& B, long long c, double* s) {
    const long long P = B.pitch;
    double di0 = FLD(d, SHCFD_F_DSI, 0, c - 1), di1 = FLD(d, SHCFD_F_DSI, 0, c);
    double dj0 = FLD(d, SHCFD_F_DSJ, 0, c - P), dj1 = FLD(d, SHCFD_F_DSJ, 0, c);
    const bool ax = d.p.dimension == SHCFD_DIM_AXISYMMETRIC;
    const double area = ax ? FLD(d, SHCFD_F_AREA, 0, c) : 0.0;
    const double src = ax ? radial_source<NY>(d, B, c) : 0.0;
    // TGv known to be zero (inviscid): x - (+0) is x, so skipping the loads
    // leaves every value unchanged
    const bool tv = !d.tgv_zero;
#pragma unroll
    for (int k = 0; k < NY + 3; k++) {
        const double vi0 = tv ? FLD(d, SHCFD_F_TGVI, k, c - 1) : 0.0, vi1 = tv ? FLD(d, SHCFD_F_TGVI, k, c) : 0.0;
        const double vj0 = tv ? FLD(d, SHCFD_F_TGVJ, k, c - P) : 0.0, vj1 = tv ? FLD(d, SHCFD_F_TGVJ, k, c) : 0.0;
        double v = ((((di0 * (FLD(d, SHCFD_F_TGI, k, c - 1) - vi0)) - (di1 * (FLD(d, SHCFD_F_TGI, k, c) - vi1))) +
                     (dj0 * (FLD(d, SHCFD_F_TGJ, k, c - P) - vj0))) -
                    (dj1 * (FLD(d, SHCFD_F_TGJ, k, c) - vj1)));
        s[k] = ax ? v + (area * (k == NY + 1 ? src : 0.0)) : v + 0.0;
    }
}

// Cell-centre split flux Jacobians and LU-SGS scalars of one cell:
// LU-SGS/2d/sch_lusgs.raw.f90:26-136, NR/2d/sch_NR.raw.f90:26-137,
// precon/2d/sch_precon.raw.f90:27-197, preconLU-SGS/2d/sch_precon.raw.f90:28-197,
// dual/2d/sch_dual.raw.f90:28-204.
template <int NY>
struct CellJac {
    static constexpr int dq = NY + 3;
    double Ap[dq * dq], Am[dq * dq], Bp[dq * dq], Bm[dq * dq];
    double alpha, phi, dtm;
    double nuA, nuB;  // spectral radii of the split: A+ - A- = nuA I, B+ - B- = nuB I
    double pre1[dq], pre2[dq], ps[4];
    bool precond;
};

template <int NY>
__device__ inline void cell_jacobian(const DevCtx& d, long long c, long long P, CellJac<NY>& J) {
    constexpr int dq = NY + 3, ig = dq + 1, ih = dq + 2, imu = dq + 4;
    const int scheme = d.p.scheme;
    double wc[NY + 8], dhi[NY];
#pragma unroll
    for (int k = 0; k < NY + 8; k++) wc[k] = FLD(d, SHCFD_F_W, k, c);
#pragma unroll
    for (int k = 0; k < NY; k++) dhi[k] = FLD(d, SHCFD_F_DHI, k, c);
    double rho = wc[0], u = wc[1], v = wc[2], p = wc[3], gamm = wc[ig], H = wc[ih];
    double cs = sqrt((gamm * p) / rho);
    double gmmad = gamm - 1.0;
    double Vs = (u * u) + (v * v);
    double n1a = FLD(d, SHCFD_F_VNCI, 0, c), n2a = FLD(d, SHCFD_F_VNCI, 1, c);
    double n1b = FLD(d, SHCFD_F_VNCJ, 0, c), n2b = FLD(d, SHCFD_F_VNCJ, 1, c);
    double unA = (n1a * u) + (n2a * v);
    double unB = (n1b * u) + (n2b * v);
    double dsci = FLD(d, SHCFD_F_DSCI, 0, c), dscj = FLD(d, SHCFD_F_DSCJ, 0, c);
    double vol = FLD(d, SHCFD_F_VOL, 0, c);
    double dtv = (d.p.time_step == 0) ? d.scal[SCAL_DT_GRBL] : FLD(d, SHCFD_F_DT, 0, c);
    double nuA, nuB;
    J.precond = precond_scheme(scheme);
    double pscale = 1.0;
    J.phi = 0.0;
    J.dtm = 0.0;
    if (!J.precond) {
        nuA = fabs(unA) + cs;
        nuB = fabs(unB) + cs;
    } else {
        double dmn = (dsci < dscj) ? dsci : dscj;
        double Vv = (wc[imu] / rho) / dmn;
        double sv = sqrt(Vs);
        double Ur = (Vv > sv) ? Vv : sv;
        Ur = (Ur < cs) ? Ur : cs;
        double phi = (1.0 / (Ur * Ur)) - (1.0 / (cs * cs));
        J.phi = phi;
        double rr = Ur / cs;
        double alp = 0.5 * (1.0 - rr * rr);
        double ud = unA * (1.0 - alp);
        double cd = sqrt(((alp * alp) * (unA * unA)) + (Ur * Ur));
        nuA = fabs(ud) + cd;
        ud = unB * (1.0 - alp);
        cd = sqrt(((alp * alp) * (unB * unB)) + (Ur * Ur));
        nuB = fabs(ud) + cd;
        double uds = (nuA * (FLD(d, SHCFD_F_DSI, 0, c) + FLD(d, SHCFD_F_DSI, 0, c - 1))) +
                     (nuB * (FLD(d, SHCFD_F_DSJ, 0, c) + FLD(d, SHCFD_F_DSJ, 0, c - P)));
        double nr = wc[imu] / rho;
        double mx = (uds > nr) ? uds : nr;
        if (scheme != SHCFD_SCHEME_DUAL) {
            double dtm = (d.p.cfl * vol) / mx;
            J.dtm = dtm;
            dtv = dtm;
            double phh = -1.0 / ((cs * cs) + (1.0 / phi));
#pragma unroll
            for (int k = 0; k < NY; k++) J.pre1[k] = phh * wc[4 + k];
            J.pre1[NY] = phh * u;
            J.pre1[NY + 1] = phh * v;
            J.pre1[NY + 2] = phh * H;
        } else {
            double dtau = (d.p.cfl_tau * vol) / mx;
            double alpha = 1.0 / (((vol / dtau) + (dsci * nuA)) + (dscj * nuB));
            J.alpha = alpha;
            double beta = ((3.0 * vol) / (2.0 * dtv)) * alpha;
            double ip = 1.0 / phi;
            J.ps[0] = 1.0 + beta;
            J.ps[1] = -1.0 / (ip + (cs * cs));
            J.ps[2] = (-1.0 / (ip + (cs * cs))) / ((1.0 / beta) + 1.0);
            J.ps[3] = beta / (((1.0 + beta) / phi) + (cs * cs));
            pscale = J.ps[1];
#pragma unroll
            for (int k = 0; k < NY; k++) J.pre1[k] = wc[4 + k];
            J.pre1[NY] = u;
            J.pre1[NY + 1] = v;
            J.pre1[NY + 2] = H;
        }
#pragma unroll
        for (int k = 0; k < NY; k++) J.pre2[k] = (0.5 * gmmad) * Vs + dhi[k];
        J.pre2[NY] = -(gmmad * u);
        J.pre2[NY + 1] = -(gmmad * v);
        J.pre2[NY + 2] = gmmad;
    }
    // Euler Jacobians A, B (column-major)
    double A[dq * dq], Bm_[dq * dq], tA[dq], tB[dq];
    const int a1 = NY, a2 = NY + 1, a3 = NY + 2;
#define AM_(k, l) A[(k) + dq * (l)]
#define BM_(k, l) Bm_[(k) + dq * (l)]
    AM_(a1, a1) = u * (3.0 - gamm);
    AM_(a2, a1) = v;
    AM_(a3, a1) = H - (u * u) * gmmad;
    AM_(a1, a2) = -(v * gmmad);
    AM_(a2, a2) = u;
    AM_(a3, a2) = -((u * v) * gmmad);
    AM_(a1, a3) = gmmad;
    AM_(a2, a3) = 0.0;
    AM_(a3, a3) = u * gamm;
#pragma unroll
    for (int k = 0; k < NY; k++) {
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
#pragma unroll
    for (int k = 0; k < NY; k++) {
        tB[k] = -(wc[4 + k] * v);
        BM_(k, a1) = 0.0;
        BM_(k, a2) = wc[4 + k];
        BM_(k, a3) = 0.0;
    }
    tB[a1] = -(v * u);
    tB[a2] = -(v * v) + (0.5 * gmmad) * Vs;
    tB[a3] = v * (-H + (0.5 * gmmad) * Vs);
#pragma unroll
    for (int k = 0; k < NY; k++) {
#pragma unroll
        for (int l = 0; l < NY; l++) {
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
    if (J.precond) {
#pragma unroll
        for (int pass = 0; pass < 2; pass++) {
            double* M = pass == 0 ? A : Bm_;
#pragma unroll
            for (int k = 0; k < dq; k++) {
                double tmp = 0.0;
#pragma unroll
                for (int l = 0; l < dq; l++) tmp = tmp + J.pre2[l] * M[l + dq * k];
                if (scheme == SHCFD_SCHEME_DUAL) tmp = pscale * tmp;
#pragma unroll
                for (int l = 0; l < dq; l++) M[l + dq * k] = M[l + dq * k] + J.pre1[l] * tmp;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < dq * dq; e++) {
        double at = (n1a * A[e]) + (n2a * Bm_[e]);
        double bt = (n1b * A[e]) + (n2b * Bm_[e]);
        J.Ap[e] = at * 0.5;
        J.Am[e] = at * 0.5;
        J.Bp[e] = bt * 0.5;
        J.Bm[e] = bt * 0.5;
    }
#pragma unroll
    for (int k = 0; k < dq; k++) {
        int e = k + dq * k;
        J.Ap[e] = J.Ap[e] + nuA * 0.5;
        J.Am[e] = J.Am[e] - nuA * 0.5;
        J.Bp[e] = J.Bp[e] + nuB * 0.5;
        J.Bm[e] = J.Bm[e] - nuB * 0.5;
    }
#undef AM_
#undef BM_
    J.nuA = nuA;
    J.nuB = nuB;
    if (scheme == SHCFD_SCHEME_NR || scheme == SHCFD_SCHEME_PRECON)  // BDF2 diagonal: sch_NR.raw.f90:134-136
        J.alpha = 1.0 / ((((3.0 * vol) / (2.0 * dtv)) + (dsci * nuA)) + (dscj * nuB));
    else if (scheme != SHCFD_SCHEME_DUAL)
        J.alpha = 1.0 / (((vol / dtv) + (dsci * nuA)) + (dscj * nuB));
}

// Cells touched by one copy of section_exchange (cond/no-nV/condition.tail.f90:19-68):
// copy 0/1 fill the first/second ghost layer of plane p1 from p2, copy 2/3
// those of p2 from p1.  Returns false for an unknown geo code.
__host__ __device__ inline bool cut_map(const int* t, int inc, int copy, int& bd, int& di, int& dj, int& bs, int& si,
                                        int& sj) {
    int geo = t[0], order = t[2];
    int p1 = t[3] - 1, c1 = t[4], s1 = t[5], pm1 = t[6];
    int p2 = t[7] - 1, c2 = t[8], s2 = t[9], pm2 = t[10];
    int a1 = s1 + inc, a2 = s2 + inc * order;
    bool to1 = copy < 2;
    int lay = copy & 1;
    int g = to1 ? c1 + (lay ? 2 * pm1 : pm1) : c2 + (lay ? 2 * pm2 : pm2);
    int src = to1 ? (lay ? c2 - pm2 : c2) : (lay ? c1 - pm1 : c1);
    bd = to1 ? p1 : p2;
    bs = to1 ? p2 : p1;
    switch (geo) {
    case 1: if (to1) { di = g; dj = a1; si = src; sj = a2; } else { di = g; dj = a2; si = src; sj = a1; } return true;
    case 2: if (to1) { di = g; dj = a1; si = a2; sj = src; } else { di = a2; dj = g; si = src; sj = a1; } return true;
    case 3: if (to1) { di = a1; dj = g; si = src; sj = a2; } else { di = g; dj = a2; si = a1; sj = src; } return true;
    case 4: if (to1) { di = a1; dj = g; si = a2; sj = src; } else { di = a2; dj = g; si = a1; sj = src; } return true;
    default: return false;
    }
}

// ---- kernel declarations ----
template <int NY> __global__ void k_set_w(DevCtx, const Tile*);
template <int NY> __global__ void k_thermo_ideal(DevCtx, const Tile*);
template <int NY> __global__ void k_bc_cut(DevCtx, const int*, int, int);
template <int NY> __global__ void k_bc_side(DevCtx, const int*, int, const double*, int);
template <int NY> __global__ void k_bc_corner(DevCtx);
template <int NY> __global__ void k_set_hv(DevCtx, const Tile*);
template <int NY> __global__ void k_set_tg(DevCtx, const Tile*);
// flags of k_primitive_ideal: apply the deferred update q += omega dq first;
// store the q it reads into QP (the copy_qp that opens a step)
constexpr int PRIM_APPLY_UPDATE = 1, PRIM_COPY_QP = 2;
template <int NY> __global__ void k_primitive_ideal(DevCtx, const Tile*, int);
template <int NY> __global__ void k_set_hv_tg(DevCtx, const Tile*);
template <int NY> __global__ void k_set_tgv(DevCtx, const Tile*);
template <int NY> __global__ void k_set_dt(DevCtx, const Tile*, double*);
template <int NY> __global__ void k_rms_part(DevCtx, const Tile*, double*);
template <int NY> __global__ void k_set_abpm(DevCtx, const Tile*);
template <int NY> __global__ void k_set_abpm_scal(DevCtx, const Tile*);
// matrices of the strip-skewed Jacobian copy, in this order
enum { SKEW_AP = 0, SKEW_BP = 1, SKEW_AM = 2, SKEW_BM = 3 };
template <int NY> __global__ void k_set_abpm_skew(DevCtx, const Strip*, const long long*, double*, long long);
template <int NY> __global__ void k_set_dqdt(DevCtx, const Tile*);
template <int NY> __global__ void k_explicit(DevCtx, const Tile*, int);
template <int NY> __global__ void k_copy_q(DevCtx, const Tile*, int, int);
__global__ void k_zero_field(double*, long long);
__global__ void k_reduce_final(DevCtx, const double*, int, int, int);

// implicit update (k_implicit.hip)
template <int NY> __global__ void k_rhs(DevCtx, const Tile*, int*, int, int, int);
template <int NY, bool SKEW>
__global__ void k_sweep_fwd_plain(DevCtx, const Strip*, int*, double*, long long, long long, const double*, long long,
                                  const long long*);
template <int NY, bool SKEW>
__global__ void k_sweep_bwd_plain(DevCtx, const Strip*, int, int*, double*, long long, double*, long long,
                                  const double*, long long, const long long*);
template <int NY> __global__ void k_update(DevCtx, const Tile*);
template <int NY> __global__ void k_add_dq(DevCtx, const Tile*);
__global__ void k_reduce_res_omega(DevCtx, const double*, int);
// quad-row sweeps for nY > 1 (k_sweep16.hip; operators from k_set_abpm16)
template <int NY> __global__ void k_set_abpm16(DevCtx, const Strip16*, double*, long long);
__global__ void k_fill16(DevCtx, const Strip16*, int, double*, long long, int);
template <int NY> __global__ void k_set_jd16(DevCtx, const Strip16*, double2*, double*);
template <int NY>
__global__ void k_sweep16j_fwd(DevCtx, const Strip16*, int, double*, long long, long long, const double2*,
                               const double*);
template <int NY>
__global__ void k_sweep16j_bwd(DevCtx, const Strip16*, int, double*, long long, double*, long long, const double2*,
                               const double*);
template <int NY>
__global__ void k_sweep16_fwd(DevCtx, const Strip16*, int, double*, long long, long long, const double*, long long);
template <int NY>
__global__ void k_sweep16_bwd(DevCtx, const Strip16*, int, double*, long long, double*, long long, const double*,
                              long long);
__global__ void k_copy_scal(DevCtx, int, int);
// chemistry (k_chem.hip)
enum { CHEM_OP_FEX = 0, CHEM_OP_FEXVAR = 1, CHEM_OP_JEX = 2, CHEM_OP_CALCT = 3, CHEM_OP_CHEMEQ2 = 4,
       CHEM_OP_POINT_IMPLICIT = 5, CHEM_OP_DVODE = 6 };
template <int NS> __global__ void k_thermo_chem(DevCtx, const Tile*, const shcfd_mech*);
template <int NS, int MODE> __global__ void k_reaction(DevCtx, const Tile*, const shcfd_mech*, double*, int);
template <int NS> __global__ void k_reaction_dvode(DevCtx, const shcfd_mech*, const double*, int);
__global__ void k_set_iflag(DevCtx, int, int);
template <int NS> __global__ void k_reaction_lds(DevCtx, const Tile*, int, const shcfd_mech*);
template <int NS, int OP> __global__ void k_chem0d(const shcfd_mech*, int, const double*, int, double*, int, int*, double);
// geometry (k_geom.hip)
__global__ void k_geom_a(DevCtx, const Tile*);
__global__ void k_geom_ghost(DevCtx, int);
__global__ void k_geojac(DevCtx, const Tile*);

// k_nasa9.hip
__global__ void k_thermo_cold(DevCtx d, const Tile* tiles, const shcfd_nasa9* __restrict__ m);

// right-hand side of the implicit update of one cell from its flux balance s:
// the dq/dt term (formed and stored here when dqdt_inline), then the scheme's
// form (time_dual_viscous.f90:37-43, time_precon_viscous.f90 of both
// preconditioned schemes, time_NR_viscous.f90, time_lusgs_viscous.f90);
// pre1(k), pre2(k) and ps2() give set_ABpm's scalars of the cell
template <int NY, class P1, class P2, class PS>
__device__ inline void rhs_of_cell(const DevCtx& d, long long c, const double* s, bool dqdt_inline, P1 pre1, P2 pre2,
                                   PS ps2, double* x) {
    constexpr int dq = NY + 3;
    const int scheme = d.p.scheme;
    if (dqdt_inline) {
        const double dtv = (d.p.time_step == 0) ? d.scal[SCAL_DT_GRBL] : FLD(d, SHCFD_F_DT, 0, c);
        const bool first = d.scal[SCAL_FIRST_DUAL] != 0.0;
#pragma unroll
        for (int k = 0; k < dq; k++) {
            const double q = FLD(d, SHCFD_F_Q, k, c), qp = FLD(d, SHCFD_F_QP, k, c);
            FLD(d, SHCFD_F_DQDT, k, c) =
                first ? (q - qp) / dtv : (((3.0 * q) - (4.0 * qp)) + FLD(d, SHCFD_F_QPP, k, c)) / (2.0 * dtv);
        }
    }
    if (scheme == SHCFD_SCHEME_DUAL) {
        double vol = FLD(d, SHCFD_F_VOL, 0, c);
        double y[dq];
#pragma unroll
        for (int k = 0; k < dq; k++) y[k] = s[k] - vol * FLD(d, SHCFD_F_DQDT, k, c);
        double dt = 0.0;
#pragma unroll
        for (int k = 0; k < dq; k++) dt = dt + pre2(k) * y[k];
        const double p2 = ps2();
#pragma unroll
        for (int k = 0; k < dq; k++) x[k] = y[k] + (p2 * dt) * pre1(k);
    } else if (scheme == SHCFD_SCHEME_PRECON_LUSGS || scheme == SHCFD_SCHEME_PRECON) {
        const bool inner = scheme == SHCFD_SCHEME_PRECON;
        const double vol = inner ? FLD(d, SHCFD_F_VOL, 0, c) : 0.0;
        double dt = 0.0;
#pragma unroll
        for (int k = 0; k < dq; k++) dt = dt + pre2(k) * s[k];
#pragma unroll
        for (int k = 0; k < dq; k++) {
            const double y = s[k] + pre1(k) * dt;
            x[k] = inner ? y - vol * FLD(d, SHCFD_F_DQDT, k, c) : y;
        }
    } else if (scheme == SHCFD_SCHEME_NR) {
        const double vol = FLD(d, SHCFD_F_VOL, 0, c);
#pragma unroll
        for (int k = 0; k < dq; k++) x[k] = s[k] - vol * FLD(d, SHCFD_F_DQDT, k, c);
    } else {
#pragma unroll
        for (int k = 0; k < dq; k++) x[k] = s[k];
    }
}

// set_ABpm's per-cell scalars: alpha, the implicit pressure source's dp/dq
// (axisymmetric), the preconditioning vectors and scalars, the local dt of the
// preconditioned schemes
template <int NY>
__device__ inline void store_abpm_scalars(const DevCtx& d, long long c, const CellJac<NY>& J) {
    constexpr int dq = NY + 3;
    FLD(d, SHCFD_F_ALPHA, 0, c) = J.alpha;
    if (d.p.dimension == SHCFD_DIM_AXISYMMETRIC) {
        // dp/dq of the implicit pressure source, kept in pre2 (unused by the
        // axisymmetric schemes): LU-SGS/q2d/sch_lusgs.raw.f90:43-49,
        // NR/q2d/sch_NR.raw.f90:43-49
        const double u = FLD(d, SHCFD_F_W, IW_U, c), v = FLD(d, SHCFD_F_W, IW_V, c);
        const double gmmad = FLD(d, SHCFD_F_W, dq + 1, c) - 1.0;
        const double Vs = (u * u) + (v * v);
        for (int k = 0; k < NY; k++) FLD(d, SHCFD_F_PRE2, k, c) = ((0.5 * gmmad) * Vs) + FLD(d, SHCFD_F_DHI, k, c);
        FLD(d, SHCFD_F_PRE2, NY, c) = -(gmmad * u);
        FLD(d, SHCFD_F_PRE2, NY + 1, c) = -(gmmad * v);
        FLD(d, SHCFD_F_PRE2, NY + 2, c) = gmmad;
    }
    if (J.precond) {
        FLD(d, SHCFD_F_PHI, 0, c) = J.phi;
        for (int k = 0; k < dq; k++) {
            FLD(d, SHCFD_F_PRE1, k, c) = J.pre1[k];
            FLD(d, SHCFD_F_PRE2, k, c) = J.pre2[k];
        }
        if (d.p.scheme == SHCFD_SCHEME_DUAL) {
            for (int k = 0; k < 4; k++) FLD(d, SHCFD_F_PRESCA, k, c) = J.ps[k];
        } else {
            FLD(d, SHCFD_F_DT, 0, c) = J.dtm;
        }
    }
}

// ---------------------------------------------------------------------------
// set_ABpm of one cell (LU-SGS, preconLU-SGS, NR, precon, dual): the stored
// scalars (alpha; phi, pre1, pre2 and pre_sca or the local dt of the
// preconditioned schemes) and, by MATS, the split Jacobians
// ---------------------------------------------------------------------------
// MATS: 0 = no stored Jacobians, 1 = Ap/Am/Bp/Bm fields (cell SoA),
// 2 = the strip-skewed copy read by the skewed sweeps (slot given),
// 3 = the quad-row copy of the 16-row sweeps (nsl = total steps NT of all
// strips, slot = 64 x global step + 4 r): lane 4r+c of wave w owns row
// k = 4w + c; per sweep (set 0: forward, 1: backward) it reads dq + 2 pairs
// of doubles, pair p of global step g at ((((set NT + g) C + w) dq2 + p) 64 +
// lane) 2, dq2 = dq + 2, so the operators of one step of a strip are
// contiguous: elements 0..dq-1 are row k of A+ (A-), dq..2dq-1 row k of B+
// (B-), then (alpha, dsci), (dscj, 0); rows k >= dq are zero
template <int NY, int MATS>
__device__ inline void set_abpm_cell(const DevCtx& d, const DevBlock& B, int i, int j, double* skew, long long nsl,
                                     long long slot) {
    using D = Dims<NY>;
    constexpr int dq = D::dq;
    const long long c = cidx(B, i, j), P = B.pitch;
    CellJac<NY> J;
    cell_jacobian<NY>(d, c, P, J);
    if (MATS == 1)
        for (int e = 0; e < dq * dq; e++) {
            FLD(d, SHCFD_F_AP, e, c) = J.Ap[e];
            FLD(d, SHCFD_F_AM, e, c) = J.Am[e];
            FLD(d, SHCFD_F_BP, e, c) = J.Bp[e];
            FLD(d, SHCFD_F_BM, e, c) = J.Bm[e];
        }
    if (MATS == 2)
        for (int e = 0; e < dq * dq; e++) {
            skew[(long long)(SKEW_AP * dq * dq + e) * nsl + slot] = J.Ap[e];
            skew[(long long)(SKEW_BP * dq * dq + e) * nsl + slot] = J.Bp[e];
            skew[(long long)(SKEW_AM * dq * dq + e) * nsl + slot] = J.Am[e];
            skew[(long long)(SKEW_BM * dq * dq + e) * nsl + slot] = J.Bm[e];
        }
    if (MATS == 3) {
        constexpr int C = (dq + 3) / 4, dq2 = dq + 2;
        const double dsci = FLD(d, SHCFD_F_DSCI, 0, c), dscj = FLD(d, SHCFD_F_DSCJ, 0, c);
        for (int set = 0; set < 2; set++) {
            const double* MA = set == 0 ? J.Ap : J.Am;
            const double* MB = set == 0 ? J.Bp : J.Bm;
            for (int e = 0; e < 2 * dq2; e++) {
                const long long g = slot >> 6, l0 = slot & 63;
                for (int w = 0; w < C; w++) {
                    double* o = skew + ((((set * nsl + g) * C + w) * dq2 + e / 2) * 64 + l0) * 2 + (e & 1);
                    for (int cc = 0; cc < 4; cc++) {
                        const int k = 4 * w + cc;
                        double v;
                        if (e < dq) v = (k < dq) ? MA[k + dq * e] : 0.0;
                        else if (e < 2 * dq) v = (k < dq) ? MB[k + dq * (e - dq)] : 0.0;
                        else v = e == 2 * dq ? J.alpha : (e == 2 * dq + 1 ? dsci : (e == 2 * dq + 2 ? dscj : 0.0));
                        o[2 * cc] = v;
                    }
                }
            }
        }
    }
    store_abpm_scalars<NY>(d, c, J);
}

}  // namespace shcfd
