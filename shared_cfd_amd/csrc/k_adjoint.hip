// k_adjoint.hip -- discrete-adjoint (reverse) kernels of the steady
// ideal-gas residual
//     R(c) = dsi(i-1) (TGi - TGvi)(i-1) - dsi(i) (TGi - TGvi)(i)
//          + dsj(j-1) (TGj - TGvj)(j-1) - dsj(j) (TGj - TGvj)(j)
// (LU-SGS/2d/time_lusgs_viscous.f90:12-22 with Area = 0) through the SLAU
// flux (core/scheme.f90:71-153), the viscous flux when the flow is viscous
// (viscosity/viscous/no-nV/2d/sch_viscous.f90:1-178, with the corner ghosts
// of cond/no-nV/condition.tail.f90:331-356 and the face metrics from the
// cell centres), MUSCL face states with the limiter
// (high_order/muscl/muscl.f90:7-73), the boundary ghost fills
// (cond/no-nV/condition.tail.f90:2-70 and the physical side rows), the
// primitive/thermo map (core/scheme.f90:1-29,
// therm_lib/ideal/thermal_model.raw.f90:7-33) and the face metrics
// (dim/2d/geometry.f90:79-169).
//
// Given lambda (one vector per interior cell) and an objective
//     J = sum over wall faces of dsj ((TG - TGv)_u a_x + (TG - TGv)_v a_y),
// one pass computes lambda^T dR/dq + dJ/dq per cell, lambda^T dR/dX + dJ/dX
// per node and the derivative with respect to the fixed boundary states.
// Local Jacobians come from forward-mode dual evaluations of the same
// scalar-generic functions the primal kernels use (shcfd_ad.h); their
// transposes are applied here.  Every output is a gather (one writer per
// element), except the fixed-state sums, which are reduced per boundary row
// in a fixed order, so the results are deterministic.
#include "shcfd_device.h"
#include "shcfd_kernels.h"
#include "shcfd_adj.h"

namespace shcfd {

using D1 = Dims<1>;

__device__ inline bool interior(const DevBlock& B, int i, int j) { return i >= 1 && i <= B.ni && j >= 1 && j <= B.nj; }

// ---------------------------------------------------------------------------
// faces: flux bars -> face-state, normal and length bars.
// DIR 0: i-face (i,j) between (i,j) and (i+1,j), i in [0,ni], j in [1,nj];
// DIR 1: j-face (i,j) between (i,j) and (i,j+1), i in [1,ni], j in [0,nj].
// ---------------------------------------------------------------------------
constexpr int NFV = 7;  // face-state components the flux depends on: rho,u,v,p,Y,gamma,H
__device__ inline int fv_comp(int k) { return k < 5 ? k : (k == 5 ? D1::ig : D1::ih); }

template <int DIR>
__global__ void __launch_bounds__(128) k_adj_face(DevCtx d, AdjView a, int b) {
    const DevBlock B = d.blk[b];
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + (DIR == 0 ? 0 : 1);
    const int j = (int)blockIdx.y + (DIR == 0 ? 1 : 0);
    if (DIR == 0 ? (i > B.ni) : (i > B.ni || j > B.nj)) return;
    const long long c0 = cidx(B, i, j), c1 = DIR == 0 ? c0 + 1 : c0 + B.pitch;
    const int i1 = DIR == 0 ? i + 1 : i, j1 = DIR == 0 ? j : j + 1;
    const bool in0 = interior(B, i, j), in1 = interior(B, i1, j1);
    const int fds = DIR == 0 ? SHCFD_F_DSI : SHCFD_F_DSJ, fvn = DIR == 0 ? SHCFD_F_VNI : SHCFD_F_VNJ;
    const int ftg = DIR == 0 ? SHCFD_F_TGI : SHCFD_F_TGJ;
    const int fl = DIR == 0 ? SHCFD_F_WHLI : SHCFD_F_WHLJ, fr = DIR == 0 ? SHCFD_F_WHRI : SHCFD_F_WHRJ;
    const double ds = FLD(d, fds, 0, c0);
    // the face's net flux TG - TGv (TGv zero when inviscid)
    const int ftv = DIR == 0 ? SHCFD_F_TGVI : SHCFD_F_TGVJ;
    auto net = [&](int k) { return FLD(d, ftg, k, c0) - (d.p.viscous ? FLD(d, ftv, k, c0) : 0.0); };
    double tgb[4], dsb = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double dl = (in1 ? a.lam[k * d.NC + c1] : 0.0) - (in0 ? a.lam[k * d.NC + c0] : 0.0);
        tgb[k] = ds * dl;
        dsb += net(k) * dl;
    }
    // objective: momentum flux through the wall faces of the objective segment
    if (DIR == 1 && j == 0 && b == a.obj_block && i >= a.obj_i0 && i <= a.obj_i1) {
        tgb[1] += ds * a.obj_ax;
        tgb[2] += ds * a.obj_ay;
        dsb += net(1) * a.obj_ax + net(2) * a.obj_ay;
    }
    using T = Dual<2 * NFV + 2>;
    T wl[D1::dw], wr[D1::dw], tg[4];
#pragma unroll
    for (int k = 0; k < D1::dw; k++) {
        wl[k] = T(FLD(d, fl, k, c0));
        wr[k] = T(FLD(d, fr, k, c0));
    }
#pragma unroll
    for (int m = 0; m < NFV; m++) {
        wl[fv_comp(m)].d[m] = 1.0;
        wr[fv_comp(m)].d[NFV + m] = 1.0;
    }
    const T n1 = T::var(FLD(d, fvn, 0, c0), 2 * NFV), n2 = T::var(FLD(d, fvn, 1, c0), 2 * NFV + 1);
    slau_face_t<1, T>(wl, wr, n1, n2, tg);
    const int gl = DIR == 0 ? AdjView::WLI : AdjView::WLJ, gr = DIR == 0 ? AdjView::WRI : AdjView::WRJ;
#pragma unroll
    for (int k = 0; k < D1::dw; k++) {
        a.face[gl][k * d.NC + c0] = 0.0;
        a.face[gr][k * d.NC + c0] = 0.0;
    }
#pragma unroll
    for (int m = 0; m < NFV; m++) {
        double sl = 0.0, sr = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            sl += tgb[k] * tg[k].d[m];
            sr += tgb[k] * tg[k].d[NFV + m];
        }
        a.face[gl][fv_comp(m) * d.NC + c0] = sl;
        a.face[gr][fv_comp(m) * d.NC + c0] = sr;
    }
    double nb1 = 0.0, nb2 = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        nb1 += tgb[k] * tg[k].d[2 * NFV];
        nb2 += tgb[k] * tg[k].d[2 * NFV + 1];
    }
    double* nbar = DIR == 0 ? a.nbar_i : a.nbar_j;
    double* dsbar = DIR == 0 ? a.dsbar_i : a.dsbar_j;
    nbar[c0] = nb1;
    nbar[d.NC + c0] = nb2;
    dsbar[c0] = dsb;
}
template __global__ void k_adj_face<0>(DevCtx, AdjView, int);
template __global__ void k_adj_face<1>(DevCtx, AdjView, int);

// ---------------------------------------------------------------------------
// viscous faces (ideal gas, nY = 1, planar).  R carries -ds TGv, so the bars
// of TGv are minus those TG gets.  Per face:
//  - the flux's dependence on its 14 intermediates (viscous_in_t) and 6
//    metrics (geojac, unit normal), in dual numbers through viscous_out_t;
//  - the intermediates' dependence on the w vectors of the 6 stencil cells
//    (face averages, xi/eta differences, T = (p/rho)/R), applied by hand;
//  - the metrics' dependence on the stencil's centre coordinates (geojac_t
//    in dual numbers).
// Stored per face and stencil position, gathered per cell by k_adj_muscl
// and k_adj_centres.  With nY = 1 the species-diffusion terms multiply the
// differences of Y = rho/rho, zero up to rounding, so the face enthalpies
// lhi are held constant.
// ---------------------------------------------------------------------------
constexpr int NVW = 8;  // w components of the viscous flux: rho, u, v, p, Y, gamma, R, mu
// slot of w component k among the NVW (-1: H, which the flux does not use)
__device__ inline int vw_slot(int k) { return k < D1::ih ? k : (k == D1::ih ? -1 : k - 1); }
// face (fi, fj) = (i + DFI[s], j + DFJ[s]) holds cell (i, j) at stencil position s
__constant__ int VDFI[2][6] = {{0, -1, -1, 0, -1, 0}, {0, 0, -1, -1, 1, 1}};
__constant__ int VDFJ[2][6] = {{0, 0, -1, -1, 1, 1}, {0, -1, -1, 0, -1, 0}};

template <int DIR>
__global__ void __launch_bounds__(128) k_adj_vface(DevCtx d, AdjView a, int b) {
    const DevBlock B = d.blk[b];
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + (DIR == 0 ? 0 : 1);
    const int j = (int)blockIdx.y + (DIR == 0 ? 1 : 0);
    if (DIR == 0 ? (i > B.ni) : (i > B.ni || j > B.nj)) return;
    const long long c0 = cidx(B, i, j);
    long long st[6];
    viscous_stencil<DIR>(c0, B.pitch, st);
    const int i1 = DIR == 0 ? i + 1 : i, j1 = DIR == 0 ? j : j + 1;
    const bool in0 = interior(B, i, j), in1 = interior(B, i1, j1);
    const double ds = FLD(d, DIR == 0 ? SHCFD_F_DSI : SHCFD_F_DSJ, 0, c0);
    double tb[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        tb[k] = -ds * ((in1 ? a.lam[k * d.NC + st[1]] : 0.0) - (in0 ? a.lam[k * d.NC + c0] : 0.0));
    if (DIR == 1 && j == 0 && b == a.obj_block && i >= a.obj_i0 && i <= a.obj_i1) {
        tb[1] -= ds * a.obj_ax;
        tb[2] -= ds * a.obj_ay;
    }
    double ws[6][D1::dw];
    const double* wp[6];
#pragma unroll
    for (int s = 0; s < 6; s++) {
#pragma unroll
        for (int k = 0; k < D1::dw; k++) ws[s][k] = FLD(d, SHCFD_F_W, k, st[s]);
        wp[s] = ws[s];
    }
    ViscIn<1, double> z0;
    viscous_in_t<1, DIR, double>(wp, z0);
    const int fg = DIR == 0 ? SHCFD_F_GEOJACI : SHCFD_F_GEOJACJ, fvn = DIR == 0 ? SHCFD_F_VNI : SHCFD_F_VNJ;
    double zb[20];
    {
        using T = Dual<20>;
        ViscIn<1, T> z;
        z.rho = T::var(z0.rho, 0);
        z.u = T::var(z0.u, 1);
        z.v = T::var(z0.v, 2);
        z.gamm = T::var(z0.gamm, 3);
        z.Rg = T::var(z0.Rg, 4);
        z.mu = T::var(z0.mu, 5);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            z.dxi[q] = T::var(z0.dxi[q], 6 + q);
            z.deta[q] = T::var(z0.deta[q], 10 + q);
        }
        T g[4];
#pragma unroll
        for (int q = 0; q < 4; q++) g[q] = T::var(FLD(d, fg, q, c0), 14 + q);
        const T n1 = T::var(FLD(d, fvn, 0, c0), 18), n2 = T::var(FLD(d, fvn, 1, c0), 19);
        const T lhi[1] = {T((FLD(d, SHCFD_F_VHI, 0, c0) + FLD(d, SHCFD_F_VHI, 0, st[1])) * 0.5)};
        T out[4];
        viscous_out_t<1, T>(z, lhi, g, n1, n2, false, T(1.0), out);
#pragma unroll
        for (int m = 0; m < 20; m++) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) acc += tb[k] * out[k].d[m];
            zb[m] = acc;
        }
    }
    double* nbar = DIR == 0 ? a.nbar_i : a.nbar_j;
    nbar[c0] += zb[18];
    nbar[d.NC + c0] += zb[19];
    // the c1 - c differences are the xi ones for DIR 0, the eta ones for DIR 1
    const int on = DIR == 0 ? 6 : 10, oc = DIR == 0 ? 10 : 6;
    double* vw = DIR == 0 ? a.vw_i : a.vw_j;
#pragma unroll
    for (int s = 0; s < 6; s++) {
        double wb[NVW];
#pragma unroll
        for (int m = 0; m < NVW; m++) wb[m] = 0.0;
        if (s < 2) {  // face averages of c and c1
            wb[0] = 0.5 * zb[0];
            wb[1] = 0.5 * zb[1];
            wb[2] = 0.5 * zb[2];
            wb[5] = 0.5 * zb[3];
            wb[6] = 0.5 * zb[4];
            wb[7] = 0.5 * zb[5];
        }
        const double fn = s == 0 ? -1.0 : (s == 1 ? 1.0 : 0.0);
        const double fc = s < 2 ? 0.0 : (s < 4 ? 0.25 : -0.25);
        const double ub = fn * zb[on] + fc * zb[oc], vb = fn * zb[on + 1] + fc * zb[oc + 1];
        const double Tb = fn * zb[on + 2] + fc * zb[oc + 2], Yb = fn * zb[on + 3] + fc * zb[oc + 3];
        wb[1] += ub;
        wb[2] += vb;
        wb[4] += Yb;
        // T = (p / rho) / R
        const double rho = ws[s][IW_RHO], R = ws[s][D1::ir];
        const double Tv = (ws[s][IW_P] / rho) / R;
        wb[3] += Tb / (rho * R);
        wb[0] -= Tb * Tv / rho;
        wb[6] -= Tb * Tv / R;
#pragma unroll
        for (int m = 0; m < NVW; m++) vw[(s * NVW + m) * d.NC + c0] = wb[m];
    }
    // metric bars -> centre-coordinate bars of the stencil
    using G = Dual<12>;
    G xc[6], rc[6], gg[4];
#pragma unroll
    for (int s = 0; s < 6; s++) {
        xc[s] = G::var(FLD(d, SHCFD_F_X, 0, st[s]), s);
        rc[s] = G::var(FLD(d, SHCFD_F_R, 0, st[s]), 6 + s);
    }
    geojac_t<DIR, G>(xc, rc, gg);
    double* vg = DIR == 0 ? a.vg_i : a.vg_j;
#pragma unroll
    for (int s = 0; s < 6; s++) {
        double xb = 0.0, rb = 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            xb += zb[14 + q] * gg[q].d[s];
            rb += zb[14 + q] * gg[q].d[6 + s];
        }
        vg[(2 * s) * d.NC + c0] = xb;
        vg[(2 * s + 1) * d.NC + c0] = rb;
    }
}
template __global__ void k_adj_vface<0>(DevCtx, AdjView, int);
template __global__ void k_adj_vface<1>(DevCtx, AdjView, int);

// sum over the viscous faces holding cell (i, j) of their stored bars
// (component slot m of nper per position)
__device__ inline double viscous_gather(const DevCtx& d, const DevBlock& B, const double* fi_arr,
                                        const double* fj_arr, int nper, int m, int i, int j) {
    double s = 0.0;
#pragma unroll
    for (int p = 0; p < 6; p++) {
        const int fi = i + VDFI[0][p], fj = j + VDFJ[0][p];
        if (fi >= 0 && fi <= B.ni && fj >= 1 && fj <= B.nj)
            s += fi_arr[(long long)(p * nper + m) * d.NC + cidx(B, fi, fj)];
    }
#pragma unroll
    for (int p = 0; p < 6; p++) {
        const int fi = i + VDFI[1][p], fj = j + VDFJ[1][p];
        if (fi >= 1 && fi <= B.ni && fj >= 0 && fj <= B.nj)
            s += fj_arr[(long long)(p * nper + m) * d.NC + cidx(B, fi, fj)];
    }
    return s;
}

// ---------------------------------------------------------------------------
// MUSCL: face-state bars -> w bars of the stencil cells (gather per cell,
// ghosts included).  For face f of a line: wL(f) = w_f + hp du(w_{f+1}-w_f,
// w_f-w_{f-1}), wR(f) = w_{f+1} - hp dl(w_{f+2}-w_{f+1}, w_{f+1}-w_f).
// ---------------------------------------------------------------------------
struct MusclD {
    double l[3];  // dwL(f)/dw_{f-1}, dw_f, dw_{f+1}
    double r[3];  // dwR(f)/dw_f, dw_{f+1}, dw_{f+2}
};
__device__ inline MusclD muscl_derivs(double wm, double w0, double w1, double w2, const DevCtx& d) {
    MusclD m;
    if (!d.p.high_order) {
        m.l[0] = 0.0; m.l[1] = 1.0; m.l[2] = 0.0;
        m.r[0] = 0.0; m.r[1] = 1.0; m.r[2] = 0.0;
        return m;
    }
    const double hp = d.p.phi_mus * 0.5, kk = d.p.k_mus;
    const int lim = d.p.limiter;
    using T = Dual<2>;
    T du, dl;
    muscl_delta_t<T>(T::var(w1 - w0, 0), T::var(w0 - wm, 1), kk, lim, du, dl);
    // du over (Du, Dl) of cell f
    m.l[0] = -hp * du.d[1];
    m.l[1] = 1.0 + hp * (du.d[1] - du.d[0]);
    m.l[2] = hp * du.d[0];
    T du1, dl1;
    muscl_delta_t<T>(T::var(w2 - w1, 0), T::var(w1 - w0, 1), kk, lim, du1, dl1);
    // dl over (Du', Dl') of cell f+1
    m.r[0] = hp * dl1.d[1];
    m.r[1] = 1.0 - hp * (dl1.d[1] - dl1.d[0]);
    m.r[2] = -hp * dl1.d[0];
    return m;
}

__global__ void __launch_bounds__(128) k_adj_muscl(DevCtx d, AdjView a, int b) {
    const DevBlock B = d.blk[b];
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 1;
    const int j = (int)blockIdx.y - 1;
    if (i > B.ni + 2 || j > B.nj + 2) return;
    const long long c = cidx(B, i, j), P = B.pitch;
    const bool row = j >= 1 && j <= B.nj, col = i >= 1 && i <= B.ni;
    for (int k = 0; k < D1::dw; k++) {
        const double* W = d.f[SHCFD_F_W] + (long long)k * d.NC;
        double s = 0.0;
        if (row) {
            // faces f = i-2 .. i+1 of row j (f in [0, ni])
            for (int f = i - 2; f <= i + 1; f++) {
                if (f < 0 || f > B.ni) continue;
                const long long cf = cidx(B, f, j);
                const MusclD m = muscl_derivs(W[cf - 1], W[cf], W[cf + 1], W[cf + 2], d);
                const double bl = a.face[AdjView::WLI][k * d.NC + cf], br = a.face[AdjView::WRI][k * d.NC + cf];
                const int o = i - f;  // position of this cell relative to f
                if (o >= -1 && o <= 1) s += bl * m.l[o + 1];
                if (o >= 0 && o <= 2) s += br * m.r[o];
            }
        }
        if (col) {
            for (int f = j - 2; f <= j + 1; f++) {
                if (f < 0 || f > B.nj) continue;
                const long long cf = cidx(B, i, f);
                const MusclD m = muscl_derivs(W[cf - P], W[cf], W[cf + P], W[cf + 2 * P], d);
                const double bl = a.face[AdjView::WLJ][k * d.NC + cf], br = a.face[AdjView::WRJ][k * d.NC + cf];
                const int o = j - f;
                if (o >= -1 && o <= 1) s += bl * m.l[o + 1];
                if (o >= 0 && o <= 2) s += br * m.r[o];
            }
        }
        if (d.p.viscous && i >= 0 && i <= B.ni + 1 && j >= 0 && j <= B.nj + 1 && vw_slot(k) >= 0)
            s += viscous_gather(d, B, a.vw_i, a.vw_j, NVW, vw_slot(k), i, j);
        a.wbar[k * d.NC + c] = s;
    }
}

// corner ghosts (set_corners, cond/no-nV/condition.tail.f90:331-356: the
// mean of the corner cell and its two side ghosts), reversed before the
// side rows; one thread per block
__global__ void k_adj_bc_corner(DevCtx d, AdjView a) {
    const int b = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= d.nb) return;
    const DevBlock B = d.blk[b];
    for (int q = 0; q < 4; q++) {
        const int pi = (q & 1) ? 1 : -1, pj = (q & 2) ? 1 : -1;
        const int i = pi < 0 ? 1 : B.ni, j = pj < 0 ? 1 : B.nj;
        const long long cc = cidx(B, i + pi, j + pj), c0 = cidx(B, i, j), cj = cidx(B, i, j + pj),
                        ci = cidx(B, i + pi, j);
        for (int k = 0; k < D1::dw; k++) {
            const double x = a.wbar[k * d.NC + cc] * (1.0 / 3.0);
            a.wbar[k * d.NC + cc] = 0.0;
            a.wbar[k * d.NC + c0] += x;
            a.wbar[k * d.NC + cj] += x;
            a.wbar[k * d.NC + ci] += x;
        }
    }
}

// centre-coordinate bars per cell (ghost centres included) from the faces
__global__ void __launch_bounds__(128) k_adj_centres(DevCtx d, AdjView a, int b) {
    const DevBlock B = d.blk[b];
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)blockIdx.y;
    if (i > B.ni + 1 || j > B.nj + 1) return;
    const long long c = cidx(B, i, j);
    a.cbar[c] = viscous_gather(d, B, a.vg_i, a.vg_j, 2, 0, i, j);
    a.cbar[d.NC + c] = viscous_gather(d, B, a.vg_i, a.vg_j, 2, 1, i, j);
}

// ghost centres (k_geom_ghost: x(0) = 2 x(1) - x(2) per side; pass 1 the
// i-sides, then pass 2 the j-sides over i = 0..ni+1) reversed: pass 2 first
__global__ void k_adj_ghost_centres(DevCtx d, AdjView a, int b, int pass) {
    const DevBlock B = d.blk[b];
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    auto fold = [&](long long g, long long c1, long long c2) {
        for (int q = 0; q < 2; q++) {
            double* cb = a.cbar + q * d.NC;
            const double x = cb[g];
            cb[g] = 0.0;
            cb[c1] += 2.0 * x;
            cb[c2] -= x;
        }
    };
    if (pass == 2) {
        const int i = t;
        if (i > B.ni + 1) return;
        fold(cidx(B, i, 0), cidx(B, i, 1), cidx(B, i, 2));
        fold(cidx(B, i, B.nj + 1), cidx(B, i, B.nj), cidx(B, i, B.nj - 1));
    } else {
        const int j = t + 1;
        if (j > B.nj) return;
        fold(cidx(B, 0, j), cidx(B, 1, j), cidx(B, 2, j));
        fold(cidx(B, B.ni + 1, j), cidx(B, B.ni, j), cidx(B, B.ni - 1, j));
    }
}

// ---------------------------------------------------------------------------
// boundary rows: ghost bars -> source cells (one launch per row, one thread
// per position along the side, one workgroup so the fixed-state sums reduce
// in a fixed order)
// ---------------------------------------------------------------------------
// one side row, by all threads of the workgroup (red: 256 doubles of LDS)
__device__ inline void adj_bc_row(const DevCtx& d, AdjView& a, const int* row, const double* pool, double* poolbar,
                                  double* red) {
    const int b = row[0] - 1, side = row[1], kind = row[4], off = row[5];
    const DevBlock B = d.blk[b];
    const int n = (side <= SHCFD_SIDE_IMAX) ? B.nj : B.ni;
    const int s = row[2] < 1 ? 1 : row[2], e = row[3] > n ? n : row[3];
    double pacc[D1::dw];
#pragma unroll
    for (int k = 0; k < D1::dw; k++) pacc[k] = 0.0;
    for (int tpos = s + (int)threadIdx.x; tpos <= e; tpos += blockDim.x) {
        int ii, jj, ig1, jg1, ig2, jg2;
        switch (side) {
        case SHCFD_SIDE_IMIN: ii = 1; jj = tpos; ig1 = 0; jg1 = tpos; ig2 = -1; jg2 = tpos; break;
        case SHCFD_SIDE_IMAX: ii = B.ni; jj = tpos; ig1 = B.ni + 1; jg1 = tpos; ig2 = B.ni + 2; jg2 = tpos; break;
        case SHCFD_SIDE_JMIN: ii = tpos; jj = 1; ig1 = tpos; jg1 = 0; ig2 = tpos; jg2 = -1; break;
        default: ii = tpos; jj = B.nj; ig1 = tpos; jg1 = B.nj + 1; ig2 = tpos; jg2 = B.nj + 2; break;
        }
        const long long ci = cidx(B, ii, jj), c1 = cidx(B, ig1, jg1), c2 = cidx(B, ig2, jg2);
        double gb[D1::dw];
#pragma unroll
        for (int k = 0; k < D1::dw; k++) {
            gb[k] = a.wbar[k * d.NC + c1] + a.wbar[k * d.NC + c2];
            a.wbar[k * d.NC + c1] = 0.0;
            a.wbar[k * d.NC + c2] = 0.0;
        }
        if (kind == SHCFD_BC_FIXED) {
#pragma unroll
            for (int k = 0; k < D1::dw; k++) pacc[k] += gb[k];
            continue;
        }
        if (kind == SHCFD_BC_REFLECT) {
            const int kv = (side <= SHCFD_SIDE_IMAX) ? IW_U : IW_V;
            gb[kv] = -gb[kv];
        } else if (kind == SHCFD_BC_NOSLIP) {
            gb[IW_U] = -gb[IW_U];
            gb[IW_V] = -gb[IW_V];
        } else if (kind == SHCFD_BC_OUTFLOW_P) {
            gb[IW_P] = 0.0;
        } else if (kind == SHCFD_BC_INFLOW_TUV) {
            // ghost (p/(R T), u, v, p, 1, kappa, cp T + (u^2 + v^2)/2, R, nu p/(R T)) with (T, u, v) from
            // the pool and p from the cell
            const double* pv = pool + off;
            const double Tq = pv[0], uq = pv[1], vq = pv[2], rg = d.p.rgas, kap = d.p.kappa;
            const double p = FLD(d, SHCFD_F_W, IW_P, ci), rt = rg * Tq;
            const double rb = gb[IW_RHO] + B.nu * gb[D1::imu];  // bar of rho_ghost (mu = nu rho)
            pacc[0] += -rb * p / (rt * Tq) + gb[D1::ih] * (kap / (kap - 1.0)) * rg;
            pacc[1] += gb[IW_U] + gb[D1::ih] * uq;
            pacc[2] += gb[IW_V] + gb[D1::ih] * vq;
            const double pb = gb[IW_P] + rb / rt;
#pragma unroll
            for (int k = 0; k < D1::dw; k++) gb[k] = 0.0;
            gb[IW_P] = pb;
        } else if (kind == SHCFD_BC_SLIP_NORMAL) {
            long long cf;
            double* nbar;
            if (side == SHCFD_SIDE_IMIN) { cf = cidx(B, 0, jj); nbar = a.nbar_i; }
            else if (side == SHCFD_SIDE_IMAX) { cf = cidx(B, B.ni, jj); nbar = a.nbar_i; }
            else if (side == SHCFD_SIDE_JMIN) { cf = cidx(B, ii, 0); nbar = a.nbar_j; }
            else { cf = cidx(B, ii, B.nj); nbar = a.nbar_j; }
            const int fv = side <= SHCFD_SIDE_IMAX ? SHCFD_F_VNI : SHCFD_F_VNJ;
            const double nx = FLD(d, fv, 0, cf), ny = FLD(d, fv, 1, cf);
            const double u = FLD(d, SHCFD_F_W, IW_U, ci), v = FLD(d, SHCFD_F_W, IW_V, ci);
            // ghost velocity g = u - 2 (n.u) n
            const double ub = gb[IW_U], vb = gb[IW_V];
            const double un = nx * u + ny * v, bn = nx * ub + ny * vb;
            gb[IW_U] = ub - 2.0 * bn * nx;
            gb[IW_V] = vb - 2.0 * bn * ny;
            nbar[cf] += -2.0 * (bn * u + un * ub);
            nbar[d.NC + cf] += -2.0 * (bn * v + un * vb);
        }
#pragma unroll
        for (int k = 0; k < D1::dw; k++) a.wbar[k * d.NC + ci] += gb[k];
    }
    if (kind != SHCFD_BC_FIXED && kind != SHCFD_BC_INFLOW_TUV) return;
    const int npv = kind == SHCFD_BC_FIXED ? D1::dw : 3;  // pool entries with bars
    for (int k = 0; k < npv; k++) {
        red[threadIdx.x] = pacc[k];
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) poolbar[off + k] += red[0];
        __syncthreads();
    }
}

// one cut row (t), positions over the workgroup's threads; each position's
// copies in the reverse of the forward fill
__device__ inline void adj_cut_row(const DevCtx& d, AdjView& a, const int* t) {
    for (int inc = (int)threadIdx.x; inc < t[1]; inc += (int)blockDim.x)
        for (int copy = 3; copy >= 0; copy--) {
            int bd, di, dj, bs, si, sj;
            if (!cut_map(t, inc, copy, bd, di, dj, bs, si, sj)) break;
            const long long g = cidx(d.blk[bd], di, dj), src = cidx(d.blk[bs], si, sj);
#pragma unroll
            for (int k = 0; k < D1::dw; k++) {
                a.wbar[k * d.NC + src] += a.wbar[k * d.NC + g];
                a.wbar[k * d.NC + g] = 0.0;
            }
        }
}

// the boundary rows of the reverse pass in one workgroup, in the reverse of
// the forward fill: the side rows, last row first, then the cut rows, last
// first.  Rows that meet at a block corner add into the same interior cell,
// so each row starts after a barrier.  The pool bars are cleared first.
__global__ void __launch_bounds__(256) k_adj_bc_rows(DevCtx d, AdjView a, const int* bc, int nbc, const double* pool,
                                                     double* poolbar, int npool, const int* cut, int ncut) {
    __shared__ double red[256];
    for (int k = (int)threadIdx.x; k < npool; k += (int)blockDim.x) poolbar[k] = 0.0;
    __syncthreads();
    for (int r = nbc - 1; r >= 0; r--) {
        adj_bc_row(d, a, bc + SHCFD_BC_ROW * r, pool, poolbar, red);
        __syncthreads();
    }
    for (int r = ncut - 1; r >= 0; r--) {
        adj_cut_row(d, a, cut + SHCFD_CUT_ROW * r);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// cells: w bars -> q bars through set_w and the ideal-gas thermo map
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline void ideal_w(const T* q, double kap, double rg, double nu, T* w) {
    T rho = 0.0 + q[0];
    T inv = 1.0 / rho;
    T u = q[1] * inv, v = q[2] * inv;
    w[IW_RHO] = rho;
    w[IW_U] = u;
    w[IW_V] = v;
    w[IW_Y(0)] = q[0] * inv;
    T ei = q[3] / rho - 0.5 * ((u * u) + (v * v));
    T Tt = (ei * (kap - 1.0)) / rg;
    T p = (rho * rg) * Tt;
    w[IW_P] = p;
    w[D1::ig] = T(kap);
    w[D1::ih] = (q[3] + p) / rho;
    w[D1::imu] = rho * nu;
    w[D1::ir] = T(rg);
}

__global__ void __launch_bounds__(128) k_adj_cell(DevCtx d, AdjView a, int b) {
    const DevBlock B = d.blk[b];
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1, j = (int)blockIdx.y + 1;
    if (i > B.ni || j > B.nj) return;
    const long long c = cidx(B, i, j);
    using T = Dual<4>;
    T q[4], w[D1::dw];
#pragma unroll
    for (int k = 0; k < 4; k++) q[k] = T::var(FLD(d, SHCFD_F_Q, k, c), k);
    ideal_w<T>(q, d.p.kappa, d.p.rgas, B.nu, w);
    // unsteady: the BDF2 term -Vol a0/dt q^n of U^n and the coupling to the
    // later steps' residuals
    const double sh = a.unsteady ? FLD(d, SHCFD_F_VOL, 0, c) * (a.unsteady == 2 ? 1.0 : 1.5) / d.scal[SCAL_DT_GRBL]
                                 : 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < D1::dw; m++) s += a.wbar[m * d.NC + c] * w[m].d[k];
        if (a.unsteady) s = (s + a.src[k * d.NC + c]) - sh * a.lam[k * d.NC + c];
        a.qbar[k * d.NC + c] = s;
    }
}

// start of an unsteady adjoint step: src = Vol/dt (2 lambda^(n+1) -
// lambda^(n+2)/2) from lam (= lambda^(n+1)) and lam2 (= lambda^(n+2)), then
// lam2 = lambda^(n+1); lam stays as the initial guess of lambda^n
__global__ void __launch_bounds__(128) k_adj_unsteady_begin(DevCtx d, AdjView a, int b) {
    const DevBlock B = d.blk[b];
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1, j = (int)blockIdx.y + 1;
    if (i > B.ni || j > B.nj) return;
    const long long c = cidx(B, i, j);
    const double f = FLD(d, SHCFD_F_VOL, 0, c) / d.scal[SCAL_DT_GRBL];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double l1 = a.lam[k * d.NC + c], l2 = a.lam2[k * d.NC + c];
        a.src[k * d.NC + c] = f * (2.0 * l1 - 0.5 * l2);
        a.lam2[k * d.NC + c] = l1;
    }
}

// d Vol / d(node) of cell (ci, cj) for its corner (ci-1+oi, cj-1+oj): Vol is
// the signed area of triangles (A, B, D) + (B, C, D) with A = (ci-1, cj-1),
// B = (ci, cj-1), C = (ci, cj), D = (ci-1, cj) as k_geom_a computes it
__device__ inline void vol_node_grad(const DevCtx& d, const DevBlock& B, int ci, int cj, int oi, int oj, double& gx,
                                     double& gy) {
    auto X = [&](int ii, int jj) { return FLD(d, SHCFD_F_XH, 0, cidx(B, ii, jj)); };
    auto Y = [&](int ii, int jj) { return FLD(d, SHCFD_F_RH, 0, cidx(B, ii, jj)); };
    const double ax = X(ci - 1, cj - 1), ay = Y(ci - 1, cj - 1), bx = X(ci, cj - 1), by = Y(ci, cj - 1);
    const double cx = X(ci, cj), cy = Y(ci, cj), dx = X(ci - 1, cj), dy = Y(ci - 1, cj);
    // T(a, b, c) = (a1 (b2 - c2) + b1 (c2 - a2) + c1 (a2 - b2)) / 2
    if (oi == 0 && oj == 0) {         // A: first vertex of T(A, B, D)
        gx = 0.5 * (by - dy); gy = 0.5 * (dx - bx);
    } else if (oi == 1 && oj == 0) {  // B: second of T(A, B, D), first of T(B, C, D)
        gx = 0.5 * (dy - ay) + 0.5 * (cy - dy); gy = 0.5 * (ax - dx) + 0.5 * (dx - cx);
    } else if (oi == 1 && oj == 1) {  // C: second of T(B, C, D)
        gx = 0.5 * (dy - by); gy = 0.5 * (bx - dx);
    } else {                          // D: third of both
        gx = 0.5 * (ay - by) + 0.5 * (by - cy); gy = 0.5 * (bx - ax) + 0.5 * (cx - bx);
    }
}

// end of an unsteady adjoint step: node bars of this step (XBAR of the last
// reverse pass plus the volume term -lambda.dq/dt dVol/dX of U^n) added to
// the sums
__global__ void __launch_bounds__(128) k_adj_unsteady_accum(DevCtx d, AdjView a, int b) {
    const DevBlock B = d.blk[b];
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)blockIdx.y;
    if (i > B.ni || j > B.nj) return;
    const long long n = cidx(B, i, j);
    double sx = a.xbar[n], sy = a.ybar[n];
    // the four cells around node (i, j): cell (ci, cj) sees it as corner (i-ci+1, j-cj+1)
#pragma unroll
    for (int oj = 0; oj < 2; oj++)
#pragma unroll
        for (int oi = 0; oi < 2; oi++) {
            const int ci = i + 1 - oi, cj = j + 1 - oj;
            if (ci < 1 || ci > B.ni || cj < 1 || cj > B.nj) continue;
            const long long c = cidx(B, ci, cj);
            double vb = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) vb -= a.lam[k * d.NC + c] * FLD(d, SHCFD_F_DQDT, k, c);
            double gx, gy;
            vol_node_grad(d, B, ci, cj, oi, oj, gx, gy);
            sx += vb * gx;
            sy += vb * gy;
        }
    a.xacc[n] += sx;
    a.yacc[n] += sy;
}

__global__ void k_adj_pool_acc(const double* poolbar, double* poolacc, int n) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k < n) poolacc[k] += poolbar[k];
}

// ---------------------------------------------------------------------------
// nodes: face length and normal bars -> node coordinate bars.
// j-face (i,j): A = node (i-1,j), B = node (i,j), n = (-dy, dx)/|d|;
// i-face (i,j): A = node (i,j-1), B = node (i,j), n = (dy, -dx)/|d|.
// ---------------------------------------------------------------------------
__device__ inline void face_geom_bar(double ax, double ay, double bx, double by, double nb1, double nb2, double dsb,
                                     bool jface, double& gx, double& gy) {
    const double dx = bx - ax, dy = by - ay;
    const double L = sqrt(dx * dx + dy * dy);
    const double n1 = jface ? -dy / L : dy / L, n2 = jface ? dx / L : -dx / L;
    const double dot = nb1 * n1 + nb2 * n2;
    const double m1 = (nb1 - dot * n1) / L, m2 = (nb2 - dot * n2) / L;
    gx = dsb * dx / L + (jface ? m2 : -m2);
    gy = dsb * dy / L + (jface ? -m1 : m1);
}

__global__ void __launch_bounds__(128) k_adj_nodes(DevCtx d, AdjView a, int b) {
    const DevBlock B = d.blk[b];
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)blockIdx.y;
    if (i > B.ni || j > B.nj) return;
    auto X = [&](int ii, int jj) { return FLD(d, SHCFD_F_XH, 0, cidx(B, ii, jj)); };
    auto Y = [&](int ii, int jj) { return FLD(d, SHCFD_F_RH, 0, cidx(B, ii, jj)); };
    double sx = 0.0, sy = 0.0, gx, gy;
    // j-face (i,j): this node is B
    if (i >= 1) {
        const long long f = cidx(B, i, j);
        face_geom_bar(X(i - 1, j), Y(i - 1, j), X(i, j), Y(i, j), a.nbar_j[f], a.nbar_j[d.NC + f], a.dsbar_j[f], true,
                      gx, gy);
        sx += gx; sy += gy;
    }
    // j-face (i+1,j): this node is A
    if (i + 1 <= B.ni) {
        const long long f = cidx(B, i + 1, j);
        face_geom_bar(X(i, j), Y(i, j), X(i + 1, j), Y(i + 1, j), a.nbar_j[f], a.nbar_j[d.NC + f], a.dsbar_j[f], true,
                      gx, gy);
        sx -= gx; sy -= gy;
    }
    // i-face (i,j): this node is B
    if (j >= 1) {
        const long long f = cidx(B, i, j);
        face_geom_bar(X(i, j - 1), Y(i, j - 1), X(i, j), Y(i, j), a.nbar_i[f], a.nbar_i[d.NC + f], a.dsbar_i[f], false,
                      gx, gy);
        sx += gx; sy += gy;
    }
    // i-face (i,j+1): this node is A
    if (j + 1 <= B.nj) {
        const long long f = cidx(B, i, j + 1);
        face_geom_bar(X(i, j), Y(i, j), X(i, j + 1), Y(i, j + 1), a.nbar_i[f], a.nbar_i[d.NC + f], a.dsbar_i[f], false,
                      gx, gy);
        sx -= gx; sy -= gy;
    }
    if (d.p.viscous) {
        // cell centres are the means of their four nodes: node (i, j) is a
        // corner of cells (i..i+1, j..j+1)
        for (int cj = j; cj <= j + 1; cj++)
            for (int ci = i; ci <= i + 1; ci++)
                if (interior(B, ci, cj)) {
                    const long long cc = cidx(B, ci, cj);
                    sx += 0.25 * a.cbar[cc];
                    sy += 0.25 * a.cbar[d.NC + cc];
                }
    }
    const long long c = cidx(B, i, j);
    a.xbar[c] = sx;
    a.ybar[c] = sy;
}

// objective value (one workgroup, fixed-order reduction)
__global__ void __launch_bounds__(256) k_adj_objective(DevCtx d, AdjView a, double* out) {
    __shared__ double red[256];
    const DevBlock B = d.blk[a.obj_block];
    double s = 0.0;
    for (int i = a.obj_i0 + (int)threadIdx.x; i <= a.obj_i1; i += blockDim.x) {
        const long long c = cidx(B, i, 0);
        const double tu = FLD(d, SHCFD_F_TGJ, 1, c) - (d.p.viscous ? FLD(d, SHCFD_F_TGVJ, 1, c) : 0.0);
        const double tv = FLD(d, SHCFD_F_TGJ, 2, c) - (d.p.viscous ? FLD(d, SHCFD_F_TGVJ, 2, c) : 0.0);
        s += FLD(d, SHCFD_F_DSJ, 0, c) * (tu * a.obj_ax + tv * a.obj_ay);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// sum of squares of the q bars over interior cells: per-workgroup partials
__global__ void __launch_bounds__(256) k_adj_norm_part(DevCtx d, AdjView a, int b, double* part) {
    __shared__ double red[256];
    const DevBlock B = d.blk[b];
    const long long n = (long long)B.ni * B.nj;
    double s = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long c = cidx(B, (int)(t % B.ni) + 1, (int)(t / B.ni) + 1);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double x = a.qbar[k * d.NC + c];
            s += x * x;
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

}  // namespace shcfd
