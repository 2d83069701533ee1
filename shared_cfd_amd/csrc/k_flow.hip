// k_flow.hip -- per-routine kernels of the flow path (the "plain" drop-in:
// one kernel per reference subroutine, face states and Jacobians stored in
// HBM exactly as the reference stores them in its module arrays).
//
// Each kernel's arithmetic follows the cited reference routine in the same
// evaluation order; the library is built with -ffp-contract=off, so no
// multiply-add is fused behind the source's back.
#include "shcfd_device.h"
#include "shcfd_kernels.h"

namespace shcfd {

// ---------------------------------------------------------------------------
// set_w: core/scheme.f90:1-29 (interior cells)
// set_thermo_prop, ideal gas: therm_lib/ideal/thermal_model.raw.f90:7-33
// ---------------------------------------------------------------------------
template <int NY>
__global__ void __launch_bounds__(256) k_set_w(DevCtx d, const Tile* tiles) {
    using D = Dims<NY>;
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    if (t.i < 1 || t.i > B.ni || t.j < 1 || t.j > B.nj) return;
    long long c = cidx(B, t.i, t.j);
    double q[D::dq];
#pragma unroll
    for (int k = 0; k < D::dq; k++) q[k] = FLD(d, SHCFD_F_Q, k, c);
    double tmp = 0.0;
#pragma unroll
    for (int k = 0; k < NY; k++) tmp = tmp + q[k];
    FLD(d, SHCFD_F_W, IW_RHO, c) = tmp;
    tmp = 1.0 / tmp;
    FLD(d, SHCFD_F_W, IW_U, c) = q[NY] * tmp;
    FLD(d, SHCFD_F_W, IW_V, c) = q[NY + 1] * tmp;
#pragma unroll
    for (int k = 0; k < NY; k++) FLD(d, SHCFD_F_W, IW_Y(k), c) = q[k] * tmp;
}

template <int NY>
__global__ void __launch_bounds__(256) k_thermo_ideal(DevCtx d, const Tile* tiles) {
    using D = Dims<NY>;
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    if (t.i < 1 || t.i > B.ni || t.j < 1 || t.j > B.nj) return;
    long long c = cidx(B, t.i, t.j);
    const double kap = d.p.kappa, rg = d.p.rgas, nu = B.nu;
    double rho = FLD(d, SHCFD_F_W, IW_RHO, c), u = FLD(d, SHCFD_F_W, IW_U, c), v = FLD(d, SHCFD_F_W, IW_V, c);
    double qe = FLD(d, SHCFD_F_Q, NY + 2, c);
    double ei = qe / rho - 0.5 * ((u * u) + (v * v));
    double T = (ei * (kap - 1.0)) / rg;
    double p = (rho * rg) * T;
    FLD(d, SHCFD_F_W, IW_P, c) = p;
    FLD(d, SHCFD_F_W, D::ig, c) = kap;
    FLD(d, SHCFD_F_W, D::ih, c) = (qe + p) / rho;
    FLD(d, SHCFD_F_W, D::imu, c) = rho * nu;
    FLD(d, SHCFD_F_W, D::ir, c) = rg;
    FLD(d, SHCFD_F_DHI, 0, c) = 0.0;
    FLD(d, SHCFD_F_VHI, 0, c) = ((kap / (kap - 1.0)) * rg) * T;
}

// set_w followed by the ideal-gas set_thermo_prop inside a whole step: one
// pass over the cells, the same expressions as k_set_w and k_thermo_ideal.
// PRIM_COPY_QP: the step's copy_qp (QP = q, as k_copy_q) of the q read here;
// PRIM_APPLY_UPDATE: then the previous inner iteration's q += omega Dq, with
// k_update's expression (time_dual_viscous.f90:113-124)
template <int NY>
__global__ void __launch_bounds__(256) k_primitive_ideal(DevCtx d, const Tile* tiles, int flags) {
    using D = Dims<NY>;
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    if (t.i < 1 || t.i > B.ni || t.j < 1 || t.j > B.nj) return;
    long long c = cidx(B, t.i, t.j);
    double q[D::dq];
#pragma unroll
    for (int k = 0; k < D::dq; k++) q[k] = FLD(d, SHCFD_F_Q, k, c);
    if (flags & PRIM_COPY_QP) {  // copy_qp of the step start: the q read here
#pragma unroll
        for (int k = 0; k < D::dq; k++) FLD(d, SHCFD_F_QP, k, c) = q[k];
    }
    if ((flags & PRIM_APPLY_UPDATE) && B.active) {
        const double om = dmin(d.p.omega_max, d.scal[SCAL_OMEGA]);
#pragma unroll
        for (int k = 0; k < D::dq; k++) {
            q[k] = q[k] + om * FLD(d, SHCFD_F_DQ, k, c);
            FLD(d, SHCFD_F_Q, k, c) = q[k];
        }
    }
    double rho = 0.0;
#pragma unroll
    for (int k = 0; k < NY; k++) rho = rho + q[k];
    const double inv = 1.0 / rho;
    const double u = q[NY] * inv, v = q[NY + 1] * inv;
    FLD(d, SHCFD_F_W, IW_RHO, c) = rho;
    FLD(d, SHCFD_F_W, IW_U, c) = u;
    FLD(d, SHCFD_F_W, IW_V, c) = v;
#pragma unroll
    for (int k = 0; k < NY; k++) FLD(d, SHCFD_F_W, IW_Y(k), c) = q[k] * inv;
    const double kap = d.p.kappa, rg = d.p.rgas, nu = B.nu;
    const double qe = q[NY + 2];
    const double ei = qe / rho - 0.5 * ((u * u) + (v * v));
    const double T = (ei * (kap - 1.0)) / rg;
    const double p = (rho * rg) * T;
    FLD(d, SHCFD_F_W, IW_P, c) = p;
    FLD(d, SHCFD_F_W, D::ig, c) = kap;
    FLD(d, SHCFD_F_W, D::ih, c) = (qe + p) / rho;
    FLD(d, SHCFD_F_W, D::imu, c) = rho * nu;
    FLD(d, SHCFD_F_W, D::ir, c) = rg;
    FLD(d, SHCFD_F_DHI, 0, c) = 0.0;
    FLD(d, SHCFD_F_VHI, 0, c) = ((kap / (kap - 1.0)) * rg) * T;
}

// ---------------------------------------------------------------------------
// boundary conditions
// ---------------------------------------------------------------------------
template <int NY>
__device__ inline void copy_wv(const DevCtx& d, long long dst, long long src) {
    using D = Dims<NY>;
#pragma unroll
    for (int k = 0; k < D::dw; k++) FLD(d, SHCFD_F_W, k, dst) = FLD(d, SHCFD_F_W, k, src);
#pragma unroll
    for (int k = 0; k < NY; k++) FLD(d, SHCFD_F_VHI, k, dst) = FLD(d, SHCFD_F_VHI, k, src);
}

// section_exchange: cond/no-nV/condition.tail.f90:2-70.  One thread per
// (cut row, inc, copy); the four copies of one inc touch disjoint cells.
template <int NY>
__global__ void k_bc_cut(DevCtx d, const int* cut, int ncut, int maxw) {
    long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int copy = (int)(gid & 3);
    long long rest = gid >> 2;
    int inc = (int)(rest % maxw);
    int r = (int)(rest / maxw);
    if (r >= ncut) return;
    const int* t = cut + SHCFD_CUT_ROW * r;
    if (inc >= t[1]) return;
    int bd, di, dj, bs, si, sj;
    if (!cut_map(t, inc, copy, bd, di, dj, bs, si, sj)) return;
    const DevBlock Bd = d.blk[bd], Bs = d.blk[bs];
    copy_wv<NY>(d, cidx(Bd, di, dj), cidx(Bs, si, sj));
}

// physical BC segments (the user set_BC loops of the samples).  One thread
// per (segment row, position along the side).
template <int NY>
__global__ void k_bc_side(DevCtx d, const int* bc, int nbc, const double* pool, int maxlen) {
    using D = Dims<NY>;
    long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int r = (int)(gid / maxlen);
    if (r >= nbc) return;
    const int* row = bc + SHCFD_BC_ROW * r;
    int b = row[0] - 1, side = row[1], s = row[2], e = row[3], kind = row[4], off = row[5];
    const DevBlock B = d.blk[b];
    int n = (side <= SHCFD_SIDE_IMAX) ? B.nj : B.ni;
    if (s < 1) s = 1;
    if (e > n) e = n;
    int tpos = s + (int)(gid % maxlen);
    if (tpos > e) return;
    int ii, jj, ig1, jg1, ig2, jg2;
    switch (side) {
    case SHCFD_SIDE_IMIN: ii = 1; jj = tpos; ig1 = 0; jg1 = tpos; ig2 = -1; jg2 = tpos; break;
    case SHCFD_SIDE_IMAX: ii = B.ni; jj = tpos; ig1 = B.ni + 1; jg1 = tpos; ig2 = B.ni + 2; jg2 = tpos; break;
    case SHCFD_SIDE_JMIN: ii = tpos; jj = 1; ig1 = tpos; jg1 = 0; ig2 = tpos; jg2 = -1; break;
    default: ii = tpos; jj = B.nj; ig1 = tpos; jg1 = B.nj + 1; ig2 = tpos; jg2 = B.nj + 2; break;
    }
    long long ci = cidx(B, ii, jj), c1 = cidx(B, ig1, jg1), c2 = cidx(B, ig2, jg2);
    const double* pv = pool + off;
    double wv[D::dw], vh[NY];
    switch (kind) {
    case SHCFD_BC_FIXED:
#pragma unroll
        for (int k = 0; k < D::dw; k++) wv[k] = pv[k];
#pragma unroll
        for (int k = 0; k < NY; k++) vh[k] = pv[D::dw + k];
        break;
    case SHCFD_BC_INFLOW_TUV: {
        const double kap = d.p.kappa, rg = d.p.rgas;
        double T = pv[0], u = pv[1], v = pv[2];
        double pin = FLD(d, SHCFD_F_W, IW_P, ci);
        double rho = pin / (rg * T);
        wv[IW_RHO] = rho; wv[IW_U] = u; wv[IW_V] = v; wv[IW_P] = pin;
#pragma unroll
        for (int k = 0; k < NY; k++) wv[IW_Y(k)] = 1.0;
        wv[D::ig] = kap;
        wv[D::ih] = ((kap / (kap - 1.0)) * rg) * T + 0.5 * ((u * u) + (v * v));
        wv[D::ir] = rg;
        wv[D::imu] = B.nu * rho;
#pragma unroll
        for (int k = 0; k < NY; k++) vh[k] = ((kap / (kap - 1.0)) * rg) * T;
        break;
    }
    case SHCFD_BC_INFLOW_MIX_P:
    case SHCFD_BC_INJECT_WALL: {
        // sample/no27.chemkin_reaction/condition.f90:53-66, :120-127
#pragma unroll
        for (int k = 0; k < D::dw; k++) wv[k] = pv[k];
#pragma unroll
        for (int k = 0; k < NY; k++) vh[k] = pv[D::dw + k];
        const double pin = FLD(d, SHCFD_F_W, IW_P, ci);
        wv[IW_RHO] = (pin * pv[D::dw + NY]) / pv[D::dw + NY + 1];
        if (kind == SHCFD_BC_INJECT_WALL) {
            const int kn = (side <= SHCFD_SIDE_IMAX) ? IW_U : IW_V, kt = (side <= SHCFD_SIDE_IMAX) ? IW_V : IW_U;
            wv[kt] = -FLD(d, SHCFD_F_W, kt, ci);
            wv[kn] = (pv[D::dw + NY + 2] * 2.0) - FLD(d, SHCFD_F_W, kn, ci);
        }
        wv[IW_P] = pin;
        wv[D::ih] = pv[D::ih] + 0.5 * ((wv[IW_U] * wv[IW_U]) + (wv[IW_V] * wv[IW_V]));
        break;
    }
    default: {
#pragma unroll
        for (int k = 0; k < D::dw; k++) wv[k] = FLD(d, SHCFD_F_W, k, ci);
#pragma unroll
        for (int k = 0; k < NY; k++) vh[k] = FLD(d, SHCFD_F_VHI, k, ci);
        if (kind == SHCFD_BC_REFLECT) {
            int kv = (side <= SHCFD_SIDE_IMAX) ? IW_U : IW_V;
            wv[kv] = -wv[kv];
        } else if (kind == SHCFD_BC_NOSLIP) {
            wv[IW_U] = -wv[IW_U];
            wv[IW_V] = -wv[IW_V];
        } else if (kind == SHCFD_BC_OUTFLOW_P) {
            wv[IW_P] = pv[0];
        } else if (kind == SHCFD_BC_SLIP_NORMAL) {
            long long cf;
            int fv;
            if (side == SHCFD_SIDE_IMIN) { cf = cidx(B, 0, jj); fv = SHCFD_F_VNI; }
            else if (side == SHCFD_SIDE_IMAX) { cf = cidx(B, B.ni, jj); fv = SHCFD_F_VNI; }
            else if (side == SHCFD_SIDE_JMIN) { cf = cidx(B, ii, 0); fv = SHCFD_F_VNJ; }
            else { cf = cidx(B, ii, B.nj); fv = SHCFD_F_VNJ; }
            double nx = FLD(d, fv, 0, cf), ny = FLD(d, fv, 1, cf);
            double u = wv[IW_U], v = wv[IW_V];
            double un = (nx * u) + (ny * v);
            wv[IW_U] = u - (2.0 * un) * nx;
            wv[IW_V] = v - (2.0 * un) * ny;
        }
        break;
    }
    }
#pragma unroll
    for (int k = 0; k < D::dw; k++) { FLD(d, SHCFD_F_W, k, c1) = wv[k]; FLD(d, SHCFD_F_W, k, c2) = wv[k]; }
#pragma unroll
    for (int k = 0; k < NY; k++) { FLD(d, SHCFD_F_VHI, k, c1) = vh[k]; FLD(d, SHCFD_F_VHI, k, c2) = vh[k]; }
}

// set_corners: cond/no-nV/condition.tail.f90:331-356
template <int NY>
__device__ inline void bc_corner_cell(const DevCtx& d, int gid) {
    using D = Dims<NY>;
    int b = gid >> 2;
    if (b >= d.nb) return;
    const DevBlock B = d.blk[b];
    int pi = (gid & 1) ? 1 : -1, pj = (gid & 2) ? 1 : -1;
    int i = (pi < 0) ? 1 : B.ni, j = (pj < 0) ? 1 : B.nj;
    long long cc = cidx(B, i + pi, j + pj), c0 = cidx(B, i, j), cj = cidx(B, i, j + pj), ci = cidx(B, i + pi, j);
#pragma unroll
    for (int k = 0; k < D::dw; k++)
        FLD(d, SHCFD_F_W, k, cc) = (1.0 / 3.0) * ((FLD(d, SHCFD_F_W, k, c0) + FLD(d, SHCFD_F_W, k, cj)) + FLD(d, SHCFD_F_W, k, ci));
#pragma unroll
    for (int k = 0; k < NY; k++)
        FLD(d, SHCFD_F_VHI, k, cc) =
            (1.0 / 3.0) * ((FLD(d, SHCFD_F_VHI, k, c0) + FLD(d, SHCFD_F_VHI, k, cj)) + FLD(d, SHCFD_F_VHI, k, ci));
}
template <int NY>
__global__ void k_bc_corner(DevCtx d) {
    bc_corner_cell<NY>(d, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// ---------------------------------------------------------------------------
// MUSCL face states: high_order/muscl/muscl.f90:7-73 (upwind: upwind.f90)
// ---------------------------------------------------------------------------
template <int NY>
__global__ void __launch_bounds__(256) k_set_hv(DevCtx d, const Tile* tiles) {
    using D = Dims<NY>;
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    const int i = t.i, j = t.j;
    const double hp = d.p.phi_mus * 0.5, kk = d.p.k_mus;
    const int lim = d.p.limiter, muscl = d.p.high_order;
    const long long c = cidx(B, i, j);
    // i-face (i,j): i in [0,ni], j in [1,nj]
    if (i <= B.ni && j >= 1 && j <= B.nj) {
        const long long cm = c - 1, cp = c + 1, cpp = c + 2;
#pragma unroll
        for (int k = 0; k < D::dw; k++) {
            double w0 = FLD(d, SHCFD_F_W, k, c), w1 = FLD(d, SHCFD_F_W, k, cp);
            double l, r;
            if (muscl) {
                double du0, dl0, du1, dl1;
                muscl_delta(w1 - w0, w0 - FLD(d, SHCFD_F_W, k, cm), kk, lim, du0, dl0);
                muscl_delta(FLD(d, SHCFD_F_W, k, cpp) - w1, w1 - w0, kk, lim, du1, dl1);
                l = w0 + hp * du0;
                r = w1 - hp * dl1;
            } else {
                l = w0;
                r = w1;
            }
            FLD(d, SHCFD_F_WHLI, k, c) = l;
            FLD(d, SHCFD_F_WHRI, k, c) = r;
        }
    }
    // j-face (i,j): i in [1,ni], j in [0,nj]
    if (i >= 1 && i <= B.ni && j <= B.nj) {
        const long long P = B.pitch;
        const long long cm = c - P, cp = c + P, cpp = c + 2 * P;
#pragma unroll
        for (int k = 0; k < D::dw; k++) {
            double w0 = FLD(d, SHCFD_F_W, k, c), w1 = FLD(d, SHCFD_F_W, k, cp);
            double l, r;
            if (muscl) {
                double du0, dl0, du1, dl1;
                muscl_delta(w1 - w0, w0 - FLD(d, SHCFD_F_W, k, cm), kk, lim, du0, dl0);
                muscl_delta(FLD(d, SHCFD_F_W, k, cpp) - w1, w1 - w0, kk, lim, du1, dl1);
                l = w0 + hp * du0;
                r = w1 - hp * dl1;
            } else {
                l = w0;
                r = w1;
            }
            FLD(d, SHCFD_F_WHLJ, k, c) = l;
            FLD(d, SHCFD_F_WHRJ, k, c) = r;
        }
    }
}

// ---------------------------------------------------------------------------
// SLAU flux: core/scheme.f90:31-259 (per face, i then j)
// ---------------------------------------------------------------------------
template <int NY>
__global__ void __launch_bounds__(256) k_set_tg(DevCtx d, const Tile* tiles) {
    using D = Dims<NY>;
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    const int i = t.i, j = t.j;
    const long long c = cidx(B, i, j);
    double wl[D::dw], wr[D::dw], tg[D::dq];
    if (i <= B.ni && j >= 1 && j <= B.nj) {
#pragma unroll
        for (int k = 0; k < D::dw; k++) { wl[k] = FLD(d, SHCFD_F_WHLI, k, c); wr[k] = FLD(d, SHCFD_F_WHRI, k, c); }
        slau_face<NY>(wl, wr, FLD(d, SHCFD_F_VNI, 0, c), FLD(d, SHCFD_F_VNI, 1, c), tg);
#pragma unroll
        for (int k = 0; k < D::dq; k++) FLD(d, SHCFD_F_TGI, k, c) = tg[k];
    }
    if (i >= 1 && i <= B.ni && j <= B.nj) {
#pragma unroll
        for (int k = 0; k < D::dw; k++) { wl[k] = FLD(d, SHCFD_F_WHLJ, k, c); wr[k] = FLD(d, SHCFD_F_WHRJ, k, c); }
        slau_face<NY>(wl, wr, FLD(d, SHCFD_F_VNJ, 0, c), FLD(d, SHCFD_F_VNJ, 1, c), tg);
#pragma unroll
        for (int k = 0; k < D::dq; k++) FLD(d, SHCFD_F_TGJ, k, c) = tg[k];
    }
}

// ---------------------------------------------------------------------------
// set_HV followed by set_TG inside a whole step: the face states of one face
// stay in registers and go straight into the SLAU flux (the same expressions
// as k_set_hv and k_set_tg, so TGi/TGj are identical); wHl/wHr are not
// stored.  The per-routine entry points keep the two separate kernels.
// ---------------------------------------------------------------------------
// left/right states of the face between cell c and c + s (s = 1: i-face,
// s = pitch: j-face), as k_set_hv forms them
template <int NY>
__device__ inline void face_states(const DevCtx& d, long long c, long long s, double hp, double kk, int lim,
                                   int muscl, double* wl, double* wr) {
#pragma unroll
    for (int k = 0; k < NY + 8; k++) {
        double w0 = FLD(d, SHCFD_F_W, k, c), w1 = FLD(d, SHCFD_F_W, k, c + s);
        if (muscl) {
            double du0, dl0, du1, dl1;
            muscl_delta(w1 - w0, w0 - FLD(d, SHCFD_F_W, k, c - s), kk, lim, du0, dl0);
            muscl_delta(FLD(d, SHCFD_F_W, k, c + 2 * s) - w1, w1 - w0, kk, lim, du1, dl1);
            wl[k] = w0 + hp * du0;
            wr[k] = w1 - hp * dl1;
        } else {
            wl[k] = w0;
            wr[k] = w1;
        }
    }
}

template <int NY>
__global__ void __launch_bounds__(256) k_set_hv_tg(DevCtx d, const Tile* tiles) {
    using D = Dims<NY>;
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    const int i = t.i, j = t.j;
    const double hp = d.p.phi_mus * 0.5, kk = d.p.k_mus;
    const int lim = d.p.limiter, muscl = d.p.high_order;
    const long long c = cidx(B, i, j);
    double wl[D::dw], wr[D::dw], tg[D::dq];
    {
        // set_corners (cond/no-nV/condition.tail.f90:331-356): the corner ghosts are
        // read by no face of this kernel, only by later ones
        const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
        if (g < 4 * d.nb) bc_corner_cell<NY>(d, g);
    }
    if (i <= B.ni && j >= 1 && j <= B.nj) {
        face_states<NY>(d, c, 1, hp, kk, lim, muscl, wl, wr);
        slau_face<NY>(wl, wr, FLD(d, SHCFD_F_VNI, 0, c), FLD(d, SHCFD_F_VNI, 1, c), tg);
#pragma unroll
        for (int k = 0; k < D::dq; k++) FLD(d, SHCFD_F_TGI, k, c) = tg[k];
    }
    if (i >= 1 && i <= B.ni && j <= B.nj) {
        face_states<NY>(d, c, B.pitch, hp, kk, lim, muscl, wl, wr);
        slau_face<NY>(wl, wr, FLD(d, SHCFD_F_VNJ, 0, c), FLD(d, SHCFD_F_VNJ, 1, c), tg);
#pragma unroll
        for (int k = 0; k < D::dq; k++) FLD(d, SHCFD_F_TGJ, k, c) = tg[k];
    }
}

// ---------------------------------------------------------------------------
// viscous flux: viscosity/viscous/no-nV/2d/sch_viscous.f90:56-272
// ---------------------------------------------------------------------------
template <int NY>
__global__ void __launch_bounds__(256) k_set_tgv(DevCtx d, const Tile* tiles) {
    using D = Dims<NY>;
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    const int i = t.i, j = t.j;
    const long long c = cidx(B, i, j);
    double out[D::dq];
    if (i <= B.ni && j >= 1 && j <= B.nj) {
        viscous_face<NY, 0>(d, B, i, j, out);
#pragma unroll
        for (int k = 0; k < D::dq; k++) FLD(d, SHCFD_F_TGVI, k, c) = out[k];
    }
    if (i >= 1 && i <= B.ni && j <= B.nj) {
        viscous_face<NY, 1>(d, B, i, j, out);
#pragma unroll
        for (int k = 0; k < D::dq; k++) FLD(d, SHCFD_F_TGVJ, k, c) = out[k];
    }
}

#if SHCFD_INST_GROUP == 0
__global__ void k_zero_field(double* f, long long n) {
    long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; gid < n; gid += stride) f[gid] = 0.0;
}
#endif

// ---------------------------------------------------------------------------
// set_dt: time_step/{local,global}/set_dt.f90 -- dt_mat per cell and a
// per-tile min; the global min is taken by k_reduce_final.
// ---------------------------------------------------------------------------
__device__ inline double block_reduce_min(double v, double* sh) {
    // fixed-shape tree over 256 threads
    for (int off = 32; off >= 1; off >>= 1) v = dmin(v, __shfl_xor(v, off));
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = sh[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) r = dmin(r, sh[k]);
    return r;
}
__device__ inline double block_reduce_sum(double v, double* sh) {
    for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = sh[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) r = r + sh[k];
    return r;
}

template <int NY>
__global__ void __launch_bounds__(256) k_set_dt(DevCtx d, const Tile* tiles, double* part) {
    using D = Dims<NY>;
    __shared__ double sh[8];
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    const int i = t.i, j = t.j;
    double dt = 1e300;
    if (i >= 1 && i <= B.ni && j >= 1 && j <= B.nj) {
        const long long c = cidx(B, i, j), P = B.pitch;
        double rho = FLD(d, SHCFD_F_W, IW_RHO, c), u = FLD(d, SHCFD_F_W, IW_U, c), v = FLD(d, SHCFD_F_W, IW_V, c);
        double p = FLD(d, SHCFD_F_W, IW_P, c);
        double a = sqrt((FLD(d, SHCFD_F_W, D::ig, c) * p) / rho);
        double uds = 0.0, un;
        un = (FLD(d, SHCFD_F_VNI, 0, c) * u) + (FLD(d, SHCFD_F_VNI, 1, c) * v);
        uds = uds + (fabs(un) + a) * FLD(d, SHCFD_F_DSI, 0, c);
        un = (FLD(d, SHCFD_F_VNI, 0, c - 1) * u) + (FLD(d, SHCFD_F_VNI, 1, c - 1) * v);
        uds = uds + (fabs(un) + a) * FLD(d, SHCFD_F_DSI, 0, c - 1);
        un = (FLD(d, SHCFD_F_VNJ, 0, c) * u) + (FLD(d, SHCFD_F_VNJ, 1, c) * v);
        uds = uds + (fabs(un) + a) * FLD(d, SHCFD_F_DSJ, 0, c);
        un = (FLD(d, SHCFD_F_VNJ, 0, c - P) * u) + (FLD(d, SHCFD_F_VNJ, 1, c - P) * v);
        uds = uds + (fabs(un) + a) * FLD(d, SHCFD_F_DSJ, 0, c - P);
        double nr = FLD(d, SHCFD_F_W, D::imu, c) / rho;
        dt = (d.p.cfl * FLD(d, SHCFD_F_VOL, 0, c)) / dmax(uds, nr);
        FLD(d, SHCFD_F_DT, 0, c) = dt;
    }
    double m = block_reduce_min(dt, sh);
    if (threadIdx.x == 0) part[tiles[blockIdx.x].slot] = m;
}

// final fixed-order reduction of per-tile partials (one workgroup)
// mode 0: min -> scal[slot] (and tt += for dt); mode 1: sum -> scal[slot]
#if SHCFD_INST_GROUP == 0
__global__ void __launch_bounds__(256) k_reduce_final(DevCtx d, const double* part, int n, int mode, int slot) {
    __shared__ double sh[8];
    double v = (mode == 0) ? 1e300 : 0.0;
    // contiguous chunks per thread, ascending order within a chunk
    int chunk = (n + blockDim.x - 1) / blockDim.x;
    int lo = threadIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    for (int k = lo; k < hi; k++) v = (mode == 0) ? dmin(v, part[k]) : v + part[k];
    double r = (mode == 0) ? block_reduce_min(v, sh) : block_reduce_sum(v, sh);
    if (threadIdx.x == 0) {
        d.scal[slot] = r;
        if (mode == 0 && slot == SCAL_DT_GRBL) d.scal[SCAL_TT] = d.scal[SCAL_TT] + r;
    }
}
#endif

// ---------------------------------------------------------------------------
// energy RMS: core/check_convergence.f90:17-28 (per-tile partial sums)
// ---------------------------------------------------------------------------
template <int NY>
__global__ void __launch_bounds__(256) k_rms_part(DevCtx d, const Tile* tiles, double* part) {
    __shared__ double sh[8];
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    double v = 0.0;
    if (t.i >= 1 && t.i <= B.ni && t.j >= 1 && t.j <= B.nj) {
        long long c = cidx(B, t.i, t.j);
        double e = FLD(d, SHCFD_F_Q, NY + 2, c) - FLD(d, SHCFD_F_QP, NY + 2, c);
        v = e * e;
    }
    double s = block_reduce_sum(v, sh);
    if (threadIdx.x == 0) part[tiles[blockIdx.x].slot] = s;
}

// ---------------------------------------------------------------------------
// set_ABpm (LU-SGS, preconLU-SGS, dual; set_abpm_cell) and set_dqdt
// ---------------------------------------------------------------------------
template <int NY, int MATS>
__device__ inline void set_abpm_body(const DevCtx& d, const Tile* tiles) {
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    if (t.i < 1 || t.i > B.ni || t.j < 1 || t.j > B.nj) return;
    set_abpm_cell<NY, MATS>(d, B, t.i, t.j, nullptr, 0, 0);
}

template <int NY>
__global__ void __launch_bounds__(256) k_set_abpm(DevCtx d, const Tile* tiles) { set_abpm_body<NY, 1>(d, tiles); }
template <int NY>
__global__ void __launch_bounds__(256) k_set_abpm_scal(DevCtx d, const Tile* tiles) { set_abpm_body<NY, 0>(d, tiles); }

// set_ABpm writing the split Jacobians in the strip-skewed layout of the
// row-per-lane sweeps: strip s, sweep step t, lane L (row j0+L) holds cell
// (i = t-L+1, j0+L) at slot sbase[s] + 64 t + L, so every step of both sweeps
// reads 64 consecutive slots per matrix element.  One wavefront per step t
// (4 per workgroup); grid = (strips, ceil((max ni + 63)/4)).
template <int NY>
__global__ void __launch_bounds__(256) k_set_abpm_skew(DevCtx d, const Strip* strips, const long long* sbase,
                                                       double* skew, long long nsl) {
    const int s = blockIdx.x;
    const int t = blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    const int L = (int)(threadIdx.x & 63);
    const Strip S = strips[s];
    const DevBlock B = d.blk[S.b];
    if (t >= B.ni + 63) return;
    const int i = t - L + 1, j = S.j0 + L;
    if (i < 1 || i > B.ni || j > B.nj) return;
    set_abpm_cell<NY, 2>(d, B, i, j, skew, nsl, sbase[s] + 64LL * t + L);
}

// set_ABpm writing the quad-row copy of the 16-row sweeps: thread = row r of
// step t of a strip (16 steps x 16 rows per workgroup); grid = (strips,
// ceil((max ni + 15)/16))
template <int NY>
__global__ void __launch_bounds__(256) k_set_abpm16(DevCtx d, const Strip16* strips, double* jac, long long nsl) {
    const Strip16 S = strips[blockIdx.x];
    const int t = (int)blockIdx.y * 16 + (int)(threadIdx.x >> 4), r = (int)(threadIdx.x & 15);
    const DevBlock B = d.blk[S.b];
    if (t >= B.ni + 15) return;
    const int i = t - r + 1, j = S.j0 + r;
    if (i < 1 || i > B.ni || j > B.nj) return;
    set_abpm_cell<NY, 3>(d, B, i, j, jac, nsl, 64LL * (S.base + t) + 4 * r);
}

template <int NY>
__global__ void __launch_bounds__(256) k_set_dqdt(DevCtx d, const Tile* tiles) {
    using D = Dims<NY>;
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    if (t.i < 1 || t.i > B.ni || t.j < 1 || t.j > B.nj) return;
    const long long c = cidx(B, t.i, t.j);
    const double dtv = (d.p.time_step == 0) ? d.scal[SCAL_DT_GRBL] : FLD(d, SHCFD_F_DT, 0, c);
    const bool first = d.scal[SCAL_FIRST_DUAL] != 0.0;
#pragma unroll
    for (int k = 0; k < D::dq; k++) {
        double q = FLD(d, SHCFD_F_Q, k, c), qp = FLD(d, SHCFD_F_QP, k, c);
        FLD(d, SHCFD_F_DQDT, k, c) = first ? (q - qp) / dtv : (((3.0 * q) - (4.0 * qp)) + FLD(d, SHCFD_F_QPP, k, c)) / (2.0 * dtv);
    }
}

// ---------------------------------------------------------------------------
// explicit schemes: euler/time_euler_viscous.raw.f90, RK2/time_RK2_viscous.raw.f90
// stage 0: Euler; 1: RK2 first; 2: RK2 second
// ---------------------------------------------------------------------------
template <int NY>
__global__ void __launch_bounds__(256) k_explicit(DevCtx d, const Tile* tiles, int stage) {
    using D = Dims<NY>;
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    // a frozen block (active = 0) keeps its state
    if (t.i < 1 || t.i > B.ni || t.j < 1 || t.j > B.nj || !B.active) return;
    const long long c = cidx(B, t.i, t.j);
    double s[D::dq];
    flux_balance<NY>(d, B, c, s);
    double dtv = (d.p.time_step == 0) ? d.scal[SCAL_DT_GRBL] : FLD(d, SHCFD_F_DT, 0, c);
    double vol = FLD(d, SHCFD_F_VOL, 0, c);
    double f = (stage == 1) ? (0.5 * dtv) / vol : dtv / vol;
#pragma unroll
    for (int k = 0; k < D::dq; k++) {
        double qn = FLD(d, SHCFD_F_QP, k, c) + f * s[k];
        if (stage == 2 && k < NY) qn = (qn > 1e-40) ? qn : 1e-40;
        FLD(d, SHCFD_F_Q, k, c) = qn;
    }
}

template <int NY>
__global__ void __launch_bounds__(256) k_copy_q(DevCtx d, const Tile* tiles, int dst, int src) {
    using D = Dims<NY>;
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    if (t.i < 1 || t.i > B.ni || t.j < 1 || t.j > B.nj) return;
    const long long c = cidx(B, t.i, t.j);
#pragma unroll
    for (int k = 0; k < D::dq; k++) FLD(d, dst, k, c) = FLD(d, src, k, c);
}

// ---------------------------------------------------------------------------
// instantiation
// ---------------------------------------------------------------------------
#define INST(NY)                                                                                         \
    template __global__ void k_set_w<NY>(DevCtx, const Tile*);                                           \
    template __global__ void k_thermo_ideal<NY>(DevCtx, const Tile*);                                    \
    template __global__ void k_primitive_ideal<NY>(DevCtx, const Tile*, int);                                 \
    template __global__ void k_bc_cut<NY>(DevCtx, const int*, int, int);                                 \
    template __global__ void k_bc_side<NY>(DevCtx, const int*, int, const double*, int);                 \
    template __global__ void k_bc_corner<NY>(DevCtx);                                                    \
    template __global__ void k_set_hv<NY>(DevCtx, const Tile*);                                          \
    template __global__ void k_set_tg<NY>(DevCtx, const Tile*);                                          \
    template __global__ void k_set_hv_tg<NY>(DevCtx, const Tile*);                                       \
    template __global__ void k_set_tgv<NY>(DevCtx, const Tile*);                                         \
    template __global__ void k_set_dt<NY>(DevCtx, const Tile*, double*);                                 \
    template __global__ void k_rms_part<NY>(DevCtx, const Tile*, double*);                               \
    template __global__ void k_set_abpm<NY>(DevCtx, const Tile*);                                        \
    template __global__ void k_set_abpm_scal<NY>(DevCtx, const Tile*);                                   \
    template __global__ void k_set_abpm_skew<NY>(DevCtx, const Strip*, const long long*, double*, long long); \
    template __global__ void k_set_abpm16<NY>(DevCtx, const Strip16*, double*, long long);                 \
    template __global__ void k_set_dqdt<NY>(DevCtx, const Tile*);                                        \
    template __global__ void k_explicit<NY>(DevCtx, const Tile*, int);                                   \
    template __global__ void k_copy_q<NY>(DevCtx, const Tile*, int, int);
SHCFD_FOR_EACH_NY(INST)
#undef INST

}  // namespace shcfd
