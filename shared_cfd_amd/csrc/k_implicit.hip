// k_implicit.hip -- right-hand side, LU-SGS hyperplane sweeps and update of
// calc_next_step_implicit for LU-SGS (LU-SGS/2d/time_lusgs_viscous.f90),
// preconditioned LU-SGS (preconLU-SGS/2d/time_precon_viscous.f90), the
// dual-time inner loops with the plain and the preconditioned operator
// (NR/2d/time_NR_viscous.f90, precon/2d/time_precon_viscous.f90) and the
// dual-time three-stage form (dual/2d/time_dual_viscous.f90).
//
// Sweep schedule ("strip wavefront").  The reference walks hyperplanes
// sm = i+j in order; a cell depends only on (i-1,j) and (i,j-1) (forward)
// or (i+1,j) and (i,j+1) (backward).  Here one 64-lane wavefront owns a
// strip of 64 rows of one block, lane l = row j0+l, and at step t lane l
// processes column i = t-l+1: its (i-1,j) input is its own previous result
// and its (i,j-1) input arrives from lane l-1 through a DPP wave shift.  The
// strip below hands its top row to lane 0 through an agent-coherent buffer
// guarded by a per-strip progress counter.  Every cell therefore evaluates
// the reference's per-cell formula on the same inputs, in an order that is a
// topological order of the same dependency graph, so the result equals the
// hyperplane order exactly.  Δq outside a block is zero (the reference's
// ghost zeroing, time_lusgs_viscous.f90:25-26, :41-42).
#include "shcfd_device.h"
#include "shcfd_kernels.h"

namespace shcfd {

// ---------------------------------------------------------------------------
// RHS: time_lusgs_viscous.f90:12-22, time_precon_viscous.f90:12-21
// (preconLU-SGS), time_NR_viscous.f90:16-25, time_precon_viscous.f90:16-27
// (precon), time_dual_viscous.f90:16-28.  Tile 0 also clears the sweep
// progress flags.  With dqdt_inline set (whole steps of the inner-loop
// schemes) the cell's dq/dt of set_dqdt (sch_dual.raw.f90:211-240) is formed
// here, with set_dqdt's expressions, stored and used, instead of in a pass
// of its own.  With abpm_inline set (the set_ABpm pass was left out, as for
// the optimised sweeps, whose prep pass k_fast_rhs_prep forms the scalars and
// this right-hand side itself) the cell's set_ABpm scalars are formed and
// stored here first, by the same code as the set_ABpm pass.
// ---------------------------------------------------------------------------
template <int NY>
__global__ void __launch_bounds__(256) k_rhs(DevCtx d, const Tile* tiles, int* prog, int nprog, int dqdt_inline,
                                             int abpm_inline) {
    using D = Dims<NY>;
    constexpr int dq = D::dq;
    if (blockIdx.x == 0)
        for (int k = threadIdx.x; k < nprog; k += blockDim.x) prog[k] = 0;
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    if (t.i < 1 || t.i > B.ni || t.j < 1 || t.j > B.nj) return;
    if (abpm_inline) set_abpm_cell<NY, 0>(d, B, t.i, t.j, nullptr, 0, 0);
    const long long c = cidx(B, t.i, t.j);
    double s[dq], x[dq];
    flux_balance<NY>(d, B, c, s);
    rhs_of_cell<NY>(
        d, c, s, dqdt_inline != 0, [&](int k) { return FLD(d, SHCFD_F_PRE1, k, c); },
        [&](int k) { return FLD(d, SHCFD_F_PRE2, k, c); }, [&]() { return FLD(d, SHCFD_F_PRESCA, 1, c); }, x);
#pragma unroll
    for (int k = 0; k < dq; k++) FLD(d, SHCFD_F_DQ, k, c) = x[k];
}

template <int DQN>
__device__ inline void matvec_scaled(const DevCtx& d, int F, long long c, double scale, const double* x, double* y) {
    // y = scale * (M x), M column-major dq x dq stored as components of F
#pragma unroll
    for (int k = 0; k < DQN; k++) {
        double s = 0.0;
#pragma unroll
        for (int l = 0; l < DQN; l++) s = s + FLD(d, F, k + DQN * l, c) * x[l];
        y[k] = scale * s;
    }
}

// the same product from the strip-skewed Jacobian copy (k_set_abpm_skew):
// element e of matrix m at skew[(m dq^2 + e) nsl + slot]
template <int DQN>
__device__ inline void matvec_skew(const double* __restrict__ skew, long long nsl, int m, long long slot, double scale,
                                   const double* x, double* y) {
    const double* base = skew + (long long)m * DQN * DQN * nsl + slot;
#pragma unroll
    for (int k = 0; k < DQN; k++) {
        double s = 0.0;
#pragma unroll
        for (int l = 0; l < DQN; l++) s = s + base[(long long)(k + DQN * l) * nsl] * x[l];
        y[k] = scale * s;
    }
}

__device__ inline void set_error(const DevCtx& d, int code) {
    atomicCAS(d.iflag + IFLAG_ERR, 0, code);
}

// wait until the progress counter of strip `other` reaches `need`
// (wave-uniform); returns false on timeout
__device__ inline bool wait_progress(const int* prog, int other, int need, int& known, long long spin_limit) {
    if (known >= need) return true;
    int v = load_flag(prog + other);
    long long it = 0;
    while (v < need && it < spin_limit) {
        __builtin_amdgcn_s_sleep(4);
        v = load_flag(prog + other);
        it++;
    }
    known = v;
    return v >= need;
}

// ---------------------------------------------------------------------------
// forward sweep (plain: Jacobians read from the stored Ap/Bp arrays)
// ---------------------------------------------------------------------------
template <int NY, bool SKEW>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) k_sweep_fwd_plain(DevCtx d, const Strip* strips, int* prog, double* bnd,
                                                        long long bstride, long long spin_limit, const double* skew,
                                                        long long nsl, const long long* sbase) {
    constexpr int dq = NY + 3;
    const int s = blockIdx.x;
    const int lane = threadIdx.x;
    const Strip S = strips[s];
    const DevBlock B = d.blk[S.b];
    const int ni = B.ni, nj = B.nj;
    const int j = S.j0 + lane;
    const bool rowok = j <= nj;
    const bool dual = d.p.scheme == SHCFD_SCHEME_DUAL;
    const bool ax = d.p.dimension == SHCFD_DIM_AXISYMMETRIC;
    const bool has_prev = S.prev >= 0, has_next = S.next >= 0;
    double a_in[dq], bo[dq];
#pragma unroll
    for (int k = 0; k < dq; k++) { a_in[k] = 0.0; bo[k] = 0.0; }
    int known = 0;
    const int nsteps = ni + 63;
    for (int t = 0; t < nsteps; t++) {
        const int i = t - lane + 1;
        const bool act = rowok && i >= 1 && i <= ni;
        double b_in[dq];
#pragma unroll
        for (int k = 0; k < dq; k++) b_in[k] = wave_shr1(bo[k]);
        const int i0 = t + 1;  // column of lane 0 at this step
        if (has_prev && i0 <= ni) {
            if (!wait_progress(prog, S.prev, i0, known, spin_limit)) { set_error(d, SHCFD_E_SPIN); return; }
            if (lane == 0) {
                const double* src = bnd + (long long)S.prev * bstride + (long long)i0 * dq;
#pragma unroll
                for (int k = 0; k < dq; k++) b_in[k] = load_agent(src + k);
            }
        } else if (lane == 0) {
#pragma unroll
            for (int k = 0; k < dq; k++) b_in[k] = 0.0;
        }
        if (act) {
            const long long c = cidx(B, i, j);
            const double al = FLD(d, SHCFD_F_ALPHA, 0, c);
            double x[dq];
#pragma unroll
            for (int k = 0; k < dq; k++) x[k] = al * (FLD(d, SHCFD_F_DQ, k, c) + (a_in[k] + b_in[k]));
            if (dual) {
                double dd = 0.0;
#pragma unroll
                for (int k = 0; k < dq; k++) dd = dd + FLD(d, SHCFD_F_PRE2, k, c) * x[k];
                const double ps4 = FLD(d, SHCFD_F_PRESCA, 3, c), ips1 = 1.0 / FLD(d, SHCFD_F_PRESCA, 0, c);
#pragma unroll
                for (int k = 0; k < dq; k++) x[k] = x[k] + (ps4 * dd) * FLD(d, SHCFD_F_PRE1, k, c);
#pragma unroll
                for (int k = 0; k < dq; k++) x[k] = ips1 * x[k];
            }
            if (ax) {
                // implicit pressure source on the radial momentum:
                // LU-SGS/q2d/time_lusgs_viscous.f90:38-44, NR/q2d/time_NR_viscous.f90:43-49
                double tmp = 0.0;
#pragma unroll
                for (int l = 0; l <= NY; l++) tmp = tmp + (x[l] * FLD(d, SHCFD_F_PRE2, l, c));
                tmp = tmp + (x[NY + 2] * FLD(d, SHCFD_F_PRE2, NY + 2, c));
                const double aa = al * FLD(d, SHCFD_F_AREA, 0, c);
                tmp = x[NY + 1] + (aa * tmp);
                x[NY + 1] = tmp / (1.0 - (aa * FLD(d, SHCFD_F_PRE2, NY + 1, c)));
            }
#pragma unroll
            for (int k = 0; k < dq; k++) FLD(d, SHCFD_F_DQ, k, c) = x[k];
            if (SKEW) {
                const long long slot = sbase[s] + 64LL * t + lane;
                matvec_skew<dq>(skew, nsl, SKEW_AP, slot, FLD(d, SHCFD_F_DSCI, 0, c), x, a_in);
                matvec_skew<dq>(skew, nsl, SKEW_BP, slot, FLD(d, SHCFD_F_DSCJ, 0, c), x, bo);
            } else {
                matvec_scaled<dq>(d, SHCFD_F_AP, c, FLD(d, SHCFD_F_DSCI, 0, c), x, a_in);
                matvec_scaled<dq>(d, SHCFD_F_BP, c, FLD(d, SHCFD_F_DSCJ, 0, c), x, bo);
            }
        } else {
#pragma unroll
            for (int k = 0; k < dq; k++) bo[k] = 0.0;
        }
        if (has_next) {
            const int i63 = t - 62;  // column of lane 63
            if (lane == 63 && act) {
                double* dst = bnd + (long long)s * bstride + (long long)i * dq;
#pragma unroll
                for (int k = 0; k < dq; k++) store_agent(dst + k, bo[k]);
            }
            if (i63 >= 1 && ((i63 & 7) == 0 || i63 == ni)) {
                wait_vmem();
                if (lane == 63) store_flag(prog + s, i63);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// backward sweep (plain).  Workgroup g handles strip nstrips-1-g so that the
// strip it waits on (the one above) has a lower workgroup id.  For LU-SGS and
// preconditioned LU-SGS the update q += Δq is applied here; for dual time the
// per-strip (res, omega) partials are written and k_update applies q += ωΔq
// (also for the NR and precon inner loops, without the S-scaling stages).
// ---------------------------------------------------------------------------
template <int NY, bool SKEW>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) k_sweep_bwd_plain(DevCtx d, const Strip* strips, int nstrips, int* prog,
                                                        double* bnd, long long bstride, double* part,
                                                        long long spin_limit, const double* skew, long long nsl,
                                                        const long long* sbase) {
    constexpr int dq = NY + 3;
    const int s = nstrips - 1 - (int)blockIdx.x;
    const int lane = threadIdx.x;
    const Strip S = strips[s];
    const DevBlock B = d.blk[S.b];
    const int ni = B.ni, nj = B.nj;
    const int j = S.j0 + lane;
    const bool rowok = j <= nj;
    const bool dual = d.p.scheme == SHCFD_SCHEME_DUAL;
    const bool inner = inner_loop_scheme(d.p.scheme);
    const bool has_prev = S.prev >= 0, has_next = S.next >= 0;
    const double dqmax = d.p.dqmax;
    double a_in[dq], bo[dq];
#pragma unroll
    for (int k = 0; k < dq; k++) { a_in[k] = 0.0; bo[k] = 0.0; }
    double res = 0.0, omega = 1e300;
    int known = 0;
    const int nsteps = ni + 63;
    for (int t = 0; t < nsteps; t++) {
        const int i = ni - t + 63 - lane;
        const bool act = rowok && i >= 1 && i <= ni;
        double b_in[dq];
#pragma unroll
        for (int k = 0; k < dq; k++) b_in[k] = wave_shl1(bo[k]);
        const int i63 = ni - t;  // column of lane 63
        if (has_next && i63 >= 1) {
            if (!wait_progress(prog, S.next, t + 1, known, spin_limit)) { set_error(d, SHCFD_E_SPIN); return; }
            if (lane == 63) {
                const double* src = bnd + (long long)S.next * bstride + (long long)i63 * dq;
#pragma unroll
                for (int k = 0; k < dq; k++) b_in[k] = load_agent(src + k);
            }
        } else if (lane == 63) {
#pragma unroll
            for (int k = 0; k < dq; k++) b_in[k] = 0.0;
        }
        if (act) {
            const long long c = cidx(B, i, j);
            const double al = FLD(d, SHCFD_F_ALPHA, 0, c);
            double x[dq];
#pragma unroll
            for (int k = 0; k < dq; k++) x[k] = FLD(d, SHCFD_F_DQ, k, c);
            if (dual) {
                // second stage, time_dual_viscous.f90:53-60
                double dd = 0.0;
#pragma unroll
                for (int k = 0; k < dq; k++) dd = dd + FLD(d, SHCFD_F_PRE2, k, c) * x[k];
                const double ps3 = FLD(d, SHCFD_F_PRESCA, 2, c), ps1 = FLD(d, SHCFD_F_PRESCA, 0, c);
#pragma unroll
                for (int k = 0; k < dq; k++) x[k] = x[k] + (ps3 * dd) * FLD(d, SHCFD_F_PRE1, k, c);
#pragma unroll
                for (int k = 0; k < dq; k++) x[k] = ps1 * x[k];
            }
#pragma unroll
            for (int k = 0; k < dq; k++) x[k] = x[k] - al * (a_in[k] + b_in[k]);
            if (dual) {
                double dd = 0.0;
#pragma unroll
                for (int k = 0; k < dq; k++) dd = dd + FLD(d, SHCFD_F_PRE2, k, c) * x[k];
                const double ps4 = FLD(d, SHCFD_F_PRESCA, 3, c), ips1 = 1.0 / FLD(d, SHCFD_F_PRESCA, 0, c);
#pragma unroll
                for (int k = 0; k < dq; k++) x[k] = x[k] + (ps4 * dd) * FLD(d, SHCFD_F_PRE1, k, c);
#pragma unroll
                for (int k = 0; k < dq; k++) x[k] = ips1 * x[k];
            }
            if (inner) {
                res = res + x[NY + 2] * x[NY + 2];
#pragma unroll
                for (int k = 0; k < NY; k++)
                    if (x[k] < 0.0) omega = dmin(omega, (-FLD(d, SHCFD_F_Q, k, c) / x[k]) * dqmax);
            } else {
                if (B.active)
#pragma unroll
                    for (int k = 0; k < dq; k++) FLD(d, SHCFD_F_Q, k, c) = FLD(d, SHCFD_F_Q, k, c) + x[k];
            }
#pragma unroll
            for (int k = 0; k < dq; k++) FLD(d, SHCFD_F_DQ, k, c) = x[k];
            if (SKEW) {
                // cell (i, j0+lane) sits at forward step i + lane - 1 = ni + 62 - t
                const long long slot = sbase[s] + 64LL * (ni + 62 - t) + lane;
                matvec_skew<dq>(skew, nsl, SKEW_AM, slot, FLD(d, SHCFD_F_DSCI, 0, c), x, a_in);
                matvec_skew<dq>(skew, nsl, SKEW_BM, slot, FLD(d, SHCFD_F_DSCJ, 0, c), x, bo);
            } else {
                matvec_scaled<dq>(d, SHCFD_F_AM, c, FLD(d, SHCFD_F_DSCI, 0, c), x, a_in);
                matvec_scaled<dq>(d, SHCFD_F_BM, c, FLD(d, SHCFD_F_DSCJ, 0, c), x, bo);
            }
        } else {
#pragma unroll
            for (int k = 0; k < dq; k++) bo[k] = 0.0;
        }
        if (has_prev) {
            const int il0 = ni - t + 63;  // column of lane 0
            if (lane == 0 && act) {
                double* dst = bnd + (long long)s * bstride + (long long)i * dq;
#pragma unroll
                for (int k = 0; k < dq; k++) store_agent(dst + k, bo[k]);
            }
            const int done = ni - il0 + 1;  // columns finished by lane 0
            if (il0 <= ni && il0 >= 1 && ((done & 7) == 0 || il0 == 1)) {
                wait_vmem();
                if (lane == 0) store_flag(prog + s, done);
            }
        }
    }
    if (inner) {
        // fixed-order tree over the 64 rows of the strip
        for (int off = 32; off >= 1; off >>= 1) {
            res = res + __shfl_xor(res, off);
            omega = dmin(omega, __shfl_xor(omega, off));
        }
        if (lane == 0) {
            part[2 * s] = res;
            part[2 * s + 1] = omega;
        }
    }
}

// final (res, omega) reduction over strips; one workgroup; sets error flags
#if SHCFD_INST_GROUP == 0
__global__ void __launch_bounds__(256) k_reduce_res_omega(DevCtx d, const double* part, int n) {
    __shared__ double shr[4], sho[4];
    int chunk = (n + blockDim.x - 1) / blockDim.x;
    int lo = threadIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    double r = 0.0, o = 1e300;
    for (int k = lo; k < hi; k++) { r = r + part[2 * k]; o = dmin(o, part[2 * k + 1]); }
    for (int off = 32; off >= 1; off >>= 1) {
        r = r + __shfl_xor(r, off);
        o = dmin(o, __shfl_xor(o, off));
    }
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { shr[w] = r; sho[w] = o; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double R = shr[0], O = sho[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) { R = R + shr[k]; O = dmin(O, sho[k]); }
        d.scal[SCAL_RES] = R;
        d.scal[SCAL_OMEGA] = O;
        if (R != R && d.p.scheme == SHCFD_SCHEME_DUAL) set_error(d, SHCFD_E_NAN);  // time_dual_viscous.f90:104-107
        if (O < d.p.omega_min) set_error(d, SHCFD_E_OMEGA);
    }
}
#endif

#if SHCFD_INST_GROUP == 0
__global__ void k_copy_scal(DevCtx d, int dst, int src) { d.scal[dst] = d.scal[src]; }
#endif

// q += min(omega_max, omega) * Δq : time_dual_viscous.f90:113-124, time_NR_viscous.f90:90-101
template <int NY>
__global__ void __launch_bounds__(256) k_update(DevCtx d, const Tile* tiles) {
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    if (t.i < 1 || t.i > B.ni || t.j < 1 || t.j > B.nj || !B.active) return;
    const long long c = cidx(B, t.i, t.j);
    const double om = dmin(d.p.omega_max, d.scal[SCAL_OMEGA]);
#pragma unroll
    for (int k = 0; k < NY + 3; k++) FLD(d, SHCFD_F_Q, k, c) = FLD(d, SHCFD_F_Q, k, c) + om * FLD(d, SHCFD_F_DQ, k, c);
}

// q += dq over the active blocks' cells (LU-SGS after the Jacobian-free
// backward sweep: the same addition time_lusgs_viscous.f90 makes per cell)
template <int NY>
__global__ void __launch_bounds__(256) k_add_dq(DevCtx d, const Tile* tiles) {
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    if (t.i < 1 || t.i > B.ni || t.j < 1 || t.j > B.nj || !B.active) return;
    const long long c = cidx(B, t.i, t.j);
#pragma unroll
    for (int k = 0; k < NY + 3; k++) FLD(d, SHCFD_F_Q, k, c) = FLD(d, SHCFD_F_Q, k, c) + FLD(d, SHCFD_F_DQ, k, c);
}

#define INST(NY)                                                                                               \
    template __global__ void k_add_dq<NY>(DevCtx, const Tile*);                                                \
    template __global__ void k_rhs<NY>(DevCtx, const Tile*, int*, int, int, int);                                        \
    template __global__ void k_sweep_fwd_plain<NY, false>(DevCtx, const Strip*, int*, double*, long long, long long, \
                                                          const double*, long long, const long long*);          \
    template __global__ void k_sweep_fwd_plain<NY, true>(DevCtx, const Strip*, int*, double*, long long, long long,  \
                                                         const double*, long long, const long long*);           \
    template __global__ void k_sweep_bwd_plain<NY, false>(DevCtx, const Strip*, int, int*, double*, long long, double*, \
                                                          long long, const double*, long long, const long long*); \
    template __global__ void k_sweep_bwd_plain<NY, true>(DevCtx, const Strip*, int, int*, double*, long long, double*, \
                                                         long long, const double*, long long, const long long*);  \
    template __global__ void k_update<NY>(DevCtx, const Tile*);
SHCFD_FOR_EACH_NY(INST)
#undef INST

}  // namespace shcfd
