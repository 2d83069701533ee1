// k_chem.hip -- Chemkin thermochemistry kernels: the reactive
// set_thermo_prop (calc_T per cell), proceed_reaction (CHEMEQ2 or the
// point-implicit LU update per cell) and a batched 0-D harness (Fex, Jex,
// calc_T, CHEMEQ2, point-implicit on arrays of independent states).
//
// One lane owns one cell; the per-cell species vectors and the 10x10
// Jacobian stay in registers (shcfd_chem.h).  The mechanism tables are read
// through uniform (scalar) loads.
#include "shcfd_chem.h"
#include "shcfd_device.h"
#include "shcfd_dvode.h"
#include "shcfd_kernels.h"

namespace shcfd {

// store/therm_lib/chemkin/flow/reactive/thermal_model.f90:13-42
template <int NS>
__global__ void __launch_bounds__(256) k_thermo_chem(DevCtx d, const Tile* tiles, const shcfd_mech* __restrict__ m) {
    using D = Dims<NS>;
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    if (t.i < 1 || t.i > B.ni || t.j < 1 || t.j > B.nj) return;
    long long c = cidx(B, t.i, t.j);
    double rho = FLD(d, SHCFD_F_W, IW_RHO, c), u = FLD(d, SHCFD_F_W, IW_U, c), v = FLD(d, SHCFD_F_W, IW_V, c);
    double qe = FLD(d, SHCFD_F_Q, NS + 2, c);
    double ei = qe / rho - 0.5 * (u * u + v * v);
    double T = FLD(d, SHCFD_F_W, IW_P, c) / (rho * FLD(d, SHCFD_F_W, D::ir, c));
    double vrho[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) vrho[k] = FLD(d, SHCFD_F_Q, k, c);
    chem::ThermoOut<NS> o;
    if (chem::calc_t<NS>(m, vrho, rho, ei, T, o) < 0) atomicCAS(d.iflag + IFLAG_ERR, 0, SHCFD_E_TNEWTON);
    double rg = chem::R_UNI / o.mwave;
    double p = rho * rg * T;
    FLD(d, SHCFD_F_W, D::ir, c) = rg;
    FLD(d, SHCFD_F_W, IW_P, c) = p;
    FLD(d, SHCFD_F_W, D::ig, c) = o.kappa;
    FLD(d, SHCFD_F_W, D::ih, c) = (qe + p) / rho;
    FLD(d, SHCFD_F_W, D::imu, c) = o.mu;
#pragma unroll
    for (int k = 0; k < NS; k++) {
        FLD(d, SHCFD_F_DHI, k, c) = o.dhi[k];
        FLD(d, SHCFD_F_VHI, k, c) = o.vhi[k];
    }
}

// one cell of proceed_reaction; returns the Newton iterations / CHEMEQ2
// attempts it took.  A CHEMEQ2 cell handed to DVODE is appended to the list
// defer (cell index, time reached, n(ns), T) and finished by
// k_reaction_dvode.
template <int NS, int MODE>
__device__ inline int reaction_cell(const DevCtx& d, const DevBlock& B, const TileCoord& t, const shcfd_mech* m,
                                    double* defer, int cap) {
    using D = Dims<NS>;
    long long c = cidx(B, t.i, t.j);
    double dt = d.scal[SCAL_DT_GRBL];
    double T = FLD(d, SHCFD_F_W, IW_P, c) / FLD(d, SHCFD_F_W, IW_RHO, c) / FLD(d, SHCFD_F_W, D::ir, c);
    double vrho[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) vrho[k] = FLD(d, SHCFD_F_Q, k, c);
    int e, att = 0;
    if constexpr (MODE == SHCFD_REACTION_POINT_IMPLICIT) {
        double T1;
        e = chem::point_implicit<NS>(m, T, vrho, dt, att, T1);
    } else {
        double n[NS + 1], tt;
        e = chem::reaction_chemeq2<NS>(m, T, vrho, dt, att, n, tt);
        if (e == chem::CHEM_DEFER) {
            const int k = atomicAdd(d.iflag + IFLAG_DEFER, 1);
            if (k < cap) {
                double* r = defer + (long long)k * (NS + 3);
                r[0] = (double)c;
                r[1] = tt;
#pragma unroll
                for (int i = 0; i <= NS; i++) r[2 + i] = n[i];
            } else {
                atomicCAS(d.iflag + IFLAG_ERR, 0, SHCFD_E_CHEM);
            }
            return att;
        }
    }
    if (e) {
        atomicCAS(d.iflag + IFLAG_ERR, 0, e);
        return att;
    }
#pragma unroll
    for (int k = 0; k < NS; k++) FLD(d, SHCFD_F_Q, k, c) = vrho[k];
    return att;
}

// proceed_reaction, thermal_model.f90:45-73 (MODE 1: reaction_chemeq2;
// MODE 2: point-implicit backward Euler), over dt_grbl.  One instantiation
// per method, so each holds only its own registers.  The per-cell iteration
// counts are summed into SCAL_CHEM_ITERS (one atomic per wave).
template <int NS, int MODE>
__global__ void __launch_bounds__(256) k_reaction(DevCtx d, const Tile* tiles, const shcfd_mech* __restrict__ m,
                                                  double* defer, int cap) {
    TileCoord t = tile_coord(tiles);
    const DevBlock B = d.blk[t.b];
    bool inside = !(t.i < 1 || t.i > B.ni || t.j < 1 || t.j > B.nj) && B.active;  // frozen blocks keep their state
    int att = 0;
    if (inside) att = reaction_cell<NS, MODE>(d, B, t, m, defer, cap);
    double w = (double)att;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off);
    if ((threadIdx.x & 63) == 0 && w > 0.0) atomicAdd(d.scal + SCAL_CHEM_ITERS, w);
}

// the DVODE hand-over of the CHEMEQ2 cells listed by k_reaction
// (sub_chem.f90:1751-1770): one listed cell per lane, grid-stride over the
// list; the DVODE steps are added to SCAL_CHEM_ITERS.  The integrator's
// history and matrices are per-lane private arrays; the pass is separate so
// that the CHEMEQ2 kernel keeps its register budget.
template <int NS>
__global__ void __launch_bounds__(64) k_reaction_dvode(DevCtx d, const shcfd_mech* __restrict__ m,
                                                       const double* __restrict__ defer, int cap) {
    const int cnt = min(d.iflag[IFLAG_DEFER], cap);
    const double dt = d.scal[SCAL_DT_GRBL];
    double steps = 0.0;
    for (int k = blockIdx.x * 64 + (int)threadIdx.x; k < cnt; k += gridDim.x * 64) {
        const double* r = defer + (long long)k * (NS + 3);
        const long long c = (long long)r[0];
        double n[NS + 1];
        for (int i = 0; i <= NS; i++) n[i] = r[2 + i];
        int st = 0;
        const int e = chem::chemeq2_dvode<NS>(m, n, r[1], dt, m->rtol, m->atol, st);
        steps += (double)st;
        if (e) {
            atomicCAS(d.iflag + IFLAG_ERR, 0, e);
            continue;
        }
        double vrho[NS];
        chem::n_to_vrho<NS>(m, n, vrho);
#pragma unroll
        for (int i = 0; i < NS; i++) FLD(d, SHCFD_F_Q, i, c) = vrho[i];
    }
    if (steps > 0.0) atomicAdd(d.scal + SCAL_CHEM_ITERS, steps);
}

#if SHCFD_INST_GROUP == 0
__global__ void k_set_iflag(DevCtx d, int slot, int v) {
    if (threadIdx.x == 0) d.iflag[slot] = v;
}
#endif

// proceed_reaction with the point-implicit update (MODE 2), one wavefront per
// workgroup: the (NS+1)^2 Newton matrix of each lane lives in LDS
// (lane-interleaved) instead of registers, so the runtime species rows of
// the Jacobian and the LU's pivot rows are indexed directly.  Workgroup k
// handles row k % tile_rows of tile k / tile_rows.
template <int NS>
__global__ void __launch_bounds__(64) k_reaction_lds(DevCtx d, const Tile* tiles, int tile_rows,
                                                      const shcfd_mech* __restrict__ m) {
    constexpr int N = NS + 1;
    __shared__ double jsh[N * N * 64];
    const Tile tl = tiles[blockIdx.x / tile_rows];
    const int i = tl.i0 + (int)threadIdx.x, j = tl.j0 + (int)(blockIdx.x % tile_rows);
    const DevBlock B = d.blk[tl.b];
    int att = 0;
    if (i >= 1 && i <= B.ni && j >= 1 && j <= B.nj && B.active) {  // frozen blocks keep their state
        using D = Dims<NS>;
        const long long c = cidx(B, i, j);
        const double dt = d.scal[SCAL_DT_GRBL];
        const double T = FLD(d, SHCFD_F_W, IW_P, c) / FLD(d, SHCFD_F_W, IW_RHO, c) / FLD(d, SHCFD_F_W, D::ir, c);
        double vrho[NS];
#pragma unroll
        for (int k = 0; k < NS; k++) vrho[k] = FLD(d, SHCFD_F_Q, k, c);
        double T1;
        const int e = chem::point_implicit_lds<NS>(m, T, vrho, dt, att, T1, jsh + threadIdx.x);
        if (e) {
            atomicCAS(d.iflag + IFLAG_ERR, 0, e);
        } else {
#pragma unroll
            for (int k = 0; k < NS; k++) FLD(d, SHCFD_F_Q, k, c) = vrho[k];
        }
    }
    double w = (double)att;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off);
    if (threadIdx.x == 0 && w > 0.0) atomicAdd(d.scal + SCAL_CHEM_ITERS, w);
}

// batched 0-D harness: cell q's input row starts at in + q*ldi, output row at
// out + q*ldo; status[q] receives an iteration/attempt count or -code.
template <int NS, int op>
__global__ void __launch_bounds__(256) k_chem0d(const shcfd_mech* __restrict__ m, int n, const double* in, int ldi,
                                                double* out, int ldo, int* status, double dt) {
    constexpr int N = NS + 1;
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double* x = in + (long long)q * ldi;
    double* y = out + (long long)q * ldo;
    int st = 0;
    if constexpr (op == CHEM_OP_FEX || op == CHEM_OP_FEXVAR || op == CHEM_OP_JEX) {
        double v[N];
#pragma unroll
        for (int k = 0; k < N; k++) v[k] = x[k];
        if constexpr (op == CHEM_OP_FEX) {
            double f[N];
            chem::fex<NS>(m, v, f);
#pragma unroll
            for (int k = 0; k < N; k++) y[k] = f[k];
        } else if constexpr (op == CHEM_OP_FEXVAR) {
            double p[N], mm[N];
            chem::fex_var<NS>(m, v, p, mm);
#pragma unroll
            for (int k = 0; k < N; k++) {
                y[k] = p[k];
                y[N + k] = mm[k];
            }
        } else {
            double J[N][N], f[N];
            chem::jex<NS>(m, v, J, f);
            // column-major, as the reference's dndn(ns+1, ns+1)
#pragma unroll
            for (int b = 0; b < N; b++)
#pragma unroll
                for (int a = 0; a < N; a++) y[a + N * b] = J[a][b];
        }
    } else if constexpr (op == CHEM_OP_CALCT) {
        // in: vrho(NS), rho, E, T start; out: T, kappa, MWave, mu, DHi(NS), vhi(NS)
        double vrho[NS];
#pragma unroll
        for (int k = 0; k < NS; k++) vrho[k] = x[k];
        double T = x[NS + 2];
        chem::ThermoOut<NS> o;
        st = chem::calc_t<NS>(m, vrho, x[NS], x[NS + 1], T, o);
        y[0] = T;
        y[1] = o.kappa;
        y[2] = o.mwave;
        y[3] = o.mu;
#pragma unroll
        for (int k = 0; k < NS; k++) {
            y[4 + k] = o.dhi[k];
            y[4 + NS + k] = o.vhi[k];
        }
    } else if constexpr (op == CHEM_OP_POINT_IMPLICIT) {
        // in: vrho(NS), T; out: vrho(NS), backward-Euler T
        double vrho[NS];
#pragma unroll
        for (int k = 0; k < NS; k++) vrho[k] = x[k];
        int it = 0;
        double T1 = x[NS];
        int e = chem::point_implicit<NS>(m, x[NS], vrho, dt, it, T1);
        st = e ? -e : it;
#pragma unroll
        for (int k = 0; k < NS; k++) y[k] = vrho[k];
        y[NS] = T1;
    } else if constexpr (op == CHEM_OP_DVODE) {
        // in: vrho(NS), T; out: vrho(NS); status: DVODE steps
        double vrho[NS];
        for (int k = 0; k < NS; k++) vrho[k] = x[k];
        int steps = 0;
        int e = chem::reaction_dvode<NS>(m, x[NS], vrho, dt, steps);
        st = e ? -e : steps;
        for (int k = 0; k < NS; k++) y[k] = vrho[k];
    } else {
        // in: vrho(NS), T; out: vrho(NS); status: CHEMEQ2 attempts (11: DVODE finished the interval)
        double vrho[NS];
        for (int k = 0; k < NS; k++) vrho[k] = x[k];
        int att = 0;
        int e = chem::reaction_chemeq2_full<NS>(m, x[NS], vrho, dt, att);
        st = e ? -e : att;
        for (int k = 0; k < NS; k++) y[k] = vrho[k];
    }
    status[q] = st;
}

#define SHCFD_CHEM_FLOW_INST(NS)                                                                                \
    template __global__ void k_thermo_chem<NS>(DevCtx, const Tile*, const shcfd_mech*);                         \
    template __global__ void k_reaction<NS, SHCFD_REACTION_CHEMEQ2>(DevCtx, const Tile*, const shcfd_mech*,     \
                                                                    double*, int);                              \
    template __global__ void k_reaction<NS, SHCFD_REACTION_POINT_IMPLICIT>(DevCtx, const Tile*,                 \
                                                                           const shcfd_mech*, double*, int);    \
    template __global__ void k_reaction_dvode<NS>(DevCtx, const shcfd_mech*, const double*, int);               \
    template __global__ void k_reaction_lds<NS>(DevCtx, const Tile*, int, const shcfd_mech*);
#define SHCFD_CHEM0D_OP(NS, OP) \
    template __global__ void k_chem0d<NS, OP>(const shcfd_mech*, int, const double*, int, double*, int, int*, double);
#define SHCFD_CHEM0D_INST(NS)                                                                     \
    SHCFD_CHEM0D_OP(NS, CHEM_OP_FEX) SHCFD_CHEM0D_OP(NS, CHEM_OP_FEXVAR) SHCFD_CHEM0D_OP(NS, CHEM_OP_JEX) \
    SHCFD_CHEM0D_OP(NS, CHEM_OP_CALCT) SHCFD_CHEM0D_OP(NS, CHEM_OP_CHEMEQ2)                          \
    SHCFD_CHEM0D_OP(NS, CHEM_OP_POINT_IMPLICIT) SHCFD_CHEM0D_OP(NS, CHEM_OP_DVODE)
SHCFD_FOR_EACH_FLOW_NS(SHCFD_CHEM_FLOW_INST)
SHCFD_FOR_EACH_NS(SHCFD_CHEM0D_INST)

}  // namespace shcfd
