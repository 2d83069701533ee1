// k_sweep16.hip -- LU-SGS hyperplane sweeps for nY > 1 on 16-row strips
// (the species-extended calc_next_step_implicit of LU-SGS/2d/
// time_lusgs_viscous.f90:24-63, NR/2d/time_NR_viscous.f90, precon/2d/ and
// preconLU-SGS/2d/time_precon_viscous.f90; dual time with its S scaling keeps
// the row-per-lane sweeps of k_implicit.hip).
//
// Layout.  A workgroup owns 16 rows of a block and has C = ceil(dq/4) waves;
// lane 4r + c of wave w holds component k = 4w + c of row j0 + r, so the
// 16 x dq unknowns of a strip occupy one lane each.  At step t row r works on
// column i = t - r + 1 (forward) or i = ni - t + 15 - r (backward):
//   * the i-neighbour's contribution stays in the lane (same row, previous
//     column);
//   * the j-neighbour's contribution comes from row r -/+ 1 of the same wave
//     by a 4-lane DPP shift, and for the edge row from the neighbouring
//     strip through a hand-off row: the producer stores each column's value
//     into a slot that k_fill16 set to a NaN sentinel, the consumer keeps a
//     ring of PH loads ahead and polls a slot only while it still holds the
//     sentinel (no flags, so no wave ever waits for its own stores);
//   * the dq components of x needed by a lane's matrix rows are exchanged
//     through LDS (double-buffered, one barrier per step).
// Every lane evaluates the reference's per-cell formula with the same
// operation order as the row-per-lane sweeps, on the same inputs, in a
// topological order of the same dependency graph, so the results are
// identical.  A lane reads its two matrix rows (2 dq doubles) and the cell's
// alpha, dsci, dscj of each step from the quad-row copy written by
// k_set_abpm16 as dq + 2 pairs of doubles, pair p of global step g = base + t
// at ((((set NT + g) C + w) (dq+2) + p) 64 + lane) 2: every 16-byte load of
// a wave reads 1 KB of consecutive bytes, and one step of a strip is one
// contiguous block (few pages per step).  Loads run PD steps ahead of their use in
// straight-line code (the last steps re-read a valid step), so the
// compiler's memory waits count exactly the loads issued after the one a
// step consumes.
#include "shcfd_device.h"
#include "shcfd_kernels.h"

namespace shcfd {

namespace {

constexpr int PD = 3;  // prefetch distance of the operator rows, in steps
constexpr int PH = 4;  // hand-off loads in flight (PD + 1: one unrolled round)
constexpr int HO = 24; // hand-off column offset: entry (col + HO) dq + k, col >= -HO
constexpr unsigned long long SENT16 = 0x7FF4DEADBEEFCAFEULL;  // never produced by arithmetic

__device__ inline bool is_sent16(double x) { return (unsigned long long)__double_as_longlong(x) == SENT16; }

__device__ inline void set_error16(const DevCtx& d, int code) { atomicCAS(d.iflag + IFLAG_ERR, 0, code); }

// workgroup barrier after this wave's LDS writes have completed; global
// loads in flight are not waited for (the prefetched operator rows)
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// rare path of the hand-off: poll the slot until the lanes `mine` see a
// value; past the bound raise SHCFD_E_SPIN and go on with zero.  Then wait
// for the wave's outstanding loads so the common path inherits no partial
// wait state.
__device__ inline double poll16(const double* slot, double x, bool mine, long long spin_limit,
                                                   int* iflag, bool& dead) {
    long long it = 0;
    while (!dead && __ballot(mine && is_sent16(x)) != 0ULL) {
        if (it++ >= spin_limit) {
            atomicCAS(iflag + IFLAG_ERR, 0, SHCFD_E_SPIN);
            dead = true;  // no further waits in this sweep: it runs to its end with zeros
            break;
        }
        __builtin_amdgcn_s_sleep(2);
        x = load_agent(slot);
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);
    return (mine && !dead) ? x : 0.0;
}

// per-lane inputs of one step: the lane's rows of the two matrices, the
// cell values of its row and the lane's components of Δq* (and q)
template <int DQ>
struct StepIn {
    double a[DQ], b[DQ];
    double al, fi, fj, r, q;
};

// the pair-interleaved operator rows and scalars of one step: `base` is the
// lane's pair 0 of that step
template <int DQ>
__device__ inline void load_ops(const double* __restrict__ base, StepIn<DQ>& in) {
    constexpr int dq2 = DQ + 2;
#pragma unroll
    for (int p = 0; p < dq2; p++) {
        const double2 v = *reinterpret_cast<const double2*>(base + 128LL * p);
        const int e0 = 2 * p, e1 = 2 * p + 1;
        double* dst0 = e0 < DQ ? &in.a[e0] : (e0 < 2 * DQ ? &in.b[e0 - DQ] : (e0 == 2 * DQ ? &in.al : &in.fj));
        *dst0 = v.x;
        if (e1 < DQ) in.a[e1] = v.y;
        else if (e1 < 2 * DQ) in.b[e1 - DQ] = v.y;
        else if (e1 == 2 * DQ + 1) in.fi = v.y;
    }
}

}  // namespace

// hand-off rows of both sweeps: the sentinel for the columns 1..ni of the
// strip's block, zero elsewhere and in the final row (read by the strips
// that have no neighbour)
#if SHCFD_INST_GROUP == 0
__global__ void __launch_bounds__(256) k_fill16(DevCtx d, const Strip16* strips, int nstrips, double* hand,
                                                long long hstride, int dq) {
    const long long n = 2LL * (nstrips + 1) * hstride;
    const double sv = __longlong_as_double((long long)SENT16);
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int row = (int)((e / hstride) % (nstrips + 1));
        const long long col = (e % hstride) / dq - HO;
        const bool live = row < nstrips && col >= 1 && col <= d.blk[strips[row].b].ni;
        hand[e] = live ? sv : 0.0;
    }
}
#endif

// ---------------------------------------------------------------------------
// forward sweep: x = alpha (R + A+(i-1) x(i-1) dsci + B+(j-1) x(j-1) dscj)
// ---------------------------------------------------------------------------
// one wave per 4 components: up to 4 waves for dq <= 16, 5 above
template <int NY>
__global__ void __launch_bounds__(NY + 3 <= 16 ? 256 : 320) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_sweep16_fwd(DevCtx d, const Strip16* strips, int nstrips, double* hand, long long hstride, long long spin_limit,
              const double* __restrict__ jac, long long nsl) {
    constexpr int dq = NY + 3, C = (dq + 3) / 4, DQP = 4 * C;
    __shared__ double xs[2][16][DQP];
    const int s = blockIdx.x;
    const int w = (int)(threadIdx.x >> 6), L = (int)(threadIdx.x & 63), r = L >> 2, k = 4 * w + (L & 3);
    const Strip16 S = strips[s];
    const DevBlock B = d.blk[S.b];
    const int ni = B.ni, nj = B.nj, j = S.j0 + r;
    const bool rowok = j <= nj, kok = k < dq;
    const bool has_next = S.next >= 0;
    const bool mine = r == 0 && kok;  // lanes that take the neighbour strip's value
    const int nsteps = ni + 15;
    constexpr int dq2 = dq + 2;
    // lane's pair 0 of step 0 (set 0); one step = C dq2 64 pairs
    const double* jl = jac + (((S.base * C + w) * dq2) * 64 + L) * 2;
    const int kk = kok ? k : 0;
    const double* hsrc = hand + (long long)(S.prev >= 0 ? S.prev : nstrips) * hstride + (long long)HO * dq + kk;
    double* hdst = hand + (long long)s * hstride + (long long)HO * dq + kk;
    double* dqf = d.f[SHCFD_F_DQ] + (long long)kk * d.NC;
    auto load = [&](int t, StepIn<dq>& in) {
        t = t < nsteps ? t : nsteps - 1;  // past the end: re-read the last step
        load_ops<dq>(jl + 128LL * C * dq2 * t, in);
        const int i = t - r + 1;
        const int ic = i < 1 ? 1 : (i > ni ? ni : i);
        in.r = dqf[cidx(B, ic, rowok ? j : nj)];
    };
    StepIn<dq> buf[PD + 1];
    double hr[PH];
#pragma unroll
    for (int p = 0; p < PH; p++) hr[p] = load_agent(hsrc + (long long)(p + 1) * dq);  // columns 1..PH
#pragma unroll
    for (int p = 0; p < PD; p++) load(p, buf[p]);
    double a_in = 0.0, bo = 0.0;
    bool dead = false;
    const int tend = ((nsteps + PD) / (PD + 1)) * (PD + 1);  // whole rounds; the extra steps are idle
    for (int t0 = 0; t0 < tend; t0 += PD + 1) {
#pragma unroll
        for (int u = 0; u <= PD; u++) {
            const int t = t0 + u;
            {
                // neighbour value of column t + 1; refill the slot with column t + 1 + PH
                double hv = hr[u];
                hr[u] = load_agent(hsrc + (long long)(t + 1 + PH) * dq);
                load(t + PD, buf[(u + PD) % (PD + 1)]);
                const StepIn<dq>& in = buf[u];
                const int i = t - r + 1;
                const bool act = rowok && i >= 1 && i <= ni;
                double b_in = shift_up4(bo);
                if (__builtin_expect(__ballot(mine && is_sent16(hv)) != 0ULL, 0))
                    hv = poll16(hsrc + (long long)(t + 1) * dq, hv, mine, spin_limit, d.iflag, dead);
                if (r == 0) b_in = kok ? hv : 0.0;
                double x = 0.0;
                if (act && kok) {
                    x = in.al * (in.r + (a_in + b_in));
                    dqf[cidx(B, i, j)] = x;
                }
                xs[t & 1][r][k] = x;
                lds_barrier();
                double xa[dq];
#pragma unroll
                for (int l = 0; l < dq; l++) xa[l] = xs[t & 1][r][l];
                if (act) {
                    double sa = 0.0, sb = 0.0;
#pragma unroll
                    for (int l = 0; l < dq; l++) sa = sa + in.a[l] * xa[l];
#pragma unroll
                    for (int l = 0; l < dq; l++) sb = sb + in.b[l] * xa[l];
                    a_in = in.fi * sa;
                    bo = in.fj * sb;
                } else {
                    bo = 0.0;
                }
                if (has_next && r == 15 && act && kok) store_agent(hdst + (long long)i * dq, bo);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// backward sweep: x = x* - alpha (A-(i+1) x(i+1) dsci + B-(j+1) x(j+1) dscj),
// then q += x (LU-SGS, preconLU-SGS) or the (res, omega) partials of the
// inner-loop schemes.  Workgroup g handles strip nstrips-1-g, so the strip
// it waits on (the one above) has a lower workgroup id.
// ---------------------------------------------------------------------------
template <int NY>
__global__ void __launch_bounds__(NY + 3 <= 16 ? 256 : 320) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_sweep16_bwd(DevCtx d, const Strip16* strips, int nstrips, double* hand, long long hstride, double* part,
              long long spin_limit, const double* __restrict__ jac, long long nsl) {
    constexpr int dq = NY + 3, C = (dq + 3) / 4, DQP = 4 * C;
    __shared__ double xs[2][16][DQP];
    __shared__ double red[2][4];
    const int s = nstrips - 1 - (int)blockIdx.x;
    const int w = (int)(threadIdx.x >> 6), L = (int)(threadIdx.x & 63), r = L >> 2, k = 4 * w + (L & 3);
    const Strip16 S = strips[s];
    const DevBlock B = d.blk[S.b];
    const int ni = B.ni, nj = B.nj, j = S.j0 + r;
    const bool rowok = j <= nj, kok = k < dq;
    const bool has_prev = S.prev >= 0;
    const bool mine = r == 15 && kok;
    const bool inner = inner_loop_scheme(d.p.scheme);
    const double dqmax = d.p.dqmax;
    const int nsteps = ni + 15;
    constexpr int dq2 = dq + 2;
    // lane's pair 0 of step 0 of set 1 (nsl = total steps of all strips)
    const double* jl = jac + ((((nsl + S.base) * C + w) * dq2) * 64 + L) * 2;
    const int kk = kok ? k : 0;
    const double* hsrc = hand + (long long)(S.next >= 0 ? S.next : nstrips) * hstride + (long long)HO * dq + kk;
    double* hdst = hand + (long long)s * hstride + (long long)HO * dq + kk;
    double* dqf = d.f[SHCFD_F_DQ] + (long long)kk * d.NC;
    double* qf = d.f[SHCFD_F_Q] + (long long)kk * d.NC;
    auto load = [&](int t, StepIn<dq>& in) {
        t = t < nsteps ? t : nsteps - 1;  // past the end: re-read the last step
        // cell (i, j0+r) sits at forward step i + r - 1 = ni + 14 - t
        load_ops<dq>(jl + 128LL * C * dq2 * (ni + 14 - t), in);
        const int i = ni - t + 15 - r;
        const int ic = i < 1 ? 1 : (i > ni ? ni : i);
        const long long c = cidx(B, ic, rowok ? j : nj);
        in.r = dqf[c];
        in.q = qf[c];
    };
    StepIn<dq> buf[PD + 1];
    double hr[PH];
#pragma unroll
    for (int p = 0; p < PH; p++) hr[p] = load_agent(hsrc + (long long)(ni - p) * dq);  // columns ni..ni-PH+1
#pragma unroll
    for (int p = 0; p < PD; p++) load(p, buf[p]);
    double a_in = 0.0, bo = 0.0, res = 0.0, omega = 1e300;
    bool dead = false;
    const int tend = ((nsteps + PD) / (PD + 1)) * (PD + 1);  // whole rounds; the extra steps are idle
    for (int t0 = 0; t0 < tend; t0 += PD + 1) {
#pragma unroll
        for (int u = 0; u <= PD; u++) {
            const int t = t0 + u;
            {
                // neighbour value of column ni - t (row 15); refill with column ni - t - PH
                double hv = hr[u];
                hr[u] = load_agent(hsrc + (long long)(ni - t - PH) * dq);
                load(t + PD, buf[(u + PD) % (PD + 1)]);
                const StepIn<dq>& in = buf[u];
                const int i = ni - t + 15 - r;
                const bool act = rowok && i >= 1 && i <= ni;
                double b_in = shift_down4(bo);
                if (__builtin_expect(__ballot(mine && is_sent16(hv)) != 0ULL, 0))
                    hv = poll16(hsrc + (long long)(ni - t) * dq, hv, mine, spin_limit, d.iflag, dead);
                if (r == 15) b_in = kok ? hv : 0.0;
                double x = 0.0;
                if (act && kok) {
                    const long long c = cidx(B, i, j);
                    x = in.r - in.al * (a_in + b_in);
                    if (inner) {
                        if (k == NY + 2) res = res + x * x;
                        if (k < NY && x < 0.0) omega = dmin(omega, (-in.q / x) * dqmax);
                    } else if (B.active) {
                        qf[c] = in.q + x;
                    }
                    dqf[c] = x;
                }
                xs[t & 1][r][k] = x;
                lds_barrier();
                double xa[dq];
#pragma unroll
                for (int l = 0; l < dq; l++) xa[l] = xs[t & 1][r][l];
                if (act) {
                    double sa = 0.0, sb = 0.0;
#pragma unroll
                    for (int l = 0; l < dq; l++) sa = sa + in.a[l] * xa[l];
#pragma unroll
                    for (int l = 0; l < dq; l++) sb = sb + in.b[l] * xa[l];
                    a_in = in.fi * sa;
                    bo = in.fj * sb;
                } else {
                    bo = 0.0;
                }
                if (has_prev && r == 0 && act && kok) store_agent(hdst + (long long)i * dq, bo);
            }
        }
    }
    if (inner) {
        // fixed-order reduction over the strip: lanes, then waves
        for (int off = 32; off >= 1; off >>= 1) {
            res = res + __shfl_xor(res, off);
            omega = dmin(omega, __shfl_xor(omega, off));
        }
        if (L == 0) { red[0][w] = res; red[1][w] = omega; }
        __syncthreads();
        if (threadIdx.x == 0) {
            double R = red[0][0], O = red[1][0];
            for (int q = 1; q < C; q++) { R = R + red[0][q]; O = dmin(O, red[1][q]); }
            part[2 * s] = R;
            part[2 * s + 1] = O;
        }
    }
}

#define INST(NY)                                                                                              \
    template __global__ void k_sweep16_fwd<NY>(DevCtx, const Strip16*, int, double*, long long, long long,      \
                                               const double*, long long);                                     \
    template __global__ void k_sweep16_bwd<NY>(DevCtx, const Strip16*, int, double*, long long, double*,        \
                                               long long, const double*, long long);
SHCFD_FOR_EACH_NY(INST)
#undef INST

}  // namespace shcfd
