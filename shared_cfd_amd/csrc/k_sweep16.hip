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

// ---------------------------------------------------------------------------
// Jacobian-free quad-row sweeps: LU-SGS and NR at nY > 1.
//
// The split Jacobians of these schemes are A+- = Atilde/2 +- (nuA/2) I with
// Atilde = n_x A + n_y B (LU-SGS/2d/sch_lusgs.raw.f90:112-129; NR likewise),
// A and B the species-extended Euler flux Jacobians of :39-110.  Their
// product with x = (x_k (k < nY), X, V, E) follows in closed form from those
// entries:
//     (Atilde x)_k    = U x_k + Y_k e                    (k < nY)
//     (Atilde x)_nY   = U X + u e + n_x dp
//     (Atilde x)_nY+1 = U V + v e + n_y dp
//     (Atilde x)_nY+2 = U E + H e + U dp
// with U = n_x u + n_y v, s = sum_k x_k, e = n_x X + n_y V - U s and
// dp = sum_k DHi_k x_k + (gamma-1) ((u^2+v^2)/2 s - u X - v V + E).
// A step therefore needs the cell's state (2 nY + 13 numbers) instead of
// the 2 dq^2 entries of the two matrices: every lane reads one pair (c1, DHi)
// (c1 = Y_k, u, v or H for its component) and the 13 cell scalars come
// through LDS, written by the lanes of components 0..7 from their prefetched
// share.  The sums s and dp are taken in the same order by every lane.  The
// algebra is exact; the rounding differs from the matrix products
// (agreement with them to ~1e-13 relative per sweep).
// ---------------------------------------------------------------------------

namespace {
constexpr int JPD = 11;   // prefetch distance of the cell data, in steps (JPD + 1 = 12 steps per round)
// a round of JPD + 1 steps must cycle whole rings: the hand-off loads (PH),
// the scalar slots (3) and the LDS exchange buffers (2)
static_assert((JPD + 1) % PH == 0 && (JPD + 1) % 6 == 0, "round length");
constexpr int JSC = 16;   // cell scalars per (step, row): u v g H nxi nyi nxj nyj nuA/2 nuB/2 alpha dsci dscj
constexpr int JST = 17;   // their row stride in LDS (odd: the 16 rows of a read fall in distinct banks)
enum { JS_U = 0, JS_V, JS_G, JS_H, JS_NXI, JS_NYI, JS_NXJ, JS_NYJ, JS_HNA, JS_HNB, JS_AL, JS_DSCI, JS_DSCJ };

template <int CTRL>
__device__ inline double quad_perm(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
// (x0 + x1) + (x2 + x3) of the four lanes of each quad, in every lane
__device__ inline double quad_sum(double x) {
    const double y = x + quad_perm<1 | (0 << 2) | (3 << 4) | (2 << 6)>(x);   // quad_perm [1,0,3,2]
    return y + quad_perm<2 | (3 << 2) | (0 << 4) | (1 << 6)>(y);             // quad_perm [2,3,0,1]
}

// Buffer views: base in scalar registers, a per-lane byte offset fixed for
// the sweep and a per-step byte offset in a scalar register, so an access
// costs one instruction; a lane that must not store gives an offset beyond
// the view (the hardware's range check drops it).
typedef unsigned int jv_u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int jv_u32x2 __attribute__((ext_vector_type(2)));
constexpr int JV_DROP = 0x7FFFFF00;
constexpr int JV_AGENT = 16;   // sc1: device-coherent access
__device__ inline __amdgpu_buffer_rsrc_t jview(const void* base, long long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
__device__ inline double jv_d(unsigned lo, unsigned hi) {
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline double2 jv_ld2(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    const jv_u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
    return make_double2(jv_d(v.x, v.y), jv_d(v.z, v.w));
}
template <int AUX = 0>
__device__ inline double jv_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    const jv_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX);
    return jv_d(v.x, v.y);
}
template <int AUX = 0>
__device__ inline void jv_st(__amdgpu_buffer_rsrc_t r, double x, int voff, int soff) {
    const unsigned long long a = (unsigned long long)__double_as_longlong(x);
    const jv_u32x2 v = {(unsigned)a, (unsigned)(a >> 32)};
    __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, AUX);
}

// per-lane inputs of one step of the Jacobian-free sweeps
struct JStepIn {
    double2 cp;    // (c1, DHi) of the lane's component
    double2 sp;    // the lane's share of the next step's cell scalars
    double r, q;
};

// the lane's outgoing terms dsci A+-(x) and dscj B+-(x) (sgn = +1 forward,
// -1 backward) from its row's sums s = sum_k x_k and qd = sum_k DHi_k x_k,
// the momentum and energy components X, V, E, its own component xk and the
// cell scalars sc (LDS, row stride JST)
template <int NY, bool FWD>
__device__ inline void jvp_out(double s, double qd, double X, double V, double E, double xk, const double* sc, int k,
                               double c1, double& a_out, double& b_out) {
#pragma clang fp contract(fast)
    const double u = sc[JS_U], v = sc[JS_V], g = sc[JS_G];
    const double nxi = sc[JS_NXI], nyi = sc[JS_NYI], nxj = sc[JS_NXJ], nyj = sc[JS_NYJ];
    const double Ui = (nxi * u) + (nyi * v), Uj = (nxj * u) + (nyj * v);
    const double ei = ((nxi * X) + (nyi * V)) - (Ui * s);
    const double ej = ((nxj * X) + (nyj * V)) - (Uj * s);
    const double dp = qd + g * ((((0.5 * ((u * u) + (v * v))) * s) - ((u * X) + (v * V))) + E);
    const double cai = k < NY ? 0.0 : (k == NY ? nxi : (k == NY + 1 ? nyi : Ui));
    const double caj = k < NY ? 0.0 : (k == NY ? nxj : (k == NY + 1 ? nyj : Uj));
    const double oA = ((Ui * xk) + (c1 * ei)) + (cai * dp);
    const double oB = ((Uj * xk) + (c1 * ej)) + (caj * dp);
    if (FWD) {
        a_out = sc[JS_DSCI] * ((0.5 * oA) + (sc[JS_HNA] * xk));
        b_out = sc[JS_DSCJ] * ((0.5 * oB) + (sc[JS_HNB] * xk));
    } else {
        a_out = sc[JS_DSCI] * ((0.5 * oA) - (sc[JS_HNA] * xk));
        b_out = sc[JS_DSCJ] * ((0.5 * oB) - (sc[JS_HNB] * xk));
    }
}
}  // namespace

// the cell data of the Jacobian-free sweeps, in their strip-step layout:
// the pair (c1, DHi) of (global step g, wave w, lane L) at cp[(g C + w) 64 + L]
// and scalar p of (step g, row r) at sc[(g 16 + r) JSC + p].  Slots of no
// cell stay zero (the arrays are cleared once at allocation).  Also stores
// alpha, as set_ABpm does.
template <int NY>
__global__ void __launch_bounds__(256) k_set_jd16(DevCtx d, const Strip16* strips, double2* cp, double* sc) {
    using D = Dims<NY>;
    constexpr int dq = D::dq, C = (dq + 3) / 4;
    const Strip16 S = strips[blockIdx.x];
    const int t = (int)blockIdx.y * 16 + (int)(threadIdx.x >> 4), r = (int)(threadIdx.x & 15);
    const DevBlock B = d.blk[S.b];
    if (t >= B.ni + 15) return;
    const int i = t - r + 1, j = S.j0 + r;
    if (i < 1 || i > B.ni || j > B.nj) return;
    const long long c = cidx(B, i, j), g = S.base + t;
    // the scalars of cell_jacobian for the non-preconditioned schemes
    // (sch_lusgs.raw.f90:26-38, :131-133; sch_NR.raw.f90:134-136)
    const double rho = FLD(d, SHCFD_F_W, 0, c), u = FLD(d, SHCFD_F_W, IW_U, c), v = FLD(d, SHCFD_F_W, IW_V, c);
    const double p = FLD(d, SHCFD_F_W, IW_P, c), gamm = FLD(d, SHCFD_F_W, D::ig, c), H = FLD(d, SHCFD_F_W, D::ih, c);
    const double cs = sqrt((gamm * p) / rho);
    const double n1a = FLD(d, SHCFD_F_VNCI, 0, c), n2a = FLD(d, SHCFD_F_VNCI, 1, c);
    const double n1b = FLD(d, SHCFD_F_VNCJ, 0, c), n2b = FLD(d, SHCFD_F_VNCJ, 1, c);
    const double unA = (n1a * u) + (n2a * v), unB = (n1b * u) + (n2b * v);
    const double nuA = fabs(unA) + cs, nuB = fabs(unB) + cs;
    const double dsci = FLD(d, SHCFD_F_DSCI, 0, c), dscj = FLD(d, SHCFD_F_DSCJ, 0, c);
    const double vol = FLD(d, SHCFD_F_VOL, 0, c);
    const double dtv = (d.p.time_step == 0) ? d.scal[SCAL_DT_GRBL] : FLD(d, SHCFD_F_DT, 0, c);
    const double alpha = (d.p.scheme == SHCFD_SCHEME_NR)
                             ? 1.0 / ((((3.0 * vol) / (2.0 * dtv)) + (dsci * nuA)) + (dscj * nuB))
                             : 1.0 / (((vol / dtv) + (dsci * nuA)) + (dscj * nuB));
    FLD(d, SHCFD_F_ALPHA, 0, c) = alpha;
#pragma unroll
    for (int w = 0; w < C; w++)
#pragma unroll
        for (int cc = 0; cc < 4; cc++) {
            const int k = 4 * w + cc;
            double c1 = 0.0, dh = 0.0;
            if (k < NY) {
                c1 = FLD(d, SHCFD_F_W, 4 + k, c);
                dh = FLD(d, SHCFD_F_DHI, k, c);
            } else if (k == NY) {
                c1 = u;
            } else if (k == NY + 1) {
                c1 = v;
            } else if (k == NY + 2) {
                c1 = H;
            }
            cp[(g * C + w) * 64 + 4 * r + cc] = make_double2(c1, dh);
        }
    double* o = sc + (g * 16 + r) * JSC;
    o[JS_U] = u;
    o[JS_V] = v;
    o[JS_G] = gamm - 1.0;
    o[JS_H] = H;
    o[JS_NXI] = n1a;
    o[JS_NYI] = n2a;
    o[JS_NXJ] = n1b;
    o[JS_NYJ] = n2b;
    o[JS_HNA] = nuA * 0.5;
    o[JS_HNB] = nuB * 0.5;
    o[JS_AL] = alpha;
    o[JS_DSCI] = dsci;
    o[JS_DSCJ] = dscj;
}

// One step of a Jacobian-free sweep, per lane: x of its component from the
// incoming terms and alpha (read one step ahead); then x, the quad partial
// sums of x_k and DHi_k x_k over the species lanes, and the next step's cell
// scalars go to LDS; after the barrier the lane reads X, V, E, the partial
// sums of its row and the scalars, and forms its outgoing terms.  LDS rows
// have odd strides, so the 16 rows of a wave's read fall in distinct banks.

// forward sweep, Jacobian-free: x = alpha (R + A+(i-1) x(i-1) dsci + B+(j-1) x(j-1) dscj)
template <int NY>
__global__ void __launch_bounds__(NY + 3 <= 16 ? 256 : 320) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_sweep16j_fwd(DevCtx d, const Strip16* strips, int nstrips, double* hand, long long hstride, long long spin_limit,
               const double2* __restrict__ jcp, const double* __restrict__ jsc) {
    constexpr int dq = NY + 3, C = (dq + 3) / 4, XST = (4 * C) | 1, PST = 2 * C + 1;
    constexpr int CP_STEP = C * 64 * 16, SC_STEP = 16 * JSC * 8;
    __shared__ double xs[2][16][XST];
    __shared__ double pq[2][16][PST];
    __shared__ double scs[3][16][JST];
    const int s = blockIdx.x;
    const int w = (int)(threadIdx.x >> 6), L = (int)(threadIdx.x & 63), r = L >> 2, k = 4 * w + (L & 3);
    const Strip16 S = strips[s];
    const DevBlock B = d.blk[S.b];
    const int ni = B.ni, nj = B.nj, j = S.j0 + r;
    const bool rowok = j <= nj, kok = k < dq;
    const bool has_next = S.next >= 0;
    const bool mine = r == 0 && kok;
    const int nsteps = ni + 15;
    const int sp_slot = 4 * w + (L & 3);   // the scalar pair this lane carries (< 8 only)
    const bool carrier = sp_slot < 8;
    const int kk = kok ? k : 0;
    const long long prow = S.prev >= 0 ? S.prev : nstrips;
    const double* hpoll = hand + prow * hstride + (long long)HO * dq + kk;
    const __amdgpu_buffer_rsrc_t vin = jview(hand + prow * hstride, hstride * 8);
    const __amdgpu_buffer_rsrc_t vout = jview(hand + (long long)s * hstride, hstride * 8);
    const int h_lane = (HO * dq + kk) * 8;                                     // column c at + c dq 8
    const int h_store = (has_next && r == 15 && kok) ? h_lane - 14 * dq * 8 : JV_DROP;  // column t - 14
    const __amdgpu_buffer_rsrc_t vcp = jview(jcp + S.base * C * 64, (long long)nsteps * CP_STEP);
    const int cp_lane = (w * 64 + L) * 16;
    const __amdgpu_buffer_rsrc_t vsc = jview(jsc + S.base * 16 * JSC, (long long)nsteps * SC_STEP);
    const int sc_lane = (r * JSC + 2 * (carrier ? sp_slot : 0)) * 8;
    // one view of all dq components (uniform across the wave: a per-lane base
    // would make every access a loop over the lanes' distinct bases); the
    // lane's component and cell in the lane offset
    const __amdgpu_buffer_rsrc_t vdq = jview(d.f[SHCFD_F_DQ], (long long)dq * d.NC * 8);
    const int dq_lane = (int)((kk * d.NC + cidx(B, 0, rowok ? j : nj) + 1 - r) * 8);  // the cell of step t at + 8 t
    auto load = [&](int t, JStepIn& in) {
        const int tc = t < nsteps ? t : nsteps - 1;  // past the end: re-read the last step
        const int tn = t + 1 < nsteps ? t + 1 : nsteps - 1;
        in.cp = jv_ld2(vcp, cp_lane, tc * CP_STEP);
        in.sp = jv_ld2(vsc, sc_lane, tn * SC_STEP);
        in.r = jv_ld(vdq, dq_lane, tc * 8);
    };
    {
        const double2 v0 = jv_ld2(vsc, sc_lane, 0);
        if (carrier) {
            scs[0][r][2 * sp_slot] = v0.x;
            scs[0][r][2 * sp_slot + 1] = v0.y;
        }
    }
    JStepIn buf[JPD + 1];
    double hr[PH];
#pragma unroll
    for (int p = 0; p < PH; p++) hr[p] = jv_ld<JV_AGENT>(vin, h_lane, (p + 1) * dq * 8);  // columns 1..PH
#pragma unroll
    for (int p = 0; p < JPD; p++) load(p, buf[p]);
    lds_barrier();
    double al = scs[0][r][JS_AL];
    double a_in = 0.0, bo = 0.0;
    bool dead = false;
    const int tend = ((nsteps + JPD) / (JPD + 1)) * (JPD + 1);  // whole rounds; the extra steps are idle
    for (int t0 = 0; t0 < tend; t0 += JPD + 1) {
#pragma unroll
        for (int u = 0; u <= JPD; u++) {
            const int t = t0 + u;
            double hv = hr[u % PH];
            hr[u % PH] = jv_ld<JV_AGENT>(vin, h_lane, (t + 1 + PH) * dq * 8);
            load(t + JPD, buf[(u + JPD) % (JPD + 1)]);
            const JStepIn& in = buf[u];
            const int i = t - r + 1;
            const bool act = rowok && i >= 1 && i <= ni;
            double b_in = shift_up4(bo);
            if (__builtin_expect(__ballot(mine && is_sent16(hv)) != 0ULL, 0))
                hv = poll16(hpoll + (long long)(t + 1) * dq, hv, mine, spin_limit, d.iflag, dead);
            if (r == 0) b_in = kok ? hv : 0.0;
            const double xc = al * (in.r + (a_in + b_in));
            const double x = (act && kok) ? xc : 0.0;
            jv_st(vdq, x, (act && kok) ? dq_lane : JV_DROP, t * 8);
            xs[u & 1][r][k] = x;
            const double ps = quad_sum(k < NY ? x : 0.0);
            const double pd = quad_sum(in.cp.y * x);
            if ((L & 3) == 0) {
                pq[u & 1][r][2 * w] = ps;
                pq[u & 1][r][2 * w + 1] = pd;
            }
            if (carrier) {
                scs[(u + 1) % 3][r][2 * sp_slot] = in.sp.x;
                scs[(u + 1) % 3][r][2 * sp_slot + 1] = in.sp.y;
            }
            lds_barrier();
            const double* sct = scs[u % 3][r];
            const double* xr = xs[u & 1][r];
            const double* pr = pq[u & 1][r];
            double sx = pr[0], qd = pr[1];
#pragma unroll
            for (int q = 1; q < C; q++) {
                sx = sx + pr[2 * q];
                qd = qd + pr[2 * q + 1];
            }
            double ao, bn;
            jvp_out<NY, true>(sx, qd, xr[NY], xr[NY + 1], xr[NY + 2], x, sct, kk, in.cp.x, ao, bn);
            al = scs[(u + 1) % 3][r][JS_AL];
            a_in = act ? ao : a_in;
            bo = act ? bn : 0.0;
            jv_st<JV_AGENT>(vout, bo, act ? h_store : JV_DROP, t * dq * 8);
        }
    }
}

// backward sweep, Jacobian-free: x = x* - alpha (A-(i+1) x(i+1) dsci + B-(j+1) x(j+1) dscj),
// and the (res, omega) partials (NR); LU-SGS's q += x is k_add_dq's pass
// after the sweep (the scattered q accesses of each step, 16 rows x 4
// components of 8 bytes, cost more than the pass)
template <int NY>
__global__ void __launch_bounds__(NY + 3 <= 16 ? 256 : 320) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_sweep16j_bwd(DevCtx d, const Strip16* strips, int nstrips, double* hand, long long hstride, double* part,
               long long spin_limit, const double2* __restrict__ jcp, const double* __restrict__ jsc) {
    constexpr int dq = NY + 3, C = (dq + 3) / 4, XST = (4 * C) | 1, PST = 2 * C + 1;
    constexpr int CP_STEP = C * 64 * 16, SC_STEP = 16 * JSC * 8;
    __shared__ double xs[2][16][XST];
    __shared__ double pq[2][16][PST];
    __shared__ double scs[3][16][JST];
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
    const int sp_slot = 4 * w + (L & 3);
    const bool carrier = sp_slot < 8;
    const int kk = kok ? k : 0;
    const long long nrow = S.next >= 0 ? S.next : nstrips;
    const double* hpoll = hand + nrow * hstride + (long long)HO * dq + kk;
    const __amdgpu_buffer_rsrc_t vin = jview(hand + nrow * hstride, hstride * 8);
    const __amdgpu_buffer_rsrc_t vout = jview(hand + (long long)s * hstride, hstride * 8);
    const int h_lane = (HO * dq + kk) * 8;
    const int h_store = (has_prev && r == 0 && kok) ? h_lane : JV_DROP;   // column ni + 15 - t
    // backward step t of row r visits the cell of forward step ni + 14 - t
    const __amdgpu_buffer_rsrc_t vcp = jview(jcp + S.base * C * 64, (long long)nsteps * CP_STEP);
    const int cp_lane = (w * 64 + L) * 16;
    const __amdgpu_buffer_rsrc_t vsc = jview(jsc + S.base * 16 * JSC, (long long)nsteps * SC_STEP);
    const int sc_lane = (r * JSC + 2 * (carrier ? sp_slot : 0)) * 8;
    // views of all dq and q components, as in the forward sweep (q read for
    // the omega limit of the inner-loop schemes only)
    const __amdgpu_buffer_rsrc_t vdq = jview(d.f[SHCFD_F_DQ], (long long)dq * d.NC * 8);
    const __amdgpu_buffer_rsrc_t vq = jview(d.f[SHCFD_F_Q], (long long)dq * d.NC * 8);
    const int c_lane = (int)((kk * d.NC + cidx(B, 0, rowok ? j : nj) + 1 - r) * 8);  // the cell of step t at + 8 (ni + 14 - t)
    auto load = [&](int t, JStepIn& in) {
        const int tc = t < nsteps ? t : nsteps - 1;
        const int tn = t + 1 < nsteps ? t + 1 : nsteps - 1;
        in.cp = jv_ld2(vcp, cp_lane, (ni + 14 - tc) * CP_STEP);
        in.sp = jv_ld2(vsc, sc_lane, (ni + 14 - tn) * SC_STEP);
        in.r = jv_ld(vdq, c_lane, (ni + 14 - tc) * 8);
        in.q = jv_ld(vq, inner ? c_lane : JV_DROP, (ni + 14 - tc) * 8);  // omega's only
    };
    {
        const double2 v0 = jv_ld2(vsc, sc_lane, (ni + 14) * SC_STEP);
        if (carrier) {
            scs[0][r][2 * sp_slot] = v0.x;
            scs[0][r][2 * sp_slot + 1] = v0.y;
        }
    }
    JStepIn buf[JPD + 1];
    double hr[PH];
#pragma unroll
    for (int p = 0; p < PH; p++) hr[p] = jv_ld<JV_AGENT>(vin, h_lane, (ni - p) * dq * 8);  // columns ni..ni-PH+1
#pragma unroll
    for (int p = 0; p < JPD; p++) load(p, buf[p]);
    lds_barrier();
    double al = scs[0][r][JS_AL];
    double a_in = 0.0, bo = 0.0, res = 0.0, omega = 1e300;
    bool dead = false;
    const int tend = ((nsteps + JPD) / (JPD + 1)) * (JPD + 1);
    for (int t0 = 0; t0 < tend; t0 += JPD + 1) {
#pragma unroll
        for (int u = 0; u <= JPD; u++) {
            const int t = t0 + u;
            double hv = hr[u % PH];
            hr[u % PH] = jv_ld<JV_AGENT>(vin, h_lane, (ni - t - PH) * dq * 8);
            load(t + JPD, buf[(u + JPD) % (JPD + 1)]);
            const JStepIn& in = buf[u];
            const int i = ni - t + 15 - r;
            const bool act = rowok && i >= 1 && i <= ni;
            double b_in = shift_down4(bo);
            if (__builtin_expect(__ballot(mine && is_sent16(hv)) != 0ULL, 0))
                hv = poll16(hpoll + (long long)(ni - t) * dq, hv, mine, spin_limit, d.iflag, dead);
            if (r == 15) b_in = kok ? hv : 0.0;
            const double xc = in.r - al * (a_in + b_in);
            const bool live = act && kok;
            const double x = live ? xc : 0.0;
            if (inner && live) {
                if (k == NY + 2) res = res + x * x;
                if (k < NY && x < 0.0) omega = dmin(omega, (-in.q / x) * dqmax);
            }
            jv_st(vdq, x, live ? c_lane : JV_DROP, (ni + 14 - t) * 8);
            xs[u & 1][r][k] = x;
            const double ps = quad_sum(k < NY ? x : 0.0);
            const double pd = quad_sum(in.cp.y * x);
            if ((L & 3) == 0) {
                pq[u & 1][r][2 * w] = ps;
                pq[u & 1][r][2 * w + 1] = pd;
            }
            if (carrier) {
                scs[(u + 1) % 3][r][2 * sp_slot] = in.sp.x;
                scs[(u + 1) % 3][r][2 * sp_slot + 1] = in.sp.y;
            }
            lds_barrier();
            const double* sct = scs[u % 3][r];
            const double* xr = xs[u & 1][r];
            const double* pr = pq[u & 1][r];
            double sx = pr[0], qd = pr[1];
#pragma unroll
            for (int q = 1; q < C; q++) {
                sx = sx + pr[2 * q];
                qd = qd + pr[2 * q + 1];
            }
            double ao, bn;
            jvp_out<NY, false>(sx, qd, xr[NY], xr[NY + 1], xr[NY + 2], x, sct, kk, in.cp.x, ao, bn);
            al = scs[(u + 1) % 3][r][JS_AL];
            a_in = act ? ao : a_in;
            bo = act ? bn : 0.0;
            jv_st<JV_AGENT>(vout, bo, act ? h_store : JV_DROP, (ni + 15 - t) * dq * 8);
        }
    }
    if (inner) {
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
                                               long long, const double*, long long);                          \
    template __global__ void k_set_jd16<NY>(DevCtx, const Strip16*, double2*, double*);                         \
    template __global__ void k_sweep16j_fwd<NY>(DevCtx, const Strip16*, int, double*, long long, long long,     \
                                                const double2*, const double*);                               \
    template __global__ void k_sweep16j_bwd<NY>(DevCtx, const Strip16*, int, double*, long long, double*,       \
                                                long long, const double2*, const double*);
SHCFD_FOR_EACH_NY(INST)
#undef INST

}  // namespace shcfd
