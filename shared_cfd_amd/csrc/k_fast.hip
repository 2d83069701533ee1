// k_fast.hip -- the optimised implicit update (LU-SGS / preconditioned LU-SGS
// / dual time) for the single-species ideal gas (dimq = 4).
//
// Reformulation (exact algebra, different rounding from the plain path):
// with S = post-scaling of the dual-time sweeps, S y = (y + ps4 (pre2.y) pre1)/ps1
// (S = I for LU-SGS and preconditioned LU-SGS), the forward sweep of
// time_dual_viscous.f90:37-50 becomes
//     v(i,j) = R(i,j) + GA(i-1,j) v(i-1,j) + GB(i,j-1) v(i,j-1),
//     GA = dsci*alpha*Ap*S,  GB = dscj*alpha*Bp*S,  Dq* = alpha*S*v,
// and the backward sweep (:67-93), with z = y/alpha and x2 the second-stage
// result (:53-60),
//     z(i,j) = x2/alpha - (HA(i+1,j) z(i+1,j) + HB(i,j+1) z(i,j+1)),
//     HA = dsci*alpha*Am*S,  HB = dscj*alpha*Bm*S,  Dq = alpha*S*z.
// The dependency graph is the reference's: cell (i,j) still waits for
// (i-1,j) and (i,j-1) (forward) or (i+1,j) and (i,j+1) (backward).
//
// Layout: a "quad strip" is 16 rows of one block owned by one wavefront;
// lane L = 4r+c holds component c of row r.  Row r processes column
// i = d-r+1 at step d, so the 16 rows of a step lie on an anti-diagonal.
// Per-cell coefficients are stored by (strip, d, lane) -- the "skewed
// layout" -- so every step of both sweeps reads 64 consecutive slots.
//
// Each strip's steps are padded to whole chunks of CH steps, with one chunk
// of guard steps on either side, so the sweeps prefetch and store without a
// bounds test.  Slots that hold no cell keep zero coefficients (the arrays
// are zeroed once and k_fast_prep writes only cells), which makes every
// contribution out of them exactly zero: the sweeps need no per-cell mask.
// Arrays with one double per lane and step (R, V, X, Z) interleave pairs of
// steps, (dd even, dd odd) per lane, so one 16-byte access covers two steps.
#include "shcfd_device.h"
#include "shcfd_kernels.h"

namespace shcfd {

constexpr int QR = 16;  // rows per quad strip
constexpr int CH = 16;  // steps per unrolled sweep chunk (padding unit)

struct QStrip {
    int b, j0, prev, next;
    long long base;  // slot of (step 0, lane 0); multiple of 128
    int nch, pad;    // chunks of CH steps covering the ni + QR - 1 steps
};

// slot of the one-double arrays for (step dd, lane L): pair-interleaved
__device__ inline long long slot1(long long base, int dd, int L) {
    return base + 64LL * (dd & ~1) + 2 * L + (dd & 1);
}

// quad broadcast: every lane of a 4-lane group gets lane k of the group
template <int K>
__device__ inline double quad_bcast(double x) {
    constexpr int ctrl = K | (K << 2) | (K << 4) | (K << 6);
    int lo = __double2loint(x), hi = __double2hiint(x);
    // every lane has a source within its quad: bound_ctrl set, no old value
    lo = __builtin_amdgcn_mov_dpp(lo, ctrl, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp(hi, ctrl, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
// lane L receives lane L-4 (0 for L < 4) / lane L+4 (0 for L > 59)
__device__ inline double shift_up4(double x) {
    int lo = __double2loint(x), hi = __double2hiint(x);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        lo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xF, 0xF, true);
        hi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xF, 0xF, true);
    }
    return __hiloint2double(hi, lo);
}
__device__ inline double shift_down4(double x) {
    int lo = __double2loint(x), hi = __double2hiint(x);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        lo = __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xF, 0xF, true);
        hi = __builtin_amdgcn_update_dpp(0, hi, 0x130, 0xF, 0xF, true);
    }
    return __hiloint2double(hi, lo);
}

// 4x4 helpers (column-major)
__device__ inline void mm4(const double* A, const double* B, double* C) {
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int l = 0; l < 4; l++) {
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < 4; m++) s = fma(A[k + 4 * m], B[m + 4 * l], s);
            C[k + 4 * l] = s;
        }
}

// post-scaling matrix S (dual) or identity
__device__ inline void post_matrix(const CellJac<1>& J, bool dual, double* S) {
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int l = 0; l < 4; l++) S[k + 4 * l] = (k == l) ? 1.0 : 0.0;
    if (dual) {
        const double ips1 = 1.0 / J.ps[0], ps4 = J.ps[3];
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int l = 0; l < 4; l++) S[k + 4 * l] = ips1 * (S[k + 4 * l] + ps4 * J.pre1[k] * J.pre2[l]);
    }
}

// cell of slot (strip S, step d, row r)
__device__ inline bool slot_cell(const QStrip& S, const DevBlock& B, int d, int r, int& i, int& j) {
    i = d - r + 1;
    j = S.j0 + r;
    return i >= 1 && i <= B.ni && j <= B.nj;
}

// ---------------------------------------------------------------------------
// prep: per cell, the folded sweep operators in the skewed layout.
// Workgroup = 16 steps x 16 rows of one quad strip; thread = one cell.
// ---------------------------------------------------------------------------
constexpr unsigned long long SENTINEL = 0x7FF4DEADBEEFCAFEULL;

// Hand-off rows: one per strip and sweep plus a final row of zeros (read by
// the strips with no neighbour); entry 4*(HOFF+i)+c holds component c of
// column i, for i from -HOFF on.
constexpr int HOFF = 48;

__global__ void __launch_bounds__(256) k_fast_prep(DevCtx d, const QStrip* strips, const int2* work, double* R_sl,
                                                   double* GA, double* GB, double* HA, double* HB, double* hand,
                                                   int nstrips, long long hstride) {
    // refill the hand-off rows of both sweeps: the sentinel for the columns
    // 1..ni of the strip's block, zero elsewhere
    {
        const long long stride = (long long)gridDim.x * blockDim.x;
        const long long nhand = 2LL * (nstrips + 1) * hstride;
        const double sv = __longlong_as_double((long long)SENTINEL);
        for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nhand; k += stride) {
            const int row = (int)((k / hstride) % (nstrips + 1));
            const long long col = (k % hstride) / 4 - HOFF;
            const bool live = row < nstrips && col >= 1 && col <= d.blk[strips[row].b].ni;
            hand[k] = live ? sv : 0.0;
        }
    }
    const int2 wk = work[blockIdx.x];
    const QStrip S = strips[wk.x];
    const DevBlock B = d.blk[S.b];
    const int r = threadIdx.x & 15;
    const int dd = wk.y + (threadIdx.x >> 4);
    int i, j;
    if (!slot_cell(S, B, dd, r, i, j)) return;
    const long long c = cidx(B, i, j);
    const bool dual = d.p.scheme == SHCFD_SCHEME_DUAL;
    CellJac<1> J;
    cell_jacobian<1>(d, c, B.pitch, J);
    FLD(d, SHCFD_F_ALPHA, 0, c) = J.alpha;
    if (d.p.scheme == SHCFD_SCHEME_PRECON_LUSGS) FLD(d, SHCFD_F_DT, 0, c) = J.dtm;
    double Sm[16], T[16];
    post_matrix(J, dual, Sm);
    const double dsci = FLD(d, SHCFD_F_DSCI, 0, c), dscj = FLD(d, SHCFD_F_DSCJ, 0, c);
    const double fa = dsci * J.alpha, fb = dscj * J.alpha;
    const long long slot = S.base + (long long)dd * 64 + 4 * r;  // lane of component 0
    // R (component c in lane 4r+c)
#pragma unroll
    for (int k = 0; k < 4; k++) R_sl[slot1(S.base, dd, 4 * r + k)] = FLD(d, SHCFD_F_DQ, k, c);
    // rows of the four operators: lane 4r+c holds row c (4 doubles)
    double* outs[4] = {GA, GB, HA, HB};
    const double* mats[4] = {J.Ap, J.Bp, J.Am, J.Bm};
    const double fs[4] = {fa, fb, fa, fb};
#pragma unroll
    for (int m = 0; m < 4; m++) {
        mm4(mats[m], Sm, T);
        double* o = outs[m] + 4 * slot;
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int l = 0; l < 4; l++) o[4 * k + l] = fs[m] * T[k + 4 * l];
    }
}

// ---------------------------------------------------------------------------
// Hand-off between quad strips.  The producer's boundary row stores its
// outgoing contribution with an agent-coherent 64-bit store into a slot that
// holds a signalling-NaN sentinel (filled by k_fast_prep before every
// sweep); the consumer polls the slot until the sentinel is gone.  No flag
// and no store-completion wait is needed, so the producer never drains its
// load pipeline.  Arithmetic never produces this NaN payload.
// ---------------------------------------------------------------------------
__device__ inline bool is_sentinel(double x) { return (unsigned long long)__double_as_longlong(x) == SENTINEL; }

// wave-uniform: poll until the lanes `mine` read a non-sentinel value from
// their slot.  Past the bound it raises SHCFD_E_SPIN and goes on with zero,
// so the sweep still runs to its end (no early exit from the step loop).
__device__ inline void handoff_wait(double& x, const double* slotp, bool mine, long long spin_limit, int* iflag) {
    long long it = 0;
    while (__ballot(mine && is_sentinel(x)) != 0ULL) {
        if (it++ >= spin_limit) {
            atomicCAS(iflag + IFLAG_ERR, 0, SHCFD_E_SPIN);
            x = 0.0;
            break;
        }
        __builtin_amdgcn_s_sleep(2);
        x = load_agent(slotp);
    }
}

// wait for every outstanding vector memory access of the wave (vmcnt(0),
// other counters left alone)
__device__ inline void wait_all_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

constexpr int PF = 8;   // coefficient prefetch depth (steps); divides CH
constexpr int PH = 8;   // hand-off prefetch depth (steps); divides CH
constexpr int HM = 8;   // extra hand-off lead (columns) taken after a miss

// Sweep schedule.  Per step a wave issues 2+2 16-byte coefficient loads,
// half a 16-byte load of R/X, one hand-off load, half a 16-byte store of
// V/Z and one hand-off store: 7 vector memory instructions, so PF = 8 steps
// of prefetch stay within the 63 accesses a wave can keep in flight.  The
// chunk body is straight-line code (no bounds tests, no masks), so the
// compiler waits only for the data a step consumes.  A hand-off miss takes a
// rare uniform branch that polls, refills the ring and then drains all
// outstanding accesses, so it leaves no pending state behind for the common
// path.

// ---------------------------------------------------------------------------
// forward sweep: one wavefront per quad strip
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_fast_fwd(DevCtx d, const QStrip* strips, int nstrips, const double* R_sl, const double* GA, const double* GB,
           double* V_sl, double* hand, long long hstride, long long spin_limit) {
    const int s = blockIdx.x;
    const int L = threadIdx.x;
    const QStrip S = strips[s];
    const int ni = d.blk[S.b].ni;
    const int nch = S.nch;
    const bool has_next = S.next >= 0;
    const bool cons = L < 4;  // lanes of row 0: they take the hand-off
    const double4* ga_p = reinterpret_cast<const double4*>(GA) + S.base + L;
    const double4* gb_p = reinterpret_cast<const double4*>(GB) + S.base + L;
    const double2* r_p = reinterpret_cast<const double2*>(R_sl) + S.base / 2 + L;
    double2* v_p = reinterpret_cast<double2*>(V_sl) + S.base / 2 + L;
    // a strip with no neighbour above reads the row of zeros
    const double* hin = hand + (long long)(S.prev >= 0 ? S.prev : nstrips) * hstride + 4 * HOFF + (L & 3);
    double* hout = hand + (long long)s * hstride + 4 * HOFF + (L & 3);
    double4 gab[PF], gbb[PF];
    double2 rb[PF / 2];
    double hb[PH];
#pragma unroll
    for (int q = 0; q < PF; q++) {
        gab[q] = ga_p[64LL * q];
        gbb[q] = gb_p[64LL * q];
    }
#pragma unroll
    for (int q = 0; q < PF / 2; q++) rb[q] = r_p[64LL * q];
#pragma unroll
    for (int q = 0; q < PH; q++) hb[q] = load_agent(hin + 4LL * (q + 1));
    double a = 0.0, bo = 0.0, vlo = 0.0;
    for (int ch = 0; ch < nch; ch++) {
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int dd = ch * CH + u;
            const double4 ga = gab[u % PF], gb = gbb[u % PF];
            const double Rc = (u & 1) ? rb[(u / 2) % (PF / 2)].y : rb[(u / 2) % (PF / 2)].x;
            gab[u % PF] = ga_p[64LL * (dd + PF)];
            gbb[u % PF] = gb_p[64LL * (dd + PF)];
            if (u & 1) rb[(u / 2) % (PF / 2)] = r_p[64LL * ((dd + PF) >> 1)];
            const int i0 = dd + 1;  // column of row 0
            double x = hb[u % PH];
            if ((__builtin_amdgcn_ballot_w64(is_sentinel(x)) & 0xFULL) != 0ULL) {
                // the producer was not far enough ahead when the ring was
                // filled: let it get PH+HM columns ahead, refill the ring,
                // then drain so the common path inherits no pending loads
                const int ahead = min(i0 + PH + HM, ni);
                double y = load_agent(hin + 4LL * ahead);
                handoff_wait(y, hin + 4LL * ahead, cons, spin_limit, d.iflag);
                x = load_agent(hin + 4LL * i0);
#pragma unroll
                for (int m = 1; m < PH; m++) hb[(u + m) % PH] = load_agent(hin + 4LL * (i0 + m));
                handoff_wait(x, hin + 4LL * i0, cons, spin_limit, d.iflag);
                wait_all_vmem();
            }
            double bin = shift_up4(bo);
            if (cons) bin = x;
            const double v = (Rc + a) + bin;
            // refill the ring slot once x is dead, so the load lands in place
            hb[u % PH] = load_agent(hin + 4LL * (i0 + PH));
            if (u & 1)
                v_p[64LL * (dd >> 1)] = make_double2(vlo, v);
            else
                vlo = v;
            const double v0 = quad_bcast<0>(v), v1 = quad_bcast<1>(v), v2 = quad_bcast<2>(v), v3 = quad_bcast<3>(v);
            a = fma(ga.w, v3, fma(ga.z, v2, fma(ga.y, v1, ga.x * v0)));
            bo = fma(gb.w, v3, fma(gb.z, v2, fma(gb.y, v1, gb.x * v0)));
            // row 15 hands column dd-14 to the strip below
            if (has_next && L >= 60) store_agent(hout + 4LL * (dd - (QR - 2)), bo);
        }
    }
}

// ---------------------------------------------------------------------------
// between the sweeps (dual only): x2/alpha from v, in the skewed layout
// (time_dual_viscous.f90:44-60 applied to Dq* = alpha*S*v)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fast_mid(DevCtx d, const QStrip* strips, const int2* work, const double* V_sl,
                                                  double* X_sl) {
    const int2 wk = work[blockIdx.x];
    const QStrip S = strips[wk.x];
    const DevBlock B = d.blk[S.b];
    const int r = threadIdx.x & 15;
    const int dd = wk.y + (threadIdx.x >> 4);
    int i, j;
    if (!slot_cell(S, B, dd, r, i, j)) return;
    const long long c = cidx(B, i, j);
    CellJac<1> J;
    cell_jacobian<1>(d, c, B.pitch, J);
    double v[4], x[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = V_sl[slot1(S.base, dd, 4 * r + k)];
    // Dq* = alpha * S v
    double dv = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) dv = fma(J.pre2[k], v[k], dv);
    const double ips1 = 1.0 / J.ps[0];
#pragma unroll
    for (int k = 0; k < 4; k++) x[k] = J.alpha * (ips1 * fma(J.ps[3] * dv, J.pre1[k], v[k]));
    // second stage: x2 = ps1*(x + ps3 (pre2.x) pre1); store x2/alpha
    double dx = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) dx = fma(J.pre2[k], x[k], dx);
#pragma unroll
    for (int k = 0; k < 4; k++)
        X_sl[slot1(S.base, dd, 4 * r + k)] = (J.ps[0] * fma(J.ps[2] * dx, J.pre1[k], x[k])) / J.alpha;
}

// ---------------------------------------------------------------------------
// backward sweep: workgroup g handles strip (n-1-g), so the strip it waits
// for (the one above) has a lower workgroup id
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_fast_bwd(DevCtx d, const QStrip* strips, int nstrips, const double* X_sl, const double* HA, const double* HB,
           double* Z_sl, double* hand, long long hstride, long long spin_limit) {
    const int s = nstrips - 1 - (int)blockIdx.x;
    const int L = threadIdx.x;
    const QStrip S = strips[s];
    const int nch = S.nch;
    const bool has_prev = S.prev >= 0;
    const bool cons = L >= 60;  // lanes of row 15: they take the hand-off
    const double4* ha_p = reinterpret_cast<const double4*>(HA) + S.base + L;
    const double4* hb_p = reinterpret_cast<const double4*>(HB) + S.base + L;
    const double2* x_p = reinterpret_cast<const double2*>(X_sl) + S.base / 2 + L;
    double2* z_p = reinterpret_cast<double2*>(Z_sl) + S.base / 2 + L;
    // a strip with no neighbour below reads the row of zeros
    const double* hin = hand + (long long)(S.next >= 0 ? S.next : nstrips) * hstride + 4 * HOFF + (L & 3);
    double* hout = hand + (long long)s * hstride + 4 * HOFF + (L & 3);
    const int top = nch * CH - 1;  // first step of the sweep (steps descend)
    double4 hab[PF], hbb[PF];
    double2 xb[PF / 2];
    double hb[PH];
#pragma unroll
    for (int q = 0; q < PF; q++) {
        hab[q] = ha_p[64LL * (top - q)];
        hbb[q] = hb_p[64LL * (top - q)];
    }
#pragma unroll
    for (int q = 0; q < PF / 2; q++) xb[q] = x_p[64LL * ((top >> 1) - q)];
    // row 15's column at step dd is dd-14
#pragma unroll
    for (int q = 0; q < PH; q++) hb[q] = load_agent(hin + 4LL * (top - q - (QR - 2)));
    double a = 0.0, bo = 0.0, zhi = 0.0;
    for (int ch = 0; ch < nch; ch++) {
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int dd = top - (ch * CH + u);  // odd for even u
            const double4 ha = hab[u % PF], hbm = hbb[u % PF];
            const double Xc = (u & 1) ? xb[(u / 2) % (PF / 2)].x : xb[(u / 2) % (PF / 2)].y;
            hab[u % PF] = ha_p[64LL * (dd - PF)];
            hbb[u % PF] = hb_p[64LL * (dd - PF)];
            if (u & 1) xb[(u / 2) % (PF / 2)] = x_p[64LL * ((dd - PF) >> 1)];
            const int i15 = dd - (QR - 2);  // column of row 15
            double x = hb[u % PH];
            if ((__builtin_amdgcn_ballot_w64(is_sentinel(x)) >> 60) != 0ULL) {
                // resynchronise as in the forward sweep (columns descend)
                const int ahead = max(i15 - PH - HM, 1);
                double y = load_agent(hin + 4LL * ahead);
                handoff_wait(y, hin + 4LL * ahead, cons, spin_limit, d.iflag);
                x = load_agent(hin + 4LL * i15);
#pragma unroll
                for (int m = 1; m < PH; m++) hb[(u + m) % PH] = load_agent(hin + 4LL * (i15 - m));
                handoff_wait(x, hin + 4LL * i15, cons, spin_limit, d.iflag);
                wait_all_vmem();
            }
            double bin = shift_down4(bo);
            if (cons) bin = x;
            const double z = Xc - (a + bin);
            hb[u % PH] = load_agent(hin + 4LL * (i15 - PH));
            if (u & 1)
                z_p[64LL * (dd >> 1)] = make_double2(z, zhi);
            else
                zhi = z;
            const double z0 = quad_bcast<0>(z), z1 = quad_bcast<1>(z), z2 = quad_bcast<2>(z), z3 = quad_bcast<3>(z);
            a = fma(ha.w, z3, fma(ha.z, z2, fma(ha.y, z1, ha.x * z0)));
            bo = fma(hbm.w, z3, fma(hbm.z, z2, fma(hbm.y, z1, hbm.x * z0)));
            // row 0 hands column dd+1 to the strip above
            if (has_prev && L < 4) store_agent(hout + 4LL * (dd + 1), bo);
        }
    }
}

// ---------------------------------------------------------------------------
// post: Dq = alpha*S*z into the natural DQ field; dual: (res, omega) partials
// per workgroup; otherwise q += Dq
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fast_post(DevCtx d, const QStrip* strips, const int2* work, const double* Z_sl,
                                                   double* part) {
    __shared__ double shr[4], sho[4];
    const int2 wk = work[blockIdx.x];
    const QStrip S = strips[wk.x];
    const DevBlock B = d.blk[S.b];
    const int r = threadIdx.x & 15;
    const int dd = wk.y + (threadIdx.x >> 4);
    const bool dual = d.p.scheme == SHCFD_SCHEME_DUAL;
    double res = 0.0, om = 1e300;
    int i, j;
    if (slot_cell(S, B, dd, r, i, j)) {
        const long long c = cidx(B, i, j);
        double z[4], x[4];
#pragma unroll
        for (int k = 0; k < 4; k++) z[k] = Z_sl[slot1(S.base, dd, 4 * r + k)];
        if (dual) {
            CellJac<1> J;
            cell_jacobian<1>(d, c, B.pitch, J);
            double dz = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) dz = fma(J.pre2[k], z[k], dz);
            const double ips1 = 1.0 / J.ps[0];
#pragma unroll
            for (int k = 0; k < 4; k++) x[k] = J.alpha * (ips1 * fma(J.ps[3] * dz, J.pre1[k], z[k]));
            res = x[3] * x[3];
            if (x[0] < 0.0) om = (-FLD(d, SHCFD_F_Q, 0, c) / x[0]) * d.p.dqmax;
#pragma unroll
            for (int k = 0; k < 4; k++) FLD(d, SHCFD_F_DQ, k, c) = x[k];
        } else {
            const double al = FLD(d, SHCFD_F_ALPHA, 0, c);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                x[k] = al * z[k];
                FLD(d, SHCFD_F_DQ, k, c) = x[k];
                if (B.active) FLD(d, SHCFD_F_Q, k, c) = FLD(d, SHCFD_F_Q, k, c) + x[k];
            }
        }
    }
    if (!dual) return;
    for (int off = 32; off >= 1; off >>= 1) {
        res = res + __shfl_xor(res, off);
        om = dmin(om, __shfl_xor(om, off));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { shr[w] = res; sho[w] = om; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = ((shr[0] + shr[1]) + (shr[2] + shr[3]));
        part[2 * blockIdx.x + 1] = dmin(dmin(sho[0], sho[1]), dmin(sho[2], sho[3]));
    }
}

}  // namespace shcfd

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
#include "shcfd_fast.h"

namespace shcfd {

int FastPath::init(Ctx& c) {
    if (c.ny != 1) return SHCFD_E_STATE;
    std::vector<QStrip> hs;
    std::vector<int2> hw;
    long long base = 0;
    int maxcols = 0;
    for (int b = 0; b < c.nb; b++) {
        const DevBlock& B = c.hblk[b];
        int ns = (B.nj + QR - 1) / QR;
        int first = (int)hs.size();
        const int nd = B.ni + QR - 1;
        const int nch = (nd + CH - 1) / CH;
        maxcols = std::max(maxcols, nch * CH);
        for (int k = 0; k < ns; k++) {
            QStrip S{};
            S.b = b;
            S.j0 = 1 + QR * k;
            S.prev = k > 0 ? first + k - 1 : -1;
            S.next = k < ns - 1 ? first + k + 1 : -1;
            S.nch = nch;
            // one chunk of guard steps before and after the padded steps
            S.base = base + 64LL * CH;
            for (int d0 = 0; d0 < nd; d0 += 16) hw.push_back(make_int2((int)hs.size(), d0));
            base += 64LL * CH * (nch + 2);
            hs.push_back(S);
        }
    }
    nstrips = (int)hs.size();
    nwork = (int)hw.size();
    nslots = base;
    // columns -HOFF .. maxcols + PH + HM of every hand-off row
    hstride = 4LL * (HOFF + maxcols + PH + HM + 2);
    auto al = [&](void** p, size_t bytes) {
        if (hipMalloc(p, bytes) != hipSuccess) return false;
        return hipMemset(*p, 0, bytes) == hipSuccess;
    };
    bool ok = al((void**)&strips, sizeof(QStrip) * hs.size()) && al((void**)&work, sizeof(int2) * hw.size()) &&
              al((void**)&R, sizeof(double) * nslots) && al((void**)&V, sizeof(double) * nslots) &&
              al((void**)&X, sizeof(double) * nslots) && al((void**)&Z, sizeof(double) * nslots) &&
              al((void**)&GA, sizeof(double) * 4 * nslots) && al((void**)&GB, sizeof(double) * 4 * nslots) &&
              al((void**)&HA, sizeof(double) * 4 * nslots) && al((void**)&HB, sizeof(double) * 4 * nslots) &&
              al((void**)&hand, sizeof(double) * 2 * hstride * (nstrips + 1)) &&
              al((void**)&part, sizeof(double) * 2 * nwork);
    if (!ok) return SHCFD_E_NOMEM;
    hipMemcpy(strips, hs.data(), sizeof(QStrip) * hs.size(), hipMemcpyHostToDevice);
    hipMemcpy(work, hw.data(), sizeof(int2) * hw.size(), hipMemcpyHostToDevice);
    ready = true;
    return SHCFD_OK;
}

void FastPath::release() {
    void* ptrs[] = {strips, work, R, V, X, Z, GA, GB, HA, HB, hand, part};
    for (void* p : ptrs)
        if (p) hipFree(p);
    *this = FastPath();
}

int FastPath::implicit(Ctx& c, int step_internal, void (*pb)(Ctx*, int), void (*pe)(Ctx*, int)) {
    DevCtx d = c.dev();
    const bool dual = c.p.scheme == SHCFD_SCHEME_DUAL;
    hipStream_t st = c.stream;
    double* hand_f = hand;
    double* hand_b = hand + hstride * (nstrips + 1);
    pb(&c, PROF_PREP);
    hipLaunchKernelGGL(k_fast_prep, dim3(nwork), dim3(256), 0, st, d, strips, work, R, GA, GB, HA, HB, hand, nstrips,
                       hstride);
    pe(&c, PROF_PREP);
    pb(&c, PROF_SWEEP_FWD);
    hipLaunchKernelGGL(k_fast_fwd, dim3(nstrips), dim3(64), 0, st, d, strips, nstrips, R, GA, GB, V, hand_f, hstride,
                       c.spin_limit);
    pe(&c, PROF_SWEEP_FWD);
    const double* Xin = V;
    if (dual) {
        pb(&c, PROF_PREP);
        hipLaunchKernelGGL(k_fast_mid, dim3(nwork), dim3(256), 0, st, d, strips, work, V, X);
        pe(&c, PROF_PREP);
        Xin = X;
    }
    pb(&c, PROF_SWEEP_BWD);
    hipLaunchKernelGGL(k_fast_bwd, dim3(nstrips), dim3(64), 0, st, d, strips, nstrips, Xin, HA, HB, Z, hand_b, hstride,
                       c.spin_limit);
    pe(&c, PROF_SWEEP_BWD);
    pb(&c, PROF_UPDATE);
    hipLaunchKernelGGL(k_fast_post, dim3(nwork), dim3(256), 0, st, d, strips, work, Z, part);
    if (dual) {
        hipLaunchKernelGGL(k_reduce_res_omega, dim3(1), dim3(256), 0, st, d, part, nwork);
        if (step_internal == 1) hipLaunchKernelGGL(k_copy_scal, dim3(1), dim3(1), 0, st, d, (int)SCAL_RES1, (int)SCAL_RES);
        hipLaunchKernelGGL(k_update<1>, dim3((unsigned)c.htiles.size()), dim3(64 * c.tile_rows), 0, st, d, c.dtiles);
    }
    pe(&c, PROF_UPDATE);
    return hipGetLastError() == hipSuccess ? SHCFD_OK : SHCFD_E_HIP;
}

}  // namespace shcfd
