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
#include "shcfd_device.h"
#include "shcfd_kernels.h"

namespace shcfd {

constexpr int QR = 16;  // rows per quad strip

struct QStrip {
    int b, j0, prev, next;
    long long base;  // first slot of this strip in the skewed layout
};

// quad broadcast: every lane of a 4-lane group gets lane k of the group
template <int K>
__device__ inline double quad_bcast(double x) {
    constexpr int ctrl = K | (K << 2) | (K << 4) | (K << 6);
    int lo = __double2loint(x), hi = __double2hiint(x);
    lo = __builtin_amdgcn_update_dpp(0, lo, ctrl, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, ctrl, 0xF, 0xF, false);
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

__global__ void __launch_bounds__(256) k_fast_prep(DevCtx d, const QStrip* strips, const int2* work, double* R_sl,
                                                   double* GA, double* GB, double* HA, double* HB, double* hand,
                                                   long long nhand) {
    // refill the strip hand-off slots with the sentinel (both sweeps)
    {
        const long long stride = (long long)gridDim.x * blockDim.x;
        const double sv = __longlong_as_double((long long)SENTINEL);
        for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nhand; k += stride) hand[k] = sv;
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
    for (int k = 0; k < 4; k++) R_sl[slot + k] = FLD(d, SHCFD_F_DQ, k, c);
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

// wave-uniform: wait until lanes in `mask_lanes` hold a non-sentinel value
// read from their slot; returns false when the bound is exceeded
__device__ inline bool handoff_wait(double& x, const double* slotp, bool mine, long long spin_limit) {
    long long it = 0;
    while (__ballot(mine && is_sentinel(x)) != 0ULL) {
        if (it++ >= spin_limit) return false;
        __builtin_amdgcn_s_sleep(2);
        if (mine) x = load_agent(slotp);
    }
    return true;
}

constexpr int PF = 8;   // coefficient prefetch depth (steps)
constexpr int PH = 16;  // hand-off prefetch depth (steps)
constexpr int UNR = 16; // unroll (multiple of PF and PH)
constexpr int HM = 16;  // extra hand-off lead (columns) kept after a miss

// ---------------------------------------------------------------------------
// forward sweep: one wavefront per quad strip
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_fast_fwd(DevCtx d, const QStrip* strips, const double* R_sl, const double* GA,
                                                 const double* GB, double* V_sl, double* hand, long long hstride,
                                                 long long spin_limit) {
    const int s = blockIdx.x;
    const int L = threadIdx.x;
    const int r = L >> 2;
    const QStrip S = strips[s];
    const DevBlock B = d.blk[S.b];
    const int ni = B.ni;
    const bool rowok = S.j0 + r <= B.nj;
    const bool has_prev = S.prev >= 0, has_next = S.next >= 0;
    const int nd = ni + QR - 1;
    const double4* GA4 = reinterpret_cast<const double4*>(GA);
    const double4* GB4 = reinterpret_cast<const double4*>(GB);
    const bool cons = has_prev && L < 4;  // lanes of row 0 that receive
    const double* hin = hand + (has_prev ? (long long)S.prev * hstride : 0) + (L & 3);
    double* hout = hand + (long long)s * hstride + (L & 3);
    double Rb[PF];
    double4 gab[PF], gbb[PF];
    double hb[PH];
    const long long base = S.base + L;
#pragma unroll
    for (int q = 0; q < PF; q++) {
        const long long sl = base + 64LL * (q < nd ? q : 0);
        Rb[q] = R_sl[sl];
        gab[q] = GA4[sl];
        gbb[q] = GB4[sl];
    }
#pragma unroll
    for (int q = 0; q < PH; q++) hb[q] = (cons && q + 1 <= ni) ? load_agent(hin + 4LL * (q + 1)) : 0.0;
    double a = 0.0, bo = 0.0;
    for (int d0 = 0; d0 < nd; d0 += UNR) {
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int dd = d0 + u;
            if (dd < nd) {
                const double Rc = Rb[u % PF];
                const double4 ga = gab[u % PF], gb = gbb[u % PF];
                if (dd + PF < nd) {
                    const long long sl = base + 64LL * (dd + PF);
                    Rb[u % PF] = R_sl[sl];
                    gab[u % PF] = GA4[sl];
                    gbb[u % PF] = GB4[sl];
                }
                const int i = dd - r + 1;
                const bool act = rowok && i >= 1 && i <= ni;
                double bin = shift_up4(bo);
                const int i0 = dd + 1;  // column of row 0
                if (has_prev && i0 <= ni) {
                    double x = hb[u % PH];
                    if (__ballot(cons && is_sentinel(x)) != 0ULL) {
                        // the producer was not far enough ahead when the ring was
                        // filled: let it get PH+HM columns ahead, then refill the
                        // whole ring so the following steps do not miss again
                        const int ahead = min(i0 + PH + HM, ni);
                        double y = cons ? load_agent(hin + 4LL * ahead) : 0.0;
                        if (!handoff_wait(y, hin + 4LL * ahead, cons, spin_limit)) {
                            atomicCAS(d.iflag + IFLAG_ERR, 0, SHCFD_E_SPIN);
                            return;
                        }
                        x = cons ? load_agent(hin + 4LL * i0) : 0.0;
#pragma unroll
                        for (int m = 1; m < PH; m++)
                            hb[(u + m) % PH] = (cons && i0 + m <= ni) ? load_agent(hin + 4LL * (i0 + m)) : 0.0;
                    }
                    if (!handoff_wait(x, hin + 4LL * i0, cons, spin_limit)) {
                        atomicCAS(d.iflag + IFLAG_ERR, 0, SHCFD_E_SPIN);
                        return;
                    }
                    if (L < 4) bin = x;
                    const int ip = i0 + PH;
                    hb[u % PH] = (cons && ip <= ni) ? load_agent(hin + 4LL * ip) : 0.0;
                } else if (L < 4) {
                    bin = 0.0;
                }
                const double v = (Rc + a) + bin;
                if (act) V_sl[base + 64LL * dd] = v;
                const double v0 = quad_bcast<0>(v), v1 = quad_bcast<1>(v), v2 = quad_bcast<2>(v), v3 = quad_bcast<3>(v);
                const double na = fma(ga.w, v3, fma(ga.z, v2, fma(ga.y, v1, ga.x * v0)));
                const double nb = fma(gb.w, v3, fma(gb.z, v2, fma(gb.y, v1, gb.x * v0)));
                a = act ? na : 0.0;
                bo = act ? nb : 0.0;
                if (has_next && L >= 60 && act) store_agent(hout + 4LL * i, bo);
            }
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
    const long long slot = S.base + (long long)dd * 64 + 4 * r;
    double v[4], x[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = V_sl[slot + k];
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
    for (int k = 0; k < 4; k++) X_sl[slot + k] = (J.ps[0] * fma(J.ps[2] * dx, J.pre1[k], x[k])) / J.alpha;
}

// ---------------------------------------------------------------------------
// backward sweep: workgroup g handles strip (n-1-g), so the strip it waits
// for (the one above) has a lower workgroup id
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_fast_bwd(DevCtx d, const QStrip* strips, int nstrips, const double* X_sl,
                                                 const double* HA, const double* HB, double* Z_sl, double* hand,
                                                 long long hstride, long long spin_limit) {
    const int s = nstrips - 1 - (int)blockIdx.x;
    const int L = threadIdx.x;
    const int r = L >> 2;
    const QStrip S = strips[s];
    const DevBlock B = d.blk[S.b];
    const int ni = B.ni;
    const bool rowok = S.j0 + r <= B.nj;
    const bool has_prev = S.prev >= 0, has_next = S.next >= 0;
    const int nd = ni + QR - 1;
    const double4* HA4 = reinterpret_cast<const double4*>(HA);
    const double4* HB4 = reinterpret_cast<const double4*>(HB);
    const bool cons = has_next && L >= 60;  // lanes of row 15 that receive
    const double* hin = hand + (has_next ? (long long)S.next * hstride : 0) + (L & 3);
    double* hout = hand + (long long)s * hstride + (L & 3);
    double Xb[PF];
    double4 hab[PF], hbb[PF];
    double hb[PH];
    const long long base = S.base + L;
    // step t processes diagonal dd = nd-1-t; row 15's column is i15 = dd-14
#pragma unroll
    for (int q = 0; q < PF; q++) {
        const long long sl = base + 64LL * (q < nd ? nd - 1 - q : 0);
        Xb[q] = X_sl[sl];
        hab[q] = HA4[sl];
        hbb[q] = HB4[sl];
    }
#pragma unroll
    for (int q = 0; q < PH; q++) {
        const int ic = nd - 1 - q - (QR - 2);
        hb[q] = (cons && ic >= 1 && ic <= ni) ? load_agent(hin + 4LL * ic) : 0.0;
    }
    double a = 0.0, bo = 0.0;
    for (int t0 = 0; t0 < nd; t0 += UNR) {
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int t = t0 + u;
            if (t < nd) {
                const int dd = nd - 1 - t;
                const double Xc = Xb[u % PF];
                const double4 ha = hab[u % PF], hbm = hbb[u % PF];
                if (t + PF < nd) {
                    const long long sl = base + 64LL * (dd - PF);
                    Xb[u % PF] = X_sl[sl];
                    hab[u % PF] = HA4[sl];
                    hbb[u % PF] = HB4[sl];
                }
                const int i = dd - r + 1;
                const bool act = rowok && i >= 1 && i <= ni;
                double bin = shift_down4(bo);
                const int i15 = dd - (QR - 2);  // column of row 15
                if (has_next && i15 >= 1 && i15 <= ni) {
                    double x = hb[u % PH];
                    if (__ballot(cons && is_sentinel(x)) != 0ULL) {
                        // resynchronise as in the forward sweep (columns descend)
                        const int ahead = max(i15 - PH - HM, 1);
                        double y = cons ? load_agent(hin + 4LL * ahead) : 0.0;
                        if (!handoff_wait(y, hin + 4LL * ahead, cons, spin_limit)) {
                            atomicCAS(d.iflag + IFLAG_ERR, 0, SHCFD_E_SPIN);
                            return;
                        }
                        x = cons ? load_agent(hin + 4LL * i15) : 0.0;
#pragma unroll
                        for (int m = 1; m < PH; m++)
                            hb[(u + m) % PH] = (cons && i15 - m >= 1) ? load_agent(hin + 4LL * (i15 - m)) : 0.0;
                    }
                    if (!handoff_wait(x, hin + 4LL * i15, cons, spin_limit)) {
                        atomicCAS(d.iflag + IFLAG_ERR, 0, SHCFD_E_SPIN);
                        return;
                    }
                    if (L >= 60) bin = x;
                    const int ip = i15 - PH;
                    hb[u % PH] = (cons && ip >= 1) ? load_agent(hin + 4LL * ip) : 0.0;
                } else if (L >= 60) {
                    bin = 0.0;
                }
                const double z = Xc - (a + bin);
                if (act) Z_sl[base + 64LL * dd] = z;
                const double z0 = quad_bcast<0>(z), z1 = quad_bcast<1>(z), z2 = quad_bcast<2>(z), z3 = quad_bcast<3>(z);
                const double na = fma(ha.w, z3, fma(ha.z, z2, fma(ha.y, z1, ha.x * z0)));
                const double nb = fma(hbm.w, z3, fma(hbm.z, z2, fma(hbm.y, z1, hbm.x * z0)));
                a = act ? na : 0.0;
                bo = act ? nb : 0.0;
                if (has_prev && L < 4 && act) store_agent(hout + 4LL * i, bo);
            }
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
        const long long slot = S.base + (long long)dd * 64 + 4 * r;
        double z[4], x[4];
#pragma unroll
        for (int k = 0; k < 4; k++) z[k] = Z_sl[slot + k];
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
    int maxni = 0;
    for (int b = 0; b < c.nb; b++) {
        const DevBlock& B = c.hblk[b];
        maxni = std::max(maxni, B.ni);
        int ns = (B.nj + QR - 1) / QR;
        int first = (int)hs.size();
        for (int k = 0; k < ns; k++) {
            QStrip S{};
            S.b = b;
            S.j0 = 1 + QR * k;
            S.prev = k > 0 ? first + k - 1 : -1;
            S.next = k < ns - 1 ? first + k + 1 : -1;
            S.base = base;
            int nd = B.ni + QR - 1;
            for (int d0 = 0; d0 < nd; d0 += 16) hw.push_back(make_int2((int)hs.size(), d0));
            base += (long long)nd * 64;
            hs.push_back(S);
        }
    }
    nstrips = (int)hs.size();
    nwork = (int)hw.size();
    nslots = base;
    hstride = 4LL * (maxni + 2);
    auto al = [&](void** p, size_t bytes) {
        if (hipMalloc(p, bytes) != hipSuccess) return false;
        return hipMemset(*p, 0, bytes) == hipSuccess;
    };
    bool ok = al((void**)&strips, sizeof(QStrip) * hs.size()) && al((void**)&work, sizeof(int2) * hw.size()) &&
              al((void**)&R, sizeof(double) * nslots) && al((void**)&V, sizeof(double) * nslots) &&
              al((void**)&X, sizeof(double) * nslots) && al((void**)&Z, sizeof(double) * nslots) &&
              al((void**)&GA, sizeof(double) * 4 * nslots) && al((void**)&GB, sizeof(double) * 4 * nslots) &&
              al((void**)&HA, sizeof(double) * 4 * nslots) && al((void**)&HB, sizeof(double) * 4 * nslots) &&
              al((void**)&hand, sizeof(double) * 2 * hstride * nstrips) && al((void**)&part, sizeof(double) * 2 * nwork);
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
    double* hand_b = hand + hstride * nstrips;
    pb(&c, PROF_PREP);
    hipLaunchKernelGGL(k_fast_prep, dim3(nwork), dim3(256), 0, st, d, strips, work, R, GA, GB, HA, HB, hand,
                       2 * hstride * nstrips);
    pe(&c, PROF_PREP);
    pb(&c, PROF_SWEEP_FWD);
    hipLaunchKernelGGL(k_fast_fwd, dim3(nstrips), dim3(64), 0, st, d, strips, R, GA, GB, V, hand_f, hstride,
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
