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
// are zeroed once and k_fast_rhs_prep writes only cells), which makes every
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

// alpha and the dual-time vectors/scalars of a cell as set_ABpm stored them
// (ALPHA, PRE1, PRE2, PRESCA: the values cell_jacobian gives for the same
// state, which the stage kernels read instead of recomputing)
struct PostScal {
    double alpha, pre1[4], pre2[4], ps[4];
};
__device__ inline PostScal post_scal(const DevCtx& d, long long c) {
    PostScal J;
    J.alpha = FLD(d, SHCFD_F_ALPHA, 0, c);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        J.pre1[k] = FLD(d, SHCFD_F_PRE1, k, c);
        J.pre2[k] = FLD(d, SHCFD_F_PRE2, k, c);
        J.ps[k] = FLD(d, SHCFD_F_PRESCA, k, c);
    }
    return J;
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

// Operator arrays G = (GA, GB) and H = (HA, HB): per step a 4 KB block of
// four 1 KB parts, part p holding doubles (2p', 2p'+1) of every lane's row,
// p = 2*(0 for A, 1 for B) + p'; lane L's pair sits at 16 L in its part.
// Each 16-byte load of a sweep step then reads 1 KB of consecutive bytes.
__device__ inline long long op_index(const QStrip& S, int dd, int L, int p, int e) {
    return (S.base / 64 + dd) * 512 + p * 128 + 2 * L + e;
}

// the folded operators of one cell c (slot step dd, row r) in its skewed
// slots from its cell Jacobian J: (GA, GB) into G for the forward sweep when
// G is given, (HA, HB) into H for the backward sweep when H is given, and
// with S = I the backward sweep's diagonal terms into Dn when Dn is given:
// A- = A+ - nuA I and B- = B+ - nuB I, so HA = GA - dsci alpha nuA I and
// HB = GB - dscj alpha nuB I; Dn holds (-dsci alpha nuA, -dscj alpha nuB) per
// row and step, 256 bytes per step
__device__ inline void prep_ops(const DevCtx& d, const QStrip& S, int dd, int r, long long c, const CellJac<1>& J,
                                double* G, double* H, double* Dn) {
    const bool dual = d.p.scheme == SHCFD_SCHEME_DUAL;
    double Sm[16], T[16];
    post_matrix(J, dual, Sm);
    const double dsci = FLD(d, SHCFD_F_DSCI, 0, c), dscj = FLD(d, SHCFD_F_DSCJ, 0, c);
    const double fa = dsci * J.alpha, fb = dscj * J.alpha;
    if (Dn)
        *reinterpret_cast<double2*>(&Dn[(S.base / 64 + dd) * 32 + 2 * r]) =
            make_double2(-(fa * J.nuA), -(fb * J.nuB));
    // rows of the operators: lane 4r+k holds row k (4 doubles)
    const double* mats[4] = {J.Ap, J.Bp, J.Am, J.Bm};
    const double fs[2] = {fa, fb};
#pragma unroll
    for (int m = 0; m < 4; m++) {
        double* o = m < 2 ? G : H;
        if (!o) continue;
        mm4(mats[m], Sm, T);
        // doubles (2p', 2p'+1) of a lane's row are adjacent: one 16-byte store
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int lp = 0; lp < 2; lp++)
                *reinterpret_cast<double2*>(&o[op_index(S, dd, 4 * r + k, 2 * (m & 1) + lp, 0)]) =
                    make_double2(fs[m & 1] * T[k + 4 * (2 * lp)], fs[m & 1] * T[k + 4 * (2 * lp + 1)]);
    }
}

// one cell's backward operators (HA, HB) from a new cell Jacobian
__device__ inline void prep_cell_h(const DevCtx& d, const QStrip& S, const DevBlock& B, int dd, int r, double* H) {
    int i, j;
    if (!slot_cell(S, B, dd, r, i, j)) return;
    const long long c = cidx(B, i, j);
    CellJac<1> J;
    cell_jacobian<1>(d, c, B.pitch, J);
    prep_ops(d, S, dd, r, c, J, nullptr, H, nullptr);
}

// XCD-aware work order.  The dispatcher hands a launch's workgroups to the
// 8 XCDs in turn (workgroup b runs on XCD b % 8).  Giving XCD x the
// contiguous work items x*per .. x*per + per - 1 keeps neighbouring items,
// which share cache lines at their edges, in one L2.  A launch then has
// 8*per workgroups; those past the last item do nothing but the refill.
constexpr int NXCD = 8;
__host__ __device__ inline int xcd_groups(int n) { return NXCD * ((n + NXCD - 1) / NXCD); }
__device__ inline int xcd_item(int b, int n) { return (b % NXCD) * ((n + NXCD - 1) / NXCD) + b / NXCD; }

// refill of the hand-off rows of both sweeps: the sentinel for the columns
// 1..ni of each strip's block, zero elsewhere (grid-stride over all rows)
__device__ inline void hand_refill(const DevCtx& d, const QStrip* strips, double* hand, int nstrips, long long hstride) {
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

// The right-hand side and prep of the optimised update in one pass (k_rhs
// followed by the operators' prep): per cell one cell Jacobian gives
// set_ABpm's scalars (stored when abpm_inline, as k_rhs stores them), the
// right-hand side of the scheme goes straight into R, and the folded
// operators follow; the hand-off rows are refilled.  The natural DQ field is
// left to k_fast_post, which stores the update there.  (HA, HB) are formed
// here too when H is given, or their diagonal terms when Dn is given (S =
// I), else by the forward sweep's launch on the CUs its strips leave idle
// (k_fast_fwd).  Workgroup = one work item (16 steps x 16 rows of one quad
// strip, in XCD-aware order); a wave takes 4 rows x 16 steps, so each load of a cell field reads
// 4 runs of 16 consecutive cells.
__global__ void __launch_bounds__(256) k_fast_rhs_prep(DevCtx d, const QStrip* strips, const int2* work, int nwork,
                                                       int dqdt_inline, int abpm_inline, double* R_sl, double* G,
                                                       double* H, double* Dn, double* hand, int nstrips,
                                                       long long hstride) {
    hand_refill(d, strips, hand, nstrips, hstride);
    const int w = xcd_item((int)blockIdx.x, nwork);
    if (w >= nwork) return;
    const int2 wk = work[w];
    const QStrip S = strips[wk.x];
    const DevBlock B = d.blk[S.b];
    const int dd = wk.y + (int)(threadIdx.x & 15), r = (int)(threadIdx.x >> 4);
    int i, j;
    if (!slot_cell(S, B, dd, r, i, j)) return;
    const long long c = cidx(B, i, j);
    CellJac<1> J;
    cell_jacobian<1>(d, c, B.pitch, J);
    double s[4], x[4];
    if (abpm_inline) {
        store_abpm_scalars<1>(d, c, J);
        flux_balance<1>(d, B, c, s);
        rhs_of_cell<1>(
            d, c, s, dqdt_inline != 0, [&](int k) { return J.pre1[k]; }, [&](int k) { return J.pre2[k]; },
            [&]() { return J.ps[1]; }, x);
    } else {
        flux_balance<1>(d, B, c, s);
        rhs_of_cell<1>(
            d, c, s, dqdt_inline != 0, [&](int k) { return FLD(d, SHCFD_F_PRE1, k, c); },
            [&](int k) { return FLD(d, SHCFD_F_PRE2, k, c); }, [&]() { return FLD(d, SHCFD_F_PRESCA, 1, c); }, x);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) R_sl[slot1(S.base, dd, 4 * r + k)] = x[k];
    prep_ops(d, S, dd, r, c, J, G, H, Dn);
}

// ---------------------------------------------------------------------------
// Hand-off between quad strips.  The producer's boundary row stores its
// outgoing contribution with an agent-coherent 64-bit store into a slot that
// holds a signalling-NaN sentinel (filled by k_fast_rhs_prep before every
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


// Buffer views.  Each sweep addresses its strip through buffer resources
// (base in scalar registers, 32-bit lane offset, per-step offsets in scalar
// registers or the instruction's immediate), so a step spends one
// instruction per access instead of 64-bit address arithmetic per lane.
// Lanes that must not store give an offset beyond the view's size: the
// hardware's range check drops the access.
typedef unsigned int shcfd_u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int shcfd_u32x2 __attribute__((ext_vector_type(2)));
constexpr int BUF_DROP = 0x7FFFFF00;     // lane offset beyond every view
constexpr int AUX_AGENT = 16;            // sc1: device-coherent access
__device__ inline __amdgpu_buffer_rsrc_t view(const double* base, long long ndoubles) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), 0, (int)(ndoubles * 8), 0x00020000);
}
__device__ inline double u2d(unsigned lo, unsigned hi) {
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline double2 ld_pair(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    const shcfd_u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
    return make_double2(u2d(v.x, v.y), u2d(v.z, v.w));
}
__device__ inline void st_pair(__amdgpu_buffer_rsrc_t r, double2 x, int voff, int soff) {
    const unsigned long long a = (unsigned long long)__double_as_longlong(x.x);
    const unsigned long long b = (unsigned long long)__double_as_longlong(x.y);
    const shcfd_u32x4 v = {(unsigned)a, (unsigned)(a >> 32), (unsigned)b, (unsigned)(b >> 32)};
    __builtin_amdgcn_raw_buffer_store_b128(v, r, voff, soff, 0);
}
__device__ inline double ld_agent(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    const shcfd_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX_AGENT);
    return u2d(v.x, v.y);
}
__device__ inline void st_agent(__amdgpu_buffer_rsrc_t r, double x, int voff, int soff) {
    const unsigned long long a = (unsigned long long)__double_as_longlong(x);
    const shcfd_u32x2 v = {(unsigned)a, (unsigned)(a >> 32)};
    __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, AUX_AGENT);
}
// byte offsets of step dd (dd >= -CH) inside a strip's views: operator
// blocks of 4 KB per step, step pairs of 1 KB (dd even), hand-off columns of
// 32 bytes
__device__ inline int op_off(int dd) { return (dd + CH) * 4096; }
__device__ inline int pair_off(int dd) { return (dd + CH) * 512; }
__device__ inline int dn_off(int dd) { return (dd + CH) * 256; }
__device__ inline int col_off(int col) { return (col + HOFF) * 32; }

// Sweep schedule.  A sweep workgroup has two wavefronts on one CU: wave 0
// computes, wave 1 streams the strip's operator blocks and R/X pairs into LDS
// with LDS-DMA loads (buffer_load ... lds), CS steps per chunk, NSLOT chunk
// slots.  The vector-memory path of a CU moves about 64 bytes per cycle, and
// a step needs 4.5 KB of operators: issued by the computing wave, those loads
// stall its dependency chain; issued by a second wave they overlap it.  The
// waves meet at one s_barrier per chunk: before it the loader has waited for
// the chunk the compute wave reads next, and the compute wave has finished
// reading the slot the loader refills next.  The compute wave keeps the
// hand-off ring (PH steps of device-coherent loads), its V/Z stores and the
// hand-off stores.  Its step body is straight-line code; a hand-off miss takes
// a rare uniform branch that polls the one column and then drains all
// outstanding accesses, so it leaves no pending state for the common path.
constexpr int CS = 4;                                   // steps per LDS chunk
constexpr int NSLOT = 3;                                // LDS chunk slots
constexpr int SLOT_G = CS * 4096;                       // operator bytes per slot
constexpr int SLOT_DN = SLOT_G + (CS / 2) * 1024;       // + R/X step pairs
constexpr int SLOT_BYTES = SLOT_DN + CS * 256;          // + diagonal terms (backward sweep with S = I)
constexpr int LD_PER_CHUNK = 4 * CS + CS / 2;           // LDS-DMA loads per chunk, + 1 with Dn
static_assert(CS * 256 == 1024, "one LDS-DMA load carries a chunk's diagonal terms");

typedef __attribute__((address_space(3))) void lds_void;

__device__ inline void lds_dma(__amdgpu_buffer_rsrc_t r, char* lds, int voff, int soff) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_void*)lds, 16, voff, soff, 0, 0);
}
// s_barrier that the compiler may not move memory accesses across
__device__ inline void wg_barrier() { asm volatile("s_barrier" ::: "memory"); }
// wait until at most N vector memory accesses of this wave are outstanding
template <int N>
__device__ inline void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | 0x0F70);
}
__device__ inline void wait_lgkm0() { __builtin_amdgcn_s_waitcnt(0xC07F); }

// loader wave: chunk j = aligned block m of CS steps (steps CS*m .. CS*m+CS-1)
// into slot j % NSLOT; the caller orders blocks (ascending or descending)
template <bool DN>
__device__ inline void load_block(__amdgpu_buffer_rsrc_t rop, __amdgpu_buffer_rsrc_t rpair, __amdgpu_buffer_rsrc_t rdn,
                                  char* lds, int j, int m, int vl) {
    char* slot = lds + (j % NSLOT) * SLOT_BYTES;
#pragma unroll
    for (int t = 0; t < CS; t++)
#pragma unroll
        for (int p = 0; p < 4; p++) lds_dma(rop, slot + t * 4096 + p * 1024, vl, op_off(CS * m + t) + 1024 * p);
#pragma unroll
    for (int q = 0; q < CS / 2; q++) lds_dma(rpair, slot + SLOT_G + q * 1024, vl, pair_off(CS * m + 2 * q));
    if constexpr (DN) lds_dma(rdn, slot + SLOT_DN, vl, dn_off(CS * m));
}

// loader side of the chunk protocol: nck chunks, block of chunk j = blk(j);
// DN: the chunks also carry the backward sweep's diagonal terms
template <bool DN, typename BLK>
__device__ inline void loader(__amdgpu_buffer_rsrc_t rop, __amdgpu_buffer_rsrc_t rpair, __amdgpu_buffer_rsrc_t rdn,
                              char* lds, int nck, int vl, BLK blk) {
    constexpr int LD = LD_PER_CHUNK + (DN ? 1 : 0);
    load_block<DN>(rop, rpair, rdn, lds, 0, blk(0), vl);
    load_block<DN>(rop, rpair, rdn, lds, 1, blk(1), vl);
    load_block<DN>(rop, rpair, rdn, lds, 2, blk(2), vl);
    wait_vmcnt<2 * LD>();  // chunk 0 in LDS
    wg_barrier();
    wait_vmcnt<LD>();      // chunk 1
    wg_barrier();
    for (int j = 0; j < nck; j++) {
        // chunk j+3 takes the slot of chunk j, read before the last barrier;
        // the blocks past the strip's end fall in its guard steps
        load_block<DN>(rop, rpair, rdn, lds, j + 3, blk(j + 3), vl);
        wait_vmcnt<LD>();  // chunk j+2
        wg_barrier();
    }
}

// compute side: one chunk's operators and R/X pair(s) from LDS
struct Chunk {
    double2 g[CS][4];
    double2 r[CS / 2];
    double2 dn[CS];  // diagonal terms of the lane's row (DN only)
};
template <bool DN = false>
__device__ inline void read_chunk(Chunk& c, const char* lds, int j, int vl) {
    const char* base = lds + (j % NSLOT) * SLOT_BYTES;
    const char* slot = base + vl;
#pragma unroll
    for (int t = 0; t < CS; t++)
#pragma unroll
        for (int p = 0; p < 4; p++) c.g[t][p] = *reinterpret_cast<const double2*>(slot + t * 4096 + p * 1024);
#pragma unroll
    for (int q = 0; q < CS / 2; q++) c.r[q] = *reinterpret_cast<const double2*>(slot + SLOT_G + q * 1024);
    if constexpr (DN) {
        // row r = lane/4 of step t: 16 bytes at t*256 + 16 r (four lanes share it)
        const char* dn = base + SLOT_DN + 4 * (vl / 16 & ~3);
#pragma unroll
        for (int t = 0; t < CS; t++) c.dn[t] = *reinterpret_cast<const double2*>(dn + t * 256);
    }
}

constexpr int PH = 4;   // hand-off prefetch depth (steps); divides 2*CS
constexpr int HM = 4;   // extra hand-off lead (columns) awaited before the start

__shared__ __attribute__((aligned(16))) char sweep_lds[NSLOT * SLOT_BYTES];

// one forward step (lane L: component L&3 of row L>>2): step dd, slot step t
// of chunk c, ring position u (compile-time after unrolling)
__device__ inline void fwd_step(const Chunk& c, int t, int u, int dd, double& a, double& bo, double& vlo,
                                double (&hb)[PH],
                                bool cons, const double* hin, __amdgpu_buffer_rsrc_t rin,
                                __amdgpu_buffer_rsrc_t rv, __amdgpu_buffer_rsrc_t rout, int vl, int vc, int vo,
                                long long spin_limit, int* iflag) {
    const double2 ga01 = c.g[t][0], ga23 = c.g[t][1], gb01 = c.g[t][2], gb23 = c.g[t][3];
    const double Rc = (t & 1) ? c.r[t >> 1].y : c.r[t >> 1].x;
    const int i0 = dd + 1;  // column of row 0
    double x = hb[u % PH];
    if (__builtin_expect((__builtin_amdgcn_ballot_w64(is_sentinel(x)) & 0xFULL) != 0ULL, 0)) {
        // the producer fell behind after the start: poll this column only
        // (a ring refill here would make the compiler copy the ring on the
        // common path), then drain so the common path inherits no pending
        // loads
        handoff_wait(x, hin + 4LL * i0, cons, spin_limit, iflag);
        wait_all_vmem();
    }
    double bin = shift_up4(bo);
    if (cons) bin = x;
    const double v = (Rc + a) + bin;
    // refill the ring slot once x is dead, so the load lands in place
    hb[u % PH] = ld_agent(rin, vc, col_off(i0 + PH));
    if (t & 1)
        st_pair(rv, make_double2(vlo, v), vl, pair_off(dd - 1));
    else
        vlo = v;
    const double v0 = quad_bcast<0>(v), v1 = quad_bcast<1>(v), v2 = quad_bcast<2>(v), v3 = quad_bcast<3>(v);
    a = fma(ga23.y, v3, fma(ga23.x, v2, fma(ga01.y, v1, ga01.x * v0)));
    bo = fma(gb23.y, v3, fma(gb23.x, v2, fma(gb01.y, v1, gb01.x * v0)));
    // row 15 hands column dd-14 to the strip below
    st_agent(rout, bo, vo, col_off(dd - (QR - 2)));
}

// ---------------------------------------------------------------------------
// forward sweep: one workgroup (compute wave + loader wave) per quad strip
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_fast_fwd(DevCtx d, const QStrip* strips, int nstrips, const double* R_sl, const double* G, double* V_sl,
           double* hand, long long hstride, long long spin_limit, const int2* work, int nwork, double* H) {
    if ((int)blockIdx.x >= nstrips) {
        // workgroups past the strips form the backward operators (HA, HB) of
        // the work items, concurrently with the sweep: 16 steps x 16 rows per
        // item, two cells per thread.  They wait on nothing, and the strips
        // (lower workgroup ids) are dispatched before them.
        const int nh = (int)gridDim.x - nstrips;
        for (int w = (int)blockIdx.x - nstrips; w < nwork; w += nh) {
            const int2 wk = work[w];
            const QStrip S = strips[wk.x];
            const DevBlock B = d.blk[S.b];
#pragma unroll 1
            for (int h = 0; h < 2; h++) {
                const int t = (int)threadIdx.x + 128 * h;
                prep_cell_h(d, S, B, wk.y + (t & 15), t >> 4, H);
            }
        }
        return;
    }
    const int s = blockIdx.x;
    const int L = threadIdx.x & 63;
    const QStrip S = strips[s];
    const int ni = d.blk[S.b].ni;
    const int nck = S.nch * (CH / CS);  // chunks, processed in ascending order
    const long long nstep = (long long)(S.nch + 2) * CH;
    const int vl = 16 * L;
    if (threadIdx.x >= 64) {
        const __amdgpu_buffer_rsrc_t rg = view(G + (S.base / 64 - CH) * 512, nstep * 512);
        const __amdgpu_buffer_rsrc_t rr = view(R_sl + S.base - 64 * CH, nstep * 64);
        loader<false>(rg, rr, rr, sweep_lds, nck, vl, [](int j) { return j; });
        return;
    }
    const bool cons = L < 4;  // lanes of row 0: they take the hand-off
    const __amdgpu_buffer_rsrc_t rv = view(V_sl + S.base - 64 * CH, nstep * 64);
    // a strip with no neighbour above reads the row of zeros
    const long long prow = S.prev >= 0 ? S.prev : nstrips;
    const __amdgpu_buffer_rsrc_t rin = view(hand + prow * hstride, hstride);
    const __amdgpu_buffer_rsrc_t rout = view(hand + (long long)s * hstride, hstride);
    const double* hin = hand + prow * hstride + 4 * HOFF + (L & 3);  // for polling
    const int vc = 8 * (L & 3);
    const int vo = (S.next >= 0 && L >= 60) ? vc : BUF_DROP;  // row 15 hands off
    // start once the strip above is PH+HM columns ahead, so that the ring
    // loads find their columns already written
    {
        const int ahead = min(PH + HM, ni);
        double y = load_agent(hin + 4LL * ahead);
        handoff_wait(y, hin + 4LL * ahead, cons, spin_limit, d.iflag);
        wait_all_vmem();
    }
    double hb[PH];
#pragma unroll
    for (int q = 0; q < PH; q++) hb[q] = ld_agent(rin, vc, col_off(q + 1));
    // drain before the loop: the wait counts at the loop head then follow the
    // back edge alone (in-order vmcnt: with these loads still in flight on
    // entry, every step would also wait for the stores of the last steps)
    wait_all_vmem();
    double a = 0.0, bo = 0.0, vlo = 0.0;
    Chunk A, B;
    wg_barrier();  // chunk 0 in LDS
    read_chunk(A, sweep_lds, 0, vl);
    wait_lgkm0();
    wg_barrier();  // chunk 1 in LDS
    for (int j = 0; j < nck; j += 2) {
        read_chunk(B, sweep_lds, j + 1, vl);
#pragma unroll
        for (int t = 0; t < CS; t++)
            fwd_step(A, t, t, CS * j + t, a, bo, vlo, hb, cons, hin, rin, rv, rout, vl, vc, vo, spin_limit,
                     d.iflag);
        wait_lgkm0();
        wg_barrier();
        read_chunk(A, sweep_lds, j + 2, vl);
#pragma unroll
        for (int t = 0; t < CS; t++)
            fwd_step(B, t, CS + t, CS * (j + 1) + t, a, bo, vlo, hb, cons, hin, rin, rv, rout, vl, vc, vo,
                     spin_limit, d.iflag);
        wait_lgkm0();
        wg_barrier();
    }
}

// ---------------------------------------------------------------------------
// between the sweeps (dual only): x2/alpha from v, in the skewed layout
// (time_dual_viscous.f90:44-60 applied to Dq* = alpha*S*v)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fast_mid(DevCtx d, const QStrip* strips, const int2* work, int nwork,
                                                  const double* V_sl, double* X_sl) {
    // work items in XCD-aware order; a wave takes 4 rows x 16 steps
    const int w = xcd_item((int)blockIdx.x, nwork);
    if (w >= nwork) return;
    const int2 wk = work[w];
    const QStrip S = strips[wk.x];
    const DevBlock B = d.blk[S.b];
    const int r = (int)(threadIdx.x >> 4);
    const int dd = wk.y + (int)(threadIdx.x & 15);
    int i, j;
    if (!slot_cell(S, B, dd, r, i, j)) return;
    const long long c = cidx(B, i, j);
    // the dual-time scalars and vectors set_ABpm stored for this state
    const PostScal J = post_scal(d, c);
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
// one backward step: step dd, slot step t of chunk c, ring position u.  DN:
// the operators are (GA, GB) and the step subtracts the diagonal terms
// (HA z = GA z - dsci alpha nuA z, the term entering as the first addend)
template <bool DN>
__device__ inline void bwd_step(const Chunk& c, int t, int u, int dd, double& a, double& bo, double& zhi,
                                double (&hb)[PH], bool cons, const double* hin, __amdgpu_buffer_rsrc_t rin,
                                __amdgpu_buffer_rsrc_t rz, __amdgpu_buffer_rsrc_t rout, int vl, int vc, int vo,
                                long long spin_limit, int* iflag) {
    const double2 ha01 = c.g[t][0], ha23 = c.g[t][1], hb01 = c.g[t][2], hb23 = c.g[t][3];
    const double Xc = (t & 1) ? c.r[t >> 1].y : c.r[t >> 1].x;
    const int i15 = dd - (QR - 2);  // column of row 15
    double x = hb[u % PH];
    if (__builtin_expect((__builtin_amdgcn_ballot_w64(is_sentinel(x)) >> 60) != 0ULL, 0)) {
        // the producer fell behind: poll this column only
        handoff_wait(x, hin + 4LL * i15, cons, spin_limit, iflag);
        wait_all_vmem();
    }
    double bin = shift_down4(bo);
    if (cons) bin = x;
    const double z = Xc - (a + bin);
    hb[u % PH] = ld_agent(rin, vc, col_off(i15 - PH));
    if (t & 1)
        zhi = z;
    else
        st_pair(rz, make_double2(z, zhi), vl, pair_off(dd));
    const double z0 = quad_bcast<0>(z), z1 = quad_bcast<1>(z), z2 = quad_bcast<2>(z), z3 = quad_bcast<3>(z);
    if constexpr (DN) {
        const double pa = c.dn[t].x * z, pb = c.dn[t].y * z;
        a = fma(ha23.y, z3, fma(ha23.x, z2, fma(ha01.y, z1, fma(ha01.x, z0, pa))));
        bo = fma(hb23.y, z3, fma(hb23.x, z2, fma(hb01.y, z1, fma(hb01.x, z0, pb))));
    } else {
        a = fma(ha23.y, z3, fma(ha23.x, z2, fma(ha01.y, z1, ha01.x * z0)));
        bo = fma(hb23.y, z3, fma(hb23.x, z2, fma(hb01.y, z1, hb01.x * z0)));
    }
    // row 0 hands column dd+1 to the strip above
    st_agent(rout, bo, vo, col_off(dd + 1));
}

// DN: H holds (GA, GB) and Dn the diagonal terms (schemes with S = I)
template <bool DN>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_fast_bwd(DevCtx d, const QStrip* strips, int nstrips, const double* X_sl, const double* H, const double* Dn,
           double* Z_sl, double* hand, long long hstride, long long spin_limit) {
    const int s = nstrips - 1 - (int)blockIdx.x;
    const int L = threadIdx.x & 63;
    const QStrip S = strips[s];
    const int nck = S.nch * (CH / CS);  // chunk j holds block nck-1-j (steps descend)
    const long long nstep = (long long)(S.nch + 2) * CH;
    const int vl = 16 * L;
    if (threadIdx.x >= 64) {
        const __amdgpu_buffer_rsrc_t rh = view(H + (S.base / 64 - CH) * 512, nstep * 512);
        const __amdgpu_buffer_rsrc_t rx = view(X_sl + S.base - 64 * CH, nstep * 64);
        const __amdgpu_buffer_rsrc_t rd = DN ? view(Dn + (S.base / 64 - CH) * 32, nstep * 32) : rx;
        loader<DN>(rh, rx, rd, sweep_lds, nck, vl, [nck](int j) { return nck - 1 - j; });
        return;
    }
    const bool cons = L >= 60;  // lanes of row 15: they take the hand-off
    const __amdgpu_buffer_rsrc_t rz = view(Z_sl + S.base - 64 * CH, nstep * 64);
    // a strip with no neighbour below reads the row of zeros
    const long long nrow = S.next >= 0 ? S.next : nstrips;
    const __amdgpu_buffer_rsrc_t rin = view(hand + nrow * hstride, hstride);
    const __amdgpu_buffer_rsrc_t rout = view(hand + (long long)s * hstride, hstride);
    const double* hin = hand + nrow * hstride + 4 * HOFF + (L & 3);  // for polling
    const int vc = 8 * (L & 3);
    const int vo = (S.prev >= 0 && L < 4) ? vc : BUF_DROP;  // row 0 hands off
    const int top = nck * CS - 1;  // first step of the sweep
    // start once the strip below is PH+HM columns ahead (columns descend
    // from ni; the padded columns above ni are zero)
    {
        const int ahead = max(d.blk[S.b].ni - PH - HM, 1);
        double y = load_agent(hin + 4LL * ahead);
        handoff_wait(y, hin + 4LL * ahead, cons, spin_limit, d.iflag);
        wait_all_vmem();
    }
    // row 15's column at step dd is dd-14
    double hb[PH];
#pragma unroll
    for (int q = 0; q < PH; q++) hb[q] = ld_agent(rin, vc, col_off(top - q - (QR - 2)));
    wait_all_vmem();  // as in the forward sweep
    double a = 0.0, bo = 0.0, zhi = 0.0;
    Chunk A, B;
    wg_barrier();
    read_chunk<DN>(A, sweep_lds, 0, vl);
    wait_lgkm0();
    wg_barrier();
    for (int j = 0; j < nck; j += 2) {
        read_chunk<DN>(B, sweep_lds, j + 1, vl);
        const int m0 = nck - 1 - j;  // block of chunk j
#pragma unroll
        for (int u = 0; u < CS; u++)
            bwd_step<DN>(A, CS - 1 - u, u, CS * m0 + CS - 1 - u, a, bo, zhi, hb, cons, hin, rin, rz, rout, vl, vc, vo,
                     spin_limit, d.iflag);
        wait_lgkm0();
        wg_barrier();
        read_chunk<DN>(A, sweep_lds, j + 2, vl);
#pragma unroll
        for (int u = 0; u < CS; u++)
            bwd_step<DN>(B, CS - 1 - u, CS + u, CS * (m0 - 1) + CS - 1 - u, a, bo, zhi, hb, cons, hin, rin, rz, rout, vl,
                     vc, vo, spin_limit, d.iflag);
        wait_lgkm0();
        wg_barrier();
    }
}

// ---------------------------------------------------------------------------
// post: Dq = alpha*S*z into the natural DQ field; inner-loop schemes: (res,
// omega) partials per workgroup; otherwise q += Dq
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fast_post(DevCtx d, const QStrip* strips, const int2* work, int nwork,
                                                   const double* Z_sl, double* part) {
    __shared__ double shr[4], sho[4];
    // work items in XCD-aware order; the partials are stored per item, so
    // their summation order does not depend on it
    const int wi = xcd_item((int)blockIdx.x, nwork);
    if (wi >= nwork) return;
    const int2 wk = work[wi];
    const QStrip S = strips[wk.x];
    const DevBlock B = d.blk[S.b];
    const bool dual = d.p.scheme == SHCFD_SCHEME_DUAL;
    const bool inner = inner_loop_scheme(d.p.scheme);
    // without the (res, omega) partials a wave takes 4 rows x 16 steps, as in
    // k_fast_rhs_prep; with them it keeps 16 rows x 4 steps, the layout their
    // summation order follows
    const int r = inner ? (int)(threadIdx.x & 15) : (int)(threadIdx.x >> 4);
    const int dd = wk.y + (inner ? (int)(threadIdx.x >> 4) : (int)(threadIdx.x & 15));
    double res = 0.0, om = 1e300;
    int i, j;
    if (slot_cell(S, B, dd, r, i, j)) {
        const long long c = cidx(B, i, j);
        double z[4], x[4];
#pragma unroll
        for (int k = 0; k < 4; k++) z[k] = Z_sl[slot1(S.base, dd, 4 * r + k)];
        if (dual) {
            const PostScal J = post_scal(d, c);
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
                if (B.active && !inner) FLD(d, SHCFD_F_Q, k, c) = FLD(d, SHCFD_F_Q, k, c) + x[k];
            }
            if (inner) {
                res = x[3] * x[3];
                if (x[0] < 0.0) om = (-FLD(d, SHCFD_F_Q, 0, c) / x[0]) * d.p.dqmax;
            }
        }
    }
    if (!inner) return;
    for (int off = 32; off >= 1; off >>= 1) {
        res = res + __shfl_xor(res, off);
        om = dmin(om, __shfl_xor(om, off));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { shr[w] = res; sho[w] = om; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * wi] = ((shr[0] + shr[1]) + (shr[2] + shr[3]));
        part[2 * wi + 1] = dmin(dmin(sho[0], sho[1]), dmin(sho[2], sho[3]));
    }
}

// ---------------------------------------------------------------------------
// Discrete adjoint: transposed LU-SGS.  With the folded operators above,
// one LU-SGS update is Dq = diag(alpha) (I + U)^-1 (I - L)^-1 R, where L holds
// GA, GB at the slots of the upwind neighbours and U holds HA, HB.  Its
// transpose applied to r is
//     y(c) = alpha(c) r(c) - HA(c)^T y(c - e_i) - HB(c)^T y(c - e_j)
//     x(c) = y(c) + GA(c)^T x(c + e_i) + GB(c)^T x(c + e_j),
// the same two sweeps with operators stored at the neighbour's slot:
//     GA'(c) = -HA(c + e_i)^T,  GB'(c) = -HB(c + e_j)^T,
//     HA'(c) = -GA(c - e_i)^T,  HB'(c) = -GB(c - e_j)^T,
// zero where the neighbour is outside the block.  LU-SGS and NR (S = I) only.
// ---------------------------------------------------------------------------
// one transposed operator of neighbour cn: -(fs * M)^T in lane rows
__device__ inline void put_transposed(double* o, const QStrip& S, int dd, int r, int pbase, const double* M, double fs,
                                      bool present) {
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int l = 0; l < 4; l++)
            o[op_index(S, dd, 4 * r + k, pbase + (l >> 1), l & 1)] = present ? -(fs * M[l + 4 * k]) : 0.0;
}

__global__ void __launch_bounds__(256) k_fast_adj_prep(DevCtx d, const QStrip* strips, const int2* work, int nwork,
                                                       double* G, double* H) {
    // work items in XCD-aware order; a wave takes 4 rows x 16 steps
    const int w = xcd_item((int)blockIdx.x, nwork);
    if (w >= nwork) return;
    const int2 wk = work[w];
    const QStrip S = strips[wk.x];
    const DevBlock B = d.blk[S.b];
    const int r = (int)(threadIdx.x >> 4);
    const int dd = wk.y + (int)(threadIdx.x & 15);
    int i, j;
    if (!slot_cell(S, B, dd, r, i, j)) return;
    const long long c = cidx(B, i, j), P = B.pitch;
    {
        CellJac<1> J;
        cell_jacobian<1>(d, c, P, J);
        FLD(d, SHCFD_F_ALPHA, 0, c) = J.alpha;
    }
    // GA' from HA(c+e_i): dsci alpha Am of the right neighbour
    {
        CellJac<1> J;
        const bool ok = i < B.ni;
        const long long cn = ok ? c + 1 : c;
        cell_jacobian<1>(d, cn, P, J);
        put_transposed(G, S, dd, r, 0, J.Am, FLD(d, SHCFD_F_DSCI, 0, cn) * J.alpha, ok);
        // HA' from GA(c-e_i)
    }
    {
        CellJac<1> J;
        const bool ok = j < B.nj;
        const long long cn = ok ? c + P : c;
        cell_jacobian<1>(d, cn, P, J);
        put_transposed(G, S, dd, r, 2, J.Bm, FLD(d, SHCFD_F_DSCJ, 0, cn) * J.alpha, ok);
    }
    {
        CellJac<1> J;
        const bool ok = i > 1;
        const long long cn = ok ? c - 1 : c;
        cell_jacobian<1>(d, cn, P, J);
        put_transposed(H, S, dd, r, 0, J.Ap, FLD(d, SHCFD_F_DSCI, 0, cn) * J.alpha, ok);
    }
    {
        CellJac<1> J;
        const bool ok = j > 1;
        const long long cn = ok ? c - P : c;
        cell_jacobian<1>(d, cn, P, J);
        put_transposed(H, S, dd, r, 2, J.Bp, FLD(d, SHCFD_F_DSCJ, 0, cn) * J.alpha, ok);
    }
}

// right-hand side of the transposed sweep: alpha r into the skewed layout
__global__ void __launch_bounds__(256) k_fast_adj_rhs(DevCtx d, const QStrip* strips, const int2* work, int nwork,
                                                      const double* r_cell, double* R_sl, double* hand, int nstrips,
                                                      long long hstride) {
    // the transposed sweeps' hand-off rows are refilled here too
    hand_refill(d, strips, hand, nstrips, hstride);
    // work items in XCD-aware order; a wave takes 4 rows x 16 steps
    const int w = xcd_item((int)blockIdx.x, nwork);
    if (w >= nwork) return;
    const int2 wk = work[w];
    const QStrip S = strips[wk.x];
    const DevBlock B = d.blk[S.b];
    const int r = (int)(threadIdx.x >> 4);
    const int dd = wk.y + (int)(threadIdx.x & 15);
    int i, j;
    if (!slot_cell(S, B, dd, r, i, j)) return;
    const long long c = cidx(B, i, j);
    const double al = FLD(d, SHCFD_F_ALPHA, 0, c);
#pragma unroll
    for (int k = 0; k < 4; k++) R_sl[slot1(S.base, dd, 4 * r + k)] = al * r_cell[k * d.NC + c];
}

// lambda += x (no alpha: the transposed product ends with the forward sweep's L^T)
__global__ void __launch_bounds__(256) k_fast_adj_post(DevCtx d, const QStrip* strips, const int2* work, int nwork,
                                                       const double* Z_sl, double* lam) {
    // work items in XCD-aware order; a wave takes 4 rows x 16 steps
    const int w = xcd_item((int)blockIdx.x, nwork);
    if (w >= nwork) return;
    const int2 wk = work[w];
    const QStrip S = strips[wk.x];
    const DevBlock B = d.blk[S.b];
    const int r = (int)(threadIdx.x >> 4);
    const int dd = wk.y + (int)(threadIdx.x & 15);
    int i, j;
    if (!slot_cell(S, B, dd, r, i, j)) return;
    const long long c = cidx(B, i, j);
#pragma unroll
    for (int k = 0; k < 4; k++) lam[k * d.NC + c] += Z_sl[slot1(S.base, dd, 4 * r + k)];
}

}  // namespace shcfd

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
#include "shcfd_fast.h"

namespace shcfd {

int FastPath::init(Ctx& c) {
    if (c.ny != 1) return SHCFD_E_STATE;
    {
        int dev = 0;
        hipDeviceProp_t prop;
        ncu = (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
                  ? prop.multiProcessorCount
                  : 0;
    }
    // strips (b, k) of every block; launch order: block by block, or, with
    // more strips than CUs, strip k of every block before strip k+1 of any,
    // so that a strip is launched well after the one it waits for instead
    // of holding a CU through the pipeline fill of its block.  Either way a
    // strip's neighbours in the sweep order have lower workgroup ids in the
    // sweep that waits for them.
    std::vector<int> nsb(c.nb), first(c.nb);
    int total = 0, maxns = 0;
    for (int b = 0; b < c.nb; b++) {
        nsb[b] = (c.hblk[b].nj + QR - 1) / QR;
        total += nsb[b];
        maxns = std::max(maxns, nsb[b]);
    }
    std::vector<int2> order;  // (b, k) in launch order
    if (ncu > 0 && total > ncu) {
        for (int k = 0; k < maxns; k++)
            for (int b = 0; b < c.nb; b++)
                if (k < nsb[b]) order.push_back(make_int2(b, k));
    } else {
        for (int b = 0; b < c.nb; b++)
            for (int k = 0; k < nsb[b]; k++) order.push_back(make_int2(b, k));
    }
    std::vector<int> idx(total);  // strip index of (b, k) at first[b] + k
    for (int b = 0, o = 0; b < c.nb; o += nsb[b], b++) first[b] = o;
    for (int s = 0; s < total; s++) idx[first[order[s].x] + order[s].y] = s;
    std::vector<QStrip> hs;
    std::vector<int2> hw;
    long long base = 0;
    int maxcols = 0;
    for (int s = 0; s < total; s++) {
        const int b = order[s].x, k = order[s].y;
        const DevBlock& B = c.hblk[b];
        const int nd = B.ni + QR - 1;
        const int nch = (nd + CH - 1) / CH;
        maxcols = std::max(maxcols, nch * CH);
        QStrip S{};
        S.b = b;
        S.j0 = 1 + QR * k;
        S.prev = k > 0 ? idx[first[b] + k - 1] : -1;
        S.next = k < nsb[b] - 1 ? idx[first[b] + k + 1] : -1;
        S.nch = nch;
        // one chunk of guard steps before and after the padded steps
        S.base = base + 64LL * CH;
        for (int d0 = 0; d0 < nd; d0 += 16) hw.push_back(make_int2((int)hs.size(), d0));
        base += 64LL * CH * (nch + 2);
        hs.push_back(S);
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
              al((void**)&G, sizeof(double) * 8 * nslots) && al((void**)&H, sizeof(double) * 8 * nslots) &&
              al((void**)&Dn, sizeof(double) * (nslots / 2)) &&
              al((void**)&hand, sizeof(double) * 2 * hstride * (nstrips + 1)) &&
              al((void**)&part, sizeof(double) * 2 * nwork);
    if (!ok) return SHCFD_E_NOMEM;
    hipMemcpy(strips, hs.data(), sizeof(QStrip) * hs.size(), hipMemcpyHostToDevice);
    hipMemcpy(work, hw.data(), sizeof(int2) * hw.size(), hipMemcpyHostToDevice);
    ready = true;
    return SHCFD_OK;
}

void FastPath::release() {
    void* ptrs[] = {strips, work, R, V, X, Z, G, H, Dn, hand, part, aR, aV, aZ, aG, aH, ahand};
    for (void* p : ptrs)
        if (p) hipFree(p);
    *this = FastPath();
}

int FastPath::implicit(Ctx& c, int step_internal, int dqdt_inline, int abpm_inline, void (*pb)(Ctx*, int),
                       void (*pe)(Ctx*, int)) {
    DevCtx d = c.dev();
    const bool dual = c.p.scheme == SHCFD_SCHEME_DUAL;
    hipStream_t st = c.stream;
    double* hand_f = hand;
    double* hand_b = hand + hstride * (nstrips + 1);
    // Without the dual-time post-scaling (S = I) the backward operators are
    // the forward ones less diagonal terms: the backward sweep streams (GA,
    // GB) and the 16 bytes of those terms per cell that prep writes, instead
    // of 256 bytes of (HA, HB).  In dual time, when the strips leave CUs idle,
    // workgroups of the forward launch past the strips form (HA, HB) there,
    // and the launch asks for 28 KB of dynamic LDS beside the sweep's 57 KB
    // so that no CU holds two of its workgroups: an operator workgroup next
    // to a strip would take issue slots and vector-memory bandwidth from the
    // strip's waves (measured on the 258k-cell forward step: 195 us per
    // forward sweep with them side by side, 178 us apart, 167 us with (HA,
    // HB) formed in prep instead, where prep costs 14 us more).  With more
    // strips than CUs (batches of many blocks) no CU is idle: prep forms (HA,
    // HB) from the same cell Jacobians as (GA, GB), and the forward launch
    // holds the strips alone.
    const bool diag = !dual;
    const bool busy = ncu > 0 && nstrips > ncu;
    const unsigned FWD_PAD_LDS = busy ? 0u : 28672u;
    const int nhwg = (busy || diag) ? 0 : std::min(nwork, 512);
    pb(&c, PROF_PREP);
    hipLaunchKernelGGL(k_fast_rhs_prep, dim3(xcd_groups(nwork)), dim3(256), 0, st, d, strips, work, nwork, dqdt_inline,
                       abpm_inline, R, G,
                       (busy && !diag) ? H : nullptr, diag ? Dn : nullptr, hand, nstrips, hstride);
    pe(&c, PROF_PREP);
    pb(&c, PROF_SWEEP_FWD);
    hipLaunchKernelGGL(k_fast_fwd, dim3(nstrips + nhwg), dim3(128), FWD_PAD_LDS, st, d, strips, nstrips, R, G, V,
                       hand_f, hstride, c.spin_limit, work, nwork, H);
    pe(&c, PROF_SWEEP_FWD);
    const double* Xin = V;
    if (dual) {
        pb(&c, PROF_PREP);
        hipLaunchKernelGGL(k_fast_mid, dim3(xcd_groups(nwork)), dim3(256), 0, st, d, strips, work, nwork, V, X);
        pe(&c, PROF_PREP);
        Xin = X;
    }
    pb(&c, PROF_SWEEP_BWD);
    if (diag)
        hipLaunchKernelGGL(k_fast_bwd<true>, dim3(nstrips), dim3(128), 0, st, d, strips, nstrips, Xin, G, Dn, Z,
                           hand_b, hstride, c.spin_limit);
    else
        hipLaunchKernelGGL(k_fast_bwd<false>, dim3(nstrips), dim3(128), 0, st, d, strips, nstrips, Xin, H, nullptr, Z,
                           hand_b, hstride, c.spin_limit);
    pe(&c, PROF_SWEEP_BWD);
    pb(&c, PROF_UPDATE);
    hipLaunchKernelGGL(k_fast_post, dim3(xcd_groups(nwork)), dim3(256), 0, st, d, strips, work, nwork, Z, part);
    if (inner_loop_scheme(c.p.scheme)) {
        hipLaunchKernelGGL(k_reduce_res_omega, dim3(1), dim3(256), 0, st, d, part, nwork);
        if (step_internal == 1) hipLaunchKernelGGL(k_copy_scal, dim3(1), dim3(1), 0, st, d, (int)SCAL_RES1, (int)SCAL_RES);
        if (c.defer_update)
            c.update_pending = 1;  // applied by the next primitive pass
        else
            hipLaunchKernelGGL(k_update<1>, dim3((unsigned)c.htiles.size()), dim3(64 * c.tile_rows), 0, st, d, c.dtiles);
    }
    pe(&c, PROF_UPDATE);
    return hipGetLastError() == hipSuccess ? SHCFD_OK : SHCFD_E_HIP;
}

int FastPath::adjoint_prep(Ctx& c) {
    const int sch = c.p.scheme;
    if (sch != SHCFD_SCHEME_LUSGS && sch != SHCFD_SCHEME_NR && sch != SHCFD_SCHEME_DUAL) return SHCFD_E_STATE;
    if (!aG) {
        auto al = [&](void** p, size_t bytes) {
            if (hipMalloc(p, bytes) != hipSuccess) return false;
            return hipMemset(*p, 0, bytes) == hipSuccess;
        };
        bool ok = al((void**)&aR, sizeof(double) * nslots) && al((void**)&aV, sizeof(double) * nslots) &&
                  al((void**)&aZ, sizeof(double) * nslots) && al((void**)&aG, sizeof(double) * 8 * nslots) &&
                  al((void**)&aH, sizeof(double) * 8 * nslots) &&
                  al((void**)&ahand, sizeof(double) * 2 * hstride * (nstrips + 1));
        if (!ok) return SHCFD_E_NOMEM;
    }
    // The transposed iteration is preconditioned with the LU-SGS operator of
    // the residual it solves: for the BDF2 steps of dual time that is the NR
    // operator (plain Jacobians, diagonal 3 Vol/(2 dt) + ...), whatever
    // pseudo-time preconditioning the primal inner loop used.
    DevCtx dp = c.dev();
    if (dp.p.scheme == SHCFD_SCHEME_DUAL) dp.p.scheme = SHCFD_SCHEME_NR;
    hipLaunchKernelGGL(k_fast_adj_prep, dim3(xcd_groups(nwork)), dim3(256), 0, c.stream, dp, strips, work, nwork, aG,
                       aH);
    return hipGetLastError() == hipSuccess ? SHCFD_OK : SHCFD_E_HIP;
}

int FastPath::adjoint_apply(Ctx& c, const double* r_cell, double* lam) {
    DevCtx d = c.dev();
    hipStream_t st = c.stream;
    hipLaunchKernelGGL(k_fast_adj_rhs, dim3(xcd_groups(nwork)), dim3(256), 0, st, d, strips, work, nwork, r_cell, aR,
                       ahand, nstrips, hstride);
    hipLaunchKernelGGL(k_fast_fwd, dim3(nstrips), dim3(128), 0, st, d, strips, nstrips, aR, aG, aV, ahand, hstride,
                       c.spin_limit, work, 0, nullptr);
    hipLaunchKernelGGL(k_fast_bwd<false>, dim3(nstrips), dim3(128), 0, st, d, strips, nstrips, aV, aH, nullptr, aZ,
                       ahand + hstride * (nstrips + 1), hstride, c.spin_limit);
    hipLaunchKernelGGL(k_fast_adj_post, dim3(xcd_groups(nwork)), dim3(256), 0, st, d, strips, work, nwork, aZ, lam);
    return hipGetLastError() == hipSuccess ? SHCFD_OK : SHCFD_E_HIP;
}

}  // namespace shcfd
