// shcfd_adjoint.hip -- C-ABI of the discrete adjoint (ideal gas, inviscid or
// viscous, on the optimised implicit path: steady LU-SGS, and the BDF2
// dual-time steps of the NR scheme with a fixed time step).  The reference has no adjoint; these
// entry points drive the reverse kernels of k_adjoint.hip and the transposed
// sweeps of k_fast.hip.
#include <cmath>
#include <new>
#include <vector>

#include "../../include/shcfd.h"
#include "shcfd_kernels.h"
#include "shcfd_adj.h"
#include "shcfd_fast.h"

// the opaque handle of the C ABI is the context itself (as in shcfd_api.hip)
struct shcfd_ctx : public shcfd::Ctx {};

namespace shcfd {

// k_adjoint.hip
template <int DIR> __global__ void k_adj_face(DevCtx, AdjView, int);
template <int DIR> __global__ void k_adj_vface(DevCtx, AdjView, int);
__global__ void k_adj_muscl(DevCtx, AdjView, int);
__global__ void k_adj_bc_corner(DevCtx, AdjView);
__global__ void k_adj_centres(DevCtx, AdjView, int);
__global__ void k_adj_ghost_centres(DevCtx, AdjView, int, int);
__global__ void k_adj_bc_rows(DevCtx, AdjView, const int*, int, const double*, double*, int, const int*, int);
__global__ void k_adj_cell(DevCtx, AdjView, int);
__global__ void k_adj_nodes(DevCtx, AdjView, int);
__global__ void k_adj_objective(DevCtx, AdjView, double*);
__global__ void k_adj_norm_part(DevCtx, AdjView, int, double*);
__global__ void k_adj_unsteady_begin(DevCtx, AdjView, int);
__global__ void k_adj_unsteady_accum(DevCtx, AdjView, int);
__global__ void k_adj_pool_acc(const double*, double*, int);
// shcfd_api.hip
bool fast_ready(Ctx* ctx);

__global__ void k_adj_set_dt(DevCtx d, double dt) { d.scal[SCAL_DT_GRBL] = dt; }

int Adjoint::init(Ctx& c) {
    const long long NC = c.NC;
    const bool vis = c.p.viscous != 0;
    const long long n = NC * (4 + 4 + 9 + 4 * 9 + 2 + 2 + 1 + 1 + 1 + 1 + 4 + 4 + 1 + 1 + (vis ? 2 * 48 + 2 * 12 + 2 : 0));
    const size_t npb = sizeof(double) * (c.npool > 0 ? c.npool : 1);
    npart = 256;
    if (hipMalloc((void**)&mem, sizeof(double) * n) != hipSuccess) return SHCFD_E_NOMEM;
    if (hipMalloc((void**)&poolbar, npb) != hipSuccess) return SHCFD_E_NOMEM;
    if (hipMalloc((void**)&poolacc, npb) != hipSuccess) return SHCFD_E_NOMEM;
    if (hipMemset(poolacc, 0, npb) != hipSuccess) return SHCFD_E_HIP;
    if (hipMalloc((void**)&part, sizeof(double) * (npart + 1)) != hipSuccess) return SHCFD_E_NOMEM;
    objval = part + npart;
    if (hipMemset(mem, 0, sizeof(double) * n) != hipSuccess) return SHCFD_E_HIP;
    if (hipMemset(poolbar, 0, sizeof(double) * (c.npool > 0 ? c.npool : 1)) != hipSuccess) return SHCFD_E_HIP;
    double* p = mem;
    auto take = [&](long long k) { double* r = p; p += k * NC; return r; };
    v.lam = take(4);
    v.qbar = take(4);
    v.wbar = take(9);
    for (int f = 0; f < 4; f++) v.face[f] = take(9);
    v.nbar_i = take(2);
    v.nbar_j = take(2);
    v.dsbar_i = take(1);
    v.dsbar_j = take(1);
    v.xbar = take(1);
    v.ybar = take(1);
    v.src = take(4);
    v.lam2 = take(4);
    v.xacc = take(1);
    v.yacc = take(1);
    v.vw_i = vis ? take(48) : nullptr;
    v.vw_j = vis ? take(48) : nullptr;
    v.vg_i = vis ? take(12) : nullptr;
    v.vg_j = vis ? take(12) : nullptr;
    v.cbar = vis ? take(2) : nullptr;
    v.unsteady = 0;
    v.obj_block = -1;
    ready = true;
    return SHCFD_OK;
}

void Adjoint::release() {
    void* ptrs[] = {mem, poolbar, poolacc, part};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    *this = Adjoint();
}

void adjoint_release(Ctx* c) {
    if (c->adj) {
        c->adj->release();
        delete c->adj;
        c->adj = nullptr;
    }
}

}  // namespace shcfd

using namespace shcfd;

// ---- restarted GMRES over adjoint fields --------------------------------
// Vectors are cell fields of 4 components (component k of cell c at k NC + c);
// only interior cells take part, block by block.
__global__ void __launch_bounds__(256) k_gm_axpby(DevCtx d, int b, double* y, const double* x, double a, double c) {
    const DevBlock B = d.blk[b];
    const long long n = (long long)B.ni * B.nj;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long cc = cidx(B, (int)(t % B.ni) + 1, (int)(t / B.ni) + 1);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const long long e = k * d.NC + cc;
            y[e] = (c == 0.0 ? 0.0 : c * y[e]) + (a == 0.0 ? 0.0 : a * x[e]);
        }
    }
}

// partial sums of <w, V_i> for i = blockIdx.y (V_i at V + i ld), one per
// workgroup of the grid's x dimension: part[i gridDim.x + blockIdx.x]
__global__ void __launch_bounds__(256) k_gm_dots(DevCtx d, int b, const double* w, const double* V, long long ld,
                                                 double* part) {
    __shared__ double red[256];
    const DevBlock B = d.blk[b];
    const double* v = V + blockIdx.y * ld;
    const long long n = (long long)B.ni * B.nj;
    double s = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long cc = cidx(B, (int)(t % B.ni) + 1, (int)(t / B.ni) + 1);
#pragma unroll
        for (int k = 0; k < 4; k++) s += w[k * d.NC + cc] * v[k * d.NC + cc];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// w += sum_i coef[i] V_i (i < nv)
__global__ void __launch_bounds__(256) k_gm_combine(DevCtx d, int b, double* w, const double* V, long long ld,
                                                    const double* coef, int nv) {
    const DevBlock B = d.blk[b];
    const long long n = (long long)B.ni * B.nj;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long cc = cidx(B, (int)(t % B.ni) + 1, (int)(t / B.ni) + 1);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const long long e = k * d.NC + cc;
            double acc = w[e];
            for (int i = 0; i < nv; i++) acc += coef[i] * V[i * ld + e];
            w[e] = acc;
        }
    }
}

namespace {

int supported(shcfd_ctx* ctx) {
    if (ctx->ny != 1 || ctx->p.thermal != SHCFD_THERMAL_IDEAL ||
        ctx->p.dimension != SHCFD_DIM_PLANAR ||
        (ctx->p.scheme != SHCFD_SCHEME_LUSGS && ctx->p.scheme != SHCFD_SCHEME_NR &&
         ctx->p.scheme != SHCFD_SCHEME_DUAL))
        return SHCFD_E_STATE;
    // the reverse boundary kernel covers the kinds up to the curved slip wall
    for (int r = 0; r < ctx->nbc; r++)
        if (ctx->hbc[SHCFD_BC_ROW * r + 4] > SHCFD_BC_SLIP_NORMAL) return SHCFD_E_STATE;
    return SHCFD_OK;
}

int ensure(shcfd_ctx* ctx) {
    int rc = supported(ctx);
    if (rc) return rc;
    // the viscous arrays exist only if the flow was viscous at set-up
    if (ctx->adj && ctx->p.viscous && !ctx->adj->v.vw_i) adjoint_release(ctx);
    if (!ctx->adj) {
        ctx->adj = new (std::nothrow) Adjoint();
        if (!ctx->adj) return SHCFD_E_NOMEM;
        rc = ctx->adj->init(*ctx);
        if (rc) { adjoint_release(ctx); return rc; }
    }
    return SHCFD_OK;
}

// residual pipeline at the current state (face states, fluxes)
int residual_state(shcfd_ctx* ctx) {
    int rc = shcfd_set_w(ctx);
    if (!rc) rc = shcfd_set_thermo_prop(ctx, 0);
    if (!rc) rc = shcfd_set_bc(ctx, 0);
    if (!rc) rc = shcfd_set_hv(ctx);
    if (!rc) rc = shcfd_set_tg(ctx);
    if (!rc && ctx->p.viscous) rc = shcfd_set_tgv(ctx);
    return rc;
}

void metric_bars_async(shcfd_ctx* ctx);

// one reverse pass with the current lambda (launches only): QBAR =
// (dR/dq)^T lam + dJ/dq, the pool bars and, when metrics, the node bars of
// the metrics
void vjp_async(shcfd_ctx* ctx, bool metrics = true) {
    DevCtx d = ctx->dev();
    AdjView& a = ctx->adj->v;
    hipStream_t st = ctx->stream;
    const bool vis = ctx->p.viscous != 0;
    for (int b = 0; b < ctx->nb; b++) {
        const DevBlock& B = ctx->hblk[b];
        const dim3 gi((unsigned)((B.ni + 1 + 127) / 128), (unsigned)B.nj), gj((unsigned)((B.ni + 127) / 128),
                                                                                (unsigned)(B.nj + 1));
        hipLaunchKernelGGL(k_adj_face<0>, gi, dim3(128), 0, st, d, a, b);
        hipLaunchKernelGGL(k_adj_face<1>, gj, dim3(128), 0, st, d, a, b);
        if (vis) {  // after k_adj_face: they add to its normal bars
            hipLaunchKernelGGL(k_adj_vface<0>, gi, dim3(128), 0, st, d, a, b);
            hipLaunchKernelGGL(k_adj_vface<1>, gj, dim3(128), 0, st, d, a, b);
        }
    }
    for (int b = 0; b < ctx->nb; b++) {
        const DevBlock& B = ctx->hblk[b];
        hipLaunchKernelGGL(k_adj_muscl, dim3((unsigned)((B.ni + 4 + 127) / 128), (unsigned)(B.nj + 4)), dim3(128), 0, st,
                           d, a, b);
    }
    // reverse of the forward fill order: corners (read by the viscous
    // fluxes only), then side rows and cut rows in one workgroup (which
    // clears the pool bars first)
    if (vis) hipLaunchKernelGGL(k_adj_bc_corner, dim3((unsigned)((ctx->nb + 63) / 64)), dim3(64), 0, st, d, a);
    hipLaunchKernelGGL(k_adj_bc_rows, dim3(1), dim3(256), 0, st, d, a, ctx->dbc, ctx->nbc, ctx->dpool,
                       ctx->adj->poolbar, ctx->npool > 0 ? ctx->npool : 1, ctx->dcut, ctx->ncut);
    for (int b = 0; b < ctx->nb; b++) {
        const DevBlock& B = ctx->hblk[b];
        hipLaunchKernelGGL(k_adj_cell, dim3((unsigned)((B.ni + 127) / 128), (unsigned)B.nj), dim3(128), 0, st, d, a, b);
    }
    if (metrics) metric_bars_async(ctx);
}

// the node bars of the metrics (XBAR, and with viscous fluxes the cell-centre
// bars folded into them) from the face and centre bars of the last product;
// only the shape gradient reads them
void metric_bars_async(shcfd_ctx* ctx) {
    DevCtx d = ctx->dev();
    AdjView& a = ctx->adj->v;
    hipStream_t st = ctx->stream;
    const bool vis = ctx->p.viscous != 0;
    for (int b = 0; b < ctx->nb; b++) {
        const DevBlock& B = ctx->hblk[b];
        if (vis) {  // centre-coordinate bars, ghost centres folded back, before the nodes
            hipLaunchKernelGGL(k_adj_centres, dim3((unsigned)((B.ni + 2 + 127) / 128), (unsigned)(B.nj + 2)), dim3(128),
                               0, st, d, a, b);
            hipLaunchKernelGGL(k_adj_ghost_centres, dim3((unsigned)((B.ni + 2 + 127) / 128)), dim3(128), 0, st, d, a, b,
                               2);
            hipLaunchKernelGGL(k_adj_ghost_centres, dim3((unsigned)((B.nj + 127) / 128)), dim3(128), 0, st, d, a, b, 1);
        }
        hipLaunchKernelGGL(k_adj_nodes, dim3((unsigned)((B.ni + 1 + 127) / 128), (unsigned)(B.nj + 1)), dim3(128), 0,
                           st, d, a, b);
    }
}

// RMS of the q bars over all interior cells (synchronises)
int qbar_rms(shcfd_ctx* ctx, double* out) {
    DevCtx d = ctx->dev();
    Adjoint& A = *ctx->adj;
    double total = 0.0;
    long long cells = 0;
    std::vector<double> h(A.npart);
    for (int b = 0; b < ctx->nb; b++) {
        hipLaunchKernelGGL(k_adj_norm_part, dim3(A.npart), dim3(256), 0, ctx->stream, d, A.v, b, A.part);
        if (hipMemcpyAsync(h.data(), A.part, sizeof(double) * A.npart, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            return SHCFD_E_HIP;
        for (int k = 0; k < A.npart; k++) total += h[k];
        cells += (long long)ctx->hblk[b].ni * ctx->hblk[b].nj;
    }
    *out = std::sqrt(total / (double)cells);
    return std::isfinite(*out) ? SHCFD_OK : SHCFD_E_NAN;
}

// lambda += M^-T (dR/dq^T lambda + g) from the current lambda until the RMS
// of the bracket falls below rtol times ref (its first value when ref <= 0;
// the operators of M are those of the last adjoint_prep); QBAR, XBAR and the
// pool bars belong to the final lambda on return
int transposed_iterations(shcfd_ctx* ctx, int max_iter, double rtol, int check_every, double* res0, double* res,
                          int* iters, double ref = 0.0) {
    AdjView& a = ctx->adj->v;
    double r0 = 0.0, r = 0.0;
    int it = 0;
    vjp_async(ctx);
    int rc = qbar_rms(ctx, &r0);
    r = r0;
    if (ref <= 0.0) ref = r0;
    while (!rc && it < max_iter && r > rtol * ref) {
        rc = ctx->fast->adjoint_apply(*ctx, a.qbar, a.lam);
        if (rc) break;
        vjp_async(ctx, false);  // the node bars once, after the loop
        it++;
        if (it % check_every == 0 || it == max_iter) rc = qbar_rms(ctx, &r);
    }
    if (!rc && it > 0) metric_bars_async(ctx);  // of the final lambda
    if (!rc) rc = qbar_rms(ctx, &r);
    if (res0) *res0 = ref;
    if (res) *res = r;
    if (iters) *iters = it;
    if (!rc) rc = shcfd_check_error(ctx);
    return rc;
}

// Restarted GMRES for (dR/dq)^T lambda = -dJ/dq, right-preconditioned by
// the transposed LU-SGS sweeps P = M^-T: the fixed-point iteration above
// converges only where every eigenvalue of I + P (dR/dq)^T lies inside the
// unit circle, which the linearised second-order (MUSCL) scheme does not
// guarantee; GMRES minimises the residual over the Krylov space instead.
// Classical Gram-Schmidt run twice (one multi-dot kernel per pass).  The
// residual reported is the RMS of dR/dq^T lambda + dJ/dq over the interior
// cells, as in the fixed-point solve; on return QBAR, XBAR and the pool bars
// belong to the final lambda.
struct Gmres {
    shcfd_ctx* ctx;
    int m;
    long long ld;           // doubles per vector (4 NC)
    double* V = nullptr;    // m + 1 basis vectors
    double* x = nullptr;    // the iterate
    double* g = nullptr;    // dJ/dq
    double* z = nullptr;    // P v
    double* w = nullptr;
    double* coef = nullptr; // device copy of combination coefficients (m + 1)
    double* part = nullptr; // dot partials: (m + 1) x NP
    static constexpr int NP = 64;
    long long cells = 0;

    int alloc() {
        const size_t n = (size_t)ld * (m + 5) + (m + 1) + (size_t)(m + 1) * NP;
        if (hipMalloc((void**)&V, sizeof(double) * n) != hipSuccess) return SHCFD_E_NOMEM;
        if (hipMemset(V, 0, sizeof(double) * n) != hipSuccess) return SHCFD_E_HIP;
        x = V + ld * (m + 1);
        g = x + ld;
        z = g + ld;
        w = z + ld;
        coef = w + ld;
        part = coef + (m + 1);
        for (int b = 0; b < ctx->nb; b++) cells += (long long)ctx->hblk[b].ni * ctx->hblk[b].nj;
        return SHCFD_OK;
    }
    void release() { if (V) (void)hipFree(V); V = nullptr; }
    unsigned grid(int b) const {
        const long long n = (long long)ctx->hblk[b].ni * ctx->hblk[b].nj;
        return (unsigned)std::min<long long>((n + 255) / 256, 1024);
    }
    // y = a x + c y
    void axpby(double* y, const double* xx, double a, double c) {
        for (int b = 0; b < ctx->nb; b++)
            hipLaunchKernelGGL(k_gm_axpby, dim3(grid(b)), dim3(256), 0, ctx->stream, ctx->dev(), b, y, xx, a, c);
    }
    // h[i] = <v, V_i>, i < nv (fixed-order host sum of the partials)
    int dots(const double* v, const double* basis, int nv, double* h) {
        std::vector<double> acc(nv, 0.0), hp((size_t)nv * NP);
        for (int b = 0; b < ctx->nb; b++) {
            hipLaunchKernelGGL(k_gm_dots, dim3(NP, (unsigned)nv), dim3(256), 0, ctx->stream, ctx->dev(), b, v, basis,
                               ld, part);
            if (hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, ctx->stream) !=
                    hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess)
                return SHCFD_E_HIP;
            for (int i = 0; i < nv; i++)
                for (int q = 0; q < NP; q++) acc[i] += hp[(size_t)i * NP + q];
        }
        for (int i = 0; i < nv; i++) h[i] = acc[i];
        return SHCFD_OK;
    }
    int norm(const double* v, double* out) {
        int rc = dots(v, v, 1, out);
        *out = std::sqrt(*out);
        return rc;
    }
    // w += sum_i c[i] V_i
    int combine(double* dst, const double* c, int nv) {
        if (hipMemcpyAsync(coef, c, sizeof(double) * nv, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            return SHCFD_E_HIP;
        for (int b = 0; b < ctx->nb; b++)
            hipLaunchKernelGGL(k_gm_combine, dim3(grid(b)), dim3(256), 0, ctx->stream, ctx->dev(), b, dst, V, ld, coef,
                               nv);
        return SHCFD_OK;
    }
    // qbar = (dR/dq)^T lam + dJ/dq with lam = src
    void vjp_of(const double* src) {
        AdjView& a = ctx->adj->v;
        (void)hipMemcpyAsync(a.lam, src, sizeof(double) * ld, hipMemcpyDeviceToDevice, ctx->stream);
        vjp_async(ctx);
    }

    int solve(int max_iter, double rtol, double* res0, double* res, int* iters) {
        AdjView& a = ctx->adj->v;
        const double rms = 1.0 / std::sqrt((double)cells);
        int rc = SHCFD_OK, it = 0;
        // dJ/dq, then the residual of the starting lambda
        (void)hipMemcpyAsync(x, a.lam, sizeof(double) * ld, hipMemcpyDeviceToDevice, ctx->stream);
        (void)hipMemsetAsync(w, 0, sizeof(double) * ld, ctx->stream);
        vjp_of(w);
        (void)hipMemcpyAsync(g, a.qbar, sizeof(double) * ld, hipMemcpyDeviceToDevice, ctx->stream);
        vjp_of(x);
        double beta = 0.0, beta0 = 0.0;
        axpby(V, a.qbar, -1.0, 0.0);   // r = -(A x + g)
        if ((rc = norm(V, &beta0))) return rc;
        beta = beta0;
        if (res0) *res0 = beta0 * rms;
        std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), gv(m + 1), h(m + 1), h2(m + 1), y(m);
        while (it < max_iter && beta > rtol * beta0 && std::isfinite(beta)) {
            axpby(V, V, 1.0 / beta - 1.0, 1.0);   // V_0 = r / beta
            std::fill(gv.begin(), gv.end(), 0.0);
            gv[0] = beta;
            int kdim = 0;
            for (int j = 0; j < m && it < max_iter; j++) {
                double* vj = V + (long long)j * ld;
                double* vn = V + (long long)(j + 1) * ld;
                (void)hipMemsetAsync(z, 0, sizeof(double) * ld, ctx->stream);
                if ((rc = ctx->fast->adjoint_apply(*ctx, vj, z))) return rc;
                vjp_of(z);
                axpby(vn, a.qbar, 1.0, 0.0);
                axpby(vn, g, -1.0, 1.0);   // A z
                // classical Gram-Schmidt, twice
                if ((rc = dots(vn, V, j + 1, h.data()))) return rc;
                for (int i = 0; i <= j; i++) h2[i] = -h[i];
                if ((rc = combine(vn, h2.data(), j + 1))) return rc;
                if ((rc = dots(vn, V, j + 1, h2.data()))) return rc;
                for (int i = 0; i <= j; i++) h[i] += h2[i], h2[i] = -h2[i];
                if ((rc = combine(vn, h2.data(), j + 1))) return rc;
                double hn = 0.0;
                if ((rc = norm(vn, &hn))) return rc;
                if (!std::isfinite(hn)) return SHCFD_E_NAN;
                if (hn > 0.0) axpby(vn, vn, 1.0 / hn - 1.0, 1.0);
                // Givens rotations of the new column of H
                for (int i = 0; i < j; i++) {
                    const double t = cs[i] * h[i] + sn[i] * h[i + 1];
                    h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
                    h[i] = t;
                }
                const double den = std::hypot(h[j], hn);
                cs[j] = den > 0.0 ? h[j] / den : 1.0;
                sn[j] = den > 0.0 ? hn / den : 0.0;
                h[j] = den;
                gv[j + 1] = -sn[j] * gv[j];
                gv[j] = cs[j] * gv[j];
                for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = h[i];
                kdim = j + 1;
                it++;
                if (std::fabs(gv[j + 1]) <= rtol * beta0 || hn == 0.0) break;
            }
            // y = H^-1 gv (upper triangular), x += P (V y)
            for (int i = kdim - 1; i >= 0; i--) {
                double t = gv[i];
                for (int l = i + 1; l < kdim; l++) t -= H[(size_t)i * m + l] * y[l];
                y[i] = t / H[(size_t)i * m + i];
            }
            (void)hipMemsetAsync(w, 0, sizeof(double) * ld, ctx->stream);
            if ((rc = combine(w, y.data(), kdim))) return rc;
            if ((rc = ctx->fast->adjoint_apply(*ctx, w, x))) return rc;
            // the true residual of the new iterate (also leaves the bars of x)
            if ((rc = residual_state(ctx))) return rc;
            vjp_of(x);
            axpby(V, a.qbar, -1.0, 0.0);
            if ((rc = norm(V, &beta))) return rc;
        }
        if (res) *res = beta * rms;
        if (iters) *iters = it;
        if (!std::isfinite(beta)) return SHCFD_E_NAN;
        return shcfd_check_error(ctx);
    }
};

// host <-> device copy of K cell-indexed components of one block in the
// reference layout (k fastest, then i = -1..ni+2, then j = -1..nj+2)
int copy_block(shcfd_ctx* ctx, double* const* dev, int K, int block, double* host, bool to_device) {
    const DevBlock& B = ctx->hblk[block];
    const size_t rows = (size_t)B.nj + 4, w = (size_t)B.ni + 4;
    std::vector<double> plane(rows * B.pitch, 0.0);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return SHCFD_E_HIP;
    for (int k = 0; k < K; k++) {
        double* base = dev[k] + B.off;
        if (to_device) {
            for (size_t j = 0; j < rows; j++)
                for (size_t i = 0; i < w; i++) plane[j * B.pitch + i] = host[k + (size_t)K * (i + w * j)];
            if (hipMemcpy(base, plane.data(), sizeof(double) * plane.size(), hipMemcpyHostToDevice) != hipSuccess)
                return SHCFD_E_HIP;
        } else {
            if (hipMemcpy(plane.data(), base, sizeof(double) * plane.size(), hipMemcpyDeviceToHost) != hipSuccess)
                return SHCFD_E_HIP;
            for (size_t j = 0; j < rows; j++)
                for (size_t i = 0; i < w; i++) host[k + (size_t)K * (i + w * j)] = plane[j * B.pitch + i];
        }
    }
    return SHCFD_OK;
}

}  // namespace

extern "C" {

int shcfd_adjoint_objective(shcfd_ctx* ctx, int block, int i0, int i1, double ax, double ay) {
    if (!ctx || block < 0 || block >= ctx->nb || i0 < 1 || i1 > ctx->hblk[block].ni || i0 > i1) return SHCFD_E_ARG;
    int rc = ensure(ctx);
    if (rc) return rc;
    AdjView& a = ctx->adj->v;
    a.obj_block = block;
    a.obj_i0 = i0;
    a.obj_i1 = i1;
    a.obj_ax = ax;
    a.obj_ay = ay;
    return SHCFD_OK;
}

int shcfd_objective(shcfd_ctx* ctx, double* J) {
    if (!ctx || !J) return SHCFD_E_ARG;
    int rc = ensure(ctx);
    if (!rc && ctx->adj->v.obj_block < 0) rc = SHCFD_E_STATE;
    if (!rc) rc = residual_state(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_adj_objective, dim3(1), dim3(256), 0, ctx->stream, ctx->dev(), ctx->adj->v, ctx->adj->objval);
    if (hipMemcpyAsync(J, ctx->adj->objval, sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return SHCFD_E_HIP;
    return shcfd_check_error(ctx);
}

int shcfd_adjoint_upload(shcfd_ctx* ctx, int what, int block, const double* host) {
    if (!ctx || !host || block < 0 || block >= ctx->nb || what != SHCFD_ADJ_LAMBDA) return SHCFD_E_ARG;
    int rc = ensure(ctx);
    if (rc) return rc;
    double* dev[4];
    for (int k = 0; k < 4; k++) dev[k] = ctx->adj->v.lam + (long long)k * ctx->NC;
    return copy_block(ctx, dev, 4, block, const_cast<double*>(host), true);
}

int shcfd_adjoint_download(shcfd_ctx* ctx, int what, int block, double* host) {
    if (!ctx || !host || block < 0 || block >= ctx->nb) return SHCFD_E_ARG;
    int rc = ensure(ctx);
    if (rc) return rc;
    AdjView& a = ctx->adj->v;
    double* dev[4];
    switch (what) {
    case SHCFD_ADJ_LAMBDA:
        for (int k = 0; k < 4; k++) dev[k] = a.lam + (long long)k * ctx->NC;
        return copy_block(ctx, dev, 4, block, host, false);
    case SHCFD_ADJ_QBAR:
        for (int k = 0; k < 4; k++) dev[k] = a.qbar + (long long)k * ctx->NC;
        return copy_block(ctx, dev, 4, block, host, false);
    case SHCFD_ADJ_XBAR:
        dev[0] = a.xbar;
        dev[1] = a.ybar;
        return copy_block(ctx, dev, 2, block, host, false);
    case SHCFD_ADJ_XBAR_SUM:
        dev[0] = a.xacc;
        dev[1] = a.yacc;
        return copy_block(ctx, dev, 2, block, host, false);
    case SHCFD_ADJ_LAMBDA_LATER:
        for (int k = 0; k < 4; k++) dev[k] = a.lam2 + (long long)k * ctx->NC;
        return copy_block(ctx, dev, 4, block, host, false);
    default:
        return SHCFD_E_ARG;
    }
}

int shcfd_adjoint_poolbar(shcfd_ctx* ctx, double* host) {
    if (!ctx || !host) return SHCFD_E_ARG;
    int rc = ensure(ctx);
    if (rc) return rc;
    if (ctx->npool <= 0) return SHCFD_OK;
    if (hipMemcpyAsync(host, ctx->adj->poolbar, sizeof(double) * ctx->npool, hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return SHCFD_E_HIP;
    return SHCFD_OK;
}

int shcfd_residual_vjp(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    int rc = ensure(ctx);
    if (!rc && ctx->adj->v.obj_block < 0) rc = SHCFD_E_STATE;
    if (!rc) ctx->adj->v.unsteady = 0;
    if (!rc) rc = residual_state(ctx);
    if (rc) return rc;
    vjp_async(ctx);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) return SHCFD_E_HIP;
    return shcfd_check_error(ctx);
}

int shcfd_adjoint_solve(shcfd_ctx* ctx, int max_iter, double rtol, int check_every, double* res0, double* res,
                        int* iters) {
    if (!ctx || max_iter < 0 || check_every < 1) return SHCFD_E_ARG;
    int rc = ensure(ctx);
    if (!rc && (ctx->adj->v.obj_block < 0 || ctx->p.scheme != SHCFD_SCHEME_LUSGS)) rc = SHCFD_E_STATE;
    if (!rc && !fast_ready(ctx)) rc = SHCFD_E_STATE;
    if (rc) return rc;
    ctx->adj->v.unsteady = 0;
    rc = residual_state(ctx);
    if (!rc) rc = shcfd_set_dt(ctx, 0, nullptr);
    if (!rc) rc = ctx->fast->adjoint_prep(*ctx);
    if (!rc) rc = transposed_iterations(ctx, max_iter, rtol, check_every, res0, res, iters);
    return rc;
}

int shcfd_adjoint_solve_gmres(shcfd_ctx* ctx, int restart, int max_iter, double rtol, double* res0, double* res,
                              int* iters) {
    if (!ctx || restart < 1 || restart > 500 || max_iter < 0) return SHCFD_E_ARG;
    int rc = ensure(ctx);
    if (!rc && (ctx->adj->v.obj_block < 0 || ctx->p.scheme != SHCFD_SCHEME_LUSGS)) rc = SHCFD_E_STATE;
    if (!rc && !fast_ready(ctx)) rc = SHCFD_E_STATE;
    if (rc) return rc;
    ctx->adj->v.unsteady = 0;
    rc = residual_state(ctx);
    if (!rc) rc = shcfd_set_dt(ctx, 0, nullptr);
    if (!rc) rc = ctx->fast->adjoint_prep(*ctx);
    if (rc) return rc;
    Gmres G{ctx, restart, 4LL * ctx->NC};
    rc = G.alloc();
    if (!rc) rc = G.solve(max_iter, rtol, res0, res, iters);
    G.release();
    return rc;
}

int shcfd_adjoint_unsteady_reset(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    int rc = ensure(ctx);
    if (rc) return rc;
    const AdjView& a = ctx->adj->v;
    const size_t cell = sizeof(double) * (size_t)ctx->NC;
    hipStream_t st = ctx->stream;
    if (hipMemsetAsync(a.lam, 0, 4 * cell, st) != hipSuccess || hipMemsetAsync(a.lam2, 0, 4 * cell, st) != hipSuccess ||
        hipMemsetAsync(a.xacc, 0, cell, st) != hipSuccess || hipMemsetAsync(a.yacc, 0, cell, st) != hipSuccess ||
        hipMemsetAsync(ctx->adj->poolacc, 0, sizeof(double) * (ctx->npool > 0 ? ctx->npool : 1), st) != hipSuccess)
        return SHCFD_E_HIP;
    return SHCFD_OK;
}

int shcfd_adjoint_unsteady_step(shcfd_ctx* ctx, int max_iter, double rtol, int check_every, double* res0,
                                double* res, int* iters) {
    if (!ctx || max_iter < 0 || check_every < 1) return SHCFD_E_ARG;
    int rc = ensure(ctx);
    if (!rc && (ctx->adj->v.obj_block < 0 || ctx->p.scheme == SHCFD_SCHEME_LUSGS || ctx->p.time_step != 0 ||
                !(ctx->dt_fixed > 0.0)))
        rc = SHCFD_E_STATE;
    if (!rc && !fast_ready(ctx)) rc = SHCFD_E_STATE;
    if (rc) return rc;
    DevCtx d = ctx->dev();
    AdjView& a = ctx->adj->v;
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_adj_set_dt, dim3(1), dim3(1), 0, st, d, ctx->dt_fixed);
    rc = residual_state(ctx);
    if (!rc) rc = shcfd_set_dqdt(ctx);
    if (rc) return rc;
    a.unsteady = ctx->first_dual ? 2 : 1;
    for (int b = 0; b < ctx->nb; b++) {
        const DevBlock& B = ctx->hblk[b];
        hipLaunchKernelGGL(k_adj_unsteady_begin, dim3((unsigned)((B.ni + 127) / 128), (unsigned)B.nj), dim3(128), 0, st,
                           d, a, b);
    }
    rc = ctx->fast->adjoint_prep(*ctx);
    if (rc) return rc;
    // the right-hand side: the residual at lambda = 0 (lambda^(n+1), the
    // starting guess, is kept in the later lambda meanwhile)
    const size_t fld = sizeof(double) * 4 * (size_t)ctx->NC;
    double rhs = 0.0;
    if (hipMemsetAsync(a.lam, 0, fld, st) != hipSuccess) return SHCFD_E_HIP;
    vjp_async(ctx);
    if ((rc = qbar_rms(ctx, &rhs))) return rc;
    if (hipMemcpyAsync(a.lam, a.lam2, fld, hipMemcpyDeviceToDevice, st) != hipSuccess) return SHCFD_E_HIP;
    rc = transposed_iterations(ctx, max_iter, rtol, check_every, res0, res, iters, rhs > 0.0 ? rhs : -1.0);
    if (rc) return rc;
    for (int b = 0; b < ctx->nb; b++) {
        const DevBlock& B = ctx->hblk[b];
        hipLaunchKernelGGL(k_adj_unsteady_accum, dim3((unsigned)((B.ni + 1 + 127) / 128), (unsigned)(B.nj + 1)),
                           dim3(128), 0, st, d, a, b);
    }
    if (ctx->npool > 0)
        hipLaunchKernelGGL(k_adj_pool_acc, dim3((unsigned)((ctx->npool + 255) / 256)), dim3(256), 0, st,
                           ctx->adj->poolbar, ctx->adj->poolacc, ctx->npool);
    if (hipGetLastError() != hipSuccess) return SHCFD_E_HIP;
    return shcfd_check_error(ctx);
}

int shcfd_adjoint_unsteady_poolbar(shcfd_ctx* ctx, double* host) {
    if (!ctx || !host) return SHCFD_E_ARG;
    int rc = ensure(ctx);
    if (rc) return rc;
    if (ctx->npool <= 0) return SHCFD_OK;
    if (hipMemcpyAsync(host, ctx->adj->poolacc, sizeof(double) * ctx->npool, hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return SHCFD_E_HIP;
    return SHCFD_OK;
}

}  // extern "C"
