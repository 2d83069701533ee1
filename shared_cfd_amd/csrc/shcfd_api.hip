// shcfd_api.hip -- C ABI of libshcfd (include/shcfd.h): context set-up, data
// movement, tables, and the per-routine / per-step drivers that enqueue the
// HIP kernels on the context's stream.
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

#include "../../include/shcfd.h"
#include "shcfd_kernels.h"
#include "shcfd_fast.h"

using namespace shcfd;

struct shcfd_ctx : public Ctx {};

namespace shcfd {
int field_ncomp(int fld, int ny) {
    int dimq = ny + 3, dimw = dimq + 5;
    switch (fld) {
    case SHCFD_F_Q: case SHCFD_F_QP: case SHCFD_F_QPP: case SHCFD_F_TGI: case SHCFD_F_TGJ:
    case SHCFD_F_TGVI: case SHCFD_F_TGVJ: case SHCFD_F_PRE1: case SHCFD_F_PRE2: case SHCFD_F_DQDT:
    case SHCFD_F_DQ:
        return dimq;
    case SHCFD_F_W: case SHCFD_F_WHLI: case SHCFD_F_WHRI: case SHCFD_F_WHLJ: case SHCFD_F_WHRJ:
        return dimw;
    case SHCFD_F_VHI: case SHCFD_F_DHI:
        return ny;
    case SHCFD_F_VNI: case SHCFD_F_VNJ: case SHCFD_F_VNCI: case SHCFD_F_VNCJ:
        return 2;
    case SHCFD_F_GEOJACI: case SHCFD_F_GEOJACJ: case SHCFD_F_GEOJACC: case SHCFD_F_PRESCA:
        return 4;
    case SHCFD_F_AP: case SHCFD_F_AM: case SHCFD_F_BP: case SHCFD_F_BM:
        return dimq * dimq;
    default:
        return 1;
    }
}
}  // namespace shcfd

#define HIPCHK(x)                                                                             \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) {                                                               \
            if (ctx) ctx->last_error = std::string(#x) + ": " + hipGetErrorString(e_);        \
            return SHCFD_E_HIP;                                                               \
        }                                                                                     \
    } while (0)
#define LAUNCHCHK()                                                                           \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess) {                                                               \
            ctx->last_error = std::string("launch: ") + hipGetErrorString(e_);                \
            return SHCFD_E_HIP;                                                               \
        }                                                                                     \
    } while (0)

// species counts 1..SHCFD_MAX_NY (kernels instantiated per group, see
// shcfd_kernels.h)
#define DISPATCH_NY(NYV, CALL)           \
    switch (NYV) {                       \
    case 1: { constexpr int NY = 1; CALL; } break; \
    case 2: { constexpr int NY = 2; CALL; } break; \
    case 3: { constexpr int NY = 3; CALL; } break; \
    case 4: { constexpr int NY = 4; CALL; } break; \
    case 5: { constexpr int NY = 5; CALL; } break; \
    case 6: { constexpr int NY = 6; CALL; } break; \
    case 7: { constexpr int NY = 7; CALL; } break; \
    case 8: { constexpr int NY = 8; CALL; } break; \
    case 9: { constexpr int NY = 9; CALL; } break; \
    case 10: { constexpr int NY = 10; CALL; } break; \
    case 11: { constexpr int NY = 11; CALL; } break; \
    case 12: { constexpr int NY = 12; CALL; } break; \
    case 13: { constexpr int NY = 13; CALL; } break; \
    case 14: { constexpr int NY = 14; CALL; } break; \
    case 15: { constexpr int NY = 15; CALL; } break; \
    case 16: { constexpr int NY = 16; CALL; } break; \
    default: return SHCFD_E_ARG;         \
    }
#define DISPATCH_FLOW_NS(NSV, CALL)       \
    switch (NSV) {                       \
    case 1: { constexpr int NS = 1; CALL; } break; \
    case 2: { constexpr int NS = 2; CALL; } break; \
    case 3: { constexpr int NS = 3; CALL; } break; \
    case 4: { constexpr int NS = 4; CALL; } break; \
    case 5: { constexpr int NS = 5; CALL; } break; \
    case 6: { constexpr int NS = 6; CALL; } break; \
    case 7: { constexpr int NS = 7; CALL; } break; \
    case 8: { constexpr int NS = 8; CALL; } break; \
    case 9: { constexpr int NS = 9; CALL; } break; \
    case 10: { constexpr int NS = 10; CALL; } break; \
    case 11: { constexpr int NS = 11; CALL; } break; \
    case 12: { constexpr int NS = 12; CALL; } break; \
    case 13: { constexpr int NS = 13; CALL; } break; \
    case 14: { constexpr int NS = 14; CALL; } break; \
    case 15: { constexpr int NS = 15; CALL; } break; \
    case 16: { constexpr int NS = 16; CALL; } break; \
    default: return SHCFD_E_ARG;         \
    }

static bool ny_supported(int ny) { return ny >= 1 && ny <= SHCFD_MAX_NY; }

static void drop_graph(shcfd::Ctx* ctx) {
    if (ctx->graph) hipGraphExecDestroy(ctx->graph);
    ctx->graph = nullptr;
    ctx->graph_steps = 0;
}

// tiles of 64 x rows cells covering i in [0, ni+1], j in [0, nj+1] of every
// block, block by block, rows then columns; one workgroup of 64*rows threads each
static void build_tiles(shcfd::Ctx* ctx, int rows) {
    ctx->htiles.clear();
    ctx->tile_rows = rows;
    for (int b = 0; b < ctx->nb; b++) {
        const DevBlock& B = ctx->hblk[b];
        for (int j0 = 0; j0 <= B.nj + 1; j0 += rows)
            for (int i0 = 0; i0 <= B.ni + 1; i0 += TI)
                ctx->htiles.push_back(Tile{b, i0, j0, (int)ctx->htiles.size()});
    }
}

// The tiles as the device list holds them: in XCD-aware launch order.  The
// dispatcher hands a launch's workgroups to the 8 XCDs in turn (workgroup p
// runs on XCD p % 8), so position p = 8k + x holds tile x*per + k of the
// block order: each XCD takes a contiguous range of tiles, and the rows a
// stencil shares between neighbouring tiles stay in one L2.  The per-tile
// partials are stored at each tile's slot, so their order is unchanged.
static std::vector<Tile> tiles_in_launch_order(const std::vector<Tile>& tiles) {
    const int n = (int)tiles.size(), per = (n + 7) / 8;
    std::vector<Tile> out;
    out.reserve(tiles.size());
    for (int k = 0; k < per; k++)
        for (int x = 0; x < 8; x++)
            if (x * per + k < n) out.push_back(tiles[x * per + k]);
    return out;
}

extern "C" {

const char* shcfd_strerror(int code) {
    switch (code) {
    case SHCFD_OK: return "ok";
    case SHCFD_E_ARG: return "invalid argument";
    case SHCFD_E_HIP: return "HIP runtime error";
    case SHCFD_E_NOMEM: return "device allocation failed";
    case SHCFD_E_NAN: return "NaN in residual";
    case SHCFD_E_OMEGA: return "Omega becomes too small";
    case SHCFD_E_RESRATE: return "ERROR. EXCEED MAXIMUM RESIDUAL.";
    case SHCFD_E_TNEWTON: return "temperature Newton iteration did not converge";
    case SHCFD_E_CHEM: return "chemistry integrator failure";
    case SHCFD_E_SPIN: return "sweep wavefront wait exceeded its bound";
    case SHCFD_E_STATE: return "call not valid for this scheme/state";
    case SHCFD_E_TOO_LARGE: return "request exceeds the fixed device scratch of a no-allocation entry point";
    default: return "unknown error";
    }
}

int shcfd_version(void) { return 100; }

// the scheme / time-step pairs the reference's checkout accepts
// (store/checkout/checkout_flow.py:120-185): precon and preconLU-SGS need
// local time steps, dual time a global one
// checkout_flow.py:106-185: the choices the reference accepts.  Axisymmetric
// cases run the explicit schemes, LU-SGS and the NR inner loop
// (time_scheme/{LU-SGS,NR}/q2d); dual time is refused there as in the
// reference, and so are the two preconditioned schemes, whose q2d sweeps
// assign to the whole Dq array through a case-insensitive name clash and then
// read an unset pressure increment (preconLU-SGS/q2d/time_precon_viscous.f90
// and precon/q2d/time_precon_viscous.f90, forward sweep).
static bool params_valid(const shcfd_params* p) {
    if (p->time_step != 0 && p->time_step != 1) return false;
    if (p->dimension != SHCFD_DIM_PLANAR && p->dimension != SHCFD_DIM_AXISYMMETRIC) return false;
    if (p->dimension == SHCFD_DIM_AXISYMMETRIC &&
        (p->scheme == SHCFD_SCHEME_DUAL || p->scheme == SHCFD_SCHEME_PRECON || p->scheme == SHCFD_SCHEME_PRECON_LUSGS))
        return false;
    switch (p->scheme) {
    case SHCFD_SCHEME_EULER:
    case SHCFD_SCHEME_RK2:
    case SHCFD_SCHEME_LUSGS:
    case SHCFD_SCHEME_NR:
        return true;
    case SHCFD_SCHEME_PRECON:
    case SHCFD_SCHEME_PRECON_LUSGS:
        return p->time_step == 1;
    case SHCFD_SCHEME_DUAL:
        return p->time_step == 0;
    default:
        return false;
    }
}

int shcfd_create(int nblocks, const int* ni, const int* nj, int ny, const shcfd_params* params, shcfd_ctx** out) {
    if (!out || nblocks <= 0 || !ni || !nj || !params || !ny_supported(ny) || !params_valid(params)) return SHCFD_E_ARG;
    *out = nullptr;
    for (int b = 0; b < nblocks; b++)
        if (ni[b] < 2 || nj[b] < 2) return SHCFD_E_ARG;
    shcfd_ctx* ctx = new (std::nothrow) shcfd_ctx();
    if (!ctx) return SHCFD_E_NOMEM;
    ctx->nb = nblocks;
    ctx->ny = ny;
    ctx->dimq = ny + 3;
    ctx->dimw = ny + 8;
    ctx->p = *params;
    long long off = 0;
    int maxni = 0;
    for (int b = 0; b < nblocks; b++) {
        DevBlock B{};
        B.ni = ni[b];
        B.nj = nj[b];
        B.pitch = ((ni[b] + 4) + 7) / 8 * 8;
        B.off = off;
        B.active = 1;
        B.nu = params->nu;
        off += (long long)B.pitch * (nj[b] + 4);
        off = (off + 63) / 64 * 64;
        ctx->hblk.push_back(B);
        maxni = std::max(maxni, ni[b]);
    }
    ctx->NC = off;
    build_tiles(ctx, TJ);
    // sweep strips, ordered by block then ascending j0
    for (int b = 0; b < nblocks; b++) {
        int first = (int)ctx->hstrips.size();
        int ns = (nj[b] + 63) / 64;
        for (int k = 0; k < ns; k++) {
            Strip S{};
            S.b = b;
            S.j0 = 1 + 64 * k;
            S.prev = k > 0 ? first + k - 1 : -1;
            S.next = k < ns - 1 ? first + k + 1 : -1;
            ctx->hstrips.push_back(S);
        }
    }
    ctx->bnd_stride = (long long)(maxni + 2) * ctx->dimq;
    ctx->npart = (int)std::max(ctx->htiles.size(), ctx->hstrips.size());
    hipError_t e;
    e = hipStreamCreate(&ctx->stream);
    if (e != hipSuccess) { delete ctx; return SHCFD_E_HIP; }
    auto al = [&](void** p, size_t bytes) -> bool {
        if (hipMalloc(p, bytes) != hipSuccess) return false;
        return hipMemset(*p, 0, bytes) == hipSuccess;
    };
    bool ok = al((void**)&ctx->dblk, sizeof(DevBlock) * nblocks) &&
              al((void**)&ctx->dtiles, sizeof(Tile) * ctx->htiles.size()) &&
              al((void**)&ctx->dstrips, sizeof(Strip) * ctx->hstrips.size()) &&
              al((void**)&ctx->dprog, sizeof(int) * 2 * ctx->hstrips.size()) &&
              al((void**)&ctx->dbnd, sizeof(double) * ctx->bnd_stride * ctx->hstrips.size()) &&
              al((void**)&ctx->dpart, sizeof(double) * 2 * ctx->npart) &&
              al((void**)&ctx->scal, sizeof(double) * SCAL_COUNT) &&
              al((void**)&ctx->iflag, sizeof(int) * IFLAG_COUNT);
    for (int fld = 0; ok && fld < SHCFD_F_COUNT; fld++)
        ok = al((void**)&ctx->f[fld], sizeof(double) * (size_t)field_ncomp(fld, ny) * (size_t)ctx->NC);
    if (!ok) { shcfd_destroy(ctx); return SHCFD_E_NOMEM; }
    hipMemcpy(ctx->dblk, ctx->hblk.data(), sizeof(DevBlock) * nblocks, hipMemcpyHostToDevice);
    {
        const std::vector<Tile> lt = tiles_in_launch_order(ctx->htiles);
        hipMemcpy(ctx->dtiles, lt.data(), sizeof(Tile) * lt.size(), hipMemcpyHostToDevice);
    }
    hipMemcpy(ctx->dstrips, ctx->hstrips.data(), sizeof(Strip) * ctx->hstrips.size(), hipMemcpyHostToDevice);
    double one = 1.0;
    hipMemcpy(ctx->scal + SCAL_FIRST_DUAL, &one, sizeof(double), hipMemcpyHostToDevice);
    if (hipDeviceSynchronize() != hipSuccess) { shcfd_destroy(ctx); return SHCFD_E_HIP; }
    *out = ctx;
    return SHCFD_OK;
}

int shcfd_destroy(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    drop_graph(ctx);
    if (ctx->fast) { ctx->fast->release(); delete ctx->fast; ctx->fast = nullptr; }
    adjoint_release(ctx);
    for (double* sp : ctx->snaps) if (sp) hipFree(sp);
    for (auto& e : ctx->prof_free) { hipEventDestroy(e.a); hipEventDestroy(e.b); }
    for (int k = 0; k < PROF_COUNT; k++)
        for (auto& e : ctx->prof_pending[k]) { hipEventDestroy(e.a); hipEventDestroy(e.b); }
    for (int fld = 0; fld < SHCFD_F_COUNT; fld++) if (ctx->f[fld]) hipFree(ctx->f[fld]);
    void* ptrs[] = {ctx->dblk, ctx->dtiles, ctx->dstrips, ctx->dprog, ctx->dbnd, ctx->dpart, ctx->scal,
                    ctx->iflag, ctx->dcut, ctx->dbc, ctx->dpool, ctx->dmech, ctx->dskew, ctx->dsbase,
                    ctx->djac16, ctx->dstrips16, ctx->dbnd16, ctx->dnasa9, ctx->dhist, ctx->ddefer,
                    ctx->djcp16, ctx->djsc16};
    for (void* p : ptrs) if (p) hipFree(p);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return SHCFD_OK;
}

int shcfd_set_params(shcfd_ctx* ctx, const shcfd_params* params) {
    if (!ctx || !params || !params_valid(params)) return SHCFD_E_ARG;
    if (params->dimension != ctx->p.dimension) return SHCFD_E_STATE;  // the metrics belong to the geometry
    drop_graph(ctx);  // a captured graph holds the parameters by value
    ctx->p = *params;
    return SHCFD_OK;
}

static int check_field(shcfd_ctx* ctx, int field, int block) {
    if (!ctx || field < 0 || field >= SHCFD_F_COUNT || block < 0 || block >= ctx->nb) return SHCFD_E_ARG;
    return SHCFD_OK;
}

int shcfd_upload_field(shcfd_ctx* ctx, int field, int block, const double* host) {
    int rc = check_field(ctx, field, block);
    if (rc || !host) return SHCFD_E_ARG;
    if (field == SHCFD_F_TGVI || field == SHCFD_F_TGVJ) {
        ctx->tgv_zero = 0;
        drop_graph(ctx);  // captured launches may skip the viscous fluxes
    }
    const DevBlock& B = ctx->hblk[block];
    int K = field_ncomp(field, ctx->ny);
    size_t rows = (size_t)B.nj + 4, w = (size_t)B.ni + 4;
    std::vector<double> plane(rows * B.pitch, 0.0);
    for (int k = 0; k < K; k++) {
        for (size_t j = 0; j < rows; j++)
            for (size_t i = 0; i < w; i++) plane[j * B.pitch + i] = host[k + (size_t)K * (i + w * j)];
        HIPCHK(hipMemcpyAsync(ctx->f[field] + (long long)k * ctx->NC + B.off, plane.data(),
                              sizeof(double) * plane.size(), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return SHCFD_OK;
}

int shcfd_download_field(shcfd_ctx* ctx, int field, int block, double* host) {
    int rc = check_field(ctx, field, block);
    if (rc || !host) return SHCFD_E_ARG;
    const DevBlock& B = ctx->hblk[block];
    int K = field_ncomp(field, ctx->ny);
    size_t rows = (size_t)B.nj + 4, w = (size_t)B.ni + 4;
    std::vector<double> plane(rows * B.pitch, 0.0);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < K; k++) {
        HIPCHK(hipMemcpy(plane.data(), ctx->f[field] + (long long)k * ctx->NC + B.off, sizeof(double) * plane.size(),
                         hipMemcpyDeviceToHost));
        for (size_t j = 0; j < rows; j++)
            for (size_t i = 0; i < w; i++) host[k + (size_t)K * (i + w * j)] = plane[j * B.pitch + i];
    }
    return SHCFD_OK;
}

int shcfd_set_cut_table(shcfd_ctx* ctx, const int* rows, int nrows) {
    if (!ctx || nrows < 0 || (nrows > 0 && !rows)) return SHCFD_E_ARG;
    for (int r = 0; r < nrows; r++) {
        const int* t = rows + SHCFD_CUT_ROW * r;
        int geo = t[0], width = t[1], order = t[2], p1 = t[3], p2 = t[7];
        if (geo < 1 || geo > 4 || width < 0 || (order != 1 && order != -1) || p1 < 1 || p1 > ctx->nb || p2 < 1 ||
            p2 > ctx->nb || std::abs(t[6]) != 1 || std::abs(t[10]) != 1)
            return SHCFD_E_ARG;
        // every cell a copy touches must lie inside its block's ghosted extents
        auto inb = [&](int b, int i, int j) {
            const DevBlock& B = ctx->hblk[b];
            return i >= -1 && i <= B.ni + 2 && j >= -1 && j <= B.nj + 2;
        };
        int incs[2] = {0, width > 0 ? width - 1 : 0};
        for (int q = 0; q < 2 && width > 0; q++)
            for (int copy = 0; copy < 4; copy++) {
                int bd, di, dj, bs, si, sj;
                if (!cut_map(t, incs[q], copy, bd, di, dj, bs, si, sj)) return SHCFD_E_ARG;
                if (!inb(bd, di, dj) || !inb(bs, si, sj)) return SHCFD_E_ARG;
            }
    }
    // a captured graph holds the table's device address and row count
    drop_graph(ctx);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->dcut) { hipFree(ctx->dcut); ctx->dcut = nullptr; }
    ctx->hcut.assign(rows, rows + SHCFD_CUT_ROW * nrows);
    ctx->ncut = nrows;
    if (nrows > 0) {
        HIPCHK(hipMalloc((void**)&ctx->dcut, sizeof(int) * SHCFD_CUT_ROW * nrows));
        HIPCHK(hipMemcpy(ctx->dcut, rows, sizeof(int) * SHCFD_CUT_ROW * nrows, hipMemcpyHostToDevice));
    }
    return SHCFD_OK;
}

int shcfd_set_bc_table(shcfd_ctx* ctx, const int* rows, int nrows, const double* pool, int npool) {
    if (!ctx || nrows < 0 || npool < 0 || (nrows > 0 && !rows) || (npool > 0 && !pool)) return SHCFD_E_ARG;
    int maxlen = 1;
    // per block side: which positions are covered (segments may not overlap)
    std::vector<std::vector<char>> cover(ctx->nb * 4);
    for (int b = 0; b < ctx->nb; b++)
        for (int s = 0; s < 4; s++) cover[b * 4 + s].assign((s <= 1 ? ctx->hblk[b].nj : ctx->hblk[b].ni) + 2, 0);
    for (int r = 0; r < nrows; r++) {
        const int* t = rows + SHCFD_BC_ROW * r;
        int b = t[0], side = t[1], s = t[2], e = t[3], kind = t[4], off = t[5];
        if (b < 1 || b > ctx->nb || side < 0 || side > 3 || kind < 0 || kind > SHCFD_BC_INJECT_WALL || off < 0)
            return SHCFD_E_ARG;
        const int mix = ctx->dimw + ctx->ny;
        int need = kind == SHCFD_BC_FIXED ? mix : kind == SHCFD_BC_INFLOW_TUV ? 3 : kind == SHCFD_BC_OUTFLOW_P ? 1
                 : kind == SHCFD_BC_INFLOW_MIX_P ? mix + 2 : kind == SHCFD_BC_INJECT_WALL ? mix + 3 : 0;
        if (need > 0 && off + need > npool) return SHCFD_E_ARG;
        const DevBlock& B = ctx->hblk[b - 1];
        int n = side <= 1 ? B.nj : B.ni;
        s = std::max(s, 1);
        e = std::min(e, n);
        auto& cv = cover[(b - 1) * 4 + side];
        for (int k = s; k <= e; k++) {
            if (cv[k]) return SHCFD_E_ARG;
            cv[k] = 1;
        }
        maxlen = std::max(maxlen, e - s + 1);
    }
    drop_graph(ctx);  // a captured graph holds the table and pool addresses
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->dbc) { hipFree(ctx->dbc); ctx->dbc = nullptr; }
    if (ctx->dpool) { hipFree(ctx->dpool); ctx->dpool = nullptr; }
    ctx->hbc.assign(rows, rows + SHCFD_BC_ROW * nrows);
    ctx->nbc = nrows;
    ctx->npool = npool;
    ctx->bc_maxlen = maxlen;
    if (nrows > 0) {
        HIPCHK(hipMalloc((void**)&ctx->dbc, sizeof(int) * SHCFD_BC_ROW * nrows));
        HIPCHK(hipMemcpy(ctx->dbc, rows, sizeof(int) * SHCFD_BC_ROW * nrows, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc((void**)&ctx->dpool, sizeof(double) * std::max(npool, 1)));
    if (npool > 0) HIPCHK(hipMemcpy(ctx->dpool, pool, sizeof(double) * npool, hipMemcpyHostToDevice));
    return SHCFD_OK;
}

static int ntiles(shcfd_ctx* ctx) { return (int)ctx->htiles.size(); }
static bool fast_enabled(shcfd_ctx* ctx);

int shcfd_init_geometry(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    DevCtx d = ctx->dev();
    hipLaunchKernelGGL(k_geom_a, dim3(ntiles(ctx)), dim3(256), 0, ctx->stream, d, ctx->dtiles);
    int maxn = 0;
    for (auto& B : ctx->hblk) maxn = std::max(maxn, std::max(B.ni, B.nj) + 2);
    hipLaunchKernelGGL(k_geom_ghost, dim3((maxn + 255) / 256, ctx->nb), dim3(256), 0, ctx->stream, d, 1);
    hipLaunchKernelGGL(k_geom_ghost, dim3((maxn + 255) / 256, ctx->nb), dim3(256), 0, ctx->stream, d, 2);
    hipLaunchKernelGGL(k_geojac, dim3(ntiles(ctx)), dim3(256), 0, ctx->stream, d, ctx->dtiles);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SHCFD_OK;
}

__global__ void k_set_scal(DevCtx d, int slot, double v) { d.scal[slot] = v; }
__global__ void k_store_scal(DevCtx d, int slot, double* dst) { *dst = d.scal[slot]; }

int shcfd_set_time(shcfd_ctx* ctx, double tt, int first_dual) {
    if (!ctx) return SHCFD_E_ARG;
    DevCtx d = ctx->dev();
    hipLaunchKernelGGL(k_set_scal, dim3(1), dim3(1), 0, ctx->stream, d, (int)SCAL_TT, tt);
    hipLaunchKernelGGL(k_set_scal, dim3(1), dim3(1), 0, ctx->stream, d, (int)SCAL_FIRST_DUAL, first_dual ? 1.0 : 0.0);
    LAUNCHCHK();
    ctx->first_dual = first_dual;
    return SHCFD_OK;
}

int shcfd_get_time(shcfd_ctx* ctx, double* dt_grbl, double* tt) {
    if (!ctx) return SHCFD_E_ARG;
    double s[SCAL_COUNT];
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(s, ctx->scal, sizeof(s), hipMemcpyDeviceToHost));
    if (dt_grbl) *dt_grbl = s[SCAL_DT_GRBL];
    if (tt) *tt = s[SCAL_TT];
    return SHCFD_OK;
}

int shcfd_synchronize(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SHCFD_OK;
}

int shcfd_check_error(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    int flags[IFLAG_COUNT];
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(flags, ctx->iflag, sizeof(flags), hipMemcpyDeviceToHost));
    if (flags[IFLAG_ERR]) {
        int z = 0;
        HIPCHK(hipMemcpy(ctx->iflag + IFLAG_ERR, &z, sizeof(int), hipMemcpyHostToDevice));
        return flags[IFLAG_ERR];
    }
    return SHCFD_OK;
}

// ---------------------------------------------------------------------------
// profiler
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {
struct ProfScope {
    Ctx* c;
    int id;
    EventPair ev{};
    ProfScope(Ctx* c_, int id_) : c(c_), id(id_) {
        if (!((c->prof >> id) & 1u)) return;
        if (c->prof_free.empty()) {
            EventPair e{};
            hipEventCreate(&e.a);
            hipEventCreate(&e.b);
            c->prof_free.push_back(e);
        }
        ev = c->prof_free.back();
        c->prof_free.pop_back();
        hipEventRecord(ev.a, c->stream);
    }
    ~ProfScope() {
        if (!((c->prof >> id) & 1u)) return;
        hipEventRecord(ev.b, c->stream);
        c->prof_pending[id].push_back(ev);
    }
};
void prof_collect(Ctx* c) {
    for (int k = 0; k < PROF_COUNT; k++) {
        for (auto& e : c->prof_pending[k]) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, e.a, e.b);
            c->prof_ms[k] += ms;
            c->prof_n[k] += 1;
            c->prof_free.push_back(e);
        }
        c->prof_pending[k].clear();
    }
}
}  // namespace

extern "C" {

int shcfd_profile_enable(shcfd_ctx* ctx, int on) {
    if (!ctx) return SHCFD_E_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    ctx->prof = (unsigned)on;
    return SHCFD_OK;
}

int shcfd_profile_read(shcfd_ctx* ctx, int kernel_class, double* total_ms, long long* launches, int reset) {
    if (!ctx || kernel_class < 0 || kernel_class >= PROF_COUNT) return SHCFD_E_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    if (total_ms) *total_ms = ctx->prof_ms[kernel_class];
    if (launches) *launches = ctx->prof_n[kernel_class];
    if (reset) {
        ctx->prof_ms[kernel_class] = 0.0;
        ctx->prof_n[kernel_class] = 0;
    }
    return SHCFD_OK;
}

// ---------------------------------------------------------------------------
// per-routine drop-ins
// ---------------------------------------------------------------------------
#define TILE_LAUNCH(KERNEL, ...)                                                                   \
    hipLaunchKernelGGL(KERNEL, dim3(ntiles(ctx)), dim3(64 * ctx->tile_rows), 0, ctx->stream, ctx->dev(), \
                       ctx->dtiles, ##__VA_ARGS__)

int shcfd_set_w(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    ProfScope ps(ctx, PROF_PRIMITIVE);
    DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_set_w<NY>));
    LAUNCHCHK();
    return SHCFD_OK;
}

int shcfd_set_thermo_prop(shcfd_ctx* ctx, int step) {
    (void)step;
    if (!ctx) return SHCFD_E_ARG;
    if (ctx->p.thermal == SHCFD_THERMAL_CHEMKIN) {
        if (!ctx->has_mech) return SHCFD_E_STATE;
        ProfScope ps(ctx, PROF_THERMO_CHEM);
        DISPATCH_FLOW_NS(ctx->ny, TILE_LAUNCH(k_thermo_chem<NS>, ctx->dmech));
        LAUNCHCHK();
        return SHCFD_OK;
    }
    if (ctx->p.thermal == SHCFD_THERMAL_COLD_NASA9) {
        if (!ctx->dnasa9) return SHCFD_E_STATE;
        ProfScope ps(ctx, PROF_THERMO_CHEM);
        TILE_LAUNCH(k_thermo_cold, ctx->dnasa9);
        LAUNCHCHK();
        return SHCFD_OK;
    }
    if (ctx->p.thermal != SHCFD_THERMAL_IDEAL) return SHCFD_E_STATE;
    ProfScope ps(ctx, PROF_PRIMITIVE);
    DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_thermo_ideal<NY>));
    LAUNCHCHK();
    return SHCFD_OK;
}

int shcfd_set_mechanism(shcfd_ctx* ctx, const shcfd_mech* mech) {
    if (!ctx || !mech || mech->ns != ctx->ny || mech->nr < 1 || mech->nr > SHCFD_CHEM_MAX_NR ||
        mech->nt < 0 || mech->nt > mech->ns)
        return SHCFD_E_ARG;
    if (!ctx->dmech) HIPCHK(hipMalloc((void**)&ctx->dmech, sizeof(shcfd_mech)));
    HIPCHK(hipMemcpy(ctx->dmech, mech, sizeof(shcfd_mech), hipMemcpyHostToDevice));
    ctx->has_mech = 1;
    return SHCFD_OK;
}

// the cold-flow model's tables (init_pack_cold_flow, NASA/core/sub_chem.f90:
// 380-387): frozen mixing of two streams, nY = 2
int shcfd_set_nasa9(shcfd_ctx* ctx, const shcfd_nasa9* tables) {
    if (!ctx || !tables || ctx->ny != 2 || ctx->p.thermal != SHCFD_THERMAL_COLD_NASA9) return SHCFD_E_ARG;
    const shcfd_nasa9& m = *tables;
    if (m.ns < 1 || m.ns > SHCFD_NASA9_MAX_NS || m.nt < 0 || m.nt > m.ns) return SHCFD_E_ARG;
    for (int i = 0; i < m.ns; i++)
        if (m.nsec[i] < 1 || m.nsec[i] > SHCFD_NASA9_MAX_SEC || !(m.mw[i] > 0.0)) return SHCFD_E_ARG;
    for (int t = 0; t < m.nt; t++)
        if (m.nsec_trans[t] < 1 || m.nsec_trans[t] > SHCFD_NASA9_MAX_TSEC || m.tr2th[t] < 0 || m.tr2th[t] >= m.ns)
            return SHCFD_E_ARG;
    if (!ctx->dnasa9) HIPCHK(hipMalloc((void**)&ctx->dnasa9, sizeof(shcfd_nasa9)));
    HIPCHK(hipMemcpy(ctx->dnasa9, tables, sizeof(shcfd_nasa9), hipMemcpyHostToDevice));
    return SHCFD_OK;
}

// the list CHEMEQ2 cells are handed to DVODE through: room for every
// interior cell, (ns + 3) doubles each; allocated outside graph capture
static int defer_buffer(shcfd_ctx* ctx) {
    if (ctx->ddefer) return SHCFD_OK;
    long long cells = 0;
    for (int b = 0; b < ctx->nb; b++) cells += (long long)ctx->hblk[b].ni * ctx->hblk[b].nj;
    if (cells > INT_MAX) return SHCFD_E_ARG;
    HIPCHK(hipMalloc((void**)&ctx->ddefer, sizeof(double) * (size_t)cells * (size_t)(ctx->ny + 3)));
    ctx->defer_cap = (int)cells;
    return SHCFD_OK;
}

int shcfd_proceed_reaction(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    if (!ctx->has_mech || ctx->p.thermal != SHCFD_THERMAL_CHEMKIN ||
        (ctx->p.reaction != SHCFD_REACTION_CHEMEQ2 && ctx->p.reaction != SHCFD_REACTION_POINT_IMPLICIT))
        return SHCFD_E_STATE;
    if (ctx->p.reaction == SHCFD_REACTION_CHEMEQ2) {
        const int rc = defer_buffer(ctx);
        if (rc) return rc;
    }
    ProfScope ps(ctx, PROF_REACTION);
    if (ctx->p.reaction == SHCFD_REACTION_POINT_IMPLICIT && ctx->ny <= SHCFD_LDS_NEWTON_MAX_NS)
        DISPATCH_FLOW_NS(ctx->ny, hipLaunchKernelGGL(k_reaction_lds<NS>, dim3((unsigned)(ntiles(ctx) * ctx->tile_rows)),
                                                     dim3(64), 0, ctx->stream, ctx->dev(), ctx->dtiles,
                                                     ctx->tile_rows, ctx->dmech))
    else if (ctx->p.reaction == SHCFD_REACTION_POINT_IMPLICIT)
        DISPATCH_FLOW_NS(ctx->ny, TILE_LAUNCH((k_reaction<NS, SHCFD_REACTION_POINT_IMPLICIT>), ctx->dmech, nullptr, 0))
    else {
        // CHEMEQ2, then DVODE for the cells it hands over (sub_chem.f90:1751-1770):
        // the list is emptied, filled by the CHEMEQ2 pass and drained by a
        // grid-stride pass that integrates one listed cell per lane
        hipLaunchKernelGGL(k_set_iflag, dim3(1), dim3(64), 0, ctx->stream, ctx->dev(), (int)IFLAG_DEFER, 0);
        DISPATCH_FLOW_NS(ctx->ny, TILE_LAUNCH((k_reaction<NS, SHCFD_REACTION_CHEMEQ2>), ctx->dmech, ctx->ddefer,
                                              ctx->defer_cap))
        const unsigned g = (unsigned)std::min(((long long)ctx->defer_cap + 63) / 64, 1024LL);
        DISPATCH_FLOW_NS(ctx->ny, hipLaunchKernelGGL(k_reaction_dvode<NS>, dim3(g), dim3(64), 0, ctx->stream,
                                                     ctx->dev(), ctx->dmech, ctx->ddefer, ctx->defer_cap))
    }
    LAUNCHCHK();
    return SHCFD_OK;
}

int shcfd_chem_iterations(shcfd_ctx* ctx, double* total, int reset) {
    if (!ctx || !total) return SHCFD_E_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(total, ctx->scal + SCAL_CHEM_ITERS, sizeof(double), hipMemcpyDeviceToHost));
    if (reset) {
        hipLaunchKernelGGL(k_set_scal, dim3(1), dim3(1), 0, ctx->stream, ctx->dev(), (int)SCAL_CHEM_ITERS, 0.0);
        LAUNCHCHK();
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return SHCFD_OK;
}

// section_exchange and the BC segments; set_corners when `corners` (whole
// steps leave the corners to the fused face-flux kernel, which reads none)
static int bc_async(shcfd_ctx* ctx, bool corners) {
    DevCtx d = ctx->dev();
    ProfScope ps(ctx, PROF_BC);
    if (ctx->ncut > 0) {
        int maxw = 1;
        for (int r = 0; r < ctx->ncut; r++) maxw = std::max(maxw, ctx->hcut[SHCFD_CUT_ROW * r + 1]);
        long long n = 4LL * maxw * ctx->ncut;
        DISPATCH_NY(ctx->ny, hipLaunchKernelGGL(k_bc_cut<NY>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                                                ctx->stream, d, ctx->dcut, ctx->ncut, maxw));
    }
    if (ctx->nbc > 0) {
        long long n = (long long)ctx->bc_maxlen * ctx->nbc;
        DISPATCH_NY(ctx->ny, hipLaunchKernelGGL(k_bc_side<NY>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                                                ctx->stream, d, ctx->dbc, ctx->nbc, ctx->dpool, ctx->bc_maxlen));
    }
    if (corners)
        DISPATCH_NY(ctx->ny, hipLaunchKernelGGL(k_bc_corner<NY>, dim3((4 * ctx->nb + 255) / 256), dim3(256), 0,
                                                ctx->stream, d));
    LAUNCHCHK();
    return SHCFD_OK;
}

int shcfd_set_bc(shcfd_ctx* ctx, int step) {
    (void)step;
    if (!ctx) return SHCFD_E_ARG;
    return bc_async(ctx, true);
}

int shcfd_set_hv(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    ProfScope ps(ctx, PROF_HV);
    DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_set_hv<NY>));
    LAUNCHCHK();
    return SHCFD_OK;
}

int shcfd_set_tg(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    ProfScope ps(ctx, PROF_TG);
    DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_set_tg<NY>));
    LAUNCHCHK();
    return SHCFD_OK;
}

int shcfd_set_tgv(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    ProfScope ps(ctx, PROF_TGV);
    if (!ctx->p.viscous) {
        // init_TGv: viscosity/non-viscous/sch_viscous.f90 (TGv = 0); nothing
        // else writes TGv of an inviscid run, so once is enough
        if (!ctx->tgv_zero) {
            long long n = (long long)ctx->dimq * ctx->NC;
            hipLaunchKernelGGL(k_zero_field, dim3(1024), dim3(256), 0, ctx->stream, ctx->f[SHCFD_F_TGVI], n);
            hipLaunchKernelGGL(k_zero_field, dim3(1024), dim3(256), 0, ctx->stream, ctx->f[SHCFD_F_TGVJ], n);
            ctx->tgv_zero = 1;
        }
    } else {
        DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_set_tgv<NY>));
        ctx->tgv_zero = 0;
    }
    LAUNCHCHK();
    return SHCFD_OK;
}

__global__ void k_fixed_dt(DevCtx d, double dt) {
    d.scal[SCAL_DT_GRBL] = dt;
    d.scal[SCAL_TT] = d.scal[SCAL_TT] + dt;
}

int shcfd_set_dt(shcfd_ctx* ctx, int step, double* dt_grbl_out) {
    (void)step;
    if (!ctx) return SHCFD_E_ARG;
    {
    ProfScope ps(ctx, PROF_DT);
    if (ctx->p.time_step == 0 && ctx->dt_fixed > 0.0) {
        hipLaunchKernelGGL(k_fixed_dt, dim3(1), dim3(1), 0, ctx->stream, ctx->dev(), ctx->dt_fixed);
    } else {
        DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_set_dt<NY>, ctx->dpart));
        if (ctx->p.time_step == 0)
            hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, ctx->stream, ctx->dev(), ctx->dpart, ntiles(ctx),
                               0, (int)SCAL_DT_GRBL);
    }
    }
    LAUNCHCHK();
    if (dt_grbl_out) {
        double s;
        HIPCHK(hipMemcpyAsync(&s, ctx->scal + SCAL_DT_GRBL, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *dt_grbl_out = s;
    }
    return SHCFD_OK;
}

int shcfd_set_fixed_dt(shcfd_ctx* ctx, double dt) {
    if (!ctx || !(dt >= 0.0) || (dt > 0.0 && ctx->p.time_step != 0)) return SHCFD_E_ARG;
    ctx->dt_fixed = dt;
    drop_graph(ctx);
    return SHCFD_OK;
}

// device snapshots of the q-sized state fields
static int snap_field(int field) {
    return field == SHCFD_F_Q || field == SHCFD_F_QP || field == SHCFD_F_QPP;
}

int shcfd_snapshot_save(shcfd_ctx* ctx, int slot, int field) {
    if (!ctx || slot < 0 || slot >= SHCFD_MAX_SNAPSHOTS || !snap_field(field)) return SHCFD_E_ARG;
    const size_t bytes = sizeof(double) * (size_t)ctx->dimq * (size_t)ctx->NC;
    if ((int)ctx->snaps.size() <= slot) ctx->snaps.resize(slot + 1, nullptr);
    if (!ctx->snaps[slot] && hipMalloc((void**)&ctx->snaps[slot], bytes) != hipSuccess) {
        ctx->snaps[slot] = nullptr;
        return SHCFD_E_NOMEM;
    }
    HIPCHK(hipMemcpyAsync(ctx->snaps[slot], ctx->f[field], bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return SHCFD_OK;
}

int shcfd_snapshot_load(shcfd_ctx* ctx, int slot, int field) {
    if (!ctx || slot < 0 || slot >= (int)ctx->snaps.size() || !ctx->snaps[slot] || !snap_field(field))
        return SHCFD_E_ARG;
    const size_t bytes = sizeof(double) * (size_t)ctx->dimq * (size_t)ctx->NC;
    HIPCHK(hipMemcpyAsync(ctx->f[field], ctx->snaps[slot], bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return SHCFD_OK;
}

int shcfd_snapshot_free(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (double* sp : ctx->snaps) if (sp) hipFree(sp);
    ctx->snaps.clear();
    return SHCFD_OK;
}

int shcfd_set_abpm(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    ProfScope ps(ctx, PROF_ABPM);
    DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_set_abpm<NY>));
    LAUNCHCHK();
    ctx->skew_active = 0;
    ctx->q16_active = 0;
    ctx->jd16_active = 0;
    return SHCFD_OK;
}

// the strip-skewed Jacobian copy is used inside whole steps for nY > 1
// (where the optimised dimq = 4 sweeps do not apply); allocated on first use
static bool skew_enabled(shcfd_ctx* ctx) {
    if (!ctx->use_skew || ctx->ny == 1 || ctx->p.dimension != SHCFD_DIM_PLANAR) return false;
    if (ctx->dskew) return true;
    const int ns = (int)ctx->hstrips.size();
    std::vector<long long> sb(ns);
    long long nsl = 0;
    for (int s = 0; s < ns; s++) {
        sb[s] = nsl;
        nsl += 64LL * (ctx->hblk[ctx->hstrips[s].b].ni + 63);
    }
    const long long dq = ctx->dimq;
    if (hipMalloc((void**)&ctx->dskew, sizeof(double) * 4 * dq * dq * nsl) != hipSuccess ||
        hipMalloc((void**)&ctx->dsbase, sizeof(long long) * ns) != hipSuccess ||
        hipMemcpy(ctx->dsbase, sb.data(), sizeof(long long) * ns, hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(ctx->dskew);
        hipFree(ctx->dsbase);
        ctx->dskew = nullptr;
        ctx->dsbase = nullptr;
        ctx->use_skew = 0;
        (void)hipGetLastError();
        return false;
    }
    ctx->skew_nsl = nsl;
    return true;
}

// the 16-row sweeps apply the split Jacobians in closed form (no operator
// copy) for the schemes whose A+-, B+- are the unpreconditioned ones
static bool jvp16_scheme(const shcfd_ctx* ctx) {
    // the Jacobian-free sweeps address the dq and q fields through one buffer
    // view of all their components, whose size must stay below their drop
    // offset (2^31 bytes)
    const bool fits = (long long)(ctx->ny + 3) * ctx->NC * 8 < 0x7FF00000LL;
    return ctx->use_jvp16 && fits && (ctx->p.scheme == SHCFD_SCHEME_LUSGS || ctx->p.scheme == SHCFD_SCHEME_NR);
}

static bool q16_layout(shcfd_ctx* ctx);

// the cell data of the Jacobian-free 16-row sweeps; allocated (and cleared,
// so slots of no cell hold zeros) on first use
static bool jd16_ready(shcfd_ctx* ctx) {
    if (ctx->djcp16) return true;
    const int C = (ctx->dimq + 3) / 4;
    const size_t ncp = (size_t)ctx->nsl16 * C * 64, nsc = (size_t)ctx->nsl16 * 16 * 16;
    if (hipMalloc((void**)&ctx->djcp16, sizeof(double2) * ncp) != hipSuccess ||
        hipMalloc((void**)&ctx->djsc16, sizeof(double) * nsc) != hipSuccess ||
        hipMemset(ctx->djcp16, 0, sizeof(double2) * ncp) != hipSuccess ||
        hipMemset(ctx->djsc16, 0, sizeof(double) * nsc) != hipSuccess) {
        hipFree(ctx->djcp16);
        hipFree(ctx->djsc16);
        ctx->djcp16 = nullptr;
        ctx->djsc16 = nullptr;
        (void)hipGetLastError();
        return false;
    }
    return true;
}

// the 16-row sweeps (nY > 1, every implicit scheme but dual time, whose S
// scaling stays on the row-per-lane sweeps): their strips and hand-off rows,
// and the cell data (LU-SGS, NR) or the quad-row operator copy (the
// preconditioned schemes, or with the closed form switched off); allocated
// on first use
static bool q16_enabled(shcfd_ctx* ctx) {
    if (!ctx->use_skew || ctx->ny == 1 || ctx->p.scheme == SHCFD_SCHEME_DUAL || ctx->p.dimension != SHCFD_DIM_PLANAR)
        return false;
    if (!q16_layout(ctx)) return false;
    if (jvp16_scheme(ctx)) {
        if (jd16_ready(ctx)) return true;
        ctx->use_jvp16 = 0;  // no room: the operator copy below
    }
    if (ctx->djac16) return true;
    const int dq = ctx->dimq, C = (dq + 3) / 4;
    if (hipMalloc((void**)&ctx->djac16, sizeof(double) * 2 * 2 * 64 * C * (dq + 2) * (size_t)ctx->nsl16) != hipSuccess ||
        hipMemset(ctx->djac16, 0, sizeof(double) * 2 * 2 * 64 * C * (dq + 2) * (size_t)ctx->nsl16) != hipSuccess) {
        hipFree(ctx->djac16);
        ctx->djac16 = nullptr;
        ctx->use_skew = 0;
        (void)hipGetLastError();
        return false;
    }
    return true;
}

static bool q16_layout(shcfd_ctx* ctx) {
    if (ctx->dstrips16) return true;
    const int dq = ctx->dimq;
    std::vector<Strip16> st;
    long long nsl = 0;
    int maxni = 0;
    for (int b = 0; b < ctx->nb; b++) {
        const DevBlock& B = ctx->hblk[b];
        maxni = std::max(maxni, B.ni);
        const int first = (int)st.size();
        for (int j0 = 1; j0 <= B.nj; j0 += 16) {
            Strip16 S{};
            S.b = b;
            S.j0 = j0;
            S.prev = (j0 == 1) ? -1 : (int)st.size() - 1;
            S.next = -1;
            S.base = nsl;  // in steps
            if (S.prev >= 0) st[S.prev].next = (int)st.size();
            nsl += B.ni + 15;
            st.push_back(S);
        }
        (void)first;
    }
    const int ns = (int)st.size();
    if (2 * ns > 2 * ctx->npart) return false;  // (res, omega) partials per strip
    // hand-off row: columns -24 .. maxni + 40 (the consumers' loads run ahead)
    const long long bstride = (long long)(maxni + 64) * dq;
    auto al = [](void** p, size_t bytes) {
        if (hipMalloc(p, bytes) != hipSuccess) return false;
        return hipMemset(*p, 0, bytes) == hipSuccess;
    };
    bool ok = al((void**)&ctx->dstrips16, sizeof(Strip16) * ns) &&
              al((void**)&ctx->dbnd16, sizeof(double) * (size_t)(2 * bstride * (ns + 1))) &&
              hipMemcpy(ctx->dstrips16, st.data(), sizeof(Strip16) * ns, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        void* ptrs[] = {ctx->dstrips16, ctx->dbnd16};
        for (void* p : ptrs) if (p) hipFree(p);
        ctx->dstrips16 = nullptr;
        ctx->dbnd16 = nullptr;
        ctx->use_skew = 0;
        (void)hipGetLastError();
        return false;
    }
    ctx->nstrips16 = ns;
    ctx->nsl16 = nsl;
    ctx->bstride16 = bstride;
    return true;
}

// set_ABpm inside a whole step: without stored Jacobians when the optimised
// sweeps run (they need alpha, the pre-conditioner vectors/scalars and the
// local dt); at nY > 1 with the quad-row copy for the 16-row sweeps, or (dual
// time) the skewed copy for the row-per-lane sweeps
static int abpm_for_step(shcfd_ctx* ctx) {
    if (!fast_enabled(ctx)) {
        if (q16_enabled(ctx)) {
            ProfScope ps(ctx, PROF_ABPM);
            int maxni = 0;
            for (const auto& B : ctx->hblk) maxni = std::max(maxni, B.ni);
            dim3 grid((unsigned)ctx->nstrips16, (unsigned)((maxni + 15 + 15) / 16));
            if (jvp16_scheme(ctx)) {
                DISPATCH_NY(ctx->ny, hipLaunchKernelGGL(k_set_jd16<NY>, grid, dim3(256), 0, ctx->stream, ctx->dev(),
                                                        ctx->dstrips16, ctx->djcp16, ctx->djsc16));
                LAUNCHCHK();
                ctx->jd16_active = 1;
                return SHCFD_OK;
            }
            DISPATCH_NY(ctx->ny, hipLaunchKernelGGL(k_set_abpm16<NY>, grid, dim3(256), 0, ctx->stream, ctx->dev(),
                                                    ctx->dstrips16, ctx->djac16, ctx->nsl16));
            LAUNCHCHK();
            ctx->q16_active = 1;
            return SHCFD_OK;
        }
        if (!skew_enabled(ctx)) return shcfd_set_abpm(ctx);
        ProfScope ps(ctx, PROF_ABPM);
        int maxni = 0;
        for (const auto& B : ctx->hblk) maxni = std::max(maxni, B.ni);
        dim3 grid((unsigned)ctx->hstrips.size(), (unsigned)((maxni + 63 + 3) / 4));
        DISPATCH_NY(ctx->ny, hipLaunchKernelGGL(k_set_abpm_skew<NY>, grid, dim3(256), 0, ctx->stream, ctx->dev(),
                                                ctx->dstrips, ctx->dsbase, ctx->dskew, ctx->skew_nsl));
        LAUNCHCHK();
        ctx->skew_active = 1;
        return SHCFD_OK;
    }
    // the optimised sweeps need only set_ABpm's scalars: the RHS pass of the
    // implicit update forms and stores them
    ctx->abpm_inline = 1;
    return SHCFD_OK;
}

int shcfd_set_dqdt(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    ProfScope ps(ctx, PROF_DQDT);
    DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_set_dqdt<NY>));
    LAUNCHCHK();
    return SHCFD_OK;
}

static void prof_begin(Ctx* c, int id) {
    if (!((c->prof >> id) & 1u)) return;
    if (c->prof_free.empty()) {
        EventPair e{};
        hipEventCreate(&e.a);
        hipEventCreate(&e.b);
        c->prof_free.push_back(e);
    }
    c->prof_open[id] = c->prof_free.back();
    c->prof_free.pop_back();
    hipEventRecord(c->prof_open[id].a, c->stream);
}
static void prof_end(Ctx* c, int id) {
    if (!((c->prof >> id) & 1u)) return;
    hipEventRecord(c->prof_open[id].b, c->stream);
    c->prof_pending[id].push_back(c->prof_open[id]);
}

static bool fast_enabled(shcfd_ctx* ctx) {
    if (!ctx->use_fast || ctx->ny != 1 || ctx->p.dimension != SHCFD_DIM_PLANAR) return false;
    if (!ctx->fast) {
        ctx->fast = new FastPath();
        if (ctx->fast->init(*ctx) != SHCFD_OK) {
            ctx->fast->release();
            delete ctx->fast;
            ctx->fast = nullptr;
            ctx->use_fast = 0;
            return false;
        }
    }
    return true;
}

extern "C++" {
namespace shcfd {
// the optimised implicit path of this context, created on first use
bool fast_ready(Ctx* ctx) { return fast_enabled(static_cast<shcfd_ctx*>(ctx)); }
}  // namespace shcfd
}

static int implicit_async(shcfd_ctx* ctx, int step_internal) {
    const int ns = (int)ctx->hstrips.size();
    DevCtx d = ctx->dev();
    int* prog_f = ctx->dprog;
    int* prog_b = ctx->dprog + ns;
    const bool jd16 = ctx->jd16_active && ctx->djcp16 && !fast_enabled(ctx);
    const bool q16 = jd16 || (ctx->q16_active && ctx->djac16 && !fast_enabled(ctx));
    ctx->q16_active = 0;
    ctx->jd16_active = 0;
    const int C16 = (ctx->dimq + 3) / 4, ns16 = ctx->nstrips16;
    if (fast_enabled(ctx)) {
        // the optimised update forms the right-hand side in its prep pass
        int rc = ctx->fast->implicit(*ctx, step_internal, ctx->dqdt_inline, ctx->abpm_inline, prof_begin, prof_end);
        ctx->dqdt_inline = 0;
        ctx->abpm_inline = 0;
        LAUNCHCHK();
        return rc;
    }
    { ProfScope ps(ctx, PROF_RHS);
    DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_rhs<NY>, ctx->dprog, 2 * ns, ctx->dqdt_inline, ctx->abpm_inline)); }
    ctx->dqdt_inline = 0;
    ctx->abpm_inline = 0;
    if (q16) {
        const dim3 blk((unsigned)(64 * C16));
        double* hand_f = ctx->dbnd16;
        double* hand_b = ctx->dbnd16 + ctx->bstride16 * (ns16 + 1);
        { ProfScope ps(ctx, PROF_SWEEP_FWD);
        hipLaunchKernelGGL(k_fill16, dim3(256), dim3(256), 0, ctx->stream, d, ctx->dstrips16, ns16, ctx->dbnd16,
                           ctx->bstride16, ctx->dimq);
        if (jd16)
            DISPATCH_NY(ctx->ny, hipLaunchKernelGGL(k_sweep16j_fwd<NY>, dim3(ns16), blk, 0, ctx->stream, d,
                                                    ctx->dstrips16, ns16, hand_f, ctx->bstride16, ctx->spin_limit,
                                                    ctx->djcp16, ctx->djsc16))
        else
            DISPATCH_NY(ctx->ny, hipLaunchKernelGGL(k_sweep16_fwd<NY>, dim3(ns16), blk, 0, ctx->stream, d,
                                                    ctx->dstrips16, ns16, hand_f, ctx->bstride16, ctx->spin_limit,
                                                    ctx->djac16, ctx->nsl16)) }
        { ProfScope ps(ctx, PROF_SWEEP_BWD);
        if (jd16)
            DISPATCH_NY(ctx->ny, hipLaunchKernelGGL(k_sweep16j_bwd<NY>, dim3(ns16), blk, 0, ctx->stream, d,
                                                    ctx->dstrips16, ns16, hand_b, ctx->bstride16, ctx->dpart,
                                                    ctx->spin_limit, ctx->djcp16, ctx->djsc16))
        else
            DISPATCH_NY(ctx->ny, hipLaunchKernelGGL(k_sweep16_bwd<NY>, dim3(ns16), blk, 0, ctx->stream, d,
                                                    ctx->dstrips16, ns16, hand_b, ctx->bstride16, ctx->dpart,
                                                    ctx->spin_limit, ctx->djac16, ctx->nsl16)) }
        if (jd16 && !inner_loop_scheme(ctx->p.scheme)) {
            ProfScope ps(ctx, PROF_UPDATE);
            DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_add_dq<NY>));
        }
        if (inner_loop_scheme(ctx->p.scheme)) {
            { ProfScope ps(ctx, PROF_REDUCE);
            hipLaunchKernelGGL(k_reduce_res_omega, dim3(1), dim3(256), 0, ctx->stream, d, ctx->dpart, ns16); }
            if (step_internal == 1)
                hipLaunchKernelGGL(k_copy_scal, dim3(1), dim3(1), 0, ctx->stream, d, (int)SCAL_RES1, (int)SCAL_RES);
            if (ctx->defer_update) {
                ctx->update_pending = 1;
            } else {
                ProfScope ps(ctx, PROF_UPDATE);
                DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_update<NY>));
            }
        }
        LAUNCHCHK();
        return SHCFD_OK;
    }
    const bool sk = ctx->skew_active && ctx->dskew;
    ctx->skew_active = 0;
    { ProfScope ps(ctx, PROF_SWEEP_FWD);
    if (sk)
        DISPATCH_NY(ctx->ny, hipLaunchKernelGGL((k_sweep_fwd_plain<NY, true>), dim3(ns), dim3(64), 0, ctx->stream, d,
                                                ctx->dstrips, prog_f, ctx->dbnd, ctx->bnd_stride, ctx->spin_limit,
                                                ctx->dskew, ctx->skew_nsl, ctx->dsbase))
    else
        DISPATCH_NY(ctx->ny, hipLaunchKernelGGL((k_sweep_fwd_plain<NY, false>), dim3(ns), dim3(64), 0, ctx->stream, d,
                                                ctx->dstrips, prog_f, ctx->dbnd, ctx->bnd_stride, ctx->spin_limit,
                                                nullptr, 0LL, nullptr)) }
    { ProfScope ps(ctx, PROF_SWEEP_BWD);
    if (sk)
        DISPATCH_NY(ctx->ny, hipLaunchKernelGGL((k_sweep_bwd_plain<NY, true>), dim3(ns), dim3(64), 0, ctx->stream, d,
                                                ctx->dstrips, ns, prog_b, ctx->dbnd, ctx->bnd_stride, ctx->dpart,
                                                ctx->spin_limit, ctx->dskew, ctx->skew_nsl, ctx->dsbase))
    else
        DISPATCH_NY(ctx->ny, hipLaunchKernelGGL((k_sweep_bwd_plain<NY, false>), dim3(ns), dim3(64), 0, ctx->stream, d,
                                                ctx->dstrips, ns, prog_b, ctx->dbnd, ctx->bnd_stride, ctx->dpart,
                                                ctx->spin_limit, nullptr, 0LL, nullptr)) }
    if (inner_loop_scheme(ctx->p.scheme)) {
        { ProfScope ps(ctx, PROF_REDUCE);
        hipLaunchKernelGGL(k_reduce_res_omega, dim3(1), dim3(256), 0, ctx->stream, d, ctx->dpart, ns); }
        if (step_internal == 1)
            hipLaunchKernelGGL(k_copy_scal, dim3(1), dim3(1), 0, ctx->stream, d, (int)SCAL_RES1, (int)SCAL_RES);
        if (ctx->defer_update) {
            ctx->update_pending = 1;
        } else {
            ProfScope ps(ctx, PROF_UPDATE);
            DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_update<NY>));
        }
    }
    LAUNCHCHK();
    return SHCFD_OK;
}

int shcfd_set_option(shcfd_ctx* ctx, int option, int value) {
    if (!ctx) return SHCFD_E_ARG;
    switch (option) {
    case SHCFD_OPT_FAST_IMPLICIT:
        ctx->use_fast = value ? 1 : 0;
        drop_graph(ctx);
        return SHCFD_OK;
    case SHCFD_OPT_SPIN_LIMIT:
        if (value < 1) return SHCFD_E_ARG;
        ctx->spin_limit = value;
        drop_graph(ctx);  // captured sweep launches hold the bound by value
        return SHCFD_OK;
    case SHCFD_OPT_SKEW_SWEEP:
        ctx->use_skew = value ? 1 : 0;
        drop_graph(ctx);
        return SHCFD_OK;
    case SHCFD_OPT_JVP_SWEEP:
        ctx->use_jvp16 = value ? 1 : 0;
        drop_graph(ctx);
        return SHCFD_OK;
    case SHCFD_OPT_TILE_ROWS: {
        if (value != 1 && value != 2 && value != 4) return SHCFD_E_ARG;
        if (value == ctx->tile_rows) return SHCFD_OK;
        HIPCHK(hipStreamSynchronize(ctx->stream));
        build_tiles(ctx, value);
        Tile* nt = nullptr;
        double* np = nullptr;
        int npart = (int)std::max(ctx->htiles.size(), ctx->hstrips.size());
        if (hipMalloc((void**)&nt, sizeof(Tile) * ctx->htiles.size()) != hipSuccess ||
            hipMalloc((void**)&np, sizeof(double) * 2 * npart) != hipSuccess) {
            hipFree(nt);
            return SHCFD_E_NOMEM;
        }
        const std::vector<Tile> lt = tiles_in_launch_order(ctx->htiles);
        HIPCHK(hipMemcpy(nt, lt.data(), sizeof(Tile) * lt.size(), hipMemcpyHostToDevice));
        hipFree(ctx->dtiles);
        hipFree(ctx->dpart);
        ctx->dtiles = nt;
        ctx->dpart = np;
        ctx->npart = npart;
        drop_graph(ctx);
        return SHCFD_OK;
    }
    default:
        return SHCFD_E_ARG;
    }
}

int shcfd_calc_next_step_implicit(shcfd_ctx* ctx, int step_internal, double* res_out, double* omega_out) {
    if (!ctx) return SHCFD_E_ARG;
    int s = ctx->p.scheme;
    if (s != SHCFD_SCHEME_LUSGS && s != SHCFD_SCHEME_PRECON_LUSGS && !inner_loop_scheme(s)) return SHCFD_E_STATE;
    int rc = implicit_async(ctx, step_internal);
    if (rc) return rc;
    if (res_out || omega_out) {
        double sc[SCAL_COUNT];
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipMemcpy(sc, ctx->scal, sizeof(sc), hipMemcpyDeviceToHost));
        if (res_out) *res_out = sc[SCAL_RES];
        if (omega_out) *omega_out = sc[SCAL_OMEGA];
        return shcfd_check_error(ctx);
    }
    return SHCFD_OK;
}

int shcfd_calc_next_step(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    ProfScope ps(ctx, PROF_EXPLICIT);
    DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_explicit<NY>, 0));
    LAUNCHCHK();
    return SHCFD_OK;
}

int shcfd_calc_next_step_rk2(shcfd_ctx* ctx, int stage) {
    if (!ctx || (stage != 1 && stage != 2)) return SHCFD_E_ARG;
    ProfScope ps(ctx, PROF_EXPLICIT);
    DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_explicit<NY>, stage));
    LAUNCHCHK();
    return SHCFD_OK;
}

int shcfd_copy_state(shcfd_ctx* ctx, int dst, int src) {
    if (!ctx) return SHCFD_E_ARG;
    auto okf = [](int f) { return f == SHCFD_F_Q || f == SHCFD_F_QP || f == SHCFD_F_QPP; };
    if (!okf(dst) || !okf(src) || dst == src) return SHCFD_E_ARG;
    DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_copy_q<NY>, dst, src));
    LAUNCHCHK();
    return SHCFD_OK;
}

int shcfd_energy_rms(shcfd_ctx* ctx, double* rms_out) {
    if (!ctx || !rms_out) return SHCFD_E_ARG;
    DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_rms_part<NY>, ctx->dpart));
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, ctx->stream, ctx->dev(), ctx->dpart, ntiles(ctx), 1,
                       (int)SCAL_RMS);
    LAUNCHCHK();
    double s;
    HIPCHK(hipMemcpyAsync(&s, ctx->scal + SCAL_RMS, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double cells = 0.0;
    for (auto& B : ctx->hblk) cells = cells + (double)(B.ni * B.nj);
    *rms_out = sqrt(s) / cells;
    return SHCFD_OK;
}

// ---------------------------------------------------------------------------
// whole steps
// ---------------------------------------------------------------------------
// q += omega Dq of an inner iteration whose update was deferred
static int apply_pending_update(shcfd_ctx* ctx) {
    if (!ctx->update_pending) return SHCFD_OK;
    ctx->update_pending = 0;
    ProfScope ps(ctx, PROF_UPDATE);
    DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_update<NY>));
    LAUNCHCHK();
    return SHCFD_OK;
}

// set_w + set_thermo_prop (ideal gas: one fused pass, which also applies a
// deferred inner-iteration update)
// copy_qp: the step opens with QP = Q, which the ideal-gas pass stores as it
// reads q
static int part_primitive(shcfd_ctx* ctx, bool copy_qp = false) {
    if (ctx->p.thermal == SHCFD_THERMAL_IDEAL) {
        const int flags = (ctx->update_pending ? PRIM_APPLY_UPDATE : 0) | (copy_qp ? PRIM_COPY_QP : 0);
        ctx->update_pending = 0;
        ProfScope ps(ctx, PROF_PRIMITIVE);
        DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_primitive_ideal<NY>, flags));
        LAUNCHCHK();
        return SHCFD_OK;
    }
    if (copy_qp) {
        const int rc = shcfd_copy_state(ctx, SHCFD_F_QP, SHCFD_F_Q);
        if (rc) return rc;
    }
    int rc = apply_pending_update(ctx);
    if (rc) return rc;
    rc = shcfd_set_w(ctx);
    if (!rc) rc = shcfd_set_thermo_prop(ctx, 0);
    return rc;
}
// set_BC, set_HV + set_TG (one fused kernel: the face states are not
// stored; it also fills the corner ghosts, set_corners), set_TGv
static int part_main(shcfd_ctx* ctx) {
    int rc = bc_async(ctx, false);
    if (rc) return rc;
    {
        ProfScope ps(ctx, PROF_HV);
        DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_set_hv_tg<NY>));
        LAUNCHCHK();
    }
    return shcfd_set_tgv(ctx);
}

int shcfd_prepare(shcfd_ctx* ctx) {
    if (!ctx) return SHCFD_E_ARG;
    int rc = part_primitive(ctx);
    if (!rc) rc = shcfd_set_bc(ctx, 0);
    return rc;
}

// one physical step of a scheme without host-side checks inside the step
// (all but dual time): only kernel launches on the context stream, so it can
// be captured into a graph
static int enqueue_step(shcfd_ctx* ctx, int step) {
    int rc = SHCFD_OK;
#define TRY(x) do { rc = (x); if (rc) return rc; } while (0)
    switch (ctx->p.scheme) {
    case SHCFD_SCHEME_LUSGS:
        if (ctx->p.reaction) {  // LU-SGS/2d/main.main.point_implicit.f90:1-9
            TRY(shcfd_set_w(ctx));
            TRY(shcfd_set_thermo_prop(ctx, step));
            TRY(shcfd_set_dt(ctx, step, nullptr));
            TRY(shcfd_proceed_reaction(ctx));
            TRY(part_primitive(ctx, true));  // QP = Q, then set_w
            TRY(part_main(ctx));
            TRY(abpm_for_step(ctx));
            TRY(implicit_async(ctx, 0));
            break;
        }
        TRY(part_primitive(ctx, true));  // QP = Q, then set_w
        TRY(shcfd_set_dt(ctx, step, nullptr));
        TRY(part_main(ctx));
        TRY(abpm_for_step(ctx));
        TRY(implicit_async(ctx, 0));
        break;
    case SHCFD_SCHEME_PRECON_LUSGS:
        TRY(part_primitive(ctx, true));  // QP = Q, then set_w
        TRY(part_main(ctx));
        TRY(abpm_for_step(ctx));
        TRY(implicit_async(ctx, 0));
        break;
    case SHCFD_SCHEME_EULER:
        TRY(part_primitive(ctx, true));  // QP = Q, then set_w
        TRY(shcfd_set_dt(ctx, step, nullptr));
        TRY(part_main(ctx));
        TRY(shcfd_calc_next_step(ctx));
        break;
    case SHCFD_SCHEME_RK2:
        if (ctx->p.reaction) {  // RK2/main.main.point_implicit.f90:1-11
            TRY(shcfd_set_w(ctx));
            TRY(shcfd_set_thermo_prop(ctx, step));
            TRY(shcfd_set_dt(ctx, step, nullptr));
            TRY(shcfd_proceed_reaction(ctx));
            TRY(part_primitive(ctx, true));  // QP = Q, then set_w
        } else {
            TRY(part_primitive(ctx, true));  // QP = Q, then set_w
            TRY(shcfd_set_dt(ctx, step, nullptr));
        }
        TRY(part_main(ctx));
        TRY(shcfd_calc_next_step_rk2(ctx, 1));
        TRY(part_primitive(ctx));
        TRY(part_main(ctx));
        TRY(shcfd_calc_next_step_rk2(ctx, 2));
        break;
    default:
        return SHCFD_E_STATE;
    }
#undef TRY
    return rc;
}

// The launches of one physical step of an inner-loop scheme (NR, precon,
// dual): NR/2d/main.main.raw.f90:1-20, precon/2d/main.main.raw.f90:1-19 (no
// set_dt: set_ABpm sets dt_mat), dual/2d/main.main.raw.f90:1-20.  Only kernel
// launches on the context stream, so the whole step -- every inner iteration,
// its (res, omega) reductions and the deferred update -- can be captured into
// one graph.  hist (may be NULL) receives (res1, res) of the step, for the
// host's check_internal_loop_convergence after a graph replay.  The step ends
// by clearing the BDF1 flag (first_dual) on the device.
static int enqueue_inner_step(shcfd_ctx* ctx, double* hist) {
    int rc = SHCFD_OK;
#define TRY(x) do { rc = (x); if (rc) return rc; } while (0)
    TRY(part_primitive(ctx));
    if (ctx->p.scheme != SHCFD_SCHEME_PRECON) TRY(shcfd_set_dt(ctx, 0, nullptr));
    TRY(shcfd_copy_state(ctx, SHCFD_F_QPP, SHCFD_F_QP));
    TRY(shcfd_copy_state(ctx, SHCFD_F_QP, SHCFD_F_Q));
    for (int j = 1; j < ctx->p.num_internal_loop; j++) {
        TRY(part_primitive(ctx));
        TRY(part_main(ctx));
        TRY(abpm_for_step(ctx));
        ctx->dqdt_inline = 1;   // set_dqdt, formed in the RHS pass
        ctx->defer_update = 1;  // q += omega Dq, in the next primitive pass
        rc = implicit_async(ctx, j);
        ctx->defer_update = 0;
        if (rc) return rc;
    }
    TRY(apply_pending_update(ctx));
#undef TRY
    if (hist) {
        hipLaunchKernelGGL(k_store_scal, dim3(1), dim3(1), 0, ctx->stream, ctx->dev(), (int)SCAL_RES1, hist);
        hipLaunchKernelGGL(k_store_scal, dim3(1), dim3(1), 0, ctx->stream, ctx->dev(), (int)SCAL_RES, hist + 1);
    }
    hipLaunchKernelGGL(k_set_scal, dim3(1), dim3(1), 0, ctx->stream, ctx->dev(), (int)SCAL_FIRST_DUAL, 0.0);
    LAUNCHCHK();
    return SHCFD_OK;
}

// check_internal_loop_convergence: sch_dual.raw.f90:284-299 (the same routine
// in sch_NR.raw.f90:216-231 and sch_precon.raw.f90:276-291)
static int check_inner_rate(shcfd_ctx* ctx, int step, int step_res, double res1, double res) {
    if (res != res || res1 != res1) return SHCFD_E_NAN;
    const double rate = res / res1;
    if (rate > ctx->p.res_rate_warning) {
        fprintf(stderr, "WARNING:NOT CONVERGED AT INTERNAL LOOP (step,res_rate) = %12d %11.3E\n", step, rate);
        if (rate > ctx->p.res_rate_error && step != step_res + 1) return SHCFD_E_RESRATE;
    }
    return SHCFD_OK;
}

int shcfd_step(shcfd_ctx* ctx, int step, int step_res, double* res_first, double* res_last) {
    if (!ctx) return SHCFD_E_ARG;
    if (!inner_loop_scheme(ctx->p.scheme)) {
        int rc = enqueue_step(ctx, step);
        return rc ? rc : shcfd_check_error(ctx);
    }
    int rc = enqueue_inner_step(ctx, nullptr);
    if (rc) return rc;
    ctx->first_dual = 0;
    double sc[SCAL_COUNT];
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(sc, ctx->scal, sizeof(sc), hipMemcpyDeviceToHost));
    if (res_first) *res_first = sc[SCAL_RES1];
    if (res_last) *res_last = sc[SCAL_RES];
    rc = shcfd_check_error(ctx);
    if (rc) return rc;
    return check_inner_rate(ctx, step, step_res, sc[SCAL_RES1], sc[SCAL_RES]);
}

// ---------------------------------------------------------------------------
// batches of independent blocks
// ---------------------------------------------------------------------------
static int upload_blocks(shcfd_ctx* ctx) {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(ctx->dblk, ctx->hblk.data(), sizeof(DevBlock) * ctx->nb, hipMemcpyHostToDevice));
    return SHCFD_OK;
}

int shcfd_set_block_nu(shcfd_ctx* ctx, const double* nu) {
    if (!ctx || !nu) return SHCFD_E_ARG;
    for (int b = 0; b < ctx->nb; b++) {
        if (!(nu[b] >= 0.0)) return SHCFD_E_ARG;
        ctx->hblk[b].nu = nu[b];
    }
    return upload_blocks(ctx);
}

int shcfd_set_block_active(shcfd_ctx* ctx, const int* active) {
    if (!ctx || !active) return SHCFD_E_ARG;
    for (int b = 0; b < ctx->nb; b++) ctx->hblk[b].active = active[b] ? 1 : 0;
    return upload_blocks(ctx);
}

int shcfd_energy_rms_blocks(shcfd_ctx* ctx, double* rms) {
    if (!ctx || !rms) return SHCFD_E_ARG;
    DISPATCH_NY(ctx->ny, TILE_LAUNCH(k_rms_part<NY>, ctx->dpart));
    LAUNCHCHK();
    std::vector<double> part(ctx->htiles.size());
    HIPCHK(hipMemcpyAsync(part.data(), ctx->dpart, sizeof(double) * part.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::vector<double> acc(ctx->nb, 0.0);
    for (size_t k = 0; k < part.size(); k++) acc[ctx->htiles[k].b] = acc[ctx->htiles[k].b] + part[k];
    for (int b = 0; b < ctx->nb; b++)
        rms[b] = sqrt(acc[b]) / (double)(ctx->hblk[b].ni * ctx->hblk[b].nj);
    return SHCFD_OK;
}

// allocate everything a step uses, outside any graph capture: the optimised
// nY = 1 path, else the quad-row copy of the 16-row sweeps (nY > 1, all
// implicit schemes but dual time), else the strip-skewed copy (nY > 1 dual
// time, or when the quad-row copy could not be allocated)
static void warm_allocations(shcfd_ctx* ctx) {
    if (ctx->p.thermal == SHCFD_THERMAL_CHEMKIN && ctx->p.reaction == SHCFD_REACTION_CHEMEQ2) defer_buffer(ctx);
    const int s = ctx->p.scheme;
    const bool implicit = s == SHCFD_SCHEME_LUSGS || s == SHCFD_SCHEME_PRECON_LUSGS || inner_loop_scheme(s);
    if (!implicit || fast_enabled(ctx)) return;
    if (!q16_enabled(ctx)) skew_enabled(ctx);
}

int shcfd_iterate(shcfd_ctx* ctx, int first_step, int nsteps, int graph_steps) {
    if (!ctx || nsteps < 0) return SHCFD_E_ARG;
    if (graph_steps < 0) graph_steps = ctx->tuned_graph_steps;
    const bool inner = inner_loop_scheme(ctx->p.scheme);
    const int step_res = first_step - 1;
    warm_allocations(ctx);
    int done = 0;
    if (graph_steps > 0 && nsteps >= graph_steps) {
        if (ctx->graph_steps != graph_steps) {
            drop_graph(ctx);
            if (inner && ctx->nhist < graph_steps) {
                if (ctx->dhist) hipFree(ctx->dhist);
                ctx->dhist = nullptr;
                ctx->nhist = 0;
                HIPCHK(hipMalloc((void**)&ctx->dhist, sizeof(double) * 2 * graph_steps));
                ctx->nhist = graph_steps;
            }
            const unsigned prof = ctx->prof;
            ctx->prof = 0;  // no event records inside the graph
            HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            int rc = SHCFD_OK;
            for (int k = 0; k < graph_steps && !rc; k++)
                rc = inner ? enqueue_inner_step(ctx, ctx->dhist + 2 * k) : enqueue_step(ctx, first_step + k);
            hipGraph_t g = nullptr;
            hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
            ctx->prof = prof;
            if (rc || ec != hipSuccess) {
                if (g) hipGraphDestroy(g);
                return rc ? rc : SHCFD_E_HIP;
            }
            hipError_t ei = hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (ei != hipSuccess) {
                ctx->graph = nullptr;
                return SHCFD_E_HIP;
            }
            ctx->graph_steps = graph_steps;
        }
        for (; done + graph_steps <= nsteps; done += graph_steps) {
            HIPCHK(hipGraphLaunch(ctx->graph, ctx->stream));
            if (!inner) continue;
            // the host checks of each replayed step, in step order
            ctx->first_dual = 0;
            std::vector<double> h(2 * (size_t)graph_steps);
            HIPCHK(hipMemcpyAsync(h.data(), ctx->dhist, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                                  ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            int rc = shcfd_check_error(ctx);
            if (rc) return rc;
            for (int k = 0; k < graph_steps; k++) {
                rc = check_inner_rate(ctx, first_step + done + k, step_res, h[2 * k], h[2 * k + 1]);
                if (rc) return rc;
            }
        }
    }
    for (; done < nsteps; done++) {
        int rc = inner ? shcfd_step(ctx, first_step + done, step_res, nullptr, nullptr)
                       : enqueue_step(ctx, first_step + done);
        if (rc) return rc;
    }
    return shcfd_check_error(ctx);
}

// Time each candidate schedule on the context's own blocks and keep the
// fastest: optimised or per-routine sweeps (nY = 1), tile rows 1/2/4 of the
// per-cell kernels, and eager launches or a graph of 8 steps.  The flow state
// (q, qp, qpp and the time scalars) is saved before and restored after, so
// tuning does not advance the solution.  choice[4] = fast, tile rows, graph
// steps, best ms per step (may be NULL).
int shcfd_autotune(shcfd_ctx* ctx, int trial_steps, double* choice) {
    if (!ctx || trial_steps < 1) return SHCFD_E_ARG;
    const size_t nq = (size_t)ctx->dimq * (size_t)ctx->NC;
    const int first_dual = ctx->first_dual;
    double* save = nullptr;
    double* ssave = nullptr;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (hipMalloc((void**)&save, sizeof(double) * 3 * nq) != hipSuccess ||
        hipMalloc((void**)&ssave, sizeof(double) * SCAL_COUNT) != hipSuccess) {
        hipFree(save);
        return SHCFD_E_NOMEM;
    }
    const int fields[3] = {SHCFD_F_Q, SHCFD_F_QP, SHCFD_F_QPP};
    for (int k = 0; k < 3; k++)
        HIPCHK(hipMemcpy(save + k * nq, ctx->f[fields[k]], sizeof(double) * nq, hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(ssave, ctx->scal, sizeof(double) * SCAL_COUNT, hipMemcpyDeviceToDevice));
    auto restore = [&]() -> int {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        for (int k = 0; k < 3; k++)
            HIPCHK(hipMemcpy(ctx->f[fields[k]], save + k * nq, sizeof(double) * nq, hipMemcpyDeviceToDevice));
        HIPCHK(hipMemcpy(ctx->scal, ssave, sizeof(double) * SCAL_COUNT, hipMemcpyDeviceToDevice));
        ctx->first_dual = first_dual;
        return SHCFD_OK;
    };
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 1e300;
    int bf = ctx->use_fast, br = ctx->tile_rows, bg = 0, rc = SHCFD_OK;
    const int fasts[2] = {1, 0};
    const int nfast = ctx->ny == 1 ? 2 : 1;
    for (int fi = 0; fi < nfast && !rc; fi++)
        for (int rows = 1; rows <= 4 && !rc; rows *= 2)
            for (int gs = 0; gs <= 8 && !rc; gs += 8) {
                rc = shcfd_set_option(ctx, SHCFD_OPT_FAST_IMPLICIT, fasts[fi]);
                if (!rc) rc = shcfd_set_option(ctx, SHCFD_OPT_TILE_ROWS, rows);
                if (!rc) rc = shcfd_iterate(ctx, 1, gs > 0 ? gs : 1, gs);  // warm-up (and capture)
                if (rc) break;
                const int n = gs > 0 ? ((trial_steps + gs - 1) / gs) * gs : trial_steps;
                hipEventRecord(e0, ctx->stream);
                rc = shcfd_iterate(ctx, 1, n, gs);
                hipEventRecord(e1, ctx->stream);
                hipEventSynchronize(e1);
                float ms = 0.0f;
                hipEventElapsedTime(&ms, e0, e1);
                const double per = (double)ms / n;
                if (!rc && per < best) {
                    best = per;
                    bf = fasts[fi];
                    br = rows;
                    bg = gs;
                }
                int r2 = restore();
                if (!rc) rc = r2;
            }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(save);
    hipFree(ssave);
    if (rc) return rc;
    rc = shcfd_set_option(ctx, SHCFD_OPT_FAST_IMPLICIT, bf);
    if (!rc) rc = shcfd_set_option(ctx, SHCFD_OPT_TILE_ROWS, br);
    ctx->tuned_graph_steps = bg;
    if (choice) {
        choice[0] = bf;
        choice[1] = br;
        choice[2] = bg;
        choice[3] = best;
    }
    return rc;
}

}  // extern "C"
