// shcfd_internal.h -- device data layout and shared helpers of libshcfd.
//
// Device layout (all blocks resident in one GPU's HBM):
//   * one global "cell index space": block b owns the (ni+4) x (nj+4)
//     ghosted cells i in [-1, ni+2], j in [-1, nj+2], row pitch `pitch`
//     (>= ni+4, multiple of 8 doubles), starting at `off` (multiple of 64);
//     NC is the total size.  Cell, i-face, j-face and node quantities all
//     use this index space (face (i,j) = the face between cell i and i+1).
//   * every field is structure-of-arrays: component k of field F lives at
//     f[F][k*NC + cell], so consecutive i are consecutive addresses.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>
#include <string>

#include "../../include/shcfd_types.h"

namespace shcfd {

struct DevBlock {
    int ni, nj, pitch, active;  // active = 0: the updates and the chemistry leave q unchanged
    long long off;              // index of cell (-1,-1)
    double nu;                  // kinematic viscosity of the ideal gas (mu = rho nu)
};

// Device-side view passed by value to every kernel.
struct DevCtx {
    int nb, ny, dimq, dimw;
    long long NC;
    const DevBlock* blk;
    double* f[SHCFD_F_COUNT];
    double* scal;     // device scalars, see SCAL_* below
    int* iflag;       // device int flags (error codes, counters)
    int tgv_zero;     // TGVI/TGVJ hold zeros: the flux balance need not read them
    shcfd_params p;
};

enum {
    SCAL_DT_GRBL = 0,
    SCAL_TT = 1,
    SCAL_OMEGA = 2,     // min omega of the last implicit update
    SCAL_RES = 3,       // res of the last implicit update
    SCAL_RES1 = 4,
    SCAL_RMS = 5,
    SCAL_FIRST_DUAL = 6,
    SCAL_CHEM_ITERS = 7,  // Newton iterations / CHEMEQ2 attempts summed over cells
    SCAL_COUNT = 16
};
enum { IFLAG_ERR = 0, IFLAG_DEFER = 1, IFLAG_COUNT = 8 };  // IFLAG_DEFER: records in the DVODE hand-over list

// Schemes whose physical step runs the inner loop of dual time: dq/dt from
// q, qp, qpp (set_dqdt), an omega-limited update and the residual norm res
// (NR/2d, precon/2d and dual/2d time_*_viscous.f90).
__host__ __device__ inline bool inner_loop_scheme(int s) {
    return s == SHCFD_SCHEME_NR || s == SHCFD_SCHEME_PRECON || s == SHCFD_SCHEME_DUAL;
}
// Schemes whose set_ABpm builds the low-Mach preconditioned operator
// (sch_precon.raw.f90, sch_dual.raw.f90).
__host__ __device__ inline bool precond_scheme(int s) {
    return s == SHCFD_SCHEME_PRECON || s == SHCFD_SCHEME_PRECON_LUSGS || s == SHCFD_SCHEME_DUAL;
}

__host__ __device__ inline long long cidx(const DevBlock& B, int i, int j) {
    return B.off + (long long)(j + 1) * B.pitch + (i + 1);
}

// 0-based component indices of w
#define IW_RHO 0
#define IW_U 1
#define IW_V 2
#define IW_P 3
#define IW_Y(k) (4 + (k))

// Tiles of 64 x 4 cells covering i in [0, ni+1], j in [0, nj+1] of each
// block; one 256-thread workgroup per tile.
constexpr int TI = 64;
constexpr int TJ = 4;

struct Tile {
    int b, i0, j0;
    int slot;  // the tile's position in block order: its slot in per-tile partial arrays
};

// One wavefront-strip of the LU-SGS sweeps: 64 consecutive rows of a block.
struct Strip {
    int b;       // block
    int j0;      // first row (1-based)
    int prev;    // strip index of the row range below (forward) or -1
    int next;    // strip index of the row range above (backward) or -1
};

// 16-row strip of the quad-row sweeps for nY > 1 (k_sweep16.hip): lane
// 4r + c of wave w holds component 4w + c of row j0 + r; sweep step t
// processes column t - r + 1 of row r (forward).  base = the strip's first
// global step in the operator copy.
struct Strip16 {
    int b, j0, prev, next;
    long long base;
};

// kernel classes timed by the optional HIP-event profiler
enum {
    PROF_PRIMITIVE = 0, PROF_BC, PROF_HV, PROF_TG, PROF_TGV, PROF_DT, PROF_ABPM, PROF_DQDT, PROF_RHS,
    PROF_SWEEP_FWD, PROF_SWEEP_BWD, PROF_REDUCE, PROF_UPDATE, PROF_EXPLICIT, PROF_RESIDUAL, PROF_PREP,
    PROF_THERMO_CHEM, PROF_REACTION, PROF_COUNT
};

struct EventPair {
    hipEvent_t a, b;
};

// ---------------------------------------------------------------------
// host-side context
// ---------------------------------------------------------------------
struct Ctx {
    // profiler: events recorded around each launch of a kernel class
    unsigned prof = 0;             // bit k set: class k is timed
    std::vector<EventPair> prof_pending[PROF_COUNT];
    std::vector<EventPair> prof_free;
    double prof_ms[PROF_COUNT] = {};
    long long prof_n[PROF_COUNT] = {};
    EventPair prof_open[PROF_COUNT] = {};
    // optimised implicit update (k_fast.hip); owned, created on first use
    struct FastPath* fast = nullptr;
    // discrete adjoint (shcfd_adjoint.hip); owned, created on first use
    struct Adjoint* adj = nullptr;
    int use_fast = 1;              // 1: optimised sweeps when available
    int nb = 0, ny = 0, dimq = 0, dimw = 0;
    shcfd_params p{};
    std::vector<DevBlock> hblk;
    long long NC = 0;
    DevBlock* dblk = nullptr;
    double* f[SHCFD_F_COUNT] = {};
    double* scal = nullptr;
    int* iflag = nullptr;
    // tables
    std::vector<Tile> htiles;
    Tile* dtiles = nullptr;
    std::vector<Strip> hstrips;
    Strip* dstrips = nullptr;
    int* dprog = nullptr;          // per-strip progress counters (fwd, bwd)
    double* dbnd = nullptr;        // per-strip boundary rows (dimq per column)
    long long bnd_stride = 0;      // doubles per strip in dbnd
    int* dcut = nullptr;
    int ncut = 0;
    int* dbc = nullptr;
    int nbc = 0;
    double* dpool = nullptr;
    int npool = 0;
    int bc_maxlen = 0;
    std::vector<int> hbc, hcut;
    // reduction scratch
    double* dpart = nullptr;       // per-tile / per-strip partials (2 per entry)
    int npart = 0;
    // host mirrors of device scalars
    double dt_grbl = 0.0, tt = 0.0;
    int first_dual = 1;
    double res1 = 0.0;
    double dt_fixed = 0.0;         // > 0: global time step held at this value
    int tgv_zero = 0;              // TGVI/TGVJ hold zeros (inviscid set_TGv done)
    std::vector<double*> snaps;    // device snapshots of q-sized fields (dimq*NC doubles each)
    hipStream_t stream = nullptr;
    std::string last_error;
    int sweep_impl = 0;            // 0 = plain (stored Jacobians)
    long long spin_limit = 1LL << 23;
    // Chemkin mechanism (module chem of the reference), device copy
    shcfd_mech* dmech = nullptr;
    shcfd_nasa9* dnasa9 = nullptr;  // cold-flow tables (SHCFD_THERMAL_COLD_NASA9)
    int has_mech = 0;
    int tile_rows = TJ;            // rows per tile of the per-cell kernels (1, 2 or 4)
    // captured whole-step graphs (shcfd_iterate)
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    int tuned_graph_steps = 0;     // schedule chosen by shcfd_autotune
    // strip-skewed Jacobian copy for the row-per-lane sweeps (nY > 1 steps)
    int use_skew = 1;
    // quad-row sweeps for nY > 1 and schemes without the dual-time S scaling
    Strip16* dstrips16 = nullptr;
    int nstrips16 = 0;
    double* djac16 = nullptr;      // operator copy: 2 sweeps x nsl16 steps x C x (dq + 2) x 64 pairs
    long long nsl16 = 0;           // total steps of all strips
    double* dbnd16 = nullptr;      // hand-off rows of both sweeps: 2 x (nstrips16 + 1) x bstride16
    long long bstride16 = 0;
    int dqdt_inline = 0;           // the next implicit update forms dq/dt in its RHS pass
    int abpm_inline = 0;           // the next implicit update forms set_ABpm's scalars in its RHS pass
    int defer_update = 0;          // the next inner iteration leaves q += omega Dq to the next primitive pass
    int update_pending = 0;        // q += omega Dq is still to be applied
    double* dhist = nullptr;       // (res1, res) per physical step of an inner-loop graph
    int nhist = 0;
    double* ddefer = nullptr;      // CHEMEQ2 cells handed to DVODE: (cell, t, n(ns), T) per record
    int defer_cap = 0;             // records: one per interior cell
    int q16_active = 0;            // the last set_ABpm of a step wrote the quad-row copy
    int use_jvp16 = 1;             // LU-SGS / NR at nY > 1: Jacobian-free 16-row sweeps
    int jd16_active = 0;           // the last set_ABpm of a step wrote their cell data
    double2* djcp16 = nullptr;     // their per-lane (c1, DHi) pairs: nsl16 x C x 64
    double* djsc16 = nullptr;      // their cell scalars: nsl16 x 16 rows x 16
    int skew_active = 0;           // the last set_ABpm of a step wrote the skewed copy
    double* dskew = nullptr;
    long long skew_nsl = 0;        // slots: sum over strips of 64 (ni + 63)
    long long* dsbase = nullptr;   // first slot of each strip

    DevCtx dev() const {
        DevCtx d{};
        d.nb = nb; d.ny = ny; d.dimq = dimq; d.dimw = dimw; d.NC = NC;
        d.blk = dblk;
        for (int i = 0; i < SHCFD_F_COUNT; i++) d.f[i] = f[i];
        d.scal = scal; d.iflag = iflag; d.p = p;
        d.tgv_zero = tgv_zero;
        return d;
    }
};

int field_ncomp(int fld, int ny);

// frees the adjoint state of a context (shcfd_adjoint.hip)
void adjoint_release(Ctx* c);

}  // namespace shcfd
