// shcfd_adj.h -- device view and host state of the discrete adjoint
// (k_adjoint.hip, the transposed sweeps in k_fast.hip, shcfd_adjoint.hip).
#pragma once
#include "shcfd_internal.h"

namespace shcfd {

// Arrays indexed like the cell fields (component k of cell c at k*NC + c);
// face arrays are stored at the index of the face's left/lower cell, node
// arrays at the index of the node.
struct AdjView {
    enum { WLI = 0, WRI = 1, WLJ = 2, WRJ = 3 };
    double* lam;       // adjoint state (4 per interior cell)
    double* qbar;      // lambda^T dR/dq + dJ/dq (4 per interior cell)
    double* wbar;      // bars of w (dimw per cell, ghosts included)
    double* face[4];   // bars of the MUSCL face states (dimw per face)
    double* nbar_i;    // bars of the i-face unit normals (2 per face)
    double* nbar_j;
    double* dsbar_i;   // bars of the face lengths
    double* dsbar_j;
    double* xbar;      // bars of the node coordinates
    double* ybar;
    int obj_block, obj_i0, obj_i1;  // wall faces j = 0, i in [i0, i1] of a block
    double obj_ax, obj_ay;          // objective weights of the momentum fluxes
    // unsteady (BDF2 dual time): 0 steady, 1 a BDF2 step, 2 the first (BDF1) step
    int unsteady;
    double* src;       // Vol/dt (2 lambda^(n+1) - lambda^(n+2)/2) (4 per interior cell)
    double* lam2;      // lambda of the later step (4 per interior cell)
    double* xacc;      // node bars summed over the steps
    double* yacc;
    // viscous (allocated when the flow is viscous, else null): per face and
    // stencil position s (c, c1, a, b, e, f), the w bars (8 components per
    // position, at (8 s + m) NC + face) and the centre-coordinate bars (at
    // (2 s + {0: x, 1: r}) NC + face) of the face's viscous flux; and the
    // centre-coordinate bars per cell (x at c, r at NC + c)
    double* vw_i;
    double* vw_j;
    double* vg_i;
    double* vg_j;
    double* cbar;
};

struct Adjoint {
    bool ready = false;
    double* mem = nullptr;     // one allocation for all arrays of AdjView
    double* poolbar = nullptr; // bars of the fixed boundary states (npool)
    double* poolacc = nullptr; // their sum over unsteady steps (npool)
    double* part = nullptr;    // reduction partials
    double* objval = nullptr;  // objective value
    int npart = 0;
    AdjView v{};
    int init(Ctx& c);
    void release();
};

}  // namespace shcfd
