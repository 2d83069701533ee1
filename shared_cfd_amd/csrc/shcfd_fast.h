// shcfd_fast.h -- host side of the optimised implicit update (k_fast.hip)
#pragma once
#include "shcfd_internal.h"

namespace shcfd {

struct QStrip;

struct FastPath {
    bool ready = false;
    int nstrips = 0, nwork = 0;
    int ncu = 0;  // compute units of the device (hipDeviceProp_t::multiProcessorCount)
    long long nslots = 0, hstride = 0;
    QStrip* strips = nullptr;
    int2* work = nullptr;
    double *R = nullptr, *V = nullptr, *X = nullptr, *Z = nullptr;
    double *G = nullptr, *H = nullptr;  // (GA, GB) and (HA, HB), split-half layout
    double* Dn = nullptr;               // diagonal terms of the backward sweep with S = I
    double* hand = nullptr;
    double* part = nullptr;
    // transposed sweeps of the discrete adjoint (allocated on first use)
    double *aR = nullptr, *aV = nullptr, *aZ = nullptr, *aG = nullptr, *aH = nullptr, *ahand = nullptr;
    int init(Ctx& c);
    void release();
    // right-hand side (with dq/dt and set_ABpm's scalars formed when asked),
    // operators, both sweeps and the update
    int implicit(Ctx& c, int step_internal, int dqdt_inline, int abpm_inline, void (*prof_begin)(Ctx*, int),
                 void (*prof_end)(Ctx*, int));
    int adjoint_prep(Ctx& c);                                      // transposed operators at the current state
    int adjoint_apply(Ctx& c, const double* r_cell, double* lam);  // lam += M^-T r
};

}  // namespace shcfd
