// Host build of the library's DVODE restatement (shared_cfd_amd/csrc/
// shcfd_dvode.h) side by side with the reference's DVODE (dvode.f, built
// into oracle/_ref), on small stiff test problems with closed-form
// right-hand sides, so both integrate the same f bit for bit:
//   0  Robertson's kinetics (N = 3)
//   1  van der Pol, mu = 1000 (N = 2)
//   2  y' = -50 (y - cos t) (N = 1)
// Test infrastructure only (tests/test_dvode.py compiles and loads it).
#define SHCFD_HD inline
#define __device__
#define __host__
#include <cmath>
#include <cstring>
#include <dlfcn.h>

#include "../../shared_cfd_amd/csrc/shcfd_dvode.h"

namespace {

void rhs(int prob, double t, const double* y, double* d) {
    if (prob == 0) {
        d[0] = -0.04 * y[0] + 1e4 * y[1] * y[2];
        d[2] = 3e7 * y[1] * y[1];
        d[1] = -d[0] - d[2];
    } else if (prob == 1) {
        d[0] = y[1];
        d[1] = 1000.0 * (1.0 - y[0] * y[0]) * y[1] - y[0];
    } else {
        d[0] = -50.0 * (y[0] - std::cos(t));
    }
}
void jac(int prob, double t, const double* y, double* pd) {  // column-major
    if (prob == 0) {
        pd[0] = -0.04;
        pd[1] = 0.04;
        pd[2] = 0.0;
        pd[3] = 1e4 * y[2];
        pd[4] = -1e4 * y[2] - 6e7 * y[1];
        pd[5] = 6e7 * y[1];
        pd[6] = 1e4 * y[1];
        pd[7] = -1e4 * y[1];
        pd[8] = 0.0;
    } else if (prob == 1) {
        pd[0] = 0.0;
        pd[1] = -2000.0 * y[0] * y[1] - 1.0;
        pd[2] = 1.0;
        pd[3] = 1000.0 * (1.0 - y[0] * y[0]);
    } else {
        pd[0] = -50.0;
    }
}
int neq_of(int prob) { return prob == 0 ? 3 : prob == 1 ? 2 : 1; }
void y0_of(int prob, double* y) {
    if (prob == 0) {
        y[0] = 1.0;
        y[1] = 0.0;
        y[2] = 0.0;
    } else if (prob == 1) {
        y[0] = 2.0;
        y[1] = 0.0;
    } else {
        y[0] = 0.0;
    }
}

int g_prob;
void ref_f(const int*, const double* t, const double* y, double* d, double*, int*) { rhs(g_prob, *t, y, d); }
void ref_j(const int*, const double* t, const double* y, const int*, const int*, double* pd, const int*, double*,
           int*) {
    jac(g_prob, *t, y, pd);
}
typedef void (*F_t)(const int*, const double*, const double*, double*, double*, int*);
typedef void (*J_t)(const int*, const double*, const double*, const int*, const int*, double*, const int*, double*,
                    int*);
typedef void (*DV_t)(F_t, const int*, double*, double*, const double*, const int*, const double*, const double*,
                     const int*, int*, const int*, double*, const int*, int*, const int*, J_t, const int*, double*,
                     int*);

struct Sys {
    int prob;
    void f(double t, const double* y, double* d) const { rhs(prob, t, y, d); }
    void jac(double t, const double* y, double* pd) const { ::jac(prob, t, y, pd); }
};

template <int N>
int mine_steps_n(int prob, double rtol, double atol, double tout, double* ts, double* hs, int* qs, int cap) {
    // DVODE's set-up for ISTATE = 1, then one DVSTEP per entry, as ITASK = 2
    // does
    shcfd::dvode::Dvode<N> d;
    Sys s{prob};
    double y[N];
    y0_of(prob, y);
    if (d.start(s, y, 0.0, tout, rtol, atol)) return -3;
    int k = 0;
    for (bool first = true; k < cap; first = false) {
        if (!first && !d.ewset(rtol, atol)) return -6;
        d.step(s);
        if (d.kflag) return d.kflag == -1 ? -4 : -5;
        if ((d.tn - tout) * d.h >= 0.0) break;
        ts[k] = d.tn;
        hs[k] = d.hu;
        qs[k] = d.nq;
        k++;
    }
    return k;
}

template <int N>
int mine_run_n(int prob, double rtol, double atol, double tout, double* y, int* nst, int* nrestart) {
    shcfd::dvode::Dvode<N> d;
    Sys s{prob};
    double t = 0.0;
    y0_of(prob, y);
    *nst = 0;
    *nrestart = 0;
    for (;;) {
        const int istate = d.run(s, y, t, tout, rtol, atol);
        *nst += d.nst;
        if (istate > 0) return istate;
        if (istate != -1) return istate;
        *nrestart += 1;
        if (*nrestart > 50) return -1;
    }
}

}  // namespace

extern "C" {

// the reference, one step per call (ITASK = 2 takes the same steps as
// ITASK = 1): per step the time reached, the step size and order used
int ref_steps(const char* so, int prob, double rtol, double atol, double tout, double* ts, double* hs, int* qs,
              int cap) {
    void* lib = dlopen(so, RTLD_NOW);
    if (!lib) return -100;
    DV_t dv = (DV_t)dlsym(lib, "dvode_");
    if (!dv) return -101;
    g_prob = prob;
    int neq = neq_of(prob), itol = 1, itask = 2, istate = 1, iopt = 0, mf = 21;
    int lrw = 22 + 9 * neq + 2 * neq * neq, liw = 30 + neq, ipar[1], iw[40];
    double rw[100], rpar[1], y[3], t = 0.0;
    std::memset(iw, 0, sizeof(iw));
    std::memset(rw, 0, sizeof(rw));
    y0_of(prob, y);
    int k = 0;
    while (k < cap) {
        dv(ref_f, &neq, y, &t, &tout, &itol, &rtol, &atol, &itask, &istate, &iopt, rw, &lrw, iw, &liw, ref_j, &mf,
           rpar, ipar);
        if (istate < 0) return istate;
        if (t >= tout) break;
        ts[k] = t;
        hs[k] = rw[10];  // RWORK(11) = HU
        qs[k] = iw[13];  // IWORK(14) = NQU
        k++;
    }
    return k;
}

// the reference with ITASK = 1 to tout, restarting after ISTATE = -1 as
// chemeq2 does (sub_chem.f90:1751-1770): final y, total steps, restarts
int ref_run(const char* so, int prob, double rtol, double atol, double tout, double* y, int* nst, int* nrestart) {
    void* lib = dlopen(so, RTLD_NOW);
    if (!lib) return -100;
    DV_t dv = (DV_t)dlsym(lib, "dvode_");
    if (!dv) return -101;
    g_prob = prob;
    int neq = neq_of(prob), itol = 1, itask = 1, istate = 1, iopt = 0, mf = 21;
    int lrw = 22 + 9 * neq + 2 * neq * neq, liw = 30 + neq, ipar[1], iw[40];
    double rw[100], rpar[1], t = 0.0;
    y0_of(prob, y);
    *nst = 0;
    *nrestart = 0;
    for (;;) {
        std::memset(iw, 0, sizeof(iw));
        dv(ref_f, &neq, y, &t, &tout, &itol, &rtol, &atol, &itask, &istate, &iopt, rw, &lrw, iw, &liw, ref_j, &mf,
           rpar, ipar);
        *nst += iw[10];  // IWORK(11) = NST
        if (istate > 0) return istate;
        if (istate != -1) return istate;
        istate = 1;
        *nrestart += 1;
        if (*nrestart > 50) return -1;
    }
}

int mine_steps(int prob, double rtol, double atol, double tout, double* ts, double* hs, int* qs, int cap) {
    switch (neq_of(prob)) {
    case 3: return mine_steps_n<3>(prob, rtol, atol, tout, ts, hs, qs, cap);
    case 2: return mine_steps_n<2>(prob, rtol, atol, tout, ts, hs, qs, cap);
    default: return mine_steps_n<1>(prob, rtol, atol, tout, ts, hs, qs, cap);
    }
}

int mine_run(int prob, double rtol, double atol, double tout, double* y, int* nst, int* nrestart) {
    switch (neq_of(prob)) {
    case 3: return mine_run_n<3>(prob, rtol, atol, tout, y, nst, nrestart);
    case 2: return mine_run_n<2>(prob, rtol, atol, tout, y, nst, nrestart);
    default: return mine_run_n<1>(prob, rtol, atol, tout, y, nst, nrestart);
    }
}

}  // extern "C"
