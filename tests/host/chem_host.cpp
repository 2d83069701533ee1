// Host build of the library's per-cell chemistry (shcfd_chem.h, shcfd_dvode.h):
// Fex beside the dn/dt of Jex, and the integrators CHEMEQ2 with its DVODE
// hand-over and the DVODE reaction,
// with the 0-D harness's rows (in: vrho(ns), T; out: vrho(ns); status:
// attempts or DVODE steps, or -code).  The GPU tests check that the device
// build of the same code gives the same bits.  Test infrastructure only.
#define SHCFD_HD inline
#define __device__
#define __host__
#include "../../shared_cfd_amd/csrc/shcfd_dvode.h"

namespace {

template <int NS>
void rows(int op, const shcfd_mech* m, int n, const double* in, double* out, int* status, double dt) {
    for (int q = 0; q < n; q++) {
        const double* x = in + (long long)q * (NS + 1);
        double vrho[NS];
        for (int k = 0; k < NS; k++) vrho[k] = x[k];
        int it = 0, e;
        if (op == SHCFD_CHEM_DVODE)
            e = shcfd::chem::reaction_dvode<NS>(m, x[NS], vrho, dt, it);
        else
            e = shcfd::chem::reaction_chemeq2_full<NS>(m, x[NS], vrho, dt, it);
        status[q] = e ? -e : it;
        for (int k = 0; k < NS; k++) out[(long long)q * NS + k] = vrho[k];
    }
}

// f of Fex and the dn/dt that Jex forms beside the Jacobian, per state (in:
// n(ns) mol/cm^3, T)
template <int NS>
void fj_rows(const shcfd_mech* m, int n, const double* in, double* f, double* dndt) {
    for (int q = 0; q < n; q++) {
        double y[NS + 1], a[NS + 1], b[NS + 1], J[NS + 1][NS + 1];
        for (int k = 0; k <= NS; k++) y[k] = in[(long long)q * (NS + 1) + k];
        shcfd::chem::fex<NS>(m, y, a);
        shcfd::chem::jex<NS>(m, y, J, b);
        for (int k = 0; k <= NS; k++) {
            f[(long long)q * (NS + 1) + k] = a[k];
            dndt[(long long)q * (NS + 1) + k] = b[k];
        }
    }
}

}  // namespace

extern "C" int host_fex_jex(const shcfd_mech* m, int n, const double* in, double* f, double* dndt) {
    switch (m->ns) {
    case 8: fj_rows<8>(m, n, in, f, dndt); return 0;
    case 9: fj_rows<9>(m, n, in, f, dndt); return 0;
    case 12: fj_rows<12>(m, n, in, f, dndt); return 0;
    default: return SHCFD_E_ARG;
    }
}

extern "C" int host_chem_eval(const shcfd_mech* m, int op, int n, const double* in, double* out, int* status,
                              double dt) {
    if (op != SHCFD_CHEM_DVODE && op != SHCFD_CHEM_CHEMEQ2) return SHCFD_E_ARG;
    switch (m->ns) {
    case 8: rows<8>(op, m, n, in, out, status, dt); return 0;
    case 12: rows<12>(op, m, n, in, out, status, dt); return 0;
    case 16: rows<16>(op, m, n, in, out, status, dt); return 0;
    default: return SHCFD_E_ARG;
    }
}
