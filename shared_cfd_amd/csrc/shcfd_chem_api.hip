// shcfd_chem_api.hip -- C ABI of the chemistry: host-evaluated boundary
// states and the batched 0-D harness on the GPU.
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/shcfd.h"
#include "shcfd_chem.h"
#include "shcfd_kernels.h"

using namespace shcfd;

namespace {

int row_in(int op, int ns) {
    switch (op) {
    case SHCFD_CHEM_FEX: case SHCFD_CHEM_FEXVAR: case SHCFD_CHEM_JEX: return ns + 1;
    case SHCFD_CHEM_CALCT: return ns + 3;
    case SHCFD_CHEM_CHEMEQ2: case SHCFD_CHEM_POINT_IMPLICIT: return ns + 1;
    default: return -1;
    }
}
int row_out(int op, int ns) {
    switch (op) {
    case SHCFD_CHEM_FEX: return ns + 1;
    case SHCFD_CHEM_FEXVAR: return 2 * (ns + 1);
    case SHCFD_CHEM_JEX: return (ns + 1) * (ns + 1);
    case SHCFD_CHEM_CALCT: return 4 + 2 * ns;
    case SHCFD_CHEM_CHEMEQ2: return ns;
    case SHCFD_CHEM_POINT_IMPLICIT: return ns + 1;
    default: return -1;
    }
}

template <int NS>
void launch0d(int op, const shcfd_mech* dm, int n, const double* din, int ldi, double* dout, int ldo, int* dst,
              double dt) {
    dim3 g((unsigned)((n + 255) / 256)), b(256);
    switch (op) {
    case SHCFD_CHEM_FEX: hipLaunchKernelGGL((k_chem0d<NS, CHEM_OP_FEX>), g, b, 0, 0, dm, n, din, ldi, dout, ldo, dst, dt); break;
    case SHCFD_CHEM_FEXVAR: hipLaunchKernelGGL((k_chem0d<NS, CHEM_OP_FEXVAR>), g, b, 0, 0, dm, n, din, ldi, dout, ldo, dst, dt); break;
    case SHCFD_CHEM_JEX: hipLaunchKernelGGL((k_chem0d<NS, CHEM_OP_JEX>), g, b, 0, 0, dm, n, din, ldi, dout, ldo, dst, dt); break;
    case SHCFD_CHEM_CALCT: hipLaunchKernelGGL((k_chem0d<NS, CHEM_OP_CALCT>), g, b, 0, 0, dm, n, din, ldi, dout, ldo, dst, dt); break;
    case SHCFD_CHEM_CHEMEQ2: hipLaunchKernelGGL((k_chem0d<NS, CHEM_OP_CHEMEQ2>), g, b, 0, 0, dm, n, din, ldi, dout, ldo, dst, dt); break;
    default: hipLaunchKernelGGL((k_chem0d<NS, CHEM_OP_POINT_IMPLICIT>), g, b, 0, 0, dm, n, din, ldi, dout, ldo, dst, dt); break;
    }
}

template <int NS>
int boundary_state(const shcfd_mech* m, double p, double T, const double* vrho_rel, double* w, double* vhi,
                   double* dhi) {
    // calc_boundary, store/therm_lib/chemkin/flow/reactive/thermal_model.f90:326-356
    double vrho[NS];
    for (int k = 0; k < NS; k++) vrho[k] = vrho_rel[k];
    double rho, E;
    chem::rho_rel2abs<NS>(m, p, T, vrho, rho, E);
    chem::ThermoOut<NS> o;
    double Tn = T;
    if (chem::calc_t<NS>(m, vrho, rho, E, Tn, o) < 0) return SHCFD_E_TNEWTON;
    const int dq = NS + 3;
    w[0] = rho;
    w[1] = 0.0;
    w[2] = 0.0;
    w[3] = p;
    for (int j = 0; j < NS; j++) w[4 + j] = vrho[j] / rho;
    w[dq + 1] = o.kappa;
    w[dq + 2] = E + p / rho;
    w[dq + 3] = chem::R_UNI / o.mwave;
    w[dq + 4] = o.mu;
    for (int j = 0; j < NS; j++) {
        vhi[j] = o.vhi[j];
        if (dhi) dhi[j] = o.dhi[j];
    }
    return SHCFD_OK;
}

}  // namespace

extern "C" {

int shcfd_chem_boundary_state(const shcfd_mech* m, double p, double T, const double* vrho_rel, double* w,
                              double* vhi, double* dhi) {
    if (!m || !vrho_rel || !w || !vhi || !(p > 0.0) || !(T > 0.0)) return SHCFD_E_ARG;
    switch (m->ns) {
    case 8: return boundary_state<8>(m, p, T, vrho_rel, w, vhi, dhi);
    case 9: return boundary_state<9>(m, p, T, vrho_rel, w, vhi, dhi);
    default: return SHCFD_E_ARG;
    }
}

int shcfd_load_mechanism(shcfd_ctx* ctx, const char* chem_inp, const char* mw_inp, const char* trans_inp, double rtol,
                         double atol) {
    if (!ctx || !chem_inp || !mw_inp) return SHCFD_E_ARG;
    shcfd_mech* m = new (std::nothrow) shcfd_mech;
    if (!m) return SHCFD_E_NOMEM;
    char err[256];
    int rc = shcfd_mech_parse(chem_inp, mw_inp, (trans_inp && trans_inp[0]) ? trans_inp : nullptr, m, err, 256);
    if (rc == SHCFD_OK) {
        m->rtol = rtol;
        m->atol = atol;
        rc = shcfd_set_mechanism(ctx, m);
    } else {
        std::fprintf(stderr, "shcfd_load_mechanism: %s\n", err);
    }
    delete m;
    return rc;
}

// init_pack_cold_flow for hosts that do not hold the tables (the Fortran
// driver): shcfd_nasa9_parse of the four files, then shcfd_set_nasa9
int shcfd_load_nasa9(shcfd_ctx* ctx, const char* chem_inp, const char* thermo_inp, const char* trans_inp,
                     const char* control_chem_inp) {
    if (!ctx || !chem_inp || !thermo_inp || !control_chem_inp) return SHCFD_E_ARG;
    shcfd_nasa9* t = new (std::nothrow) shcfd_nasa9;
    if (!t) return SHCFD_E_NOMEM;
    char err[256];
    int rc = shcfd_nasa9_parse(chem_inp, thermo_inp, (trans_inp && trans_inp[0]) ? trans_inp : nullptr,
                               control_chem_inp, t, nullptr, err, 256);
    if (rc == SHCFD_OK)
        rc = shcfd_set_nasa9(ctx, t);
    else
        std::fprintf(stderr, "shcfd_load_nasa9: %s\n", err);
    delete t;
    return rc;
}

int shcfd_chem_eval(const shcfd_mech* m, int op, int n, const double* in, double* out, int* status, double dt) {
    if (!m || n < 0 || !in || !out || (m->ns != 8 && m->ns != 9) || m->nr < 1 || m->nr > SHCFD_CHEM_MAX_NR)
        return SHCFD_E_ARG;
    int ldi = row_in(op, m->ns), ldo = row_out(op, m->ns);
    if (ldi < 0) return SHCFD_E_ARG;
    if (n == 0) return SHCFD_OK;
    shcfd_mech* dm = nullptr;
    double *din = nullptr, *dout = nullptr;
    int* dst = nullptr;
    int rc = SHCFD_OK;
    if (hipMalloc((void**)&dm, sizeof(shcfd_mech)) != hipSuccess ||
        hipMalloc((void**)&din, sizeof(double) * (size_t)n * ldi) != hipSuccess ||
        hipMalloc((void**)&dout, sizeof(double) * (size_t)n * ldo) != hipSuccess ||
        hipMalloc((void**)&dst, sizeof(int) * (size_t)n) != hipSuccess) {
        rc = SHCFD_E_NOMEM;
    } else if (hipMemcpy(dm, m, sizeof(shcfd_mech), hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(din, in, sizeof(double) * (size_t)n * ldi, hipMemcpyHostToDevice) != hipSuccess) {
        rc = SHCFD_E_HIP;
    } else {
        if (m->ns == 8)
            launch0d<8>(op, dm, n, din, ldi, dout, ldo, dst, dt);
        else
            launch0d<9>(op, dm, n, din, ldi, dout, ldo, dst, dt);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(out, dout, sizeof(double) * (size_t)n * ldo, hipMemcpyDeviceToHost) != hipSuccess ||
            (status && hipMemcpy(status, dst, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess))
            rc = SHCFD_E_HIP;
    }
    hipFree(dm);
    hipFree(din);
    hipFree(dout);
    hipFree(dst);
    return rc;
}

}  // extern "C"
