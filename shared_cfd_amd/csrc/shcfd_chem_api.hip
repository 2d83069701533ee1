// shcfd_chem_api.hip -- C ABI of the chemistry: host-evaluated boundary
// states and the batched 0-D harness on the GPU.
#include <cstdio>
#include <cstring>
#include <mutex>
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
    case SHCFD_CHEM_CHEMEQ2: case SHCFD_CHEM_POINT_IMPLICIT: case SHCFD_CHEM_DVODE: return ns + 1;
    default: return -1;
    }
}
int row_out(int op, int ns) {
    switch (op) {
    case SHCFD_CHEM_FEX: return ns + 1;
    case SHCFD_CHEM_FEXVAR: return 2 * (ns + 1);
    case SHCFD_CHEM_JEX: return (ns + 1) * (ns + 1);
    case SHCFD_CHEM_CALCT: return 4 + 2 * ns;
    case SHCFD_CHEM_CHEMEQ2: case SHCFD_CHEM_DVODE: return ns;
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
    case SHCFD_CHEM_DVODE: hipLaunchKernelGGL((k_chem0d<NS, CHEM_OP_DVODE>), g, b, 0, 0, dm, n, din, ldi, dout, ldo, dst, dt); break;
    default: hipLaunchKernelGGL((k_chem0d<NS, CHEM_OP_POINT_IMPLICIT>), g, b, 0, 0, dm, n, din, ldi, dout, ldo, dst, dt); break;
    }
}

// the 0-D launch for the mechanism's species count (1..SHCFD_MAX_NY)
void launch0d_ns(int ns, int op, const shcfd_mech* dm, int n, const double* din, int ldi, double* dout, int ldo,
                 int* dst, double dt) {
    switch (ns) {
#define SHCFD_NS_CASE(N) case N: launch0d<N>(op, dm, n, din, ldi, dout, ldo, dst, dt); break;
    SHCFD_NS_CASE(1) SHCFD_NS_CASE(2) SHCFD_NS_CASE(3) SHCFD_NS_CASE(4) SHCFD_NS_CASE(5) SHCFD_NS_CASE(6)
    SHCFD_NS_CASE(7) SHCFD_NS_CASE(8) SHCFD_NS_CASE(9) SHCFD_NS_CASE(10) SHCFD_NS_CASE(11) SHCFD_NS_CASE(12)
    SHCFD_NS_CASE(13) SHCFD_NS_CASE(14) SHCFD_NS_CASE(15) SHCFD_NS_CASE(16)
#undef SHCFD_NS_CASE
    default: break;
    }
}

bool mech_ok(const shcfd_mech* m) {
    return m->ns >= 1 && m->ns <= SHCFD_MAX_NY && m->nr >= 1 && m->nr <= SHCFD_CHEM_MAX_NR;
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

// ---- device scratch of shcfd_chem_eval_noalloc -------------------------------
// One file-scope array per GPU, held until the process ends.  A call stages
// the mechanism table, the input rows, the output rows and the status words
// in it, each part starting on a 256-byte boundary.  Host code never uses the
// array's name as an address: the per-device address comes from
// hipGetSymbolAddress, taken once per device on the first call made with that
// device current.
static __device__ __align__(256) unsigned char scratch[SHCFD_CHEM_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

int scratch_address(unsigned char** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return SHCFD_E_HIP;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return SHCFD_E_HIP;
        }
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return SHCFD_OK;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }
}  // namespace

extern "C" {

int shcfd_chem_eval_noalloc(const shcfd_mech* m, int op, int n, const double* in, double* out, int* status,
                            double dt) {
    if (!m || n < 0 || !in || !out || !mech_ok(m)) return SHCFD_E_ARG;
    const int ldi = row_in(op, m->ns), ldo = row_out(op, m->ns);
    if (ldi < 0) return SHCFD_E_ARG;
    if (n == 0) return SHCFD_OK;
    const size_t o_mech = 0;
    const size_t o_in = o_mech + align256(sizeof(shcfd_mech));
    const size_t o_out = o_in + align256(sizeof(double) * (size_t)n * ldi);
    const size_t o_st = o_out + align256(sizeof(double) * (size_t)n * ldo);
    const size_t need = o_st + align256(sizeof(int) * (size_t)n);
    if (need > (size_t)SHCFD_CHEM_SCRATCH_BYTES) return SHCFD_E_TOO_LARGE;
    unsigned char* base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    shcfd_mech* dm = reinterpret_cast<shcfd_mech*>(base + o_mech);
    double* din = reinterpret_cast<double*>(base + o_in);
    double* dout = reinterpret_cast<double*>(base + o_out);
    int* dst = reinterpret_cast<int*>(base + o_st);
    // every output row and status word is written by the kernel, so nothing
    // left by an earlier call needs clearing
    if (hipMemcpy(dm, m, sizeof(shcfd_mech), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(din, in, sizeof(double) * (size_t)n * ldi, hipMemcpyHostToDevice) != hipSuccess)
        return SHCFD_E_HIP;
    launch0d_ns(m->ns, op, dm, n, din, ldi, dout, ldo, dst, dt);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(0) != hipSuccess ||
        hipMemcpy(out, dout, sizeof(double) * (size_t)n * ldo, hipMemcpyDeviceToHost) != hipSuccess ||
        (status && hipMemcpy(status, dst, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess))
        return SHCFD_E_HIP;
    return SHCFD_OK;
}

int shcfd_chem_boundary_state(const shcfd_mech* m, double p, double T, const double* vrho_rel, double* w,
                              double* vhi, double* dhi) {
    if (!m || !vrho_rel || !w || !vhi || !(p > 0.0) || !(T > 0.0)) return SHCFD_E_ARG;
    switch (m->ns) {
#define SHCFD_NS_CASE(N) case N: return boundary_state<N>(m, p, T, vrho_rel, w, vhi, dhi);
    SHCFD_NS_CASE(1) SHCFD_NS_CASE(2) SHCFD_NS_CASE(3) SHCFD_NS_CASE(4) SHCFD_NS_CASE(5) SHCFD_NS_CASE(6)
    SHCFD_NS_CASE(7) SHCFD_NS_CASE(8) SHCFD_NS_CASE(9) SHCFD_NS_CASE(10) SHCFD_NS_CASE(11) SHCFD_NS_CASE(12)
    SHCFD_NS_CASE(13) SHCFD_NS_CASE(14) SHCFD_NS_CASE(15) SHCFD_NS_CASE(16)
#undef SHCFD_NS_CASE
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
    if (!m || n < 0 || !in || !out || !mech_ok(m)) return SHCFD_E_ARG;
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
        launch0d_ns(m->ns, op, dm, n, din, ldi, dout, ldo, dst, dt);
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
