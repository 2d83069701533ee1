// shcfd_nasa9.h -- the frozen-composition ("cold flow") thermodynamics of
// the NASA-9 library: cold_flow (store/therm_lib/NASA/core/sub_chem.f90:
// 489-610) with the polynomial helpers of func_chem.f90.  Host and device
// code: the flow kernel evaluates it per cell, the host evaluates it for
// the static states of the two streams (initialize_cold_flow).
//
// The mixture is Y0 kg of the fuel stream and Y1 kg of the oxidizer stream
// per kg, n = Y0 nf + Y1 no mol/kg.  T follows from the internal energy by
// the reference's under-relaxed Newton iteration (omega = 0.5, at most 100
// steps, stopping when |E/Ru - T sum n_i(h_i/RT - 1)| <= 1e-11 cv/Ru or the
// step is below 1e-9 K); kappa = 1 + Ru sum n / cv; the stream derivatives
// DHi = Ru T sum_i n_i (kappa - (kappa-1) h_i/(Ru T)) and enthalpies vhi; the
// viscosity by Wilke's rule over the species with a transport fit.  The
// arithmetic follows the Fortran expression by expression; exp and log are
// the table-free methods of shcfd_chem.h (bit-identical on host and device)
// and x**0.25 is sqrt(sqrt(x)), as in the Chemkin kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/shcfd_types.h"
#include "shcfd_chem.h"

namespace shcfd {
namespace nasa9 {

constexpr double RU = 8.3144621;        // J/(mol K), mod_chem.raw.f90:6
constexpr double R_UNI = 8.3144621e3;   // J/(kmol K), core/prmtr.f90:4
constexpr double OMEGA = 0.5;           // mod_chem.raw.f90:8
constexpr double EPS = 1e-9;            // :9
constexpr int MAX_ITER = 100;

// check_section_number, func_chem.f90:41-62
__host__ __device__ inline int section(const shcfd_nasa9* m, int s, double T) {
    if (m->trange[s][0][0] > T) return 0;
    for (int l = 0; l < m->nsec[s]; l++)
        if (m->trange[s][l][1] > T) return l;
    return m->nsec[s] - 1;
}

// check_section_number_trans, func_chem.f90:173-192
__host__ __device__ inline int section_trans(const shcfd_nasa9* m, int t, double T) {
    for (int l = 0; l < m->nsec_trans[t]; l++)
        if (m->trange_trans[t][l][1] > T) return l;
    return m->nsec_trans[t] - 1;
}

// calc_vThrt (func_chem.f90:64-90): h/RT = sum_k co_k v_k
__host__ __device__ inline void v_thrt(double T, double logT, double* v) {
    double tt = 1.0 / (T * T);
    v[0] = -tt;
    tt = tt * T;
    v[1] = tt * logT;
    v[7] = tt;
    tt = tt * T;
    v[2] = tt;
    tt = tt * T;
    v[3] = tt / 2.0;
    tt = tt * T;
    v[4] = tt / 3.0;
    tt = tt * T;
    v[5] = tt / 4.0;
    tt = tt * T;
    v[6] = tt / 5.0;
    v[8] = 0.0;
}

// calc_vTcpr (func_chem.f90:91-116): cp/R = sum_k co_k v_k
__host__ __device__ inline void v_tcpr(double T, double* v) {
    double tt = 1.0 / (T * T);
    v[0] = tt;
    tt = tt * T;
    v[1] = tt;
    tt = tt * T;
    v[2] = tt;
    tt = tt * T;
    v[3] = tt;
    tt = tt * T;
    v[4] = tt;
    tt = tt * T;
    v[5] = tt;
    tt = tt * T;
    v[6] = tt;
    v[7] = 0.0;
    v[8] = 0.0;
}

struct ColdOut {
    double mwave, kappa, mu;
    double dhi[2], vhi[2];
    int iters;
};

// E/Ru T and cv/Ru per kg (sums of n_i (h_i/RT - 1) and n_i (cp_i/R - 1))
__host__ __device__ inline void mixture_sums(const shcfd_nasa9* m, const double* n, double T, double& ert, double& cvr) {
    double vh[9], vc[9];
    v_thrt(T, chem::dlog(T), vh);
    v_tcpr(T, vc);
    ert = 0.0;
    cvr = 0.0;
    for (int i = 0; i < m->ns; i++) {
        const int sec = section(m, i, T);
        const double* co = m->co[i][sec];
        double a = -1.0, b = -1.0;
        for (int j = 0; j < 9; j++) {
            a = a + (co[j] * vh[j]);
            b = b + (co[j] * vc[j]);
        }
        ert = ert + (a * n[i]);
        cvr = cvr + (b * n[i]);
    }
}

// cold_flow (sub_chem.f90:489-610); T in: the start value, out: the solution
__host__ __device__ inline void cold_flow(const shcfd_nasa9* m, double y0, double y1, double E, double& T, ColdOut& o) {
    const int ns = m->ns;
    double n[SHCFD_NASA9_MAX_NS];
    double sn = 0.0;
    for (int i = 0; i < ns; i++) {
        n[i] = (y0 * m->nf[i]) + (y1 * m->no[i]);
        sn = sn + n[i];
    }
    o.mwave = 1e3 / sn;
    double ert, cvr;
    mixture_sums(m, n, T, ert, cvr);
    int k = 0;
    double t_old = T * 0.5;
    while (fabs((E / RU) - (ert * T)) > ((cvr * EPS) * 1e-2) && fabs(T - t_old) > EPS && k < MAX_ITER) {
        t_old = T;
        T = T + ((OMEGA * ((E / RU) - (ert * T))) / cvr);
        mixture_sums(m, n, T, ert, cvr);
        k = k + 1;
    }
    o.iters = k;
    const double kappa = 1.0 + (sn / cvr);
    o.kappa = kappa;
    const double logT = chem::dlog(T);
    double vh[9];
    v_thrt(T, logT, vh);
    double dhs[SHCFD_NASA9_MAX_NS], vhs[SHCFD_NASA9_MAX_NS];
    for (int i = 0; i < ns; i++) {
        const int sec = section(m, i, T);
        const double* co = m->co[i][sec];
        double a = 0.0;
        for (int j = 0; j < 9; j++) a = a + (co[j] * vh[j]);
        dhs[i] = 1.0 * (kappa - ((kappa - 1.0) * a));
        vhs[i] = 1.0 * a;
    }
    const double rt = RU * T;
    double d0 = 0.0, d1 = 0.0, h0 = 0.0, h1 = 0.0;
    for (int i = 0; i < ns; i++) {
        d0 = d0 + (m->nf[i] * dhs[i]);
        d1 = d1 + (m->no[i] * dhs[i]);
        h0 = h0 + (m->nf[i] * vhs[i]);
        h1 = h1 + (m->no[i] * vhs[i]);
    }
    o.dhi[0] = d0 * rt;
    o.dhi[1] = d1 * rt;
    o.vhi[0] = h0 * rt;
    o.vhi[1] = h1 * rt;
    // calc_vTmu (func_chem.f90:143-157) and Wilke's mixing rule
    const double tt = 1.0 / (T * T);
    const double vmu[4] = {logT, tt * T, tt, 1.0};
    double mun[SHCFD_NASA9_MAX_NS];
    for (int t = 0; t < m->nt; t++) {
        const double* a = m->trans[t][section_trans(m, t, T)];
        double s = 0.0;
        for (int j = 0; j < 4; j++) s = s + (a[j] * vmu[j]);
        mun[t] = chem::dexp(s);
    }
    double mu = 0.0;
    for (int i = 0; i < m->nt; i++) {
        double denom = 0.0;
        const double mwi = m->mw[m->tr2th[i]];
        for (int j = 0; j < m->nt; j++) {
            const double mwj = m->mw[m->tr2th[j]];
            const double x = 1.0 + (sqrt(mun[i] / mun[j]) * sqrt(sqrt(mwj / mwi)));
            const double phi = (0.25 * (x * x)) * sqrt((2.0 * mwj) / (mwi + mwj));
            denom = denom + (n[m->tr2th[j]] * phi);
        }
        mu = mu + ((n[m->tr2th[i]] * mun[i]) / denom);
    }
    o.mu = mu * 1e-7;
}

}  // namespace nasa9
}  // namespace shcfd
