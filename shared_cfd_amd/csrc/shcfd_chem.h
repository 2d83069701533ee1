// shcfd_chem.h -- Chemkin thermochemistry for one cell, usable on the device
// (one cell per lane) and on the host (boundary states at set-up).
//
// Follows store/therm_lib/chemkin/core/sub_chem.f90 and
// store/therm_lib/chemkin/uvhp/uv/FJ.f90; each function names the lines it
// covers.  NS (the species count) is a template parameter so per-species
// vectors live in registers.  The mechanism's species indices are run-time
// but uniform across a wavefront; they are applied with unrolled
// compare-and-select over the NS slots (sel_get / sel_add), which keeps
// every array statically indexed (no scratch memory) and each update in the
// reference's order.
#pragma once
#include "../../include/shcfd_types.h"

#ifndef SHCFD_HD
#define SHCFD_HD __host__ __device__ inline
#endif

namespace shcfd {
namespace chem {

// module const_chem (mod_chem.raw.f90:8-14)
constexpr double RU = 8.3144621e7;   // erg/(mol K)
constexpr double PST = 1.01325e6;    // dyn/cm^2
constexpr double N_EPS = 1e-40;
constexpr double DD = 0.14;
constexpr double LOGE210 = 0.43429448190325;
constexpr double R_UNI = 8.3144621e3;  // J/(kmol K), chem_var

// ---- exp / log --------------------------------------------------------------
// exp: Cody-Waite reduction x = k ln2 + r with ln2 split into a 32-bit head
// and a tail, |r| <= ln2/2, Taylor series of e^r to degree 13 by fused
// Horner steps, scaled by 2^k in two factors.  log: x = 2^k m with
// m in [sqrt(1/2), sqrt(2)), f = (m-1)/(m+1), ln m = 2 atanh f summed to
// f^21.  Both are within 2 ulp of a correctly rounded result, and since
// every step is a correctly rounded IEEE operation, host and device
// builds of this code give the same bits.
constexpr double LN2_HI = 6.93147180369123816490e-01;
constexpr double LN2_LO = 1.90821492927058770002e-10;
constexpr double INV_LN2 = 1.44269504088896338700e+00;
constexpr double INV_LN10 = 4.34294481903251827651e-01;

SHCFD_HD double pow2i(int k) {
    unsigned long long u = (unsigned long long)(k + 1023) << 52;
    double d;
    __builtin_memcpy(&d, &u, 8);
    return d;
}

SHCFD_HD double dexp(double x) {
    if (x != x) return x + x;
    if (x > 709.782712893384) return __builtin_huge_val();
    if (x < -745.1332191019412) return 0.0;
    double kd = __builtin_rint(x * INV_LN2);
    int k = (int)kd;
    double r = __builtin_fma(-kd, LN2_HI, x);
    r = __builtin_fma(-kd, LN2_LO, r);
    double p = 1.0 / 6227020800.0;
    p = __builtin_fma(p, r, 1.0 / 479001600.0);
    p = __builtin_fma(p, r, 1.0 / 39916800.0);
    p = __builtin_fma(p, r, 1.0 / 3628800.0);
    p = __builtin_fma(p, r, 1.0 / 362880.0);
    p = __builtin_fma(p, r, 1.0 / 40320.0);
    p = __builtin_fma(p, r, 1.0 / 5040.0);
    p = __builtin_fma(p, r, 1.0 / 720.0);
    p = __builtin_fma(p, r, 1.0 / 120.0);
    p = __builtin_fma(p, r, 1.0 / 24.0);
    p = __builtin_fma(p, r, 1.0 / 6.0);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    int k1 = k / 2, k2 = k - k1;
    return (p * pow2i(k1)) * pow2i(k2);
}

SHCFD_HD double dlog(double x) {
    if (x != x || x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_huge_val();
    if (x == __builtin_huge_val()) return x;
    int kadj = 0;
    if (x < 2.2250738585072014e-308) {
        x = x * 18014398509481984.0;
        kadj = -54;
    }
    unsigned long long u;
    __builtin_memcpy(&u, &x, 8);
    int k = (int)((u >> 52) & 0x7ff) - 1023 + kadj;
    u = (u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m;
    __builtin_memcpy(&m, &u, 8);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        k = k + 1;
    }
    double f = (m - 1.0) / (m + 1.0);
    double s = f * f;
    double p = 1.0 / 21.0;
    p = __builtin_fma(p, s, 1.0 / 19.0);
    p = __builtin_fma(p, s, 1.0 / 17.0);
    p = __builtin_fma(p, s, 1.0 / 15.0);
    p = __builtin_fma(p, s, 1.0 / 13.0);
    p = __builtin_fma(p, s, 1.0 / 11.0);
    p = __builtin_fma(p, s, 1.0 / 9.0);
    p = __builtin_fma(p, s, 1.0 / 7.0);
    p = __builtin_fma(p, s, 1.0 / 5.0);
    p = __builtin_fma(p, s, 1.0 / 3.0);
    double t = s * p;
    double f2 = f + f;
    double lm = __builtin_fma(f2, t, f2);
    double kd = (double)k;
    return kd * LN2_HI + __builtin_fma(kd, LN2_LO, lm);
}

SHCFD_HD double dlog10(double x) { return dlog(x) * INV_LN10; }
SHCFD_HD double dsqrt(double x) { return __builtin_sqrt(x); }
SHCFD_HD double fmax2(double a, double b) { return (a > b || b != b) ? a : b; }
SHCFD_HD double fmin2(double a, double b) { return (a < b || b != b) ? a : b; }

// ---- uniform run-time index into a register vector --------------------------
template <int N>
SHCFD_HD double sel_get(const double (&a)[N], int idx) {
    double x = a[0];
#pragma unroll
    for (int k = 1; k < N; k++)
        if (k == idx) x = a[k];
    return x;
}
template <int N>
SHCFD_HD void sel_add(double (&a)[N], int idx, double v) {
#pragma unroll
    for (int k = 0; k < N; k++)
        if (k == idx) a[k] = a[k] + v;
}
template <int N>
SHCFD_HD void sel_sub(double (&a)[N], int idx, double v) {
#pragma unroll
    for (int k = 0; k < N; k++)
        if (k == idx) a[k] = a[k] - v;
}

// ---- NASA-7 basis vectors (sub_chem.f90:20-150) ------------------------------
struct Basis {
    double cmurt[7], chrt[7], ccpr[7], cdmurt[7], cdcpr[7];
};
template <bool DERIV>
SHCFD_HD void set_coeff(double T, double logT, Basis& B) {
    double Ti1 = 1.0 / T, Ti2 = Ti1 * Ti1, T2 = T * T, T3 = T2 * T, T4 = T3 * T;
    B.cmurt[0] = 1.0 - logT;
    B.cmurt[1] = -T / 2.0;
    B.cmurt[2] = -T2 / 6.0;
    B.cmurt[3] = -T3 / 12.0;
    B.cmurt[4] = -T4 / 20.0;
    B.cmurt[5] = Ti1;
    B.cmurt[6] = -1.0;
    B.chrt[0] = 1.0;
    B.chrt[1] = T / 2.0;
    B.chrt[2] = T2 / 3.0;
    B.chrt[3] = T3 / 4.0;
    B.chrt[4] = T4 / 5.0;
    B.chrt[5] = Ti1;
    B.chrt[6] = 0.0;
    B.ccpr[0] = 1.0;
    B.ccpr[1] = T;
    B.ccpr[2] = T2;
    B.ccpr[3] = T3;
    B.ccpr[4] = T4;
    B.ccpr[5] = 0.0;
    B.ccpr[6] = 0.0;
    if (DERIV) {
        B.cdmurt[0] = -Ti1;
        B.cdmurt[1] = -1.0 / 2.0;
        B.cdmurt[2] = -T / 3.0;
        B.cdmurt[3] = -T2 / 4.0;
        B.cdmurt[4] = -T3 / 5.0;
        B.cdmurt[5] = -Ti2;
        B.cdmurt[6] = 0.0;
        B.cdcpr[0] = 0.0;
        B.cdcpr[1] = 1.0;
        B.cdcpr[2] = 2.0 * T;
        B.cdcpr[3] = 3.0 * T2;
        B.cdcpr[4] = 4.0 * T3;
        B.cdcpr[5] = 0.0;
        B.cdcpr[6] = 0.0;
    }
}

// polynomial range of species k at T (check_section_number, :3-19)
SHCFD_HD const double* coeff_at(const shcfd_mech* m, int k, double T) {
    return m->coeff[k][(T <= m->tthre[k][2]) ? 0 : 1];
}

// species sums of FJ.f90:185-197 / :516-540
template <int NS, bool DERIV>
struct SpeciesThermo {
    double vmu0rt[NS], vert[NS], vcvr[NS], vdmu0rt[DERIV ? NS : 1], vdcvrdT[DERIV ? NS : 1];
};
template <int NS, bool DERIV>
SHCFD_HD void species_thermo(const shcfd_mech* m, double T, double logT, SpeciesThermo<NS, DERIV>& S) {
    Basis B;
    set_coeff<DERIV>(T, logT, B);
#pragma unroll
    for (int j = 0; j < NS; j++) {
        const double* c = coeff_at(m, j, T);
        double t1 = 0.0, t2 = DERIV ? 0.0 : -1.0, t3 = -1.0, t4 = -1.0, t5 = 0.0;
        if (DERIV) {
#pragma unroll
            for (int i = 0; i < 7; i++) {
                double tmp = c[i];
                t1 = t1 + tmp * B.cmurt[i];
                t2 = t2 + tmp * B.cdmurt[i];
                t3 = t3 + tmp * B.chrt[i];
                t4 = t4 + tmp * B.ccpr[i];
                t5 = t5 + tmp * B.cdcpr[i];
            }
            S.vmu0rt[j] = t1;
            S.vdmu0rt[j] = t2;
            S.vert[j] = t3;
            S.vcvr[j] = t4;
            S.vdcvrdT[j] = t5;
        } else {
#pragma unroll
            for (int i = 0; i < 7; i++) {
                double tmp = c[i];
                t1 = t1 + tmp * B.cmurt[i];
                t2 = t2 + tmp * B.chrt[i];
                t3 = t3 + tmp * B.ccpr[i];
            }
            S.vmu0rt[j] = t1;
            S.vert[j] = t2;
            S.vcvr[j] = t3;
        }
    }
}

// log rate constants of reaction i (FJ.f90:212-256 and :549-633)
template <int NS, bool DERIV>
SHCFD_HD void rate_constants(const shcfd_mech* m, int i, double T, double logT, double Tinv, double logPRT,
                             double vlogM, const SpeciesThermo<NS, DERIV>& S, double* vlogk, double* dlnkdT,
                             double& dlnkdlnM) {
    const double* abe = m->abe[i];
    const double* cabe = m->cabe[i];
    int kind = m->kind[i];
    vlogk[0] = abe[0] + abe[1] * logT - abe[2] * Tinv;
    vlogk[1] = 0.0;
    if (DERIV) {
        dlnkdT[0] = (abe[1] + abe[2] * Tinv) * Tinv;
        dlnkdT[1] = 0.0;
    }
    if (kind >= SHCFD_RXN_M) {
        vlogk[0] = vlogk[0] + vlogM;
        if (DERIV) dlnkdlnM = 1.0;
    }
    if (m->reverse_off[i] != 0) return;
    if (kind == SHCFD_RXN_REV || kind == SHCFD_RXN_M_REV) {
        vlogk[1] = cabe[0] + cabe[1] * logT - cabe[2] * Tinv;
        if (DERIV) dlnkdT[1] = (cabe[1] + cabe[2] * Tinv) * Tinv;
        if (kind == SHCFD_RXN_M_REV) vlogk[1] = vlogk[1] + vlogM;
        return;
    }
    if (kind == SHCFD_RXN_LOW || kind == SHCFD_RXN_TROE) {
        double logk0 = cabe[0] + cabe[1] * logT - cabe[2] * Tinv;
        double logkinf = vlogk[0];
        double log_Pr = logk0 - logkinf + vlogM;
        double Pr = dexp(log_Pr);
        vlogk[0] = logkinf + dlog(Pr / (1.0 + Pr));
        double dlnk0dT = 0.0, dlnkinfdT = 0.0;
        if (DERIV) {
            dlnk0dT = (cabe[1] + cabe[2] * Tinv) * Tinv;
            dlnkinfdT = dlnkdT[0];
            dlnkdT[0] = (dlnk0dT + Pr * dlnkinfdT) / (1.0 + Pr);
            dlnkdlnM = 1.0 / (1.0 + Pr);
        }
        if (kind == SHCFD_RXN_TROE) {
            const double* tr = m->troe[i];
            double aa = tr[0];
            if (!DERIV) {
                // FJ.f90:235-243
                double logFcent = dlog10((1.0 - aa) * dexp(-T * tr[1]) + aa * dexp(-T * tr[2]) + dexp(-Tinv * tr[3]));
                double cc = -0.4 - 0.67 * logFcent;
                double nn = 0.75 - 1.27 * logFcent;
                double logPrc = log_Pr * LOGE210 + cc;
                double t = logPrc / (nn - DD * logPrc);
                double logF = (1.0 / (1.0 + t * t)) * logFcent;
                vlogk[0] = vlogk[0] + logF / LOGE210;
            } else {
                // FJ.f90:593-612
                double dlnPrdT = dlnk0dT - dlnkinfdT;
                double Fcent = (1.0 - aa) * dexp(-T * tr[1]) + aa * dexp(-T * tr[2]) + dexp(-tr[3] * Tinv);
                double dFcentdT = -(1.0 - aa) * tr[1] * dexp(-T * tr[1]) - aa * tr[2] * dexp(-T * tr[2]) +
                                  tr[3] * (Tinv * Tinv) * dexp(-tr[3] * Tinv);
                double dlnFcentdT = dFcentdT / Fcent;
                double logFcent = dlog10(Fcent);
                double cc = -0.4 - 0.67 * logFcent;
                double nn = 0.75 - 1.27 * logFcent;
                double logPrc = log_Pr * LOGE210 + cc;
                double temp1 = logPrc / (nn - DD * logPrc);
                double temp2 = 1.0 / (1.0 + temp1 * temp1);
                double logF = temp2 * logFcent;
                double den = nn - DD * logPrc;
                vlogk[0] = vlogk[0] + logF / LOGE210;
                dlnkdT[0] = dlnkdT[0] + temp2 * (-2.0 * logF * temp1 / (den * den) *
                                                     (nn * dlnPrdT + (1.27 * logPrc - 0.67 * nn) * dlnFcentdT) +
                                                 dlnFcentdT);
                dlnkdlnM = dlnkdlnM - 2.0 * temp2 * logF * temp1 / (den * den) * nn;
            }
        }
    }
    // reverse rate through K_c (FJ.f90:246-256, :615-630)
    double numu0rt = 0.0, nudmu0rt = 0.0;
    for (int k = 0; k < m->num_nu[i][0]; k++) {
        int it = m->ind_nu[i][0][k];
        numu0rt = numu0rt - sel_get(S.vmu0rt, it);
        if (DERIV) nudmu0rt = nudmu0rt - sel_get(S.vdmu0rt, it);
    }
    for (int k = 0; k < m->num_nu[i][1]; k++) {
        int it = m->ind_nu[i][1][k];
        numu0rt = numu0rt + sel_get(S.vmu0rt, it);
        if (DERIV) nudmu0rt = nudmu0rt + sel_get(S.vdmu0rt, it);
    }
    vlogk[1] = vlogk[0] - m->snu[i] * logPRT + numu0rt;
    if (DERIV) dlnkdT[1] = dlnkdT[0] + m->snu[i] * Tinv + nudmu0rt;
}

// third-body concentration (FJ.f90:199-210)
template <int N>
SHCFD_HD double third_body(const shcfd_mech* m, int i, const double (&n)[N], double sn, double logsn, double& vM) {
    double tmp = sn;
    for (int j = 0; j < m->num_m[i]; j++) tmp = tmp + m->men[i][j] * sel_get(n, m->ind_m[i][j]);
    vM = tmp;
    return (tmp == sn) ? logsn : dlog(tmp);
}

// forward minus reverse rate of reaction i (FJ.f90:272-284)
template <int NS>
SHCFD_HD void rates(const shcfd_mech* m, int i, const double (&n)[NS + 1], const double* vlogk, double& rf, double& rr) {
    double r = dexp(vlogk[0]);
    for (int k = 0; k < m->num_nu[i][0]; k++) r = r * sel_get(n, m->ind_nu[i][0][k]);
    rf = r;
    rr = 0.0;
    if (m->reverse_off[i] == 0) {
        r = dexp(vlogk[1]);
        for (int k = 0; k < m->num_nu[i][1]; k++) r = r * sel_get(n, m->ind_nu[i][1][k]);
        rr = r;
    }
}

// Fex: FJ.f90:152-300.  n[0..NS) mol/cm^3, n[NS] = T.
template <int NS>
SHCFD_HD void fex(const shcfd_mech* m, const double (&n)[NS + 1], double (&dndt)[NS + 1]) {
    double T = n[NS], logT = dlog(T), Tinv = 1.0 / T, logPRT = dlog(PST / RU) - logT;
    double sn = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) sn = sn + n[j];
    double logsn = dlog(sn);
    SpeciesThermo<NS, false> S;
    species_thermo<NS, false>(m, T, logT, S);
#pragma unroll
    for (int j = 0; j <= NS; j++) dndt[j] = 0.0;
    double vlogM = 0.0, vM = 0.0, unused = 0.0;
    for (int i = 0; i < m->nr; i++) {
        if (m->exist_m[i]) vlogM = third_body(m, i, n, sn, logsn, vM);
        double vlogk[2];
        rate_constants<NS, false>(m, i, T, logT, Tinv, logPRT, vlogM, S, vlogk, nullptr, unused);
        double rf, rr;
        rates<NS>(m, i, n, vlogk, rf, rr);
        double Dr = (m->reverse_off[i] == 0) ? rf - rr : rf;
        for (int k = 0; k < m->num_nu[i][0]; k++) sel_sub(dndt, m->ind_nu[i][0][k], Dr);
        for (int k = 0; k < m->num_nu[i][1]; k++) sel_add(dndt, m->ind_nu[i][1][k], Dr);
    }
    double de = 0.0, cv = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) {
        de = de + S.vert[j] * dndt[j];
        cv = cv + S.vcvr[j] * n[j];
    }
    dndt[NS] = -(de * T / cv);
}

// Fex_var: FJ.f90:301-453 (production >= 0, destruction <= 0)
template <int NS>
SHCFD_HD void fex_var(const shcfd_mech* m, const double (&n)[NS + 1], double (&dp)[NS + 1], double (&dm)[NS + 1]) {
    double T = n[NS], logT = dlog(T), Tinv = 1.0 / T, logPRT = dlog(PST / RU) - logT;
    double sn = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) sn = sn + n[j];
    double logsn = dlog(sn);
    SpeciesThermo<NS, false> S;
    species_thermo<NS, false>(m, T, logT, S);
#pragma unroll
    for (int j = 0; j <= NS; j++) {
        dp[j] = 0.0;
        dm[j] = 0.0;
    }
    double vlogM = 0.0, vM = 0.0, unused = 0.0;
    for (int i = 0; i < m->nr; i++) {
        if (m->exist_m[i]) vlogM = third_body(m, i, n, sn, logsn, vM);
        double vlogk[2];
        rate_constants<NS, false>(m, i, T, logT, Tinv, logPRT, vlogM, S, vlogk, nullptr, unused);
        double rf, rr;
        rates<NS>(m, i, n, vlogk, rf, rr);
        for (int k = 0; k < m->num_nu[i][0]; k++) sel_sub(dm, m->ind_nu[i][0][k], rf);
        for (int k = 0; k < m->num_nu[i][1]; k++) sel_add(dp, m->ind_nu[i][1][k], rf);
        if (m->reverse_off[i] == 0) {
            for (int k = 0; k < m->num_nu[i][0]; k++) sel_add(dp, m->ind_nu[i][0][k], rr);
            for (int k = 0; k < m->num_nu[i][1]; k++) sel_sub(dm, m->ind_nu[i][1][k], rr);
        }
    }
    double de = 0.0, cv = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) {
        de = de + S.vert[j] * (dp[j] + dm[j]);
        cv = cv + S.vcvr[j] * n[j];
    }
    double dT = -(de * T / cv);
    dp[NS] = fmax2(dT, 0.0);
    dm[NS] = fmin2(dT, 0.0);
}

// Storage of the (NS+1)^2 Jacobian for jex_impl and the LU: registers
// (indices compile-time after unrolling; a runtime row is selected by
// comparison) or per-lane LDS (lane-interleaved, runtime row indexing).
template <int N>
struct RegMat {
    double (&J)[N][N];
    SHCFD_HD double get(int a, int b) const { return J[a][b]; }
    SHCFD_HD void set(int a, int b, double v) { J[a][b] = v; }
    template <int NR>
    SHCFD_HD void sub_row(int it, const double (&v)[N]) {
#pragma unroll
        for (int a = 0; a < NR; a++)
            if (a == it) {
#pragma unroll
                for (int b = 0; b < N; b++) J[a][b] = J[a][b] - v[b];
            }
    }
    template <int NR>
    SHCFD_HD void add_row(int it, const double (&v)[N]) {
#pragma unroll
        for (int a = 0; a < NR; a++)
            if (a == it) {
#pragma unroll
                for (int b = 0; b < N; b++) J[a][b] = J[a][b] + v[b];
            }
    }
};
template <int N>
struct LdsMat {
    double* p;  // this lane's element (0,0); element (a,b) at p[(a N + b) 64]
    __device__ double get(int a, int b) const { return p[(a * N + b) * 64]; }
    __device__ void set(int a, int b, double v) { p[(a * N + b) * 64] = v; }
    template <int NR>
    __device__ void sub_row(int it, const double (&v)[N]) {
#pragma unroll
        for (int b = 0; b < N; b++) p[(it * N + b) * 64] = p[(it * N + b) * 64] - v[b];
    }
    template <int NR>
    __device__ void add_row(int it, const double (&v)[N]) {
#pragma unroll
        for (int b = 0; b < N; b++) p[(it * N + b) * 64] = p[(it * N + b) * 64] + v[b];
    }
};

// Jex: FJ.f90:454-708.  J(a,b) = d(dndt_a)/d(n_b) (row a, column b).
template <int NS, class Mat>
SHCFD_HD void jex_impl(const shcfd_mech* m, const double (&n)[NS + 1], Mat& J, double (&dndt)[NS + 1]) {
    constexpr int N = NS + 1;
    double T = n[NS], logT = dlog(T), Tinv = 1.0 / T, logPRT = dlog(PST / RU) - logT;
    double sn = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) sn = sn + n[j];
    sn = fmax2(sn, 1e-300);
    double logsn = dlog(sn);
    SpeciesThermo<NS, true> S;
    species_thermo<NS, true>(m, T, logT, S);
#pragma unroll
    for (int a = 0; a < N; a++) {
        dndt[a] = 0.0;
#pragma unroll
        for (int b = 0; b < N; b++) J.set(a, b, 0.0);
    }
    double vlogM = 0.0, vM = 0.0, dlnkdlnM = 0.0;
    for (int i = 0; i < m->nr; i++) {
        if (m->exist_m[i]) vlogM = third_body(m, i, n, sn, logsn, vM);
        double vlogk[2], dlnkdT[2];
        rate_constants<NS, true>(m, i, T, logT, Tinv, logPRT, vlogM, S, vlogk, dlnkdT, dlnkdlnM);
        double drdn[N];
#pragma unroll
        for (int k = 0; k < NS; k++) drdn[k] = 0.0;
        double r = dexp(vlogk[0]);
        for (int k = 0; k < m->num_nu[i][0]; k++) r = r * sel_get(n, m->ind_nu[i][0][k]);
        double Dr = r;
        drdn[NS] = dlnkdT[0] * r;
        for (int k = 0; k < m->num_nu[i][0]; k++) {
            int it = m->ind_nu[i][0][k];
            sel_add(drdn, it, r / sel_get(n, it));
        }
        if (m->reverse_off[i] == 0) {
            r = dexp(vlogk[1]);
            for (int k = 0; k < m->num_nu[i][1]; k++) r = r * sel_get(n, m->ind_nu[i][1][k]);
            Dr = Dr - r;
            drdn[NS] = drdn[NS] - r * dlnkdT[1];
            for (int k = 0; k < m->num_nu[i][1]; k++) {
                int it = m->ind_nu[i][1][k];
                sel_sub(drdn, it, r / sel_get(n, it));
            }
        }
        if (m->exist_m[i]) {
            double t1 = Dr * dlnkdlnM / vM;
#pragma unroll
            for (int k = 0; k < NS; k++) drdn[k] = drdn[k] + t1;
            for (int k = 0; k < m->num_m[i]; k++) sel_add(drdn, m->ind_m[i][k], t1 * m->men[i][k]);
        }
        for (int k = 0; k < m->num_nu[i][0]; k++) {
            int it = m->ind_nu[i][0][k];
            sel_sub(dndt, it, Dr);
            J.template sub_row<NS>(it, drdn);
        }
        for (int k = 0; k < m->num_nu[i][1]; k++) {
            int it = m->ind_nu[i][1][k];
            sel_add(dndt, it, Dr);
            J.template add_row<NS>(it, drdn);
        }
    }
    // FJ.f90:680-697
    double de = 0.0, cv = 0.0, dcv = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) {
        t1 = t1 + S.vert[j] * J.get(j, NS);
        de = de + S.vert[j] * dndt[j];
        dcv = dcv + S.vcvr[j] * dndt[j];
        cv = cv + S.vcvr[j] * n[j];
        t2 = t2 + S.vdcvrdT[j] * n[j];
    }
    double dTdt = -(de * T / cv);
    dndt[NS] = dTdt;
    J.set(NS, NS, -((T * t1 + dcv + dTdt * t2) / cv));
#pragma unroll
    for (int j = 0; j < NS; j++) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < NS; k++) s = s + S.vert[k] * J.get(k, j);
        J.set(NS, j, -((T * s + dTdt * S.vcvr[j]) / cv));
    }
}

template <int NS>
SHCFD_HD void jex(const shcfd_mech* m, const double (&n)[NS + 1], double (&J)[NS + 1][NS + 1], double (&dndt)[NS + 1]) {
    RegMat<NS + 1> R{J};
    jex_impl<NS>(m, n, R, dndt);
}

// rho_rel2abs: sub_chem.f90:1036-1085 (vrho relative in, kg/m^3 out)
template <int NS>
SHCFD_HD void rho_rel2abs(const shcfd_mech* m, double p, double T, double* vrho, double& rho, double& E) {
    double n[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) n[j] = vrho[j] * m->inv_mw[j];
    double tmp = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) tmp = tmp + n[j];
    tmp = tmp * RU * T;
    tmp = p * 1e1 / tmp;
#pragma unroll
    for (int j = 0; j < NS; j++) n[j] = n[j] * tmp;
    double r = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) r = r + n[j] * m->mw[j];
    r = r * 1e3;
#pragma unroll
    for (int j = 0; j < NS; j++) vrho[j] = n[j] * m->mw[j] * 1e3;
    Basis B;
    set_coeff<false>(T, dlog(T), B);
    double e = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) {
        const double* c = coeff_at(m, j, T);
        double dh = 0.0;
#pragma unroll
        for (int i = 0; i < 7; i++) dh = dh + c[i] * B.chrt[i];
        e = e + (dh - 1.0) * n[j];
    }
    E = e * RU * T / r / 1e1;
    rho = r;
}

// calc_T: sub_chem.f90:1507-1645 with ns_tocalc = ns.  T is the Newton start
// on entry.  Returns the iteration count, or -1 past max_T_loop = 1000.
template <int NS>
struct ThermoOut {
    double kappa, mwave, mu, dhi[NS], vhi[NS];
};
template <int NS>
SHCFD_HD int calc_t(const shcfd_mech* m, const double* vrho, double rho, double E, double& T_io, ThermoOut<NS>& out) {
    double n[NS];
    double tmp = rho * 1e-11;
#pragma unroll
    for (int j = 0; j < NS; j++) n[j] = (vrho[j] > tmp) ? vrho[j] * m->inv_mw[j] * 1e-3 : 0.0;
    double Eobj = E * 1e1 * rho / RU;
    double T = T_io;
    double chrt[6], cvnow = 0.0;
    int k = 1;
    for (;;) {
        double Ti1 = 1.0 / T, T2 = T * T, T3 = T2 * T, T4 = T3 * T;  // set_coeff_less, :20-49
        chrt[0] = 1.0;
        chrt[1] = T / 2.0;
        chrt[2] = T2 / 3.0;
        chrt[3] = T3 / 4.0;
        chrt[4] = T4 / 5.0;
        chrt[5] = Ti1;
        double ccpr[6] = {1.0, T, T2, T3, T4, 0.0};
        double Enow = 0.0;
        cvnow = 0.0;
#pragma unroll
        for (int j = 0; j < NS; j++) {
            if (n[j] > 0.0) {
                const double* c = coeff_at(m, j, T);
                double de = -1.0, dcv = -1.0;
#pragma unroll
                for (int i = 0; i < 6; i++) {
                    de = de + c[i] * chrt[i];
                    dcv = dcv + c[i] * ccpr[i];
                }
                Enow = Enow + de * n[j];
                cvnow = cvnow + dcv * n[j];
            }
        }
        double dT = (Eobj - Enow * T) / cvnow;
        double adT = dT < 0.0 ? -dT : dT, dT1 = T - 1e3;
        if (dT1 < 0.0) dT1 = -dT1;
        if (adT < 1e-8 || (adT < 1e-4 && dT1 < 1e-4) || k > 1000) break;
        T = fmax2(100.0, T + dT);
        k = k + 1;
    }
    T_io = T;
    if (k > 1000) return -1;
    double sn = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) sn = sn + n[j];
    double kappa = 1.0 + sn / cvnow;
    out.kappa = kappa;
    out.mwave = rho / sn * 1e-3;
    tmp = RU * 1e-4 * T;
#pragma unroll
    for (int j = 0; j < NS; j++) {
        const double* c = coeff_at(m, j, T);
        double dh = 0.0;
#pragma unroll
        for (int i = 0; i < 6; i++) dh = dh + c[i] * chrt[i];
        out.dhi[j] = tmp * m->inv_mw[j] * (kappa - (kappa - 1.0) * dh);
        out.vhi[j] = tmp * m->inv_mw[j] * dh;
    }
    // Wilke mixture viscosity (:1600-1633)
    double logT = dlog(T), invT = 1.0 / T;
    double cmu[4] = {logT, invT, invT * invT, 1.0};
    double muN[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) muN[i] = 0.0;
    for (int i = 0; i < m->nt; i++) {
        int s = -1;  // check_section_number_trans, :151-173
        if (m->trange[i][0][0] > T) s = 0;
        for (int l = 0; l < m->nsec_trans[i]; l++)
            if (m->trange[i][l][1] > T) {
                s = l;
                break;
            }
        if (s == -1) s = m->nsec_trans[i] - 1;
        double v = 0.0;
        if (sel_get(n, m->tr2th[i]) > 0.0) {
            double t = 0.0;
#pragma unroll
            for (int j = 0; j < 4; j++) t = t + cmu[j] * m->trans[i][s][j];
            v = dexp(t);
        }
#pragma unroll
        for (int q = 0; q < NS; q++)
            if (q == i) muN[q] = v;
    }
    double mu = 0.0;
    for (int i = 0; i < m->nt; i++) {
        double mui = sel_get(muN, i);
        if (mui > 0.0) {
            int it = m->tr2th[i];
            double denom = 0.0;
            for (int j = 0; j < m->nt; j++) {
                double muj = sel_get(muN, j);
                if (muj > 0.0) {
                    int jt = m->tr2th[j];
                    double x = 1.0 + dsqrt(mui / muj) * dsqrt(dsqrt(m->mw[jt] / m->mw[it]));
                    double phi = 0.25 * (x * x) * dsqrt(2.0 * m->mw[jt] / (m->mw[it] + m->mw[jt]));
                    denom = denom + sel_get(n, jt) * phi;
                }
            }
            mu = mu + sel_get(n, it) * mui / denom;
        }
    }
    out.mu = mu * 1e-7;
    return k;
}

// chemeq2's return value when it hands the rest of the interval to DVODE
constexpr int CHEM_DEFER = 1 << 20;

// chemeq2: sub_chem.f90:1714-1819.  y[0..NS) mol/cm^3, y[NS] = T, from 0 to
// tout.  Once more than 10 attempts were made the reference hands the rest
// of the interval, from the start of the current substep, to DVODE
// (:1751-1770): this returns CHEM_DEFER with y holding that substep's
// initial state and tt_out its time, and the caller continues with
// chemeq2_dvode (shcfd_dvode.h).  The reference's "not converted at
// CHEMEQ2" stop (:1810) returns SHCFD_E_CHEM.
template <int NS>
SHCFD_HD int chemeq2(const shcfd_mech* m, double (&y)[NS + 1], double tout, double rtol, int& attempts,
                     double& tt_out) {
    constexpr int N = NS + 1;
    double y0[N], yp[N], p[N], q[N], p0[N], q0[N];
    double tt = 0.0, dt = tout - tt, ratio = 1e300;
    int countersum = 0;
    for (;;) {
#pragma unroll
        for (int i = 0; i < N; i++) y0[i] = y[i];
        double counter = 0.0;
        if (tout - tt < dt) dt = tout - tt;
        for (;;) {
#pragma unroll
            for (int i = 0; i < N; i++) y[i] = y0[i];
            countersum = countersum + 1;
            if (countersum > 10) {
                attempts = countersum;
                tt_out = tt;
                return CHEM_DEFER;
            }
            fex_var<NS>(m, y, q0, p0);
#pragma unroll
            for (int i = 0; i < NS; i++) {
                p0[i] = -p0[i] / (y[i] + 1e-300);
                double var = p0[i] * dt, var2 = var * var, var3 = var * var2;
                double alp = (180.0 + 60.0 * var + 11.0 * var2 + var3) / (360.0 + 60.0 * var + 12.0 * var2 + var3);
                yp[i] = y0[i] + dt * (q0[i] - p0[i] * y0[i]) / (1.0 + alp * p0[i] * dt);
            }
            yp[NS] = y0[NS] + dt * (q0[NS] + p0[NS]);
            ratio = 1e300;
            fex_var<NS>(m, yp, q, p);
#pragma unroll
            for (int i = 0; i < NS; i++) {
                p[i] = -p[i] / (yp[i] + 1e-300);
                p[i] = 0.5 * (p0[i] + p[i]);
                double var = p[i] * dt, var2 = var * var, var3 = var * var2;
                double alp = (180.0 + 60.0 * var + 11.0 * var2 + var3) / (360.0 + 60.0 * var + 12.0 * var2 + var3);
                q[i] = alp * q[i] + (1.0 - alp) * q0[i];
                y[i] = y0[i] + dt * (q[i] - p[i] * y0[i]) / (1.0 + alp * p[i] * dt);
                double d = y[i] - yp[i] + 1e-300;
                var = (y[i] * rtol + rtol * 1e-5) / (d < 0.0 ? -d : d);
                ratio = fmin2(ratio, var);
            }
            y[NS] = y0[NS] + 0.5 * dt * (q[NS] + p[NS] + q0[NS] + p0[NS]);
            if (1.0 < ratio) break;
            counter = counter + 1.0;
            dt = dt * dsqrt(ratio) / counter;
            if (counter > 10.0) {
                attempts = countersum;
                return SHCFD_E_CHEM;
            }
        }
        tt = tt + dt;
        if (tout <= tt) break;
        dt = dt * dsqrt(ratio);
    }
    attempts = countersum;
    return 0;
}

// the conversions of reaction_chemeq2 (sub_chem.f90:1696-1700, :1707-1711)
template <int NS>
SHCFD_HD void vrho_to_n(const shcfd_mech* m, const double (&vrho)[NS], double (&n)[NS + 1]) {
#pragma unroll
    for (int j = 0; j < NS; j++) {
        n[j] = vrho[j] * m->inv_mw[j] * 1e-3;
        n[j] = fmax2(n[j], N_EPS);
    }
}
template <int NS>
SHCFD_HD void n_to_vrho(const shcfd_mech* m, double (&n)[NS + 1], double (&vrho)[NS]) {
#pragma unroll
    for (int j = 0; j < NS; j++) {
        n[j] = fmax2(n[j], N_EPS);
        vrho[j] = n[j] * m->mw[j] * 1e3;
    }
}

// reaction_chemeq2: sub_chem.f90:1684-1713 (vrho kg/m^3 in/out).  On
// CHEM_DEFER, vrho is untouched and n, tt hold the state DVODE continues
// from (chemeq2_dvode in shcfd_dvode.h, then n_to_vrho).
template <int NS>
SHCFD_HD int reaction_chemeq2(const shcfd_mech* m, double T, double (&vrho)[NS], double tout, int& attempts,
                              double (&n)[NS + 1], double& tt) {
    vrho_to_n<NS>(m, vrho, n);
    attempts = 0;
    tt = 0.0;
    if (T < 0.0 || tout < 0.0) return SHCFD_E_ARG;
    n[NS] = T;
    int e = chemeq2<NS>(m, n, tout, m->rtol, attempts, tt);
    if (e) return e;
    n_to_vrho<NS>(m, n, vrho);
    return 0;
}

// LU with partial pivoting and solve, in the order of LINPACK dgefa/dgesl
// (dvode.f:3816-3931, :3932-4040): column k's pivot is the first largest
// |a(i,k)|, i >= k; multipliers are -a(i,k)/a(k,k) stored below the
// diagonal; the row swap is applied column by column.  All indices are
// compile-time after unrolling; the pivot row is selected per column.
// A[row][col].  Returns 0 or k+1 for a zero pivot.
template <int N>
SHCFD_HD int lu_solve(double (&A)[N][N], double (&b)[N]) {
    int info = 0;
    int piv[N];
#pragma unroll
    for (int k = 0; k < N - 1; k++) {
        int l = k;
        double amax = A[k][k] < 0.0 ? -A[k][k] : A[k][k];
#pragma unroll
        for (int i = k + 1; i < N; i++) {
            double a = A[i][k] < 0.0 ? -A[i][k] : A[i][k];
            if (a > amax) {
                amax = a;
                l = i;
            }
        }
        piv[k] = l;
        // pivot value and swap within column k
        double pv = A[k][k];
#pragma unroll
        for (int i = k + 1; i < N; i++)
            if (i == l) pv = A[i][k];
        if (pv == 0.0) {
            info = k + 1;
            continue;
        }
#pragma unroll
        for (int i = k + 1; i < N; i++)
            if (i == l) A[i][k] = A[k][k];
        A[k][k] = pv;
        double t = -1.0 / pv;
#pragma unroll
        for (int i = k + 1; i < N; i++) A[i][k] = t * A[i][k];
#pragma unroll
        for (int j = k + 1; j < N; j++) {
            double tj = A[k][j];
#pragma unroll
            for (int i = k + 1; i < N; i++)
                if (i == l) tj = A[i][j];
#pragma unroll
            for (int i = k + 1; i < N; i++)
                if (i == l) A[i][j] = A[k][j];
            A[k][j] = tj;
#pragma unroll
            for (int i = k + 1; i < N; i++) A[i][j] = A[i][j] + tj * A[i][k];
        }
    }
    piv[N - 1] = N - 1;
    if (A[N - 1][N - 1] == 0.0) info = N;
    if (info) return info;
    // dgesl, job = 0
#pragma unroll
    for (int k = 0; k < N - 1; k++) {
        int l = piv[k];
        double t = b[k];
#pragma unroll
        for (int i = k + 1; i < N; i++)
            if (i == l) t = b[i];
#pragma unroll
        for (int i = k + 1; i < N; i++)
            if (i == l) b[i] = b[k];
        b[k] = t;
#pragma unroll
        for (int i = k + 1; i < N; i++) b[i] = b[i] + t * A[i][k];
    }
#pragma unroll
    for (int kb = 0; kb < N; kb++) {
        int k = N - 1 - kb;
        b[k] = b[k] / A[k][k];
        double t = -b[k];
#pragma unroll
        for (int i = 0; i < k; i++) b[i] = b[i] + t * A[i][k];
    }
    return 0;
}

// Point-implicit source update over dt: the backward-Euler step
// y1 = y0 + dt F(y1) (species and T at constant rho and e) solved by Newton
// iterations y <- y - (I - dt J(y))^-1 (y - y0 - dt F(y)) with the analytic
// Jacobian and the LU above -- DVODE's first-order Newton matrix
// P = I - h l0 J, l0 = 1 (dvode.f DVJAC :2957-3140).  Concentrations are
// floored at n_eps after each iteration (as reaction_chemeq2 floors them);
// converged when every species moved by less than PI_RTOL of itself plus
// PI_ATOL of the total concentration and T by less than PI_RTOL*T; fails
// (SHCFD_E_CHEM, vrho untouched) after PI_MAXIT iterations, on a zero pivot
// or on T <= 0.  The first iterate is the classic linearised point-implicit
// update.  T is not written back to the flow (the thermo recomputes it from
// e); T_out returns the backward-Euler temperature.
constexpr int PI_MAXIT = 32;
constexpr double PI_RTOL = 1e-10;
constexpr double PI_ATOL = 1e-14;
template <int NS>
SHCFD_HD int point_implicit(const shcfd_mech* m, double T, double (&vrho)[NS], double dt, int& iters, double& T_out) {
    constexpr int N = NS + 1;
    double y0[N], y[N];
#pragma unroll
    for (int j = 0; j < NS; j++) {
        y0[j] = vrho[j] * m->inv_mw[j] * 1e-3;
        y0[j] = fmax2(y0[j], N_EPS);
        y[j] = y0[j];
    }
    y0[NS] = T;
    y[NS] = T;
    double atol = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) atol = atol + y0[j];
    atol = atol * PI_ATOL;
    int it = 0;
    bool done = false;
    while (!done) {
        if (it >= PI_MAXIT) {
            iters = it;
            return SHCFD_E_CHEM;
        }
        it = it + 1;
        // F(y) is the dn/dt Jex forms beside the Jacobian: the rates of Fex
        // (FJ.f90:152-300) in the same order, equal to Fex's result in every
        // bit for a positive total concentration (y is floored at N_EPS), so
        // each rate is evaluated once per iteration
        double f[N], J[N][N];
        jex<NS>(m, y, J, f);
#pragma unroll
        for (int a = 0; a < N; a++) {
#pragma unroll
            for (int b = 0; b < N; b++) J[a][b] = -(dt * J[a][b]);
            J[a][a] = J[a][a] + 1.0;
            f[a] = (y0[a] - y[a]) + dt * f[a];
        }
        if (lu_solve<N>(J, f) != 0) {
            iters = it;
            return SHCFD_E_CHEM;
        }
        done = true;
#pragma unroll
        for (int j = 0; j < NS; j++) {
            double yn = fmax2(y[j] + f[j], N_EPS);
            double d = yn - y[j];
            if ((d < 0.0 ? -d : d) > PI_RTOL * y[j] + atol) done = false;
            y[j] = yn;
        }
        double Tn = y[NS] + f[NS];
        double dT = Tn - y[NS];
        if ((dT < 0.0 ? -dT : dT) > PI_RTOL * y[NS]) done = false;
        y[NS] = Tn;
        if (!(y[NS] > 0.0)) {
            iters = it;
            return SHCFD_E_CHEM;
        }
    }
#pragma unroll
    for (int j = 0; j < NS; j++) vrho[j] = y[j] * m->mw[j] * 1e3;
    iters = it;
    T_out = y[NS];
    return 0;
}

// lu_solve with the matrix in per-lane LDS: the same pivot choice, swaps and
// operation order as lu_solve, with runtime row indices.  Column k, once
// pivoted and scaled, stays in registers for the updates of the columns
// right of it (the loop over k is unrolled, so the copy is indexed at
// compile time); every element still receives the same operations in the
// same order.
template <int N>
__device__ int lu_solve_lds(LdsMat<N>& A, double (&b)[N]) {
    int info = 0;
    int piv[N];
#pragma unroll
    for (int k = 0; k < N - 1; k++) {
        double col[N];
#pragma unroll
        for (int i = k; i < N; i++) col[i] = A.get(i, k);
        int l = k;
        double amax = col[k] < 0.0 ? -col[k] : col[k];
#pragma unroll
        for (int i = k + 1; i < N; i++) {
            const double a = col[i] < 0.0 ? -col[i] : col[i];
            if (a > amax) {
                amax = a;
                l = i;
            }
        }
        piv[k] = l;
        double pv = col[k];
#pragma unroll
        for (int i = k + 1; i < N; i++)
            if (i == l) pv = col[i];
        if (pv == 0.0) {
            info = k + 1;
            continue;
        }
#pragma unroll
        for (int i = k + 1; i < N; i++)
            if (i == l) col[i] = col[k];
        col[k] = pv;
        const double t = -1.0 / pv;
#pragma unroll
        for (int i = k + 1; i < N; i++) col[i] = t * col[i];
#pragma unroll
        for (int i = k; i < N; i++) A.set(i, k, col[i]);
        for (int j = k + 1; j < N; j++) {
            const double tj = A.get(l, j);
            if (l != k) A.set(l, j, A.get(k, j));
            A.set(k, j, tj);
#pragma unroll
            for (int i = k + 1; i < N; i++) A.set(i, j, A.get(i, j) + tj * col[i]);
        }
    }
    piv[N - 1] = N - 1;
    if (A.get(N - 1, N - 1) == 0.0) info = N;
    if (info) return info;
#pragma unroll
    for (int k = 0; k < N - 1; k++) {
        const int l = piv[k];
        double t = b[k];
#pragma unroll
        for (int i = k + 1; i < N; i++)
            if (i == l) t = b[i];
#pragma unroll
        for (int i = k + 1; i < N; i++)
            if (i == l) b[i] = b[k];
        b[k] = t;
#pragma unroll
        for (int i = k + 1; i < N; i++) b[i] = b[i] + t * A.get(i, k);
    }
#pragma unroll
    for (int kb = 0; kb < N; kb++) {
        const int k = N - 1 - kb;
        b[k] = b[k] / A.get(k, k);
        const double t = -b[k];
#pragma unroll
        for (int i = 0; i < k; i++) b[i] = b[i] + t * A.get(i, k);
    }
    return 0;
}

// point_implicit with the Newton matrix in per-lane LDS (Jl: this lane's
// element (0,0), stride 64 doubles): the same iteration and arithmetic
template <int NS>
__device__ int point_implicit_lds(const shcfd_mech* m, double T, double (&vrho)[NS], double dt, int& iters,
                                  double& T_out, double* Jl) {
    constexpr int N = NS + 1;
    LdsMat<N> J{Jl};
    double y0[N], y[N];
#pragma unroll
    for (int j = 0; j < NS; j++) {
        y0[j] = vrho[j] * m->inv_mw[j] * 1e-3;
        y0[j] = fmax2(y0[j], N_EPS);
        y[j] = y0[j];
    }
    y0[NS] = T;
    y[NS] = T;
    double atol = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) atol = atol + y0[j];
    atol = atol * PI_ATOL;
    int it = 0;
    bool done = false;
    while (!done) {
        if (it >= PI_MAXIT) {
            iters = it;
            return SHCFD_E_CHEM;
        }
        it = it + 1;
        double f[N];
        jex_impl<NS>(m, y, J, f);  // F(y) beside J, as in point_implicit
        for (int a = 0; a < N; a++) {
            for (int b = 0; b < N; b++) J.set(a, b, -(dt * J.get(a, b)));
            J.set(a, a, J.get(a, a) + 1.0);
        }
#pragma unroll
        for (int a = 0; a < N; a++) f[a] = (y0[a] - y[a]) + dt * f[a];
        if (lu_solve_lds<N>(J, f) != 0) {
            iters = it;
            return SHCFD_E_CHEM;
        }
        done = true;
#pragma unroll
        for (int j = 0; j < NS; j++) {
            double yn = fmax2(y[j] + f[j], N_EPS);
            double d = yn - y[j];
            if ((d < 0.0 ? -d : d) > PI_RTOL * y[j] + atol) done = false;
            y[j] = yn;
        }
        double Tn = y[NS] + f[NS];
        double dT = Tn - y[NS];
        if ((dT < 0.0 ? -dT : dT) > PI_RTOL * y[NS]) done = false;
        y[NS] = Tn;
        if (!(y[NS] > 0.0)) {
            iters = it;
            return SHCFD_E_CHEM;
        }
    }
#pragma unroll
    for (int j = 0; j < NS; j++) vrho[j] = y[j] * m->mw[j] * 1e3;
    iters = it;
    T_out = y[NS];
    return 0;
}

}  // namespace chem
}  // namespace shcfd
