/*
 * shcfd_chem_oracle.c -- CPU restatement of the reference's Chemkin
 * thermochemistry (ISASShimadaLab/shared_cfd, store/therm_lib/chemkin).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP chemistry kernels
 * and part of the CPU baseline; never linked into the product.
 *
 * Every function works on a mechanism table (shcfd_mech, 0-based indices)
 * and a per-cell state, with no globals, so OpenMP loops over cells can call
 * them.  Arithmetic follows the reference's expressions in its evaluation
 * order (Fortran: * and / left to right, ** before *), compiled with
 * -ffp-contract=off.
 */
#include <math.h>
#include <pthread.h>
#include <string.h>

#include "shcfd_oracle.h"

/* ---- transcendental functions -------------------------------------------
 * The chemistry evaluates exp/log per reaction and per substep.  Both this
 * restatement and the GPU kernels use the same published methods with the
 * same operation order, so they agree to the last bit; against a libm these
 * are within 2 ulp (checked in the tests).
 *   exp: Cody-Waite reduction x = k ln2 + r (ln2 split hi/lo), |r| <= ln2/2,
 *        Taylor series of e^r to degree 13 (Horner, fused), times 2^k.
 *   log: x = 2^k m, m in [sqrt(1/2), sqrt(2)), f = (m-1)/(m+1),
 *        ln m = 2 atanh f = 2f (1 + f^2/3 + f^4/5 + ... + f^20/21).
 * The reference calls the Fortran intrinsics exp/log/log10 and x**0.25;
 * x**0.25 is taken as sqrt(sqrt(x)). */
#define LN2_HI 6.93147180369123816490e-01
#define LN2_LO 1.90821492927058770002e-10
#define INV_LN2 1.44269504088896338700e+00
#define INV_LN10 4.34294481903251827651e-01

static double pow2i(int k) { /* 2^k for -1022 <= k <= 1023 */
    union { double d; unsigned long long u; } v;
    v.u = (unsigned long long)(k + 1023) << 52;
    return v.d;
}

double or_exp(double x) {
    if (x != x) return x + x;
    if (x > 709.782712893384) return 1.0 / 0.0;
    if (x < -745.1332191019412) return 0.0;
    double kd = rint(x * INV_LN2);
    int k = (int)kd;
    double r = fma(-kd, LN2_HI, x);
    r = fma(-kd, LN2_LO, r);
    double p = 1.0 / 6227020800.0;       /* 1/13! */
    p = fma(p, r, 1.0 / 479001600.0);    /* 1/12! */
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    int k1 = k / 2, k2 = k - k1;
    return (p * pow2i(k1)) * pow2i(k2);
}

double or_log(double x) {
    if (x != x || x < 0.0) return 0.0 / 0.0;
    if (x == 0.0) return -1.0 / 0.0;
    if (x == 1.0 / 0.0) return x;
    int kadj = 0;
    if (x < 2.2250738585072014e-308) {
        x = x * 18014398509481984.0; /* 2^54 */
        kadj = -54;
    }
    union { double d; unsigned long long u; } v;
    v.d = x;
    int k = (int)((v.u >> 52) & 0x7ff) - 1023 + kadj;
    v.u = (v.u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m = v.d;
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        k = k + 1;
    }
    double f = (m - 1.0) / (m + 1.0);
    double s = f * f;
    double p = 1.0 / 21.0;
    p = fma(p, s, 1.0 / 19.0);
    p = fma(p, s, 1.0 / 17.0);
    p = fma(p, s, 1.0 / 15.0);
    p = fma(p, s, 1.0 / 13.0);
    p = fma(p, s, 1.0 / 11.0);
    p = fma(p, s, 1.0 / 9.0);
    p = fma(p, s, 1.0 / 7.0);
    p = fma(p, s, 1.0 / 5.0);
    p = fma(p, s, 1.0 / 3.0);
    double t = s * p;
    double f2 = f + f;
    double lm = fma(f2, t, f2);
    double kd = (double)k;
    return kd * LN2_HI + fma(kd, LN2_LO, lm);
}

double or_log10(double x) { return or_log(x) * INV_LN10; }

/* module const_chem, store/therm_lib/chemkin/core/mod_chem.raw.f90:8-14 */
#define RU 8.3144621e7
#define PST 1.01325e6
#define N_EPS 1e-40
#define DD 0.14
#define LOGE210 0.43429448190325

/* sub_chem.f90:3-19 -- 0: low range (T <= Tcommon), 1: high range */
static void check_section_number(const shcfd_mech *m, double T, int *sec) {
    for (int k = 0; k < m->ns; k++) sec[k] = (T <= m->tthre[k][2]) ? 0 : 1;
}

/* sub_chem.f90:50-85 (set_coeff) and :86-150 (set_coeff_more) */
static void set_coeff_more(double T, double logT, double *cmurt, double *chrt, double *ccpr, double *cdmurt,
                           double *cdcpr) {
    double Ti1 = 1.0 / T, Ti2 = Ti1 * Ti1, T2 = T * T, T3 = T2 * T, T4 = T3 * T;
    cmurt[0] = 1.0 - logT;
    cmurt[1] = -T / 2.0;
    cmurt[2] = -T2 / 6.0;
    cmurt[3] = -T3 / 12.0;
    cmurt[4] = -T4 / 20.0;
    cmurt[5] = Ti1;
    cmurt[6] = -1.0;
    chrt[0] = 1.0;
    chrt[1] = T / 2.0;
    chrt[2] = T2 / 3.0;
    chrt[3] = T3 / 4.0;
    chrt[4] = T4 / 5.0;
    chrt[5] = Ti1;
    chrt[6] = 0.0;
    ccpr[0] = 1.0;
    ccpr[1] = T;
    ccpr[2] = T2;
    ccpr[3] = T3;
    ccpr[4] = T4;
    ccpr[5] = 0.0;
    ccpr[6] = 0.0;
    if (cdmurt) {
        cdmurt[0] = -Ti1;
        cdmurt[1] = -1.0 / 2.0;
        cdmurt[2] = -T / 3.0;
        cdmurt[3] = -T2 / 4.0;
        cdmurt[4] = -T3 / 5.0;
        cdmurt[5] = -Ti2;
        cdmurt[6] = 0.0;
        cdcpr[0] = 0.0;
        cdcpr[1] = 1.0;
        cdcpr[2] = 2.0 * T;
        cdcpr[3] = 3.0 * T2;
        cdcpr[4] = 4.0 * T3;
        cdcpr[5] = 0.0;
        cdcpr[6] = 0.0;
    }
}

/* Per-reaction log rate constants shared by Fex, Fex_var and Jex
 * (FJ.f90:212-270, :358-416, :549-633).  Returns vlogk[0..1]; with
 * derivatives requested also dlnkdT[0..1] and dlnkdlnM.  vlogM/vM are the
 * third-body concentration terms of the current reaction. */
static void rate_constants(const shcfd_mech *m, int i, double T, double logT, double Tinv, double logPRT,
                           double vlogM, const double *vmu0rt, const double *vdmu0rt, double *vlogk, double *dlnkdT,
                           double *dlnkdlnM) {
    const double *abe = m->abe[i], *cabe = m->cabe[i];
    int kind = m->kind[i];
    vlogk[0] = abe[0] + abe[1] * logT - abe[2] * Tinv;
    vlogk[1] = 0.0;
    if (dlnkdT) {
        dlnkdT[0] = (abe[1] + abe[2] * Tinv) * Tinv;
        dlnkdT[1] = 0.0;
    }
    if (kind >= SHCFD_RXN_M) {
        vlogk[0] = vlogk[0] + vlogM;
        if (dlnkdlnM) *dlnkdlnM = 1.0;
    }
    if (m->reverse_off[i] != 0) return;
    if (kind == SHCFD_RXN_REV || kind == SHCFD_RXN_M_REV) {
        vlogk[1] = cabe[0] + cabe[1] * logT - cabe[2] * Tinv;
        if (dlnkdT) dlnkdT[1] = (cabe[1] + cabe[2] * Tinv) * Tinv;
        if (kind == SHCFD_RXN_M_REV) vlogk[1] = vlogk[1] + vlogM;
        return;
    }
    if (kind == SHCFD_RXN_LOW || kind == SHCFD_RXN_TROE) {
        double logk0 = cabe[0] + cabe[1] * logT - cabe[2] * Tinv;
        double logkinf = vlogk[0];
        double log_Pr = logk0 - logkinf + vlogM;
        double Pr = or_exp(log_Pr);
        vlogk[0] = logkinf + or_log(Pr / (1.0 + Pr));
        double dlnk0dT = 0.0, dlnkinfdT = 0.0;
        if (dlnkdT) {
            dlnk0dT = (cabe[1] + cabe[2] * Tinv) * Tinv;
            dlnkinfdT = dlnkdT[0];
            dlnkdT[0] = (dlnk0dT + Pr * dlnkinfdT) / (1.0 + Pr);
            *dlnkdlnM = 1.0 / (1.0 + Pr);
        }
        if (kind == SHCFD_RXN_TROE) {
            const double *tr = m->troe[i];
            double aa = tr[0];
            if (!dlnkdT) {
                /* FJ.f90:235-243 */
                double logFcent = or_log10((1.0 - aa) * or_exp(-T * tr[1]) + aa * or_exp(-T * tr[2]) + or_exp(-Tinv * tr[3]));
                double cc = -0.4 - 0.67 * logFcent;
                double nn = 0.75 - 1.27 * logFcent;
                double logPrc = log_Pr * LOGE210 + cc;
                double t = logPrc / (nn - DD * logPrc);
                double logF = (1.0 / (1.0 + t * t)) * logFcent;
                vlogk[0] = vlogk[0] + logF / LOGE210;
            } else {
                /* FJ.f90:593-612 */
                double dlnPrdT = dlnk0dT - dlnkinfdT;
                double Fcent = (1.0 - aa) * or_exp(-T * tr[1]) + aa * or_exp(-T * tr[2]) + or_exp(-tr[3] * Tinv);
                double dFcentdT = -(1.0 - aa) * tr[1] * or_exp(-T * tr[1]) - aa * tr[2] * or_exp(-T * tr[2]) +
                                  tr[3] * (Tinv * Tinv) * or_exp(-tr[3] * Tinv);
                double dlnFcentdT = dFcentdT / Fcent;
                double logFcent = or_log10(Fcent);
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
                *dlnkdlnM = *dlnkdlnM - 2.0 * temp2 * logF * temp1 / (den * den) * nn;
            }
        }
    }
    /* reverse from the equilibrium constant (FJ.f90:246-256, :615-630) */
    double numu0rt = 0.0, nudmu0rt = 0.0;
    for (int k = 0; k < m->num_nu[i][0]; k++) {
        numu0rt = numu0rt - vmu0rt[m->ind_nu[i][0][k]];
        if (dlnkdT) nudmu0rt = nudmu0rt - vdmu0rt[m->ind_nu[i][0][k]];
    }
    for (int k = 0; k < m->num_nu[i][1]; k++) {
        numu0rt = numu0rt + vmu0rt[m->ind_nu[i][1][k]];
        if (dlnkdT) nudmu0rt = nudmu0rt + vdmu0rt[m->ind_nu[i][1][k]];
    }
    vlogk[1] = vlogk[0] - m->snu[i] * logPRT + numu0rt;
    if (dlnkdT) dlnkdT[1] = dlnkdT[0] + m->snu[i] * Tinv + nudmu0rt;
}

/* third-body concentration (FJ.f90:199-210) */
static double third_body(const shcfd_mech *m, int i, const double *n, double sn, double logsn, double *vM) {
    double tmp = sn;
    for (int j = 0; j < m->num_m[i]; j++) tmp = tmp + m->men[i][j] * n[m->ind_m[i][j]];
    if (vM) *vM = tmp;
    return (tmp == sn) ? logsn : or_log(tmp);
}

/* species thermo sums at T (FJ.f90:185-197) */
static void species_thermo(const shcfd_mech *m, double T, double logT, double *vmu0rt, double *vert, double *vcvr) {
    int sec[SHCFD_CHEM_MAX_NS];
    double cmurt[7], chrt[7], ccpr[7];
    check_section_number(m, T, sec);
    set_coeff_more(T, logT, cmurt, chrt, ccpr, NULL, NULL);
    for (int j = 0; j < m->ns; j++) {
        double t1 = 0.0, t2 = -1.0, t3 = -1.0;
        for (int i = 0; i < 7; i++) {
            double tmp = m->coeff[j][sec[j]][i];
            t1 = t1 + tmp * cmurt[i];
            t2 = t2 + tmp * chrt[i];
            t3 = t3 + tmp * ccpr[i];
        }
        vmu0rt[j] = t1;
        vert[j] = t2;
        vcvr[j] = t3;
    }
}

/* Fex: store/therm_lib/chemkin/uvhp/uv/FJ.f90:152-300.
 * n[0..ns-1] in mol/cm^3, n[ns] = T; dndt likewise. */
void or_chem_fex(const shcfd_mech *m, const double *n, double *dndt) {
    int ns = m->ns;
    double T = n[ns], logT = or_log(T), Tinv = 1.0 / T, logPRT = or_log(PST / RU) - logT;
    double sn = 0.0;
    for (int j = 0; j < ns; j++) sn = sn + n[j];
    double logsn = or_log(sn);
    double vmu0rt[SHCFD_CHEM_MAX_NS], vert[SHCFD_CHEM_MAX_NS], vcvr[SHCFD_CHEM_MAX_NS];
    species_thermo(m, T, logT, vmu0rt, vert, vcvr);
    for (int j = 0; j < ns; j++) dndt[j] = 0.0;
    double vlogM = 0.0;
    for (int i = 0; i < m->nr; i++) {
        if (m->exist_m[i]) vlogM = third_body(m, i, n, sn, logsn, NULL);
        double vlogk[2];
        rate_constants(m, i, T, logT, Tinv, logPRT, vlogM, vmu0rt, NULL, vlogk, NULL, NULL);
        double r = or_exp(vlogk[0]);
        for (int k = 0; k < m->num_nu[i][0]; k++) r = r * n[m->ind_nu[i][0][k]];
        double Dr = r;
        if (m->reverse_off[i] == 0) {
            r = or_exp(vlogk[1]);
            for (int k = 0; k < m->num_nu[i][1]; k++) r = r * n[m->ind_nu[i][1][k]];
            Dr = Dr - r;
        }
        for (int k = 0; k < m->num_nu[i][0]; k++) dndt[m->ind_nu[i][0][k]] -= Dr;
        for (int k = 0; k < m->num_nu[i][1]; k++) dndt[m->ind_nu[i][1][k]] += Dr;
    }
    double de = 0.0, cv = 0.0;
    for (int j = 0; j < ns; j++) {
        de = de + vert[j] * dndt[j];
        cv = cv + vcvr[j] * n[j];
    }
    dndt[ns] = -(de * T / cv);
}

/* Fex_var: FJ.f90:301-453 -- production (>= 0) and destruction (<= 0) parts */
void or_chem_fex_var(const shcfd_mech *m, const double *n, double *dndtp, double *dndtm) {
    int ns = m->ns;
    double T = n[ns], logT = or_log(T), Tinv = 1.0 / T, logPRT = or_log(PST / RU) - logT;
    double sn = 0.0;
    for (int j = 0; j < ns; j++) sn = sn + n[j];
    double logsn = or_log(sn);
    double vmu0rt[SHCFD_CHEM_MAX_NS], vert[SHCFD_CHEM_MAX_NS], vcvr[SHCFD_CHEM_MAX_NS];
    species_thermo(m, T, logT, vmu0rt, vert, vcvr);
    for (int j = 0; j < ns; j++) {
        dndtp[j] = 0.0;
        dndtm[j] = 0.0;
    }
    double vlogM = 0.0;
    for (int i = 0; i < m->nr; i++) {
        if (m->exist_m[i]) vlogM = third_body(m, i, n, sn, logsn, NULL);
        double vlogk[2];
        rate_constants(m, i, T, logT, Tinv, logPRT, vlogM, vmu0rt, NULL, vlogk, NULL, NULL);
        double r = or_exp(vlogk[0]);
        for (int k = 0; k < m->num_nu[i][0]; k++) r = r * n[m->ind_nu[i][0][k]];
        for (int k = 0; k < m->num_nu[i][0]; k++) dndtm[m->ind_nu[i][0][k]] -= r;
        for (int k = 0; k < m->num_nu[i][1]; k++) dndtp[m->ind_nu[i][1][k]] += r;
        if (m->reverse_off[i] == 0) {
            r = or_exp(vlogk[1]);
            for (int k = 0; k < m->num_nu[i][1]; k++) r = r * n[m->ind_nu[i][1][k]];
            for (int k = 0; k < m->num_nu[i][0]; k++) dndtp[m->ind_nu[i][0][k]] += r;
            for (int k = 0; k < m->num_nu[i][1]; k++) dndtm[m->ind_nu[i][1][k]] -= r;
        }
    }
    double de = 0.0, cv = 0.0;
    for (int j = 0; j < ns; j++) {
        de = de + vert[j] * (dndtp[j] + dndtm[j]);
        cv = cv + vcvr[j] * n[j];
    }
    double dT = -(de * T / cv);
    dndtp[ns] = fmax(dT, 0.0);
    dndtm[ns] = fmin(dT, 0.0);
}

/* Jex: FJ.f90:454-708.  J column-major (ns+1)x(ns+1): J[a + (ns+1)*b] =
 * d(dndt_a)/d(n_b). */
void or_chem_jex(const shcfd_mech *m, const double *n, double *J) {
    int ns = m->ns, N = ns + 1;
    double T = n[ns], logT = or_log(T), Tinv = 1.0 / T, logPRT = or_log(PST / RU) - logT;
    double sn = 0.0;
    for (int j = 0; j < ns; j++) sn = sn + n[j];
    sn = fmax(sn, 1e-300);
    double logsn = or_log(sn);
    int sec[SHCFD_CHEM_MAX_NS];
    double cmurt[7], chrt[7], ccpr[7], cdmurt[7], cdcpr[7];
    double vmu0rt[SHCFD_CHEM_MAX_NS], vdmu0rt[SHCFD_CHEM_MAX_NS], vert[SHCFD_CHEM_MAX_NS], vcvr[SHCFD_CHEM_MAX_NS],
        vdcvrdT[SHCFD_CHEM_MAX_NS], dndt[SHCFD_CHEM_MAX_NS + 1], drdn[SHCFD_CHEM_MAX_NS + 1];
    check_section_number(m, T, sec);
    set_coeff_more(T, logT, cmurt, chrt, ccpr, cdmurt, cdcpr);
    for (int j = 0; j < ns; j++) {
        double t1 = 0.0, t2 = 0.0, t3 = -1.0, t4 = -1.0, t5 = 0.0;
        for (int i = 0; i < 7; i++) {
            double tmp = m->coeff[j][sec[j]][i];
            t1 = t1 + tmp * cmurt[i];
            t2 = t2 + tmp * cdmurt[i];
            t3 = t3 + tmp * chrt[i];
            t4 = t4 + tmp * ccpr[i];
            t5 = t5 + tmp * cdcpr[i];
        }
        vmu0rt[j] = t1;
        vdmu0rt[j] = t2;
        vert[j] = t3;
        vcvr[j] = t4;
        vdcvrdT[j] = t5;
        dndt[j] = 0.0;
    }
    for (int a = 0; a < N * N; a++) J[a] = 0.0;
    double vlogM = 0.0, vM = 0.0, dlnkdlnM = 0.0;
    for (int i = 0; i < m->nr; i++) {
        if (m->exist_m[i]) vlogM = third_body(m, i, n, sn, logsn, &vM);
        double vlogk[2], dlnkdT[2];
        rate_constants(m, i, T, logT, Tinv, logPRT, vlogM, vmu0rt, vdmu0rt, vlogk, dlnkdT, &dlnkdlnM);
        for (int k = 0; k < ns; k++) drdn[k] = 0.0;
        double r = or_exp(vlogk[0]);
        for (int k = 0; k < m->num_nu[i][0]; k++) r = r * n[m->ind_nu[i][0][k]];
        double Dr = r;
        drdn[ns] = dlnkdT[0] * r;
        for (int k = 0; k < m->num_nu[i][0]; k++) {
            int it = m->ind_nu[i][0][k];
            drdn[it] = drdn[it] + r / n[it];
        }
        if (m->reverse_off[i] == 0) {
            r = or_exp(vlogk[1]);
            for (int k = 0; k < m->num_nu[i][1]; k++) r = r * n[m->ind_nu[i][1][k]];
            Dr = Dr - r;
            drdn[ns] = drdn[ns] - r * dlnkdT[1];
            for (int k = 0; k < m->num_nu[i][1]; k++) {
                int it = m->ind_nu[i][1][k];
                drdn[it] = drdn[it] - r / n[it];
            }
        }
        if (m->exist_m[i]) {
            double t1 = Dr * dlnkdlnM / vM;
            for (int k = 0; k < ns; k++) drdn[k] = drdn[k] + t1;
            for (int k = 0; k < m->num_m[i]; k++) {
                int it = m->ind_m[i][k];
                drdn[it] = drdn[it] + t1 * m->men[i][k];
            }
        }
        for (int k = 0; k < m->num_nu[i][0]; k++) {
            int it = m->ind_nu[i][0][k];
            dndt[it] = dndt[it] - Dr;
            for (int j = 0; j < N; j++) J[it + N * j] = J[it + N * j] - drdn[j];
        }
        for (int k = 0; k < m->num_nu[i][1]; k++) {
            int it = m->ind_nu[i][1][k];
            dndt[it] = dndt[it] + Dr;
            for (int j = 0; j < N; j++) J[it + N * j] = J[it + N * j] + drdn[j];
        }
    }
    /* FJ.f90:680-697 */
    double de = 0.0, cv = 0.0, dcv = 0.0, t1 = 0.0, t2 = 0.0;
    for (int j = 0; j < ns; j++) {
        t1 = t1 + vert[j] * J[j + N * ns];
        de = de + vert[j] * dndt[j];
        dcv = dcv + vcvr[j] * dndt[j];
        cv = cv + vcvr[j] * n[j];
        t2 = t2 + vdcvrdT[j] * n[j];
    }
    double dTdt = -(de * T / cv);
    J[ns + N * ns] = -((T * t1 + dcv + dTdt * t2) / cv);
    for (int j = 0; j < ns; j++) {
        double s = 0.0;
        for (int k = 0; k < ns; k++) s = s + vert[k] * J[k + N * j];
        J[ns + N * j] = -((T * s + dTdt * vcvr[j]) / cv);
    }
}

/* rho_rel2abs: sub_chem.f90:1036-1085.  vrho in: relative amounts, out:
 * kg/m^3; p in Pa; E in J/kg. */
void or_chem_rho_rel2abs(const shcfd_mech *m, double p, double T, double *vrho, double *rho, double *E) {
    int ns = m->ns;
    double n[SHCFD_CHEM_MAX_NS];
    for (int j = 0; j < ns; j++) n[j] = vrho[j] * m->inv_mw[j];
    double tmp = 0.0;
    for (int j = 0; j < ns; j++) tmp = tmp + n[j];
    tmp = tmp * RU * T;
    tmp = p * 1e1 / tmp;
    for (int j = 0; j < ns; j++) n[j] = n[j] * tmp;
    double r = 0.0;
    for (int j = 0; j < ns; j++) r = r + n[j] * m->mw[j];
    r = r * 1e3;
    for (int j = 0; j < ns; j++) vrho[j] = n[j] * m->mw[j] * 1e3;
    int sec[SHCFD_CHEM_MAX_NS];
    double cmurt[7], chrt[7], ccpr[7];
    check_section_number(m, T, sec);
    set_coeff_more(T, or_log(T), cmurt, chrt, ccpr, NULL, NULL);
    double e = 0.0;
    for (int j = 0; j < ns; j++) {
        double dh = 0.0;
        for (int i = 0; i < 7; i++) dh = dh + m->coeff[j][sec[j]][i] * chrt[i];
        e = e + (dh - 1.0) * n[j];
    }
    *E = e * RU * T / r / 1e1;
    *rho = r;
}

/* calc_T: sub_chem.f90:1507-1645 (ns_tocalc = ns, the reactive model).
 * out[0] = kappa, out[1] = MWave, out[2] = mu, out[3..3+ns) = DHi,
 * out[3+ns..3+2ns) = vhi.  T is the Newton start on entry.  Returns the
 * Newton iteration count, or -1 past max_T_loop (1000). */
int or_chem_calc_t(const shcfd_mech *m, const double *vrho, double rho, double E, double *T_io, double *out) {
    int ns = m->ns;
    double n[SHCFD_CHEM_MAX_NS];
    double tmp = rho * 1e-11;
    for (int j = 0; j < ns; j++) n[j] = (vrho[j] > tmp) ? vrho[j] * m->inv_mw[j] * 1e-3 : 0.0;
    double Eobj = E * 1e1 * rho / RU;
    double T = *T_io;
    int sec[SHCFD_CHEM_MAX_NS];
    double chrt[6], ccpr[6], cvnow = 0.0;
    int k = 1;
    for (;;) {
        check_section_number(m, T, sec);
        double Ti1 = 1.0 / T, T2 = T * T, T3 = T2 * T, T4 = T3 * T; /* set_coeff_less, :20-49 */
        chrt[0] = 1.0; chrt[1] = T / 2.0; chrt[2] = T2 / 3.0; chrt[3] = T3 / 4.0; chrt[4] = T4 / 5.0; chrt[5] = Ti1;
        ccpr[0] = 1.0; ccpr[1] = T; ccpr[2] = T2; ccpr[3] = T3; ccpr[4] = T4; ccpr[5] = 0.0;
        double Enow = 0.0;
        cvnow = 0.0;
        for (int j = 0; j < ns; j++) {
            if (n[j] > 0.0) {
                double de = -1.0, dcv = -1.0;
                for (int i = 0; i < 6; i++) {
                    de = de + m->coeff[j][sec[j]][i] * chrt[i];
                    dcv = dcv + m->coeff[j][sec[j]][i] * ccpr[i];
                }
                Enow = Enow + de * n[j];
                cvnow = cvnow + dcv * n[j];
            }
        }
        double dT = (Eobj - Enow * T) / cvnow;
        if (fabs(dT) < 1e-8 || (fabs(dT) < 1e-4 && fabs(T - 1e3) < 1e-4) || k > 1000) break;
        T = fmax(100.0, T + dT);
        k = k + 1;
    }
    *T_io = T;
    if (k > 1000) return -1;
    double sn = 0.0;
    for (int j = 0; j < ns; j++) sn = sn + n[j];
    double kappa = 1.0 + sn / cvnow;
    out[0] = kappa;
    out[1] = rho / sn * 1e-3;
    tmp = RU * 1e-4 * T;
    for (int j = 0; j < ns; j++) {
        double dh = 0.0;
        for (int i = 0; i < 6; i++) dh = dh + m->coeff[j][sec[j]][i] * chrt[i];
        out[3 + j] = tmp * m->inv_mw[j] * (kappa - (kappa - 1.0) * dh);
        out[3 + ns + j] = tmp * m->inv_mw[j] * dh;
    }
    /* Wilke mixture viscosity, :1600-1633 */
    double logT = or_log(T), invT = 1.0 / T;
    double cmu[4] = {logT, invT, invT * invT, 1.0};
    double muN[SHCFD_CHEM_MAX_NS];
    for (int i = 0; i < m->nt; i++) {
        int s = -1; /* check_section_number_trans, :151-173 */
        if (m->trange[i][0][0] > T) s = 0;
        for (int l = 0; l < m->nsec_trans[i]; l++)
            if (m->trange[i][l][1] > T) {
                s = l;
                break;
            }
        if (s == -1) s = m->nsec_trans[i] - 1;
        if (n[m->tr2th[i]] > 0.0) {
            double t = 0.0;
            for (int j = 0; j < 4; j++) t = t + cmu[j] * m->trans[i][s][j];
            muN[i] = or_exp(t);
        } else {
            muN[i] = 0.0;
        }
    }
    double mu = 0.0;
    for (int i = 0; i < m->nt; i++) {
        if (muN[i] > 0.0) {
            int it = m->tr2th[i];
            double denom = 0.0;
            for (int j = 0; j < m->nt; j++) {
                if (muN[j] > 0.0) {
                    int jt = m->tr2th[j];
                    double x = 1.0 + sqrt(muN[i] / muN[j]) * sqrt(sqrt(m->mw[jt] / m->mw[it]));
                    double phi = 0.25 * (x * x) * sqrt(2.0 * m->mw[jt] / (m->mw[it] + m->mw[jt]));
                    denom = denom + n[jt] * phi;
                }
            }
            mu = mu + n[it] * muN[i] / denom;
        }
    }
    out[2] = mu * 1e-7;
    return k;
}

/* ---- DVODE ---------------------------------------------------------------
 * The reference's own stiff integrator (store/therm_lib/chemkin/core/
 * dvode.f), compiled from its source into oracle/_ref/libdvode_ref.so by
 * `make ref` and handed in by the loader (or_set_dvode).  Its state lives in
 * COMMON blocks, so calls are serialised.  Fex and Jex (the callbacks of
 * sub_chem.f90:1755-1757) are this file's or_chem_fex and or_chem_jex. */
typedef void (*or_dv_f)(const int *, const double *, const double *, double *, double *, int *);
typedef void (*or_dv_jac)(const int *, const double *, const double *, const int *, const int *, double *,
                          const int *, double *, int *);
typedef void (*or_dv_fn)(or_dv_f, const int *, double *, double *, const double *, const int *, const double *,
                         const double *, const int *, int *, const int *, double *, const int *, int *, const int *,
                         or_dv_jac, const int *, double *, int *);
static or_dv_fn g_dvode;
static pthread_mutex_t g_dvode_lock = PTHREAD_MUTEX_INITIALIZER;
static const shcfd_mech *g_dvode_mech;
static long g_dvode_calls;

void or_set_dvode(void *fn) { g_dvode = (or_dv_fn)fn; }
int or_have_dvode(void) { return g_dvode != NULL; }
long or_dvode_calls(void) { return g_dvode_calls; }

static void dv_fex(const int *neq, const double *t, const double *y, double *ydot, double *rpar, int *ipar) {
    or_chem_fex(g_dvode_mech, y, ydot);
}
static void dv_jex(const int *neq, const double *t, const double *y, const int *ml, const int *mu, double *pd,
                   const int *nrowpd, double *rpar, int *ipar) {
    or_chem_jex(g_dvode_mech, y, pd);
}

/* The DVODE loops of chemeq2 (sub_chem.f90:1751-1770: ATOL fixed) and of
 * reaction (:1281-1305: ATOL atol0 on the first call, atol_restart after a
 * restart): ITOL = 1, ITASK = 1, IOPT = 0, MF = 21, integrate y from *tt to
 * tout, restarting with ISTATE = 1 after ISTATE = -1 (500 steps) at most
 * max_recalc = 50 times.  Returns 0, SHCFD_E_CHEM (the reference stops or
 * reports failure), or SHCFD_E_STATE when the reference's DVODE is not
 * loaded.  *steps (may be NULL) accumulates IWORK(11) = NST. */
static int or_dvode_loop(const shcfd_mech *m, double *y, double *tt, double tout, double rtol, double atol0,
                         double atol_restart, int *steps) {
    if (!g_dvode) return SHCFD_E_STATE;
    enum { NMAX = SHCFD_CHEM_MAX_NS + 1 };
    int neq = m->ns + 1, itol = 1, itask = 1, iopt = 0, mf = 21, istate = 1, nrecalc = 0, e = 0;
    int lrw = 22 + 9 * neq + 2 * neq * neq, liw = 30 + neq; /* LRW, LIW: mod_chem.raw.f90:24-25 */
    double rwork[22 + 9 * NMAX + 2 * NMAX * NMAX], rpar[1], atol = atol0;
    int iwork[30 + NMAX], ipar[1];
    pthread_mutex_lock(&g_dvode_lock);
    g_dvode_mech = m;
    for (;;) {
        memset(iwork, 0, sizeof(iwork));
        g_dvode(dv_fex, &neq, y, tt, &tout, &itol, &rtol, &atol, &itask, &istate, &iopt, rwork, &lrw, iwork, &liw,
                dv_jex, &mf, rpar, ipar);
        if (steps) *steps += iwork[10];
        if (istate > 0) break;
        if (istate != -1) { /* "odd istate" */
            e = SHCFD_E_CHEM;
            break;
        }
        istate = 1;
        atol = atol_restart;
        nrecalc = nrecalc + 1;
        if (nrecalc > 50) { /* "Exceed max_recalc." */
            e = SHCFD_E_CHEM;
            break;
        }
    }
    g_dvode_calls++;
    pthread_mutex_unlock(&g_dvode_lock);
    return e;
}

/* reaction: sub_chem.f90:1236-1315, the per-cell DVODE integration over
 * [0, tout] (vrho kg/m^3 in/out); on failure vrho is untouched. */
int or_chem_reaction_dvode(const shcfd_mech *m, double T, double *vrho, double tout, int *steps) {
    int ns = m->ns;
    double n[SHCFD_CHEM_MAX_NS + 1], tt = 0.0;
    for (int j = 0; j < ns; j++) {
        n[j] = vrho[j] * m->inv_mw[j] * 1e-3;
        n[j] = fmax(n[j], N_EPS);
    }
    if (steps) *steps = 0;
    if (T < 0.0 || tout < 0.0) return SHCFD_E_ARG;
    n[ns] = T;
    int e = or_dvode_loop(m, n, &tt, tout, m->rtol, 0.0, m->atol, steps);
    if (e) return e;
    for (int j = 0; j < ns; j++) {
        n[j] = fmax(n[j], N_EPS);
        vrho[j] = n[j] * m->mw[j] * 1e3;
    }
    return 0;
}

/* chemeq2: sub_chem.f90:1714-1819.  y[0..ns-1] mol/cm^3, y[ns] = T,
 * integrated from 0 to tout.  Once more than 10 attempts have been made
 * over the interval, the rest of it, from the start of the current
 * substep, goes to DVODE (:1751-1770; SHCFD_E_STATE when the reference's
 * DVODE is not loaded).  The reference's "not converted at CHEMEQ2" stop
 * (:1810) returns SHCFD_E_CHEM. */
int or_chem_chemeq2(const shcfd_mech *m, double *y, double tout, double rtol, int *attempts) {
    int ns = m->ns, N = ns + 1;
    double y0[SHCFD_CHEM_MAX_NS + 1], yp[SHCFD_CHEM_MAX_NS + 1], p[SHCFD_CHEM_MAX_NS + 1], q[SHCFD_CHEM_MAX_NS + 1],
        p0[SHCFD_CHEM_MAX_NS + 1], q0[SHCFD_CHEM_MAX_NS + 1];
    double tt = 0.0, dt = tout - tt, ratio = 1e300;
    int countersum = 0;
    for (;;) {
        for (int i = 0; i < N; i++) y0[i] = y[i];
        double counter = 0.0;
        if (tout - tt < dt) dt = tout - tt;
        for (;;) {
            for (int i = 0; i < N; i++) y[i] = y0[i];
            countersum = countersum + 1;
            if (countersum > 10) {
                if (attempts) *attempts = countersum;
                return or_dvode_loop(m, y, &tt, tout, rtol, m->atol, m->atol, NULL); /* then dt = 0: done */
            }
            or_chem_fex_var(m, y, q0, p0);
            for (int i = 0; i < ns; i++) {
                p0[i] = -p0[i] / (y[i] + 1e-300);
                double var = p0[i] * dt, var2 = var * var, var3 = var * var2;
                double alp = (180.0 + 60.0 * var + 11.0 * var2 + var3) / (360.0 + 60.0 * var + 12.0 * var2 + var3);
                yp[i] = y0[i] + dt * (q0[i] - p0[i] * y0[i]) / (1.0 + alp * p0[i] * dt);
            }
            yp[ns] = y0[ns] + dt * (q0[ns] + p0[ns]);
            ratio = 1e300;
            or_chem_fex_var(m, yp, q, p);
            for (int i = 0; i < ns; i++) {
                p[i] = -p[i] / (yp[i] + 1e-300);
                p[i] = 0.5 * (p0[i] + p[i]);
                double var = p[i] * dt, var2 = var * var, var3 = var * var2;
                double alp = (180.0 + 60.0 * var + 11.0 * var2 + var3) / (360.0 + 60.0 * var + 12.0 * var2 + var3);
                q[i] = alp * q[i] + (1.0 - alp) * q0[i];
                y[i] = y0[i] + dt * (q[i] - p[i] * y0[i]) / (1.0 + alp * p[i] * dt);
                var = (y[i] * rtol + rtol * 1e-5) / fabs(y[i] - yp[i] + 1e-300);
                ratio = fmin(ratio, var);
            }
            y[ns] = y0[ns] + 0.5 * dt * (q[ns] + p[ns] + q0[ns] + p0[ns]);
            if (1.0 < ratio) break;
            counter = counter + 1.0;
            dt = dt * sqrt(ratio) / counter;
            if (counter > 10.0) {
                if (attempts) *attempts = countersum;
                return SHCFD_E_CHEM;
            }
        }
        tt = tt + dt;
        if (tout <= tt) break;
        dt = dt * sqrt(ratio);
    }
    if (attempts) *attempts = countersum;
    return 0;
}

/* reaction_chemeq2: sub_chem.f90:1684-1713.  vrho in kg/m^3 (in/out). */
int or_chem_reaction_chemeq2(const shcfd_mech *m, double T, double *vrho, double tout, int *attempts) {
    int ns = m->ns;
    double n[SHCFD_CHEM_MAX_NS + 1];
    for (int j = 0; j < ns; j++) {
        n[j] = vrho[j] * m->inv_mw[j] * 1e-3;
        n[j] = fmax(n[j], N_EPS);
    }
    if (T < 0.0 || tout < 0.0) return SHCFD_E_ARG;
    n[ns] = T;
    int e = or_chem_chemeq2(m, n, tout, m->rtol, attempts);
    if (e) return e;
    for (int j = 0; j < ns; j++) {
        n[j] = fmax(n[j], N_EPS);
        vrho[j] = n[j] * m->mw[j] * 1e3;
    }
    return 0;
}

/* LINPACK dgefa / dgesl (Gaussian elimination with partial pivoting,
 * column-oriented), the factor/solve pair DVODE uses for its Newton matrix
 * (store/therm_lib/chemkin/core/dvode.f:3816-3931, :3932-4040).  a is
 * column-major n x n; ipvt 0-based.  Returns 0, or k+1 when U(k,k) == 0. */
int or_dgefa(double *a, int n, int *ipvt) {
    int info = 0;
    for (int k = 0; k < n - 1; k++) {
        int l = k;
        double amax = fabs(a[k + n * k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(a[i + n * k]) > amax) {
                amax = fabs(a[i + n * k]);
                l = i;
            }
        ipvt[k] = l;
        if (a[l + n * k] == 0.0) {
            info = k + 1;
            continue;
        }
        if (l != k) {
            double t = a[l + n * k];
            a[l + n * k] = a[k + n * k];
            a[k + n * k] = t;
        }
        double t = -1.0 / a[k + n * k];
        for (int i = k + 1; i < n; i++) a[i + n * k] = t * a[i + n * k];
        for (int j = k + 1; j < n; j++) {
            double tj = a[l + n * j];
            if (l != k) {
                a[l + n * j] = a[k + n * j];
                a[k + n * j] = tj;
            }
            for (int i = k + 1; i < n; i++) a[i + n * j] = a[i + n * j] + tj * a[i + n * k];
        }
    }
    ipvt[n - 1] = n - 1;
    if (a[(n - 1) + n * (n - 1)] == 0.0) info = n;
    return info;
}

void or_dgesl(const double *a, int n, const int *ipvt, double *b) {
    for (int k = 0; k < n - 1; k++) {
        int l = ipvt[k];
        double t = b[l];
        if (l != k) {
            b[l] = b[k];
            b[k] = t;
        }
        for (int i = k + 1; i < n; i++) b[i] = b[i] + t * a[i + n * k];
    }
    for (int kb = 0; kb < n; kb++) {
        int k = n - 1 - kb;
        b[k] = b[k] / a[k + n * k];
        double t = -b[k];
        for (int i = 0; i < k; i++) b[i] = b[i] + t * a[i + n * k];
    }
}

/* Point-implicit source update of one cell over dt: the backward-Euler step
 * y1 = y0 + dt F(y1) (species and T at constant rho and e) solved by Newton
 * iterations y <- y - (I - dt J(y))^-1 (y - y0 - dt F(y)), each with the
 * analytic Jacobian (Jex) and the LINPACK factor/solve above -- DVODE's
 * first-order Newton matrix P = I - h l0 J with l0 = 1 (dvode.f DVJAC,
 * :2957-3140).  After each iteration concentrations are floored at n_eps
 * (as reaction_chemeq2 floors them); the iteration stops when every species
 * moved by less than PI_RTOL of itself plus PI_ATOL of the total
 * concentration, and T by less than
 * PI_RTOL*T, and fails after PI_MAXIT iterations.  The first iterate is the
 * classic linearised point-implicit update.  T is not written back (the
 * thermo recomputes it from e).  vrho kg/m^3 in/out; *iters = iterations. */
#define PI_MAXIT 32
#define PI_RTOL 1e-10
#define PI_ATOL 1e-14
int or_chem_point_implicit(const shcfd_mech *m, double T, double *vrho, double dt, int *iters, double *T_out) {
    int ns = m->ns, N = ns + 1;
    double y0[SHCFD_CHEM_MAX_NS + 1], y[SHCFD_CHEM_MAX_NS + 1], f[SHCFD_CHEM_MAX_NS + 1],
        P[(SHCFD_CHEM_MAX_NS + 1) * (SHCFD_CHEM_MAX_NS + 1)];
    int ipvt[SHCFD_CHEM_MAX_NS + 1];
    for (int j = 0; j < ns; j++) {
        y0[j] = vrho[j] * m->inv_mw[j] * 1e-3;
        y0[j] = fmax(y0[j], N_EPS);
        y[j] = y0[j];
    }
    y0[ns] = T;
    y[ns] = T;
    double atol = 0.0;
    for (int j = 0; j < ns; j++) atol = atol + y0[j];
    atol = atol * PI_ATOL;
    int it = 0, done = 0;
    while (!done) {
        if (it >= PI_MAXIT) {
            if (iters) *iters = it;
            return SHCFD_E_CHEM;
        }
        it = it + 1;
        or_chem_fex(m, y, f);
        or_chem_jex(m, y, P);
        for (int a = 0; a < N * N; a++) P[a] = -(dt * P[a]);
        for (int a = 0; a < N; a++) P[a + N * a] = P[a + N * a] + 1.0;
        for (int a = 0; a < N; a++) f[a] = (y0[a] - y[a]) + dt * f[a];
        if (or_dgefa(P, N, ipvt) != 0) {
            if (iters) *iters = it;
            return SHCFD_E_CHEM;
        }
        or_dgesl(P, N, ipvt, f);
        done = 1;
        for (int j = 0; j < ns; j++) {
            double yn = fmax(y[j] + f[j], N_EPS);
            if (fabs(yn - y[j]) > PI_RTOL * y[j] + atol) done = 0;
            y[j] = yn;
        }
        double Tn = y[ns] + f[ns];
        if (fabs(Tn - y[ns]) > PI_RTOL * y[ns]) done = 0;
        y[ns] = Tn;
        if (!(y[ns] > 0.0)) {
            if (iters) *iters = it;
            return SHCFD_E_CHEM;
        }
    }
    for (int j = 0; j < ns; j++) vrho[j] = y[j] * m->mw[j] * 1e3;
    if (iters) *iters = it;
    if (T_out) *T_out = y[ns];
    return 0;
}

/* ======================================================================== */
/* NASA-9 frozen mixing of two streams: cold_flow,                          */
/* store/therm_lib/NASA/core/sub_chem.f90:489-610, with check_section_number */
/* (func_chem.f90:41-62), calc_vThrt (:64-90), calc_vTcpr (:91-116),         */
/* calc_vTmu (:143-157) and check_section_number_trans (:173-192).           */
/* ======================================================================== */
#define N9_RU 8.3144621 /* NASA/core/mod_chem.raw.f90:6 */
#define N9_OMEGA 0.5
#define N9_EPS 1e-9

static int n9_sec(const shcfd_nasa9 *m, int s, double T) {
    int l;
    if (m->trange[s][0][0] > T) return 0;
    for (l = 0; l < m->nsec[s]; l++)
        if (m->trange[s][l][1] > T) return l;
    return m->nsec[s] - 1;
}

static int n9_sec_trans(const shcfd_nasa9 *m, int t, double T) {
    int l, num = -1;
    if (m->trange_trans[t][0][0] > T) num = 0;
    for (l = 0; l < m->nsec_trans[t]; l++)
        if (m->trange_trans[t][l][1] > T) { num = l; break; }
    return num < 0 ? m->nsec_trans[t] - 1 : num;
}

/* h/RT = sum co_k vh_k, cp/R = sum co_k vc_k; the powers of T by the
 * reference's running product from 1/T^2 */
static void n9_basis(double T, double logT, double *vh, double *vc) {
    double p[7];
    int k;
    p[0] = 1.0 / (T * T);
    for (k = 1; k < 7; k++) p[k] = p[k - 1] * T;
    vh[0] = -p[0];
    vh[1] = p[1] * logT;
    vh[2] = p[2];
    vh[3] = p[3] / 2.0;
    vh[4] = p[4] / 3.0;
    vh[5] = p[5] / 4.0;
    vh[6] = p[6] / 5.0;
    vh[7] = p[1];
    vh[8] = 0.0;
    for (k = 0; k < 7; k++) vc[k] = p[k];
    vc[7] = 0.0;
    vc[8] = 0.0;
}

static void n9_sums(const shcfd_nasa9 *m, const double *n, double T, double *ert, double *cvr) {
    double vh[9], vc[9];
    int i, j;
    n9_basis(T, or_log(T), vh, vc);
    *ert = 0.0;
    *cvr = 0.0;
    for (i = 0; i < m->ns; i++) {
        const double *co = m->co[i][n9_sec(m, i, T)];
        double a = -1.0, b = -1.0;
        for (j = 0; j < 9; j++) {
            a = a + co[j] * vh[j];
            b = b + co[j] * vc[j];
        }
        *ert = *ert + a * n[i];
        *cvr = *cvr + b * n[i];
    }
}

/* out = (MWave, kappa, mu, DHi(1:2), vhi(1:2)); returns the iterations */
int or_cold_flow(const shcfd_nasa9 *m, double y0, double y1, double E, double *Tio, double *out) {
    double n[SHCFD_NASA9_MAX_NS], mun[SHCFD_NASA9_MAX_NS], dhs[SHCFD_NASA9_MAX_NS], vhs[SHCFD_NASA9_MAX_NS];
    double vh[9], vc[9], vmu[4];
    double sn = 0.0, ert, cvr, T = *Tio, T_old, kappa, logT, rt, mu;
    int i, j, k = 0;
    for (i = 0; i < m->ns; i++) {
        n[i] = y0 * m->nf[i] + y1 * m->no[i];
        sn = sn + n[i];
    }
    out[0] = 1e3 / sn;
    n9_sums(m, n, T, &ert, &cvr);
    T_old = T * 0.5;
    while (fabs(E / N9_RU - ert * T) > cvr * N9_EPS * 1e-2 && fabs(T - T_old) > N9_EPS && k < 100) {
        T_old = T;
        T = T + N9_OMEGA * (E / N9_RU - ert * T) / cvr;
        n9_sums(m, n, T, &ert, &cvr);
        k = k + 1;
    }
    kappa = 1.0 + sn / cvr;
    out[1] = kappa;
    logT = or_log(T);
    n9_basis(T, logT, vh, vc);
    for (i = 0; i < m->ns; i++) {
        const double *co = m->co[i][n9_sec(m, i, T)];
        double a = 0.0;
        for (j = 0; j < 9; j++) a = a + co[j] * vh[j];
        dhs[i] = 1.0 * (kappa - (kappa - 1.0) * a);
        vhs[i] = 1.0 * a;
    }
    rt = N9_RU * T;
    {
        double d0 = 0.0, d1 = 0.0, h0 = 0.0, h1 = 0.0;
        for (i = 0; i < m->ns; i++) {
            d0 = d0 + m->nf[i] * dhs[i];
            d1 = d1 + m->no[i] * dhs[i];
            h0 = h0 + m->nf[i] * vhs[i];
            h1 = h1 + m->no[i] * vhs[i];
        }
        out[3] = d0 * rt;
        out[4] = d1 * rt;
        out[5] = h0 * rt;
        out[6] = h1 * rt;
    }
    vmu[2] = 1.0 / (T * T);
    vmu[0] = logT;
    vmu[1] = vmu[2] * T;
    vmu[3] = 1.0;
    for (i = 0; i < m->nt; i++) {
        const double *a = m->trans[i][n9_sec_trans(m, i, T)];
        double s = 0.0;
        for (j = 0; j < 4; j++) s = s + a[j] * vmu[j];
        mun[i] = or_exp(s);
    }
    mu = 0.0;
    for (i = 0; i < m->nt; i++) {
        double denom = 0.0, MWi = m->mw[m->tr2th[i]];
        for (j = 0; j < m->nt; j++) {
            double MWj = m->mw[m->tr2th[j]];
            double x = 1.0 + sqrt(mun[i] / mun[j]) * sqrt(sqrt(MWj / MWi));
            double phi = 0.25 * (x * x) * sqrt(2.0 * MWj / (MWi + MWj));
            denom = denom + n[m->tr2th[j]] * phi;
        }
        mu = mu + n[m->tr2th[i]] * mun[i] / denom;
    }
    out[2] = mu * 1e-7;
    *Tio = T;
    return k;
}
