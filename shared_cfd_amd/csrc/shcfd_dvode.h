// shcfd_dvode.h -- the stiff ODE integrator the reference's CHEMEQ2 hands
// difficult cells to: DVODE (P. N. Brown, G. D. Byrne, A. C. Hindmarsh,
// "VODE: a variable-coefficient ODE solver", SIAM J. Sci. Stat. Comput. 10,
// 1989), as vendored in store/therm_lib/chemkin/core/dvode.f and called by
// chemeq2 (store/therm_lib/chemkin/core/sub_chem.f90:1751-1770) with
// MF = 21 (BDF, Newton iteration with a user-supplied dense Jacobian,
// Jacobian saved for reuse), ITOL = 1 (scalar tolerances), ITASK = 1
// (integrate to TOUT and interpolate there), IOPT = 0 (defaults: order <= 5,
// 500 steps per call, no step bounds).
//
// Restated for one system of N equations per thread, usable on the host
// and the device.  Only the paths that configuration reaches are kept; each
// routine names the dvode.f lines it follows, and the operations are in the
// same order.  Arrays are 0-based: yh[j][i] is YH(i+1, j+1), the Nordsieck
// history of order j scaled by h^j / j!.  The state of the reference's
// COMMON blocks DVOD01/DVOD02 is a member of one Dvode object, so every
// thread integrates its own system.
//
// The system type provides f(t, y, ydot) (Fex's contract) and
// jac(t, y, pd), pd column-major N x N with pd[i + N j] = d f_i / d y_j
// (Jex's contract; every entry is written).  pow is evaluated as
// exp(log(x) r) with the chemistry's own exp and log (shcfd_chem.h), so host
// and device builds of this code agree bit for bit.
#pragma once
#include "shcfd_chem.h"

namespace shcfd {
namespace dvode {

// DUMACH (dvode.f:3771-3808): the unit roundoff, 2^-52
constexpr double UROUND = 2.220446049250313080847e-16;
constexpr int MAXORD = 5;  // MORD(2), BDF
constexpr int LMAX = MAXORD + 1;
constexpr int MXSTEP = 500;  // MXSTP0

SHCFD_HD double dabs(double x) { return x < 0.0 ? -x : x; }
SHCFD_HD double dmax2(double a, double b) { return a > b ? a : b; }
SHCFD_HD double dmin2(double a, double b) { return a < b ? a : b; }
SHCFD_HD double sign_of(double a, double b) { return b >= 0.0 ? dabs(a) : -dabs(a); }
// x ** r for x > 0 (r = 1/L of the step and order selection), with the
// chemistry's exp and log.  The reference's ** is its runtime's pow; a host
// build that substitutes that pow (SHCFD_DVODE_POW) takes the reference's
// steps bit for bit.
#ifndef SHCFD_DVODE_POW
#define SHCFD_DVODE_POW(x, r) chem::dexp(chem::dlog(x) * (r))
#endif
SHCFD_HD double dpow(double x, double r) { return SHCFD_DVODE_POW(x, r); }
// x ** k for an integer k >= 0, by binary powering (the real ** integer
// of the Fortran runtime: the factors x^(2^b) of the set bits of k, in
// increasing b, multiplied into 1)
SHCFD_HD double ipow(double x, int k) {
    double r = 1.0;
    for (;;) {
        if (k & 1) r = r * x;
        k >>= 1;
        if (k == 0) break;
        x = x * x;
    }
    return r;
}

template <int N>
struct Dvode {
    // YH, EWT (holds 1/EWT after set-up, as RWORK(LEWT) does), SAVF, ACOR,
    // the work vector Y, the iteration matrix WM (column-major) with its
    // pivots, and the saved Jacobian of JSV = 1 (WM(LOCJS))
    double yh[LMAX][N], ewt[N], savf[N], acor[N], y[N], wm[N * N], wjs[N * N];
    int ipvt[N];
    double el[13], tau[13], tq[5];
    double acnrm, ccmxj, conp, crate, drc, eta, etamax, h, hmin, hmxi, hnew, hscal, prl1, rc, rl1, tn, hu;
    double etaq, etaqm1;  // SAVEd locals of DVSTEP
    int icf, ipup, jcur, jstart, kflag, kuth, l, newh, newq, nq, nqwait, nslj, nslp, nst;

    // DVNORM (dvode.f:3467-3500): weighted root-mean-square norm
    SHCFD_HD static double vnorm(const double* v, const double* w) {
        double sum = 0.0;
        for (int i = 0; i < N; i++) {
            const double t = v[i] * w[i];
            sum = sum + t * t;
        }
        return chem::dsqrt(sum / (double)N);
    }

    // DEWSET (ITOL = 1, dvode.f:3419-3445), then the inversion of the
    // caller (dvode.f driver); false when a weight is not positive
    SHCFD_HD bool ewset(double rtol, double atol) {
        for (int i = 0; i < N; i++) ewt[i] = rtol * dabs(yh[0][i]) + atol;
        for (int i = 0; i < N; i++) {
            if (ewt[i] <= 0.0) return false;
            ewt[i] = 1.0 / ewt[i];
        }
        return true;
    }

    // DGEFA (dvode.f:3816-3931): LU with partial pivoting, column-major
    SHCFD_HD int gefa() {
        int info = 0;
        for (int k = 0; k < N - 1; k++) {
            // IDAMAX over rows k.. of column k: the first largest |a|
            int lp = k;
            double dmax = dabs(wm[k + N * k]);
            for (int i = k + 1; i < N; i++) {
                const double a = dabs(wm[i + N * k]);
                if (a > dmax) {
                    dmax = a;
                    lp = i;
                }
            }
            ipvt[k] = lp;
            if (wm[lp + N * k] == 0.0) {
                info = k + 1;
                continue;
            }
            if (lp != k) {
                const double t = wm[lp + N * k];
                wm[lp + N * k] = wm[k + N * k];
                wm[k + N * k] = t;
            }
            const double t = -1.0 / wm[k + N * k];
            for (int i = k + 1; i < N; i++) wm[i + N * k] = t * wm[i + N * k];
            for (int j = k + 1; j < N; j++) {
                const double tj = wm[lp + N * j];
                if (lp != k) {
                    wm[lp + N * j] = wm[k + N * j];
                    wm[k + N * j] = tj;
                }
                if (tj != 0.0)
                    for (int i = k + 1; i < N; i++) wm[i + N * j] = wm[i + N * j] + tj * wm[i + N * k];
            }
        }
        ipvt[N - 1] = N - 1;
        if (wm[(N - 1) + N * (N - 1)] == 0.0) info = N;
        return info;
    }

    // DGESL with JOB = 0 (dvode.f:3932-4040): solve with the factors
    SHCFD_HD void gesl(double* b) const {
        for (int k = 0; k < N - 1; k++) {
            const int lp = ipvt[k];
            const double t = b[lp];
            if (lp != k) {
                b[lp] = b[k];
                b[k] = t;
            }
            if (t != 0.0)
                for (int i = k + 1; i < N; i++) b[i] = b[i] + t * wm[i + N * k];
        }
        for (int kb = 0; kb < N; kb++) {
            const int k = N - 1 - kb;
            b[k] = b[k] / wm[k + N * k];
            const double t = -b[k];
            if (t != 0.0)
                for (int i = 0; i < k; i++) b[i] = b[i] + t * wm[i + N * k];
        }
    }

    // DVSET, METH = 2 (dvode.f:2369-2566, labels 200-300): the BDF
    // coefficients EL and the error-test constants TQ for order NQ
    SHCFD_HD void set() {
        const double flotl = (double)l;
        for (int i = 2; i < l; i++) el[i] = 0.0;
        el[0] = 1.0;
        el[1] = 1.0;
        double alph0 = -1.0, ahatn0 = -1.0, hsum = h, rxi = 1.0, rxis = 1.0;
        if (nq != 1) {
            for (int j = 1; j <= nq - 2; j++) {
                hsum = hsum + tau[j - 1];
                rxi = h / hsum;
                const int jp1 = j + 1;
                alph0 = alph0 - 1.0 / (double)jp1;
                for (int iback = 1; iback <= jp1; iback++) {
                    const int i = (j + 3) - iback;
                    el[i - 1] = el[i - 1] + el[i - 2] * rxi;
                }
            }
            alph0 = alph0 - 1.0 / (double)nq;
            rxis = -el[1] - alph0;
            hsum = hsum + tau[nq - 2];
            rxi = h / hsum;
            ahatn0 = -el[1] - rxi;
            for (int iback = 1; iback <= nq; iback++) {
                const int i = (nq + 2) - iback;
                el[i - 1] = el[i - 1] + el[i - 2] * rxis;
            }
        }
        const double t1 = 1.0 - ahatn0 + alph0;
        const double t2 = 1.0 + (double)nq * t1;
        tq[1] = dabs(alph0 * t2 / t1);
        tq[4] = dabs(t2 / (el[l - 1] * rxi / rxis));
        if (nqwait == 1) {
            const double cnqm1 = rxis / el[l - 1];
            const double t3 = alph0 + 1.0 / (double)nq;
            const double t4 = ahatn0 + rxi;
            double elp = t3 / (1.0 - t4 + t3);
            tq[0] = dabs(elp / cnqm1);
            hsum = hsum + tau[nq - 1];
            rxi = h / hsum;
            const double t5 = alph0 - 1.0 / (double)(nq + 1);
            const double t6 = ahatn0 - rxi;
            elp = t2 / (1.0 - t6 + t5);
            tq[2] = dabs(elp * rxi * (flotl + 1.0) * t5);
        }
        tq[3] = 0.1 * tq[1];  // CORTES
    }

    // DVJUST, METH = 2 (dvode.f:2568-2726, labels 200-370): adjust the
    // history array for an order change by iord = +1 or -1
    SHCFD_HD void just(int iord) {
        if (nq == 2 && iord != 1) return;
        if (iord != 1) {
            for (int j = 0; j < LMAX; j++) el[j] = 0.0;
            el[2] = 1.0;
            double hsum = 0.0;
            for (int j = 1; j <= nq - 2; j++) {
                hsum = hsum + tau[j - 1];
                const double xi = hsum / hscal;
                const int jp1 = j + 1;
                for (int iback = 1; iback <= jp1; iback++) {
                    const int i = (j + 4) - iback;
                    el[i - 1] = el[i - 1] * xi + el[i - 2];
                }
            }
            for (int j = 3; j <= nq; j++)
                for (int i = 0; i < N; i++) yh[j - 1][i] = yh[j - 1][i] - yh[l - 1][i] * el[j - 1];
            return;
        }
        for (int j = 0; j < LMAX; j++) el[j] = 0.0;
        el[2] = 1.0;
        double alph0 = -1.0, alph1 = 1.0, prod = 1.0, xiold = 1.0, hsum = hscal;
        if (nq != 1) {
            for (int j = 1; j <= nq - 1; j++) {
                const int jp1 = j + 1;
                hsum = hsum + tau[jp1 - 1];
                const double xi = hsum / hscal;
                prod = prod * xi;
                alph0 = alph0 - 1.0 / (double)jp1;
                alph1 = alph1 + 1.0 / xi;
                for (int iback = 1; iback <= jp1; iback++) {
                    const int i = (j + 4) - iback;
                    el[i - 1] = el[i - 1] * xiold + el[i - 2];
                }
                xiold = xi;
            }
        }
        const double t1 = (-alph0 - alph1) / prod;
        const int lp1 = l + 1;
        for (int i = 0; i < N; i++) yh[lp1 - 1][i] = t1 * yh[LMAX - 1][i];
        for (int j = 3; j <= nq + 1; j++) {
            const double e = el[j - 1];
            if (e != 0.0)
                for (int i = 0; i < N; i++) yh[j - 1][i] = yh[j - 1][i] + e * yh[lp1 - 1][i];
        }
    }

    // DVJAC, MITER = 1 with JSV = 1 (dvode.f:2957-3105): a new or the saved
    // Jacobian, P = I - h rl1 J, and its LU; returns IERPJ
    template <typename Sys>
    SHCFD_HD int jac(Sys& sys) {
        const double hrl1 = h * rl1;
        int jok = 1;
        if (nst == 0 || nst > nslj + 50) jok = -1;  // MSBJ
        if (icf == 1 && drc < ccmxj) jok = -1;
        if (icf == 2) jok = -1;
        if (jok == -1) {
            nslj = nst;
            jcur = 1;
            for (int i = 0; i < N * N; i++) wm[i] = 0.0;
            sys.jac(tn, y, wm);
            for (int i = 0; i < N * N; i++) wjs[i] = wm[i];
        } else {
            jcur = 0;
            for (int i = 0; i < N * N; i++) wm[i] = wjs[i];
        }
        const double con = -hrl1;
        for (int i = 0; i < N * N; i++) wm[i] = con * wm[i];
        for (int i = 0; i < N; i++) wm[i + N * i] = wm[i + N * i] + 1.0;
        return gefa() != 0 ? 1 : 0;
    }

    // DVNLSD, MITER = 1 (dvode.f:2728-2955): the corrector's modified
    // Newton iteration; nflag in: 0 or -1 / -2 after a failure, out: 0
    // (converged) or -1
    template <typename Sys>
    SHCFD_HD int nlsd(Sys& sys, int nflag) {
        if (jstart == 0) nslp = 0;
        if (nflag == 0) icf = 0;
        if (nflag == -2) ipup = 1;
        if (jstart == 0 || jstart == -1) ipup = 1;
        drc = dabs(rc - 1.0);
        if (drc > 0.3 || nst >= nslp + 20) ipup = 1;  // CCMAX, MSBP
        for (;;) {  // label 220
            int m = 0;
            double delp = 0.0, del = 0.0;
            for (int i = 0; i < N; i++) y[i] = yh[0][i];
            sys.f(tn, y, savf);
            if (ipup > 0) {
                const int ierpj = jac(sys);
                ipup = 0;
                rc = 1.0;
                drc = 0.0;
                crate = 1.0;
                nslp = nst;
                if (ierpj != 0) {  // label 430
                    icf = 2;
                    ipup = 1;
                    return -1;
                }
            }
            for (int i = 0; i < N; i++) acor[i] = 0.0;
            bool retry = false;
            for (;;) {  // label 270 / 350
                for (int i = 0; i < N; i++) y[i] = (rl1 * h) * savf[i] - (rl1 * yh[1][i] + acor[i]);
                gesl(y);
                if (rc != 1.0) {
                    const double cscale = 2.0 / (1.0 + rc);
                    for (int i = 0; i < N; i++) y[i] = cscale * y[i];
                }
                del = vnorm(y, ewt);
                for (int i = 0; i < N; i++) acor[i] = acor[i] + y[i];
                for (int i = 0; i < N; i++) y[i] = yh[0][i] + acor[i];
                // label 400
                if (m != 0) crate = dmax2(0.3 * crate, del / delp);
                const double dcon = del * dmin2(1.0, crate) / tq[3];
                if (dcon <= 1.0) {  // label 450
                    jcur = 0;
                    icf = 0;
                    acnrm = (m == 0) ? del : vnorm(acor, ewt);
                    return 0;
                }
                m = m + 1;
                if (m == 3 || (m >= 2 && del > 2.0 * delp)) {  // MAXCOR, RDIV: label 410
                    if (jcur == 1) {  // label 430
                        icf = 2;
                        ipup = 1;
                        return -1;
                    }
                    icf = 1;
                    ipup = 1;
                    retry = true;
                    break;
                }
                delp = del;
                sys.f(tn, y, savf);
            }
            if (!retry) break;
        }
        return -1;
    }

    // the Pascal-triangle prediction of the history array (dvode.f DVSTEP
    // label 200) and its retraction (labels 420-430, 510-520)
    SHCFD_HD void predict() {
        for (int jb = 1; jb <= nq; jb++)
            for (int j = nq - jb + 1; j <= nq; j++)
                for (int i = 0; i < N; i++) yh[j - 1][i] = yh[j - 1][i] + yh[j][i];
    }
    SHCFD_HD void retract() {
        for (int jb = 1; jb <= nq; jb++)
            for (int j = nq - jb + 1; j <= nq; j++)
                for (int i = 0; i < N; i++) yh[j - 1][i] = yh[j - 1][i] - yh[j][i];
    }
    // rescale the history by eta (label 150)
    SHCFD_HD void rescale() {
        double r = 1.0;
        for (int j = 2; j <= l; j++) {
            r = r * eta;
            for (int i = 0; i < N; i++) yh[j - 1][i] = r * yh[j - 1][i];
        }
        h = hscal * eta;
        hscal = h;
        rc = rc * eta;
    }

    // DVSTEP (dvode.f:1934-2367): one step, with order and step size
    // selection; sets kflag (0, -1 error test failed repeatedly, -2
    // corrector failed repeatedly)
    template <typename Sys>
    SHCFD_HD void step(Sys& sys) {
        kflag = 0;
        const double told = tn;
        int ncf = 0;
        jcur = 0;
        int nflag = 0;
        const double hmin_ = hmin;
        int go;  // 150: rescale then 200; 200: advance
        if (jstart > 0) {
            if (kuth == 1) {
                eta = dmin2(eta, h / hscal);
                newh = 1;
            }
            go = 200;
            if (newh != 0) {  // label 50
                if (newq != nq) {
                    if (newq < nq) {
                        just(-1);
                        nq = newq;
                        l = nq + 1;
                        nqwait = l;
                    } else {
                        just(1);
                        nq = newq;
                        l = nq + 1;
                        nqwait = l;
                    }
                }
                go = 150;
            }
        } else {
            // jstart == 0: the first call
            nq = 1;
            l = 2;
            tau[0] = h;
            prl1 = 1.0;
            rc = 0.0;
            etamax = 1.0e4;  // ETAMX1
            nqwait = 2;
            hscal = h;
            go = 200;
        }
        for (;;) {
            if (go == 150) rescale();
            // label 200
            tn = tn + h;
            predict();
            set();
            rl1 = 1.0 / el[1];
            rc = rc * (rl1 / prl1);
            prl1 = rl1;
            nflag = nlsd(sys, nflag);
            if (nflag != 0) {
                ncf = ncf + 1;
                etamax = 1.0;
                tn = told;
                retract();
                if (dabs(h) <= hmin_ * 1.00001 || ncf == 10) {  // ONEPSM, MXNCF: label 670
                    kflag = -2;
                    break;
                }
                eta = 0.25;  // ETACF
                eta = dmax2(eta, hmin_ / dabs(h));
                nflag = -1;
                go = 150;
                continue;
            }
            // label 450
            const double dsm = acnrm / tq[1];
            if (dsm <= 1.0) {
                // the step is accepted
                kflag = 0;
                nst = nst + 1;
                hu = h;
                for (int iback = 1; iback <= nq; iback++) {
                    const int i = l - iback;
                    tau[i] = tau[i - 1];
                }
                tau[0] = h;
                for (int j = 0; j < l; j++) {
                    const double e = el[j];
                    if (e != 0.0)
                        for (int i = 0; i < N; i++) yh[j][i] = yh[j][i] + e * acor[i];
                }
                nqwait = nqwait - 1;
                if (!(l == LMAX || nqwait != 1)) {
                    for (int i = 0; i < N; i++) yh[LMAX - 1][i] = acor[i];
                    conp = tq[4];
                }
                if (etamax != 1.0) {
                    choose_order_and_step(dsm);
                } else {
                    // label 490
                    if (nqwait < 2) nqwait = 2;
                    newq = nq;
                    newh = 0;
                    eta = 1.0;
                    hnew = h;
                }
                // label 690, 700
                etamax = 10.0;  // ETAMX3, and ETAMX2 while NST <= 10: both 10
                const double r = 1.0 / tq[1];
                for (int i = 0; i < N; i++) acor[i] = r * acor[i];
                break;
            }
            // label 500: the error test failed
            kflag = kflag - 1;
            nflag = -2;
            tn = told;
            retract();
            if (dabs(h) <= hmin_ * 1.00001) {  // label 660
                kflag = -1;
                break;
            }
            etamax = 1.0;
            if (kflag > -3) {  // KFC
                const double flotl = (double)l;
                eta = 1.0 / (dpow(6.0 * dsm, 1.0 / flotl) + 1.0e-6);  // BIAS2, ADDON
                eta = dmax2(dmax2(eta, hmin_ / dabs(h)), 0.1);           // ETAMIN
                if (kflag <= -2 && eta > 0.2) eta = 0.2;                 // ETAMXF
                go = 150;
                continue;
            }
            // label 530
            if (kflag == -7) {  // KFH
                kflag = -1;
                break;
            }
            if (nq != 1) {
                eta = dmax2(0.1, hmin_ / dabs(h));
                just(-1);
                l = nq;
                nq = nq - 1;
                nqwait = l;
                go = 150;
                continue;
            }
            // label 540: restart at order 1 from a fresh derivative
            eta = dmax2(0.1, hmin_ / dabs(h));
            h = h * eta;
            hscal = h;
            tau[0] = h;
            sys.f(tn, y, savf);
            for (int i = 0; i < N; i++) yh[1][i] = h * savf[i];
            nqwait = 10;
            go = 200;
        }
        jstart = 1;  // label 720
    }

    // DVSTEP labels 560-640: after a successful step with NQWAIT reaching
    // zero, the step ratios of orders NQ-1, NQ and NQ+1 and the choice
    SHCFD_HD void choose_order_and_step(double dsm) {
        const double flotl = (double)l;
        etaq = 1.0 / (dpow(6.0 * dsm, 1.0 / flotl) + 1.0e-6);  // BIAS2
        if (nqwait != 0) {
            eta = etaq;  // label 600
            newq = nq;
        } else {
            nqwait = 2;
            etaqm1 = 0.0;
            if (nq != 1) {
                const double ddn = vnorm(yh[l - 1], ewt) / tq[0];
                etaqm1 = 1.0 / (dpow(6.0 * ddn, 1.0 / (flotl - 1.0)) + 1.0e-6);  // BIAS1
            }
            double etaqp1 = 0.0;
            if (l != LMAX) {
                const double cnquot = (tq[4] / conp) * ipow(h / tau[1], l);
                for (int i = 0; i < N; i++) savf[i] = acor[i] - cnquot * yh[LMAX - 1][i];
                const double dup = vnorm(savf, ewt) / tq[2];
                etaqp1 = 1.0 / (dpow(10.0 * dup, 1.0 / (flotl + 1.0)) + 1.0e-6);  // BIAS3
            }
            int pick;  // 0: NQ, -1: NQ-1, +1: NQ+1
            if (etaq >= etaqp1)
                pick = (etaq < etaqm1) ? -1 : 0;
            else
                pick = (etaqp1 > etaqm1) ? 1 : -1;
            if (pick == 0) {
                eta = etaq;
                newq = nq;
            } else if (pick == -1) {
                eta = etaqm1;
                newq = nq - 1;
            } else {
                eta = etaqp1;
                newq = nq + 1;
                for (int i = 0; i < N; i++) yh[LMAX - 1][i] = acor[i];
            }
        }
        // label 630
        if (eta < 1.5 || etamax == 1.0) {  // THRESH: label 640
            newq = nq;
            newh = 0;
            eta = 1.0;
            hnew = h;
            return;
        }
        eta = dmin2(eta, etamax);
        eta = eta / dmax2(1.0, dabs(h) * hmxi * eta);
        newh = 1;
        hnew = h * eta;
    }

    // DVHIN (dvode.f:1673-1809): the initial step size; false when TOUT is
    // too close to T (IER = -1)
    template <typename Sys>
    SHCFD_HD bool hin(Sys& sys, double t0, double tout, double atol, double& h0) {
        const double tdist = dabs(tout - t0);
        const double tround = UROUND * dmax2(dabs(t0), dabs(tout));
        if (tdist < 2.0 * tround) return false;
        const double hlb = 100.0 * tround;
        double hub = 0.1 * tdist;
        for (int i = 0; i < N; i++) {
            const double delyi = 0.1 * dabs(yh[0][i]) + atol;
            const double afi = dabs(yh[1][i]);
            if (afi * hub > delyi) hub = delyi / afi;
        }
        int iter = 0;
        double hg = chem::dsqrt(hlb * hub);
        if (hub < hlb) {
            h0 = sign_of(hg, tout - t0);
            return true;
        }
        double hnew_;
        for (;;) {  // label 50
            const double hh = sign_of(hg, tout - t0);
            const double t1 = t0 + hh;
            for (int i = 0; i < N; i++) y[i] = yh[0][i] + hh * yh[1][i];
            sys.f(t1, y, acor);
            for (int i = 0; i < N; i++) acor[i] = (acor[i] - yh[1][i]) / hh;
            const double yddnrm = vnorm(acor, ewt);
            if (yddnrm * hub * hub > 2.0)
                hnew_ = chem::dsqrt(2.0 / yddnrm);
            else
                hnew_ = chem::dsqrt(hg * hub);
            iter = iter + 1;
            if (iter >= 4) break;
            const double hrat = hnew_ / hg;
            if (hrat > 0.5 && hrat < 2.0) break;
            if (iter >= 2 && hnew_ > 2.0 * hg) {
                hnew_ = hg;
                break;
            }
            hg = hnew_;
        }
        h0 = hnew_ * 0.5;  // label 80
        if (h0 < hlb) h0 = hlb;
        if (h0 > hub) h0 = hub;
        h0 = sign_of(h0, tout - t0);
        return true;
    }

    // DVINDY with K = 0 (dvode.f:1811-1932): the interpolated solution at t
    SHCFD_HD void indy(double t, double* dky) const {
        const double s = (t - tn) / h;
        for (int i = 0; i < N; i++) dky[i] = 1.0 * yh[l - 1][i];
        for (int jb = 1; jb <= nq; jb++) {
            const int j = nq - jb;
            for (int i = 0; i < N; i++) dky[i] = 1.0 * yh[j][i] + s * dky[i];
        }
    }

    // DVODE with ISTATE = 1 on entry (dvode.f driver, labels 10-420,
    // 500-580): integrate yv from t to tout.  Returns ISTATE: 2 (t =
    // tout), -1 (MXSTEP steps taken; yv, t at the last step), -2 (too much
    // accuracy requested), -4 (error test failed repeatedly), -5
    // (corrector failed repeatedly), -6 (a weight became zero), -3
    // (illegal input: tout too close to t, or a nonpositive weight at start)
    template <typename Sys>
    SHCFD_HD int run(Sys& sys, double* yv, double& t, double tout, double rtol, double atol) {
        nst = 0;
        if (tout == t) return 2;
        if (start(sys, yv, t, tout, rtol, atol)) return -3;
        int nslast = 0;
        bool first = true;
        for (;;) {
            if (!first) {
                // label 250
                if (nst - nslast >= MXSTEP) {
                    for (int i = 0; i < N; i++) yv[i] = yh[0][i];
                    t = tn;
                    return -1;
                }
                if (!ewset(rtol, atol)) {
                    for (int i = 0; i < N; i++) yv[i] = yh[0][i];
                    t = tn;
                    return -6;
                }
            }
            first = false;
            // label 270
            const double tolsf = UROUND * vnorm(yh[0], ewt);
            if (tolsf > 1.0) {
                for (int i = 0; i < N; i++) yv[i] = yh[0][i];
                t = tn;
                return nst == 0 ? -3 : -2;
            }
            step(sys);
            if (kflag != 0) {
                for (int i = 0; i < N; i++) yv[i] = yh[0][i];
                t = tn;
                return kflag == -1 ? -4 : -5;
            }
            // label 300, 310 (ITASK = 1)
            kuth = 0;
            if ((tn - tout) * h < 0.0) continue;
            indy(tout, yv);
            t = tout;
            return 2;
        }
    }

    // the set-up of a DVODE call with ISTATE = 1 (dvode.f driver, labels
    // 100-180): the history array from y and f(t, y), the inverted weights
    // and the initial step; nonzero when a weight is not positive or tout
    // is too close to t (ISTATE = -3)
    template <typename Sys>
    SHCFD_HD int start(Sys& sys, const double* yv, double t, double tout, double rtol, double atol) {
        nst = 0;
        hmin = 0.0;
        hmxi = 0.0;
        tn = t;
        jstart = 0;
        ccmxj = 0.2;
        nslj = 0;
        hu = 0.0;
        icf = 0;
        ipup = 0;
        kuth = 0;
        newq = 1;
        newh = 0;
        etaq = etaqm1 = 0.0;
        sys.f(t, yv, yh[1]);
        for (int i = 0; i < N; i++) yh[0][i] = yv[i];
        nq = 1;
        h = 1.0;
        if (!ewset(rtol, atol)) return -3;
        double h0 = 0.0;
        if (!hin(sys, t, tout, atol, h0)) return -3;
        h = h0;  // HMXI = 0: no bound on h0 (label 180)
        for (int i = 0; i < N; i++) yh[1][i] = h0 * yh[1][i];
        return 0;
    }
};

}  // namespace dvode

namespace chem {

// a column-major N x N matrix, element (a, b) at p[a + N b]: jex_impl
// writes DVODE's Jacobian array through it
template <int N>
struct ColMat {
    double* p;
    SHCFD_HD double get(int a, int b) const { return p[a + N * b]; }
    SHCFD_HD void set(int a, int b, double v) { p[a + N * b] = v; }
    template <int NR>
    SHCFD_HD void sub_row(int it, const double (&v)[N]) {
        for (int b = 0; b < N; b++) p[it + N * b] = p[it + N * b] - v[b];
    }
    template <int NR>
    SHCFD_HD void add_row(int it, const double (&v)[N]) {
        for (int b = 0; b < N; b++) p[it + N * b] = p[it + N * b] + v[b];
    }
};

// the system DVODE integrates for the chemistry: y = (n(ns) mol/cm^3, T)
// at constant density and internal energy, f = Fex (FJ.f90:152-300),
// Jacobian = Jex (FJ.f90:454-708)
template <int NS>
struct ChemSys {
    const shcfd_mech* m;
    SHCFD_HD void f(double, const double* y, double* ydot) const {
        fex<NS>(m, *reinterpret_cast<const double(*)[NS + 1]>(y), *reinterpret_cast<double(*)[NS + 1]>(ydot));
    }
    SHCFD_HD void jac(double, const double* y, double* pd) const {
        ColMat<NS + 1> J{pd};
        double f[NS + 1];
        jex_impl<NS>(m, *reinterpret_cast<const double(*)[NS + 1]>(y), J, f);
    }
};

constexpr int MAX_RECALC = 50;  // mod_chem.raw.f90:27

// chemeq2's hand-over to DVODE (sub_chem.f90:1751-1770): from state y at
// time tt to tout with MF = 21, restarting (ISTATE = 1) from where a call
// stopped after its 500 steps, at most MAX_RECALC times.  The reference
// stops the run on any other failure; here that cell returns SHCFD_E_CHEM.
// steps accumulates the DVODE steps taken.
template <int NS>
SHCFD_HD int chemeq2_dvode(const shcfd_mech* m, double (&y)[NS + 1], double tt, double tout, double rtol,
                           double atol, int& steps) {
    dvode::Dvode<NS + 1> dv;
    ChemSys<NS> sys{m};
    int nrecalc = 0;
    for (;;) {
        const int istate = dv.run(sys, y, tt, tout, rtol, atol);
        steps = steps + dv.nst;
        if (istate > 0) return 0;
        if (istate != -1) return SHCFD_E_CHEM;  // "odd istate"
        nrecalc = nrecalc + 1;
        if (nrecalc > MAX_RECALC) return SHCFD_E_CHEM;  // "Exceed max_recalc."
    }
}

// reaction_chemeq2 with its DVODE hand-over done in place
// (sub_chem.f90:1684-1819)
template <int NS>
SHCFD_HD int reaction_chemeq2_full(const shcfd_mech* m, double T, double (&vrho)[NS], double tout, int& attempts) {
    double n[NS + 1], tt;
    int e = reaction_chemeq2<NS>(m, T, vrho, tout, attempts, n, tt);
    if (e != CHEM_DEFER) return e;
    int steps = 0;
    e = chemeq2_dvode<NS>(m, n, tt, tout, m->rtol, m->atol, steps);
    if (e) return e;
    n_to_vrho<NS>(m, n, vrho);
    return 0;
}

// reaction: sub_chem.f90:1236-1315, the per-cell DVODE integration over
// [0, tout] (vrho kg/m^3 in/out): ATOL = 0 on the first call and the user's
// ATOL after a restart.  On failure vrho is untouched and SHCFD_E_CHEM is
// returned (the reference returns flag = .false.).  steps receives the
// DVODE steps taken.
template <int NS>
SHCFD_HD int reaction_dvode(const shcfd_mech* m, double T, double (&vrho)[NS], double tout, int& steps) {
    double n[NS + 1];
    vrho_to_n<NS>(m, vrho, n);
    steps = 0;
    if (T < 0.0 || tout < 0.0) return SHCFD_E_ARG;
    n[NS] = T;
    double tt = 0.0, atol = 0.0;
    dvode::Dvode<NS + 1> dv;
    ChemSys<NS> sys{m};
    int nrecalc = 0;
    for (;;) {
        const int istate = dv.run(sys, n, tt, tout, m->rtol, atol);
        steps = steps + dv.nst;
        if (istate > 0) break;
        if (istate != -1) return SHCFD_E_CHEM;  // "odd istate"
        atol = m->atol;
        nrecalc = nrecalc + 1;
        if (nrecalc > MAX_RECALC) return SHCFD_E_CHEM;  // "Exceed max_recalc."
    }
    n_to_vrho<NS>(m, n, vrho);
    return 0;
}

}  // namespace chem
}  // namespace shcfd
