// linesearch.hpp -- More-Thuente line search as a resumable state machine.
//
// J. J. More and D. J. Thuente, "Line search algorithms with guaranteed sufficient
// decrease", ACM TOMS 20 (1994) 286-307: the safeguarded cubic/quadratic step of the
// paper and its two-stage driver.  SuperDCA's optimizer (CppNumericalSolvers'
// LbfgsSolver with MoreThuente, used at SuperDCA/include/plmDCA.hpp:218-221,304 and
// patched to evaluate value and gradient together,
// SuperDCA/externals/CppNumericalSolvers.linesearch_morethuente.patch) uses the same
// method; its constants are not vendored and are chosen here: ftol 1e-4, gtol 0.9,
// xtol 1e-10, extrapolation bounds 1.1 and 4.
//
// Resumable form: start() with f(0) and f'(0); for every trial step the caller
// evaluates f and f' there and calls step(), which returns Continue (evaluate at the
// new stp), Converged, or Warning (no further progress possible).
#pragma once

#include <algorithm>
#include <cmath>

namespace sdca {

struct MoreThuente {
    enum Result { Continue = 0, Converged = 1, Warning = 2 };
    double ftol = 1e-4, gtol = 0.9, xtol = 1e-10, stpmin = 1e-20, stpmax = 1e20;
    bool brackt = false;
    int stage = 1;
    double finit = 0, ginit = 0, gtest = 0, width = 0, width1 = 0;
    double stx = 0, fx = 0, gx = 0, sty = 0, fy = 0, gy = 0, stmin = 0, stmax = 0;

    void start(double f0, double g0, double stp)
    {
        brackt = false;
        stage = 1;
        finit = f0;
        ginit = g0;
        gtest = ftol * g0;
        width = stpmax - stpmin;
        width1 = 2.0 * width;
        stx = 0.0;
        fx = f0;
        gx = g0;
        sty = 0.0;
        fy = f0;
        gy = g0;
        stmin = 0.0;
        stmax = stp + 4.0 * stp;
    }

    // Safeguarded step: new trial step from the best point (stx), the other interval
    // end (sty) and the current trial (stp).
    static void cstep(double &stx, double &fx, double &dx, double &sty, double &fy, double &dy, double &stp, double fp,
                      double dp, bool &brackt, double lo, double hi)
    {
        const double sgnd = dp * (dx / std::fabs(dx));
        double stpf;
        if (fp > fx) {  // higher function value: the minimum is bracketed
            const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
            const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
            double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
            if (stp < stx) gamma = -gamma;
            const double p = (gamma - dx) + theta;
            const double q = ((gamma - dx) + gamma) + dp;
            const double stpc = stx + (p / q) * (stp - stx);
            const double stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx);
            stpf = (std::fabs(stpc - stx) < std::fabs(stpq - stx)) ? stpc : stpc + (stpq - stpc) / 2.0;
            brackt = true;
        } else if (sgnd < 0.0) {  // derivatives of opposite sign: bracketed
            const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
            const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
            double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
            if (stp > stx) gamma = -gamma;
            const double p = (gamma - dp) + theta;
            const double q = ((gamma - dp) + gamma) + dx;
            const double stpc = stp + (p / q) * (stx - stp);
            const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
            stpf = (std::fabs(stpc - stp) > std::fabs(stpq - stp)) ? stpc : stpq;
            brackt = true;
        } else if (std::fabs(dp) < std::fabs(dx)) {  // same sign, decreasing magnitude
            const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
            const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
            double gamma = s * std::sqrt(std::max(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
            if (stp > stx) gamma = -gamma;
            const double p = (gamma - dp) + theta;
            const double q = (gamma + (dx - dp)) + gamma;
            const double r = p / q;
            double stpc;
            if (r < 0.0 && gamma != 0.0) stpc = stp + r * (stx - stp);
            else if (stp > stx) stpc = hi;
            else stpc = lo;
            const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
            if (brackt) {
                stpf = (std::fabs(stpc - stp) < std::fabs(stpq - stp)) ? stpc : stpq;
                stpf = (stp > stx) ? std::min(stp + 0.66 * (sty - stp), stpf) : std::max(stp + 0.66 * (sty - stp), stpf);
            } else {
                stpf = (std::fabs(stpc - stp) > std::fabs(stpq - stp)) ? stpc : stpq;
                stpf = std::max(lo, std::min(hi, stpf));
            }
        } else {  // same sign, no decrease in magnitude
            if (brackt) {
                const double theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp;
                const double s = std::max({std::fabs(theta), std::fabs(dy), std::fabs(dp)});
                double gamma = s * std::sqrt((theta / s) * (theta / s) - (dy / s) * (dp / s));
                if (stp > sty) gamma = -gamma;
                const double p = (gamma - dp) + theta;
                const double q = ((gamma - dp) + gamma) + dy;
                stpf = stp + (p / q) * (sty - stp);
            } else {
                stpf = (stp > stx) ? hi : lo;
            }
        }
        if (fp > fx) {
            sty = stp;
            fy = fp;
            dy = dp;
        } else {
            if (sgnd < 0.0) {
                sty = stx;
                fy = fx;
                dy = dx;
            }
            stx = stp;
            fx = fp;
            dx = dp;
        }
        stp = stpf;
    }

    Result step(double f, double g, double &stp)
    {
        const double ftest = finit + stp * gtest;
        if (stage == 1 && f <= ftest && g >= 0.0) stage = 2;
        if (brackt && (stp <= stmin || stp >= stmax)) return Warning;
        if (brackt && stmax - stmin <= xtol * stmax) return Warning;
        if (stp == stpmax && f <= ftest && g <= gtest) return Warning;
        if (stp == stpmin && (f > ftest || g >= gtest)) return Warning;
        if (f <= ftest && std::fabs(g) <= gtol * (-ginit)) return Converged;
        if (stage == 1 && f <= fx && f > ftest) {
            // modified function psi(stp) = f(stp) - f(0) - ftol*stp*f'(0)
            double fm = f - stp * gtest, fxm = fx - stx * gtest, fym = fy - sty * gtest;
            double gm = g - gtest, gxm = gx - gtest, gym = gy - gtest;
            cstep(stx, fxm, gxm, sty, fym, gym, stp, fm, gm, brackt, stmin, stmax);
            fx = fxm + stx * gtest;
            fy = fym + sty * gtest;
            gx = gxm + gtest;
            gy = gym + gtest;
        } else {
            cstep(stx, fx, gx, sty, fy, gy, stp, f, g, brackt, stmin, stmax);
        }
        if (brackt) {
            if (std::fabs(sty - stx) >= 0.66 * width1) stp = stx + 0.5 * (sty - stx);
            width1 = width;
            width = std::fabs(sty - stx);
            stmin = std::min(stx, sty);
            stmax = std::max(stx, sty);
        } else {
            stmin = stp + 1.1 * (stp - stx);
            stmax = stp + 4.0 * (stp - stx);
        }
        stp = std::max(stp, stpmin);
        stp = std::min(stp, stpmax);
        if (brackt && (stp <= stmin || stp >= stmax || stmax - stmin <= xtol * stmax)) stp = stx;
        return Continue;
    }
};

}  // namespace sdca
