// lbfgs_coeffs.hpp -- the L-BFGS two-loop recursion on coefficients.
//
// Nocedal & Wright, Numerical Optimization (2nd ed.), Algorithm 7.4, with the initial
// matrix gamma I, gamma = s.y / y.y of the newest pair: q = g; for the pairs newest
// first a_p = rho_p s_p.q, q -= a_p y_p; r = gamma q; for the pairs oldest first
// b = rho_p y_p.r, r += (a_p - b) s_p; the direction is d = -r = -H g.
//
// Every vector of the recursion lies in the span of {g, s_p, y_p}, so it runs here on
// coefficient vectors over that span, from the inner products alone:
//   SY[p][q] = s_p.y_q, YY[p][q] = y_p.y_q, SG[p] = s_p.g, YG[p] = y_p.g.
// The result is d = cg g + sum_p (cs_p s_p + cy_p y_p); the stored vectors are read once,
// by the pass that forms d.
#pragma once

namespace sdca {

// pairs[0..n) are the pair slots, oldest first; m is the number of slots (cs and cy get m
// entries, zero outside pairs); LD is the leading dimension of SY and YY.
template <int LD>
inline void two_loop_coeffs(const int *pairs, int n, int m, const double (&SY)[LD][LD], const double (&YY)[LD][LD],
                            const double *SG, const double *YG, double *cs, double *cy, double &cg)
{
    for (int p = 0; p < m; ++p) cs[p] = cy[p] = 0.0;
    cg = 1.0;
    double a[LD];
    for (int t = n - 1; t >= 0; --t) {
        const int p = pairs[t];
        double sq = cg * SG[p];  // s_p . q, q = cg g + sum cy y
        for (int v = 0; v < n; ++v) sq += cy[pairs[v]] * SY[p][pairs[v]];
        a[p] = sq / SY[p][p];
        cy[p] -= a[p];
    }
    if (n > 0) {
        const int p = pairs[n - 1];
        const double gam = SY[p][p] / YY[p][p];
        cg *= gam;
        for (int u = 0; u < m; ++u) cy[u] *= gam;
    }
    for (int t = 0; t < n; ++t) {
        const int p = pairs[t];
        double yr = cg * YG[p];  // y_p . r, r = cg g + sum cy y + sum cs s
        for (int v = 0; v < n; ++v) {
            const int u = pairs[v];
            yr += cy[u] * YY[p][u] + cs[u] * SY[u][p];
        }
        cs[p] += a[p] - yr / SY[p][p];
    }
    for (int p = 0; p < m; ++p) {
        cs[p] = -cs[p];
        cy[p] = -cy[p];
    }
    cg = -cg;
}

}  // namespace sdca
