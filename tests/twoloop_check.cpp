// Checks superdca_amd/csrc/lbfgs_coeffs.hpp (the two-loop recursion on coefficients used by
// the batched L-BFGS) against the recursion on explicit vectors, Nocedal & Wright,
// Algorithm 7.4, and against the secant equation H y_k = s_k of the newest pair.
// Prints one line per case:  <case> <n_pairs> <max rel err vs vectors> <secant rel err>
// Built and run by tests/test_twoloop.py.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "lbfgs_coeffs.hpp"

namespace {

constexpr int LD = 16;
using Vec = std::vector<double>;

double dot(const Vec &a, const Vec &b)
{
    double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

double rel_err(const Vec &a, const Vec &b)
{
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num = std::fmax(num, std::fabs(a[i] - b[i]));
        den = std::fmax(den, std::fabs(b[i]));
    }
    return num / den;
}

// -H g by Algorithm 7.4 on the vectors; S, Y listed oldest first
Vec two_loop_vectors(const std::vector<Vec> &S, const std::vector<Vec> &Y, const Vec &g)
{
    const int k = (int)S.size();
    Vec q = g;
    std::vector<double> a(k), rho(k);
    for (int i = k - 1; i >= 0; --i) {
        rho[i] = 1.0 / dot(Y[i], S[i]);
        a[i] = rho[i] * dot(S[i], q);
        for (size_t j = 0; j < q.size(); ++j) q[j] -= a[i] * Y[i][j];
    }
    const double gamma = k ? dot(S[k - 1], Y[k - 1]) / dot(Y[k - 1], Y[k - 1]) : 1.0;
    Vec r(q.size());
    for (size_t j = 0; j < q.size(); ++j) r[j] = gamma * q[j];
    for (int i = 0; i < k; ++i) {
        const double b = rho[i] * dot(Y[i], r);
        for (size_t j = 0; j < r.size(); ++j) r[j] += (a[i] - b) * S[i][j];
    }
    for (double &v : r) v = -v;
    return r;
}

// the direction from the coefficients; slot[t] is the pair slot of the t-th oldest pair
Vec two_loop_coeff_dir(const std::vector<Vec> &S, const std::vector<Vec> &Y, const Vec &g, const std::vector<int> &slot,
                       int m)
{
    const int k = (int)S.size();
    double SY[LD][LD] = {}, YY[LD][LD] = {}, SG[LD] = {}, YG[LD] = {};
    for (int t = 0; t < k; ++t) {
        const int p = slot[t];
        SG[p] = dot(S[t], g);
        YG[p] = dot(Y[t], g);
        for (int u = 0; u < k; ++u) {
            SY[p][slot[u]] = dot(S[t], Y[u]);
            YY[p][slot[u]] = dot(Y[t], Y[u]);
        }
    }
    double cs[LD], cy[LD], cg;
    sdca::two_loop_coeffs(slot.data(), k, m, SY, YY, SG, YG, cs, cy, cg);
    Vec d(g.size());
    for (size_t j = 0; j < g.size(); ++j) {
        double v = cg * g[j];
        for (int t = 0; t < k; ++t) v += cs[slot[t]] * S[t][j] + cy[slot[t]] * Y[t][j];
        d[j] = v;
    }
    return d;
}

}  // namespace

int main()
{
    std::mt19937_64 rng(12345);
    std::normal_distribution<double> N01(0.0, 1.0);
    std::uniform_real_distribution<double> U(0.1, 10.0);
    const int n = 60;
    // A = Q diag(lam) Q^T, Q a product of Householder reflections
    Vec lam(n);
    for (double &v : lam) v = U(rng);
    std::vector<Vec> hh(3, Vec(n));
    for (Vec &h : hh) {
        for (double &v : h) v = N01(rng);
        const double nn = std::sqrt(dot(h, h));
        for (double &v : h) v /= nn;
    }
    auto apply_A = [&](const Vec &x) {
        Vec y = x;
        for (int r = 2; r >= 0; --r) {  // Q^T x
            const double c = 2.0 * dot(hh[r], y);
            for (int j = 0; j < n; ++j) y[j] -= c * hh[r][j];
        }
        for (int j = 0; j < n; ++j) y[j] *= lam[j];
        for (int r = 0; r < 3; ++r) {  // Q y
            const double c = 2.0 * dot(hh[r], y);
            for (int j = 0; j < n; ++j) y[j] -= c * hh[r][j];
        }
        return y;
    };
    // (number of pairs, number of slots, slot order): slots in order, and a ring that has
    // wrapped around, so the oldest pair does not sit in slot 0
    struct Case {
        int k, m;
        std::vector<int> slot;
    };
    const std::vector<Case> cases = {{0, 10, {}},
                                     {1, 10, {0}},
                                     {3, 10, {0, 1, 2}},
                                     {10, 10, {0, 1, 2, 3, 4, 5, 6, 7, 8, 9}},
                                     {5, 5, {3, 4, 0, 1, 2}},
                                     {8, 16, {7, 2, 11, 0, 15, 4, 9, 13}}};
    int id = 0;
    for (const Case &c : cases) {
        std::vector<Vec> S(c.k, Vec(n)), Y(c.k);
        // y = A s plus a perturbation: pairs of a non-quadratic function, whose s_p.y_q
        // matrix is not symmetric (curvature s.y > 0 kept)
        for (int t = 0; t < c.k; ++t) {
            do {
                for (double &v : S[t]) v = N01(rng);
                Y[t] = apply_A(S[t]);
                for (double &v : Y[t]) v += 0.5 * N01(rng);
            } while (dot(S[t], Y[t]) <= 0.0);
        }
        Vec g(n);
        for (double &v : g) v = N01(rng);
        const Vec dv = two_loop_vectors(S, Y, g);
        const Vec dc = two_loop_coeff_dir(S, Y, g, c.slot, c.m);
        double secant = 0.0;
        if (c.k > 0) {  // -H y_newest = -s_newest
            const Vec ds = two_loop_coeff_dir(S, Y, Y[c.k - 1], c.slot, c.m);
            Vec ms(n);
            for (int j = 0; j < n; ++j) ms[j] = -S[c.k - 1][j];
            secant = rel_err(ds, ms);
        }
        std::printf("%d %d %.3e %.3e\n", id++, c.k, rel_err(dc, dv), secant);
    }
    return 0;
}
