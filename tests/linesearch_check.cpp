// Drives superdca_amd/csrc/linesearch.hpp (the host line search of the batched L-BFGS)
// on the one-dimensional test functions of More and Thuente, ACM TOMS 20 (1994) 286-307,
// section 5, and prints one line per run:
//   <function> <alpha0> <stp> <f(stp)> <f'(stp)> <f(0)> <f'(0)> <evaluations> <result>
// Built and run by tests/test_linesearch.py.
#include <cmath>
#include <cstdio>

#include "linesearch.hpp"

namespace {

const double kPi = 3.14159265358979323846;

// phi(a) = -a / (a^2 + b), b = 2
void f1(double a, double &f, double &g)
{
    const double b = 2.0;
    f = -a / (a * a + b);
    g = (a * a - b) / ((a * a + b) * (a * a + b));
}

// phi(a) = (a + b)^5 - 2 (a + b)^4, b = 0.004
void f2(double a, double &f, double &g)
{
    const double b = 0.004, t = a + b;
    f = std::pow(t, 5) - 2.0 * std::pow(t, 4);
    g = 5.0 * std::pow(t, 4) - 8.0 * std::pow(t, 3);
}

// phi(a) = phi0(a) + 2 (1 - b) / (l pi) sin(l pi a / 2), b = 0.01, l = 39
void f3(double a, double &f, double &g)
{
    const double b = 0.01, l = 39.0;
    double f0, g0;
    if (a <= 1.0 - b) {
        f0 = 1.0 - a;
        g0 = -1.0;
    } else if (a >= 1.0 + b) {
        f0 = a - 1.0;
        g0 = 1.0;
    } else {
        f0 = (a - 1.0) * (a - 1.0) / (2.0 * b) + b / 2.0;
        g0 = (a - 1.0) / b;
    }
    f = f0 + 2.0 * (1.0 - b) / (l * kPi) * std::sin(l * kPi * a / 2.0);
    g = g0 + (1.0 - b) * std::cos(l * kPi * a / 2.0);
}

// Yanai, Ozawa and Kaneko: phi(a) = gamma(b1) sqrt((1 - a)^2 + b2^2) + gamma(b2) sqrt(a^2 + b1^2),
// gamma(b) = sqrt(1 + b^2) - b
void yok(double a, double b1, double b2, double &f, double &g)
{
    const double g1 = std::sqrt(1.0 + b1 * b1) - b1, g2 = std::sqrt(1.0 + b2 * b2) - b2;
    const double s1 = std::sqrt((1.0 - a) * (1.0 - a) + b2 * b2), s2 = std::sqrt(a * a + b1 * b1);
    f = g1 * s1 + g2 * s2;
    g = -g1 * (1.0 - a) / s1 + g2 * a / s2;
}
void f4(double a, double &f, double &g) { yok(a, 0.001, 0.001, f, g); }
void f5(double a, double &f, double &g) { yok(a, 0.01, 0.001, f, g); }
void f6(double a, double &f, double &g) { yok(a, 0.001, 0.01, f, g); }

struct Case {
    int id;
    void (*fn)(double, double &, double &);
    double ftol, gtol;
};

}  // namespace

int main()
{
    // the constants of the paper's tables: (mu, eta) = (0.001, 0.1), (0.1, 0.1), (0.1, 0.1),
    // (0.001, 0.001) for functions 4-6
    const Case cases[] = {{1, f1, 1e-3, 0.1}, {2, f2, 0.1, 0.1},     {3, f3, 0.1, 0.1},
                          {4, f4, 1e-3, 1e-3}, {5, f5, 1e-3, 1e-3}, {6, f6, 1e-3, 1e-3}};
    const double starts[] = {1e-3, 1e-1, 1e1, 1e3};
    for (const Case &c : cases) {
        for (double a0 : starts) {
            sdca::MoreThuente ls;
            ls.ftol = c.ftol;
            ls.gtol = c.gtol;
            ls.xtol = 1e-16;
            double f0, g0;
            c.fn(0.0, f0, g0);
            ls.start(f0, g0, a0);
            double stp = a0, f = 0, g = 0;
            int nev = 0, res = sdca::MoreThuente::Continue;
            while (res == sdca::MoreThuente::Continue && nev < 100) {
                c.fn(stp, f, g);
                ++nev;
                res = ls.step(f, g, stp);
            }
            std::printf("%d %.17g %.17g %.17g %.17g %.17g %.17g %d %d\n", c.id, a0, stp, f, g, f0, g0, nev, res);
        }
    }
    return 0;
}
