// splines.cpp -- radial basis functions and their FP64 cubic-Hermite lookup tables.
//
// Radial bases restate radbase/calcCheb/radcore (src/radial.F90:6-33, 430-606, 630-668);
// radial functions R = sum crad * g as radfunc (src/radial.F90:610-626).  Tables follow
// setuplookupRadspline_* (src/radial.F90:164-426): nodes x_n = n*rc/nlut for n = nlut..1,
// coefficients c0 = f(x_n), c1 = h f'(x_n), c2 = 3(f1-f0) - f1' - 2 f0', c3 = -2(f1-f0) + f1' + f0'
// with f, f' at x_{nlut+1} taken as 0.  Rows are node-major so one lookup reads one row;
// within a row, functions are stored l-major (R_{l,n}) for the device loops.
#include <cmath>

#include "magace.hpp"

namespace magace {

namespace {
constexpr double kPi = 3.14159265358979323846264338327950288419;

void chebyshev(int n, double x, double* t, double* dt) {
    double u[40];
    t[0] = 1.0; dt[0] = 0.0; u[0] = 1.0;
    if (n <= 0) return;
    t[1] = x; u[1] = 2.0 * x;
    for (int m = 1; m < n; ++m) {
        t[m + 1] = 2.0 * x * t[m] - t[m - 1];
        u[m + 1] = 2.0 * x * u[m] - u[m - 1];
    }
    for (int m = 1; m <= n; ++m) dt[m] = (double)m * u[m - 1];  // T_m' = m U_{m-1}
}
}  // namespace

void radial_basis(int radtype, double lam, int nb, double cut, double dcut, double r, double* g,
                  double* dg) {
    if (r >= cut) {
        for (int n = 0; n < nb; ++n) g[n] = dg[n] = 0.0;
        return;
    }
    double t[40], dt[40];
    const double x0 = r / cut;
    if (radtype == 3 || radtype == 4) {
        // x = 1 - 2 (e^{-lam r/rc} - e^{-lam}) / (1 - e^{-lam}); g1 = 1, gn = (1 - T_{n-1}(x))/2
        const double y1 = std::exp(-lam * x0), y2 = std::exp(-lam);
        const double x = 1.0 - 2.0 * (y1 - y2) / (1.0 - y2);
        const double dx = 2.0 * lam * y1 / (cut * (1.0 - y2));
        chebyshev(nb - 1, x, t, dt);
        g[0] = 1.0; dg[0] = 0.0;
        for (int n = 1; n < nb; ++n) { g[n] = 0.5 - 0.5 * t[n]; dg[n] = -0.5 * dt[n] * dx; }
        if (radtype == 3) {  // cosine envelope
            const double env = 0.5 * (1.0 + std::cos(kPi * x0));
            const double denv = -0.5 * std::sin(kPi * x0) * kPi / cut;
            for (int n = 0; n < nb; ++n) { dg[n] = g[n] * denv + dg[n] * env; g[n] *= env; }
        }
    } else if (radtype == 5 || radtype == 6) {
        // power-law variable x = 2(1 - (1-r/rc)^lam) - 1
        const double y0 = 1.0 - x0;
        const double p = std::pow(y0, lam - 1.0);
        const double x = 2.0 * (1.0 - p * y0) - 1.0;
        const double dx = 2.0 * lam / cut * p;
        if (radtype == 5) {
            chebyshev(nb, x, t, dt);
            for (int n = 0; n < nb; ++n) { g[n] = 0.5 - 0.5 * t[n + 1]; dg[n] = -0.5 * dt[n + 1] * dx; }
        } else {
            chebyshev(nb - 1, x, t, dt);
            g[0] = 1.0; dg[0] = 0.0;
            for (int n = 1; n < nb; ++n) { g[n] = 0.5 - 0.5 * t[n]; dg[n] = -0.5 * dt[n] * dx; }
            const double env = 0.5 * (1.0 + std::cos(kPi * x0));
            const double denv = -0.5 * std::sin(kPi * x0) * kPi / cut;
            for (int n = 0; n < nb; ++n) { dg[n] = g[n] * denv + dg[n] * env; g[n] *= env; }
        }
    } else if (radtype == 7) {
        // g_n = T_n(1 - 2 (r/rc)^2)
        const double x = 1.0 - 2.0 * x0 * x0, dx = -(4.0 / cut) * x0;
        chebyshev(nb, x, t, dt);
        for (int n = 0; n < nb; ++n) { g[n] = t[n + 1]; dg[n] = dt[n + 1] * dx; }
    } else if (radtype == 8) {
        const double x = 1.0 - 2.0 * x0, dx = -2.0 / cut;
        chebyshev(nb, x, t, dt);
        for (int n = 0; n < nb; ++n) { g[n] = t[n + 1] - 1.0; dg[n] = dt[n + 1] * dx; }
    } else {
        for (int n = 0; n < nb; ++n) g[n] = dg[n] = std::nan("");
        return;
    }
    // smooth outer cutoff over [rc - dc, rc]
    const double lo = cut - dcut;
    if (r > lo) {
        const double fc = 0.5 * (1.0 + std::cos(kPi * (r - lo) / dcut));
        const double dfc = -0.5 * std::sin(kPi * (r - lo) / dcut) * kPi / dcut;
        for (int n = 0; n < nb; ++n) { dg[n] = g[n] * dfc + dg[n] * fc; g[n] *= fc; }
    }
}

void core_repulsion(double r, double pre, double lam, double cut, double& cr, double& dcr) {
    cr = dcr = 0.0;
    if (r >= cut) return;
    pre = std::fabs(pre);
    lam = std::fabs(lam);
    const double r2 = r * r, lr2 = lam * r2;
    if (lr2 >= 50.0) return;
    const double y = std::exp(-lr2);
    double c = pre * y / r;
    double dc = -pre * y * (2.0 * lr2 + 1.0) / r2;
    const double env = 0.5 * (1.0 + std::cos(kPi * r / cut));
    const double denv = -0.5 * std::sin(kPi * r / cut) * kPi / cut;
    dcr = c * denv + dc * env;
    cr = c * env;
}

Splines build_splines(const Potential& pot) {
    const Header& h = pot.h;
    Splines S;
    const int N = S.nlut;
    S.nfa = h.nr3_a * (h.l3a + 1);
    S.nfm = h.nr2_m * (h.l2m + 1);
    S.atomic_stride = 4 * (h.nb_a + S.nfa + 1);
    S.magnetic_stride = 4 * (h.nb_m + S.nfm);
    S.rs_a = (double)N / pot.rcut;
    S.rs_m = (double)N / pot.rcut_m;
    S.atomic.assign((size_t)N * S.atomic_stride, 0.0);
    S.magnetic.assign((size_t)N * S.magnetic_stride, 0.0);

    auto fill = [&](std::vector<double>& tab, int stride, int nf, double step,
                    auto&& eval /* (x, f[nf], df[nf]) */) {
        std::vector<double> f(nf), df(nf), f1(nf, 0.0), f1d(nf, 0.0);
        for (int n = N; n >= 1; --n) {
            const double x = step * (double)n;
            eval(x, f.data(), df.data());
            double* row = tab.data() + (size_t)(n - 1) * stride;
            for (int k = 0; k < nf; ++k) {
                const double f0 = f[k], f0d = df[k] * step;
                double* c = row + 4 * k;
                c[0] = f0;
                c[1] = f0d;
                c[2] = 3.0 * (f1[k] - f0) - f1d[k] - 2.0 * f0d;
                c[3] = -2.0 * (f1[k] - f0) + f1d[k] + f0d;
                f1[k] = f0;
                f1d[k] = f0d;
            }
        }
    };
    const int nfa_tot = h.nb_a + S.nfa + 1;
    fill(S.atomic, S.atomic_stride, nfa_tot, 1.0 / S.rs_a, [&](double x, double* f, double* df) {
        double g[40], dg[40];
        radial_basis(h.radtype_a, h.lambda_a, h.nb_a, h.cut_a, h.dcut_a, x, g, dg);
        for (int k = 0; k < h.nb_a; ++k) { f[k] = g[k]; df[k] = dg[k]; }
        for (int l = 0; l <= h.l3a; ++l)
            for (int n = 0; n < h.nr3_a; ++n) {
                const double* c = pot.crad_a.data() + ((size_t)n * (h.l3a + 1) + l) * h.nb_a;
                double s = 0.0, ds = 0.0;
                for (int k = 0; k < h.nb_a; ++k) { s += c[k] * g[k]; ds += c[k] * dg[k]; }
                f[h.nb_a + l * h.nr3_a + n] = s;
                df[h.nb_a + l * h.nr3_a + n] = ds;
            }
        double cr, dcr;
        core_repulsion(x, pot.core[0], pot.core[1], h.cut_a, cr, dcr);
        f[nfa_tot - 1] = cr;
        df[nfa_tot - 1] = dcr;
    });
    const int nfm_tot = h.nb_m + S.nfm;
    fill(S.magnetic, S.magnetic_stride, nfm_tot, 1.0 / S.rs_m, [&](double x, double* f, double* df) {
        double g[40], dg[40];
        radial_basis(h.radtype_m, h.lambda_m, h.nb_m, h.cut_m, h.dcut_m, x, g, dg);
        for (int k = 0; k < h.nb_m; ++k) { f[k] = g[k]; df[k] = dg[k]; }
        for (int l = 0; l <= h.l2m; ++l)
            for (int n = 0; n < h.nr2_m; ++n) {
                const double* c = pot.crad_m.data() + ((size_t)n * (h.l2m + 1) + l) * h.nb_m;
                double s = 0.0, ds = 0.0;
                for (int k = 0; k < h.nb_m; ++k) { s += c[k] * g[k]; ds += c[k] * dg[k]; }
                f[h.nb_m + l * h.nr2_m + n] = s;
                df[h.nb_m + l * h.nr2_m + n] = ds;
            }
    });
    return S;
}

}  // namespace magace
