// fit.cpp -- chi^2 fit of a normal bath (e_k, V_k) of one orbital to a
// Matsubara function, the host-side step after Sigma in a DMFT loop
// (chi2_fitgf_normal_normal, ED_CHI2FIT/fitgf_normal_normal.f90:18-156).
//
//   chi2(a) = sum_n |fg(n) - F_a(iw_n)|^2 / W_n,   n = 1..L
//   F = Delta_a(iw) = sum_k V_k^2/(iw - e_k)                  (cg_scheme delta)
//   F = 1/(iw + xmu - hloc - Delta_a(iw))                     (cg_scheme weiss)
//   W_n = L (cg_weight 0), 1 (1), n (2), w_n (3)   (fitgf_normal_normal.f90:56-65)
//
// Minimiser: Polak-Ribiere nonlinear conjugate gradients with a bracketing
// Brent line search and the analytic gradient, stopping when two successive
// function values satisfy 2|f1-f0| <= ftol (|f1|+|f0|+1e-10).  The reference
// calls SciFortran's fmin_cg, which is not available; the iteration history
// therefore differs and only the fitted function is comparable.
#include <cmath>
#include <complex>
#include <vector>

#include "dmed_internal.h"

namespace {

using cd = std::complex<double>;

// a chi^2 with its gradient over the packed bath parameters
struct Cost {
    virtual ~Cost() = default;
    virtual double eval(const std::vector<double> &a, std::vector<double> *grad) const = 0;
};

struct Chi2 : Cost {
    int L, nb, scheme, weight;
    double beta, xmu, hloc;
    std::vector<cd> fg;
    std::vector<double> wm, W;

    double eval(const std::vector<double> &a, std::vector<double> *grad) const override
    {
        const double *e = a.data(), *v = a.data() + nb;
        double f = 0.0;
        if (grad) grad->assign(2 * nb, 0.0);
        for (int n = 0; n < L; ++n) {
            const cd iw(0.0, wm[n]);
            cd delta = 0.0;
            for (int k = 0; k < nb; ++k) delta += v[k] * v[k] / (iw - e[k]);
            cd F, g0 = 0.0;
            if (scheme == 1) {
                F = delta;
            } else {
                g0 = 1.0 / (iw + xmu - hloc - delta);
                F = g0;
            }
            const cd r = fg[n] - F;
            f += std::norm(r) / W[n];
            if (grad) {
                // dF/da: delta scheme dDelta, weiss scheme g0^2 dDelta
                for (int k = 0; k < nb; ++k) {
                    const cd den = iw - e[k];
                    cd de = v[k] * v[k] / (den * den);
                    cd dv = 2.0 * v[k] / den;
                    if (scheme != 1) {
                        de *= g0 * g0;
                        dv *= g0 * g0;
                    }
                    (*grad)[k] += -2.0 * (r.real() * de.real() + r.imag() * de.imag()) / W[n];
                    (*grad)[nb + k] += -2.0 * (r.real() * dv.real() + r.imag() * dv.imag()) / W[n];
                }
            }
        }
        return f;
    }
};

// minimise f(p + t d) over t; returns the minimum value and moves p
double line_min(const Cost &c, std::vector<double> &p, const std::vector<double> &d)
{
    const size_t n = p.size();
    std::vector<double> q(n);
    auto f1 = [&](double t) {
        for (size_t i = 0; i < n; ++i) q[i] = p[i] + t * d[i];
        return c.eval(q, nullptr);
    };
    // bracket a minimum: a < b < cc with f(b) <= f(a), f(cc)
    const double gold = 1.618033988749895;
    double a = 0.0, b = 1e-3;
    double fa = f1(a), fb = f1(b);
    if (fb > fa) {
        // shrink until f(b) < f(a) or give up
        int it = 0;
        while (fb > fa && it++ < 60) {
            b *= 0.1;
            fb = f1(b);
        }
        if (fb > fa) return fa;
    }
    double cc = b + gold * (b - a), fc = f1(cc);
    int it = 0;
    while (fc < fb && it++ < 200) {
        a = b; fa = fb;
        b = cc; fb = fc;
        cc = b + gold * (b - a);
        fc = f1(cc);
    }
    // Brent's method on [a, cc] seeded at b
    const double cgold = 0.3819660112501051, zeps = 1e-18, tol = 3e-8;
    double lo = std::min(a, cc), hi = std::max(a, cc);
    double x = b, w = b, v = b, fx = fb, fw = fb, fv = fb, dd = 0.0, e = 0.0;
    for (int iter = 0; iter < 200; ++iter) {
        const double xm = 0.5 * (lo + hi);
        const double tol1 = tol * std::fabs(x) + zeps, tol2 = 2.0 * tol1;
        if (std::fabs(x - xm) <= tol2 - 0.5 * (hi - lo)) break;
        bool golden = true;
        if (std::fabs(e) > tol1) {
            double r = (x - w) * (fx - fv), qq = (x - v) * (fx - fw);
            double pp = (x - v) * qq - (x - w) * r;
            qq = 2.0 * (qq - r);
            if (qq > 0.0) pp = -pp;
            qq = std::fabs(qq);
            const double etemp = e;
            e = dd;
            if (!(std::fabs(pp) >= std::fabs(0.5 * qq * etemp) || pp <= qq * (lo - x) ||
                  pp >= qq * (hi - x))) {
                dd = pp / qq;
                const double u = x + dd;
                if (u - lo < tol2 || hi - u < tol2) dd = (xm - x) >= 0 ? tol1 : -tol1;
                golden = false;
            }
        }
        if (golden) {
            e = (x >= xm) ? lo - x : hi - x;
            dd = cgold * e;
        }
        const double u = std::fabs(dd) >= tol1 ? x + dd : x + (dd >= 0 ? tol1 : -tol1);
        const double fu = f1(u);
        if (fu <= fx) {
            if (u >= x) lo = x; else hi = x;
            v = w; fv = fw;
            w = x; fw = fx;
            x = u; fx = fu;
        } else {
            if (u < x) lo = u; else hi = u;
            if (fu <= fw || w == x) {
                v = w; fv = fw;
                w = u; fw = fu;
            } else if (fu <= fv || v == x || v == w) {
                v = u; fv = fu;
            }
        }
    }
    for (size_t i = 0; i < n; ++i) p[i] += x * d[i];
    return fx;
}

// chi2_fitgf_hybrid_normal (ED_CHI2FIT/fitgf_hybrid_normal.f90:18-156,
// 230-333): parameters a = [e_k (Nbath), V_lk (Norb x Nbath, l-major)],
//   Delta_ab(z) = sum_k V_ak V_bk/(z - e_k),
//   F = Delta (delta) or G0 = [(z + xmu) 1 - Hloc - Delta]^-1 (weiss),
//   chi2 = sum_{a <= b} sum_n |fg_ab(n) - F_ab(iw_n)|^2 / W_n.
// Gradient: dF = dDelta, or G0 dDelta G0 for weiss, with
// dDelta/de_k = s u u^T (s = 1/(z-e_k)^2, u = V_{.k}) and
// dDelta/dV_lk = t (e_l u^T + u e_l^T) (t = 1/(z-e_k)).
struct Chi2Hybrid : Cost {
    int L, nb, no, scheme;
    double xmu;
    std::vector<double> hloc;        // [a*no + b]
    std::vector<cd> fg;              // [(a*no + b)*L + n]
    std::vector<double> wm, W;

    double eval(const std::vector<double> &a, std::vector<double> *grad) const override
    {
        const double *e = a.data(), *v = a.data() + nb;   // v[l*nb + k]
        double f = 0.0;
        const size_t np = a.size();
        if (grad) grad->assign(np, 0.0);
        std::vector<cd> D(no * no), G(no * no), F(no * no), Gu(no), uG(no);
        for (int n = 0; n < L; ++n) {
            const cd z(0.0, wm[n]);
            for (int i = 0; i < no * no; ++i) D[i] = 0.0;
            for (int k = 0; k < nb; ++k) {
                const cd t = 1.0 / (z - e[k]);
                for (int i = 0; i < no; ++i)
                    for (int j = 0; j < no; ++j) D[i * no + j] += v[i * nb + k] * v[j * nb + k] * t;
            }
            if (scheme == 1) {
                F = D;
            } else {
                for (int i = 0; i < no; ++i)
                    for (int j = 0; j < no; ++j)
                        G[i * no + j] = (i == j ? z + xmu : cd(0.0)) - hloc[i * no + j] - D[i * no + j];
                if (!dmed::invert_small(no, G.data())) return HUGE_VAL;
                F = G;
            }
            for (int i = 0; i < no; ++i)
                for (int j = i; j < no; ++j) {
                    const cd r = fg[(size_t)(i * no + j) * L + n] - F[i * no + j];
                    f += std::norm(r) / W[n];
                }
            if (!grad) continue;
            for (int k = 0; k < nb; ++k) {
                const cd t = 1.0 / (z - e[k]), s2 = t * t;
                // G0 u and u^T G0 (weiss); u itself (delta)
                for (int i = 0; i < no; ++i) {
                    Gu[i] = 0.0;
                    uG[i] = 0.0;
                    for (int j = 0; j < no; ++j) {
                        const cd gij = scheme == 1 ? cd(i == j ? 1.0 : 0.0) : G[i * no + j];
                        const cd gji = scheme == 1 ? cd(i == j ? 1.0 : 0.0) : G[j * no + i];
                        Gu[i] += gij * v[j * nb + k];
                        uG[i] += v[j * nb + k] * gji;
                    }
                }
                for (int i = 0; i < no; ++i)
                    for (int j = i; j < no; ++j) {
                        const cd r = fg[(size_t)(i * no + j) * L + n] - F[i * no + j];
                        auto acc = [&](size_t idx, cd dF) {
                            (*grad)[idx] += -2.0 * (r.real() * dF.real() + r.imag() * dF.imag()) / W[n];
                        };
                        acc(k, s2 * Gu[i] * uG[j]);
                        for (int l = 0; l < no; ++l) {
                            // (G0 e_l)_i (u^T G0)_j + (G0 u)_i (e_l^T G0)_j
                            const cd gil = scheme == 1 ? cd(i == l ? 1.0 : 0.0) : G[i * no + l];
                            const cd glj = scheme == 1 ? cd(l == j ? 1.0 : 0.0) : G[l * no + j];
                            acc(nb + (size_t)l * nb + k, t * (gil * uG[j] + Gu[i] * glj));
                        }
                    }
            }
        }
        return f;
    }
};

// Polak-Ribiere conjugate gradients over p with the line search above,
// stopping when 2|f1 - f0| <= ftol (|f1| + |f0| + 1e-10); returns the last
// chi^2 and the iteration count
double cg_minimize(const Cost &c, std::vector<double> &p, int niter, double ftol, int *iters)
{
    std::vector<double> g, h, xi;
    double fp = c.eval(p, &xi);
    for (auto &t : xi) t = -t;
    g = xi;
    h = xi;
    int its = 0;
    for (its = 1; its <= niter; ++its) {
        const double fret = line_min(c, p, xi);
        if (2.0 * std::fabs(fret - fp) <= ftol * (std::fabs(fret) + std::fabs(fp) + 1e-10)) {
            fp = fret;
            break;
        }
        fp = fret;
        c.eval(p, &xi);
        double gg = 0.0, dgg = 0.0;
        for (size_t i = 0; i < p.size(); ++i) {
            gg += g[i] * g[i];
            dgg += (xi[i] + g[i]) * xi[i];
        }
        if (gg == 0.0) break;
        const double gam = std::max(0.0, dgg / gg);
        for (size_t i = 0; i < p.size(); ++i) {
            g[i] = -xi[i];
            h[i] = g[i] + gam * h[i];
            xi[i] = h[i];
        }
    }
    *iters = std::min(its, niter);
    return fp;
}

double fit_weight(int weight, int n, int L, double wn)
{
    switch (weight) {
        case 1: return 1.0;
        case 2: return (double)(n + 1);
        case 3: return wn;
        default: return (double)L;
    }
}

}  // namespace

extern "C" int dmed_chi2_fit_hybrid(int32_t scheme, int32_t weight, const double *fg, int32_t L,
                                    double beta, double xmu, const double *hloc, int32_t norb,
                                    int32_t nbath, double *e, double *v, int32_t niter,
                                    double ftol, double *chi2, int32_t *iter)
{
    if (!fg || !hloc || !e || !v || !chi2 || L < 1 || nbath < 0 || beta <= 0 || norb < 1 ||
        norb > 3)
        return dmed::fail(DMED_EINVAL, "dmed_chi2_fit_hybrid: bad argument");
    Chi2Hybrid c;
    c.L = L;
    c.nb = nbath;
    c.no = norb;
    c.scheme = scheme;
    c.xmu = xmu;
    c.hloc.assign(norb * norb, 0.0);
    c.fg.assign((size_t)norb * norb * L, 0.0);
    c.wm.resize(L);
    c.W.resize(L);
    for (int a = 0; a < norb; ++a)
        for (int b = 0; b < norb; ++b) {
            c.hloc[a * norb + b] = hloc[a + norb * b];
            for (int n = 0; n < L; ++n) {
                const size_t f = (size_t)a + norb * ((size_t)b + (size_t)norb * n);
                c.fg[(size_t)(a * norb + b) * L + n] = cd(fg[2 * f], fg[2 * f + 1]);
            }
        }
    for (int n = 0; n < L; ++n) {
        c.wm[n] = M_PI / beta * (2 * (n + 1) - 1);
        c.W[n] = fit_weight(weight, n, L, c.wm[n]);
    }
    std::vector<double> p(nbath + (size_t)norb * nbath);
    for (int k = 0; k < nbath; ++k) p[k] = e[k];
    for (int l = 0; l < norb; ++l)
        for (int k = 0; k < nbath; ++k) p[nbath + (size_t)l * nbath + k] = v[k + (size_t)l * nbath];
    int its = 0;
    *chi2 = cg_minimize(c, p, niter, ftol, &its);
    for (int k = 0; k < nbath; ++k) e[k] = p[k];
    for (int l = 0; l < norb; ++l)
        for (int k = 0; k < nbath; ++k) v[k + (size_t)l * nbath] = p[nbath + (size_t)l * nbath + k];
    if (iter) *iter = its;
    return 0;
}

extern "C" int dmed_chi2_fit(int32_t scheme, int32_t weight, const double *fg, int32_t L,
                             double beta, double xmu, double hloc_aa, int32_t nbath, double *e,
                             double *v, int32_t niter, double ftol, double *chi2, int32_t *iter)
{
    if (!fg || !e || !v || !chi2 || L < 1 || nbath < 0 || beta <= 0)
        return dmed::fail(DMED_EINVAL, "dmed_chi2_fit: bad argument");
    Chi2 c;
    c.L = L;
    c.nb = nbath;
    c.scheme = scheme;
    c.weight = weight;
    c.beta = beta;
    c.xmu = xmu;
    c.hloc = hloc_aa;
    c.fg.resize(L);
    c.wm.resize(L);
    c.W.resize(L);
    for (int n = 0; n < L; ++n) {
        c.fg[n] = cd(fg[2 * n], fg[2 * n + 1]);
        c.wm[n] = M_PI / beta * (2 * (n + 1) - 1);
        c.W[n] = fit_weight(weight, n, L, c.wm[n]);
    }
    std::vector<double> p(2 * nbath);
    for (int k = 0; k < nbath; ++k) {
        p[k] = e[k];
        p[nbath + k] = v[k];
    }
    int its = 0;
    const double fp = cg_minimize(c, p, niter, ftol, &its);
    for (int k = 0; k < nbath; ++k) {
        e[k] = p[k];
        v[k] = p[nbath + k];
    }
    *chi2 = fp;
    if (iter) *iter = its;
    return 0;
}
