// hostla.cpp -- small dense linear algebra and pole sums on the host.
//
// host_syev: symmetric eigenproblem of an n x n matrix (n up to a few
// hundred: Krylov projections and sectors of dim <= lanc_dim_threshold).
// Householder reduction to tridiagonal form, then implicit QL with Wilkinson
// shifts (the iteration of tql2, ED_GREENS_FUNCTIONS.f90:361-499) applied to
// the accumulated orthogonal matrix.
//
// tql2_first_row: the reference's tql2 on (alpha, beta) with Z = I, carrying
// only the first row of Z -- the pole weights Z(1,j)^2 of
// add_to_lanczos_gf_normal (build_gf_normal.f90:470-487) -- in O(n^2).
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "dmed_internal.h"

namespace dmed {

static double hyp(double a, double b)
{
    // sqrt(a^2+b^2) without destructive underflow/overflow (Moler-Morrison,
    // the pythag of ED_GREENS_FUNCTIONS.f90:513-539)
    double p = std::max(std::fabs(a), std::fabs(b));
    if (std::isnan(a) || std::isnan(b)) return NAN;
    if (p != 0.0) {
        double r = std::min(std::fabs(a), std::fabs(b)) / p;
        r = r * r;
        for (;;) {
            const double t = 4.0 + r;
            if (t == 4.0) break;
            const double s = r / t;
            const double u = 1.0 + 2.0 * s;
            p = u * p;
            r = (s / u) * (s / u) * r;
        }
    }
    return p;
}

// Implicit QL on the tridiagonal (d[0..n), e[1..n) = sub-diagonal; e[0]
// unused), rotations applied to the `nrow` x n row-major... column-major
// matrix z (leading dimension ldz).  Ascending sort of d with the columns.
static int ql_implicit(int n, double *d, double *e, double *z, int nrow, int ldz)
{
    if (n <= 1) return 0;
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    e[n - 1] = 0.0;
    double f = 0.0, tst1 = 0.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        tst1 = std::max(tst1, std::fabs(d[l]) + std::fabs(e[l]));
        int m = l;
        for (; m < n; ++m)
            if (tst1 + std::fabs(e[m]) == tst1) break;
        if (m >= n) m = n - 1;
        if (m != l) {
            for (;;) {
                if (iter >= 30) return l + 1;
                ++iter;
                const int l1 = l + 1, l2 = l1 + 1;
                double g = d[l];
                double p = (d[l1] - g) / (2.0 * e[l]);
                double r = hyp(p, 1.0);
                const double sr = p >= 0.0 ? std::fabs(r) : -std::fabs(r);
                d[l] = e[l] / (p + sr);
                d[l1] = e[l] * (p + sr);
                const double dl1 = d[l1];
                double h = g - d[l];
                for (int i = l2; i < n; ++i) d[i] -= h;
                f += h;
                p = d[m];
                double c = 1.0, c2 = 1.0, c3 = 1.0, s = 0.0, s2 = 0.0;
                const double el1 = e[l1];
                for (int i = m - 1; i >= l; --i) {
                    c3 = c2;
                    c2 = c;
                    s2 = s;
                    g = c * e[i];
                    h = c * p;
                    r = hyp(p, e[i]);
                    e[i + 1] = s * r;
                    s = e[i] / r;
                    c = p / r;
                    p = c * d[i] - s * g;
                    d[i + 1] = h + s * (c * g + s * d[i]);
                    for (int k = 0; k < nrow; ++k) {
                        double *zi = z + (size_t)i * ldz, *zi1 = z + (size_t)(i + 1) * ldz;
                        h = zi1[k];
                        zi1[k] = s * zi[k] + c * h;
                        zi[k] = c * zi[k] - s * h;
                    }
                }
                p = -s * s2 * c3 * el1 * e[l] / dl1;
                e[l] = s * p;
                d[l] = c * p;
                if (!(tst1 + std::fabs(e[l]) > tst1)) break;
            }
        }
        d[l] += f;
    }
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        double p = d[i];
        for (int j = i + 1; j < n; ++j)
            if (d[j] < p) { k = j; p = d[j]; }
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
            for (int r = 0; r < nrow; ++r) std::swap(z[(size_t)i * ldz + r], z[(size_t)k * ldz + r]);
        }
    }
    return 0;
}

int host_syev(int n, double *a, double *w)
{
    if (n <= 0) return 0;
    auto A = [&](int i, int j) -> double & { return a[(size_t)i + (size_t)j * n]; };
    std::vector<double> Q((size_t)n * n, 0.0), v(n), p(n), q(n), e(n, 0.0);
    for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
    for (int k = 0; k + 2 < n; ++k) {
        double sig = 0.0;
        for (int i = k + 1; i < n; ++i) sig += A(i, k) * A(i, k);
        const double x0 = A(k + 1, k);
        const double tail = sig - x0 * x0;
        if (tail == 0.0) continue;                       // column already reduced
        const double nrm = std::sqrt(sig);
        const double alpha = x0 > 0 ? -nrm : nrm;
        for (int i = 0; i < n; ++i) v[i] = 0.0;
        for (int i = k + 1; i < n; ++i) v[i] = A(i, k);
        v[k + 1] -= alpha;
        const double vn2 = tail + (x0 - alpha) * (x0 - alpha);
        const double c = 2.0 / vn2;
        // p = c B v over the trailing block
        double vp = 0.0;
        for (int i = k + 1; i < n; ++i) {
            double s = 0.0;
            for (int j = k + 1; j < n; ++j) s += A(i, j) * v[j];
            p[i] = c * s;
            vp += v[i] * p[i];
        }
        const double half = 0.5 * c * vp;
        for (int i = k + 1; i < n; ++i) q[i] = p[i] - half * v[i];
        for (int j = k + 1; j < n; ++j)
            for (int i = k + 1; i < n; ++i) A(i, j) -= v[i] * q[j] + q[i] * v[j];
        A(k + 1, k) = alpha;
        A(k, k + 1) = alpha;
        for (int i = k + 2; i < n; ++i) A(i, k) = A(k, i) = 0.0;
        // Q <- Q H
        for (int r = 0; r < n; ++r) {
            double s = 0.0;
            for (int i = k + 1; i < n; ++i) s += Q[(size_t)i * n + r] * v[i];
            s *= c;
            for (int i = k + 1; i < n; ++i) Q[(size_t)i * n + r] -= s * v[i];
        }
    }
    for (int i = 0; i < n; ++i) w[i] = A(i, i);
    for (int i = 1; i < n; ++i) e[i] = A(i, i - 1);
    const int rc = ql_implicit(n, w, e.data(), Q.data(), n, n);
    std::copy(Q.begin(), Q.end(), a);
    return rc == 0 ? 0 : fail(DMED_ENOCONV, "QL iteration did not converge");
}

int tql2_first_row(int n, double *d, double *e, double *z1)
{
    for (int j = 0; j < n; ++j) z1[j] = j == 0 ? 1.0 : 0.0;
    return ql_implicit(n, d, e, z1, 1, 1);
}

void gf_add_poles(double peso, double ei, const double *lam, const double *w2, int n, int isign,
                  const double *wm, int lmats, const double *wr, int lreal, double eps, double *gm,
                  double *gr)
{
    // pw / (z - isign*de) with the reciprocal written out:
    //   Matsubara z = i*wm:      (-s*de - i*wm) / (de^2 + wm^2)
    //   real axis z = w + i*eps: ((w - s*de) - i*eps) / ((w - s*de)^2 + eps^2)
    for (int j = 0; j < n; ++j) {
        const double sde = (double)isign * (lam[j] - ei);
        const double pw = peso * w2[j];
        if (pw == 0.0) continue;
        const double sde2 = sde * sde;
        for (int i = 0; i < lmats; ++i) {
            const double q = pw / (sde2 + wm[i] * wm[i]);
            gm[2 * i] -= sde * q;
            gm[2 * i + 1] -= wm[i] * q;
        }
        const double eps2 = eps * eps;
        for (int i = 0; i < lreal; ++i) {
            const double x = wr[i] - sde;
            const double q = pw / (x * x + eps2);
            gr[2 * i] += x * q;
            gr[2 * i + 1] -= eps * q;
        }
    }
}

// One branch (isign = +1 or -1) of the susceptibility pole sums of the
// reference: add_to_lanczos_spinChi (build_chi_spin.f90:186-250),
// add_to_lanczos_densChi, _densChi_mix and _densChi_tot
// (build_chi_dens.f90:505-733).  The weight peso_j = (pr + i pi)*w2_j*peso_bz
// is complex for the (N_a -/+ i N_b) chains of the orbital-off-diagonal
// charge susceptibility (cnorm2 = xi*norm0, build_chi_dens.f90:263, 300);
// chi_tau is real in the reference, so it takes Re(peso_j).  The iv = 0
// term is iv0_sign*(peso_j*beta if beta*dE < 0.1 else peso_j*(1-exp(-beta dE))/dE):
// the spin sums use iv0_sign = +1 for both branches, the charge sums
// iv0_sign = -isign (build_chi_dens.f90:539-542 against 560-563).
// Reciprocals written out: 1/(x + i y) = (x - i y)/(x^2 + y^2).
void chi_add_branch(double pr, double pi, double peso_bz, double ei, const double *lam,
                    const double *w2, int n, int isign, int iv0_sign, double beta,
                    const double *vm, int nv, const double *tau, int ntau, const double *wr,
                    int lreal, double eps, double *chi_iv, double *chi_tau, double *chi_w)
{
    for (int j = 0; j < n; ++j) {
        const double de = lam[j] - ei;
        const double wr_ = pr * w2[j] * peso_bz, wi_ = pi * w2[j] * peso_bz;
        const double ex = std::exp(-beta * de);
        if (nv > 0) {
            const double f = beta * de < 1e-1 ? beta : (1.0 - ex) / de;
            chi_iv[0] += iv0_sign * wr_ * f;
            chi_iv[1] += iv0_sign * wi_ * f;
        }
        // isign +1: peso*(ex - 1)/(z - de);  -1: peso*(1 - ex)/(z + de)
        const double num = isign > 0 ? ex - 1.0 : 1.0 - ex;
        const double sde = isign > 0 ? de : -de;
        for (int i = 1; i < nv; ++i) {
            // num/(i vm - sde) = num*(-sde - i vm)/(sde^2 + vm^2)
            const double g = num / (sde * sde + vm[i] * vm[i]);
            const double xr = -sde * g, xi = -vm[i] * g;
            chi_iv[2 * i] += wr_ * xr - wi_ * xi;
            chi_iv[2 * i + 1] += wr_ * xi + wi_ * xr;
        }
        for (int i = 0; i < ntau; ++i)
            chi_tau[i] += isign > 0 ? wr_ * std::exp(-tau[i] * de)
                                    : wr_ * std::exp(-(beta - tau[i]) * de);
        for (int i = 0; i < lreal; ++i) {
            // num/(wr + i eps - sde) = num*(x - i eps)/(x^2 + eps^2), x = wr - sde
            const double x = wr[i] - sde;
            const double g = num / (x * x + eps * eps);
            const double xr = x * g, xi = -eps * g;
            chi_w[2 * i] += wr_ * xr - wi_ * xi;
            chi_w[2 * i + 1] += wr_ * xi + wi_ * xr;
        }
    }
}

// add_to_lanczos_spinChi (build_chi_spin.f90:186-250).  S_z is hermitian,
// so the particle (isign = +1) and hole (isign = -1) terms come from one
// chain, added +1 then -1.
void chi_add_poles(double peso_f, double peso_bz, double ei, const double *lam, const double *w2,
                   int n, double beta, const double *vm, int nv, const double *tau, int ntau,
                   const double *wr, int lreal, double eps, double *chi_iv, double *chi_tau,
                   double *chi_w)
{
    for (int isign = 1; isign >= -1; isign -= 2)
        chi_add_branch(peso_f, 0.0, peso_bz, ei, lam, w2, n, isign, 1, beta, vm, nv, tau, ntau,
                       wr, lreal, eps, chi_iv, chi_tau, chi_w);
}

}  // namespace dmed
