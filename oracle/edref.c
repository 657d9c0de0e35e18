/*
 * oracle/edref.c -- CPU restatement of the reference DMFT_ED hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The GPU results are tested against this file,
 * and it is the CPU timing baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library (dmft-ed_amd/) never
 * links, loads or calls it.
 *
 * It restates, in plain C99 with complex(8) arithmetic, the algorithm of the
 * reference (hutuniu/dmft-ed v5.1.1, normal mode, normal bath):
 *   - combined 2*Ns-bit state code  m = iup + idw*2^Ns      ED_SETUP.f90:861-892
 *   - c / cdg with the (-1)^{#occupied below} sign          ED_SETUP.f90:1055-1081
 *   - bdecomp                                               ED_SETUP.f90:1209-1221
 *   - binomial via a rounded real(8) product                ED_SETUP.f90:1258-1275
 *   - binary_search over the sorted map (1-based, 0=absent) ED_SETUP.f90:1282-1299
 *   - on-the-fly H*v (directMatVec_cc + build_hxv.f90 + HxVimp/HxVint/
 *     HxVbath/HxVimp_bath, normal branches)                 ED_HAMILTONIAN_MATVEC.f90:289-351
 *   - plain Lanczos tridiagonalisation (sp_lanc_tridiag)    .repo/PLAIN_LANCZOS.f90:87-180
 *   - plain Lanczos ground state (lanczos_plain_c)          .repo/PLAIN_LANCZOS.f90:286-385
 *   - EISPACK tql2 / pythag                                 ED_GREENS_FUNCTIONS.f90:361-539
 *   - add_to_lanczos_gf_normal pole sum                     build_gf_normal.f90:442-488
 *   - GF start vectors c^+|psi>, c|psi>                     build_gf_normal.f90:95-160
 *   - Delta / G0^{-1} / Sigma (normal bath)                 ED_BATH_FUNCTIONS.f90:221-258,1784-1807;
 *                                                           get_sigma_normal.f90:1-38
 *   - dens / docc of a state                                ED_OBSERVABLES.f90:118-150
 *   - local observables and energies of a state             ED_OBSERVABLES.f90:74-150, 450-669
 *
 * Deviation (stated): state codes are int64, so Ns=16 does not overflow as the
 * reference's integer(4) code does (ED_SETUP.f90:889).
 *
 * Parity pinning: tests/golden/make_golden.py builds an independent
 * Jordan-Wigner Fock-space Hamiltonian in numpy (different code path, no
 * shared formulae) and the closed forms (U=0, V=0 atomic limit, Toeplitz tql2);
 * tests/test_oracle.py checks every function here against them.
 *
 * Threading: ref_hxv_omp splits rows into contiguous blocks, mirroring the
 * reference's MPI row partition (ED_HAMILTONIAN_MATVEC.f90:411-416).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zdouble;

/* ---------------------------------------------------------------- model */
/* Module state, as the reference keeps it in ED_INPUT_VARS / ED_VARS_GLOBAL. */
static int g_norb = 1, g_nbath = 0, g_nspin = 1, g_ns = 1;
static double g_e[2][3][64];     /* dmft_bath%e(ispin,iorb,k) */
static double g_v[2][3][64];     /* dmft_bath%v(ispin,iorb,k) */
static double g_hloc[2][2][3][3];/* real(impHloc(ispin,jspin,iorb,jorb)) */
static double g_uloc[3], g_ust, g_jh, g_jx, g_jp, g_xmu;
static int g_hfmode = 1, g_jhflag = 0;
static int g_hybrid = 0;          /* bath_type: 0 normal, 1 hybrid */

/* bath: flat array in the set_dmft_bath layout (dmft_aux.f90:489-512):
 *   e(ispin,iorb,i) at io = i + (iorb-1)*Nbath + (ispin-1)*Nbath*Norb
 *   v(...)          at io + Nspin*Norb*Nbath
 * hloc: Fortran-ordered real array Hloc(Nspin,Nspin,Norb,Norb). */
/* bath_type of the next ref_set_model: 0 normal, 1 hybrid (ED_INPUT_VARS bath_type). */
void ref_set_bath_type(int hybrid) { g_hybrid = hybrid ? 1 : 0; }

int ref_set_model(int norb, int nbath, int nspin, const double *bath,
                  const double *hloc_f, const double *uloc, double ust,
                  double jh, double jx, double jp, double xmu, int hfmode,
                  int jhflag)
{
    if (norb < 1 || norb > 3 || nspin < 1 || nspin > 2 || nbath < 0 || nbath > 64)
        return -1;
    g_norb = norb; g_nbath = nbath; g_nspin = nspin;
    /* setup_ed_dimensions, ED_SETUP.f90:86-88 */
    g_ns = g_hybrid ? nbath + norb : (nbath + 1) * norb;
    if (g_ns > 30) return -2;
    memset(g_e, 0, sizeof g_e);
    memset(g_v, 0, sizeof g_v);
    if (g_hybrid) {
        /* hybrid layout (dmft_aux.f90:572-590): e(ispin,1,i) at i + (ispin-1)*Nbath,
         * v(ispin,iorb,i) at Nspin*Nbath + i + (iorb-1)*Nbath + (ispin-1)*Norb*Nbath */
        for (int s = 0; s < nspin; ++s) {
            for (int k = 0; k < nbath; ++k) g_e[s][0][k] = bath[k + s * nbath];
            for (int a = 0; a < norb; ++a)
                for (int k = 0; k < nbath; ++k)
                    g_v[s][a][k] = bath[nspin * nbath + k + a * nbath + s * norb * nbath];
        }
    } else {
        int stride = nspin * norb * nbath;
        for (int s = 0; s < nspin; ++s)
            for (int a = 0; a < norb; ++a)
                for (int k = 0; k < nbath; ++k) {
                    int io = k + a * nbath + s * nbath * norb;
                    g_e[s][a][k] = bath[io];
                    g_v[s][a][k] = bath[stride + io];
                }
    }
    memset(g_hloc, 0, sizeof g_hloc);
    for (int s1 = 0; s1 < nspin; ++s1)
        for (int s2 = 0; s2 < nspin; ++s2)
            for (int a = 0; a < norb; ++a)
                for (int b = 0; b < norb; ++b)
                    g_hloc[s1][s2][a][b] =
                        hloc_f[s1 + nspin * (s2 + nspin * (a + norb * b))];
    for (int a = 0; a < 3; ++a) g_uloc[a] = uloc[a];
    g_ust = ust; g_jh = jh; g_jx = jx; g_jp = jp; g_xmu = xmu;
    g_hfmode = hfmode; g_jhflag = jhflag;
    return 0;
}

int ref_ns(void) { return g_ns; }

/* ------------------------------------------------------ Hilbert space */
/* binomial: ED_SETUP.f90:1258-1275 (real(8) running product, rounded). */
long long ref_binomial(int n1, int n2)
{
    if (n2 < 0) return 0;
    if (n2 == 0) return 1;
    double xh = 1.0;
    for (int i = 1; i <= n2; ++i) xh = xh * (double)(n1 + 1 - i) / (double)i;
    return (long long)(xh + 0.5);
}

long long ref_sector_dim(int nup, int ndw)
{
    return ref_binomial(g_ns, nup) * ref_binomial(g_ns, ndw);
}

static int popc64(uint64_t x) { return __builtin_popcountll(x); }

/* build_sector, normal case (ED_SETUP.f90:872-889): idw ascending (outer),
 * iup ascending (inner), code = iup + idw*2^Ns. */
long long ref_build_sector(int nup, int ndw, int64_t *map)
{
    long long dim = 0;
    uint64_t top = 1ull << g_ns;
    for (uint64_t idw = 0; idw < top; ++idw) {
        if (popc64(idw) != ndw) continue;
        for (uint64_t iup = 0; iup < top; ++iup) {
            if (popc64(iup) != nup) continue;
            if (map) map[dim] = (int64_t)(iup + (idw << g_ns));
            ++dim;
        }
    }
    return dim;
}

/* c / cdg: ED_SETUP.f90:1055-1081.  pos is 1-based.  Returns 0, or -1 on a
 * Pauli violation (the reference stops). */
int ref_c(int pos, int64_t in, int64_t *out, double *fsgn)
{
    if (!((in >> (pos - 1)) & 1)) return -1;
    double s = 1.0;
    for (int l = 1; l <= pos - 1; ++l)
        if ((in >> (l - 1)) & 1) s = -s;
    *fsgn = s;
    *out = in & ~((int64_t)1 << (pos - 1));
    return 0;
}

int ref_cdg(int pos, int64_t in, int64_t *out, double *fsgn)
{
    if ((in >> (pos - 1)) & 1) return -1;
    double s = 1.0;
    for (int l = 1; l <= pos - 1; ++l)
        if ((in >> (l - 1)) & 1) s = -s;
    *fsgn = s;
    *out = in | ((int64_t)1 << (pos - 1));
    return 0;
}

/* bdecomp: ED_SETUP.f90:1209-1221, ivec(l+1) = bit l of i. */
void ref_bdecomp(int64_t i, int ntot, int *ivec)
{
    for (int l = 0; l < ntot; ++l) ivec[l] = (int)((i >> l) & 1);
}

/* binary_search: ED_SETUP.f90:1282-1299.  The recursion halves a(:) at
 * mid = size/2+1 (1-based); written here as the equivalent loop over the
 * same midpoints.  Returns the 1-based index, 0 when absent. */
long long ref_binary_search(const int64_t *a, long long n, int64_t value)
{
    long long lo = 0, len = n;   /* current slice a[lo .. lo+len-1] */
    while (len > 0) {
        long long mid = len / 2;      /* 0-based offset of mid = size/2+1 */
        int64_t am = a[lo + mid];
        if (am > value) {
            len = mid;                  /* a(:mid-1) */
        } else if (am < value) {
            lo = lo + mid + 1;          /* a(mid+1:) */
            len = len - mid - 1;
        } else {
            return lo + mid + 1;
        }
    }
    return 0;
}

/* getBathStride(iorb,k) = Norb + (iorb-1)*Nbath + k (1-based), ED_SETUP.f90:447-451 */
/* getBathStride (ED_SETUP.f90:446-455), 1-based site of bath level k0 of
 * orbital a0: Norb + (a0)*Nbath + k0 + 1 (normal), Norb + k0 + 1 (hybrid) */
static inline int bath_stride(int a0, int k0)
{
    return g_hybrid ? g_norb + k0 + 1 : g_norb + a0 * g_nbath + k0 + 1;
}

/* ----------------------------------------------------------- H * v */
static int g_abort_flag = 0;

/* Rows [first, last) of the product; row i is stored at hv[i - hoff]. */
static void hxv_rows(const int64_t *map, long long dim, const zdouble *vin,
                     zdouble *hv, long long first, long long last, long long hoff)
{
    const int Ns = g_ns, Norb = g_norb, Nbath = g_nbath, Nsp = g_nspin - 1;
    int ib[64];
    double nup[3], ndw[3];
    for (long long i = first; i < last; ++i) {
        int64_t m = map[i];
        ref_bdecomp(m, 2 * Ns, ib);
        for (int a = 0; a < Norb; ++a) { nup[a] = ib[a]; ndw[a] = ib[a + Ns]; }
        zdouble acc = 0;
        /* HxVimp.f90:2-10 -- diagonal local part */
        double htmp = 0.0, sn = 0.0;
        for (int a = 0; a < Norb; ++a) sn += nup[a] + ndw[a];
        htmp -= g_xmu * sn;
        for (int a = 0; a < Norb; ++a) {
            htmp += g_hloc[0][0][a][a] * nup[a];
            htmp += g_hloc[Nsp][Nsp][a][a] * ndw[a];
        }
        acc += htmp * vin[i];
        /* HxVimp.f90:15-49 -- intra-spin inter-orbital Hloc hops */
        for (int a = 0; a < Norb; ++a)
            for (int b = 0; b < Norb; ++b) {
                int64_t k1, k2; double sg1, sg2; long long j;
                if (g_hloc[0][0][a][b] != 0.0 && ib[b] == 1 && ib[a] == 0) {
                    ref_c(b + 1, m, &k1, &sg1); ref_cdg(a + 1, k1, &k2, &sg2);
                    j = ref_binary_search(map, dim, k2);
                    if (j == 0) { g_abort_flag = 1; continue; }
                    acc += g_hloc[0][0][a][b] * sg1 * sg2 * vin[j - 1];
                }
                if (g_hloc[Nsp][Nsp][a][b] != 0.0 && ib[b + Ns] == 1 && ib[a + Ns] == 0) {
                    ref_c(b + 1 + Ns, m, &k1, &sg1); ref_cdg(a + 1 + Ns, k1, &k2, &sg2);
                    j = ref_binary_search(map, dim, k2);
                    if (j == 0) { g_abort_flag = 1; continue; }
                    acc += g_hloc[Nsp][Nsp][a][b] * sg1 * sg2 * vin[j - 1];
                }
            }
        /* HxVint.f90:1-43 -- density-density interaction (+ Hartree shifts) */
        htmp = 0.0;
        for (int a = 0; a < Norb; ++a) htmp += g_uloc[a] * nup[a] * ndw[a];
        if (Norb > 1) {
            for (int a = 0; a < Norb; ++a)
                for (int b = a + 1; b < Norb; ++b)
                    htmp += g_ust * (nup[a] * ndw[b] + nup[b] * ndw[a]);
            for (int a = 0; a < Norb; ++a)
                for (int b = a + 1; b < Norb; ++b)
                    htmp += (g_ust - g_jh) * (nup[a] * nup[b] + ndw[a] * ndw[b]);
        }
        if (g_hfmode) {
            for (int a = 0; a < Norb; ++a)
                htmp = htmp - 0.5 * g_uloc[a] * (nup[a] + ndw[a]) + 0.25 * g_uloc[a];
            if (Norb > 1)
                for (int a = 0; a < Norb; ++a)
                    for (int b = a + 1; b < Norb; ++b) {
                        htmp = htmp - 0.5 * g_ust * (nup[a] + ndw[a] + nup[b] + ndw[b]) + 0.25 * g_ust;
                        htmp = htmp - 0.5 * (g_ust - g_jh) * (nup[a] + ndw[a] + nup[b] + ndw[b]) + 0.25 * (g_ust - g_jh);
                    }
        }
        acc += htmp * vin[i];
        /* HxVint.f90:48-97 -- spin-exchange and pair-hopping (Jhflag, Norb>1) */
        if (Norb > 1 && g_jhflag) {
            for (int a = 0; a < Norb; ++a)
                for (int b = 0; b < Norb; ++b) {
                    int64_t k1, k2, k3, k4; double s1, s2, s3, s4; long long j;
                    if (a != b && ib[b] == 1 && ib[a + Ns] == 1 && ib[b + Ns] == 0 && ib[a] == 0) {
                        ref_c(b + 1, m, &k1, &s1); ref_c(a + 1 + Ns, k1, &k2, &s2);
                        ref_cdg(b + 1 + Ns, k2, &k3, &s3); ref_cdg(a + 1, k3, &k4, &s4);
                        j = ref_binary_search(map, dim, k4);
                        if (j != 0) acc += g_jx * s1 * s2 * s3 * s4 * vin[j - 1];
                    }
                }
            for (int a = 0; a < Norb; ++a)
                for (int b = 0; b < Norb; ++b) {
                    int64_t k1, k2, k3, k4; double s1, s2, s3, s4; long long j;
                    if (a != b && ib[b] == 1 && ib[b + Ns] == 1 && ib[a + Ns] == 0 && ib[a] == 0) {
                        ref_c(b + 1, m, &k1, &s1); ref_c(b + 1 + Ns, k1, &k2, &s2);
                        ref_cdg(a + 1 + Ns, k2, &k3, &s3); ref_cdg(a + 1, k3, &k4, &s4);
                        j = ref_binary_search(map, dim, k4);
                        if (j != 0) acc += g_jp * s1 * s2 * s3 * s4 * vin[j - 1];
                    }
                }
        }
        /* HxVbath.f90:1-14 -- diagonal bath energies; the orbital loop runs
         * over size(dmft_bath%e,2), which is 1 for the hybrid bath */
        htmp = 0.0;
        for (int a = 0; a < (g_hybrid ? 1 : Norb); ++a)
            for (int k = 0; k < Nbath; ++k) {
                int alfa = bath_stride(a, k);
                htmp += g_e[0][a][k] * ib[alfa - 1];
                htmp += g_e[Nsp][a][k] * ib[alfa - 1 + Ns];
            }
        acc += htmp * vin[i];
        /* HxVimp_bath.f90:1-38 -- spin-conserving hybridisation */
        for (int a = 0; a < Norb; ++a)
            for (int k = 0; k < Nbath; ++k) {
                int ms = bath_stride(a, k);
                int64_t k1, k2; double sg1, sg2; long long j;
                double vu = g_v[0][a][k], vd = g_v[Nsp][a][k];
                if (vu != 0.0 && ib[a] == 1 && ib[ms - 1] == 0) {
                    ref_c(a + 1, m, &k1, &sg1); ref_cdg(ms, k1, &k2, &sg2);
                    j = ref_binary_search(map, dim, k2);
                    if (j == 0) { g_abort_flag = 1; continue; }
                    acc += vu * sg1 * sg2 * vin[j - 1];
                }
                if (vu != 0.0 && ib[a] == 0 && ib[ms - 1] == 1) {
                    ref_c(ms, m, &k1, &sg1); ref_cdg(a + 1, k1, &k2, &sg2);
                    j = ref_binary_search(map, dim, k2);
                    if (j == 0) { g_abort_flag = 1; continue; }
                    acc += vu * sg1 * sg2 * vin[j - 1];
                }
                if (vd != 0.0 && ib[a + Ns] == 1 && ib[ms - 1 + Ns] == 0) {
                    ref_c(a + 1 + Ns, m, &k1, &sg1); ref_cdg(ms + Ns, k1, &k2, &sg2);
                    j = ref_binary_search(map, dim, k2);
                    if (j == 0) { g_abort_flag = 1; continue; }
                    acc += vd * sg1 * sg2 * vin[j - 1];
                }
                if (vd != 0.0 && ib[a + Ns] == 0 && ib[ms - 1 + Ns] == 1) {
                    ref_c(ms + Ns, m, &k1, &sg1); ref_cdg(a + 1 + Ns, k1, &k2, &sg2);
                    j = ref_binary_search(map, dim, k2);
                    if (j == 0) { g_abort_flag = 1; continue; }
                    acc += vd * sg1 * sg2 * vin[j - 1];
                }
            }
        hv[i - hoff] = acc;
    }
}

/* directMatVec_cc (ED_HAMILTONIAN_MATVEC.f90:289-351), one thread.  vin/hv
 * are interleaved (re,im) double pairs.  Returns 0, or -3 when a hop target
 * is missing from the map (the reference reads vin(0) there). */
int ref_hxv(const int64_t *map, long long dim, const double *vin, double *hv)
{
    g_abort_flag = 0;
    hxv_rows(map, dim, (const zdouble *)vin, (zdouble *)hv, 0, dim, 0);
    return g_abort_flag ? -3 : 0;
}

/* Same product with nthreads contiguous row blocks: rank r owns rows
 * [r*Q, (r+1)*Q + R) with R the remainder on the last rank, as the MPI
 * partition at ED_HAMILTONIAN_MATVEC.f90:411-416. */
int ref_hxv_omp(const int64_t *map, long long dim, const double *vin,
                double *hv, int nthreads)
{
    g_abort_flag = 0;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        int r = omp_get_thread_num(), p = omp_get_num_threads();
        long long q = dim / p, rem = (r == p - 1) ? dim % p : 0;
        hxv_rows(map, dim, (const zdouble *)vin, (zdouble *)hv, r * q,
                 (r + 1) * q + rem, 0);
    }
#else
    (void)nthreads;
    hxv_rows(map, dim, (const zdouble *)vin, (zdouble *)hv, 0, dim, 0);
#endif
    return g_abort_flag ? -3 : 0;
}

/* Rows [first, last) of the same product with nthreads contiguous blocks,
 * stored at out[0 .. last-first): a bounded block of a large sector (the CPU
 * timing sample, and parity checks of single columns or row ranges). */
int ref_hxv_block(const int64_t *map, long long dim, const double *vin, double *out,
                  long long first, long long last, int nthreads)
{
    g_abort_flag = 0;
    if (first < 0 || last > dim || first > last) return -1;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        int r = omp_get_thread_num(), p = omp_get_num_threads();
        long long n = last - first, q = n / p, rem = (r == p - 1) ? n % p : 0;
        hxv_rows(map, dim, (const zdouble *)vin, (zdouble *)out, first + r * q,
                 first + (r + 1) * q + rem, first);
    }
#else
    (void)nthreads;
    hxv_rows(map, dim, (const zdouble *)vin, (zdouble *)out, first, last, first);
#endif
    return g_abort_flag ? -3 : 0;
}

/* ------------------------------------------------------ tql2 / pythag */
/* pythag: ED_GREENS_FUNCTIONS.f90:513-539 (Moler-Morrison iteration). */
double ref_pythag(double a, double b)
{
    double p = fmax(fabs(a), fabs(b));
    if (isnan(a) || isnan(b)) return NAN;   /* the iteration below never ends on NaN */
    if (p != 0.0) {
        double r = fmin(fabs(a), fabs(b)) / p;
        r = r * r;
        for (;;) {
            double t = 4.0 + r;
            if (t == 4.0) break;
            double s = r / t;
            double u = 1.0 + 2.0 * s;
            p = u * p;
            r = (s / u) * (s / u) * r;
        }
    }
    return p;
}

/* tql2 (ED_GREENS_FUNCTIONS.f90:361-499): implicit QL on a symmetric
 * tridiagonal.  d[n] diagonal, e[n] with the sub-diagonal in e[1..n-1]
 * (e[0] ignored, EISPACK convention), z[n*n] column-major, updated in place.
 * Returns ierr (0, or l when eigenvalue l failed to converge in 30 sweeps). */
int ref_tql2(int n, double *d, double *e, double *z)
{
#define Z(i, j) z[(i) + (size_t)(j) * n]
    if (n == 1) return 0;
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    double f = 0.0, tst1 = 0.0;
    e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int j = 0;
        double h = fabs(d[l]) + fabs(e[l]);
        tst1 = fmax(tst1, h);
        int m;
        for (m = l; m < n; ++m) {
            double tst2 = tst1 + fabs(e[m]);
            if (tst2 == tst1) break;
        }
        if (m == n) m = n - 1; /* Fortran do-loop exit value when no break: e(n)=0 always breaks */
        if (m != l) {
            for (;;) {
                if (j >= 30) return l + 1;
                ++j;
                int l1 = l + 1, l2 = l1 + 1;
                double g = d[l];
                double p = (d[l1] - g) / (2.0 * e[l]);
                double r = ref_pythag(p, 1.0);
                double sr = (p >= 0.0) ? fabs(r) : -fabs(r);
                d[l] = e[l] / (p + sr);
                d[l1] = e[l] * (p + sr);
                double dl1 = d[l1];
                h = g - d[l];
                for (int i = l2; i < n; ++i) d[i] -= h;
                f += h;
                p = d[m];
                double c = 1.0, c2 = c, c3 = c, el1 = e[l1], s = 0.0, s2 = 0.0;
                int mml = m - l;
                for (int ii = 1; ii <= mml; ++ii) {
                    c3 = c2; c2 = c; s2 = s;
                    int i = m - ii;
                    g = c * e[i];
                    h = c * p;
                    r = ref_pythag(p, e[i]);
                    e[i + 1] = s * r;
                    s = e[i] / r;
                    c = p / r;
                    p = c * d[i] - s * g;
                    d[i + 1] = h + s * (c * g + s * d[i]);
                    for (int k = 0; k < n; ++k) {
                        h = Z(k, i + 1);
                        Z(k, i + 1) = s * Z(k, i) + c * h;
                        Z(k, i) = c * Z(k, i) - s * h;
                    }
                }
                p = -s * s2 * c3 * el1 * e[l] / dl1;
                e[l] = s * p;
                d[l] = c * p;
                double tst2 = tst1 + fabs(e[l]);
                if (!(tst2 > tst1)) break;
            }
        }
        d[l] = d[l] + f;
    }
    for (int ii = 1; ii < n; ++ii) {
        int i = ii - 1, k = i;
        double p = d[i];
        for (int j = ii; j < n; ++j)
            if (d[j] < p) { k = j; p = d[j]; }
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
            for (int j = 0; j < n; ++j) {
                double t = Z(j, i); Z(j, i) = Z(j, k); Z(j, k) = t;
            }
        }
    }
    return 0;
#undef Z
}

/* ------------------------------------------------------------ Lanczos */
static zdouble zdot(long long n, const zdouble *a, const zdouble *b)
{
    /* Fortran dot_product(a,b) = sum(conjg(a)*b) */
    zdouble s = 0;
    for (long long i = 0; i < n; ++i) s += conj(a[i]) * b[i];
    return s;
}

/* lanczos_plain_iteration_c (.repo/PLAIN_LANCZOS.f90:87-118). */
static int lanczos_iter(int iter, const int64_t *map, long long dim, zdouble *vin,
                        zdouble *vout, zdouble *tmp, double *a, double *b,
                        int nthreads)
{
    if (iter == 1) {
        double norm = sqrt(creal(zdot(dim, vin, vin)));
        if (norm == 0.0) return -4;
        for (long long i = 0; i < dim; ++i) vin[i] /= norm;
        *b = 0.0;
    }
    int rc = ref_hxv_omp(map, dim, (const double *)vin, (double *)tmp, nthreads);
    if (rc) return rc;
    for (long long i = 0; i < dim; ++i) tmp[i] -= (*b) * vout[i];
    *a = creal(zdot(dim, vin, tmp));
    for (long long i = 0; i < dim; ++i) tmp[i] -= (*a) * vin[i];
    *b = sqrt(creal(zdot(dim, tmp, tmp)));
    for (long long i = 0; i < dim; ++i) { vout[i] = vin[i]; vin[i] = tmp[i] / (*b); }
    return 0;
}

/* sp_lanc_tridiag semantics (.repo/PLAIN_LANCZOS.f90:154-180): alanc(k)=a_k,
 * blanc(k+1)=b_k for k<nitermax; exit when |b|<threshold.  vin (interleaved
 * complex, length dim) is overwritten.  alpha/beta are zero-padded after an
 * early exit; *nsteps receives the number of recorded alphas. */
int ref_lanc_tridiag(const int64_t *map, long long dim, double *vin, int nitermax,
                     double *alanc, double *blanc, double threshold,
                     int nthreads, int *nsteps)
{
    zdouble *v = (zdouble *)vin;
    zdouble *vout = calloc(dim, sizeof(zdouble));
    zdouble *tmp = calloc(dim, sizeof(zdouble));
    if (!vout || !tmp) { free(vout); free(tmp); return -5; }
    for (int k = 0; k < nitermax; ++k) { alanc[k] = 0.0; blanc[k] = 0.0; }
    double a = 0.0, b = 0.0;
    int rc = 0, iter;
    for (iter = 1; iter <= nitermax; ++iter) {
        rc = lanczos_iter(iter, map, dim, v, vout, tmp, &a, &b, nthreads);
        if (rc) break;
        alanc[iter - 1] = a;
        if (iter < nitermax) blanc[iter] = b;
        if (fabs(b) < threshold) { ++iter; break; }
    }
    if (nsteps) *nsteps = (iter > nitermax ? nitermax : iter - 1);
    free(vout); free(tmp);
    return rc;
}

/* lanczos_plain_c (.repo/PLAIN_LANCZOS.f90:286-385): ground-state energy by
 * plain Lanczos with the ncheck stop rule, then the eigenvector by re-running
 * the recurrence and accumulating Z(k,1)*v_k.  The historical snapshot adds
 * the vector *after* each step (v_{k+1}, :381); this restatement adds v_k,
 * the Ritz vector that sp_eigh returns.  vect (interleaved complex) is
 * the start vector on entry and the normalised ground state on exit.
 * egs receives the nlanc Ritz values.  Returns nlanc (>0) or an error (<0). */
int ref_lanczos_gs(const int64_t *map, long long dim, double *vect, int nitermax,
                   double threshold, int ncheck, double *egs, int nthreads)
{
    zdouble *vv = (zdouble *)vect;
    zdouble *vin = malloc(dim * sizeof(zdouble));
    zdouble *vout = calloc(dim, sizeof(zdouble));
    zdouble *tmp = calloc(dim, sizeof(zdouble));
    double *al = calloc(nitermax + 1, sizeof(double));
    double *bl = calloc(nitermax + 2, sizeof(double));
    double *dg = calloc(nitermax, sizeof(double));
    double *sd = calloc(nitermax, sizeof(double));
    double *Z = calloc((size_t)nitermax * nitermax, sizeof(double));
    double *esave = calloc(nitermax, sizeof(double));
    int rc = 0, nlanc = 0;
    if (!vin || !vout || !tmp || !al || !bl || !dg || !sd || !Z || !esave) { rc = -5; goto out; }
    memcpy(vin, vv, dim * sizeof(zdouble));
    double a = 0, b = 0;
    for (int iter = 1; iter <= nitermax; ++iter) {
        rc = lanczos_iter(iter, map, dim, vin, vout, tmp, &a, &b, nthreads);
        if (rc) goto out;
        if (fabs(b) < threshold) break;
        nlanc++;
        al[iter] = a; bl[iter + 1] = b;   /* 1-based alanc(iter), blanc(iter+1) */
        for (int i = 0; i < nlanc; ++i) { dg[i] = al[i + 1]; sd[i] = 0.0; }
        for (int i = 1; i < nlanc; ++i) sd[i] = bl[i + 1];
        memset(Z, 0, (size_t)nlanc * nlanc * sizeof(double));
        for (int i = 0; i < nlanc; ++i) Z[i + (size_t)i * nlanc] = 1.0;
        ref_tql2(nlanc, dg, sd, Z);
        if (nlanc >= ncheck) {
            esave[nlanc - ncheck] = dg[0];
            if (nlanc >= ncheck + 1) {
                double diff = esave[nlanc - ncheck] - esave[nlanc - ncheck - 1];
                if (fabs(diff) <= threshold) break;
            }
        }
    }
    for (int i = 0; i < nlanc; ++i) { dg[i] = al[i + 1]; sd[i] = 0.0; }
    for (int i = 1; i < nlanc; ++i) sd[i] = bl[i + 1];
    memset(Z, 0, (size_t)nlanc * nlanc * sizeof(double));
    for (int i = 0; i < nlanc; ++i) Z[i + (size_t)i * nlanc] = 1.0;
    ref_tql2(nlanc, dg, sd, Z);
    for (int i = 0; i < nlanc; ++i) egs[i] = dg[i];
    /* eigenvector: second pass of the recurrence */
    memcpy(vin, vv, dim * sizeof(zdouble));
    memset(vout, 0, dim * sizeof(zdouble));
    memset(vv, 0, dim * sizeof(zdouble));
    for (int iter = 1; iter <= nlanc; ++iter) {
        double aa = al[iter], bb = bl[iter];
        /* the second pass hands the stored (a,b) in and lets the iteration
         * recompute them, exactly as the reference does */
        rc = lanczos_iter(iter, map, dim, vin, vout, tmp, &aa, &bb, nthreads);
        if (rc) goto out;
        /* after lanczos_iter, vout holds v_iter (the vector the step used) */
        for (long long i = 0; i < dim; ++i) vv[i] += vout[i] * Z[(iter - 1)];
    }
    {
        double norm = sqrt(creal(zdot(dim, vv, vv)));
        for (long long i = 0; i < dim; ++i) vv[i] /= norm;
    }
    rc = nlanc;
out:
    free(vin); free(vout); free(tmp); free(al); free(bl); free(dg); free(sd);
    free(Z); free(esave);
    return rc;
}

/* ---------------------------------------------------- Green's function */
/* GF start vector (build_gf_normal.f90:105-113 / 144-152): for every basis
 * state m of sector I with the site empty (dagger=1) or full (dagger=0),
 * vvinit(j) = sgn*psi(m) with j = binary_search(HJ%map, r).  isite is the
 * 1-based spin-orbital index impIndex(iorb,ispin) = iorb + (ispin-1)*Ns.
 * psi/out are interleaved complex.  Returns 0, -3 when a target is absent. */
int ref_apply_op(const int64_t *mapI, long long dimI, const double *psi,
                 const int64_t *mapJ, long long dimJ, int isite, int dagger,
                 double *out)
{
    const zdouble *p = (const zdouble *)psi;
    zdouble *o = (zdouble *)out;
    for (long long j = 0; j < dimJ; ++j) o[j] = 0;
    for (long long m = 0; m < dimI; ++m) {
        int64_t i = mapI[m], r; double sgn;
        int occ = (int)((i >> (isite - 1)) & 1);
        if (dagger && occ == 0) {
            ref_cdg(isite, i, &r, &sgn);
        } else if (!dagger && occ == 1) {
            ref_c(isite, i, &r, &sgn);
        } else continue;
        long long j = ref_binary_search(mapJ, dimJ, r);
        if (j == 0) return -3;
        o[j - 1] = sgn * p[m];
    }
    return 0;
}

/* add_to_lanczos_gf_normal (build_gf_normal.f90:442-488).  peso_bz is the
 * already-formed Boltzmann weight vnorm2*exp(-beta(Ei-Egs))/Z.  alanc[n],
 * blanc[n] as returned by ref_lanc_tridiag (blanc[0] unused).  gmats/greal
 * are interleaved complex accumulators of length lmats / lreal. */
int ref_add_to_gf(double peso_bz, double ei, const double *alanc,
                  const double *blanc, int nlanc, int isign, const double *wm,
                  int lmats, const double *wr, int lreal, double eps,
                  double *gmats, double *greal)
{
    double *dg = malloc(nlanc * sizeof(double));
    double *sd = malloc(nlanc * sizeof(double));
    double *Z = calloc((size_t)nlanc * nlanc, sizeof(double));
    if (!dg || !sd || !Z) { free(dg); free(sd); free(Z); return -5; }
    for (int i = 0; i < nlanc; ++i) { dg[i] = alanc[i]; sd[i] = 0.0; Z[i + (size_t)i * nlanc] = 1.0; }
    for (int i = 1; i < nlanc; ++i) sd[i] = blanc[i];
    int ierr = ref_tql2(nlanc, dg, sd, Z);
    zdouble *gm = (zdouble *)gmats, *gr = (zdouble *)greal;
    for (int j = 0; j < nlanc; ++j) {
        double de = dg[j] - ei;
        double peso = peso_bz * Z[(size_t)j * nlanc] * Z[(size_t)j * nlanc];
        for (int i = 0; i < lmats; ++i) gm[i] += peso / (I * wm[i] - isign * de);
        for (int i = 0; i < lreal; ++i) gr[i] += peso / ((wr[i] + I * eps) - isign * de);
    }
    free(dg); free(sd); free(Z);
    return ierr;
}

/* ------------------------------------------------------ bath functions */
/* delta_bath (normal): Delta_aa(z) = sum_k V_k^2/(z - e_k)   ED_BATH_FUNCTIONS.f90:248-256
 * invg0: z + xmu - Hloc_aa - Delta_aa                         ED_BATH_FUNCTIONS.f90:1798-1803
 * z is interleaved complex of length L; outputs interleaved complex. */
void ref_invg0(int ispin, int iorb, const double *z, int L, double *out)
{
    const zdouble *zz = (const zdouble *)z;
    zdouble *o = (zdouble *)out;
    for (int i = 0; i < L; ++i) {
        zdouble delta = 0;
        for (int k = 0; k < g_nbath; ++k)
            delta += g_v[ispin][iorb][k] * g_v[ispin][iorb][k] / (zz[i] - g_e[ispin][iorb][k]);
        o[i] = zz[i] + g_xmu - g_hloc[ispin][ispin][iorb][iorb] - delta;
    }
}

void ref_delta(int ispin, int iorb, const double *z, int L, double *out)
{
    const zdouble *zz = (const zdouble *)z;
    zdouble *o = (zdouble *)out;
    for (int i = 0; i < L; ++i) {
        zdouble delta = 0;
        for (int k = 0; k < g_nbath; ++k)
            delta += g_v[ispin][iorb][k] * g_v[ispin][iorb][k] / (zz[i] - g_e[ispin][iorb][k]);
        o[i] = delta;
    }
}

/* delta_bath_mats, any bath: Delta_ab(z) = sum_k V_a(k) V_b(k)/(z - e(k)),
 * e(k) of the hybrid bath shared by the orbitals (ED_BATH_FUNCTIONS.f90:
 * 294-310), of orbital a for the normal bath (zero for a != b). */
void ref_delta_ab(int ispin, int a, int b, const double *z, int L, double *out)
{
    const zdouble *zz = (const zdouble *)z;
    zdouble *o = (zdouble *)out;
    for (int i = 0; i < L; ++i) {
        zdouble delta = 0;
        if (g_hybrid)
            for (int k = 0; k < g_nbath; ++k)
                delta += g_v[ispin][a][k] * g_v[ispin][b][k] / (zz[i] - g_e[ispin][0][k]);
        else if (a == b)
            for (int k = 0; k < g_nbath; ++k)
                delta += g_v[ispin][a][k] * g_v[ispin][a][k] / (zz[i] - g_e[ispin][a][k]);
        o[i] = delta;
    }
}

/* Sigma = G0^{-1} - G^{-1}   (get_sigma_normal.f90:27-38) */
void ref_sigma(int ispin, int iorb, const double *z, const double *g, int L, double *out)
{
    ref_invg0(ispin, iorb, z, L, out);
    zdouble *o = (zdouble *)out;
    const zdouble *gg = (const zdouble *)g;
    for (int i = 0; i < L; ++i) o[i] = o[i] - 1.0 / gg[i];
}

/* dens / docc contribution of one state (ED_OBSERVABLES.f90:118-150):
 * dens(a) += peso*|psi|^2*(nup+ndw), docc(a) += peso*|psi|^2*nup*ndw. */
void ref_observables(const int64_t *map, long long dim, const double *psi,
                     double peso, double *dens, double *docc)
{
    const zdouble *p = (const zdouble *)psi;
    for (long long i = 0; i < dim; ++i) {
        double w = peso * (creal(p[i]) * creal(p[i]) + cimag(p[i]) * cimag(p[i]));
        for (int a = 0; a < g_norb; ++a) {
            double nu = (double)((map[i] >> a) & 1);
            double nd = (double)((map[i] >> (a + g_ns)) & 1);
            dens[a] += (nu + nd) * w;
            docc[a] += nu * nd * w;
        }
    }
}

/* Local observables and energies of one state, accumulated into out[42]
 * (observables_impurity, ED_OBSERVABLES.f90:74-150, normal mode, and
 * local_energy_impurity, :450-669).  Layout (Norb <= 3, [a] = orbital,
 * [a*3+b] = orbital pair):
 *   0 dens[3]  3 dens_up[3]  6 dens_dw[3]  9 docc[3]  12 magz[3]
 *   15 sz2[9]  24 n2[9]  33 s2tot
 *   34 eknot  35 epot  36 ehartree  37 dust  38 dund  39 dse  40 dph  41 eint
 * The diagonal terms carry gs_weight = peso*|psi_i|^2.  As in the reference,
 * the off-diagonal Hloc terms of eknot (:536-543) and the spin-exchange and
 * pair-hopping terms of epot/dse/dph (:597-647) are summed over states
 * without peso, and epot includes ehartree (:669); eint is never assigned
 * by the reference and stays 0. */
int ref_local_obs(const int64_t *map, long long dim, const double *psi, double peso, double *out)
{
    const zdouble *p = (const zdouble *)psi;
    const int no = g_norb, ns = g_ns, sd = g_nspin - 1;
    for (long long i = 0; i < dim; ++i) {
        const int64_t m = map[i];
        const double w = peso * (creal(p[i]) * creal(p[i]) + cimag(p[i]) * cimag(p[i]));
        double nup[3], ndw[3], sz[3], nt[3], szs = 0.0;
        for (int a = 0; a < no; ++a) {
            nup[a] = (double)((m >> a) & 1);
            ndw[a] = (double)((m >> (a + ns)) & 1);
            sz[a] = 0.5 * (nup[a] - ndw[a]);
            nt[a] = nup[a] + ndw[a];
            szs += sz[a];
        }
        for (int a = 0; a < no; ++a) {
            out[0 + a] += nt[a] * w;
            out[3 + a] += nup[a] * w;
            out[6 + a] += ndw[a] * w;
            out[9 + a] += nup[a] * ndw[a] * w;
            out[12 + a] += (nup[a] - ndw[a]) * w;
            out[15 + a * 3 + a] += sz[a] * sz[a] * w;
            out[24 + a * 3 + a] += nt[a] * nt[a] * w;
            for (int b = a + 1; b < no; ++b) {
                out[15 + a * 3 + b] += sz[a] * sz[b] * w;
                out[15 + b * 3 + a] += sz[b] * sz[a] * w;
                out[24 + a * 3 + b] += nt[a] * nt[b] * w;
                out[24 + b * 3 + a] += nt[b] * nt[a] * w;
            }
        }
        out[33] += szs * szs * w;
        /* eknot: eloc(ispin,iorb) = impHloc(ispin,ispin,iorb,iorb) */
        for (int a = 0; a < no; ++a)
            out[34] += g_hloc[0][0][a][a] * nup[a] * w + g_hloc[sd][sd][a][a] * ndw[a] * w;
        for (int a = 0; a < no; ++a)
            for (int b = 0; b < no; ++b) {
                int64_t k1, k2;
                double s1, s2;
                if (nup[a] == 0.0 && nup[b] == 1.0) {
                    ref_c(b + 1, m, &k1, &s1);
                    ref_cdg(a + 1, k1, &k2, &s2);
                    long long j = ref_binary_search(map, dim, k2);
                    if (j > 0) out[34] += creal(g_hloc[0][0][a][b] * s1 * s2 * p[i] * conj(p[j - 1]));
                }
                if (ndw[a] == 0.0 && ndw[b] == 1.0) {
                    ref_c(b + 1 + ns, m, &k1, &s1);
                    ref_cdg(a + 1 + ns, k1, &k2, &s2);
                    long long j = ref_binary_search(map, dim, k2);
                    if (j > 0) out[34] += creal(g_hloc[sd][sd][a][b] * s1 * s2 * p[i] * conj(p[j - 1]));
                }
            }
        /* epot: density-density terms */
        for (int a = 0; a < no; ++a) out[35] += g_uloc[a] * nup[a] * ndw[a] * w;
        for (int a = 0; a < no; ++a)
            for (int b = a + 1; b < no; ++b) {
                out[35] += g_ust * (nup[a] * ndw[b] + nup[b] * ndw[a]) * w;
                out[37] += (nup[a] * ndw[b] + nup[b] * ndw[a]) * w;
                out[35] += (g_ust - g_jh) * (nup[a] * nup[b] + ndw[a] * ndw[b]) * w;
                out[38] += (nup[a] * nup[b] + ndw[a] * ndw[b]) * w;
            }
        if (no > 1 && g_jhflag) {
            for (int a = 0; a < no; ++a)
                for (int b = 0; b < no; ++b) {
                    if (a == b) continue;
                    int64_t k1, k2, k3, k4;
                    double s1, s2, s3, s4;
                    /* spin exchange: c_b,up c_a,dw cdg_b,dw cdg_a,up */
                    if (nup[b] == 1.0 && ndw[a] == 1.0 && ndw[b] == 0.0 && nup[a] == 0.0) {
                        ref_c(b + 1, m, &k1, &s1);
                        ref_c(a + 1 + ns, k1, &k2, &s2);
                        ref_cdg(b + 1 + ns, k2, &k3, &s3);
                        ref_cdg(a + 1, k3, &k4, &s4);
                        long long j = ref_binary_search(map, dim, k4);
                        if (j > 0) {
                            const double v = creal(s1 * s2 * s3 * s4 * p[i] * conj(p[j - 1]));
                            out[35] += g_jx * v;
                            out[39] += v;
                        }
                    }
                    /* pair hopping: c_b,up c_b,dw cdg_a,dw cdg_a,up */
                    if (nup[b] == 1.0 && ndw[b] == 1.0 && ndw[a] == 0.0 && nup[a] == 0.0) {
                        ref_c(b + 1, m, &k1, &s1);
                        ref_c(b + 1 + ns, k1, &k2, &s2);
                        ref_cdg(a + 1 + ns, k2, &k3, &s3);
                        ref_cdg(a + 1, k3, &k4, &s4);
                        long long j = ref_binary_search(map, dim, k4);
                        if (j > 0) {
                            const double v = creal(s1 * s2 * s3 * s4 * p[i] * conj(p[j - 1]));
                            out[35] += g_jp * v;
                            out[40] += v;
                        }
                    }
                }
        }
        if (g_hfmode) {
            double eh = 0.0;
            for (int a = 0; a < no; ++a)
                eh += -0.5 * g_uloc[a] * (nup[a] + ndw[a]) * w + 0.25 * g_uloc[a] * w;
            for (int a = 0; a < no; ++a)
                for (int b = a + 1; b < no; ++b) {
                    const double nn = nup[a] + ndw[a] + nup[b] + ndw[b];
                    eh += -0.5 * g_ust * nn * w + 0.25 * g_ust * w;
                    eh += -0.5 * (g_ust - g_jh) * nn * w + 0.25 * (g_ust - g_jh) * w;
                }
            out[36] += eh;
            out[35] += eh;
        }
    }
    return 0;
}

/* Dense sector Hamiltonian (real part) by applying H to unit vectors: the
 * matrix ed_buildH_c dumps for dim <= lanc_dim_threshold
 * (ED_HAMILTONIAN_MATVEC.f90:185-196).  H is column-major dim x dim. */
int ref_dense_h(const int64_t *map, long long dim, double *H)
{
    zdouble *e = calloc(dim, sizeof(zdouble));
    zdouble *h = calloc(dim, sizeof(zdouble));
    if (!e || !h) { free(e); free(h); return -5; }
    int rc = 0;
    for (long long j = 0; j < dim && rc == 0; ++j) {
        e[j] = 1.0;
        rc = ref_hxv(map, dim, (const double *)e, (double *)h);
        for (long long i = 0; i < dim; ++i) H[i + j * dim] = creal(h[i]);
        e[j] = 0.0;
    }
    free(e); free(h);
    return rc;
}
