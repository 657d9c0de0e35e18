/* Host sanitizer run of the oracle (oracle/edref.c): every entry point on
 * small sectors of a two-orbital model with Hund's coupling (normal bath,
 * two spins) and of a hybrid-bath model, built with AddressSanitizer and
 * UndefinedBehaviorSanitizer by tests/test_oracle_sanitize.py.  Prints "OK"
 * and exits 0 when every call returned its success code; the sanitizers abort
 * on the first out-of-bounds access, leak or undefined operation. */
#include <complex.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int ref_set_model(int norb, int nbath, int nspin, const double *bath, const double *hloc_f,
                  const double *uloc, double ust, double jh, double jx, double jp, double xmu,
                  int hfmode, int jhflag);
void ref_set_bath_type(int hybrid);
int ref_ns(void);
long long ref_sector_dim(int nup, int ndw);
long long ref_build_sector(int nup, int ndw, int64_t *map);
long long ref_binary_search(const int64_t *a, long long n, int64_t value);
int ref_hxv(const int64_t *map, long long dim, const double *vin, double *hv);
int ref_hxv_omp(const int64_t *map, long long dim, const double *vin, double *hv, int nthreads);
int ref_hxv_block(const int64_t *map, long long dim, const double *vin, double *out,
                  long long first, long long last, int nthreads);
int ref_lanc_tridiag(const int64_t *map, long long dim, double *vin, int nitermax, double *alanc,
                     double *blanc, double threshold, int nthreads, int *nsteps);
int ref_lanczos_gs(const int64_t *map, long long dim, double *vect, int nitermax,
                   double threshold, int ncheck, double *egs, int nthreads);
int ref_apply_op(const int64_t *mapI, long long dimI, const double *psi, const int64_t *mapJ,
                 long long dimJ, int isite, int dagger, double *out);
int ref_add_to_gf(double peso_bz, double ei, const double *alanc, const double *blanc, int nlanc,
                  int isign, const double *wm, int lmats, const double *wr, int lreal,
                  double eps, double *gmats, double *greal);
void ref_delta_ab(int ispin, int a, int b, const double *z, int L, double *out);
void ref_sigma(int ispin, int iorb, const double *z, const double *g, int L, double *out);
void ref_observables(const int64_t *map, long long dim, const double *psi, double peso,
                     double *dens, double *docc);
int ref_local_obs(const int64_t *map, long long dim, const double *psi, double peso,
                  double *out);
int ref_dense_h(const int64_t *map, long long dim, double *H);

static int fails = 0;
#define CHECK(cond, what)                                                  \
    do {                                                                   \
        if (!(cond)) {                                                     \
            fprintf(stderr, "failed: %s (line %d)\n", what, __LINE__);    \
            ++fails;                                                       \
        }                                                                  \
    } while (0)

static uint64_t rng_state = 12345u;
static double urand(void)
{
    rng_state = rng_state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(rng_state >> 11) / 9007199254740992.0 - 0.5;
}

static int64_t *sector(int nup, int ndw, long long *dim)
{
    *dim = ref_sector_dim(nup, ndw);
    int64_t *map = malloc((size_t)*dim * sizeof *map);
    CHECK(map && ref_build_sector(nup, ndw, map) == *dim, "build_sector");
    return map;
}

/* H*v in three forms, a chain, the ground state, c^+ psi, the pole sum,
 * Delta and Sigma, the observables and the dense matrix on sector (nup, ndw) */
static void exercise(int nup, int ndw, int norb, int nspin)
{
    long long dim, dimJ;
    int64_t *map = sector(nup, ndw, &dim);
    double *x = malloc(2 * dim * sizeof *x), *y = malloc(2 * dim * sizeof *y);
    double *y2 = malloc(2 * dim * sizeof *y2), *y3 = malloc(2 * dim * sizeof *y3);
    for (long long i = 0; i < 2 * dim; ++i) x[i] = (i & 1) ? 0.0 : urand();
    CHECK(ref_hxv(map, dim, x, y) == 0, "hxv");
    CHECK(ref_hxv_omp(map, dim, x, y2, 3) == 0, "hxv_omp");
    CHECK(ref_hxv_block(map, dim, x, y3, dim / 3, dim - 1, 2) == 0, "hxv_block");
    CHECK(ref_hxv_block(map, dim, x, y3, 0, dim + 1, 1) == -1, "hxv_block rejects a bad range");
    for (long long i = 0; i < 2 * dim; ++i) CHECK(y[i] == y2[i], "hxv_omp equals hxv");
    CHECK(ref_binary_search(map, dim, map[dim / 2]) == dim / 2 + 1, "binary_search");

    const int nit = 40;
    double *a = calloc(nit, sizeof *a), *b = calloc(nit, sizeof *b);
    int nsteps = 0;
    memcpy(y, x, 2 * dim * sizeof *x);
    CHECK(ref_lanc_tridiag(map, dim, y, nit, a, b, 1e-13, 2, &nsteps) == 0, "lanc_tridiag");
    CHECK(nsteps >= 1 && nsteps <= nit, "chain length");

    double *egs = calloc(200, sizeof *egs);
    memcpy(y, x, 2 * dim * sizeof *x);
    CHECK(ref_lanczos_gs(map, dim, y, 200, 1e-12, 5, egs, 2) > 0, "lanczos_gs");

    int64_t *mapJ = sector(nup + 1, ndw, &dimJ);
    double *o = malloc(2 * dimJ * sizeof *o);
    CHECK(ref_apply_op(map, dim, y, mapJ, dimJ, 1, 1, o) == 0, "apply_op c^+");

    const int L = 16, LR = 8;
    double wm[16], wr[8], z[32], gm[32], gr[16], d[32], sg[32];
    for (int i = 0; i < L; ++i) wm[i] = 0.1 * (2 * i + 1);
    for (int i = 0; i < LR; ++i) wr[i] = -2.0 + 0.5 * i;
    memset(gm, 0, sizeof gm);
    memset(gr, 0, sizeof gr);
    CHECK(ref_add_to_gf(0.7, egs[0], a, b, nsteps, 1, wm, L, wr, LR, 0.01, gm, gr) == 0,
          "add_to_gf");
    for (int i = 0; i < L; ++i) {
        z[2 * i] = 0.0;
        z[2 * i + 1] = wm[i];
    }
    for (int s = 0; s < nspin; ++s)
        for (int p = 0; p < norb; ++p) {
            for (int q = 0; q < norb; ++q) ref_delta_ab(s, p, q, z, L, d);
            ref_sigma(s, p, z, gm, L, sg);
        }

    double dens[3] = {0}, docc[3] = {0}, obs[42] = {0};
    ref_observables(map, dim, y, 1.0, dens, docc);
    CHECK(ref_local_obs(map, dim, y, 1.0, obs) == 0, "local_obs");

    if (dim <= 400) {
        double *H = malloc((size_t)dim * dim * sizeof *H);
        CHECK(H && ref_dense_h(map, dim, H) == 0, "dense_h");
        free(H);
    }
    free(map); free(mapJ); free(x); free(y); free(y2); free(y3); free(o);
    free(a); free(b); free(egs);
}

int main(void)
{
    /* two orbitals, two bath levels each (Ns = 6), two spins, Hund's J with
     * the spin-exchange and pair-hopping terms, off-diagonal Hloc */
    const int norb = 2, nbath = 2, nspin = 2;
    double bath[2 * 2 * 2 * 2];
    for (int i = 0; i < 8; ++i) bath[i] = 1.5 * urand();
    for (int i = 8; i < 16; ++i) bath[i] = 0.4 + 0.2 * urand();
    double hloc[2 * 2 * 2 * 2] = {0};
    /* Hloc(s1,s2,a,b) Fortran order: s1 + 2*(s2 + 2*(a + 2*b)) */
    hloc[0 + 2 * (0 + 2 * (0 + 2 * 1))] = 0.1;
    hloc[0 + 2 * (0 + 2 * (1 + 2 * 0))] = 0.1;
    hloc[1 + 2 * (1 + 2 * (0 + 2 * 1))] = 0.1;
    hloc[1 + 2 * (1 + 2 * (1 + 2 * 0))] = 0.1;
    const double uloc[3] = {2.0, 2.0, 0.0};
    ref_set_bath_type(0);
    CHECK(ref_set_model(norb, nbath, nspin, bath, hloc, uloc, 1.5, 0.25, 0.25, 0.25, 0.0, 1, 1) == 0,
          "set_model normal");
    CHECK(ref_ns() == 6, "Ns");
    exercise(3, 3, norb, nspin);
    exercise(2, 4, norb, nspin);

    /* hybrid bath: three shared levels (Ns = 5), one spin */
    double hbath[3 + 2 * 3];
    for (int i = 0; i < 3; ++i) hbath[i] = urand();
    for (int i = 3; i < 9; ++i) hbath[i] = 0.3 + 0.1 * urand();
    double hloc1[4] = {0};
    ref_set_bath_type(1);
    CHECK(ref_set_model(2, 3, 1, hbath, hloc1, uloc, 1.0, 0.0, 0.0, 0.0, 0.0, 1, 0) == 0,
          "set_model hybrid");
    exercise(2, 2, 2, 1);
    ref_set_bath_type(0);

    if (fails) return 1;
    printf("OK\n");
    return 0;
}
