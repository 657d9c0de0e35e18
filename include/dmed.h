/*
 * dmed.h -- C ABI of the MI355X (gfx950) DMFT-ED impurity-solver hot path.
 *
 * Plain C: pointers, sizes, status codes; no C++ or torch types.  The Fortran
 * module DMFT_ED (dmft-ed_amd/fortran/) binds these entry points with
 * ISO_C_BINDING; tests and bench.py bind them with ctypes.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative DMED_E* code on
 *     failure; dmed_last_error() then describes the failure.  The reference
 *     aborts with `stop "..."` instead (e.g. ED_HAMILTONIAN_MATVEC.f90:313,317;
 *     ED_SETUP.f90:1061,1075); the Fortran shim turns a nonzero code into stop.
 *   - One process, one GPU.  Calls are host-synchronous unless named
 *     *_async.  The caller serialises its calls: the public entry points are
 *     not meant to be called from several host threads at once.  Inside one
 *     call the library may run its own host worker threads, each with its own
 *     HIP stream and scratch (the batched solve, dmed_solve_sites); those are
 *     joined before the call returns.
 *   - State vectors are real (FP64 or FP32).  The reference's vectors are
 *     complex(8); H is real for a normal bath with real Hloc, so the real
 *     vector is the reference vector up to a global phase.
 *   - Vector element order is the reference's sector order
 *     (ED_SETUP.f90:872-889): i = r_up + r_dw*D_up (0-based), r = rank of the
 *     spin string in numeric order; i.e. a column-major D_up x D_dw matrix.
 *   - Arrays that mirror reference arrays use the reference's Fortran layout
 *     (column-major, first index fastest), e.g. Hloc(Nspin,Nspin,Norb,Norb),
 *     impGmats(Nspin,Nspin,Norb,Norb,Lmats) as interleaved (re,im) doubles.
 */
#ifndef DMED_H
#define DMED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define DMED_OK 0
#define DMED_EINVAL (-1)    /* bad argument */
#define DMED_ESTATE (-2)    /* call out of order (no init / no model / ...) */
#define DMED_EHIP (-3)      /* HIP runtime error */
#define DMED_ENOMEM (-4)    /* device or host allocation failed */
#define DMED_EUNSUPP (-5)   /* model feature not built for the GPU path */
#define DMED_ENOCONV (-6)   /* eigensolver did not reach the tolerance */
#define DMED_ENORM (-7)     /* a state that must be normalised is not */

/* precision selector for H*v and the GF chains */
#define DMED_FP64 0
#define DMED_FP32 1         /* FP32 vectors, FP64-accumulated reductions */
#define DMED_MIXED 2        /* Lanczos chains only: FP64 head steps, then FP32 */

/* ------------------------------------------------------------ setup */
/* Select the device and create the library stream.  Replaces the implicit
 * module initialisation of the reference (init_ed_structure,
 * ED_SETUP.f90:135-332). */
int dmed_init(int device);
/* Release the library stream, its work buffers and the worker contexts that
 * batched solves keep between calls (streams and grow-only work buffers). */
int dmed_finalize(void);
const char *dmed_last_error(void);

/* Impurity model, normal bath, normal ed_mode.
 *   bath : flat bath vector, set_dmft_bath layout (ED_BATH/dmft_aux.f90:489-512):
 *          e(ispin,iorb,k) at k + iorb*Nbath + ispin*Nbath*Norb (0-based),
 *          v(ispin,iorb,k) at the same index + Nspin*Norb*Nbath.
 *          Length 2*Nspin*Norb*Nbath = get_bath_dimension() (user_aux.f90:17-30).
 *   hloc : real part of impHloc(Nspin,Nspin,Norb,Norb), Fortran order.
 * Globals of ED_INPUT_VARS.f90:13-24 (Uloc, Ust, Jh, Jx, Jp, xmu, HFmode,
 * Jhflag).  Jhflag=1 with Norb>1 adds the spin-exchange (Jx) and
 * pair-hopping (Jp) terms of HxVint.f90:48-97.
 * Replaces: set_dmft_bath + set_Hloc + the module globals read by
 * build_hxv.f90 (ED_HAMILTONIAN_MATVEC.f90:289-351). */
#define DMED_BATH_NORMAL 0   /* Nbath levels per orbital, Ns = (Nbath+1)*Norb */
#define DMED_BATH_HYBRID 1   /* Nbath levels shared by all orbitals, Ns = Nbath+Norb */
typedef struct dmed_model {
    int32_t norb, nbath, nspin;
    const double *bath;
    const double *hloc;
    double uloc[3];
    double ust, jh, jx, jp, xmu;
    int32_t hfmode, jhflag;
    /* bath_type (ED_INPUT_VARS bath_type).  DMED_BATH_HYBRID uses the layout
     * of dmft_aux.f90:572-590: e(ispin,1,k) at k + (ispin-1)*Nbath, then
     * v(ispin,iorb,k) at Nspin*Nbath + k + (iorb-1)*Nbath + (ispin-1)*Norb*Nbath,
     * length Nspin*(Nbath + Norb*Nbath); level k sits on site Norb+k and
     * couples to every orbital (getBathStride, ED_SETUP.f90:452-455). */
    int32_t bath_type, pad_;
} dmed_model;
int dmed_set_model(const dmed_model *m);

/* Ns = (Nbath+1)*Norb, or Nbath+Norb for the hybrid bath
 * (setup_ed_dimensions, ED_SETUP.f90:81-128) */
int dmed_get_ns(int32_t *ns);

/* binomial(n1,n2) with the reference's rounding (ED_SETUP.f90:1258-1275) */
int64_t dmed_binomial(int32_t n1, int32_t n2);

/* --------------------------------------------------------- sectors */
/* A sector (nup, ndw) with its device tables: up/down spin strings, hop
 * tables (target rank + signed coefficient) per spin, separable diagonal.
 * Replaces build_sector / setup_Hv_sector (ED_SETUP.f90:861-892,
 * ED_HAMILTONIAN_MATVEC.f90:99-111); library-owned, freed by
 * dmed_sector_free. */
typedef struct dmed_sector dmed_sector;
int dmed_sector_create(int32_t nup, int32_t ndw, dmed_sector **out);
int dmed_sector_free(dmed_sector *s);
int dmed_sector_dims(const dmed_sector *s, int64_t *dim, int64_t *dim_up,
                     int64_t *dim_dw);
/* Reference state codes map(i) = iup + idw*2^Ns (ED_SETUP.f90:889), as
 * int64 (the reference's int32 code overflows at Ns=16).  codes: host array
 * of length dim. */
int dmed_sector_map(const dmed_sector *s, int64_t *codes);

/* --------------------------------------------------------- vectors */
/* Device vectors of a sector, library-owned. */
typedef struct dmed_vec dmed_vec;
int dmed_vec_create(const dmed_sector *s, int32_t prec, dmed_vec **out);
int dmed_vec_free(dmed_vec *v);
int dmed_vec_upload(dmed_vec *v, const double *host);     /* host FP64, length dim */
int dmed_vec_download(const dmed_vec *v, double *host);   /* host FP64, length dim */
void *dmed_vec_data(const dmed_vec *v);                   /* device pointer */
/* dst = src on the device, converting FP64 <-> FP32; both vectors belong to
 * the same sector.  A streaming copy (read + write of dim elements); the
 * reference copies vectors with Fortran assignment (ED_EIGENSPACE.f90:177-178). */
int dmed_vec_copy(dmed_vec *dst, const dmed_vec *src);

/* ------------------------------------------------------------- H*v */
/* y = H x on the sector, x and y device pointers of length dim in the given
 * precision.  Replaces the operator behind the procedure pointer
 * spHtimesV_cc(Nloc, v, Hv) (ED_VARS_GLOBAL.f90:43-49,100) -- directMatVec_cc
 * (ED_HAMILTONIAN_MATVEC.f90:289-351) and spMatVec_cc (:221-236). */
int dmed_hxv(const dmed_sector *s, const void *x_dev, void *y_dev, int32_t prec);
/* Same product on host FP64 arrays (copies in and out; for tests). */
int dmed_hxv_host(const dmed_sector *s, const double *x, double *y, int32_t prec);
/* Same product, launched without a final synchronisation (for timing). */
int dmed_hxv_async(const dmed_sector *s, const void *x_dev, void *y_dev, int32_t prec);
int dmed_sync(void);
/* Measurement entry point (no counterpart in the reference): nrep launches
 * of the H*v kernel bracketed by HIP events recorded on the library stream;
 * ms_avg receives the average duration of one launch in milliseconds. */
int dmed_hxv_timed(const dmed_sector *s, const void *x_dev, void *y_dev, int32_t prec,
                   int32_t nrep, double *ms_avg);
/* Launch shape of the sector's plain H*v in precision prec, after the
 * autotuner has run for it (no counterpart in the reference: it reports the
 * library's own choice).  kernel 1: the tile kernel (columns_per_group
 * columns staged in LDS per workgroup); kernel 2: the persistent column
 * kernel (one column in LDS, rows_per_lane rows per lane, groups_per_cu
 * workgroups per CU; row_passes 2 when each lane takes its rows in two
 * halves, the form that fits two workgroups per CU in registers).  tuned = 1
 * when the shape came from the autotuner. */
typedef struct dmed_hxv_shape {
    int32_t kernel, lanes, rows_per_lane, columns_per_group, groups_per_cu, grid, tuned;
    int32_t row_passes;
} dmed_hxv_shape;
int dmed_hxv_shape_of(const dmed_sector *s, int32_t prec, dmed_hxv_shape *out);

/* ------------------------------------------------------ eigensolver */
/* Lowest neig eigenpairs of the sector.  Replaces sp_eigh(MatVec, Dim,
 * Neigen, Nblock, Nitermax, eval, evec, tol) (call site ED_DIAG.f90:179-181)
 * and, for dim <= dense_threshold, ed_buildH_c + eigh (ED_DIAG.f90:185-193).
 *   ncv      : Krylov basis size (Nblock of the reference); 0 = automatic
 *   maxit    : maximum number of H*v products
 *   tol      : residual tolerance ||H psi - E psi|| <= tol*max(1,|E|)
 *   evals    : host array [neig], ascending
 *   evecs    : array of neig handles created by the call (FP64), or NULL
 *   resid    : host array [neig] of final residual norms, or NULL
 * Returns DMED_ENOCONV (results still filled) if tol is not reached. */
int dmed_eigh(const dmed_sector *s, int32_t neig, int32_t ncv, int32_t maxit,
              double tol, int32_t dense_threshold, double *evals,
              dmed_vec **evecs, double *resid);

/* Ground state by two-pass plain Lanczos with explicit-residual restarts
 * (lanczos_plain_c, .repo/PLAIN_LANCZOS.f90:286-385; the solver the
 * reference names for sectors where a stored Krylov basis does not fit).
 * prec = DMED_FP64, or DMED_MIXED: the first restarts run the recurrence on
 * FP32 vectors (FP64 scalars and reductions) until the FP64 residual of the
 * accumulated vector is below 1e-5*max(1,|E|); the remaining restarts run in
 * FP64.  The Rayleigh quotient and residual are always FP64.
 *   evec : handle created by the call (FP64, normalised), or NULL */
int dmed_ground_state(const dmed_sector *s, int32_t prec, int32_t maxit, double tol,
                      double *eval, dmed_vec **evec, double *resid);

/* -------------------------------------------------- Green's function */
/* out = c^+_{isite} psi (dagger=1) or c_{isite} psi (dagger=0), where
 * isite = iorb + (ispin-1)*Ns is the reference's 1-based spin-orbital index
 * (impIndex, ED_SETUP.f90:428-431) and out lives in the target sector
 * (getCDGsector / getCsector, :464-492).  norm2 = <out|out>.  Replaces the
 * start-vector loop of lanc_build_gf_normal_c (build_gf_normal.f90:105-113,
 * 144-152). */
int dmed_apply_op(const dmed_vec *psi, int32_t isite, int32_t dagger,
                  const dmed_sector *dst, dmed_vec *out, double *norm2);

/* Plain Lanczos tridiagonalisation from v0 (normalised internally, v0 is
 * not modified): alpha[k], beta[k+1] for k < nlanc-1 with the semantics of
 * sp_lanc_tridiag (build_gf_normal.f90:124-126; .repo/PLAIN_LANCZOS.f90:154-180),
 * stopping early when |b| < threshold; unset entries are zero.
 * prec = DMED_FP32 runs H*v and the vectors in FP32 with FP64 reductions;
 * prec = DMED_MIXED runs the first 32 steps in FP64 and the rest in FP32
 * (the rounding of the first steps dominates the error of an FP32 chain).
 * nsteps receives the number of alphas produced. */
int dmed_chain(const dmed_sector *s, const dmed_vec *v0, int32_t nlanc,
               int32_t prec, double threshold, double *alpha, double *beta,
               int32_t *nsteps);

/* Lanczos chain that allocates no device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 * arguments and results as dmed_chain (the caller still owns v0 and the host
 * alpha/beta/nsteps); the chain vectors and reduction partials live in a
 * library-owned static device workspace of DMED_NOALLOC_SCRATCH_BYTES per
 * GPU, held until the process ends, so the sector may have at most
 * DMED_NOALLOC_MAX_DIM states (DMED_EINVAL beyond).  The autotuner does not
 * time candidates inside this call and Lanczos steps are launched one by
 * one.  Calls that overlap on one GPU race on the shared workspace. */
#define DMED_NOALLOC_MAX_DIM 2097152LL
#define DMED_NOALLOC_SCRATCH_BYTES (96ull << 20)
int dmed_chain_noalloc(const dmed_sector *s, const dmed_vec *v0, int32_t nlanc, int32_t prec,
                       double threshold, double *alpha, double *beta, int32_t *nsteps);

/* add_to_lanczos_gf_normal (build_gf_normal.f90:442-488): accumulate
 *   G(z) += peso_bz * sum_j Z(1,j)^2 / (z - isign*(lambda_j - ei))
 * with (lambda, Z) the tql2 spectrum of the chain, on z = i*wm[n] and
 * z = wr[n] + i*eps.  gmats/greal: interleaved complex, lmats / lreal. */
int dmed_gf_add_chain(double peso_bz, double ei, const double *alpha,
                      const double *beta, int32_t nlanc, int32_t isign,
                      const double *wm, int32_t lmats, const double *wr,
                      int32_t lreal, double eps, double *gmats, double *greal);

/* S_z |psi> into out (same sector, FP64), S_z = (n_a,up - n_a,dw)/2 of
 * impurity orbital iorb (0-based), or of all impurity orbitals when
 * iorb = -1; the start vector of lanc_ed_build_spinChi_c /
 * lanc_ed_build_spinChi_tot_c (build_chi_spin.f90:55-60, 131-136).
 * norm2 receives ||S_z psi||^2. */
int dmed_apply_sz(const dmed_vec *psi, int32_t iorb, dmed_vec *out, double *norm2);

/* N |psi> into out (same sector and precision), N = sum over the impurity
 * orbitals a in mask (bit a = orbital a, 0-based) of n_a,up + n_a,dw; the
 * start vectors of lanc_ed_build_densChi_diag_c (one orbital),
 * _offdiag_c (two) and _tot_c (all) (build_chi_dens.f90:69-75, 214-221,
 * 143-149).  norm2 receives ||N psi||^2. */
int dmed_apply_occ(const dmed_vec *psi, uint32_t mask, dmed_vec *out, double *norm2);

/* add_to_lanczos_spinChi (build_chi_spin.f90:186-250), both isign = +1 and
 * -1 of one chain: with (lambda, Z) the tql2 spectrum of the chain,
 * dE_j = lambda_j - ei and peso_j = peso_f*peso_bz*Z(1,j)^2,
 *   chi_iv[0] += peso_j*beta if beta*dE_j < 0.1, else peso_j*(1-exp(-beta dE_j))/dE_j
 *   chi_iv[n] += peso_j*(exp(-beta dE_j)-1)/(i vm[n] - dE_j)
 *              + peso_j*(1-exp(-beta dE_j))/(i vm[n] + dE_j),   n >= 1
 *   chi_tau[n] += peso_j*(exp(-tau[n] dE_j) + exp(-(beta-tau[n]) dE_j))
 *   chi_w[n]  += the two terms of chi_iv with i vm[n] -> wr[n] + i eps
 * (the iv = 0 term is added once per isign).  vm: nv bosonic frequencies
 * 2*pi*n/beta, n = 0..nv-1; chi_iv, chi_w interleaved complex. */
int dmed_chi_add_chain(double peso_f, double peso_bz, double ei, const double *alpha,
                       const double *beta_l, int32_t nlanc, double beta, const double *vm,
                       int32_t nv, const double *tau, int32_t ntau, const double *wr,
                       int32_t lreal, double eps, double *chi_iv, double *chi_tau,
                       double *chi_w);

/* dens/docc contribution of one state (ED_OBSERVABLES.f90:118-150):
 * dens[a] += peso*sum|psi|^2 (n_a,up + n_a,dw), docc[a] += peso*sum|psi|^2 n_a,up n_a,dw */
int dmed_observables(const dmed_vec *psi, double peso, double *dens, double *docc);

/* Local observables and energies (observables_impurity, ED_OBSERVABLES.f90:74-150,
 * and local_energy_impurity, :450-669, normal mode).  Orbital arrays [a],
 * orbital pairs [a*3+b], Norb <= 3.  epot includes ehartree; eint is never
 * assigned by the reference and stays 0; dse and dph (spin exchange and pair
 * hopping, Jhflag with Norb > 1) are summed without the state weight peso, as
 * the reference does (ED_OBSERVABLES.f90:597-647), and epot adds Jx*dse +
 * Jp*dph. */
typedef struct dmed_local_obs {
    double dens[3], dens_up[3], dens_dw[3], docc[3], magz[3];
    double sz2[9], n2[9];
    double s2tot;
    double eknot, epot, ehartree, dust, dund, dse, dph, eint;
} dmed_local_obs;

/* Accumulate one state into acc: the terms the reference weights with
 * gs_weight = peso*|psi_i|^2 get peso; the off-diagonal Hloc terms of eknot
 * get weight 1, as the reference sums them without peso
 * (ED_OBSERVABLES.f90:536-543).  psi must be normalised. */
int dmed_local_observables(const dmed_vec *psi, double peso, dmed_local_obs *acc);

/* ------------------------------------------------------ whole solve */
/* Parameters of ed_solve (names and defaults of ED_INPUT_VARS.f90:111-176). */
typedef struct dmed_solve_params {
    int32_t lanc_nstates_sector;   /* 6 */
    int32_t lanc_nstates_total;    /* 1  (1 => T=0) */
    int32_t lanc_nstates_step;     /* 2 */
    int32_t lanc_niter;            /* 512 */
    int32_t lanc_ngfiter;          /* 200 */
    int32_t lanc_ncv_factor;       /* 3 */
    int32_t lanc_ncv_add;          /* 5 */
    int32_t lanc_dim_threshold;    /* 256 */
    double lanc_tolerance;         /* 1e-12 */
    double gs_threshold;           /* 1e-9 */
    double cutoff;                 /* 1e-9 */
    double beta;                   /* 1000 */
    double eps;                    /* 0.01 */
    double wini, wfin;             /* -5, 5 */
    int32_t lmats, lreal;          /* 5000, 5000 */
    int32_t prec_gf;               /* DMED_FP64 / DMED_FP32 / DMED_MIXED chains */
    int32_t batched;               /* 1: batched sector/chain engine */
    int32_t verbose;
    int32_t gf_fp64_head;          /* 32: FP64 steps of a DMED_MIXED chain */
    int32_t prec_gs;               /* DMED_FP64 / DMED_MIXED plain-Lanczos ground states */
    int32_t chi_spin;              /* 0; 1: spin susceptibility (CHIFLAG, ED_MAIN.f90:282) */
    int32_t ltau;                  /* 1000: imaginary-time points, raised to beta (ED_INPUT_VARS.f90:191) */
    int32_t chi_dens;              /* 0; 1: charge susceptibility (CHIFLAG, build_chi_dens.f90:4-41) */
} dmed_solve_params;
void dmed_solve_params_default(dmed_solve_params *p);

/* Results of one solve; arrays are caller-allocated in the reference layouts:
 *   dens, docc        : [Norb]
 *   gmats, smats      : impGmats/impSmats(Nspin,Nspin,Norb,Norb,Lmats) complex
 *   greal, sreal      : (Nspin,Nspin,Norb,Norb,Lreal) complex
 *   neigen_sector     : [(Ns+1)^2] in/out (ED_DIAG.f90:328-343), may be NULL
 *   lanc_nstates_total: in/out (ED_DIAG.f90:345-396), may be NULL */
typedef struct dmed_solve_out {
    double egs, zeta;
    int32_t nstates, numgs;
    double *dens, *docc;
    double *gmats, *greal, *smats, *sreal;
    int32_t *neigen_sector;
    int32_t *lanc_nstates_total;
    double t_diag, t_gf;           /* seconds spent in the two phases */
    dmed_local_obs *obs;           /* local observables and energies, or NULL */
    /* spin susceptibility (chi_spin = 1), rows iorb = 0..Norb-1 and Norb =
     * the total (Norb > 1), Fortran order like spinChi_* (ED_VARS_GLOBAL.f90:
     * 136-138): chi_iv (Norb+1, 0:Lmats) complex, chi_tau (Norb+1, 0:Ltau)
     * real, chi_w (Norb+1, Lreal) complex; Ltau as raised; may be NULL */
    double *spin_chi_iv, *spin_chi_tau, *spin_chi_w;
    int32_t ltau_used, pad2_;
    /* charge susceptibility (chi_dens = 1), Fortran order like densChi_*
     * (ED_VARS_GLOBAL.f90:143-157); each may be NULL:
     *   dens_chi_iv/_mix_iv (Norb, Norb, 0:Lmats) complex,
     *   dens_chi_tau/_mix_tau (Norb, Norb, 0:Ltau) real,
     *   dens_chi_w/_mix_w (Norb, Norb, Lreal) complex,
     *   dens_chi_tot_iv (0:Lmats), _tot_tau (0:Ltau), _tot_w (Lreal).
     * Only a < b of dens and the diagonal are assigned by the reference's
     * offdiag/diag builders; mix and tot are filled when Norb > 1 (zero
     * otherwise); dens is divided by zeta twice, mix and tot once, as the
     * reference does (build_chi_dens.f90:36-38). */
    double *dens_chi_iv, *dens_chi_tau, *dens_chi_w;
    double *dens_chi_mix_iv, *dens_chi_mix_tau, *dens_chi_mix_w;
    double *dens_chi_tot_iv, *dens_chi_tot_tau, *dens_chi_tot_w;
} dmed_solve_out;

/* ed_solve_single backend (ED_MAIN.f90:251-289): diagonalize_impurity
 * (ed_diag_c + ed_analysis), observables (dens/docc), buildgf_impurity
 * (build_gf_normal + get_sigma_normal), on the current model. */
int dmed_solve(const dmed_solve_params *p, dmed_solve_out *out);

/* ed_solve_lattice (ED_MAIN.f90:364-468): dmed_solve for nsites independent
 * impurity models (inequivalent sites; per-site baths, Hloc and
 * interactions in models[i]), results in outs[i] (each with its own
 * neigen_sector / lanc_nstates_total in/out, as neigen_sectorii and
 * neigen_totalii of the reference).  With p->batched and nsites > 1 the
 * sites run concurrently on host workers with their own streams (each
 * site's chains serially); otherwise one after another.  Results do not
 * depend on the schedule.  The model set by dmed_set_model is unchanged. */
int dmed_solve_sites(int32_t nsites, const dmed_model *models, const dmed_solve_params *p,
                     dmed_solve_out *outs);

/* Sigma = G0^-1 - G^-1 of the normal bath on arbitrary points z
 * (get_sigma_normal.f90:27-38; invg0_bath_mats, ED_BATH_FUNCTIONS.f90:1784-1807).
 * ispin, iorb: 0-based.  z, g, sigma: interleaved complex of length L. */
int dmed_sigma(int32_t ispin, int32_t iorb, const double *z, const double *g,
               int32_t L, double *sigma);
/* Delta(z) = sum_k V_k^2/(z-e_k) (delta_bath_mats, ED_BATH_FUNCTIONS.f90:221-258);
 * ispin, iorb 0-based */
int dmed_delta(int32_t ispin, int32_t iorb, const double *z, int32_t L, double *delta);

/* chi^2 fit of one spin of a hybrid bath (e_k; V_ak for all orbitals), in
 * place, to the orbital matrix fg(Norb,Norb,L) (Fortran order, complex) on
 * the first L Matsubara points, summed over a <= b: scheme 0 fits
 * G0 = [(iw + xmu) 1 - Hloc - Delta]^-1, 1 fits Delta, with
 * Delta_ab = sum_k V_ak V_bk/(iw - e_k); hloc(Norb,Norb) real, Fortran order;
 * e[Nbath], v[k + a*Nbath]; weights as dmed_chi2_fit.  Host-only.  Replaces
 * chi2_fitgf_hybrid_normal (ED_CHI2FIT/fitgf_hybrid_normal.f90:18-156) with
 * the same minimiser as dmed_chi2_fit (parity of the optimiser unpinned). */
int dmed_chi2_fit_hybrid(int32_t scheme, int32_t weight, const double *fg, int32_t L,
                         double beta, double xmu, const double *hloc, int32_t norb,
                         int32_t nbath, double *e, double *v, int32_t niter, double ftol,
                         double *chi2, int32_t *iter);

/* chi^2 fit of one orbital's normal bath (e_k, V_k), in place, to fg on the
 * first L Matsubara points w_n = pi/beta*(2n-1): scheme 0 fits
 * G0 = 1/(iw + xmu - hloc_aa - Delta) ("weiss"), 1 fits Delta ("delta");
 * weight = cg_weight (0: L, 1: 1, 2: n, 3: w_n).  Host-only.  Replaces
 * chi2_fitgf_normal_normal (ED_CHI2FIT/fitgf_normal_normal.f90:18-156) and
 * its fmin_cg call (:86-90). */
int dmed_chi2_fit(int32_t scheme, int32_t weight, const double *fg, int32_t L, double beta,
                  double xmu, double hloc_aa, int32_t nbath, double *e, double *v,
                  int32_t niter, double ftol, double *chi2, int32_t *iter);

#ifdef __cplusplus
}
#endif
#endif /* DMED_H */
