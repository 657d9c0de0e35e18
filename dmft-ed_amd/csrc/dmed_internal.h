// dmed_internal.h -- shared declarations of the gfx950 ED hot-path library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <complex>
#include <string>
#include <vector>

#include "../../include/dmed.h"

namespace dmed {

constexpr int kMaxNs = 30;         // spin strings are uint32_t
constexpr int kMaxTerms = 256;     // one-body hop terms per spin

// -------------------------------------------------------------- errors
int fail(int code, const std::string &msg);
#define DMED_HIP(call)                                                      \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess)                                               \
            return ::dmed::fail(DMED_EHIP, std::string(#call) + ": " +      \
                                           hipGetErrorString(e_));          \
    } while (0)

// --------------------------------------------------------------- model
struct Model {
    int norb = 0, nbath = 0, nspin = 1, ns = 0;
    std::vector<double> e, v;          // [spin][orb][k], spin = 0 (up), 1 (down)
    double hloc[2][3][3] = {};          // spin-diagonal real Hloc
    double uloc[3] = {}, ust = 0, jh = 0, jx = 0, jp = 0, xmu = 0;
    int hfmode = 1, jhflag = 0;
    int hybrid = 0;                     // bath_type: levels shared by all orbitals
    int generation = 0;                 // bumped by every dmed_set_model
    // e: [spin][orb][k] (normal) or [spin][k] (hybrid); v: [spin][orb][k]
    double E(int s, int a, int k) const
    {
        return hybrid ? e[s * nbath + k] : e[(s * norb + a) * nbath + k];
    }
    double V(int s, int a, int k) const { return v[(s * norb + a) * nbath + k]; }
    // 0-based site of bath level k of orbital a (getBathStride, ED_SETUP.f90:446-455)
    int bath_bit(int a, int k) const { return hybrid ? norb + k : norb + a * nbath + k; }
};

// one-body hop term c^+_i c_j (bit positions, i != j) with coefficient t
struct HopTerm {
    int32_t i, j;
    double t;
};

// Per-spin tables of a sector: strings of n electrons in Ns orbitals in
// numeric order, their single-spin diagonal energy, impurity bits and hop
// lists (target rank, signed coefficient) laid out hop-major [h][r].
struct SpinTables {
    int n = 0;
    int64_t dim = 0;
    int nhop = 0;
    uint32_t *str = nullptr;
    double *diag = nullptr;
    uint8_t *imp = nullptr;
    int32_t *hidx = nullptr;
    double *hval = nullptr;
    float *hval32 = nullptr;
    // packed form for the persistent column kernel: word = rank<<8 | term<<1 | sign,
    // with |coefficient| = coef[term]; valid when packed == true
    bool packed = false;
    int ncoef = 0;
    uint32_t *hpk = nullptr;            // [nhop][dim]
    double *coef = nullptr;             // [ncoef]
    // impurity pair hops for the spin-exchange / pair-hopping terms
    // (Jhflag, Norb > 1): ipair[(a*norb + b)*dim + r] = (rank << 1) | sign of
    // c^+_a c_b applied to string r (b occupied, a empty), -1 otherwise
    int32_t *ipair = nullptr;
    std::vector<uint32_t> h_str;        // host copies (small)
    std::vector<int32_t> h_ipair;
    std::vector<double> h_diag;
    std::vector<uint8_t> h_imp;
    std::vector<int32_t> h_hidx;
    std::vector<double> h_hval;
};

struct Context;

}  // namespace dmed

struct dmed_sector {
    int nup = 0, ndw = 0, ns = 0, norb = 0;
    int64_t dim = 0;
    int model_generation = -1;
    dmed::SpinTables up, dw;
    int nimp = 1;                       // 2^Norb
    int jterms = 0;                     // spin-exchange / pair-hopping present
    double jx = 0.0, jp = 0.0;
    double *W = nullptr;                // device [nimp*nimp] interaction table W[iu*nimp+id]
    std::vector<double> h_W;
};

struct dmed_vec {
    const dmed_sector *sec = nullptr;
    int prec = DMED_FP64;
    int64_t dim = 0;
    void *data = nullptr;
};

namespace dmed {

// ------------------------------------------------------------- context
constexpr int kScalCap = 4096;

// An execution context: one HIP stream with its own reduction scratch,
// device scalars (Lanczos chain state) and grow-only work buffers.  The
// library's calls run on the context bound to the calling host thread
// (the default one unless a worker bound its own), so independent sectors
// and chains can run concurrently on separate streams.
// Work-buffer slots of an execution context.
constexpr int kSlotChain = 0;           // 0..2: Lanczos chain vectors (v_prev, v, w)
constexpr int kSlotTrlBasis = 3;        // thick-restart basis
constexpr int kSlotTrlW = 4;            // thick-restart work vector
constexpr int kSlotTrlSmall = 5;        // projected matrix and coefficients
constexpr int kSlotDeflY = 6;           // locked eigenvectors of a deflation round
constexpr int kSlotDeflX = 7;           // eigenvectors collected over rounds
constexpr int kSlotOpTable = 8;         // c / c^+ source-target tables
constexpr int kSlotStart = 9;           // GF chain start vector
constexpr int kSlotMid = 10;            // intermediate vector of a two-operator start vector
constexpr int kSlotChainLo = 12;        // 12..14: FP32 tail of a mixed chain
constexpr int kExecSlots = 16;
constexpr int kMixedHead = 32;          // FP64 steps at the head of a mixed chain

struct Exec {
    hipStream_t stream = nullptr;
    double *partials = nullptr;
    size_t partials_cap = 0;            // in doubles
    double *dscal = nullptr;            // kScalCap doubles
    void *buf[kExecSlots] = {};
    size_t buf_cap[kExecSlots] = {};
};
int exec_create(Exec &e);
void exec_destroy(Exec &e);
void exec_pool_clear();                 // destroys the idle worker contexts (solve.cpp)
Exec &ex();                             // context of the calling thread
void bind_exec(Exec *e);                // nullptr: back to the default context
int exec_buf(Exec &e, int slot, size_t bytes, void **p);

struct Context {
    bool initialized = false;
    int device = 0;
    int num_cu = 256;
    size_t lds_max = 160 * 1024;
    Exec exec;                          // default execution context
    Model model;
    bool has_model = false;
    std::string last_error;
};
Context &ctx();
// Model of the calling thread: the one bound with bind_model (site-parallel
// solves bind one per worker), else the context's (dmed_set_model).
Model &cur_model();
bool has_cur_model();
void bind_model(Model *m);                // nullptr: back to the context's model
int model_fill(const dmed_model *m, Model &md);
int upload_binom();                       // binomial table for rank_of, once per process

int ensure_partials(size_t n);          // on ex()
int copy_h2d(void *dst, const void *src, size_t bytes);   // on ex().stream, synchronous

// ------------------------------------------------------------- sectors
int sector_build(dmed_sector *s, int nup, int ndw);
// Sector (nup, ndw) stacked twice along the down-spin index: vectors of
// 2*dim entries [x; y] and H*[x; y] = [H x; H y] (H (+) H).  Built from
// sector_build with the down-spin tables repeated and their column indices
// shifted by D_dw in the second copy; for chains of complex start vectors
// x - i y under the real H, whose Lanczos coefficients are those of [x; y].
int sector_build_stacked(dmed_sector *s, int nup, int ndw);
// out = (wu*n_up + wd*n_dw)(orbitals in mask) psi on the calling context's
// stream, norm2 = ||out||^2 (api.hip)
int apply_occ(const dmed_sector *s, int prec, const void *psi, unsigned mask, double wu,
              double wd, void *out, double *norm2);
void sector_release(dmed_sector *s);
int64_t binom(int n, int k);   // exact, n <= 62

// ---------------------------------------------------------------- H*v
// Epilogue of the H*v kernel.
//   kPlain   : y = Hx
//   kLanczos : y = sx*Hx - b*svp*vp, and block partials of <sx*x, y>
enum HxvMode { kPlain = 0, kLanczos = 1 };

struct ChainScalars {      // device-resident (inside Context::dscal)
    double sx;             // scale of the current Lanczos vector
    double svp;            // scale of the previous one
    double b;              // previous beta
    double alpha;          // current alpha
    double thr;            // early-exit threshold on beta
    double q;              // <x_new, x_old> of the stored vectors after the last update
    int32_t step;          // 0-based step index
    int32_t done;          // 1 once |beta| < thr
    int32_t nlanc;
    int32_t pad;
};

// the plain-mode launch shape of s (after tuning)
int hxv_shape(const dmed_sector *s, int prec, dmed_hxv_shape *out);

// mode kLanczos: y = sx*H*x with the block partials of <x, y> (cs->sx)
int launch_hxv(const dmed_sector *s, const void *x, void *y, int prec, int mode,
               ChainScalars *cs, double *partials, int *nparts, hipStream_t st);

// While a NoAllocScope lives on a thread, library code on that thread
// allocates no device memory: the autotuner keeps recorded or fixed launch
// shapes instead of timing candidates on scratch vectors.
bool thread_no_alloc();
struct NoAllocScope {
    NoAllocScope();
    ~NoAllocScope();
    bool prev;
};

// ---------------------------------------------------------- vector ops
int launch_zero(void *p, size_t bytes, hipStream_t st);   // kernel fill, no memset node
int launch_reduce_to(const double *partials, int n, double *out, int op, hipStream_t st);
int launch_dot(const void *x, const void *y, int64_t n, int prec, double *partials,
               int *nparts, hipStream_t st);
int launch_fill_random(void *x, int64_t n, int prec, uint64_t seed, hipStream_t st);
int launch_scale_copy(void *dst, int dprec, const void *src, int sprec, int64_t n,
                      double a, hipStream_t st);
int launch_axpy(void *y, int yprec, const void *x, int xprec, int64_t n, double a,
                hipStream_t st);
// dst (cols x rows) = transpose of src (rows x cols), column-major FP64
int launch_transpose(const double *src, int64_t rows, int64_t cols, double *dst, hipStream_t st);

int dot_host(const double *x, const double *y, int64_t n, double *out);

// --------------------------------------------------------- Krylov
// prec: DMED_FP64, DMED_FP32, or DMED_MIXED (`head` FP64 steps, then FP32;
// head < 0 selects kMixedHead)
int chain_run(const dmed_sector *s, const void *v0, int v0prec, int nlanc, int prec,
              double threshold, double *alpha, double *beta, int *nsteps, int head = -1);
// chain_run with its vectors and partial sums in the library's static device
// workspace (dmed_chain_noalloc); noalloc_setup takes that workspace's
// address on a device, once, with the device current (dmed_init)
int noalloc_setup(int device);
int chain_run_noalloc(const dmed_sector *s, const void *v0, int v0prec, int nlanc, int prec,
                      double threshold, double *alpha, double *beta, int *nsteps);
int eigh_sector(const dmed_sector *s, int neig, int ncv, int maxit, double tol,
                int dense_threshold, double *evals, dmed_vec **evecs, double *resid);
// lowest_multiplet: the lowest level and its copies within `degen` (at most
// neig); otherwise the neig lowest with multiplets completed.  *nout pairs.
// gs_prec (DMED_FP64 / DMED_MIXED) applies where the ground state comes
// from the plain Lanczos solver.
int eigh_sector_ex(const dmed_sector *s, int neig, int ncv, int maxit, double tol,
                   int dense_threshold, double degen, bool lowest_multiplet, int *nout,
                   double *evals, dmed_vec **evecs, double *resid, int gs_prec = DMED_FP64);
// plain two-pass Lanczos ground state with residual restarts; DMED_MIXED runs
// the first restarts on FP32 vectors and finishes in FP64
int ground_state(const dmed_sector *s, int prec, int maxit, double tol, double *eval,
                 dmed_vec **evec, double *resid);

// --------------------------------------------- persistent sector queue
// Small Krylov sectors of a batched solve, diagonalised by one launch of a
// persistent work-queue kernel (sectorq.hip): each workgroup takes a sector
// from a device counter and runs its whole thick-restart Lanczos with the
// multiplet-completing deflation rounds, with no host round trip.
constexpr int64_t kQueueMaxDim = 16384;
struct QueueItem {
    const dmed_sector *s = nullptr;
    int neig = 1, ncv = 0, maxit = 0;
    double tol = 1e-12, degen = 0.0;
    bool lowest = false;                    // lowest level and its copies only
    std::vector<dmed_vec *> vecs;           // out: nout eigenvectors, ascending
    std::vector<double> evals, resid;       // out
    int nout = 0, rc = 0, used = 0;         // out: rc 0 or DMED_ENOCONV
};
struct SectorQueue {
    std::vector<QueueItem> items;
    float kernel_ms = 0.f;                  // out: duration of the launch
    double *work = nullptr, *res = nullptr;
    double **ptrs = nullptr;
    int *counter = nullptr;
    void *tasks = nullptr;
    hipStream_t stream = nullptr;           // set once launched
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};
// dimension limit of the queue: kQueueMaxDim, or DMED_SECTOR_QUEUE_MAXDIM
int64_t queue_max_dim();
bool queue_takes(const dmed_sector *s, int neig, int ncv);
// launch on st (asynchronous); queue_finish waits, fills the items and frees
// the queue's buffers (also after a failed launch)
int queue_launch(SectorQueue &q, hipStream_t st);
int queue_finish(SectorQueue &q);

// ------------------------------------------------------------ host LA
// symmetric eigen-decomposition of a dense n x n matrix (column-major, in
// place -> eigenvectors), eigenvalues ascending
int host_syev(int n, double *a, double *w);
// in-place inverse of a small complex n x n matrix (row-major); false if singular
bool invert_small(int n, std::complex<double> *A);
// tql2 of the reference restricted to the first row of Z (the pole weights)
int tql2_first_row(int n, double *d, double *e, double *z1);
void gf_add_poles(double peso, double ei, const double *lam, const double *w2, int n,
                  int isign, const double *wm, int lmats, const double *wr, int lreal,
                  double eps, double *gm, double *gr);
// one branch of the susceptibility pole sums (build_chi_spin.f90:186-250,
// build_chi_dens.f90:505-733): complex weight pr + i pi, iv = 0 term signed
// by iv0_sign; vm[0..nv), tau[0..ntau), wr[0..lreal); chi_iv, chi_w
// interleaved complex, chi_tau real (takes the real part)
void chi_add_branch(double pr, double pi, double peso_bz, double ei, const double *lam,
                    const double *w2, int n, int isign, int iv0_sign, double beta,
                    const double *vm, int nv, const double *tau, int ntau, const double *wr,
                    int lreal, double eps, double *chi_iv, double *chi_tau, double *chi_w);
// add_to_lanczos_spinChi (build_chi_spin.f90:186-250) for isign = +1 then -1;
// vm[0..nv), tau[0..ntau), wr[0..lreal); chi_iv, chi_w interleaved complex
void chi_add_poles(double peso_f, double peso_bz, double ei, const double *lam, const double *w2,
                   int n, double beta, const double *vm, int nv, const double *tau, int ntau,
                   const double *wr, int lreal, double eps, double *chi_iv, double *chi_tau,
                   double *chi_w);

}  // namespace dmed
