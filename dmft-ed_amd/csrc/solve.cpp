// solve.cpp -- the ed_solve_single sequence (ED_MAIN.f90:251-289) over the
// device engine:
//   diagonalize_impurity : ed_diag_c (ED_DIAG.f90:93-228) + ed_analysis (:248-398)
//   observables_impurity : dens / docc (ED_OBSERVABLES.f90:74-168)
//   buildgf_impurity     : build_gf_normal (build_gf_normal.f90:4-180) and
//                          get_sigma_normal (get_sigma_normal.f90:1-38)
// The state list keeps device vectors (handles), never host copies of them.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <functional>
#include <string>
#include <thread>
#include <cmath>
#include <complex>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "dmed_internal.h"

namespace dmed {

namespace {

struct State {
    double e;
    int sector;          // 0-based isector = nup*(Ns+1)+ndw
    dmed_vec *vec;
};

// Energy-ordered list with the insertion/pop rules of ED_EIGENSPACE.f90:
// a new state goes before existing states of equal energy (:160-163), pop
// removes the last state, emax is only raised by inserts (es_pop_state
// leaves it, :284-290).
struct StateList {
    std::vector<State> items;
    double emax = -HUGE_VAL, emin = HUGE_VAL;
    void free_all()
    {
        for (auto &s : items) dmed_vec_free(s.vec);
        items.clear();
        emax = -HUGE_VAL;
        emin = HUGE_VAL;
    }
    void insert(const State &s)
    {
        size_t k = 0;
        while (k < items.size() && !(s.e <= items[k].e)) ++k;
        items.insert(items.begin() + k, s);
        emax = std::max(emax, s.e);
        emin = std::min(emin, s.e);
    }
    // es_add_state_c (:126-148); returns false if the state was not kept
    bool add(const State &s, int size)
    {
        if (size <= 0 || (int)items.size() < size) {
            insert(s);
            return true;
        }
        if (s.e < items.back().e) {
            pop();
            insert(s);
            return true;
        }
        return false;
    }
    void pop()
    {
        dmed_vec_free(items.back().vec);
        items.pop_back();
    }
};

struct SectorCache {
    std::map<int, dmed_sector *> m;
    ~SectorCache()
    {
        for (auto &kv : m) dmed_sector_free(kv.second);
    }
    int get(int nup, int ndw, int ns, dmed_sector **out)
    {
        const int key = nup * (ns + 1) + ndw;
        auto it = m.find(key);
        if (it != m.end()) {
            *out = it->second;
            return 0;
        }
        dmed_sector *s = nullptr;
        if (int rc = dmed_sector_create(nup, ndw, &s)) return rc;
        m[key] = s;
        *out = s;
        return 0;
    }
    void drop(int key)
    {
        auto it = m.find(key);
        if (it != m.end()) {
            dmed_sector_free(it->second);
            m.erase(it);
        }
    }
};

inline size_t gidx(int s1, int s2, int a, int b, int l, int nspin, int norb)
{
    // (Nspin,Nspin,Norb,Norb,L) Fortran order
    return (size_t)s1 + nspin * ((size_t)s2 + nspin * ((size_t)a + norb * ((size_t)b + (size_t)norb * l)));
}

// Worker execution contexts, kept between calls.  Their streams and
// grow-only buffers are reused, so once the buffers have reached the sizes a
// workload needs, a batched phase allocates and frees no device memory while
// other workers run (hipMalloc and hipFree synchronise with the device).
// Contexts are handed out exclusively, so nested or concurrent callers get
// distinct ones; at most kPoolKeep idle contexts are kept, and the rest are
// destroyed on release.  exec_pool_clear (dmed_finalize) destroys them all.
constexpr int kPoolKeep = 16;
std::mutex g_pool_mu;
std::vector<Exec *> g_pool;

int pool_acquire(int nw, std::vector<Exec *> &out)
{
    std::lock_guard<std::mutex> lk(g_pool_mu);
    while ((int)out.size() < nw) {
        if (!g_pool.empty()) {
            out.push_back(g_pool.back());
            g_pool.pop_back();
            continue;
        }
        Exec *e = new Exec;
        if (int rc = exec_create(*e)) {
            exec_destroy(*e);
            delete e;
            return rc;
        }
        out.push_back(e);
    }
    return 0;
}

void pool_release(std::vector<Exec *> &v)
{
    std::lock_guard<std::mutex> lk(g_pool_mu);
    for (Exec *e : v) {
        if ((int)g_pool.size() < kPoolKeep) {
            g_pool.push_back(e);
        } else {
            exec_destroy(*e);
            delete e;
        }
    }
    v.clear();
}

}  // namespace

void exec_pool_clear()
{
    std::lock_guard<std::mutex> lk(g_pool_mu);
    for (Exec *e : g_pool) {
        exec_destroy(*e);
        delete e;
    }
    g_pool.clear();
}

namespace {

// Run ntask tasks: serially on the calling thread, or (batched) on host
// workers, each bound to its own stream/scratch context from the pool above,
// taking tasks from a shared counter.  Task results must not depend on which
// worker ran them.
int run_tasks(int batched, int ntask, const std::function<void(int)> &fn)
{
    if (!batched || ntask <= 1) {
        for (int k = 0; k < ntask; ++k) fn(k);
        return 0;
    }
    int nw = 8;
    if (const char *e = std::getenv("DMED_WORKERS")) nw = std::max(1, std::atoi(e));
    nw = std::min(nw, ntask);
    std::vector<Exec *> execs;
    if (int rc = pool_acquire(nw, execs)) {
        pool_release(execs);
        return rc;
    }
    std::atomic<int> next(0);
    const int dev = ctx().device;
    std::vector<std::thread> th;
    for (int w = 0; w < nw; ++w)
        th.emplace_back([&, w]() {
            hipSetDevice(dev);
            bind_exec(execs[w]);
            for (;;) {
                const int k = next.fetch_add(1);
                if (k >= ntask) break;
                fn(k);
            }
            hipStreamSynchronize(execs[w]->stream);
            bind_exec(nullptr);
        });
    for (auto &t : th) t.join();
    pool_release(execs);
    return 0;
}

// Host-only parallel loop (no device context): tasks write disjoint outputs,
// so the result does not depend on the thread count.
void host_parallel(int ntask, const std::function<void(int)> &fn)
{
    int nw = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    nw = std::min(nw, ntask);
    if (nw <= 1) {
        for (int k = 0; k < ntask; ++k) fn(k);
        return;
    }
    std::atomic<int> next(0);
    std::vector<std::thread> th;
    for (int w = 0; w < nw; ++w)
        th.emplace_back([&]() {
            for (int k; (k = next.fetch_add(1)) < ntask;) fn(k);
        });
    for (auto &t : th) t.join();
}

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// In-place inverse of a small complex n x n matrix (row-major), Gauss-Jordan
// with partial pivoting; false if singular.
bool invert_small(int n, std::complex<double> *A)
{
    using cd = std::complex<double>;
    std::vector<cd> M(2 * (size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) M[i * 2 * n + j] = A[i * n + j];
        M[i * 2 * n + n + i] = 1.0;
    }
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::abs(M[r * 2 * n + c]) > std::abs(M[piv * 2 * n + c])) piv = r;
        if (std::abs(M[piv * 2 * n + c]) == 0.0) return false;
        if (piv != c)
            for (int j = 0; j < 2 * n; ++j) std::swap(M[c * 2 * n + j], M[piv * 2 * n + j]);
        const cd inv = 1.0 / M[c * 2 * n + c];
        for (int j = 0; j < 2 * n; ++j) M[c * 2 * n + j] *= inv;
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const cd f = M[r * 2 * n + c];
            if (f == 0.0) continue;
            for (int j = 0; j < 2 * n; ++j) M[r * 2 * n + j] -= f * M[c * 2 * n + j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) A[i * n + j] = M[i * 2 * n + n + j];
    return true;
}

}  // namespace dmed

using namespace dmed;

extern "C" void dmed_solve_params_default(dmed_solve_params *p)
{
    p->lanc_nstates_sector = 6;
    p->lanc_nstates_total = 1;
    p->lanc_nstates_step = 2;
    p->lanc_niter = 512;
    p->lanc_ngfiter = 200;
    p->lanc_ncv_factor = 3;
    p->lanc_ncv_add = 5;
    p->lanc_dim_threshold = 256;
    p->lanc_tolerance = 1e-12;
    p->gs_threshold = 1e-9;
    p->cutoff = 1e-9;
    p->beta = 1000.0;
    p->eps = 0.01;
    p->wini = -5.0;
    p->wfin = 5.0;
    p->lmats = 5000;
    p->lreal = 5000;
    p->prec_gf = DMED_FP64;
    p->batched = 0;
    p->verbose = 0;
    p->gf_fp64_head = kMixedHead;
    p->prec_gs = DMED_FP64;
    p->chi_spin = 0;
    p->ltau = 1000;
    p->chi_dens = 0;
}

extern "C" int dmed_solve(const dmed_solve_params *p, dmed_solve_out *out)
{
    Context &c = ctx();
    if (!c.initialized) return fail(DMED_ESTATE, "dmed_init not called");
    if (!has_cur_model()) return fail(DMED_ESTATE, "dmed_set_model not called");
    if (!p || !out) return fail(DMED_EINVAL, "null argument");
    if (p->prec_gf != DMED_FP64 && p->prec_gf != DMED_FP32 && p->prec_gf != DMED_MIXED)
        return fail(DMED_EINVAL, "prec_gf must be DMED_FP64, DMED_FP32 or DMED_MIXED");
    if (p->prec_gs != DMED_FP64 && p->prec_gs != DMED_MIXED)
        return fail(DMED_EINVAL, "prec_gs must be DMED_FP64 or DMED_MIXED");
    const Model &m = cur_model();
    const int ns = m.ns, norb = m.norb, nspin = m.nspin;
    const int nsectors = (ns + 1) * (ns + 1);
    // init_ed_structure: finite T iff lanc_nstates_total != 1, and then the
    // per-sector and total counts are made even (ED_SETUP.f90:243-269)
    int nst_total = out->lanc_nstates_total ? *out->lanc_nstates_total : p->lanc_nstates_total;
    const bool finite_t = nst_total != 1;
    int nst_sector = p->lanc_nstates_sector;
    if (finite_t) {
        if (nst_sector % 2) ++nst_sector;
        if (nst_total % 2) ++nst_total;
    }
    std::vector<int> neigen(nsectors);
    for (int nu = 0; nu <= ns; ++nu)
        for (int nd = 0; nd <= ns; ++nd) {
            const int is = nu * (ns + 1) + nd;
            const int64_t dim = binom(ns, nu) * binom(ns, nd);
            neigen[is] = out->neigen_sector && out->neigen_sector[is] > 0
                             ? out->neigen_sector[is]
                             : (int)std::min<int64_t>(dim, nst_sector);
        }

    const double t0 = now();
    SectorCache cache;
    StateList sl;
    double oldzero = 1000.0;
    int rc = 0;
    // ------------------------------------------------------ ed_diag_c
    // Eigenpairs of every sector (serially, or concurrently on per-thread
    // streams when batched), then the state-list rules applied in the
    // reference's sector order, so both forms keep the same states.
    struct SectorRes {
        int nout = 0, rc = 0;
        std::vector<double> ev;
        std::vector<dmed_vec *> vecs;
        std::string err;
    };
    std::vector<SectorRes> sres(nsectors);
    std::vector<dmed_sector *> secs(nsectors, nullptr);
    const auto tsec0 = std::chrono::steady_clock::now();
    for (int nu = 0; nu <= ns && !rc; ++nu)
        for (int nd = 0; nd <= ns && !rc; ++nd)
            rc = cache.get(nu, nd, ns, &secs[nu * (ns + 1) + nd]);
    if (rc) return rc;
    if (p->verbose > 1)
        std::fprintf(stderr, "sector tables: %.3f s\n",
                     std::chrono::duration<double>(std::chrono::steady_clock::now() - tsec0).count());
    auto solve_sector = [&](int is) {
        const auto t0s = std::chrono::steady_clock::now();
        dmed_sector *sec = secs[is];
        const int64_t dim = sec->dim;
        const int ne = (int)std::min<int64_t>(dim, neigen[is]);
        const int nblock = (int)std::min<int64_t>(dim, p->lanc_ncv_factor * ne + p->lanc_ncv_add);
        SectorRes &r = sres[is];
        r.ev.assign(ne, 0.0);
        r.vecs.assign(ne, nullptr);
        std::vector<double> res(ne);
        // T=0 keeps only states within gs_threshold of the running minimum
        // (ED_DIAG.f90:202-213), so only the lowest level of a sector and its
        // degenerate copies can enter; finite T needs the ne lowest.
        r.rc = eigh_sector_ex(sec, ne, std::max(nblock, 2 * ne + 10), 50 * std::max(p->lanc_niter, 100),
                              p->lanc_tolerance, p->lanc_dim_threshold, 10.0 * p->gs_threshold,
                              !finite_t, &r.nout, r.ev.data(), r.vecs.data(), res.data(),
                              p->prec_gs);
        if (r.rc == DMED_ENOCONV) r.rc = 0;      // ARPACK returns its best pairs too
        if (r.rc) r.err = dmed_last_error();
        if (p->verbose > 1)
            std::fprintf(stderr, "sector %d dim=%lld neig=%d nout=%d %.3f ms\n", is, (long long)dim,
                         ne, r.nout,
                         1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now() - t0s).count());
    };
    // Nspin = 1 makes H symmetric under the spin flip, which maps sector
    // (nu, nd) onto (nd, nu) and an eigenvector psi(iu, id) onto the transposed
    // array (times the global sign (-1)^(nu*nd), immaterial for an
    // eigenvector).  A sector with nu < nd then takes the eigenvalues of its
    // twin and the transposed vectors instead of being diagonalised; this is
    // the twin-sector symmetry the reference offers behind ed_twin
    // (ED_SETUP.f90:429-436, ED_EIGENSPACE.f90:182-188), here always on for
    // Nspin = 1 unless DMED_TWIN=0.  Both sectors keep their own entries in
    // the state list, in the reference's sector order.
    const char *tw_env = std::getenv("DMED_TWIN");
    const bool twin = nspin == 1 && !(tw_env && tw_env[0] == '0');
    auto twin_of = [&](int is) { return (is % (ns + 1)) * (ns + 1) + is / (ns + 1); };
    std::vector<char> derived(nsectors, 0);
    for (int is = 0; is < nsectors && twin; ++is)
        if (is / (ns + 1) < is % (ns + 1) && neigen[is] == neigen[twin_of(is)]) derived[is] = 1;
    {
        std::vector<int> order;
        for (int i = 0; i < nsectors; ++i)
            if (!derived[i]) order.push_back(i);
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return secs[x]->dim > secs[y]->dim; });
        // Batched form: the small Krylov sectors go to one launch of the
        // persistent work-queue solver on a stream of their own, while the
        // host workers take the rest (largest first) on theirs.
        SectorQueue q;
        std::vector<int> qsec, rest;
        const char *q_env = std::getenv("DMED_SECTOR_QUEUE");
        const bool use_queue = p->batched && !(q_env && q_env[0] == '0');
        for (int is : order) {
            const int64_t dim = secs[is]->dim;
            const int ne = (int)std::min<int64_t>(dim, neigen[is]);
            const int nblock = (int)std::min<int64_t>(dim, p->lanc_ncv_factor * ne + p->lanc_ncv_add);
            const int ncv = std::max(nblock, 2 * ne + 10);
            if (use_queue && dim > p->lanc_dim_threshold && ne < dim &&
                queue_takes(secs[is], ne, ncv)) {
                QueueItem it;
                it.s = secs[is];
                it.neig = ne;
                it.ncv = ncv;
                it.maxit = 50 * std::max(p->lanc_niter, 100);
                it.tol = p->lanc_tolerance;
                it.degen = 10.0 * p->gs_threshold;
                it.lowest = !finite_t;
                q.items.push_back(std::move(it));
                qsec.push_back(is);
            } else {
                rest.push_back(is);
            }
        }
        hipStream_t qs = nullptr;
        const double tq0 = now();
        int qrc = 0;
        if (!q.items.empty()) {
            if (hipStreamCreateWithFlags(&qs, hipStreamNonBlocking) != hipSuccess)
                qrc = fail(DMED_EHIP, "sector queue: stream");
            if (!qrc) qrc = queue_launch(q, qs);
        }
        rc = run_tasks(p->batched, (int)rest.size(), [&](int k) { solve_sector(rest[k]); });
        const int frc = queue_finish(q);
        if (!qrc) qrc = frc;
        if (qs) hipStreamDestroy(qs);
        if (rc) return rc;
        if (p->verbose > 1 && !q.items.empty())
            std::fprintf(stderr, "sector queue: %d sectors, kernel %.3f ms; %.3f ms with the "
                                 "workers\n", (int)q.items.size(), q.kernel_ms, 1e3 * (now() - tq0));
        for (size_t k = 0; k < qsec.size(); ++k) {
            QueueItem &it = q.items[k];
            SectorRes &r = sres[qsec[k]];
            if (qrc) {
                // the queue could not run: these sectors take the host-driven path
                solve_sector(qsec[k]);
                continue;
            }
            r.nout = it.nout;
            r.ev = it.evals;
            r.vecs = it.vecs;
            r.rc = 0;                      // DMED_ENOCONV: best pairs kept, as ARPACK's
            if (p->verbose > 1)
                std::fprintf(stderr, "sector %d dim=%lld neig=%d nout=%d queue, %d H*v\n",
                             qsec[k], (long long)it.s->dim, it.neig, it.nout, it.used);
        }
    }
    for (int is = 0; is < nsectors && !rc; ++is) {
        if (!derived[is]) continue;
        const int it = twin_of(is);
        const SectorRes &t = sres[it];
        SectorRes &r = sres[is];
        if (t.rc) continue;                     // reported with the twin below
        r.nout = t.nout;
        r.ev = t.ev;
        r.vecs.assign(t.vecs.size(), nullptr);
        for (int i = 0; i < t.nout && !rc; ++i) {
            const dmed_sector *T = secs[it];
            if (t.vecs[i]->prec != DMED_FP64) {
                rc = fail(DMED_ESTATE, "twin sector: FP64 eigenvectors expected");
                break;
            }
            if ((rc = dmed_vec_create(secs[is], DMED_FP64, &r.vecs[i]))) break;
            rc = launch_transpose((const double *)t.vecs[i]->data, T->up.dim, T->dw.dim,
                                  (double *)r.vecs[i]->data, ex().stream);
        }
    }
    if (!rc && twin) DMED_HIP(hipStreamSynchronize(ex().stream));
    for (int is = 0; is < nsectors && !rc; ++is) {
        SectorRes &r = sres[is];
        if (r.rc) {
            rc = fail(r.rc, r.err);
            break;
        }
        if (p->verbose > 2 && r.nout > 0)
            std::fprintf(stderr, "sector %d (%d,%d) dim=%lld E0=%.12f\n", is + 1, is / (ns + 1),
                         is % (ns + 1), (long long)secs[is]->dim, r.ev[0]);
        for (int i = 0; i < r.nout; ++i) {
            State st{r.ev[i], is, r.vecs[i]};
            bool kept = false;
            if (finite_t) {
                kept = sl.add(st, nst_total);
            } else if (r.ev[i] < oldzero - 10.0 * p->gs_threshold) {
                oldzero = r.ev[i];
                sl.free_all();
                sl.insert(st);
                kept = true;
            } else if (std::fabs(r.ev[i] - oldzero) <= p->gs_threshold) {
                oldzero = std::min(oldzero, r.ev[i]);
                sl.insert(st);
                kept = true;
            }
            if (!kept) dmed_vec_free(r.vecs[i]);
            r.vecs[i] = nullptr;
        }
    }
    for (auto &r : sres)
        for (auto *v : r.vecs) dmed_vec_free(v);
    if (rc) {
        sl.free_all();
        return rc;
    }
    // ----------------------------------------------------- ed_analysis
    const double beta = p->beta;
    const double egs = sl.emin;
    double zeta = 0.0;
    if (finite_t)
        for (auto &s : sl.items) zeta += std::exp(-beta * (s.e - egs));
    else
        zeta = (double)sl.items.size();
    int numgs = 0;
    for (auto &s : sl.items)
        if (std::fabs(s.e - egs) <= p->gs_threshold) ++numgs;
    if (finite_t) {
        std::vector<int> count(nsectors, 0);
        for (auto &s : sl.items) count[s.sector]++;
        for (int i = 0; i < nsectors; ++i) {
            neigen[i] += count[i] > 0 ? 1 : -1;
            if (neigen[i] > count[i]) neigen[i] = count[i] + 1;
            if (neigen[i] <= 0) neigen[i] = 1;
        }
        if (std::exp(-beta * (sl.emax - egs)) > p->cutoff) {
            nst_total += p->lanc_nstates_step;
        } else {
            while (!sl.items.empty() && std::exp(-beta * (sl.items.back().e - egs)) <= p->cutoff)
                sl.pop();
            nst_total = std::max((int)sl.items.size(), p->lanc_nstates_step) + p->lanc_nstates_step;
        }
    }
    const double t1 = now();
    // ---------------------------------------------------- observables
    // observables_impurity / local_energy_impurity over the kept states
    dmed_local_obs obs{};
    for (auto &s : sl.items) {
        const double peso = (finite_t ? std::exp(-beta * (s.e - egs)) : 1.0) / zeta;
        if ((rc = dmed_local_observables(s.vec, peso, &obs))) break;
    }
    std::vector<double> dens(obs.dens, obs.dens + norb), docc(obs.docc, obs.docc + norb);
    // ------------------------------------------------ Green's functions
    const int lm = p->lmats, lr = p->lreal;
    std::vector<double> wm(lm), wr(lr);
    for (int i = 0; i < lm; ++i) wm[i] = M_PI / beta * (double)(2 * (i + 1) - 1);
    for (int i = 0; i < lr; ++i)
        wr[i] = lr > 1 ? p->wini + (p->wfin - p->wini) * (double)i / (double)(lr - 1) : p->wini;
    const size_t ng = (size_t)nspin * nspin * norb * norb;
    if (out->gmats) std::fill(out->gmats, out->gmats + 2 * ng * lm, 0.0);
    if (out->greal) std::fill(out->greal, out->greal + 2 * ng * lr, 0.0);
    std::vector<double> gm(2 * (size_t)lm), gr(2 * (size_t)lr);
    const int ngf = std::max(p->lanc_ngfiter, 1);
    // one task per (spin, orbital, state, particle/hole) chain, in the order
    // of lanc_build_gf_normal_c (build_gf_normal.f90:83-176)
    // kind 0: c_a psi (build_gf_normal.f90:83-176); for the hybrid bath and
    // a < b also (lanc_build_gf_normal_mix_c, :192-420) kind 1: (c_a + c_b) psi
    // and kind 2: the complex (c_a -/+ i c_b) psi, run as the stacked real
    // vector [c_a psi; c_b psi] on the stacked target sector
    struct Chain {
        int s, a, state, dagger, isign;
        int kind = 0, b = -1;
        dmed_sector *J = nullptr;
        dmed_sector *J0 = nullptr;      // kind 2: the unstacked target sector
        int nlanc = 0, rc = 0;
        double norm2 = 0.0;
        std::vector<double> al, bl;
        std::vector<double> lam, w2;    // poles: tql2 eigenvalues and Z(1,j)^2
        std::string err;
    };
    std::vector<Chain> chains;
    for (int s = 0; s < nspin && !rc; ++s)
        for (int a = 0; a < norb && !rc; ++a)
            for (int k = 0; k < (int)sl.items.size() && !rc; ++k) {
                const int nu = sl.items[k].sector / (ns + 1), nd = sl.items[k].sector % (ns + 1);
                for (int pass = 0; pass < 2; ++pass) {
                    const int dagger = pass == 0 ? 1 : 0;
                    const int ju = nu + (s == 0 ? (dagger ? 1 : -1) : 0);
                    const int jd = nd + (s == 1 ? (dagger ? 1 : -1) : 0);
                    if (ju < 0 || jd < 0 || ju > ns || jd > ns) continue;
                    Chain ch;
                    ch.s = s;
                    ch.a = a;
                    ch.state = k;
                    ch.dagger = dagger;
                    ch.isign = dagger ? 1 : -1;
                    if ((rc = cache.get(ju, jd, ns, &ch.J))) break;
                    ch.nlanc = (int)std::min<int64_t>(ch.J->dim, ngf);
                    chains.push_back(std::move(ch));
                }
            }
    const bool offdiag = m.hybrid && norb > 1;
    std::map<int, dmed_sector *> stacked_gf;
    for (int s = 0; s < nspin && !rc && offdiag; ++s)
        for (int a = 0; a < norb && !rc; ++a)
            for (int b = a + 1; b < norb && !rc; ++b)
                for (int k = 0; k < (int)sl.items.size() && !rc; ++k) {
                    const int nu = sl.items[k].sector / (ns + 1), nd = sl.items[k].sector % (ns + 1);
                    // (c^+_a + c^+_b) +1, (c_a + c_b) -1, (c^+_a + i c^+_b) +1, (c_a - i c_b) -1
                    for (int pass = 0; pass < 4 && !rc; ++pass) {
                        const int dagger = (pass % 2 == 0) ? 1 : 0;
                        const int ju = nu + (s == 0 ? (dagger ? 1 : -1) : 0);
                        const int jd = nd + (s == 1 ? (dagger ? 1 : -1) : 0);
                        if (ju < 0 || jd < 0 || ju > ns || jd > ns) continue;
                        Chain ch;
                        ch.s = s;
                        ch.a = a;
                        ch.b = b;
                        ch.state = k;
                        ch.dagger = dagger;
                        ch.isign = dagger ? 1 : -1;
                        ch.kind = pass < 2 ? 1 : 2;
                        if ((rc = cache.get(ju, jd, ns, &ch.J0))) break;
                        ch.J = ch.J0;
                        if (ch.kind == 2) {
                            const int key = ju * (ns + 1) + jd;
                            auto it = stacked_gf.find(key);
                            if (it == stacked_gf.end()) {
                                auto sec = std::make_unique<dmed_sector>();
                                if ((rc = sector_build_stacked(sec.get(), ju, jd))) {
                                    sector_release(sec.get());
                                    break;
                                }
                                it = stacked_gf.emplace(key, sec.release()).first;
                            }
                            ch.J = it->second;
                        }
                        ch.nlanc = (int)std::min<int64_t>(ch.J0->dim, ngf);
                        chains.push_back(std::move(ch));
                    }
                }
    auto run_chain = [&](int t) {
        Chain &ch = chains[t];
        const State &st = sl.items[ch.state];
        // start vector in a reusable buffer of the calling context
        dmed_vec v{ch.J, DMED_FP64, ch.J->dim, nullptr};
        if ((ch.rc = exec_buf(ex(), kSlotStart, std::max<int64_t>(ch.J->dim, 1) * 8, &v.data))) {
            ch.err = dmed_last_error();
            return;
        }
        const int isite = ch.a + 1 + ch.s * ns;
        if (ch.kind == 0) {
            ch.rc = dmed_apply_op(st.vec, isite, ch.dagger, ch.J, &v, &ch.norm2);
        } else {
            // c_a psi and c_b psi into the two halves of a 2*D buffer
            const int64_t d0 = ch.J0->dim;
            void *two = nullptr;
            if (ch.kind == 1) {
                if ((ch.rc = exec_buf(ex(), kSlotMid, std::max<int64_t>(2 * d0, 1) * 8, &two))) {
                    ch.err = dmed_last_error();
                    return;
                }
            } else {
                two = v.data;       // the stacked start vector itself (J->dim = 2*d0)
            }
            dmed_vec h1{ch.J0, DMED_FP64, d0, two};
            dmed_vec h2{ch.J0, DMED_FP64, d0, (double *)two + d0};
            double n1 = 0.0, n2 = 0.0;
            ch.rc = dmed_apply_op(st.vec, isite, ch.dagger, ch.J0, &h1, &n1);
            if (!ch.rc)
                ch.rc = dmed_apply_op(st.vec, ch.b + 1 + ch.s * ns, ch.dagger, ch.J0, &h2, &n2);
            if (!ch.rc && ch.kind == 1) {
                hipStream_t cs = ex().stream;
                ch.rc = launch_scale_copy(v.data, DMED_FP64, h1.data, DMED_FP64, d0, 1.0, cs);
                if (!ch.rc) ch.rc = launch_axpy(v.data, DMED_FP64, h2.data, DMED_FP64, d0, 1.0, cs);
                if (!ch.rc) ch.rc = ensure_partials(1 << 14);
                if (!ch.rc) ch.rc = dot_host((const double *)v.data, (const double *)v.data, d0, &ch.norm2);
            } else if (!ch.rc) {
                ch.norm2 = n1 + n2;
            }
        }
        if (!ch.rc && ch.norm2 > 0.0) {
            // zero norm: the operator annihilates the state and the chain
            // carries no weight (the reference divides by 0 here)
            ch.al.assign(ch.nlanc, 0.0);
            ch.bl.assign(ch.nlanc, 0.0);
            int nsteps = 0;
            ch.rc = chain_run(ch.J, v.data, DMED_FP64, ch.nlanc, p->prec_gf, 1e-13, ch.al.data(),
                              ch.bl.data(), &nsteps, p->gf_fp64_head);
            // poles of the chain (add_to_lanczos_gf_normal: tql2 with Z = I)
            if (!ch.rc) {
                ch.lam = ch.al;
                std::vector<double> e(ch.nlanc, 0.0);
                for (int i = 1; i < ch.nlanc; ++i) e[i] = ch.bl[i];
                ch.w2.assign(ch.nlanc, 0.0);
                if (tql2_first_row(ch.nlanc, ch.lam.data(), e.data(), ch.w2.data()))
                    ch.rc = fail(DMED_ENOCONV, "tql2 did not converge");
                for (auto &z : ch.w2) z *= z;
            }
        }
        if (ch.rc) ch.err = dmed_last_error();
    };
    if (!rc) rc = run_tasks(p->batched, (int)chains.size(), run_chain);
    for (auto &ch : chains)
        if (!rc && ch.rc) rc = fail(ch.rc, ch.err);
    for (auto &kv : stacked_gf) dmed_sector_free(kv.second);
    // full orbital matrices G[s][a][b] (hybrid bath), interleaved complex
    std::vector<double> gfm(offdiag ? 2 * (size_t)nspin * norb * norb * lm : 0),
        gfr(offdiag ? 2 * (size_t)nspin * norb * norb * lr : 0);
    for (int s = 0; s < nspin && !rc; ++s)
        for (int a = 0; a < norb && !rc; ++a) {
            std::fill(gm.begin(), gm.end(), 0.0);
            std::fill(gr.begin(), gr.end(), 0.0);
            // pole sums of the chains of (s, a), in chain order for every
            // frequency, split over frequency blocks (deterministic either way)
            struct Term { const Chain *ch; double peso, e; };
            std::vector<Term> terms;
            for (auto &ch : chains) {
                if (ch.kind != 0 || ch.s != s || ch.a != a || !(ch.norm2 > 0.0)) continue;
                const State &st = sl.items[ch.state];
                double peso = 0.0;
                if (finite_t) {
                    if (beta * (st.e - egs) < 200) peso = ch.norm2 * std::exp(-beta * (st.e - egs)) / zeta;
                } else {
                    peso = ch.norm2 / zeta;
                }
                terms.push_back({&ch, peso, st.e});
            }
            const int nfreq = lm + lr, nblk = std::max(1, std::min(64, nfreq / 64));
            host_parallel(nblk, [&](int b) {
                const int f0 = (int)((int64_t)nfreq * b / nblk), f1 = (int)((int64_t)nfreq * (b + 1) / nblk);
                const int m0 = std::min(f0, lm), m1 = std::min(f1, lm);
                const int r0 = std::max(f0, lm) - lm, r1 = std::max(f1, lm) - lm;
                for (const Term &t : terms)
                    gf_add_poles(t.peso, t.e, t.ch->lam.data(), t.ch->w2.data(), t.ch->nlanc,
                                 t.ch->isign, wm.data() + m0, m1 - m0, wr.data() + r0, r1 - r0,
                                 p->eps, gm.data() + 2 * m0, gr.data() + 2 * r0);
            });
            for (int l = 0; l < lm && out->gmats; ++l) {
                const size_t o = gidx(s, s, a, a, l, nspin, norb);
                out->gmats[2 * o] = gm[2 * l];
                out->gmats[2 * o + 1] = gm[2 * l + 1];
            }
            for (int l = 0; l < lr && out->greal; ++l) {
                const size_t o = gidx(s, s, a, a, l, nspin, norb);
                out->greal[2 * o] = gr[2 * l];
                out->greal[2 * o + 1] = gr[2 * l + 1];
            }
            if (offdiag) {
                // Sigma is an orbital matrix: after the off-diagonal elements
                std::copy(gm.begin(), gm.end(), gfm.begin() + 2 * (size_t)lm * ((s * norb + a) * norb + a));
                std::copy(gr.begin(), gr.end(), gfr.begin() + 2 * (size_t)lr * ((s * norb + a) * norb + a));
                continue;
            }
            // Sigma = G0^-1 - G^-1 on both axes
            std::vector<double> z(2 * (size_t)std::max(lm, lr)), sg(2 * (size_t)std::max(lm, lr));
            if (out->smats) {
                for (int l = 0; l < lm; ++l) { z[2 * l] = 0.0; z[2 * l + 1] = wm[l]; }
                if ((rc = dmed_sigma(s, a, z.data(), gm.data(), lm, sg.data()))) break;
                for (int l = 0; l < lm; ++l) {
                    const size_t o = gidx(s, s, a, a, l, nspin, norb);
                    out->smats[2 * o] = sg[2 * l];
                    out->smats[2 * o + 1] = sg[2 * l + 1];
                }
            }
            if (out->sreal) {
                for (int l = 0; l < lr; ++l) { z[2 * l] = wr[l]; z[2 * l + 1] = p->eps; }
                if ((rc = dmed_sigma(s, a, z.data(), gr.data(), lr, sg.data()))) break;
                for (int l = 0; l < lr; ++l) {
                    const size_t o = gidx(s, s, a, a, l, nspin, norb);
                    out->sreal[2 * o] = sg[2 * l];
                    out->sreal[2 * o + 1] = sg[2 * l + 1];
                }
            }
        }
    // ------------------------- hybrid bath: off-diagonal G and matrix Sigma
    // add_to_lanczos_gf_normal (build_gf_normal.f90:442-488) of the kind-1
    // chains with weight norm^2 and of the kind-2 chains with weight
    // -i*norm^2, then G_ab = (G_ab - (1-i) G_aa - (1-i) G_bb)/2, G_ba = G_ab
    // (:40-53), and Sigma = G0^-1 - G^-1 as Norb x Norb matrices
    // (get_sigma_normal.f90:40-56) with G0^-1 = (z + xmu) 1 - Hloc - Delta,
    // Delta_ab = sum_k V_a(k) V_b(k)/(z - e_k) (ED_BATH_FUNCTIONS.f90:1836-1848, 294-310).
    for (int s = 0; s < nspin && !rc && offdiag; ++s) {
        for (int a = 0; a < norb; ++a)
            for (int b = a + 1; b < norb; ++b) {
                std::vector<double> am(2 * (size_t)lm, 0.0), ar(2 * (size_t)lr, 0.0),
                    tm(2 * (size_t)lm), tr(2 * (size_t)lr);
                for (auto &ch : chains) {
                    if (ch.kind == 0 || ch.s != s || ch.a != a || ch.b != b || !(ch.norm2 > 0.0))
                        continue;
                    const State &st = sl.items[ch.state];
                    double peso = 0.0;
                    if (finite_t) {
                        if (beta * (st.e - egs) < 200) peso = ch.norm2 * std::exp(-beta * (st.e - egs)) / zeta;
                    } else {
                        peso = ch.norm2 / zeta;
                    }
                    if (ch.kind == 1) {
                        gf_add_poles(peso, st.e, ch.lam.data(), ch.w2.data(), ch.nlanc, ch.isign,
                                     wm.data(), lm, wr.data(), lr, p->eps, am.data(), ar.data());
                        continue;
                    }
                    // weight -i*peso: (x + i y) += -i (tx + i ty) = ty - i tx
                    std::fill(tm.begin(), tm.end(), 0.0);
                    std::fill(tr.begin(), tr.end(), 0.0);
                    gf_add_poles(peso, st.e, ch.lam.data(), ch.w2.data(), ch.nlanc, ch.isign,
                                 wm.data(), lm, wr.data(), lr, p->eps, tm.data(), tr.data());
                    for (int l = 0; l < lm; ++l) { am[2 * l] += tm[2 * l + 1]; am[2 * l + 1] -= tm[2 * l]; }
                    for (int l = 0; l < lr; ++l) { ar[2 * l] += tr[2 * l + 1]; ar[2 * l + 1] -= tr[2 * l]; }
                }
                auto combine = [&](std::vector<double> &g, const std::vector<double> &full, int L) {
                    const double *ga = full.data() + 2 * (size_t)L * ((s * norb + a) * norb + a);
                    const double *gb = full.data() + 2 * (size_t)L * ((s * norb + b) * norb + b);
                    for (int l = 0; l < L; ++l) {
                        // (1 - i)(x + i y) = (x + y) + i (y - x)
                        const double x = ga[2 * l] + gb[2 * l], y = ga[2 * l + 1] + gb[2 * l + 1];
                        g[2 * l] = 0.5 * (g[2 * l] - (x + y));
                        g[2 * l + 1] = 0.5 * (g[2 * l + 1] - (y - x));
                    }
                };
                combine(am, gfm, lm);
                combine(ar, gfr, lr);
                for (int pr = 0; pr < 2; ++pr) {
                    const int i = pr ? b : a, j = pr ? a : b;
                    std::copy(am.begin(), am.end(), gfm.begin() + 2 * (size_t)lm * ((s * norb + i) * norb + j));
                    std::copy(ar.begin(), ar.end(), gfr.begin() + 2 * (size_t)lr * ((s * norb + i) * norb + j));
                    for (int l = 0; l < lm && out->gmats; ++l) {
                        const size_t o = gidx(s, s, i, j, l, nspin, norb);
                        out->gmats[2 * o] = am[2 * l];
                        out->gmats[2 * o + 1] = am[2 * l + 1];
                    }
                    for (int l = 0; l < lr && out->greal; ++l) {
                        const size_t o = gidx(s, s, i, j, l, nspin, norb);
                        out->greal[2 * o] = ar[2 * l];
                        out->greal[2 * o + 1] = ar[2 * l + 1];
                    }
                }
            }
        using cd = std::complex<double>;
        const int sb = m.nspin == 2 ? s : 0;
        auto sigma_axis = [&](const std::vector<double> &full, int L, bool mats, double *dst) {
            std::vector<cd> G(norb * norb), S(norb * norb);
            for (int l = 0; l < L; ++l) {
                const cd z = mats ? cd(0.0, wm[l]) : cd(wr[l], p->eps);
                for (int a = 0; a < norb; ++a)
                    for (int b = 0; b < norb; ++b) {
                        const double *g = full.data() + 2 * ((size_t)L * ((s * norb + a) * norb + b) + l);
                        G[a * norb + b] = cd(g[0], g[1]);
                    }
                if (!invert_small(norb, G.data())) {
                    rc = fail(DMED_ENOCONV, "singular G matrix in Sigma = G0^-1 - G^-1");
                    return;
                }
                for (int a = 0; a < norb; ++a)
                    for (int b = 0; b < norb; ++b) {
                        cd d = 0.0;
                        for (int k = 0; k < m.nbath; ++k)
                            d += m.V(sb, a, k) * m.V(sb, b, k) / (z - m.E(sb, a, k));
                        const cd g0 = (a == b ? z + m.xmu : cd(0.0)) - m.hloc[sb][a][b] - d;
                        const cd sg = g0 - G[a * norb + b];
                        if (dst) {
                            const size_t o = gidx(s, s, a, b, l, nspin, norb);
                            dst[2 * o] = sg.real();
                            dst[2 * o + 1] = sg.imag();
                        }
                    }
            }
        };
        sigma_axis(gfm, lm, true, out->smats);
        if (!rc) sigma_axis(gfr, lr, false, out->sreal);
    }
    // ---------------------------------------------- spin susceptibility
    // build_chi_spin (build_chi_spin.f90:4-18, ED_MAIN.f90:282): for every
    // kept state, S_z of each orbital (rows 0..Norb-1) and, when Norb > 1, of
    // all orbitals (row Norb) in the state's own sector, one chain each, the
    // poles added for isign = +1 and -1; the weights carry 1/zeta and the
    // sums are divided by zeta once more at the end, as the reference does.
    // A zero S_z psi carries no weight (the reference divides by its norm).
    // bosonic frequencies vm = 2 pi n/beta (n = 0..Lmats) and tau =
    // linspace(0, beta, Ltau+1) of allocate_grids (ED_GREENS_FUNCTIONS.f90:281-296)
    const int ltau = std::max((int)beta, p->ltau);         // ED_INPUT_VARS.f90:191
    const int nv = lm + 1, nt = ltau + 1;
    std::vector<double> vm, tau;
    if (p->chi_spin || p->chi_dens) {
        vm.resize(nv);
        tau.resize(nt);
        for (int i = 0; i < nv; ++i) vm[i] = M_PI / beta * 2.0 * (double)i;
        for (int i = 0; i < nt; ++i) tau[i] = ltau > 0 ? beta * (double)i / (double)ltau : 0.0;
    }
    for (auto &it : sl.items)
        if ((p->chi_spin || p->chi_dens) && !rc && it.vec->prec != DMED_FP64)
            rc = fail(DMED_ESTATE, "susceptibility start vectors need FP64 states");
    if (!rc && p->chi_spin) {
        const int nrow = norb + 1;
        struct ChiTask {
            int state, row;
            int nlanc = 0, rc = 0;
            double norm2 = 0.0;
            std::vector<double> lam, w2;
            std::string err;
        };
        std::vector<ChiTask> tasks;
        for (int k = 0; k < (int)sl.items.size(); ++k)
            for (int row = 0; row < (norb > 1 ? nrow : norb); ++row) {
                ChiTask t;
                t.state = k;
                t.row = row;
                tasks.push_back(std::move(t));
            }
        auto run_chi = [&](int t) {
            ChiTask &ck = tasks[t];
            const State &st = sl.items[ck.state];
            const dmed_sector *S = st.vec->sec;
            dmed_vec v{S, DMED_FP64, S->dim, nullptr};
            if ((ck.rc = exec_buf(ex(), kSlotStart, std::max<int64_t>(S->dim, 1) * 8, &v.data))) {
                ck.err = dmed_last_error();
                return;
            }
            ck.rc = dmed_apply_sz(st.vec, ck.row < norb ? ck.row : -1, &v, &ck.norm2);
            if (!ck.rc && ck.norm2 > 0.0) {
                ck.nlanc = (int)std::min<int64_t>(S->dim, ngf);
                std::vector<double> al(ck.nlanc, 0.0), bl(ck.nlanc, 0.0);
                int nsteps = 0;
                ck.rc = chain_run(S, v.data, DMED_FP64, ck.nlanc, p->prec_gf, 1e-13, al.data(),
                                  bl.data(), &nsteps, p->gf_fp64_head);
                if (!ck.rc) {
                    ck.lam = al;
                    std::vector<double> e(ck.nlanc, 0.0);
                    for (int i = 1; i < ck.nlanc; ++i) e[i] = bl[i];
                    ck.w2.assign(ck.nlanc, 0.0);
                    if (tql2_first_row(ck.nlanc, ck.lam.data(), e.data(), ck.w2.data()))
                        ck.rc = fail(DMED_ENOCONV, "tql2 did not converge");
                    for (auto &z : ck.w2) z *= z;
                }
            }
            if (ck.rc) ck.err = dmed_last_error();
        };
        rc = run_tasks(p->batched, (int)tasks.size(), run_chi);
        for (auto &ck : tasks)
            if (!rc && ck.rc) rc = fail(ck.rc, ck.err);
        if (!rc) {
            std::vector<double> civ(2 * (size_t)nrow * nv, 0.0), ctau((size_t)nrow * nt, 0.0),
                cw(2 * (size_t)nrow * lr, 0.0);
            for (auto &ck : tasks) {
                if (!(ck.norm2 > 0.0)) continue;
                const State &st = sl.items[ck.state];
                const double peso_f = ck.norm2 / zeta;
                const double peso_bz = finite_t ? std::exp(-beta * (st.e - egs)) : 1.0;
                chi_add_poles(peso_f, peso_bz, st.e, ck.lam.data(), ck.w2.data(), ck.nlanc, beta,
                              vm.data(), nv, tau.data(), nt, wr.data(), lr, p->eps,
                              civ.data() + 2 * (size_t)ck.row * nv, ctau.data() + (size_t)ck.row * nt,
                              cw.data() + 2 * (size_t)ck.row * lr);
            }
            // Fortran order (Norb+1, L): element (row, i) at row + nrow*i
            for (int row = 0; row < nrow; ++row) {
                for (int i = 0; i < nv && out->spin_chi_iv; ++i) {
                    out->spin_chi_iv[2 * ((size_t)row + nrow * i)] = civ[2 * ((size_t)row * nv + i)] / zeta;
                    out->spin_chi_iv[2 * ((size_t)row + nrow * i) + 1] = civ[2 * ((size_t)row * nv + i) + 1] / zeta;
                }
                for (int i = 0; i < nt && out->spin_chi_tau; ++i)
                    out->spin_chi_tau[(size_t)row + nrow * i] = ctau[(size_t)row * nt + i] / zeta;
                for (int i = 0; i < lr && out->spin_chi_w; ++i) {
                    out->spin_chi_w[2 * ((size_t)row + nrow * i)] = cw[2 * ((size_t)row * lr + i)] / zeta;
                    out->spin_chi_w[2 * ((size_t)row + nrow * i) + 1] = cw[2 * ((size_t)row * lr + i) + 1] / zeta;
                }
            }
            out->ltau_used = ltau;
        }
    }
    // -------------------------------------------- charge susceptibility
    // build_chi_dens (build_chi_dens.f90:4-41), chains in the order of the
    // reference, every kept state in each:
    //   diag    a:     N_a psi, isign +1 and -1 into dens(a,a)          (:49-110)
    //   offdiag a < b: (N_a + N_b) psi, +1 and -1 into dens(a,b);
    //                  (N_a - i N_b) psi, weight i*norm^2, +1 and
    //                  (N_a + i N_b) psi, weight i*norm^2, -1           (:188-323)
    //   mix     a, b:  c^+_{b,s} c_{a,s} psi, +1, then c^+_{a,s} c_{b,s} psi,
    //                  -1, for each spin s < Nspin, into mix(a,b)      (:339-490)
    //   tot:           N psi, +1 and -1 into tot                        (:118-178)
    // with N_a = n_a,up + n_a,dw; offdiag, mix and tot only when Norb > 1.
    // Under the real H the two complex chains have the Lanczos coefficients
    // of the stacked real vector [N_a psi; N_b psi] under H (+) H (the
    // Lanczos vectors are p_k(H) applied to both parts, with norms
    // ||p_k(H) N_a psi||^2 + ||p_k(H) N_b psi||^2), so one chain on the stacked
    // sector serves both.  The pole sums are add_to_lanczos_densChi(_mix,
    // _tot) (:505-733); dens(a,b) real frequencies of a < b then become
    // 0.5*(w(a,b) - (1+i) w(a,a) - (1+i) w(b,b)) (:19-23), and only the dens
    // arrays are divided by zeta a second time (:36-38).  A zero start
    // vector carries no weight (the reference divides by its norm), and a
    // mix chain whose c_{a,s} leaves the sector range is skipped (the
    // reference's getCsector returns 0 there and the loop has no valid
    // vector).
    if (!rc && p->chi_dens) {
        enum { kOcc = 0, kStacked = 1, kMix = 2 };
        enum { kDens = 0, kMixArr = 1, kTot = 2 };
        struct DTask {
            int kind, state, target, a, b;
            int spin = 0, isign = 0;        // mix: operator spin; isign 0 = both branches
            bool imag = false;              // weight i*norm^2
            unsigned mask = 0;
            dmed_sector *S = nullptr;       // sector of the chain
            dmed_sector *M = nullptr;       // mix: sector after c_{.,s}
            int nlanc = 0, rc = 0;
            double norm2 = 0.0;
            std::vector<double> lam, w2;
            std::string err;
        };
        std::vector<DTask> dt;
        std::map<int, dmed_sector *> stacked;
        auto state_sector = [&](int k) { return const_cast<dmed_sector *>(sl.items[k].vec->sec); };
        const int nst = (int)sl.items.size();
        for (int a = 0; a < norb; ++a)
            for (int k = 0; k < nst; ++k) {
                DTask t{kOcc, k, kDens, a, a};
                t.mask = 1u << a;
                t.S = state_sector(k);
                dt.push_back(std::move(t));
            }
        if (norb > 1) {
            for (int a = 0; a < norb && !rc; ++a)
                for (int b = a + 1; b < norb && !rc; ++b)
                    for (int k = 0; k < nst && !rc; ++k) {
                        DTask t{kOcc, k, kDens, a, b};
                        t.mask = (1u << a) | (1u << b);
                        t.S = state_sector(k);
                        dt.push_back(std::move(t));
                        DTask u{kStacked, k, kDens, a, b};
                        u.imag = true;
                        const int key = sl.items[k].sector;
                        auto it = stacked.find(key);
                        if (it == stacked.end()) {
                            auto sec = std::make_unique<dmed_sector>();
                            if ((rc = sector_build_stacked(sec.get(), key / (ns + 1), key % (ns + 1)))) {
                                sector_release(sec.get());
                                break;
                            }
                            it = stacked.emplace(key, sec.release()).first;
                        }
                        u.S = it->second;
                        dt.push_back(std::move(u));
                    }
            for (int a = 0; a < norb && !rc; ++a)
                for (int b = 0; b < norb && !rc; ++b)
                    for (int k = 0; k < nst && !rc; ++k) {
                        const int nu = sl.items[k].sector / (ns + 1), nd = sl.items[k].sector % (ns + 1);
                        for (int pass = 0; pass < 2 && !rc; ++pass)
                            for (int sp = 0; sp < nspin && !rc; ++sp) {
                                const int mu = nu - (sp == 0), md = nd - (sp == 1);
                                if (mu < 0 || md < 0) continue;
                                // pass 0: c^+_b c_a, isign +1; pass 1: c^+_a c_b, isign -1
                                DTask t{kMix, k, kMixArr, pass == 0 ? a : b, pass == 0 ? b : a};
                                t.spin = sp;
                                t.isign = pass == 0 ? 1 : -1;
                                t.S = state_sector(k);
                                if ((rc = cache.get(mu, md, ns, &t.M))) break;
                                dt.push_back(std::move(t));
                            }
                    }
            for (int k = 0; k < nst; ++k) {
                DTask t{kOcc, k, kTot, 0, 0};
                t.mask = (1u << norb) - 1u;
                t.S = state_sector(k);
                dt.push_back(std::move(t));
            }
        }
        auto run_dchi = [&](int t) {
            DTask &dk = dt[t];
            const State &st = sl.items[dk.state];
            const dmed_sector *I = st.vec->sec;
            void *buf = nullptr;
            if ((dk.rc = exec_buf(ex(), kSlotStart, std::max<int64_t>(dk.S->dim, 1) * 8, &buf))) {
                dk.err = dmed_last_error();
                return;
            }
            if (dk.kind == kOcc) {
                dk.rc = apply_occ(I, DMED_FP64, st.vec->data, dk.mask, 1.0, 1.0, buf, &dk.norm2);
            } else if (dk.kind == kStacked) {
                double na = 0.0, nb = 0.0;
                dk.rc = apply_occ(I, DMED_FP64, st.vec->data, 1u << dk.a, 1.0, 1.0, buf, &na);
                if (!dk.rc)
                    dk.rc = apply_occ(I, DMED_FP64, st.vec->data, 1u << dk.b, 1.0, 1.0,
                                      (double *)buf + I->dim, &nb);
                dk.norm2 = na + nb;
            } else {
                dmed_vec mid{dk.M, DMED_FP64, dk.M->dim, nullptr};
                dmed_vec out{dk.S, DMED_FP64, dk.S->dim, buf};
                if ((dk.rc = exec_buf(ex(), kSlotMid, std::max<int64_t>(dk.M->dim, 1) * 8, &mid.data))) {
                    dk.err = dmed_last_error();
                    return;
                }
                double n1 = 0.0;
                dk.rc = dmed_apply_op(st.vec, dk.a + 1 + dk.spin * ns, 0, dk.M, &mid, &n1);
                if (!dk.rc) dk.rc = dmed_apply_op(&mid, dk.b + 1 + dk.spin * ns, 1, dk.S, &out, &dk.norm2);
            }
            if (!dk.rc && dk.norm2 > 0.0) {
                dk.nlanc = (int)std::min<int64_t>(dk.kind == kStacked ? I->dim : dk.S->dim, ngf);
                std::vector<double> al(dk.nlanc, 0.0), bl(dk.nlanc, 0.0);
                int nsteps = 0;
                dk.rc = chain_run(dk.S, buf, DMED_FP64, dk.nlanc, p->prec_gf, 1e-13, al.data(),
                                  bl.data(), &nsteps, p->gf_fp64_head);
                if (!dk.rc) {
                    dk.lam = al;
                    std::vector<double> e(dk.nlanc, 0.0);
                    for (int i = 1; i < dk.nlanc; ++i) e[i] = bl[i];
                    dk.w2.assign(dk.nlanc, 0.0);
                    if (tql2_first_row(dk.nlanc, dk.lam.data(), e.data(), dk.w2.data()))
                        dk.rc = fail(DMED_ENOCONV, "tql2 did not converge");
                    for (auto &z : dk.w2) z *= z;
                }
            }
            if (dk.rc) dk.err = dmed_last_error();
        };
        if (!rc) rc = run_tasks(p->batched, (int)dt.size(), run_dchi);
        for (auto &dk : dt)
            if (!rc && dk.rc) rc = fail(dk.rc, dk.err);
        for (auto &kv : stacked) dmed_sector_free(kv.second);
        if (!rc) {
            const size_t nab = (size_t)norb * norb;
            // [target][a*norb+b][freq]; tot uses slot 0
            std::vector<double> civ[3], ctau[3], cw[3];
            for (int q = 0; q < 3; ++q) {
                civ[q].assign(2 * nab * nv, 0.0);
                ctau[q].assign(nab * nt, 0.0);
                cw[q].assign(2 * nab * lr, 0.0);
            }
            for (auto &dk : dt) {
                if (!(dk.norm2 > 0.0)) continue;
                const State &st = sl.items[dk.state];
                const double pf = dk.norm2 / zeta;
                const double pbz = finite_t ? std::exp(-beta * (st.e - egs)) : 1.0;
                // mix chains run c^+_{b} c_{a} with (a, b) swapped on the -1 pass
                const size_t ab = dk.target == kTot ? 0
                                  : dk.kind == kMix && dk.isign < 0 ? (size_t)dk.b * norb + dk.a
                                                                   : (size_t)dk.a * norb + dk.b;
                const int q = dk.target;
                for (int isign = 1; isign >= -1; isign -= 2) {
                    if (dk.isign != 0 && isign != dk.isign) continue;
                    chi_add_branch(dk.imag ? 0.0 : pf, dk.imag ? pf : 0.0, pbz, st.e, dk.lam.data(),
                                   dk.w2.data(), dk.nlanc, isign, -isign, beta, vm.data(), nv,
                                   tau.data(), nt, wr.data(), lr, p->eps, civ[q].data() + 2 * ab * nv,
                                   ctau[q].data() + ab * nt, cw[q].data() + 2 * ab * lr);
                }
            }
            // w(a,b) = 0.5*(w(a,b) - (1+i) w(a,a) - (1+i) w(b,b)), a < b
            for (int a = 0; a < norb; ++a)
                for (int b = a + 1; b < norb; ++b) {
                    double *wab = cw[kDens].data() + 2 * ((size_t)a * norb + b) * lr;
                    const double *waa = cw[kDens].data() + 2 * ((size_t)a * norb + a) * lr;
                    const double *wbb = cw[kDens].data() + 2 * ((size_t)b * norb + b) * lr;
                    for (int i = 0; i < lr; ++i) {
                        const double xr = waa[2 * i] + wbb[2 * i], xi = waa[2 * i + 1] + wbb[2 * i + 1];
                        const double re = wab[2 * i] - (xr - xi), im = wab[2 * i + 1] - (xr + xi);
                        wab[2 * i] = 0.5 * re;
                        wab[2 * i + 1] = 0.5 * im;
                    }
                }
            // Fortran order (Norb, Norb, L): element (a, b, i) at a + Norb*(b + Norb*i)
            double *oiv[3] = {out->dens_chi_iv, out->dens_chi_mix_iv, out->dens_chi_tot_iv};
            double *otau[3] = {out->dens_chi_tau, out->dens_chi_mix_tau, out->dens_chi_tot_tau};
            double *ow[3] = {out->dens_chi_w, out->dens_chi_mix_w, out->dens_chi_tot_w};
            for (int q = 0; q < 3; ++q) {
                const double sc = q == kDens ? zeta : 1.0;
                const int na = q == kTot ? 1 : norb;
                for (int a = 0; a < na; ++a)
                    for (int b = 0; b < na; ++b) {
                        const size_t ab = (size_t)a * norb + b;
                        const size_t fo = (size_t)a + (size_t)na * b, fs = (size_t)na * na;
                        for (int i = 0; i < nv && oiv[q]; ++i) {
                            oiv[q][2 * (fo + fs * i)] = civ[q][2 * (ab * nv + i)] / sc;
                            oiv[q][2 * (fo + fs * i) + 1] = civ[q][2 * (ab * nv + i) + 1] / sc;
                        }
                        for (int i = 0; i < nt && otau[q]; ++i)
                            otau[q][fo + fs * i] = ctau[q][ab * nt + i] / sc;
                        for (int i = 0; i < lr && ow[q]; ++i) {
                            ow[q][2 * (fo + fs * i)] = cw[q][2 * (ab * lr + i)] / sc;
                            ow[q][2 * (fo + fs * i) + 1] = cw[q][2 * (ab * lr + i) + 1] / sc;
                        }
                    }
            }
            out->ltau_used = ltau;
        }
    }
    const double t2 = now();
    out->egs = egs;
    out->zeta = zeta;
    out->nstates = (int)sl.items.size();
    out->numgs = numgs;
    if (out->dens) std::copy(dens.begin(), dens.end(), out->dens);
    if (out->docc) std::copy(docc.begin(), docc.end(), out->docc);
    if (out->obs) *out->obs = obs;
    if (out->neigen_sector) std::copy(neigen.begin(), neigen.end(), out->neigen_sector);
    if (out->lanc_nstates_total) *out->lanc_nstates_total = nst_total;
    out->t_diag = t1 - t0;
    out->t_gf = t2 - t1;
    sl.free_all();
    return rc;
}

// ed_solve_lattice (ED_MAIN.f90:364-468): independent impurity problems of
// inequivalent sites.  With batched set and more than one site, the sites
// run concurrently on host workers, each with its own stream, scratch and
// bound model, and each site's solve runs its chains serially; otherwise the
// sites run one after another with the given params.  A site's results do
// not depend on which worker ran it.  The context's model (dmed_set_model)
// is left as it was.
extern "C" int dmed_solve_sites(int32_t nsites, const dmed_model *models,
                                const dmed_solve_params *p, dmed_solve_out *outs)
{
    if (!ctx().initialized) return fail(DMED_ESTATE, "dmed_init not called");
    if (nsites < 1 || !models || !p || !outs) return fail(DMED_EINVAL, "null or empty argument");
    std::vector<Model> ms(nsites);
    for (int i = 0; i < nsites; ++i)
        if (int rc = model_fill(&models[i], ms[i]))
            return fail(rc, "site " + std::to_string(i) + ": " + dmed_last_error());
    if (int rc = upload_binom()) return rc;
    const bool conc = p->batched && nsites > 1;
    dmed_solve_params q = *p;
    if (conc) q.batched = 0;
    std::vector<int> rcs(nsites, 0);
    std::vector<std::string> errs(nsites);
    // concurrent sites build sectors and allocate while others capture chain
    // steps: captures are thread-local and every host-to-device copy of a
    // solve goes on the worker's own stream, so neither disturbs the other
    int rc = run_tasks(conc ? 1 : 0, nsites, [&](int i) {
        bind_model(&ms[i]);
        rcs[i] = dmed_solve(&q, &outs[i]);
        if (rcs[i]) errs[i] = dmed_last_error();
        bind_model(nullptr);
    });
    if (rc) return rc;
    for (int i = 0; i < nsites; ++i)
        if (rcs[i]) return fail(rcs[i], "site " + std::to_string(i) + ": " + errs[i]);
    return 0;
}
