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
#include <cstdio>
#include <map>
#include <memory>
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

// Run ntask tasks: serially on the calling thread, or (batched) on a pool of
// host workers, each bound to its own stream/scratch context, taking tasks
// from a shared counter.  Task results must not depend on which worker ran
// them.
int run_tasks(int batched, int ntask, const std::function<void(int)> &fn)
{
    if (!batched || ntask <= 1) {
        for (int k = 0; k < ntask; ++k) fn(k);
        return 0;
    }
    int nw = 8;
    if (const char *e = std::getenv("DMED_WORKERS")) nw = std::max(1, std::atoi(e));
    nw = std::min(nw, ntask);
    std::vector<Exec> execs(nw);
    for (auto &e : execs)
        if (int rc = exec_create(e)) {
            for (auto &x : execs) exec_destroy(x);
            return rc;
        }
    std::atomic<int> next(0);
    const int dev = ctx().device;
    std::vector<std::thread> th;
    for (int w = 0; w < nw; ++w)
        th.emplace_back([&, w]() {
            hipSetDevice(dev);
            bind_exec(&execs[w]);
            for (;;) {
                const int k = next.fetch_add(1);
                if (k >= ntask) break;
                fn(k);
            }
            hipStreamSynchronize(execs[w].stream);
            bind_exec(nullptr);
        });
    for (auto &t : th) t.join();
    for (auto &e : execs) exec_destroy(e);
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
    p->pad_ = 0;
}

extern "C" int dmed_solve(const dmed_solve_params *p, dmed_solve_out *out)
{
    Context &c = ctx();
    if (!c.initialized) return fail(DMED_ESTATE, "dmed_init not called");
    if (!c.has_model) return fail(DMED_ESTATE, "dmed_set_model not called");
    if (!p || !out) return fail(DMED_EINVAL, "null argument");
    if (p->prec_gf != DMED_FP64 && p->prec_gf != DMED_FP32 && p->prec_gf != DMED_MIXED)
        return fail(DMED_EINVAL, "prec_gf must be DMED_FP64, DMED_FP32 or DMED_MIXED");
    if (p->prec_gs != DMED_FP64 && p->prec_gs != DMED_MIXED)
        return fail(DMED_EINVAL, "prec_gs must be DMED_FP64 or DMED_MIXED");
    const Model &m = c.model;
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
    {
        std::vector<int> order(nsectors);
        for (int i = 0; i < nsectors; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return secs[x]->dim > secs[y]->dim; });
        if ((rc = run_tasks(p->batched, nsectors, [&](int k) { solve_sector(order[k]); })))
            return rc;
    }
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
    struct Chain {
        int s, a, state, dagger, isign;
        dmed_sector *J = nullptr;
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
        ch.rc = dmed_apply_op(st.vec, isite, ch.dagger, ch.J, &v, &ch.norm2);
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
    for (int s = 0; s < nspin && !rc; ++s)
        for (int a = 0; a < norb && !rc; ++a) {
            std::fill(gm.begin(), gm.end(), 0.0);
            std::fill(gr.begin(), gr.end(), 0.0);
            // pole sums of the chains of (s, a), in chain order for every
            // frequency, split over frequency blocks (deterministic either way)
            struct Term { const Chain *ch; double peso, e; };
            std::vector<Term> terms;
            for (auto &ch : chains) {
                if (ch.s != s || ch.a != a || !(ch.norm2 > 0.0)) continue;
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
    // ---------------------------------------------- spin susceptibility
    // build_chi_spin (build_chi_spin.f90:4-18, ED_MAIN.f90:282): for every
    // kept state, S_z of each orbital (rows 0..Norb-1) and, when Norb > 1, of
    // all orbitals (row Norb) in the state's own sector, one chain each, the
    // poles added for isign = +1 and -1; the weights carry 1/zeta and the
    // sums are divided by zeta once more at the end, as the reference does.
    // A zero S_z psi carries no weight (the reference divides by its norm).
    if (!rc && p->chi_spin) {
        const int ltau = std::max((int)beta, p->ltau);     // ED_INPUT_VARS.f90:191
        const int nrow = norb + 1, nv = lm + 1, nt = ltau + 1;
        std::vector<double> vm(nv), tau(nt);
        for (int i = 0; i < nv; ++i) vm[i] = M_PI / beta * 2.0 * (double)i;
        for (int i = 0; i < nt; ++i) tau[i] = ltau > 0 ? beta * (double)i / (double)ltau : 0.0;
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
