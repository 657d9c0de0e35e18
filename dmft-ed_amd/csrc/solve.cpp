// solve.cpp -- the ed_solve_single sequence (ED_MAIN.f90:251-289) over the
// device engine:
//   diagonalize_impurity : ed_diag_c (ED_DIAG.f90:93-228) + ed_analysis (:248-398)
//   observables_impurity : dens / docc (ED_OBSERVABLES.f90:74-168)
//   buildgf_impurity     : build_gf_normal (build_gf_normal.f90:4-180) and
//                          get_sigma_normal (get_sigma_normal.f90:1-38)
// The state list keeps device vectors (handles), never host copies of them.
#include <algorithm>
#include <chrono>
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
    p->reserved = 0;
}

extern "C" int dmed_solve(const dmed_solve_params *p, dmed_solve_out *out)
{
    Context &c = ctx();
    if (!c.initialized) return fail(DMED_ESTATE, "dmed_init not called");
    if (!c.has_model) return fail(DMED_ESTATE, "dmed_set_model not called");
    if (!p || !out) return fail(DMED_EINVAL, "null argument");
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
    for (int nu = 0; nu <= ns && !rc; ++nu)
        for (int nd = 0; nd <= ns && !rc; ++nd) {
            const int is = nu * (ns + 1) + nd;
            dmed_sector *sec = nullptr;
            if ((rc = cache.get(nu, nd, ns, &sec))) break;
            const int64_t dim = sec->dim;
            const int ne = (int)std::min<int64_t>(dim, neigen[is]);
            const int nblock = (int)std::min<int64_t>(dim, p->lanc_ncv_factor * ne + p->lanc_ncv_add);
            std::vector<double> ev(ne), res(ne);
            std::vector<dmed_vec *> vecs(ne, nullptr);
            // T=0 keeps only states within gs_threshold of the running minimum
            // (ED_DIAG.f90:202-213), so only the lowest level of a sector and
            // its degenerate copies can enter; finite T needs the ne lowest.
            int nout = 0;
            rc = eigh_sector_ex(sec, ne, std::max(nblock, 2 * ne + 10), 50 * std::max(p->lanc_niter, 100),
                                p->lanc_tolerance, p->lanc_dim_threshold, 10.0 * p->gs_threshold,
                                !finite_t, &nout, ev.data(), vecs.data(), res.data());
            if (rc == DMED_ENOCONV) rc = 0;      // ARPACK returns its best pairs too
            if (rc) break;
            if (p->verbose > 2)
                std::fprintf(stderr, "sector %d (%d,%d) dim=%lld E0=%.12f\n", is + 1, nu, nd,
                             (long long)dim, ev[0]);
            bool used = false;
            for (int i = 0; i < nout; ++i) {
                State st{ev[i], is, vecs[i]};
                bool kept = false;
                if (finite_t) {
                    kept = sl.add(st, nst_total);
                } else if (ev[i] < oldzero - 10.0 * p->gs_threshold) {
                    oldzero = ev[i];
                    sl.free_all();
                    sl.insert(st);
                    kept = true;
                } else if (std::fabs(ev[i] - oldzero) <= p->gs_threshold) {
                    oldzero = std::min(oldzero, ev[i]);
                    sl.insert(st);
                    kept = true;
                }
                if (!kept) dmed_vec_free(vecs[i]);
                used = used || kept;
            }
            (void)used;
        }
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
    std::vector<double> dens(norb, 0.0), docc(norb, 0.0);
    for (auto &s : sl.items) {
        const double peso = (finite_t ? std::exp(-beta * (s.e - egs)) : 1.0) / zeta;
        if ((rc = dmed_observables(s.vec, peso, dens.data(), docc.data()))) break;
    }
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
    std::vector<double> al(ngf), bl(ngf);
    for (int s = 0; s < nspin && !rc; ++s)
        for (int a = 0; a < norb && !rc; ++a) {
            std::fill(gm.begin(), gm.end(), 0.0);
            std::fill(gr.begin(), gr.end(), 0.0);
            const int isite = a + 1 + s * ns;
            for (auto &st : sl.items) {
                const int nu = st.sector / (ns + 1), nd = st.sector % (ns + 1);
                for (int pass = 0; pass < 2 && !rc; ++pass) {
                    const int dagger = pass == 0 ? 1 : 0;
                    const int isign = pass == 0 ? 1 : -1;
                    const int ju = nu + (s == 0 ? (dagger ? 1 : -1) : 0);
                    const int jd = nd + (s == 1 ? (dagger ? 1 : -1) : 0);
                    if (ju < 0 || jd < 0 || ju > ns || jd > ns) continue;
                    dmed_sector *J = nullptr;
                    if ((rc = cache.get(ju, jd, ns, &J))) break;
                    dmed_vec *v = nullptr;
                    if ((rc = dmed_vec_create(J, DMED_FP64, &v))) break;
                    double norm2 = 0.0;
                    rc = dmed_apply_op(st.vec, isite, dagger, J, v, &norm2);
                    if (!rc && norm2 > 0.0) {
                        // zero norm: the operator annihilates the state and the
                        // chain carries no weight (the reference divides by 0 here)
                        const int nlanc = (int)std::min<int64_t>(J->dim, ngf);
                        std::fill(al.begin(), al.end(), 0.0);
                        std::fill(bl.begin(), bl.end(), 0.0);
                        int nsteps = 0;
                        rc = chain_run(J, v->data, DMED_FP64, nlanc, p->prec_gf, 1e-13, al.data(),
                                       bl.data(), &nsteps);
                        double peso = 0.0;
                        if (finite_t) {
                            if (beta * (st.e - egs) < 200) peso = norm2 * std::exp(-beta * (st.e - egs)) / zeta;
                        } else {
                            peso = norm2 / zeta;
                        }
                        if (!rc)
                            rc = dmed_gf_add_chain(peso, st.e, al.data(), bl.data(), nlanc, isign,
                                                   wm.data(), lm, wr.data(), lr, p->eps, gm.data(),
                                                   gr.data());
                    }
                    dmed_vec_free(v);
                }
                if (rc) break;
            }
            if (rc) break;
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
    const double t2 = now();
    out->egs = egs;
    out->zeta = zeta;
    out->nstates = (int)sl.items.size();
    out->numgs = numgs;
    if (out->dens) std::copy(dens.begin(), dens.end(), out->dens);
    if (out->docc) std::copy(docc.begin(), docc.end(), out->docc);
    if (out->neigen_sector) std::copy(neigen.begin(), neigen.end(), out->neigen_sector);
    if (out->lanc_nstates_total) *out->lanc_nstates_total = nst_total;
    out->t_diag = t1 - t0;
    out->t_gf = t2 - t1;
    sl.free_all();
    return rc;
}
