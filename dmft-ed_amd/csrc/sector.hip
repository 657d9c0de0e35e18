// sector.hip -- device-built sector tables.
//
// A sector (n_up, n_dw) of the reference (ED_SETUP.f90:861-892) is the
// product of two spin-string spaces.  Its basis index i = r_up + r_dw*D_up,
// with r the rank of the string in numeric order; numeric order of n-bit
// strings is colex order, so r = sum_j C(p_j, j) over the set bits
// p_1 < p_2 < ... (combinatorial number system).  That makes the index of the
// reference's map/binary_search (ED_SETUP.f90:1282-1299) a closed form, and
// every table below is built on the GPU, one thread per string.
//
// Fermionic sign of c^+_i c_j on one spin string s (bits 0-based):
//   (-1)^{popcount(s & (2^j-1))} * (-1)^{popcount((s - 2^j) & (2^i-1))}
// which is c(j) then cdg(i) of ED_SETUP.f90:1055-1081 restricted to one spin.
// For the down spin the reference counts all up bits too; they enter twice in
// a same-spin hop and cancel, so per-spin tables are exact.
#include "dmed_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace dmed {

__constant__ uint32_t c_binom[32][32];   // C(p,k), p,k < 32
static std::once_flag g_binom_once;
static hipError_t g_binom_status = hipSuccess;

int64_t binom(int n, int k)
{
    if (k < 0 || k > n) return 0;
    int64_t r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

// The table is written once per process, by whichever thread gets here
// first; the others wait for it (workers build sectors concurrently).
int upload_binom()
{
    std::call_once(g_binom_once, [] {
        uint32_t h[32][32];
        for (int p = 0; p < 32; ++p)
            for (int k = 0; k < 32; ++k) {
                int64_t b = binom(p, k);
                h[p][k] = (uint32_t)std::min<int64_t>(b, 0xffffffffll);
            }
        hipStream_t st = ex().stream;
        g_binom_status = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_binom), h, sizeof h, 0,
                                                hipMemcpyHostToDevice, st);
        if (g_binom_status == hipSuccess) g_binom_status = hipStreamSynchronize(st);
    });
    if (g_binom_status != hipSuccess)
        return fail(DMED_EHIP, std::string("binomial table upload: ") +
                                   hipGetErrorString(g_binom_status));
    return 0;
}

__device__ __forceinline__ int32_t rank_of(uint32_t s)
{
    int32_t r = 0;
    int j = 1;
    while (s) {
        int p = __ffs(s) - 1;
        r += (int32_t)c_binom[p][j];
        ++j;
        s &= s - 1;
    }
    return r;
}

__global__ void k_unrank(uint32_t *__restrict__ str, int64_t dim, int ns, int n)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= dim) return;
    uint32_t s = 0;
    int64_t rr = r;
    int k = n;
    for (int p = ns - 1; p >= 0 && k > 0; --p) {
        int64_t c = c_binom[p][k];
        if (rr >= c) {
            s |= 1u << p;
            rr -= c;
            --k;
        }
    }
    str[r] = s;
}

// count pass (count_only): atomicMax of the number of applicable hops;
// fill pass: diagonal, impurity bits and the padded hop table.
__global__ void k_spin_tables(const uint32_t *__restrict__ str, int64_t dim, int norb,
                              const double *__restrict__ onsite,
                              const int32_t *__restrict__ ti, const int32_t *__restrict__ tj,
                              const double *__restrict__ tt, int nterms, int count_only,
                              int nhop, double *__restrict__ diag, uint8_t *__restrict__ imp,
                              int32_t *__restrict__ hidx, double *__restrict__ hval,
                              float *__restrict__ hval32, int *__restrict__ maxcount)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= dim) return;
    const uint32_t s = str[r];
    int cnt = 0;
    if (count_only) {
        for (int t = 0; t < nterms; ++t) {
            const int i = ti[t], j = tj[t];
            if (((s >> j) & 1u) && !((s >> i) & 1u)) ++cnt;
        }
        atomicMax(maxcount, cnt);
        return;
    }
    double d = 0.0;
    for (uint32_t q = s; q; q &= q - 1) d += onsite[__ffs(q) - 1];
    diag[r] = d;
    imp[r] = (uint8_t)(s & ((1u << norb) - 1u));
    for (int t = 0; t < nterms; ++t) {
        const int i = ti[t], j = tj[t];
        if (((s >> j) & 1u) && !((s >> i) & 1u)) {
            const uint32_t s1 = s & ~(1u << j);
            const int par = __popc(s & ((1u << j) - 1u)) + __popc(s1 & ((1u << i) - 1u));
            const uint32_t s2 = s1 | (1u << i);
            const double c = (par & 1) ? -tt[t] : tt[t];
            hidx[(int64_t)cnt * dim + r] = rank_of(s2);
            hval[(int64_t)cnt * dim + r] = c;
            hval32[(int64_t)cnt * dim + r] = (float)c;
            ++cnt;
        }
    }
    for (; cnt < nhop; ++cnt) {          // padding: self, zero coefficient
        hidx[(int64_t)cnt * dim + r] = (int32_t)r;
        hval[(int64_t)cnt * dim + r] = 0.0;
        hval32[(int64_t)cnt * dim + r] = 0.0f;
    }
}

static void release_spin(SpinTables &t)
{
    hipFree(t.str);
    hipFree(t.diag);
    hipFree(t.imp);
    hipFree(t.hidx);
    hipFree(t.hval);
    hipFree(t.hval32);
    hipFree(t.hpk);
    hipFree(t.coef);
    hipFree(t.ipair);
    t = SpinTables();
}

void sector_release(dmed_sector *s)
{
    release_spin(s->up);
    release_spin(s->dw);
    hipFree(s->W);
    s->W = nullptr;
}

// one-body terms and onsite energies of spin `sp` (0 up, 1 down) for the
// normal or hybrid bath (bath_bit): HxVimp.f90:2-49, HxVbath.f90:1-14,
// HxVimp_bath.f90:1-38
static void spin_terms(const Model &m, int sp, std::vector<HopTerm> &terms,
                       std::vector<double> &onsite)
{
    const int s = (m.nspin == 2) ? sp : 0;     // Nspin=1 reuses spin 1 for the down spin
    onsite.assign(m.ns, 0.0);
    for (int a = 0; a < m.norb; ++a) onsite[a] = -m.xmu + m.hloc[s][a][a];
    // the hybrid bath has one set of levels (HxVbath.f90:4-10 loops over size(e,2) = 1)
    for (int a = 0; a < (m.hybrid ? 1 : m.norb); ++a)
        for (int k = 0; k < m.nbath; ++k) onsite[m.bath_bit(a, k)] = m.E(s, a, k);
    terms.clear();
    for (int a = 0; a < m.norb; ++a)
        for (int b = 0; b < m.norb; ++b)
            if (a != b && m.hloc[s][a][b] != 0.0) terms.push_back({a, b, m.hloc[s][a][b]});
    for (int a = 0; a < m.norb; ++a)
        for (int k = 0; k < m.nbath; ++k) {
            const double v = m.V(s, a, k);
            if (v == 0.0) continue;
            const int ms = m.bath_bit(a, k);
            terms.push_back({ms, a, v});   // imp -> bath
            terms.push_back({a, ms, v});   // bath -> imp
        }
}

// Packed hop words from the host tables: target rank (24 bits), index of
// |coefficient| in t.coef (7 bits), sign; left unpacked (t.packed = false)
// past 2^24 ranks or 128 distinct |coefficients|.
static int pack_hops(SpinTables &t)
{
    const int64_t dim = t.dim;
    std::vector<double> coef;
    std::vector<uint32_t> pk((size_t)t.nhop * dim);
    bool ok = dim < (1 << 24);
    for (size_t q = 0; q < pk.size() && ok; ++q) {
        const double a = std::fabs(t.h_hval[q]);
        int id = -1;
        for (size_t u = 0; u < coef.size(); ++u)
            if (coef[u] == a) { id = (int)u; break; }
        if (id < 0) {
            if (coef.size() >= 128) { ok = false; break; }
            coef.push_back(a);
            id = (int)coef.size() - 1;
        }
        pk[q] = ((uint32_t)t.h_hidx[q] << 8) | ((uint32_t)id << 1) | (t.h_hval[q] < 0 ? 1u : 0u);
    }
    if (ok) {
        if (coef.empty()) coef.push_back(0.0);
        t.ncoef = (int)coef.size();
        DMED_HIP(hipMalloc(&t.hpk, std::max<size_t>(pk.size(), 1) * 4));
        DMED_HIP(hipMalloc(&t.coef, coef.size() * 8));
        if (!pk.empty())
            if (int rc = copy_h2d(t.hpk, pk.data(), pk.size() * 4)) return rc;
        if (int rc = copy_h2d(t.coef, coef.data(), coef.size() * 8)) return rc;
        t.packed = true;
    }
    return 0;
}

static int build_spin(const Model &m, int sp, int n, SpinTables &t, hipStream_t st)
{
    t.n = n;
    t.dim = binom(m.ns, n);
    const int64_t dim = t.dim;
    std::vector<HopTerm> terms;
    std::vector<double> onsite;
    spin_terms(m, sp, terms, onsite);
    if ((int)terms.size() > kMaxTerms) return fail(DMED_EINVAL, "too many one-body terms");
    const int nterms = (int)terms.size();
    // small upload block: onsite[ns] | t[nterms] | ti | tj | maxcount
    std::vector<double> hbuf(m.ns + nterms + 1, 0.0);
    std::vector<int32_t> ibuf(2 * nterms + 1, 0);
    for (int p = 0; p < m.ns; ++p) hbuf[p] = onsite[p];
    for (int q = 0; q < nterms; ++q) {
        hbuf[m.ns + q] = terms[q].t;
        ibuf[q] = terms[q].i;
        ibuf[nterms + q] = terms[q].j;
    }
    double *d_h = nullptr;
    int32_t *d_i = nullptr;
    DMED_HIP(hipMalloc(&d_h, hbuf.size() * sizeof(double)));
    DMED_HIP(hipMalloc(&d_i, ibuf.size() * sizeof(int32_t)));
    DMED_HIP(hipMemcpyAsync(d_h, hbuf.data(), hbuf.size() * sizeof(double),
                            hipMemcpyHostToDevice, st));
    DMED_HIP(hipMemcpyAsync(d_i, ibuf.data(), ibuf.size() * sizeof(int32_t),
                            hipMemcpyHostToDevice, st));
    int *d_max = d_i + 2 * nterms;
    DMED_HIP(hipMalloc(&t.str, std::max<int64_t>(dim, 1) * sizeof(uint32_t)));
    DMED_HIP(hipMalloc(&t.diag, std::max<int64_t>(dim, 1) * sizeof(double)));
    DMED_HIP(hipMalloc(&t.imp, std::max<int64_t>(dim, 1)));
    const int bs = 256;
    const unsigned nb = (unsigned)((dim + bs - 1) / bs);
    k_unrank<<<nb, bs, 0, st>>>(t.str, dim, m.ns, n);
    DMED_HIP(hipGetLastError());
    k_spin_tables<<<nb, bs, 0, st>>>(t.str, dim, m.norb, d_h, d_i, d_i + nterms, d_h + m.ns,
                                     nterms, 1, 0, nullptr, nullptr, nullptr, nullptr, nullptr,
                                     d_max);
    DMED_HIP(hipGetLastError());
    int maxcount = 0;
    DMED_HIP(hipMemcpyAsync(&maxcount, d_max, sizeof(int), hipMemcpyDeviceToHost, st));
    DMED_HIP(hipStreamSynchronize(st));
    t.nhop = maxcount;
    const int64_t hn = std::max<int64_t>((int64_t)t.nhop * dim, 1);
    DMED_HIP(hipMalloc(&t.hidx, hn * sizeof(int32_t)));
    DMED_HIP(hipMalloc(&t.hval, hn * sizeof(double)));
    DMED_HIP(hipMalloc(&t.hval32, hn * sizeof(float)));
    k_spin_tables<<<nb, bs, 0, st>>>(t.str, dim, m.norb, d_h, d_i, d_i + nterms, d_h + m.ns,
                                     nterms, 0, t.nhop, t.diag, t.imp, t.hidx, t.hval,
                                     t.hval32, d_max);
    DMED_HIP(hipGetLastError());
    // host copies (O(D_up + nhop*D_up), small next to the vectors)
    t.h_str.resize(dim);
    t.h_diag.resize(dim);
    t.h_imp.resize(dim);
    t.h_hidx.resize((size_t)t.nhop * dim);
    t.h_hval.resize((size_t)t.nhop * dim);
    if (dim) {
        DMED_HIP(hipMemcpyAsync(t.h_str.data(), t.str, dim * 4, hipMemcpyDeviceToHost, st));
        DMED_HIP(hipMemcpyAsync(t.h_diag.data(), t.diag, dim * 8, hipMemcpyDeviceToHost, st));
        DMED_HIP(hipMemcpyAsync(t.h_imp.data(), t.imp, dim, hipMemcpyDeviceToHost, st));
        if (t.nhop) {
            DMED_HIP(hipMemcpyAsync(t.h_hidx.data(), t.hidx, (size_t)t.nhop * dim * 4,
                                    hipMemcpyDeviceToHost, st));
            DMED_HIP(hipMemcpyAsync(t.h_hval.data(), t.hval, (size_t)t.nhop * dim * 8,
                                    hipMemcpyDeviceToHost, st));
        }
    }
    DMED_HIP(hipStreamSynchronize(st));
    hipFree(d_h);
    hipFree(d_i);
    return pack_hops(t);
}

// Interaction energy of the impurity occupations (HxVint.f90:1-43):
// iu, id = impurity bits of the up / down strings.
static double interaction(const Model &m, unsigned iu, unsigned id)
{
    double nup[3] = {}, ndw[3] = {};
    for (int a = 0; a < m.norb; ++a) {
        nup[a] = (iu >> a) & 1u;
        ndw[a] = (id >> a) & 1u;
    }
    double h = 0.0;
    for (int a = 0; a < m.norb; ++a) h += m.uloc[a] * nup[a] * ndw[a];
    if (m.norb > 1) {
        for (int a = 0; a < m.norb; ++a)
            for (int b = a + 1; b < m.norb; ++b)
                h += m.ust * (nup[a] * ndw[b] + nup[b] * ndw[a]);
        for (int a = 0; a < m.norb; ++a)
            for (int b = a + 1; b < m.norb; ++b)
                h += (m.ust - m.jh) * (nup[a] * nup[b] + ndw[a] * ndw[b]);
    }
    if (m.hfmode) {
        for (int a = 0; a < m.norb; ++a)
            h = h - 0.5 * m.uloc[a] * (nup[a] + ndw[a]) + 0.25 * m.uloc[a];
        if (m.norb > 1)
            for (int a = 0; a < m.norb; ++a)
                for (int b = a + 1; b < m.norb; ++b) {
                    const double nn = nup[a] + ndw[a] + nup[b] + ndw[b];
                    h = h - 0.5 * m.ust * nn + 0.25 * m.ust;
                    h = h - 0.5 * (m.ust - m.jh) * nn + 0.25 * (m.ust - m.jh);
                }
    }
    return h;
}

// Impurity pair-hop tables of one spin (host strings -> device): for every
// ordered pair (a, b) of impurity orbitals, a != b, the string with b
// emptied and a filled, and the sign of c^+_a c_b on it (the fermionic signs
// of c_b on the string, then of c^+_a on the result: ED_SETUP.f90:1055-1081).
// The spin-exchange and pair-hopping terms of HxVint.f90:48-97 are products
// of one such hop per spin: their two down-spin signs carry (-1)^n_up each,
// which cancel.
static int build_pairs(const Model &m, SpinTables &t)
{
    const int no = m.norb;
    const int64_t dim = t.dim;
    std::vector<int32_t> tab((size_t)no * no * dim, -1);
    for (int a = 0; a < no; ++a)
        for (int b = 0; b < no; ++b) {
            if (a == b) continue;
            int32_t *row = tab.data() + (size_t)(a * no + b) * dim;
            for (int64_t r = 0; r < dim; ++r) {
                const uint32_t u = t.h_str[r];
                if (!((u >> b) & 1u) || ((u >> a) & 1u)) continue;
                const uint32_t u1 = u & ~(1u << b);
                const int n1 = __builtin_popcount(u & ((1u << b) - 1u));
                const int n2 = __builtin_popcount(u1 & ((1u << a) - 1u));
                const uint32_t u2 = u1 | (1u << a);
                const auto it = std::lower_bound(t.h_str.begin(), t.h_str.end(), u2);
                if (it == t.h_str.end() || *it != u2) return fail(DMED_EINVAL, "pair hop leaves the sector");
                const int64_t rk = it - t.h_str.begin();
                row[r] = (int32_t)((rk << 1) | ((n1 + n2) & 1));
            }
        }
    DMED_HIP(hipMalloc(&t.ipair, tab.size() * 4));
    if (int rc = copy_h2d(t.ipair, tab.data(), tab.size() * 4)) return rc;
    t.h_ipair = std::move(tab);
    return 0;
}

int sector_build(dmed_sector *s, int nup, int ndw)
{
    Context &c = ctx();
    const Model &m = cur_model();
    if (nup < 0 || ndw < 0 || nup > m.ns || ndw > m.ns)
        return fail(DMED_EINVAL, "sector occupation out of range");
    if (int rc = upload_binom()) return rc;
    s->nup = nup;
    s->ndw = ndw;
    s->ns = m.ns;
    s->norb = m.norb;
    s->model_generation = m.generation;
    if (int rc = build_spin(m, 0, nup, s->up, ex().stream)) return rc;
    if (int rc = build_spin(m, 1, ndw, s->dw, ex().stream)) return rc;
    s->dim = s->up.dim * s->dw.dim;
    s->nimp = 1 << m.norb;
    s->jterms = (m.jhflag && m.norb > 1 && (m.jx != 0.0 || m.jp != 0.0)) ? 1 : 0;
    s->jx = m.jx;
    s->jp = m.jp;
    if (s->jterms) {
        if (int rc = build_pairs(m, s->up)) return rc;
        if (int rc = build_pairs(m, s->dw)) return rc;
    }
    s->h_W.resize(s->nimp * s->nimp);
    for (int iu = 0; iu < s->nimp; ++iu)
        for (int id = 0; id < s->nimp; ++id)
            s->h_W[iu * s->nimp + id] = interaction(m, iu, id);
    DMED_HIP(hipMalloc(&s->W, s->h_W.size() * sizeof(double)));
    if (int rc = copy_h2d(s->W, s->h_W.data(), s->h_W.size() * sizeof(double))) return rc;
    return 0;
}

// Down-spin tables of the sector repeated twice (see sector_build_stacked):
// column c + D_dw of the second copy is column c of the first, its hops
// shifted by D_dw; rebuilt from the host copies and uploaded again.
static int stack_spin(SpinTables &t)
{
    const int64_t d = t.dim, d2 = 2 * d;
    if (d2 >= ((int64_t)1 << 30)) return fail(DMED_EINVAL, "stacked sector too large");
    auto twice = [](auto &v) { v.insert(v.end(), v.begin(), v.end()); };
    twice(t.h_str);
    twice(t.h_diag);
    twice(t.h_imp);
    std::vector<int32_t> hidx((size_t)t.nhop * d2);
    std::vector<double> hval((size_t)t.nhop * d2);
    for (int h = 0; h < t.nhop; ++h)
        for (int64_t r = 0; r < d; ++r) {
            const size_t q = (size_t)h * d + r;
            hidx[(size_t)h * d2 + r] = t.h_hidx[q];
            hidx[(size_t)h * d2 + d + r] = t.h_hidx[q] + (int32_t)d;
            hval[(size_t)h * d2 + r] = hval[(size_t)h * d2 + d + r] = t.h_hval[q];
        }
    t.h_hidx = std::move(hidx);
    t.h_hval = std::move(hval);
    if (!t.h_ipair.empty()) {
        const size_t npair = t.h_ipair.size() / (size_t)std::max<int64_t>(d, 1);
        std::vector<int32_t> ip(npair * d2);
        for (size_t pr = 0; pr < npair; ++pr)
            for (int64_t r = 0; r < d; ++r) {
                const int32_t x = t.h_ipair[pr * d + r];
                ip[pr * d2 + r] = x;
                ip[pr * d2 + d + r] = x < 0 ? x : x + (int32_t)(2 * d);
            }
        t.h_ipair = std::move(ip);
    }
    hipFree(t.str);
    hipFree(t.diag);
    hipFree(t.imp);
    hipFree(t.hidx);
    hipFree(t.hval);
    hipFree(t.hval32);
    hipFree(t.hpk);
    hipFree(t.coef);
    hipFree(t.ipair);
    t.str = nullptr; t.diag = nullptr; t.imp = nullptr; t.hidx = nullptr; t.hval = nullptr;
    t.hval32 = nullptr; t.hpk = nullptr; t.coef = nullptr; t.ipair = nullptr;
    t.packed = false;
    t.ncoef = 0;
    t.dim = d2;
    const size_t hn = std::max<size_t>((size_t)t.nhop * d2, 1);
    std::vector<float> h32(t.h_hval.size());
    for (size_t q = 0; q < h32.size(); ++q) h32[q] = (float)t.h_hval[q];
    DMED_HIP(hipMalloc(&t.str, std::max<int64_t>(d2, 1) * sizeof(uint32_t)));
    DMED_HIP(hipMalloc(&t.diag, std::max<int64_t>(d2, 1) * sizeof(double)));
    DMED_HIP(hipMalloc(&t.imp, std::max<int64_t>(d2, 1)));
    DMED_HIP(hipMalloc(&t.hidx, hn * sizeof(int32_t)));
    DMED_HIP(hipMalloc(&t.hval, hn * sizeof(double)));
    DMED_HIP(hipMalloc(&t.hval32, hn * sizeof(float)));
    if (d2) {
        if (int rc = copy_h2d(t.str, t.h_str.data(), d2 * 4)) return rc;
        if (int rc = copy_h2d(t.diag, t.h_diag.data(), d2 * 8)) return rc;
        if (int rc = copy_h2d(t.imp, t.h_imp.data(), d2)) return rc;
    }
    if (!t.h_hidx.empty()) {
        if (int rc = copy_h2d(t.hidx, t.h_hidx.data(), t.h_hidx.size() * 4)) return rc;
        if (int rc = copy_h2d(t.hval, t.h_hval.data(), t.h_hval.size() * 8)) return rc;
        if (int rc = copy_h2d(t.hval32, h32.data(), h32.size() * 4)) return rc;
    }
    if (!t.h_ipair.empty()) {
        DMED_HIP(hipMalloc(&t.ipair, t.h_ipair.size() * 4));
        if (int rc = copy_h2d(t.ipair, t.h_ipair.data(), t.h_ipair.size() * 4)) return rc;
    }
    return pack_hops(t);
}

int sector_build_stacked(dmed_sector *s, int nup, int ndw)
{
    if (int rc = sector_build(s, nup, ndw)) return rc;
    DMED_HIP(hipStreamSynchronize(ex().stream));
    if (int rc = stack_spin(s->dw)) return rc;
    s->dim = s->up.dim * s->dw.dim;
    return 0;
}

}  // namespace dmed
