// api.hip -- the C ABI (include/dmed.h): context, model, sectors, vectors,
// H*v, eigensolver, Green's-function start vectors, chains and the whole
// ed_solve sequence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "dmed_internal.h"

#include <atomic>

namespace dmed {

Context &ctx()
{
    static Context c;
    return c;
}

static thread_local Exec *t_exec = nullptr;
static thread_local std::string t_error;
static thread_local Model *t_model = nullptr;
static std::atomic<int> g_generation{0};

Exec &ex() { return t_exec ? *t_exec : ctx().exec; }
void bind_exec(Exec *e) { t_exec = e; }
Model &cur_model() { return t_model ? *t_model : ctx().model; }
bool has_cur_model() { return t_model ? true : ctx().has_model; }
void bind_model(Model *m) { t_model = m; }

int fail(int code, const std::string &msg)
{
    if (t_exec) t_error = msg;          // worker threads keep their own message
    else ctx().last_error = msg;
    return code;
}

int exec_create(Exec &e)
{
    DMED_HIP(hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking));
    DMED_HIP(hipMalloc(&e.dscal, kScalCap * sizeof(double)));
    DMED_HIP(hipMemsetAsync(e.dscal, 0, kScalCap * sizeof(double), e.stream));
    DMED_HIP(hipMalloc(&e.partials, (1 << 16) * sizeof(double)));
    e.partials_cap = 1 << 16;
    return 0;
}

void exec_destroy(Exec &e)
{
    if (e.stream) hipStreamSynchronize(e.stream);
    hipFree(e.partials);
    hipFree(e.dscal);
    for (int i = 0; i < kExecSlots; ++i) hipFree(e.buf[i]);
    if (e.stream) hipStreamDestroy(e.stream);
    e = Exec();
}

int exec_buf(Exec &e, int slot, size_t bytes, void **p)
{
    if (slot < 0 || slot >= kExecSlots) return fail(DMED_EINVAL, "bad buffer slot");
    bytes = std::max<size_t>(bytes, 256);
    if (bytes > e.buf_cap[slot]) {
        if (e.buf[slot]) {
            hipStreamSynchronize(e.stream);
            hipFree(e.buf[slot]);
        }
        e.buf[slot] = nullptr;
        e.buf_cap[slot] = 0;
        DMED_HIP(hipMalloc(&e.buf[slot], bytes));
        e.buf_cap[slot] = bytes;
    }
    *p = e.buf[slot];
    return 0;
}

// Host-to-device copy on the calling thread's stream, finished on return.
// Never a null-stream copy: that would break a graph capture running on
// another worker's stream.
int copy_h2d(void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return 0;
    hipStream_t st = ex().stream;
    DMED_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    DMED_HIP(hipStreamSynchronize(st));
    return 0;
}

int ensure_partials(size_t n)
{
    Exec &e = ex();
    if (n <= e.partials_cap) return 0;
    if (e.partials) {
        hipStreamSynchronize(e.stream);
        hipFree(e.partials);
    }
    e.partials = nullptr;
    e.partials_cap = 0;
    DMED_HIP(hipMalloc(&e.partials, n * sizeof(double)));
    e.partials_cap = n;
    return 0;
}

}  // namespace dmed

using namespace dmed;

#define REQUIRE_INIT()                                                     \
    do {                                                                   \
        if (!ctx().initialized) return fail(DMED_ESTATE, "dmed_init not called"); \
    } while (0)
#define REQUIRE_MODEL()                                                    \
    do {                                                                   \
        REQUIRE_INIT();                                                    \
        if (!has_cur_model()) return fail(DMED_ESTATE, "dmed_set_model not called"); \
    } while (0)

extern "C" {

int dmed_init(int device)
{
    Context &c = ctx();
    if (c.initialized) {
        if (device == c.device) return 0;
        return fail(DMED_ESTATE, "already initialised on another device");
    }
    int n = 0;
    DMED_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(DMED_EINVAL, "no such device");
    DMED_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    DMED_HIP(hipGetDeviceProperties(&prop, device));
    c.device = device;
    c.num_cu = prop.multiProcessorCount;
    c.lds_max = std::max<size_t>(prop.sharedMemPerBlock, prop.sharedMemPerBlockOptin);
    if (c.lds_max == 0) c.lds_max = 65536;
    if (c.lds_max > 160 * 1024) c.lds_max = 160 * 1024;
    if (int rc = exec_create(c.exec)) return rc;
    if (int rc = noalloc_setup(device)) return rc;
    DMED_HIP(hipStreamSynchronize(c.exec.stream));
    c.initialized = true;
    return 0;
}

int dmed_finalize(void)
{
    Context &c = ctx();
    if (!c.initialized) return 0;
    exec_pool_clear();
    exec_destroy(c.exec);
    c.initialized = false;
    c.has_model = false;
    return 0;
}

const char *dmed_last_error(void) { return t_exec ? t_error.c_str() : ctx().last_error.c_str(); }

}  // extern "C"

namespace dmed {

// Validate a model description and copy it into md under a new generation
// (sectors remember the generation they were built for).
int model_fill(const dmed_model *m, Model &md)
{
    if (!m || !m->bath || !m->hloc) return fail(DMED_EINVAL, "null model");
    if (m->norb < 1 || m->norb > 3) return fail(DMED_EINVAL, "Norb must be 1..3 (ED_SETUP.f90:276)");
    if (m->nspin < 1 || m->nspin > 2) return fail(DMED_EINVAL, "Nspin must be 1..2 (ED_SETUP.f90:275)");
    if (m->nbath < 0) return fail(DMED_EINVAL, "Nbath < 0");
    if (m->bath_type != DMED_BATH_NORMAL && m->bath_type != DMED_BATH_HYBRID)
        return fail(DMED_EINVAL, "bath_type must be DMED_BATH_NORMAL or DMED_BATH_HYBRID");
    const int hyb = m->bath_type == DMED_BATH_HYBRID;
    const int ns = hyb ? m->nbath + m->norb : (m->nbath + 1) * m->norb;
    if (ns > kMaxNs) return fail(DMED_EINVAL, "Ns too large for 32-bit spin strings");
    md.norb = m->norb;
    md.nbath = m->nbath;
    md.nspin = m->nspin;
    md.ns = ns;
    md.hybrid = hyb;
    const int nv = m->nspin * m->norb * m->nbath;
    const int ne = hyb ? m->nspin * m->nbath : nv;
    md.e.assign(m->bath, m->bath + ne);
    md.v.assign(m->bath + ne, m->bath + ne + nv);
    std::memset(md.hloc, 0, sizeof md.hloc);
    for (int s = 0; s < m->nspin; ++s)
        for (int a = 0; a < m->norb; ++a)
            for (int b = 0; b < m->norb; ++b)
                md.hloc[s][a][b] = m->hloc[s + m->nspin * (s + m->nspin * (a + m->norb * b))];
    if (m->nspin == 1)
        for (int a = 0; a < m->norb; ++a)
            for (int b = 0; b < m->norb; ++b) md.hloc[1][a][b] = md.hloc[0][a][b];
    for (int a = 0; a < 3; ++a) md.uloc[a] = m->uloc[a];
    md.ust = m->ust;
    md.jh = m->jh;
    md.jx = m->jx;
    md.jp = m->jp;
    md.xmu = m->xmu;
    md.hfmode = m->hfmode;
    md.jhflag = m->jhflag;
    md.generation = ++g_generation;
    return 0;
}

}  // namespace dmed

extern "C" {

int dmed_set_model(const dmed_model *m)
{
    REQUIRE_INIT();
    if (int rc = model_fill(m, ctx().model)) return rc;
    ctx().has_model = true;
    return 0;
}

int dmed_get_ns(int32_t *ns)
{
    REQUIRE_MODEL();
    *ns = cur_model().ns;
    return 0;
}

int64_t dmed_binomial(int32_t n1, int32_t n2)
{
    // real(8) running product rounded to the nearest integer, ED_SETUP.f90:1258-1275
    if (n2 < 0) return 0;
    if (n2 == 0) return 1;
    double xh = 1.0;
    for (int i = 1; i <= n2; ++i) xh = xh * (double)(n1 + 1 - i) / (double)i;
    return (int64_t)(xh + 0.5);
}

int dmed_sector_create(int32_t nup, int32_t ndw, dmed_sector **out)
{
    REQUIRE_MODEL();
    if (!out) return fail(DMED_EINVAL, "null output");
    auto s = std::make_unique<dmed_sector>();
    if (int rc = sector_build(s.get(), nup, ndw)) {
        sector_release(s.get());
        return rc;
    }
    *out = s.release();
    return 0;
}

int dmed_sector_free(dmed_sector *s)
{
    if (!s) return 0;
    if (ctx().initialized) hipStreamSynchronize(ex().stream);
    sector_release(s);
    delete s;
    return 0;
}

int dmed_sector_dims(const dmed_sector *s, int64_t *dim, int64_t *dim_up, int64_t *dim_dw)
{
    if (!s) return fail(DMED_EINVAL, "null sector");
    if (dim) *dim = s->dim;
    if (dim_up) *dim_up = s->up.dim;
    if (dim_dw) *dim_dw = s->dw.dim;
    return 0;
}

int dmed_sector_map(const dmed_sector *s, int64_t *codes)
{
    if (!s || !codes) return fail(DMED_EINVAL, "null argument");
    const int64_t du = s->up.dim, dd = s->dw.dim;
    for (int64_t id = 0; id < dd; ++id) {
        const int64_t hi = (int64_t)s->dw.h_str[id] << s->ns;
        int64_t *o = codes + id * du;
        for (int64_t iu = 0; iu < du; ++iu) o[iu] = hi | (int64_t)s->up.h_str[iu];
    }
    return 0;
}

int dmed_vec_create(const dmed_sector *s, int32_t prec, dmed_vec **out)
{
    REQUIRE_INIT();
    if (!s || !out) return fail(DMED_EINVAL, "null argument");
    if (prec != DMED_FP64 && prec != DMED_FP32) return fail(DMED_EINVAL, "bad precision");
    auto v = std::make_unique<dmed_vec>();
    v->sec = s;
    v->prec = prec;
    v->dim = s->dim;
    const size_t b = std::max<int64_t>(s->dim, 1) * (prec == DMED_FP32 ? 4 : 8);
    DMED_HIP(hipMalloc(&v->data, b));
    DMED_HIP(hipMemsetAsync(v->data, 0, b, ex().stream));
    *out = v.release();
    return 0;
}

int dmed_vec_free(dmed_vec *v)
{
    if (!v) return 0;
    if (ctx().initialized) hipStreamSynchronize(ex().stream);
    hipFree(v->data);
    delete v;
    return 0;
}

void *dmed_vec_data(const dmed_vec *v) { return v ? v->data : nullptr; }

int dmed_vec_copy(dmed_vec *dst, const dmed_vec *src)
{
    REQUIRE_INIT();
    if (!dst || !src) return fail(DMED_EINVAL, "null argument");
    if (dst->sec != src->sec || dst->dim != src->dim)
        return fail(DMED_EINVAL, "vectors belong to different sectors");
    if (dst == src) return 0;
    hipStream_t st = ex().stream;
    if (int rc = launch_scale_copy(dst->data, dst->prec, src->data, src->prec, src->dim, 1.0, st))
        return rc;
    DMED_HIP(hipStreamSynchronize(st));
    return 0;
}

int dmed_vec_upload(dmed_vec *v, const double *host)
{
    REQUIRE_INIT();
    if (!v || !host) return fail(DMED_EINVAL, "null argument");
    hipStream_t st = ex().stream;
    if (v->prec == DMED_FP64) {
        DMED_HIP(hipMemcpyAsync(v->data, host, v->dim * 8, hipMemcpyHostToDevice, st));
    } else {
        double *tmp = nullptr;
        DMED_HIP(hipMalloc(&tmp, std::max<int64_t>(v->dim, 1) * 8));
        DMED_HIP(hipMemcpyAsync(tmp, host, v->dim * 8, hipMemcpyHostToDevice, st));
        int rc = launch_scale_copy(v->data, DMED_FP32, tmp, DMED_FP64, v->dim, 1.0, st);
        hipStreamSynchronize(st);
        hipFree(tmp);
        if (rc) return rc;
    }
    DMED_HIP(hipStreamSynchronize(st));
    return 0;
}

int dmed_vec_download(const dmed_vec *v, double *host)
{
    REQUIRE_INIT();
    if (!v || !host) return fail(DMED_EINVAL, "null argument");
    hipStream_t st = ex().stream;
    if (v->prec == DMED_FP64) {
        DMED_HIP(hipMemcpyAsync(host, v->data, v->dim * 8, hipMemcpyDeviceToHost, st));
    } else {
        double *tmp = nullptr;
        DMED_HIP(hipMalloc(&tmp, std::max<int64_t>(v->dim, 1) * 8));
        int rc = launch_scale_copy(tmp, DMED_FP64, v->data, DMED_FP32, v->dim, 1.0, st);
        if (!rc) DMED_HIP(hipMemcpyAsync(host, tmp, v->dim * 8, hipMemcpyDeviceToHost, st));
        hipStreamSynchronize(st);
        hipFree(tmp);
        if (rc) return rc;
    }
    DMED_HIP(hipStreamSynchronize(st));
    return 0;
}

static int check_sector(const dmed_sector *s)
{
    if (!s) return fail(DMED_EINVAL, "null sector");
    if (s->model_generation != cur_model().generation)
        return fail(DMED_ESTATE, "sector was built for an older model; recreate it");
    return 0;
}

int dmed_hxv_async(const dmed_sector *s, const void *x, void *y, int32_t prec)
{
    REQUIRE_MODEL();
    if (int rc = check_sector(s)) return rc;
    if (!x || !y || x == y) return fail(DMED_EINVAL, "x and y must be distinct device buffers");
    if (prec != DMED_FP64 && prec != DMED_FP32) return fail(DMED_EINVAL, "bad precision");
    return launch_hxv(s, x, y, prec, kPlain, nullptr, nullptr, nullptr, ex().stream);
}

int dmed_hxv(const dmed_sector *s, const void *x, void *y, int32_t prec)
{
    if (int rc = dmed_hxv_async(s, x, y, prec)) return rc;
    DMED_HIP(hipStreamSynchronize(ex().stream));
    return 0;
}

int dmed_hxv_timed(const dmed_sector *s, const void *x, void *y, int32_t prec, int32_t nrep,
                   double *ms_avg)
{
    REQUIRE_MODEL();
    if (int rc = check_sector(s)) return rc;
    if (!x || !y || x == y || nrep < 1 || !ms_avg) return fail(DMED_EINVAL, "bad argument");
    hipStream_t st = ex().stream;
    hipEvent_t e0, e1;
    DMED_HIP(hipEventCreate(&e0));
    DMED_HIP(hipEventCreate(&e1));
    DMED_HIP(hipEventRecord(e0, st));
    int rc = 0;
    for (int r = 0; r < nrep && !rc; ++r)
        rc = launch_hxv(s, x, y, prec, kPlain, nullptr, nullptr, nullptr, st);
    DMED_HIP(hipEventRecord(e1, st));
    DMED_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    DMED_HIP(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *ms_avg = (double)ms / nrep;
    return rc;
}

int dmed_hxv_shape_of(const dmed_sector *s, int32_t prec, dmed_hxv_shape *out)
{
    REQUIRE_MODEL();
    if (int rc = check_sector(s)) return rc;
    if (!out) return fail(DMED_EINVAL, "null argument");
    if (prec != DMED_FP64 && prec != DMED_FP32) return fail(DMED_EINVAL, "bad precision");
    return hxv_shape(s, prec, out);
}

int dmed_sync(void)
{
    REQUIRE_INIT();
    DMED_HIP(hipStreamSynchronize(ex().stream));
    return 0;
}

int dmed_hxv_host(const dmed_sector *s, const double *x, double *y, int32_t prec)
{
    REQUIRE_MODEL();
    if (int rc = check_sector(s)) return rc;
    dmed_vec *vx = nullptr, *vy = nullptr;
    int rc = dmed_vec_create(s, prec, &vx);
    if (!rc) rc = dmed_vec_create(s, prec, &vy);
    if (!rc) rc = dmed_vec_upload(vx, x);
    if (!rc) rc = dmed_hxv(s, vx->data, vy->data, prec);
    if (!rc) rc = dmed_vec_download(vy, y);
    dmed_vec_free(vx);
    dmed_vec_free(vy);
    return rc;
}

int dmed_eigh(const dmed_sector *s, int32_t neig, int32_t ncv, int32_t maxit, double tol,
              int32_t dense_threshold, double *evals, dmed_vec **evecs, double *resid)
{
    REQUIRE_MODEL();
    if (int rc = check_sector(s)) return rc;
    if (!evals) return fail(DMED_EINVAL, "null evals");
    if (maxit <= 0) maxit = 20000;
    if (tol <= 0) tol = 1e-12;
    return eigh_sector(s, neig, ncv, maxit, tol, dense_threshold, evals, evecs, resid);
}

int dmed_ground_state(const dmed_sector *s, int32_t prec, int32_t maxit, double tol,
                      double *eval, dmed_vec **evec, double *resid)
{
    REQUIRE_MODEL();
    if (int rc = check_sector(s)) return rc;
    if (!eval) return fail(DMED_EINVAL, "null eval");
    if (maxit <= 0) maxit = 20000;
    if (tol <= 0) tol = 1e-12;
    return ground_state(s, prec, maxit, tol, eval, evec, resid);
}

}  // extern "C"

// ------------------------------------------------- apply c^+ / c
namespace dmed {

static int32_t host_rank(uint32_t s)
{
    int32_t r = 0;
    int j = 1;
    while (s) {
        const int p = __builtin_ctz(s);
        r += (int32_t)binom(p, j);
        ++j;
        s &= s - 1;
    }
    return r;
}

template <typename T>
__global__ void k_apply_up(const T *__restrict__ psi, int64_t dupI, T *__restrict__ out,
                           int64_t dupJ, int64_t ddw, const int32_t *__restrict__ src,
                           const int8_t *__restrict__ sgn, double *__restrict__ partials)
{
    __shared__ double red[16];
    double s2 = 0.0;
    const int64_t n = dupJ * ddw;
    // (column, row) of the flat index advance by the fixed grid stride, so
    // the only 64-bit divisions are the three below
    const int64_t e0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t qs = stride / dupJ, rs = stride - qs * dupJ;
    int64_t c = e0 / dupJ, r = e0 - c * dupJ;
    for (int64_t e = e0; e < n; e += stride) {
        const int32_t j = src[r];
        const double v = j >= 0 ? (double)sgn[r] * (double)psi[j + c * dupI] : 0.0;
        out[e] = (T)v;
        s2 += v * v;
        c += qs;
        r += rs;
        if (r >= dupJ) {
            r -= dupJ;
            ++c;
        }
    }
    double v = s2;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
        partials[blockIdx.x] = t;
    }
}

template <typename T>
__global__ void k_apply_dw(const T *__restrict__ psi, int64_t dup, T *__restrict__ out,
                           int64_t ddwJ, const int32_t *__restrict__ src,
                           const int8_t *__restrict__ sgn, double gsign, double *__restrict__ partials)
{
    __shared__ double red[16];
    double s2 = 0.0;
    const int64_t n = dup * ddwJ;
    const int64_t e0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t qs = stride / dup, rs = stride - qs * dup;
    int64_t c = e0 / dup, r = e0 - c * dup;
    for (int64_t e = e0; e < n; e += stride) {
        const int32_t j = src[c];
        const double v = j >= 0 ? gsign * (double)sgn[c] * (double)psi[r + (int64_t)j * dup] : 0.0;
        out[e] = (T)v;
        s2 += v * v;
        c += qs;
        r += rs;
        if (r >= dup) {
            r -= dup;
            ++c;
        }
    }
    double v = s2;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
        partials[blockIdx.x] = t;
    }
}

// src/sign of every target string of spin space J from spin space I for
// c^+ (dagger) or c on bit p; -1 where the operator annihilates the state.
static void op_tables(const SpinTables &I, const SpinTables &J, int p, int dagger,
                      std::vector<int32_t> &src, std::vector<int8_t> &sgn)
{
    src.assign(J.dim, -1);
    sgn.assign(J.dim, 0);
    const uint32_t bit = 1u << p;
    for (int64_t r = 0; r < J.dim; ++r) {
        const uint32_t t = J.h_str[r];
        const bool has = (t & bit) != 0;
        if (dagger ? !has : has) continue;     // c^+ leaves the bit set, c clears it
        const uint32_t s = dagger ? (t & ~bit) : (t | bit);
        src[r] = host_rank(s);
        sgn[r] = (__builtin_popcount(s & (bit - 1u)) & 1) ? -1 : 1;
    }
}

}  // namespace dmed

extern "C" int dmed_apply_op(const dmed_vec *psi, int32_t isite, int32_t dagger,
                             const dmed_sector *dst, dmed_vec *out, double *norm2)
{
    REQUIRE_MODEL();
    if (!psi || !dst || !out) return fail(DMED_EINVAL, "null argument");
    if (int rc = check_sector(dst)) return rc;
    if (int rc = check_sector(psi->sec)) return rc;
    const dmed_sector *I = psi->sec;
    const int ns = I->ns;
    if (isite < 1 || isite > 2 * ns) return fail(DMED_EINVAL, "isite out of range");
    const bool up = isite <= ns;
    const int p = up ? isite - 1 : isite - 1 - ns;
    const int dn = dagger ? 1 : -1;
    if (dst->nup != I->nup + (up ? dn : 0) || dst->ndw != I->ndw + (up ? 0 : dn))
        return fail(DMED_EINVAL, "destination sector does not match the operator");
    if (out->sec != dst || out->prec != psi->prec)
        return fail(DMED_EINVAL, "output vector must belong to dst with psi's precision");
    Context &c = ctx();
    hipStream_t st = ex().stream;
    std::vector<int32_t> src;
    std::vector<int8_t> sgn;
    op_tables(up ? I->up : I->dw, up ? dst->up : dst->dw, p, dagger, src, sgn);
    int32_t *dsrc = nullptr;
    if (int rc = exec_buf(ex(), kSlotOpTable, std::max<size_t>(src.size(), 1) * 5, (void **)&dsrc)) return rc;
    int8_t *dsgn = reinterpret_cast<int8_t *>(dsrc + std::max<size_t>(src.size(), 1));
    DMED_HIP(hipMemcpyAsync(dsrc, src.data(), src.size() * 4, hipMemcpyHostToDevice, st));
    DMED_HIP(hipMemcpyAsync(dsgn, sgn.data(), sgn.size(), hipMemcpyHostToDevice, st));
    const int64_t n = dst->dim;
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(2048, (n + 255) / 256));
    if (int rc = ensure_partials(4096)) return rc;
    // the down-spin operator passes all up electrons: global sign (-1)^nup
    const double gsign = (I->nup & 1) ? -1.0 : 1.0;
    if (psi->prec == DMED_FP64) {
        if (up)
            hipLaunchKernelGGL(k_apply_up<double>, dim3(g), dim3(256), 0, st, (const double *)psi->data,
                               I->up.dim, (double *)out->data, dst->up.dim, dst->dw.dim, dsrc, dsgn,
                               ex().partials);
        else
            hipLaunchKernelGGL(k_apply_dw<double>, dim3(g), dim3(256), 0, st, (const double *)psi->data,
                               I->up.dim, (double *)out->data, dst->dw.dim, dsrc, dsgn, gsign,
                               ex().partials);
    } else {
        if (up)
            hipLaunchKernelGGL(k_apply_up<float>, dim3(g), dim3(256), 0, st, (const float *)psi->data,
                               I->up.dim, (float *)out->data, dst->up.dim, dst->dw.dim, dsrc, dsgn,
                               ex().partials);
        else
            hipLaunchKernelGGL(k_apply_dw<float>, dim3(g), dim3(256), 0, st, (const float *)psi->data,
                               I->up.dim, (float *)out->data, dst->dw.dim, dsrc, dsgn, gsign,
                               ex().partials);
    }
    DMED_HIP(hipGetLastError());
    double *d = ex().dscal + 10;
    if (int rc = launch_reduce_to(ex().partials, (int)g, d, 0, st)) return rc;
    double h = 0.0;
    DMED_HIP(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, st));
    DMED_HIP(hipStreamSynchronize(st));
    if (norm2) *norm2 = h;
    return 0;
}

extern "C" int dmed_chain(const dmed_sector *s, const dmed_vec *v0, int32_t nlanc, int32_t prec,
                          double threshold, double *alpha, double *beta, int32_t *nsteps)
{
    REQUIRE_MODEL();
    if (int rc = check_sector(s)) return rc;
    if (!v0 || v0->sec != s) return fail(DMED_EINVAL, "start vector must belong to the sector");
    if (!alpha || !beta) return fail(DMED_EINVAL, "null alpha/beta");
    int ns = 0;
    int rc = chain_run(s, v0->data, v0->prec, nlanc, prec, threshold, alpha, beta, &ns);
    if (nsteps) *nsteps = ns;
    return rc;
}

extern "C" int dmed_chain_noalloc(const dmed_sector *s, const dmed_vec *v0, int32_t nlanc,
                                  int32_t prec, double threshold, double *alpha, double *beta,
                                  int32_t *nsteps)
{
    REQUIRE_MODEL();
    if (int rc = check_sector(s)) return rc;
    if (!v0 || v0->sec != s) return fail(DMED_EINVAL, "start vector must belong to the sector");
    if (!alpha || !beta) return fail(DMED_EINVAL, "null alpha/beta");
    int ns = 0;
    int rc = chain_run_noalloc(s, v0->data, v0->prec, nlanc, prec, threshold, alpha, beta, &ns);
    if (nsteps) *nsteps = ns;
    return rc;
}

extern "C" int dmed_gf_add_chain(double peso_bz, double ei, const double *alpha, const double *beta,
                                 int32_t nlanc, int32_t isign, const double *wm, int32_t lmats,
                                 const double *wr, int32_t lreal, double eps, double *gmats,
                                 double *greal)
{
    if (nlanc <= 0) return 0;
    std::vector<double> d(alpha, alpha + nlanc), e(nlanc, 0.0), z(nlanc);
    for (int i = 1; i < nlanc; ++i) e[i] = beta[i];
    if (tql2_first_row(nlanc, d.data(), e.data(), z.data()))
        return fail(DMED_ENOCONV, "tql2 did not converge");
    for (auto &x : z) x = x * x;
    gf_add_poles(peso_bz, ei, d.data(), z.data(), nlanc, isign, wm, lmats, wr, lreal, eps, gmats,
                 greal);
    return 0;
}

// --------------------------------------------------------- observables
namespace dmed {

template <typename T>
__global__ void k_obs(const T *__restrict__ psi, int64_t dup, int64_t ddw,
                      const uint8_t *__restrict__ impu, const uint8_t *__restrict__ impd, int norb,
                      double *__restrict__ partials)
{
    __shared__ double red[6][16];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const int64_t n = dup * ddw;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / dup, r = e - c * dup;
        const double w = (double)psi[e] * (double)psi[e];
        const unsigned iu = impu[r], id = impd[c];
        for (int a = 0; a < norb; ++a) {
            const double nu = (iu >> a) & 1u, nd = (id >> a) & 1u;
            acc[2 * a] += (nu + nd) * w;
            acc[2 * a + 1] += nu * nd * w;
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int q = 0; q < 2 * norb; ++q) {
        double v = acc[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[q][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < 2 * norb) {
        double t = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[threadIdx.x][i];
        partials[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

// Joint weight of the impurity occupations: P[bu*nimp + bd] = sum |psi|^2 over
// the elements whose up / down strings carry impurity bits bu / bd.  A block
// owns whole columns (fixed bd); per column the lanes sum by bu in registers
// and the block reduces in a fixed order, so the result is deterministic.
// partials[bin * gridDim.x + block], nimp <= 8.
template <typename T>
__global__ void __launch_bounds__(256) k_imp_hist(const T *__restrict__ psi, int64_t dup, int64_t ddw,
                                                  const uint8_t *__restrict__ impu,
                                                  const uint8_t *__restrict__ impd, int nimp,
                                                  double *__restrict__ partials)
{
    __shared__ double red[8][4];
    __shared__ double hist[64];
    for (int t = threadIdx.x; t < 64; t += blockDim.x) hist[t] = 0.0;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int64_t c = blockIdx.x; c < ddw; c += gridDim.x) {
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const T *col = psi + c * dup;
        for (int64_t r = threadIdx.x; r < dup; r += blockDim.x) {
            const double x = (double)col[r];
            const unsigned iu = impu[r];
#pragma unroll
            for (int q = 0; q < 8; ++q) acc[q] += (iu == (unsigned)q) ? x * x : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            double v = acc[q];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if (lane == 0) red[q][wv] = v;
        }
        __syncthreads();
        if (threadIdx.x < nimp) {
            double t = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[threadIdx.x][w];
            hist[threadIdx.x * nimp + impd[c]] += t;
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < nimp * nimp; t += blockDim.x)
        partials[(int64_t)t * gridDim.x + blockIdx.x] = hist[t];
}

// <psi|O|psi> of the spin-exchange (O_x) and pair-hopping (O_p) operators,
// sum over ordered impurity pairs a != b (ED_OBSERVABLES.f90:597-647): per
// element, the up pair hop b -> a of its row with the down pair hop a -> b
// (O_x) or b -> a (O_p) of its column, signs from the pair tables.
// partials[block] (O_x) and partials[gridDim.x + block] (O_p).
template <typename T>
__global__ void __launch_bounds__(256) k_jterm_sums(const T *__restrict__ psi, int64_t dup,
                                                    int64_t ddw, const int32_t *__restrict__ ju,
                                                    const int32_t *__restrict__ jd, int no,
                                                    double *__restrict__ partials)
{
    __shared__ double red[2][4];
    double sx = 0.0, sp = 0.0;
    const int64_t n = dup * ddw;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / dup, r = e - c * dup;
        const double x = (double)psi[e];
        for (int pa = 0; pa < no; ++pa)
            for (int pb = 0; pb < no; ++pb) {
                if (pa == pb) continue;
                const int32_t wu = ju[(int64_t)(pa * no + pb) * dup + r];
                if (wu < 0) continue;
                const int64_t ru = wu >> 1;
                const double su = (wu & 1) ? -1.0 : 1.0;
                const int32_t wx = jd[(int64_t)(pb * no + pa) * ddw + c];
                if (wx >= 0) sx += x * su * ((wx & 1) ? -1.0 : 1.0) * (double)psi[(int64_t)(wx >> 1) * dup + ru];
                const int32_t wp = jd[(int64_t)(pa * no + pb) * ddw + c];
                if (wp >= 0) sp += x * su * ((wp & 1) ? -1.0 : 1.0) * (double)psi[(int64_t)(wp >> 1) * dup + ru];
            }
    }
    for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_down(sx, off, 64);
        sp += __shfl_down(sp, off, 64);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wv] = sx;
        red[1][wv] = sp;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[threadIdx.x][w];
        partials[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

// out = (wu*n_up + wd*n_dw) psi, occupations summed over the impurity
// orbitals in `mask` (bit a = orbital a): S_z with (1/2, -1/2), the charge
// N with (1, 1).  The impurity bytes of the row and the column give the
// occupations; the factor is an exact small multiple of 1/2.  Partials of
// ||out||^2 per block.
template <typename T>
__global__ void __launch_bounds__(256) k_apply_occ(const T *__restrict__ psi, int64_t dup,
                                                   int64_t ddw, const uint8_t *__restrict__ impu,
                                                   const uint8_t *__restrict__ impd, unsigned mask,
                                                   double wu, double wd, T *__restrict__ out,
                                                   double *__restrict__ partials)
{
    __shared__ double red[4];
    double s2 = 0.0;
    const int64_t n = dup * ddw;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / dup, r = e - c * dup;
        const double f = wu * (double)__popc((unsigned)impu[r] & mask) +
                         wd * (double)__popc((unsigned)impd[c] & mask);
        const T v = (T)(f * (double)psi[e]);
        out[e] = v;
        s2 += (double)v * (double)v;
    }
    for (int off = 32; off > 0; off >>= 1) s2 += __shfl_down(s2, off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = s2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
        partials[blockIdx.x] = t;
    }
}

// out = (wu*n_up + wd*n_dw)(mask) psi on the calling context's stream; psi
// and out are vectors of sector s in precision prec (out may point into a
// larger buffer, e.g. one half of a stacked vector); norm2 = ||out||^2
int apply_occ(const dmed_sector *s, int prec, const void *psi, unsigned mask, double wu,
              double wd, void *out, double *norm2)
{
    hipStream_t st = ex().stream;
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(2048, (s->dim + 255) / 256));
    if (int rc = ensure_partials(4096)) return rc;
    if (prec == DMED_FP64)
        hipLaunchKernelGGL(k_apply_occ<double>, dim3(g), dim3(256), 0, st, (const double *)psi,
                           s->up.dim, s->dw.dim, s->up.imp, s->dw.imp, mask, wu, wd,
                           (double *)out, ex().partials);
    else
        hipLaunchKernelGGL(k_apply_occ<float>, dim3(g), dim3(256), 0, st, (const float *)psi,
                           s->up.dim, s->dw.dim, s->up.imp, s->dw.imp, mask, wu, wd,
                           (float *)out, ex().partials);
    DMED_HIP(hipGetLastError());
    std::vector<double> h(g);
    DMED_HIP(hipMemcpyAsync(h.data(), ex().partials, g * 8, hipMemcpyDeviceToHost, st));
    DMED_HIP(hipStreamSynchronize(st));
    double t = 0.0;
    for (unsigned b = 0; b < g; ++b) t += h[b];
    *norm2 = t;
    return 0;
}

}  // namespace dmed

extern "C" int dmed_apply_sz(const dmed_vec *psi, int32_t iorb, dmed_vec *out, double *norm2)
{
    REQUIRE_MODEL();
    if (!psi || !out || !norm2) return fail(DMED_EINVAL, "null argument");
    const dmed_sector *s = psi->sec;
    if (int rc = check_sector(s)) return rc;
    if (iorb < -1 || iorb >= s->norb) return fail(DMED_EINVAL, "iorb out of range");
    if (out->sec != s || out->prec != psi->prec || out == psi)
        return fail(DMED_EINVAL, "output vector must be another vector of psi's sector and precision");
    const unsigned mask = iorb < 0 ? (1u << s->norb) - 1u : 1u << iorb;
    return dmed::apply_occ(s, psi->prec, psi->data, mask, 0.5, -0.5, out->data, norm2);
}

extern "C" int dmed_apply_occ(const dmed_vec *psi, uint32_t mask, dmed_vec *out, double *norm2)
{
    REQUIRE_MODEL();
    if (!psi || !out || !norm2) return fail(DMED_EINVAL, "null argument");
    const dmed_sector *s = psi->sec;
    if (int rc = check_sector(s)) return rc;
    if (mask == 0 || mask >= (1u << s->norb)) return fail(DMED_EINVAL, "orbital mask out of range");
    if (out->sec != s || out->prec != psi->prec || out == psi)
        return fail(DMED_EINVAL, "output vector must be another vector of psi's sector and precision");
    return dmed::apply_occ(s, psi->prec, psi->data, mask, 1.0, 1.0, out->data, norm2);
}

extern "C" int dmed_chi_add_chain(double peso_f, double peso_bz, double ei, const double *alpha,
                                  const double *beta_l, int32_t nlanc, double beta,
                                  const double *vm, int32_t nv, const double *tau, int32_t ntau,
                                  const double *wr, int32_t lreal, double eps, double *chi_iv,
                                  double *chi_tau, double *chi_w)
{
    if (nlanc <= 0) return 0;
    if (!alpha || !beta_l || (nv > 0 && (!vm || !chi_iv)) || (ntau > 0 && (!tau || !chi_tau)) ||
        (lreal > 0 && (!wr || !chi_w)))
        return fail(DMED_EINVAL, "null argument");
    std::vector<double> d(alpha, alpha + nlanc), e(nlanc, 0.0), z(nlanc);
    for (int i = 1; i < nlanc; ++i) e[i] = beta_l[i];
    if (tql2_first_row(nlanc, d.data(), e.data(), z.data()))
        return fail(DMED_ENOCONV, "tql2 did not converge");
    for (auto &x : z) x = x * x;
    chi_add_poles(peso_f, peso_bz, ei, d.data(), z.data(), nlanc, beta, vm, nv, tau, ntau, wr,
                  lreal, eps, chi_iv, chi_tau, chi_w);
    return 0;
}

extern "C" int dmed_observables(const dmed_vec *psi, double peso, double *dens, double *docc)
{
    REQUIRE_MODEL();
    if (!psi || !dens || !docc) return fail(DMED_EINVAL, "null argument");
    const dmed_sector *s = psi->sec;
    if (int rc = check_sector(s)) return rc;
    Context &c = ctx();
    hipStream_t st = ex().stream;
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, (s->dim + 255) / 256));
    if (int rc = ensure_partials(6 * 1024)) return rc;
    if (psi->prec == DMED_FP64)
        hipLaunchKernelGGL(k_obs<double>, dim3(g), dim3(256), 0, st, (const double *)psi->data,
                           s->up.dim, s->dw.dim, s->up.imp, s->dw.imp, s->norb, ex().partials);
    else
        hipLaunchKernelGGL(k_obs<float>, dim3(g), dim3(256), 0, st, (const float *)psi->data,
                           s->up.dim, s->dw.dim, s->up.imp, s->dw.imp, s->norb, ex().partials);
    DMED_HIP(hipGetLastError());
    std::vector<double> h((size_t)2 * s->norb * g);
    DMED_HIP(hipMemcpyAsync(h.data(), ex().partials, h.size() * 8, hipMemcpyDeviceToHost, st));
    DMED_HIP(hipStreamSynchronize(st));
    for (int a = 0; a < s->norb; ++a) {
        double d1 = 0.0, d2 = 0.0;
        for (unsigned b = 0; b < g; ++b) {
            d1 += h[(size_t)(2 * a) * g + b];
            d2 += h[(size_t)(2 * a + 1) * g + b];
        }
        dens[a] += peso * d1;
        docc[a] += peso * d2;
    }
    return 0;
}

// Local observables and energies of one normalised state (observables_impurity,
// ED_OBSERVABLES.f90:74-150; local_energy_impurity, :450-669).  Every diagonal
// quantity is a function of the impurity bits only, so it follows from the
// joint weight P(bu, bd); the off-diagonal Hloc part of eknot is
// <psi|c+_a c_b|psi> = <c_a psi, c_b psi>, from the annihilated vectors.
extern "C" int dmed_local_observables(const dmed_vec *psi, double peso, dmed_local_obs *acc)
{
    REQUIRE_MODEL();
    if (!psi || !acc) return fail(DMED_EINVAL, "null argument");
    const dmed_sector *s = psi->sec;
    if (int rc = check_sector(s)) return rc;
    const Model &md = cur_model();
    const int no = md.norb, nimp = s->nimp, sd = md.nspin - 1;
    if (nimp > 8) return fail(DMED_EUNSUPP, "local observables: Norb <= 3");
    hipStream_t st = ex().stream;
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, s->dw.dim));
    if (int rc = ensure_partials((size_t)64 * g + 16)) return rc;
    if (psi->prec == DMED_FP64)
        hipLaunchKernelGGL(dmed::k_imp_hist<double>, dim3(g), dim3(256), 0, st,
                           (const double *)psi->data, s->up.dim, s->dw.dim, s->up.imp, s->dw.imp,
                           nimp, ex().partials);
    else
        hipLaunchKernelGGL(dmed::k_imp_hist<float>, dim3(g), dim3(256), 0, st,
                           (const float *)psi->data, s->up.dim, s->dw.dim, s->up.imp, s->dw.imp,
                           nimp, ex().partials);
    DMED_HIP(hipGetLastError());
    std::vector<double> h((size_t)nimp * nimp * g);
    DMED_HIP(hipMemcpyAsync(h.data(), ex().partials, h.size() * 8, hipMemcpyDeviceToHost, st));
    DMED_HIP(hipStreamSynchronize(st));
    dmed_local_obs o{};
    for (int bu = 0; bu < nimp; ++bu)
        for (int bd = 0; bd < nimp; ++bd) {
            double P = 0.0;
            for (unsigned b = 0; b < g; ++b) P += h[(size_t)(bu * nimp + bd) * g + b];
            const double w = peso * P;
            if (w == 0.0) continue;
            double nup[3] = {}, ndw[3] = {}, sz[3] = {}, nt[3] = {}, szs = 0.0;
            for (int a = 0; a < no; ++a) {
                nup[a] = (bu >> a) & 1;
                ndw[a] = (bd >> a) & 1;
                sz[a] = 0.5 * (nup[a] - ndw[a]);
                nt[a] = nup[a] + ndw[a];
                szs += sz[a];
            }
            for (int a = 0; a < no; ++a) {
                o.dens[a] += nt[a] * w;
                o.dens_up[a] += nup[a] * w;
                o.dens_dw[a] += ndw[a] * w;
                o.docc[a] += nup[a] * ndw[a] * w;
                o.magz[a] += (nup[a] - ndw[a]) * w;
                for (int b = 0; b < no; ++b) {
                    o.sz2[a * 3 + b] += sz[a] * sz[b] * w;
                    o.n2[a * 3 + b] += nt[a] * nt[b] * w;
                }
                o.eknot += (md.hloc[0][a][a] * nup[a] + md.hloc[sd][a][a] * ndw[a]) * w;
                o.epot += md.uloc[a] * nup[a] * ndw[a] * w;
            }
            o.s2tot += szs * szs * w;
            for (int a = 0; a < no; ++a)
                for (int b = a + 1; b < no; ++b) {
                    const double du = nup[a] * ndw[b] + nup[b] * ndw[a];
                    const double dn = nup[a] * nup[b] + ndw[a] * ndw[b];
                    o.epot += md.ust * du * w + (md.ust - md.jh) * dn * w;
                    o.dust += du * w;
                    o.dund += dn * w;
                }
            if (md.hfmode) {
                double eh = 0.0;
                for (int a = 0; a < no; ++a) {
                    eh += -0.5 * md.uloc[a] * nt[a] * w + 0.25 * md.uloc[a] * w;
                    for (int b = a + 1; b < no; ++b) {
                        const double nn = nt[a] + nt[b];
                        eh += -0.5 * md.ust * nn * w + 0.25 * md.ust * w;
                        eh += -0.5 * (md.ust - md.jh) * nn * w + 0.25 * (md.ust - md.jh) * w;
                    }
                }
                o.ehartree += eh;
                o.epot += eh;
            }
        }
    // off-diagonal Hloc: sum_{a != b} Hloc_ab <c_a psi, c_b psi> per spin (weight 1)
    for (int sp = 0; sp < 2; ++sp) {
        const double(*hl)[3] = md.hloc[sp == 0 ? 0 : sd];
        bool any = false;
        for (int a = 0; a < no; ++a)
            for (int b = 0; b < no; ++b) any = any || (a != b && hl[a][b] != 0.0);
        const int nocc = sp == 0 ? s->nup : s->ndw;
        if (!any || nocc == 0) continue;
        if (psi->prec != DMED_FP64)
            return fail(DMED_EINVAL, "off-diagonal Hloc observables need an FP64 state");
        dmed_sector *t = nullptr;
        if (int rc = dmed_sector_create(s->nup - (sp == 0), s->ndw - (sp == 1), &t)) return rc;
        std::vector<dmed_vec *> cv(no, nullptr);
        int rc = 0;
        for (int a = 0; a < no && !rc; ++a) {
            rc = dmed_vec_create(t, DMED_FP64, &cv[a]);
            double n2 = 0.0;
            if (!rc) rc = dmed_apply_op(psi, a + 1 + sp * s->ns, 0, t, cv[a], &n2);
        }
        for (int a = 0; a < no && !rc; ++a)
            for (int b = 0; b < no && !rc; ++b) {
                if (a == b || hl[a][b] == 0.0) continue;
                double d = 0.0;
                rc = dot_host((const double *)cv[a]->data, (const double *)cv[b]->data, t->dim, &d);
                o.eknot += hl[a][b] * d;
            }
        for (auto *v : cv) dmed_vec_free(v);
        dmed_sector_free(t);
        if (rc) return rc;
    }
    // spin exchange and pair hopping, per state without the Boltzmann weight
    // (ED_OBSERVABLES.f90:597-647, as the reference accumulates them)
    if (s->jterms) {
        const unsigned gj = (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, (s->dim + 255) / 256));
        if (int rc = ensure_partials((size_t)2 * gj + 16)) return rc;
        if (psi->prec == DMED_FP64)
            hipLaunchKernelGGL(dmed::k_jterm_sums<double>, dim3(gj), dim3(256), 0, st,
                               (const double *)psi->data, s->up.dim, s->dw.dim, s->up.ipair,
                               s->dw.ipair, no, ex().partials);
        else
            hipLaunchKernelGGL(dmed::k_jterm_sums<float>, dim3(gj), dim3(256), 0, st,
                               (const float *)psi->data, s->up.dim, s->dw.dim, s->up.ipair,
                               s->dw.ipair, no, ex().partials);
        DMED_HIP(hipGetLastError());
        std::vector<double> hj((size_t)2 * gj);
        DMED_HIP(hipMemcpyAsync(hj.data(), ex().partials, hj.size() * 8, hipMemcpyDeviceToHost, st));
        DMED_HIP(hipStreamSynchronize(st));
        double ox = 0.0, op = 0.0;
        for (unsigned b = 0; b < gj; ++b) {
            ox += hj[b];
            op += hj[gj + b];
        }
        o.dse += ox;
        o.dph += op;
        o.epot += md.jx * ox + md.jp * op;
    }
    double *dst = reinterpret_cast<double *>(acc);
    const double *src = reinterpret_cast<const double *>(&o);
    for (size_t q = 0; q < sizeof(dmed_local_obs) / sizeof(double); ++q) dst[q] += src[q];
    return 0;
}

// ------------------------------------------------------- bath functions
extern "C" int dmed_delta(int32_t ispin, int32_t iorb, const double *z, int32_t L, double *delta)
{
    REQUIRE_MODEL();
    const Model &m = cur_model();
    if (ispin < 0 || ispin >= m.nspin || iorb < 0 || iorb >= m.norb)
        return fail(DMED_EINVAL, "spin/orbital out of range");
    using cd = std::complex<double>;
    const cd *zz = reinterpret_cast<const cd *>(z);
    cd *o = reinterpret_cast<cd *>(delta);
    for (int i = 0; i < L; ++i) {
        cd d = 0.0;
        for (int k = 0; k < m.nbath; ++k) {
            const double v = m.V(ispin, iorb, k);
            d += v * v / (zz[i] - m.E(ispin, iorb, k));
        }
        o[i] = d;
    }
    return 0;
}

extern "C" int dmed_sigma(int32_t ispin, int32_t iorb, const double *z, const double *g, int32_t L,
                          double *sigma)
{
    if (int rc = dmed_delta(ispin, iorb, z, L, sigma)) return rc;
    const Model &m = cur_model();
    if (m.hybrid && m.norb > 1)
        return fail(DMED_EUNSUPP, "hybrid bath with Norb > 1: Sigma is an orbital matrix (dmed_solve)");
    using cd = std::complex<double>;
    const cd *zz = reinterpret_cast<const cd *>(z);
    const cd *gg = reinterpret_cast<const cd *>(g);
    cd *o = reinterpret_cast<cd *>(sigma);
    for (int i = 0; i < L; ++i) {
        const cd invg0 = zz[i] + m.xmu - m.hloc[ispin][iorb][iorb] - o[i];
        o[i] = invg0 - 1.0 / gg[i];
    }
    return 0;
}
