// krylov.hip -- Lanczos engines on the device: the plain Green's-function
// chain (sp_lanc_tridiag semantics) and the ground-state eigensolver that
// replaces sp_eigh.
//
// Chain step k (.repo/PLAIN_LANCZOS.f90:87-118, with the vector scale kept in
// a device scalar instead of a rescaling pass):
//   K1  w = sx*H v - b*svp*v_prev,   partials of <sx v, w>        (H*v kernel)
//   F1  alpha_k = sum partials
//   K2  w -= alpha_k*sx*v,           partials of <w, w>
//   F2  beta_k = sqrt(sum); v_prev <- v (scale sx); v <- w (scale 1/beta_k)
// All scalars stay on the device, so a chain is a fixed kernel sequence with
// no host round trip; reductions are two-level and deterministic.
#include "dmed_internal.h"

#include <cstdlib>
#include <cstdio>
#include <atomic>
#include <map>
#include <mutex>
#include <string>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace dmed {

int hxv_nparts(const dmed_sector *s, int prec);

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ __forceinline__ double bsum(double v, double *sh)
{
    v = wsum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    if (w == 0) {
        v = (lane < nw) ? sh[lane] : 0.0;
        v = wsum(v);
    }
    return v;
}

constexpr int kVecBS = 256;
constexpr int kVecGrid = 2048;   // grid-stride vector kernels: fixed partial count

static unsigned vec_grid(int64_t n)
{
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(kVecGrid, (n + kVecBS - 1) / kVecBS));
}

template <typename TX, typename TY>
__global__ void __launch_bounds__(kVecBS) k_dot(const TX *__restrict__ x, const TY *__restrict__ y,
                                                int64_t n, double *__restrict__ partials)
{
    __shared__ double red[16];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        s += (double)x[i] * (double)y[i];
    s = bsum(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// per-thread sum of p[threadIdx.x + k*blockDim.x], k increasing (the order
// of a plain strided loop), with eight loads in flight: they are issued
// unconditionally at indices clamped into range, and only the in-range ones
// are added (a loop that loads and adds one element at a time waits for
// each load in turn)
__device__ __forceinline__ double strided_sum(const double *__restrict__ p, int n)
{
    double s = 0.0;
    const int bd = blockDim.x;
    for (int i0 = threadIdx.x; i0 < n; i0 += 8 * bd) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = p[min(i0 + k * bd, n - 1)];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (i0 + k * bd < n) s += v[k];
    }
    return s;
}

// fixed-order final sum; op 0: plain sum, 1: sqrt of sum
__global__ void __launch_bounds__(1024) k_reduce(const double *__restrict__ p, int n,
                                                 double *__restrict__ out, int op)
{
    __shared__ double red[16];
    double s = strided_sum(p, n);
    s = bsum(s, red);
    if (threadIdx.x == 0) *out = op == 1 ? sqrt(s) : s;
}

// chain state init: sx = 1/||v0||, b = 0
__global__ void __launch_bounds__(1024) k_chain_init(const double *__restrict__ p, int n,
                                                     ChainScalars *cs, double thr, int nlanc)
{
    __shared__ double red[16];
    double s = strided_sum(p, n);
    s = bsum(s, red);
    if (threadIdx.x == 0) {
        cs->sx = 1.0 / sqrt(s);
        cs->svp = 0.0;
        cs->b = 0.0;
        cs->alpha = 0.0;
        cs->q = 0.0;
        cs->thr = thr;
        cs->step = 0;
        cs->done = (s == 0.0) ? 1 : 0;
        cs->nlanc = nlanc;
    }
}

__global__ void __launch_bounds__(1024) k_chain_alpha(const double *__restrict__ p, int n,
                                                      ChainScalars *cs, double *__restrict__ alpha)
{
    __shared__ double red[16];
    if (cs->done) return;
    double s = strided_sum(p, n);
    s = bsum(s, red);
    if (threadIdx.x == 0) {
        // alpha = <v, Hv - b v_prev> = sx*(<x, y> - b*svp*<x, x_prev>)
        const double a = cs->sx * (s - cs->b * cs->svp * cs->q);
        cs->alpha = a;
        alpha[cs->step] = a;
    }
}

// 16-byte vectors of T
template <typename T> struct Vec16;
template <> struct Vec16<double> { using type = double2; };
template <> struct Vec16<float> { using type = float4; };

// w = Hv - b*v_prev - alpha*v in place of y = Hv (stored scales sx, svp), with
// the block partials of ||w||^2 (first gridDim.x) and <w, x> (next gridDim.x).
// Each thread takes the elements i = tid + k*stride in increasing order, as a
// plain grid-stride loop would, with four of them loaded before the first is
// used; the partial sums (hence beta) do not depend on that grouping.
template <typename T>
__global__ void __launch_bounds__(kVecBS) k_chain_update(T *__restrict__ w, const T *__restrict__ v,
                                                         const T *__restrict__ vp, int64_t n,
                                                         const ChainScalars *cs,
                                                         double *__restrict__ partials)
{
    constexpr int U = 4;
    __shared__ double red[16];
    if (cs->done) return;
    const double ca = cs->alpha * cs->sx, cb = cs->b * cs->svp;
    double s1 = 0.0, s2 = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += U * stride) {
        T a[U], b[U], c[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int64_t e = min(i + k * stride, n - 1);
            a[k] = w[e];
            b[k] = v[e];
            c[k] = vp[e];
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            if (i + k * stride >= n) break;
            const double vi = (double)b[k];
            const T wt = (T)((double)a[k] - cb * (double)c[k] - ca * vi);
            w[i + k * stride] = wt;
            const double wi = (double)wt;
            s1 += wi * wi;
            s2 += wi * vi;
        }
    }
    s1 = bsum(s1, red);
    __syncthreads();
    s2 = bsum(s2, red);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s1;
        partials[gridDim.x + blockIdx.x] = s2;
    }
}

__global__ void __launch_bounds__(1024) k_chain_beta(const double *__restrict__ p, int n,
                                                     ChainScalars *cs, double *__restrict__ beta)
{
    __shared__ double red[16];
    if (cs->done) return;
    double s = strided_sum(p, n), q = strided_sum(p + n, n);
    s = bsum(s, red);
    __syncthreads();
    q = bsum(q, red);
    if (threadIdx.x == 0) {
        const double b = sqrt(s);
        const int k = cs->step;
        if (k + 1 < cs->nlanc) beta[k + 1] = b;
        cs->b = b;
        cs->q = q;
        cs->svp = cs->sx;
        cs->sx = 1.0 / b;
        cs->step = k + 1;
        if (fabs(b) < cs->thr || k + 1 >= cs->nlanc) cs->done = 1;
    }
}

// psi += z_k * sx * v  (second pass of the ground-state Lanczos); 16-byte
// loads of v and of psi (psi and v 16-byte aligned), the last n % (16/sizeof(T))
// elements one per thread
template <typename T>
__global__ void __launch_bounds__(kVecBS) k_accum(double *__restrict__ psi, const T *__restrict__ v,
                                                  int64_t n, const ChainScalars *cs,
                                                  const double *__restrict__ z)
{
    using V = typename Vec16<T>::type;
    constexpr int E = 16 / sizeof(T);
    const double c = z[cs->step] * cs->sx;
    const int64_t nv = n / E;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const V *vv = reinterpret_cast<const V *>(v);
    double2 *pp = reinterpret_cast<double2 *>(psi);
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nv; j += stride) {
        const V b = vv[j];
        const T *y = reinterpret_cast<const T *>(&b);
#pragma unroll
        for (int k = 0; k < E / 2; ++k) {
            double2 p = pp[j * (E / 2) + k];
            p.x += c * (double)y[2 * k];
            p.y += c * (double)y[2 * k + 1];
            pp[j * (E / 2) + k] = p;
        }
    }
    const int64_t t = nv * E + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) psi[t] += c * (double)v[t];
}

template <typename TD, typename TS>
__global__ void __launch_bounds__(kVecBS) k_scale_copy(TD *__restrict__ d, const TS *__restrict__ s,
                                                       int64_t n, double a)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = (TD)(a * (double)s[i]);
}

template <typename TY, typename TX>
__global__ void __launch_bounds__(kVecBS) k_axpy(TY *__restrict__ y, const TX *__restrict__ x,
                                                 int64_t n, double a)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (TY)((double)y[i] + a * (double)x[i]);
}

// counter-based hash PRNG in [-1,1): SplitMix64 of (seed, index)
__device__ __forceinline__ double hash_uniform(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + 0x9e3779b97f4a7c15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

template <typename T>
__global__ void __launch_bounds__(kVecBS) k_fill_random(T *__restrict__ x, int64_t n, uint64_t seed)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (T)hash_uniform(seed, (uint64_t)i);
}

// zero n 32-bit words (a kernel rather than a memset node, so that streams a
// caller captures into a graph replay it correctly)
__global__ void __launch_bounds__(kVecBS) k_zero_words(uint32_t *__restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

int launch_zero(void *p, size_t bytes, hipStream_t st)
{
    const int64_t n = (int64_t)(bytes / 4);
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_zero_words, dim3(vec_grid(n)), dim3(kVecBS), 0, st, (uint32_t *)p, n);
    DMED_HIP(hipGetLastError());
    return 0;
}

// ----------------------------------------- block-Krylov helpers (FP64)
// One classical Gram-Schmidt pass and its norms in two launches, with no
// cross-block synchronisation: a producer writes per-block partials, and
// every block of the consumer re-sums the few it needs in index order (the
// same values in the same order in every block, hence the same result).
// Cross-block tickets would need agent-scope fences, which on a multi-XCD
// chip write back and invalidate the per-XCD L2 caches.  The grid of these
// kernels is capped so the re-summed partials stay small.
constexpr int kGsGrid = 512;
constexpr int kGsMaxRows = 160;     // Y and V rows (nl + m <= 64 + 64) plus slack

// at most this many partials per thread and row when n <= kGsGrid and the
// block has kVecBS threads: their loads are all issued before the first add
constexpr int kGsPer = (kGsGrid + kVecBS - 1) / kVecBS;

// sum of 1 <= n <= kGsGrid partials in index order, valid in every thread
// (kVecBS threads per block).  The loads are unconditional, at indices
// clamped into range (a select between a load and a constant would make the
// compiler wait for each load); only the in-range ones are added.
__device__ __forceinline__ double psum_all(const double *__restrict__ p, int n, double *red)
{
    double v[kGsPer];
#pragma unroll
    for (int k = 0; k < kGsPer; ++k) v[k] = p[min((int)threadIdx.x + k * kVecBS, n - 1)];
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kGsPer; ++k)
        if ((int)threadIdx.x + k * kVecBS < n) s += v[k];
    s = bsum(s, red);
    __shared__ double out;
    if (threadIdx.x == 0) out = s;
    __syncthreads();
    s = out;
    __syncthreads();
    return s;
}

// c[r] = sum of the 1 <= g <= kGsGrid partials of row r of P (row stride
// g), r < nr, into shared c; sixteen rows at a time with all their loads
// issued before the first add (unconditional, at clamped indices, as in
// psum_all), each thread summing its in-range partials in index order,
// reduced by wave shuffles and one pass over the waves (a fixed order for a
// given g).  kVecBS threads per block; red: shared, >= 16 * waves per block.
__device__ __forceinline__ void rowsums(const double *__restrict__ P, int nr, int g,
                                        double *c, double *red)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = kVecBS >> 6;
    for (int r0 = 0; r0 < nr; r0 += 16) {
        double v[16][kGsPer];
#pragma unroll
        for (int q = 0; q < 16; ++q)
#pragma unroll
            for (int k = 0; k < kGsPer; ++k)
                v[q][k] = P[(int64_t)min(r0 + q, nr - 1) * g + min((int)threadIdx.x + k * kVecBS, g - 1)];
        double s[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            s[q] = 0.0;
#pragma unroll
            for (int k = 0; k < kGsPer; ++k)
                if ((int)threadIdx.x + k * kVecBS < g) s[q] += v[q][k];
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) s[q] = wsum(s[q]);
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < 16; ++q) red[q * nw + w] = s[q];
        }
        __syncthreads();
        if ((int)threadIdx.x < 16 && r0 + (int)threadIdx.x < nr) {
            double t = 0.0;
            for (int k = 0; k < nw; ++k) t += red[threadIdx.x * nw + k];
            c[r0 + threadIdx.x] = t;
        }
        __syncthreads();
    }
}

// Partials of the projections of w on the rows Y_0..Y_{ny-1}, V_0..V_{nv-1}
// and, as row ny + nv, on w itself (when nrows says so): P[r*gridDim.x + block].
// Each block takes rpb <= kProjRows consecutive rows (blockIdx.y), so w is
// read once per rpb rows; every row keeps its own per-thread order and block
// reduction, as with one row per block, so rpb does not change the result.
// Second pass (dgks != 0): runs only when the first pass cancelled more than
// half of ||w||^2 (Daniel-Gragg-Kaufman-Stewart), judged from the first
// pass's partials: n0 (||w||^2 before, g0 values) and n1 (after, g0 values).
constexpr int kProjRows = 8;

// rows per block of k_proj_dot: several for long vectors, whose re-reads of w
// cost HBM traffic; one for short ones, which need the blocks for parallelism
static int proj_rpb(int64_t n) { return n >= ((int64_t)1 << 14) ? kProjRows : 1; }
static unsigned proj_groups(int nrows, int rpb) { return (unsigned)((nrows + rpb - 1) / rpb); }

__global__ void __launch_bounds__(kVecBS) k_proj_dot(
    const double *__restrict__ Y, int ny, const double *__restrict__ V, int nv,
    const double *__restrict__ w, int64_t n, double *__restrict__ P, int dgks,
    const double *__restrict__ n0, const double *__restrict__ n1, int nrows, int rpb)
{
    __shared__ double red[16];
    if (dgks) {
        const double a0 = psum_all(n0, gridDim.x, red);
        const double a1 = psum_all(n1, gridDim.x, red);
        if (!(a1 < 0.5 * a0)) return;
    }
    const int r0 = blockIdx.y * rpb;
    const int nr = min(rpb, nrows - r0);
    const double *b[kProjRows];
#pragma unroll
    for (int q = 0; q < kProjRows; ++q) {
        const int r = min(r0 + q, nrows - 1);
        b[q] = r < ny ? Y + (int64_t)r * n : (r < ny + nv ? V + (int64_t)(r - ny) * n : w);
    }
    double s[kProjRows];
#pragma unroll
    for (int q = 0; q < kProjRows; ++q) s[q] = 0.0;
    // the loads of w and of the block's rows are all issued before the first
    // product (a load and its use under one condition make the compiler
    // wait for each load in turn)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double wi = w[i];
        double x[kProjRows];
#pragma unroll
        for (int q = 0; q < kProjRows; ++q)
            if (q < nr) x[q] = b[q][i];
#pragma unroll
        for (int q = 0; q < kProjRows; ++q)
            if (q < nr) s[q] += x[q] * wi;
    }
#pragma unroll
    for (int q = 0; q < kProjRows; ++q) {
        if (q >= nr) break;
        const double t = bsum(s[q], red);
        if (threadIdx.x == 0) P[(int64_t)(r0 + q) * gridDim.x + blockIdx.x] = t;
        __syncthreads();
    }
}

// w -= sum_r Y_r cy_r + sum_r V_r cv_r, the coefficients summed from the
// row partials P of k_proj_dot; partials of the new ||w||^2 into Pn.
// Block 0 stores the V coefficients at acc (first: overwrite, else add).
// dgks: second pass, gated as in k_proj_dot.
__global__ void __launch_bounds__(kVecBS) k_proj_axpy(
    const double *__restrict__ Y, int ny, const double *__restrict__ V, int nv,
    double *__restrict__ w, int64_t n, const double *__restrict__ P, double *__restrict__ Pn,
    double *__restrict__ acc, int first, int dgks, const double *__restrict__ n0,
    const double *__restrict__ n1)
{
    __shared__ double red[16];
    __shared__ double red16[16 * (kVecBS / 64)];
    __shared__ double c[kGsMaxRows];
    if (dgks) {
        const double a0 = psum_all(n0, gridDim.x, red);
        const double a1 = psum_all(n1, gridDim.x, red);
        if (!(a1 < 0.5 * a0)) return;
    }
    const int nr = ny + nv;
    rowsums(P, nr, gridDim.x, c, red16);
    if (blockIdx.x == 0 && acc)
        for (int r = ny + (int)threadIdx.x; r < nr; r += blockDim.x)
            acc[r - ny] = first ? c[r] : acc[r - ny] + c[r];
    // rows in chunks of kProjRows, the loads of a chunk all issued before
    // its first product; the terms are added in row order
    double s2 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double wv = w[i];
        double s = 0.0;
        for (int j0 = 0; j0 < nr; j0 += kProjRows) {
            double x[kProjRows];
#pragma unroll
            for (int q = 0; q < kProjRows; ++q) {
                const int j = j0 + q;
                if (j < nr) x[q] = (j < ny ? Y + (int64_t)j * n : V + (int64_t)(j - ny) * n)[i];
            }
#pragma unroll
            for (int q = 0; q < kProjRows; ++q)
                if (j0 + q < nr) s += x[q] * c[j0 + q];
        }
        const double x = wv - s;
        w[i] = x;
        s2 += x * x;
    }
    s2 = bsum(s2, red);
    if (threadIdx.x == 0) Pn[blockIdx.x] = s2;
}

// Lanczos three-term step of a thick-restart expansion step j (not the first
// of a restart cycle): w -= alpha v_j + beta v_{j-1} before the full
// Gram-Schmidt pass, with alpha = <v_j, H v_j> summed from the np partials
// that the Lanczos-mode H*v left and beta = *bprev, the norm that normalised
// v_j.  In exact arithmetic H v_j has no other component in the basis, so
// the full pass that follows removes only rounding-level remainders and
// rarely cancels enough to need its second pass.  Block 0 stores alpha and
// beta at acc[j] and acc[j-1] (0 below); the full pass adds its
// coefficients to them.  np >= 1.
__global__ void __launch_bounds__(kVecBS) k_trl_local(double *__restrict__ w,
                                                      const double *__restrict__ vj,
                                                      const double *__restrict__ vjm1, int64_t n,
                                                      const double *__restrict__ pa, int np,
                                                      const double *__restrict__ bprev,
                                                      double *__restrict__ acc, int j)
{
    __shared__ double red[16];
    __shared__ double out;
    // alpha: each thread sums its partials in index order, four loads in
    // flight (unconditional, at clamped indices)
    double s = 0.0;
    for (int i0 = 0; i0 < np; i0 += 4 * kVecBS) {
        double v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = pa[min(i0 + (int)threadIdx.x + k * kVecBS, np - 1)];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i0 + (int)threadIdx.x + k * kVecBS < np) s += v[k];
    }
    s = bsum(s, red);
    if (threadIdx.x == 0) out = s;
    __syncthreads();
    const double al = out, b = *bprev;
    if (blockIdx.x == 0)
        for (int r = threadIdx.x; r <= j; r += blockDim.x)
            acc[r] = r == j ? al : (r == j - 1 ? b : 0.0);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double a = w[i], x = vj[i], y = vjm1[i];
        w[i] = a - al * x - b * y;
    }
}

// dst = w / ||w||, with ||w||^2 the sum of the partials p1 (after the first
// pass) or, when the second pass ran (same test as above on n0, p1), p2;
// 0 when the norm is 0.  Block 0 stores the norm at *nrm.  Without a
// projection (noproj) p1 holds the partials of ||w||^2 directly.
__global__ void __launch_bounds__(kVecBS) k_normalize(double *__restrict__ dst,
                                                      const double *__restrict__ w, int64_t n,
                                                      const double *__restrict__ n0,
                                                      const double *__restrict__ p1,
                                                      const double *__restrict__ p2, int noproj,
                                                      double *__restrict__ nrm)
{
    __shared__ double red[16];
    double q = psum_all(p1, gridDim.x, red);
    if (!noproj) {
        const double a0 = psum_all(n0, gridDim.x, red);
        if (q < 0.5 * a0) q = psum_all(p2, gridDim.x, red);
    }
    const double r = sqrt(q);
    const double a = r > 0.0 ? 1.0 / r : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *nrm = r;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = a * w[i];
}

// in place V[:, 0:k] = V[:, 0:m] S[0:m, 0:k]  (S column-major m x k)
__global__ void __launch_bounds__(128) k_rotate_basis(double *__restrict__ V, int64_t n, int m,
                                                      int k, const double *__restrict__ S)
{
    extern __shared__ double row[];   // [m][128]
    const int64_t base = (int64_t)blockIdx.x * 128;
    const int t = threadIdx.x;
    const int64_t i = base + t;
    if (i < n)
        for (int j = 0; j < m; ++j) row[j * 128 + t] = V[(int64_t)j * n + i];
    __syncthreads();
    if (i >= n) return;
    for (int q = 0; q < k; ++q) {
        double s = 0.0;
        for (int j = 0; j < m; ++j) s += row[j * 128 + t] * S[q * m + j];
        V[(int64_t)q * n + i] = s;
    }
}

// ------------------------------------------------------------ launchers
int launch_reduce_to(const double *partials, int n, double *out, int op, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, partials, n, out, op);
    DMED_HIP(hipGetLastError());
    return 0;
}

int launch_dot(const void *x, const void *y, int64_t n, int prec, double *partials, int *nparts,
               hipStream_t st)
{
    const unsigned g = vec_grid(n);
    if (prec == DMED_FP32)
        hipLaunchKernelGGL((k_dot<float, float>), dim3(g), dim3(kVecBS), 0, st,
                           (const float *)x, (const float *)y, n, partials);
    else
        hipLaunchKernelGGL((k_dot<double, double>), dim3(g), dim3(kVecBS), 0, st,
                           (const double *)x, (const double *)y, n, partials);
    DMED_HIP(hipGetLastError());
    *nparts = (int)g;
    return 0;
}

int launch_fill_random(void *x, int64_t n, int prec, uint64_t seed, hipStream_t st)
{
    const unsigned g = vec_grid(n);
    if (prec == DMED_FP32)
        hipLaunchKernelGGL(k_fill_random<float>, dim3(g), dim3(kVecBS), 0, st, (float *)x, n, seed);
    else
        hipLaunchKernelGGL(k_fill_random<double>, dim3(g), dim3(kVecBS), 0, st, (double *)x, n, seed);
    DMED_HIP(hipGetLastError());
    return 0;
}

// dst (cols x rows) = transpose of src (rows x cols), both column-major FP64:
// 64 x 64 tiles through LDS (one padding column against bank conflicts);
// tiles on the edges are guarded per element.
__global__ void __launch_bounds__(256) k_transpose(const double *__restrict__ src, int64_t rows,
                                                   int64_t cols, double *__restrict__ dst)
{
    __shared__ double tile[64][65];
    const int64_t r0 = (int64_t)blockIdx.x * 64, c0 = (int64_t)blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;     // 64 x 4 threads
    for (int k = ty; k < 64; k += 4) {
        const int64_t r = r0 + tx, c = c0 + k;
        if (r < rows && c < cols) tile[k][tx] = src[r + rows * c];
    }
    __syncthreads();
    for (int k = ty; k < 64; k += 4) {
        const int64_t c = c0 + tx, r = r0 + k;                  // dst(c, r)
        if (r < rows && c < cols) dst[c + cols * r] = tile[tx][k];
    }
}

int launch_transpose(const double *src, int64_t rows, int64_t cols, double *dst, hipStream_t st)
{
    if (rows <= 0 || cols <= 0) return 0;
    const dim3 grid((unsigned)((rows + 63) / 64), (unsigned)((cols + 63) / 64));
    if (grid.y > 65535u) return fail(DMED_EINVAL, "transpose: too many column tiles");
    hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, st, src, rows, cols, dst);
    DMED_HIP(hipGetLastError());
    return 0;
}

int launch_scale_copy(void *dst, int dprec, const void *src, int sprec, int64_t n, double a,
                      hipStream_t st)
{
    const unsigned g = vec_grid(n);
    if (dprec == DMED_FP32 && sprec == DMED_FP32)
        hipLaunchKernelGGL((k_scale_copy<float, float>), dim3(g), dim3(kVecBS), 0, st,
                           (float *)dst, (const float *)src, n, a);
    else if (dprec == DMED_FP32)
        hipLaunchKernelGGL((k_scale_copy<float, double>), dim3(g), dim3(kVecBS), 0, st,
                           (float *)dst, (const double *)src, n, a);
    else if (sprec == DMED_FP32)
        hipLaunchKernelGGL((k_scale_copy<double, float>), dim3(g), dim3(kVecBS), 0, st,
                           (double *)dst, (const float *)src, n, a);
    else
        hipLaunchKernelGGL((k_scale_copy<double, double>), dim3(g), dim3(kVecBS), 0, st,
                           (double *)dst, (const double *)src, n, a);
    DMED_HIP(hipGetLastError());
    return 0;
}

int launch_axpy(void *y, int yprec, const void *x, int xprec, int64_t n, double a, hipStream_t st)
{
    const unsigned g = vec_grid(n);
    if (yprec == DMED_FP64 && xprec == DMED_FP64)
        hipLaunchKernelGGL((k_axpy<double, double>), dim3(g), dim3(kVecBS), 0, st, (double *)y,
                           (const double *)x, n, a);
    else if (yprec == DMED_FP64)
        hipLaunchKernelGGL((k_axpy<double, float>), dim3(g), dim3(kVecBS), 0, st, (double *)y,
                           (const float *)x, n, a);
    else
        return fail(DMED_EINVAL, "axpy into FP32 not provided");
    DMED_HIP(hipGetLastError());
    return 0;
}

// ----------------------------------------------------------------- chain
namespace {

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) hipFree(p); }
};

ChainScalars *chain_scalars() { return reinterpret_cast<ChainScalars *>(ex().dscal); }
double *chain_alpha() { return ex().dscal + 16; }
double *chain_beta() { return ex().dscal + 16 + (kScalCap - 16) / 3; }
double *chain_aux() { return ex().dscal + 16 + 2 * ((kScalCap - 16) / 3); }
constexpr int kChainMax = (kScalCap - 16) / 3;

// Chain on three rotated work vectors bufs = {v_prev, v, w}; bufs[1] holds
// the (unnormalised) start vector.  chain_init normalises it through the
// device scale and zeroes v_prev; chain_steps launches `count` steps.  If
// psi != nullptr, psi += z[k] * v_k is accumulated before each step (second
// pass of the ground-state Lanczos), z read from chain_aux().
int chain_init(const dmed_sector *s, void *bufs[3], int prec, int nlanc_total, double thr,
               double *partials)
{
    hipStream_t st = ex().stream;
    const int64_t n = s->dim;
    const size_t esz = prec == DMED_FP32 ? 4 : 8;
    int np = 0;
    if (int rc = launch_zero(bufs[0], n * esz, st)) return rc;
    if (int rc = launch_dot(bufs[1], bufs[1], n, prec, partials, &np, st)) return rc;
    hipLaunchKernelGGL(k_chain_init, dim3(1), dim3(1024), 0, st, partials, np, chain_scalars(),
                       thr, nlanc_total);
    DMED_HIP(hipGetLastError());
    return 0;
}

// Lanczos steps are captured into a HIP graph when a run is long enough:
// three steps bring the (v_prev, v, w) rotation back to the start, so one
// captured three-step graph is replayed.  DMED_GRAPHS=0 turns this off.
constexpr int kGraphMinSteps = 30;

// DMED_TEST_CAPTURE_FAIL=q makes step q (0..2) of the captured three-step
// window fail before it is recorded: a test of the fallback to eager steps.
static int capture_fail_step()
{
    static const int q = [] {
        const char *e = std::getenv("DMED_TEST_CAPTURE_FAIL");
        return e ? std::atoi(e) : -1;
    }();
    return q;
}

static bool graphs_enabled()
{
    static const bool on = [] {
        const char *e = std::getenv("DMED_GRAPHS");
        return !(e && e[0] == '0');
    }();
    return on;
}

int chain_steps(const dmed_sector *s, void *bufs[3], int prec, int count, double *psi,
                double *partials, bool allow_graphs = true)
{
    hipStream_t st = ex().stream;
    const int64_t n = s->dim;
    ChainScalars *cs = chain_scalars();
    const int nph = hxv_nparts(s, prec);
    const unsigned g = vec_grid(n);
    // the kernels of one step; no memsets, no host synchronisation
    auto one = [&](void *vp, void *v, void *w) -> int {
        if (psi) {
            if (prec == DMED_FP32)
                hipLaunchKernelGGL(k_accum<float>, dim3(g), dim3(kVecBS), 0, st, psi,
                                   (const float *)v, n, cs, chain_aux());
            else
                hipLaunchKernelGGL(k_accum<double>, dim3(g), dim3(kVecBS), 0, st, psi,
                                   (const double *)v, n, cs, chain_aux());
        }
        if (int rc = launch_hxv(s, v, w, prec, kLanczos, cs, partials, nullptr, st))
            return rc;
        hipLaunchKernelGGL(k_chain_alpha, dim3(1), dim3(1024), 0, st, partials, nph, cs,
                           chain_alpha());
        if (prec == DMED_FP32)
            hipLaunchKernelGGL(k_chain_update<float>, dim3(g), dim3(kVecBS), 0, st, (float *)w,
                               (const float *)v, (const float *)vp, n, cs, partials);
        else
            hipLaunchKernelGGL(k_chain_update<double>, dim3(g), dim3(kVecBS), 0, st, (double *)w,
                               (const double *)v, (const double *)vp, n, cs, partials);
        hipLaunchKernelGGL(k_chain_beta, dim3(1), dim3(1024), 0, st, partials, (int)g, cs,
                           chain_beta());
        return hipGetLastError() == hipSuccess ? 0 : fail(DMED_EHIP, "chain step launch failed");
    };
    void *vp = bufs[0], *v = bufs[1], *w = bufs[2];
    if (((uintptr_t)vp | (uintptr_t)v | (uintptr_t)w | (uintptr_t)psi) % 16 != 0)
        return fail(DMED_EINVAL, "chain work vectors must be 16-byte aligned");
    auto rotate = [&]() {
        void *t = vp;
        vp = v;
        v = w;
        w = t;
    };
    int k = 0;
    if (count >= kGraphMinSteps && allow_graphs && graphs_enabled()) {
        // one eager step first (kernel attributes are set outside the capture)
        if (int rc = one(vp, v, w)) return rc;
        rotate();
        ++k;
        hipGraph_t gr = nullptr;
        hipGraphExec_t ge = nullptr;
        // nothing recorded during a capture has run: if the capture fails
        // part-way, the eager loop below resumes from these buffers
        void *const keep[3] = {vp, v, w};
        bool ok = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            int rc = 0;
            const int fail_at = capture_fail_step();
            for (int q = 0; q < 3 && !rc; ++q) {
                rc = q == fail_at ? fail(DMED_EHIP, "capture failure forced for a test")
                                  : one(vp, v, w);
                rotate();
            }
            // after three rotations the buffers are back where they started
            ok = hipStreamEndCapture(st, &gr) == hipSuccess && rc == 0 && gr != nullptr;
            ok = ok && hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0) == hipSuccess;
        }
        if (!ok) {
            vp = keep[0];
            v = keep[1];
            w = keep[2];
        }
        bool launched_fail = false;
        if (ok) {
            const int reps = (count - k) / 3;
            for (int r = 0; r < reps; ++r)
                if (hipGraphLaunch(ge, st) != hipSuccess) {
                    launched_fail = true;
                    break;
                }
            if (!launched_fail) k += 3 * reps;
            DMED_HIP(hipStreamSynchronize(st));
        }
        if (ge) hipGraphExecDestroy(ge);
        if (gr) hipGraphDestroy(gr);
        // a capture that could not be built leaves the chain where the eager
        // step left it and the loop below carries on; a failed replay cannot
        (void)hipGetLastError();
        if (!ok) {
            static std::atomic<bool> warned{false};
            if (!warned.exchange(true))
                std::fprintf(stderr, "dmed: Lanczos-step graph capture unavailable, "
                                     "launching the steps one by one\n");
        }
        if (launched_fail) return fail(DMED_EHIP, "chain graph launch failed");
    }
    for (; k < count; ++k) {
        if (int rc = one(vp, v, w)) return rc;
        rotate();
    }
    bufs[0] = vp;
    bufs[1] = v;
    bufs[2] = w;
    return 0;
}

}  // namespace

// Work vectors and reduction partials of one chain: FP64 (b64) and, for the
// FP32 tail of a mixed chain or an all-FP32 chain, b32.
struct ChainWork {
    void *b64[3] = {};
    void *b32[3] = {};
    double *partials = nullptr;
};

static int chain_run_ws(const dmed_sector *s, const void *v0, int v0prec, int nlanc, int prec,
                        double threshold, double *alpha, double *beta, int *nsteps, int head,
                        const ChainWork &ws, bool allow_graphs)
{
    hipStream_t st = ex().stream;
    if (nlanc <= 0) return fail(DMED_EINVAL, "nlanc must be positive");
    if (nlanc > kChainMax) return fail(DMED_EINVAL, "nlanc exceeds the chain buffer");
    if (prec != DMED_FP64 && prec != DMED_FP32 && prec != DMED_MIXED)
        return fail(DMED_EINVAL, "bad chain precision");
    if (head < 0) head = kMixedHead;
    // DMED_MIXED: the first `head` steps in FP64, the rest in FP32.  The low
    // moments of the spectral function, which fix G(iw) at all but the
    // highest resolution, come from the first steps; FP32 rounding there is
    // what dominates the error of an all-FP32 chain.
    const int nhead = prec == DMED_MIXED ? std::min(head, nlanc) : (prec == DMED_FP64 ? nlanc : 0);
    const int p0 = nhead > 0 ? DMED_FP64 : DMED_FP32;
    const int64_t n = s->dim;
    void *bufs[3];
    for (int k = 0; k < 3; ++k) bufs[k] = p0 == DMED_FP32 ? ws.b32[k] : ws.b64[k];
    if (int rc = launch_scale_copy(bufs[1], p0, v0, v0prec, n, 1.0, st)) return rc;
    if (int rc = launch_zero(chain_alpha(), kChainMax * sizeof(double), st)) return rc;
    if (int rc = launch_zero(chain_beta(), kChainMax * sizeof(double), st)) return rc;
    if (int rc = chain_init(s, bufs, p0, nlanc, threshold, ws.partials)) return rc;
    if (nhead > 0 && nhead < nlanc) {
        if (int rc = chain_steps(s, bufs, DMED_FP64, nhead, nullptr, ws.partials, allow_graphs))
            return rc;
        // the recurrence needs v and v_prev; w is overwritten by the next step
        void *lo[3] = {ws.b32[0], ws.b32[1], ws.b32[2]};
        if (int rc = launch_scale_copy(lo[0], DMED_FP32, bufs[0], DMED_FP64, n, 1.0, st)) return rc;
        if (int rc = launch_scale_copy(lo[1], DMED_FP32, bufs[1], DMED_FP64, n, 1.0, st)) return rc;
        if (int rc = chain_steps(s, lo, DMED_FP32, nlanc - nhead, nullptr, ws.partials, allow_graphs))
            return rc;
    } else {
        if (int rc = chain_steps(s, bufs, p0, nlanc, nullptr, ws.partials, allow_graphs)) return rc;
    }
    ChainScalars h;
    DMED_HIP(hipMemcpyAsync(alpha, chain_alpha(), nlanc * sizeof(double), hipMemcpyDeviceToHost, st));
    DMED_HIP(hipMemcpyAsync(beta, chain_beta(), nlanc * sizeof(double), hipMemcpyDeviceToHost, st));
    DMED_HIP(hipMemcpyAsync(&h, chain_scalars(), sizeof h, hipMemcpyDeviceToHost, st));
    DMED_HIP(hipStreamSynchronize(st));
    if (nsteps) *nsteps = h.step;
    return 0;
}

// partial sums a chain step of s may leave
static size_t chain_nparts(const dmed_sector *s)
{
    const int nparts = std::max(hxv_nparts(s, DMED_FP64), hxv_nparts(s, DMED_FP32));
    return (size_t)std::max(nparts, 2 * kVecGrid) + 16;
}

int chain_run(const dmed_sector *s, const void *v0, int v0prec, int nlanc, int prec,
              double threshold, double *alpha, double *beta, int *nsteps, int head)
{
    const size_t n1 = (size_t)std::max<int64_t>(s->dim, 1);
    ChainWork ws;
    if (int rc = ensure_partials(chain_nparts(s))) return rc;
    ws.partials = ex().partials;
    const bool any64 = prec != DMED_FP32, any32 = prec != DMED_FP64;
    for (int k = 0; k < 3; ++k) {
        if (any64)
            if (int rc = exec_buf(ex(), kSlotChain + k, n1 * 8, &ws.b64[k])) return rc;
        if (any32)
            if (int rc = exec_buf(ex(), prec == DMED_FP32 ? kSlotChain + k : kSlotChainLo + k,
                                  n1 * 4, &ws.b32[k]))
                return rc;
    }
    return chain_run_ws(s, v0, v0prec, nlanc, prec, threshold, alpha, beta, nsteps, head, ws, true);
}

// ---------------------------------------------- allocation-free chain
// The workspace of dmed_chain_noalloc: a file-scope device array (each GPU
// the library runs on holds its own copy until the process ends).  Host code
// never names it in a kernel argument or copy: its per-device address comes
// from hipGetSymbolAddress at set-up.  Every call on a device shares it, so
// calls that overlap on one GPU race.
constexpr size_t kNoAllocBytes = DMED_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
static std::mutex g_scratch_mu;
static std::map<int, unsigned char *> g_scratch;   // device -> address of scratch

int noalloc_setup(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch.count(device)) return 0;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
        return fail(DMED_EHIP, std::string("no-alloc workspace: hipGetSymbolAddress: ") +
                                   hipGetErrorString(e));
    g_scratch[device] = static_cast<unsigned char *>(ptr);
    return 0;
}

int chain_run_noalloc(const dmed_sector *s, const void *v0, int v0prec, int nlanc, int prec,
                      double threshold, double *alpha, double *beta, int *nsteps)
{
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        auto it = g_scratch.find(ctx().device);
        if (it != g_scratch.end()) base = it->second;
    }
    if (!base) return fail(DMED_ESTATE, "no-alloc workspace not set up (dmed_init)");
    const int64_t n = s->dim;
    if (n > DMED_NOALLOC_MAX_DIM)
        return fail(DMED_EINVAL, "dmed_chain_noalloc: sector dimension " + std::to_string(n) +
                                     " exceeds DMED_NOALLOC_MAX_DIM (" +
                                     std::to_string((long long)DMED_NOALLOC_MAX_DIM) + ")");
    const size_t n1 = (size_t)std::max<int64_t>(n, 1);
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    // partials, then three FP64 and three FP32 vectors
    const size_t pbytes = up(chain_nparts(s) * 8);
    const size_t need = pbytes + 3 * up(n1 * 8) + 3 * up(n1 * 4);
    if (need > kNoAllocBytes)
        return fail(DMED_EINVAL, "dmed_chain_noalloc: the sector's partial sums exceed the workspace");
    ChainWork ws;
    size_t off = 0;
    ws.partials = reinterpret_cast<double *>(base);
    off += pbytes;
    for (int k = 0; k < 3; ++k) {
        ws.b64[k] = base + off;
        off += up(n1 * 8);
    }
    for (int k = 0; k < 3; ++k) {
        ws.b32[k] = base + off;
        off += up(n1 * 4);
    }
    // no timing runs of the autotuner (they allocate) and no graph instantiation
    NoAllocScope guard;
    return chain_run_ws(s, v0, v0prec, nlanc, prec, threshold, alpha, beta, nsteps, -1, ws, false);
}

// ------------------------------------------------------ ground state
int eigh_sector_ex(const dmed_sector *s, int neig, int ncv, int maxit, double tol,
                   int dense_threshold, double degen, bool lowest_multiplet, int *nout,
                   double *evals, dmed_vec **evecs, double *resid, int gs_prec);
constexpr int64_t kTrlMaxDim = 25000000;   // above: plain Lanczos ground state

// <x, y> of two device FP64 vectors, returned to the host
int dot_host(const double *x, const double *y, int64_t n, double *out)
{
        int np = 0;
    if (int rc = launch_dot(x, y, n, DMED_FP64, ex().partials, &np, ex().stream)) return rc;
    double *d = ex().dscal + 9;
    if (int rc = launch_reduce_to(ex().partials, np, d, 0, ex().stream)) return rc;
    DMED_HIP(hipMemcpyAsync(out, d, sizeof(double), hipMemcpyDeviceToHost, ex().stream));
    DMED_HIP(hipStreamSynchronize(ex().stream));
    return 0;
}

namespace {

// lowest eigenpair of the symmetric tridiagonal (a[0..n), b[1..n)) via the
// dense host solver (n is at most a few hundred)
int tridiag_lowest(int n, const double *a, const double *b, double *theta, double *z)
{
    std::vector<double> T((size_t)n * n, 0.0), w(n);
    for (int i = 0; i < n; ++i) T[(size_t)i * n + i] = a[i];
    for (int i = 1; i < n; ++i) T[(size_t)i * n + i - 1] = T[(size_t)(i - 1) * n + i] = b[i];
    if (int rc = host_syev(n, T.data(), w.data())) return rc;
    *theta = w[0];
    for (int i = 0; i < n; ++i) z[i] = T[i];
    return 0;
}

// dense path: assemble H of a small sector on the host from the sector
// tables (ED_HAMILTONIAN_MATVEC.f90:185-196 dumps the same matrix) and
// diagonalise it
int dense_eigh(const dmed_sector *s, int neig, double *evals, dmed_vec **evecs, double *resid)
{
    const int64_t dup = s->up.dim, ddw = s->dw.dim, n = s->dim;
    std::vector<double> H((size_t)n * n, 0.0), w(n);
    for (int64_t id = 0; id < ddw; ++id)
        for (int64_t iu = 0; iu < dup; ++iu) {
            const int64_t i = iu + id * dup;
            H[(size_t)i * n + i] += s->up.h_diag[iu] + s->dw.h_diag[id] +
                                    s->h_W[s->up.h_imp[iu] * s->nimp + s->dw.h_imp[id]];
            for (int h = 0; h < s->up.nhop; ++h) {
                const int64_t j = s->up.h_hidx[(size_t)h * dup + iu] + id * dup;
                H[(size_t)j * n + i] += s->up.h_hval[(size_t)h * dup + iu];
            }
            for (int h = 0; h < s->dw.nhop; ++h) {
                const int64_t j = iu + (int64_t)s->dw.h_hidx[(size_t)h * ddw + id] * dup;
                H[(size_t)j * n + i] += s->dw.h_hval[(size_t)h * ddw + id];
            }
            // spin exchange and pair hopping: up pair hop b -> a with down
            // pair hop a -> b (Jx) or b -> a (Jp)
            if (s->jterms) {
                const int no = s->norb;
                for (int pa = 0; pa < no; ++pa)
                    for (int pb = 0; pb < no; ++pb) {
                        if (pa == pb) continue;
                        const int32_t wu = s->up.h_ipair[(size_t)(pa * no + pb) * dup + iu];
                        if (wu < 0) continue;
                        const double su = (wu & 1) ? -1.0 : 1.0;
                        const int32_t wx = s->dw.h_ipair[(size_t)(pb * no + pa) * ddw + id];
                        if (wx >= 0)
                            H[(size_t)((wu >> 1) + (int64_t)(wx >> 1) * dup) * n + i] +=
                                s->jx * su * ((wx & 1) ? -1.0 : 1.0);
                        const int32_t wp = s->dw.h_ipair[(size_t)(pa * no + pb) * ddw + id];
                        if (wp >= 0)
                            H[(size_t)((wu >> 1) + (int64_t)(wp >> 1) * dup) * n + i] +=
                                s->jp * su * ((wp & 1) ? -1.0 : 1.0);
                    }
            }
        }
    if (int rc = host_syev((int)n, H.data(), w.data())) return rc;
    for (int q = 0; q < neig; ++q) {
        evals[q] = w[q];
        if (resid) resid[q] = 0.0;
        if (evecs) {
            dmed_vec *v = new dmed_vec;
            v->sec = s;
            v->prec = DMED_FP64;
            v->dim = n;
            DMED_HIP(hipMalloc(&v->data, std::max<int64_t>(n, 1) * 8));
            if (int rc = copy_h2d(v->data, H.data() + (size_t)q * n, n * 8)) {
                hipFree(v->data);
                delete v;
                return rc;
            }
            evecs[q] = v;
        }
    }
    return 0;
}

// explicit residual ||H y - theta y|| for a device FP64 vector y
int residual(const dmed_sector *s, const double *y, double theta, double *work, double *out)
{
    hipStream_t st = ex().stream;
    const int64_t n = s->dim;
    if (int rc = launch_hxv(s, y, work, DMED_FP64, kPlain, nullptr, nullptr, nullptr, st))
        return rc;
    if (int rc = launch_axpy(work, DMED_FP64, y, DMED_FP64, n, -theta, st)) return rc;
    int np = 0;
    if (int rc = launch_dot(work, work, n, DMED_FP64, ex().partials, &np, st)) return rc;
    double *d = ex().dscal + 8;
    if (int rc = launch_reduce_to(ex().partials, np, d, 1, st)) return rc;
    DMED_HIP(hipMemcpyAsync(out, d, sizeof(double), hipMemcpyDeviceToHost, st));
    DMED_HIP(hipStreamSynchronize(st));
    return 0;
}


// Plain Lanczos ground state (one eigenpair), memory 4 vectors:
//   pass 1: recurrence until the lowest Ritz value's residual estimate
//           |beta_k z_k| <= tol*max(1,|theta|), checked every ncheck steps;
//   pass 2: identical recurrence, psi = sum_k z_k v_k (lanczos_plain_c,
//           .repo/PLAIN_LANCZOS.f90:367-384);
//   refine: explicit residual; if above tol, restart from psi.
int plain_gs(const dmed_sector *s, int maxit, double tol, double *eval, dmed_vec **evec,
             double *resid, int prec)
{
    hipStream_t st = ex().stream;
    if (prec != DMED_FP64 && prec != DMED_MIXED)
        return fail(DMED_EINVAL, "ground-state precision must be DMED_FP64 or DMED_MIXED");
    const int64_t n = s->dim;
    const size_t n1 = (size_t)std::max<int64_t>(n, 1);
    void *b64[3], *b32[3] = {nullptr, nullptr, nullptr};
    for (int k = 0; k < 3; ++k)
        if (int rc = exec_buf(ex(), kSlotChain + k, n1 * 8, &b64[k])) return rc;
    if (prec == DMED_MIXED)
        for (int k = 0; k < 3; ++k)
            if (int rc = exec_buf(ex(), kSlotChainLo + k, n1 * 4, &b32[k])) return rc;
    const int nparts = std::max(hxv_nparts(s, DMED_FP64), hxv_nparts(s, DMED_FP32));
    if (int rc = ensure_partials(std::max(nparts, 2 * kVecGrid) + 16)) return rc;
    double *psi = nullptr;
    DMED_HIP(hipMalloc(&psi, n1 * 8));
    dmed_vec *out = new dmed_vec{s, DMED_FP64, n, psi};
    auto cleanup = [&](int rc) {
        if (rc) { hipFree(out->data); delete out; }
        return rc;
    };
    if (int rc = launch_fill_random(psi, n, DMED_FP64, 1234567ull, st)) return cleanup(rc);
    const int ncheck = 10;
    int used = 0;
    double theta = 0.0, res = INFINITY;
    std::vector<double> a(kChainMax), b(kChainMax), z(kChainMax);
    // DMED_MIXED: restarts run the recurrence on FP32 vectors (FP64 scalars
    // and reductions) until the FP64 residual of the accumulated vector is
    // below kMixedSwitch.  The later restarts, which take the residual down to
    // tol, start from that vector psi0 with kRefineHead FP64 steps and go on
    // in FP32: the first step gives alpha_0 = <psi0|H|psi0> and
    // beta_1 = ||H psi0 - alpha_0 psi0|| in FP64, and every later Krylov
    // vector enters the new psi with a weight of the order of the residual
    // over the gap, so its FP32 rounding lands far below tol.  The Rayleigh
    // quotient and the residual are FP64.
    constexpr double kMixedSwitch = 1e-5;
    constexpr int kRefineHead = 2;
    enum { kFp32Stage, kRefineStage, kFp64Stage };
    int stage = prec == DMED_MIXED ? kFp32Stage : kFp64Stage;
    // (v_prev, v) of an FP64 head into the FP32 buffers; w is rewritten by
    // the next step
    auto to_fp32 = [&](void *cb[3]) -> int {
        if (int rc = launch_scale_copy(b32[0], DMED_FP32, cb[0], DMED_FP64, n, 1.0, st)) return rc;
        if (int rc = launch_scale_copy(b32[1], DMED_FP32, cb[1], DMED_FP64, n, 1.0, st)) return rc;
        cb[0] = b32[0];
        cb[1] = b32[1];
        cb[2] = b32[2];
        return 0;
    };
    for (int restart = 0; restart < 12 && used < maxit; ++restart) {
        const bool refine = stage == kRefineStage;
        const int p0 = stage == kFp32Stage ? DMED_FP32 : DMED_FP64;
        void **bb = p0 == DMED_FP32 ? b32 : b64;
        // the FP32 stage only has to hand the refinement a vector with a
        // residual near 1e-6: tighter FP32 restarts cost more steps than the
        // refinement saves
        const double stol = stage == kFp32Stage ? std::max(tol, 1e-6) : tol;
        // pass 1, in chunks of ncheck steps with a host convergence check
        if (int rc = launch_scale_copy(bb[1], p0, psi, DMED_FP64, n, 1.0, st))
            return cleanup(rc);
        DMED_HIP(hipMemsetAsync(chain_alpha(), 0, kChainMax * sizeof(double), st));
        DMED_HIP(hipMemsetAsync(chain_beta(), 0, kChainMax * sizeof(double), st));
        const int budget = std::min(kChainMax - 1, maxit - used);
        const int nhead = refine ? std::min(kRefineHead, budget) : 0;
        void *bufs[3] = {bb[0], bb[1], bb[2]};
        if (int rc = chain_init(s, bufs, p0, budget, 0.0, ex().partials)) return cleanup(rc);
        int steps = 0;
        if (nhead > 0) {
            if (int rc = chain_steps(s, bufs, DMED_FP64, nhead, nullptr, ex().partials))
                return cleanup(rc);
            if (int rc = to_fp32(bufs)) return cleanup(rc);
            steps = nhead;
        }
        const int cur = refine ? DMED_FP32 : p0;
        ChainScalars cs{};
        for (;;) {
            const int chunk = std::min(ncheck, budget - steps);
            if (chunk > 0)
                if (int rc = chain_steps(s, bufs, cur, chunk, nullptr, ex().partials))
                    return cleanup(rc);
            steps += chunk;
            DMED_HIP(hipMemcpyAsync(a.data(), chain_alpha(), steps * 8, hipMemcpyDeviceToHost, st));
            DMED_HIP(hipMemcpyAsync(b.data(), chain_beta(), (steps + 1) * 8, hipMemcpyDeviceToHost, st));
            DMED_HIP(hipMemcpyAsync(&cs, chain_scalars(), sizeof cs, hipMemcpyDeviceToHost, st));
            DMED_HIP(hipStreamSynchronize(st));
            const int m = std::min(steps, cs.step);
            if (m < 1) return cleanup(fail(DMED_ENORM, "zero start vector in the ground-state Lanczos"));
            if (int rc = tridiag_lowest(m, a.data(), b.data(), &theta, z.data())) return cleanup(rc);
            const double est = fabs(cs.b * z[m - 1]);
            if (est <= 0.1 * stol * std::max(1.0, fabs(theta)) || cs.done || steps >= budget) {
                steps = m;
                break;
            }
        }
        used += steps;
        // pass 2: rerun the recurrence from the same start, in the same
        // precisions, accumulating psi (FP64)
        DMED_HIP(hipMemcpyAsync(chain_aux(), z.data(), steps * 8, hipMemcpyHostToDevice, st));
        if (int rc = launch_scale_copy(bb[1], p0, psi, DMED_FP64, n, 1.0, st))
            return cleanup(rc);
        DMED_HIP(hipMemsetAsync(psi, 0, n * 8, st));
        void *bufs2[3] = {bb[0], bb[1], bb[2]};
        if (int rc = chain_init(s, bufs2, p0, steps, 0.0, ex().partials)) return cleanup(rc);
        const int h2 = std::min(nhead, steps);
        if (h2 > 0) {
            if (int rc = chain_steps(s, bufs2, DMED_FP64, h2, psi, ex().partials)) return cleanup(rc);
            if (int rc = to_fp32(bufs2)) return cleanup(rc);
        }
        if (steps > h2)
            if (int rc = chain_steps(s, bufs2, cur, steps - h2, psi, ex().partials)) return cleanup(rc);
        used += steps;
        // normalise, FP64 Rayleigh quotient and explicit FP64 residual
        double nn = 0.0;
        if (int rc = dot_host(psi, psi, n, &nn)) return cleanup(rc);
        if (int rc = launch_scale_copy(psi, DMED_FP64, psi, DMED_FP64, n, 1.0 / sqrt(nn), st))
            return cleanup(rc);
        if (int rc = launch_hxv(s, psi, b64[2], DMED_FP64, kPlain, nullptr, nullptr, nullptr, st))
            return cleanup(rc);
        double rq = 0.0;
        if (int rc = dot_host(psi, (const double *)b64[2], n, &rq)) return cleanup(rc);
        theta = rq;
        if (int rc = residual(s, psi, theta, (double *)b64[2], &res)) return cleanup(rc);
        used += 2;
        if (res <= tol * std::max(1.0, fabs(theta))) break;
        if (stage == kFp32Stage && (res <= kMixedSwitch * std::max(1.0, fabs(theta)) || restart >= 3))
            stage = kRefineStage;
    }
    *eval = theta;
    if (resid) *resid = res;
    if (evec) *evec = out;
    else cleanup(1);
    return res <= tol * std::max(1.0, fabs(theta)) ? 0 : DMED_ENOCONV;
}

// Thick-restart Lanczos with full reorthogonalisation in a device basis, for
// the neig lowest eigenpairs of H restricted to the orthogonal complement of
// nl locked orthonormal vectors Y (Wu & Simon, SIAM J. Matrix Anal. Appl. 22
// (2000) 602).  Each expansion step after the first of a restart cycle takes
// the three-term Lanczos step, then one classical Gram-Schmidt pass against
// Y and the basis, and a second pass when the first cancelled more than half
// of ||w||^2; the first step of a cycle takes the two passes directly.  Ritz
// vectors are written to X (neig contiguous device vectors).
int trl_run(const dmed_sector *s, int neig, int m, int maxit, double tol, const double *Y,
            int nl, uint64_t seed, double *theta_out, double *X, double *resid_out, int *used_out)
{
    hipStream_t st = ex().stream;
    const int64_t n = s->dim;
    const int64_t avail = n - nl;                 // dimension of the complement
    if (avail < neig) return fail(DMED_EINVAL, "complement smaller than neig");
    m = std::max(m, neig + 2);
    m = std::min(m, 64);                          // k_rotate_basis keeps m*128 doubles in LDS
    m = (int)std::min<int64_t>(m, avail);
    const bool full = (int64_t)m == avail;
    double *V = nullptr, *w = nullptr, *dS = nullptr;
    if (int rc = exec_buf(ex(), kSlotTrlBasis, (size_t)(m + 1) * n * 8, (void **)&V)) return rc;
    if (int rc = exec_buf(ex(), kSlotTrlW, n * 8, (void **)&w)) return rc;
    // dS: Ritz rotation (m*m), and per expansion step j the accumulated
    // coefficients (j+1) and the new norm, in rows of m+2 (kept on the
    // device until the restart); then the scalars of the Lanczos-mode H*v
    // (scale 1: y = Hx, with the partials of <x, y>)
    const size_t ldc = (size_t)m + 2;
    if (nl + m > kGsMaxRows) return fail(DMED_EINVAL, "too many projection rows");
    const size_t csoff = (size_t)m * m + (size_t)m * ldc + 8;
    if (int rc = exec_buf(ex(), kSlotTrlSmall, csoff * 8 + sizeof(ChainScalars), (void **)&dS))
        return rc;
    double *dcoef = dS + (size_t)m * m;
    ChainScalars *tcs = reinterpret_cast<ChainScalars *>(dS + csoff);
    {
        ChainScalars h{};
        h.sx = 1.0;
        if (int rc = copy_h2d(tcs, &h, sizeof h)) return rc;
    }
    // partial sums of one orthogonalisation: rows of the first pass
    // (nl + m + 1, the last one ||w||^2 before it), ||w||^2 after it, rows
    // of the second pass, ||w||^2 after that; then the <v_j, H v_j> partials
    // of the H*v
    const unsigned g = (unsigned)std::min<unsigned>(vec_grid(n), kGsGrid);
    const int nr = nl + m;
    const int nph = hxv_nparts(s, DMED_FP64);
    if (int rc = ensure_partials((size_t)(2 * nr + 3) * g + (size_t)nph + 16)) return rc;
    double *P1 = ex().partials, *Pn1 = P1 + (size_t)(nr + 1) * g, *P2 = Pn1 + g, *Pn2 = P2 + (size_t)nr * g;
    double *Pa = Pn2 + g;
    // classical Gram-Schmidt of w against Y[0..nl) and V[0..jmax), repeated
    // once when the first pass cancelled more than half of ||w||^2 (Daniel-
    // Gragg-Kaufman-Stewart criterion, decided on the device); the V
    // coefficients of both passes are stored at cacc when given (first:
    // overwrite, else add to what is there).  Four launches, no host sync;
    // normalize_into then needs no reduction launch.
    bool noproj = false;
    const double *n0_last = nullptr;
    auto orth = [&](int jmax, double *cacc, int first) -> int {
        const int rows = nl + jmax;
        noproj = rows <= 0;
        if (noproj) {
            // only ||w||^2: the w row of k_proj_dot, into Pn1
            hipLaunchKernelGGL(k_proj_dot, dim3(g, 1), dim3(kVecBS), 0, st, Y, 0, V, 0, w, n, Pn1,
                               0, (const double *)nullptr, (const double *)nullptr, 1, 1);
            DMED_HIP(hipGetLastError());
            return 0;
        }
        const double *n0 = P1 + (size_t)rows * g;   // the w row
        const int rpb = proj_rpb(n);
        hipLaunchKernelGGL(k_proj_dot, dim3(g, proj_groups(rows + 1, rpb)), dim3(kVecBS), 0, st, Y,
                           nl, V, jmax, w, n, P1, 0, (const double *)nullptr,
                           (const double *)nullptr, rows + 1, rpb);
        hipLaunchKernelGGL(k_proj_axpy, dim3(g), dim3(kVecBS), 0, st, Y, nl, V, jmax, w, n, P1, Pn1,
                           cacc, first, 0, (const double *)nullptr, (const double *)nullptr);
        hipLaunchKernelGGL(k_proj_dot, dim3(g, proj_groups(rows, rpb)), dim3(kVecBS), 0, st, Y, nl,
                           V, jmax, w, n, P2, 1, n0, Pn1, rows, rpb);
        hipLaunchKernelGGL(k_proj_axpy, dim3(g), dim3(kVecBS), 0, st, Y, nl, V, jmax, w, n, P2, Pn2,
                           cacc, 0, 1, n0, Pn1);
        DMED_HIP(hipGetLastError());
        n0_last = n0;
        return 0;
    };
    // dst = w/||w|| with the norm from the partials orth left; stored at nrm
    auto normalize_into = [&](double *dst, double *nrm) -> int {
        hipLaunchKernelGGL(k_normalize, dim3(g), dim3(kVecBS), 0, st, dst, w, n,
                           noproj ? Pn1 : n0_last, Pn1, Pn2, noproj ? 1 : 0, nrm);
        DMED_HIP(hipGetLastError());
        return 0;
    };
    std::vector<double> T((size_t)m * m, 0.0), S, theta(m), hc((size_t)m * ldc);
    // start vector in the complement of Y
    if (int rc = launch_fill_random(w, n, DMED_FP64, seed, st)) return rc;
    if (int rc = orth(0, nullptr, 1)) return rc;
    if (int rc = normalize_into(V, dcoef + m + 1)) return rc;
    int k = 0, used = 0;
    double beta_last = 0.0;
    bool converged = false;
    for (;;) {
        // expansion k..m-1 entirely on the stream; one copy of the
        // coefficients per restart
        // expansion step j > k: H v_j has, in exact arithmetic, only the
        // components alpha_j v_j and beta_{j-1} v_{j-1} in the basis; they are
        // taken off first (k_trl_local), so the full pass that follows sees no
        // large cancellation and its second pass seldom runs.  Step k (the
        // arrow of the kept Ritz vectors, or the first step) goes straight to
        // the full passes.
        for (int j = k; j < m; ++j) {
            double *vj = V + (size_t)j * n;
            const bool local = j > k;
            if (int rc = launch_hxv(s, vj, w, DMED_FP64, local ? kLanczos : kPlain,
                                    local ? tcs : nullptr, local ? Pa : nullptr, nullptr, st))
                return rc;
            ++used;
            double *cj = dcoef + (size_t)j * ldc;
            if (local) {
                hipLaunchKernelGGL(k_trl_local, dim3(g), dim3(kVecBS), 0, st, w, vj, vj - n, n, Pa,
                                   nph, dcoef + (size_t)(j - 1) * ldc + m + 1, cj, j);
                DMED_HIP(hipGetLastError());
            }
            if (int rc = orth(j + 1, cj, local ? 0 : 1)) return rc;
            if (int rc = normalize_into(V + (size_t)(j + 1) * n, cj + m + 1)) return rc;
        }
        DMED_HIP(hipMemcpyAsync(hc.data() + (size_t)k * ldc, dcoef + (size_t)k * ldc,
                                (size_t)(m - k) * ldc * 8, hipMemcpyDeviceToHost, st));
        DMED_HIP(hipStreamSynchronize(st));
        for (int j = k; j < m; ++j) {
            const double *c1 = hc.data() + (size_t)j * ldc;
            for (int q = 0; q <= j; ++q) {
                const bool arrow = (j == k) && (q < k);
                if (arrow || q >= j - 1) T[(size_t)q * m + j] = T[(size_t)j * m + q] = c1[q];
            }
            const double bn = c1[m + 1];
            if (j + 1 < m) T[(size_t)(j + 1) * m + j] = T[(size_t)j * m + j + 1] = bn;
            beta_last = bn;
        }
        S = T;
        if (int rc = host_syev(m, S.data(), theta.data())) return rc;
        bool ok = true;
        for (int q = 0; q < neig; ++q) {
            const double rn = full ? 0.0 : fabs(beta_last * S[(size_t)q * m + (m - 1)]);
            if (rn > 0.1 * tol * std::max(1.0, fabs(theta[q]))) ok = false;
        }
        if (ok || used >= maxit || full) {
            converged = ok || full;
            DMED_HIP(hipMemcpyAsync(dS, S.data(), (size_t)m * neig * 8, hipMemcpyHostToDevice, st));
            const unsigned gb = (unsigned)((n + 127) / 128);
            hipLaunchKernelGGL(k_rotate_basis, dim3(gb), dim3(128), (size_t)m * 128 * 8, st, V, n, m,
                               neig, dS);
            DMED_HIP(hipGetLastError());
            break;
        }
        // thick restart: keep kk Ritz vectors and the residual direction
        const int kk = std::min(m - 2, neig + std::max(1, (m - neig) / 2));
        DMED_HIP(hipMemcpyAsync(dS, S.data(), (size_t)m * kk * 8, hipMemcpyHostToDevice, st));
        const unsigned gb = (unsigned)((n + 127) / 128);
        hipLaunchKernelGGL(k_rotate_basis, dim3(gb), dim3(128), (size_t)m * 128 * 8, st, V, n, m, kk,
                           dS);
        DMED_HIP(hipGetLastError());
        DMED_HIP(hipMemcpyAsync(V + (size_t)kk * n, V + (size_t)m * n, n * 8,
                                hipMemcpyDeviceToDevice, st));
        std::fill(T.begin(), T.end(), 0.0);
        for (int q = 0; q < kk; ++q) {
            T[(size_t)q * m + q] = theta[q];
            const double arrow = beta_last * S[(size_t)q * m + (m - 1)];
            T[(size_t)q * m + kk] = T[(size_t)kk * m + q] = arrow;
        }
        k = kk;
    }
    for (int q = 0; q < neig; ++q) {
        theta_out[q] = theta[q];
        double r = 0.0;
        if (int rc = residual(s, V + (size_t)q * n, theta[q], w, &r)) return rc;
        resid_out[q] = r;
        if (r > tol * std::max(1.0, fabs(theta[q]))) converged = false;
        DMED_HIP(hipMemcpyAsync(X + (size_t)q * n, V + (size_t)q * n, n * 8, hipMemcpyDeviceToDevice, st));
    }
    DMED_HIP(hipStreamSynchronize(st));
    if (used_out) *used_out += used;
    return converged ? 0 : DMED_ENOCONV;
}

// Lowest neig eigenpairs with complete degenerate multiplets.  A
// single-vector Krylov space holds one direction per eigenspace, so after a
// converged run the search is repeated on the orthogonal complement of the
// kept vectors; a round that brings no new level below the current top ends
// the search.  lowest_multiplet: return only the lowest level and its copies
// within degen (absolute), at most neig of them.
int trl_deflated(const dmed_sector *s, int neig, int m, int maxit, double tol, double degen,
                 bool lowest_multiplet, int *nout, double *evals, dmed_vec **evecs, double *resid)
{
    hipStream_t st = ex().stream;
    const int64_t n = s->dim;
    double *Y = nullptr, *X = nullptr;
    if (int rc = exec_buf(ex(), kSlotDeflY, (size_t)neig * n * 8, (void **)&Y)) return rc;
    const int want = lowest_multiplet ? 1 : neig;
    if (int rc = exec_buf(ex(), kSlotDeflX, (size_t)want * n * 8, (void **)&X)) return rc;
    struct Kept { double e, r; int slot; };
    std::vector<Kept> kept;
    std::vector<double> th(want), rr(want);
    int used = 0, rcfinal = 0;
    constexpr double confirm_tol = 1e-6;
    for (int round = 0; round <= neig; ++round) {
        const int nl = (int)kept.size();
        if (n - nl < want) break;
        // rounds after the first only have to show that the complement holds
        // nothing below the kept levels: a loose solve settles that unless a
        // Ritz value, less its residual bound, comes within reach
        const double tol_r = (nl > 0 && confirm_tol > tol) ? confirm_tol : tol;
        int rc = trl_run(s, want, m, std::max(maxit - used, 4 * m), tol_r, Y, nl,
                         1234567ull + 7919ull * round, th.data(), X, rr.data(), &used);
        if (tol_r > tol && (rc == 0 || rc == DMED_ENOCONV)) {
            bool reach = false;
            for (int q = 0; q < want; ++q) {
                const double lo = th[q] - rr[q];
                if (lowest_multiplet ? (lo <= kept[0].e + degen)
                                     : ((int)kept.size() < neig || lo < kept.back().e))
                    reach = true;
            }
            if (!reach) break;
            rc = trl_run(s, want, m, std::max(maxit - used, 4 * m), tol, Y, nl,
                         1234567ull + 7919ull * round, th.data(), X, rr.data(), &used);
        }
        if (rc == DMED_ENOCONV) rcfinal = rc;
        else if (rc) return rc;
        bool added = false;
        for (int q = 0; q < want; ++q) {
            if (lowest_multiplet) {
                if (!kept.empty() && !(th[q] <= kept[0].e + degen)) continue;
                if ((int)kept.size() >= neig) continue;
            } else if ((int)kept.size() >= neig) {
                if (!(th[q] < kept.back().e)) continue;   // nothing new below the top
                // evict the highest kept level; its slot is reused
                const int slot = kept.back().slot;
                kept.pop_back();
                kept.push_back({th[q], rr[q], slot});
                DMED_HIP(hipMemcpyAsync(Y + (size_t)slot * n, X + (size_t)q * n, n * 8,
                                        hipMemcpyDeviceToDevice, st));
                std::sort(kept.begin(), kept.end(), [](const Kept &x, const Kept &y) { return x.e < y.e; });
                added = true;
                continue;
            }
            // free slot
            std::vector<bool> busy(neig, false);
            for (auto &k : kept) busy[k.slot] = true;
            int slot = 0;
            while (busy[slot]) ++slot;
            DMED_HIP(hipMemcpyAsync(Y + (size_t)slot * n, X + (size_t)q * n, n * 8,
                                    hipMemcpyDeviceToDevice, st));
            kept.push_back({th[q], rr[q], slot});
            std::sort(kept.begin(), kept.end(), [](const Kept &x, const Kept &y) { return x.e < y.e; });
            added = true;
        }
        if (!added || used >= maxit) break;
        if (lowest_multiplet && (int)kept.size() >= neig) break;
    }
    // evicted vectors may have left Y non-compact: the kept slots are what
    // is returned, in ascending energy
    *nout = (int)kept.size();
    for (size_t q = 0; q < kept.size(); ++q) {
        evals[q] = kept[q].e;
        if (resid) resid[q] = kept[q].r;
        if (evecs) {
            dmed_vec *v = new dmed_vec{s, DMED_FP64, n, nullptr};
            DMED_HIP(hipMalloc(&v->data, n * 8));
            DMED_HIP(hipMemcpyAsync(v->data, Y + (size_t)kept[q].slot * n, n * 8,
                                    hipMemcpyDeviceToDevice, st));
            evecs[q] = v;
        }
    }
    DMED_HIP(hipStreamSynchronize(st));
    return rcfinal;
}

}  // namespace

int eigh_sector(const dmed_sector *s, int neig, int ncv, int maxit, double tol,
                int dense_threshold, double *evals, dmed_vec **evecs, double *resid)
{
    int nout = 0;
    const int rc = eigh_sector_ex(s, neig, ncv, maxit, tol, dense_threshold, 0.0, false, &nout,
                                  evals, evecs, resid, DMED_FP64);
    return rc;
}

int ground_state(const dmed_sector *s, int prec, int maxit, double tol, double *eval,
                 dmed_vec **evec, double *resid)
{
    if (s->dim < 1) return fail(DMED_EINVAL, "empty sector");
    return plain_gs(s, maxit, tol, eval, evec, resid, prec);
}

int eigh_sector_ex(const dmed_sector *s, int neig, int ncv, int maxit, double tol,
                   int dense_threshold, double degen, bool lowest_multiplet, int *nout,
                   double *evals, dmed_vec **evecs, double *resid, int gs_prec)
{
    const int64_t n = s->dim;
    if (neig < 1 || neig > n) return fail(DMED_EINVAL, "neig out of range");
    if (n <= dense_threshold || neig == n) {
        if (!lowest_multiplet) {
            *nout = neig;
            return dense_eigh(s, neig, evals, evecs, resid);
        }
        // lowest level and its copies among the neig lowest
        std::vector<double> ev(neig), rs(neig);
        std::vector<dmed_vec *> vv(neig, nullptr);
        if (int rc = dense_eigh(s, neig, ev.data(), vv.data(), rs.data())) return rc;
        int cnt = 0;
        for (int q = 0; q < neig; ++q) {
            if (ev[q] <= ev[0] + degen) {
                evals[cnt] = ev[q];
                if (resid) resid[cnt] = rs[q];
                if (evecs) evecs[cnt] = vv[q];
                else dmed_vec_free(vv[q]);
                ++cnt;
            } else {
                dmed_vec_free(vv[q]);
            }
        }
        *nout = cnt;
        return 0;
    }
    if (int rc = ensure_partials(std::max(hxv_nparts(s, DMED_FP64), kVecGrid) + 16)) return rc;
    // stored-basis solver with multiplet completion while it fits comfortably;
    // past that the 4-vector plain Lanczos (ground state only, no in-sector
    // degeneracy search)
    int m = ncv > 0 ? ncv : std::max(2 * neig + 10, 20);
    size_t freeb = 0, totalb = 0;
    hipMemGetInfo(&freeb, &totalb);
    const double need = (double)(m + 2 + 2 * neig) * n * 8.0;
    if ((neig == 1 || lowest_multiplet) && (need > 0.5 * (double)freeb || n > kTrlMaxDim)) {
        *nout = 1;
        return plain_gs(s, maxit, tol, evals, evecs ? &evecs[0] : nullptr, resid, gs_prec);
    }
    if (need > 0.8 * (double)freeb) return fail(DMED_ENOMEM, "Krylov basis does not fit");
    return trl_deflated(s, neig, m, maxit, tol, degen, lowest_multiplet, nout, evals, evecs, resid);
}

}  // namespace dmed
