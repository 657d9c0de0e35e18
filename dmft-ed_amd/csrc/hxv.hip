// hxv.hip -- on-the-fly sector H*v for gfx950.
//
// The reference (build_hxv.f90 with HxVimp/HxVint/HxVbath/HxVimp_bath, called
// from directMatVec_cc, ED_HAMILTONIAN_MATVEC.f90:289-351) decodes every row
// with bdecomp, applies c/cdg bit loops and finds each hop target by binary
// search in the sector map.  Here the state vector is the D_up x D_dw
// column-major matrix X of the reference order (i = r_up + r_dw*D_up) and
//
//   Y(iu,id) = [d_up(iu) + d_dw(id) + W(imp_up(iu), imp_dw(id))] X(iu,id)
//            + sum_h tu_h(iu) X(ju_h(iu), id)          (up hops, same column)
//            + sum_h td_h(id) X(iu, jd_h(id))          (down hops, same row)
//
// with the per-spin hop tables built on the device (sector.hip).  One
// workgroup owns `cb` whole columns: it stages them in LDS with coalesced
// loads, gathers the up hops from LDS, and reads the down-hop neighbour
// columns straight from HBM with lanes running along the contiguous up
// index (coalesced sweeps).  In Lanczos mode the stored vector x carries the
// scale sx (v = sx*x): the kernel writes y = Hv = sx*H*x and the block
// partials of <x, y>; the -b*v_prev and -alpha*v terms of the step are left to
// the streaming update kernel, so no third vector is read here.
#include "dmed_internal.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace dmed {

template <typename T>
struct HxvArgs {
    const T *__restrict__ x;
    T *__restrict__ y;
    int32_t dup, ddw, nhu, nhd, nimp, cb;
    float inv_dup;
    const int32_t *__restrict__ hu_idx;   // [nhu][dup]
    const T *__restrict__ hu_val;
    const int32_t *__restrict__ hd_idx;   // [nhd][ddw]
    const T *__restrict__ hd_val;
    const double *__restrict__ du;
    const double *__restrict__ dd;
    const uint8_t *__restrict__ impu;
    const uint8_t *__restrict__ impd;
    const double *__restrict__ W;         // [nimp][nimp]
    const ChainScalars *cs;
    double *__restrict__ partials;
    // spin exchange / pair hopping (jterms): impurity pair-hop tables
    const int32_t *__restrict__ ju;       // [norb*norb][dup]
    const int32_t *__restrict__ jd;       // [norb*norb][ddw]
    int32_t jterms, norb;
    double jx, jp;
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// sum over the workgroup; result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    if (w == 0) {
        v = (lane < nw) ? sh[lane] : 0.0;
        v = wave_sum(v);
    }
    return v;
}

template <typename T, int MODE, bool LDS>
__global__ void __launch_bounds__(1024) k_hxv(const HxvArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red[16];
    const int64_t c0 = (int64_t)blockIdx.x * a.cb;
    const int ncol = (int)min((int64_t)a.cb, (int64_t)a.ddw - c0);
    const int dup = a.dup;
    const int nel = ncol * dup;
    T *col = reinterpret_cast<T *>(smem);
    const size_t colbytes = LDS ? (((size_t)a.cb * dup * sizeof(T) + 15) & ~(size_t)15) : 0;
    double *wt = reinterpret_cast<double *>(smem + colbytes);
    const T *__restrict__ xb = a.x + c0 * dup;

    double sx = 1.0;
    if (MODE == kLanczos) {
        if (a.cs->done) return;
        sx = a.cs->sx;
    }
    if (LDS)
        for (int e = threadIdx.x; e < nel; e += blockDim.x) col[e] = xb[e];
    for (int e = threadIdx.x; e < ncol * a.nimp; e += blockDim.x) {
        const int cc = e / a.nimp, iu = e - cc * a.nimp;
        wt[e] = a.dd[c0 + cc] + a.W[iu * a.nimp + a.impd[c0 + cc]];
    }
    __syncthreads();

    double dot = 0.0;
    for (int e = threadIdx.x; e < nel; e += blockDim.x) {
        int cc = 0, i = e;
        if (ncol > 1) {
            cc = (int)((float)e * a.inv_dup);
            i = e - cc * dup;
            if (i < 0) { --cc; i += dup; }
            else if (i >= dup) { ++cc; i -= dup; }
        }
        const T *xc = LDS ? (col + cc * dup) : (xb + (int64_t)cc * dup);
        const T xo = xc[i];
        T acc = (T)(a.du[i] + wt[cc * a.nimp + a.impu[i]]) * xo;
        // hops four at a time: every index and value load of a group is issued
        // before its gathers, and the gathers before the adds, which keep the
        // table order
        int h = 0;
        for (; h + 4 <= a.nhu; h += 4) {
            int32_t id[4];
            T v[4], g[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                id[q] = a.hu_idx[(int64_t)(h + q) * dup + i];
                v[q] = a.hu_val[(int64_t)(h + q) * dup + i];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) g[q] = xc[id[q]];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc += v[q] * g[q];
        }
        for (; h < a.nhu; ++h)
            acc += a.hu_val[(int64_t)h * dup + i] * xc[a.hu_idx[(int64_t)h * dup + i]];
        const int64_t c = c0 + cc;
        h = 0;
        for (; h + 4 <= a.nhd; h += 4) {
            int64_t j[4];
            T v[4], g[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                j[q] = a.hd_idx[(int64_t)(h + q) * a.ddw + c];
                v[q] = a.hd_val[(int64_t)(h + q) * a.ddw + c];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) g[q] = a.x[j[q] * dup + i];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc += v[q] * g[q];
        }
        for (; h < a.nhd; ++h) {
            const int64_t jj = a.hd_idx[(int64_t)h * a.ddw + c];
            acc += a.hd_val[(int64_t)h * a.ddw + c] * a.x[jj * dup + i];
        }
        // spin exchange Jx c^+_a,up c_b,up c^+_b,dw c_a,dw and pair hopping
        // Jp c^+_a,up c_b,up c^+_a,dw c_b,dw (HxVint.f90:48-97): an up hop
        // b -> a on the row with a down hop a -> b (Jx) or b -> a (Jp) on
        // the column
        if (a.jterms) {
            const int no = a.norb;
            for (int pa = 0; pa < no; ++pa)
                for (int pb = 0; pb < no; ++pb) {
                    if (pa == pb) continue;
                    const int32_t wu = a.ju[(int64_t)(pa * no + pb) * dup + i];
                    if (wu < 0) continue;
                    const int64_t iu2 = wu >> 1;
                    const double su = (wu & 1) ? -1.0 : 1.0;
                    const int32_t wx = a.jd[(int64_t)(pb * no + pa) * a.ddw + c];
                    if (wx >= 0)
                        acc += (T)(a.jx * su * ((wx & 1) ? -1.0 : 1.0)) *
                               a.x[(int64_t)(wx >> 1) * dup + iu2];
                    const int32_t wp = a.jd[(int64_t)(pa * no + pb) * a.ddw + c];
                    if (wp >= 0)
                        acc += (T)(a.jp * su * ((wp & 1) ? -1.0 : 1.0)) *
                               a.x[(int64_t)(wp >> 1) * dup + iu2];
                }
        }
        const int64_t o = c0 * dup + e;
        if (MODE == kPlain) {
            a.y[o] = acc;
        } else {
            const T yy = (T)(sx * (double)acc);
            a.y[o] = yy;
            dot += (double)xo * (double)yy;
        }
    }
    if (MODE == kLanczos) {
        const double s = block_sum(dot, red);
        if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
    }
}


// ------------------------------------------------------------------------
// Persistent column kernel for large columns (D_up up to 1024*RPT).
// One 1024-lane workgroup per CU (two where registers and LDS allow) walks
// columns c = xcd*nper + loc + k*grid: chunks of nper consecutive columns are
// dealt round-robin to the 8 XCDs (workgroup b runs on XCD b%8), so each
// XCD's concurrent columns are neighbours and the whole chip works on one
// window of consecutive columns.  Per column:
//   1. x(:,c) is copied into LDS by LDS DMA (global_load_lds), which needs no
//      registers and overlaps the down hops;
//   2. down hops: for each neighbour column (HDU at a time) every lane loads
//      its RPT rows, then accumulates;
//   3. wait for the DMA, barrier; up hops gather from LDS through the packed
//      4-byte hop words (target rank, coefficient index, sign), whose table
//      stays L2-resident;
//   4. the diagonal d_up + d_dw + W is added last from the LDS copy of x, and
//      y is written with non-temporal stores.
// Each lane owns rows tid + 1024*r, r < RPT.  Rows past D_up load row D_up-1
// (in bounds) and are never stored: the inner loops carry no per-row branch,
// so all loads of a hop step are in flight before the first is used.
// Summation order per row: down hops in table order, up hops in table order,
// then the diagonal.
//
// SPLIT: the down hops were already summed into Y by k_hxv_dn (the same FMA
// chain in the same order); step 2 loads that partial column instead, and the
// kernel finishes Y in place.  The result is bit-identical to the one-pass form.
template <typename T, int MODE, int RPT, bool SPLIT>
__global__ void __launch_bounds__(1024) k_hxv_col(
    const T *__restrict__ X, T *__restrict__ Y, int32_t dup, int32_t ddw, int32_t nhu, int32_t nhd,
    int32_t nimp, int32_t ncoef, const uint32_t *__restrict__ HU, const double *__restrict__ COEF,
    const int32_t *__restrict__ HDI, const double *__restrict__ HDV, const double *__restrict__ DU,
    const double *__restrict__ DD, const uint8_t *__restrict__ IMPU, const uint8_t *__restrict__ IMPD,
    const double *__restrict__ WTAB, int32_t dma16, const ChainScalars *__restrict__ CS,
    double *__restrict__ PARTIALS)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T *col = reinterpret_cast<T *>(smem);
    __shared__ T sc[128];
    __shared__ double wt[8];
    __shared__ double red[16];
    // neighbour columns in flight: two where the registers allow (FP32 up to
    // 13 rows per lane), else one; FP64 spills at two
    constexpr int HDU = (sizeof(T) == 4 && RPT <= 13) ? 2 : 1;
    double sx = 1.0;
    if (MODE == kLanczos) {
        if (CS->done) return;
        sx = CS->sx;
    }
    for (int t = threadIdx.x; t < ncoef; t += blockDim.x) sc[t] = (T)COEF[t];
    const int nper = gridDim.x / 8;                  // workgroups per XCD
    const int xcd = blockIdx.x & 7, loc = blockIdx.x >> 3;
    const int last = dup - 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int colbytes = dup * (int)sizeof(T);
    // impurity bits of the lane's rows, four per register
    constexpr int NP = (RPT + 3) / 4;
    uint32_t imp4[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) imp4[q] = 0u;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int ic = min((int)threadIdx.x + r * 1024, last);
        imp4[r / 4] |= (uint32_t)IMPU[ic] << (8 * (r % 4));
    }
    double dot = 0.0;
    for (int c = xcd * nper + loc; c < ddw; c += gridDim.x) {
        const char *xc = reinterpret_cast<const char *>(X + (int64_t)c * dup);
        // 1. x(:,c) -> LDS: 16-byte granules (1 KB per wave instruction)
        //    when every column start is 16-byte aligned, else 4-byte ones
        if (dma16) {
            for (int k = wave; k * 1024 < colbytes; k += 16) {
                const int off = k * 1024 + lane * 16;
                if (off < colbytes)
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void *)(xc + off),
                        (__attribute__((address_space(3))) void *)(smem + k * 1024), 16, 0, 0);
            }
        } else {
            for (int k = wave; k * 256 < colbytes; k += 16) {
                const int off = k * 256 + lane * 4;
                if (off < colbytes)
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void *)(xc + off),
                        (__attribute__((address_space(3))) void *)(smem + k * 256), 4, 0, 0);
            }
        }
        if ((int)threadIdx.x < nimp)
            wt[threadIdx.x] = DD[c] + WTAB[threadIdx.x * nimp + IMPD[c]];
        T acc[RPT];
        if constexpr (SPLIT) {
            // 2'. the down-hop sums of this column, written by k_hxv_dn
            const T *yc = Y + (int64_t)c * dup;
#pragma unroll
            for (int r = 0; r < RPT; ++r)
                acc[r] = __builtin_nontemporal_load(&yc[min((int)threadIdx.x + r * 1024, last)]);
        } else {
#pragma unroll
        for (int r = 0; r < RPT; ++r) acc[r] = (T)0;
        }
        // 2. down hops
        for (int h0 = 0; !SPLIT && h0 < nhd; h0 += HDU) {
            T v[HDU];
            const T *xj[HDU];
#pragma unroll
            for (int q = 0; q < HDU; ++q) {
                const int h = min(h0 + q, nhd - 1);
                v[q] = h0 + q < nhd ? (T)HDV[(int64_t)h * ddw + c] : (T)0;
                xj[q] = X + (int64_t)HDI[(int64_t)h * ddw + c] * dup;
            }
            T g[HDU][RPT];
#pragma unroll
            for (int q = 0; q < HDU; ++q)
#pragma unroll
                for (int r = 0; r < RPT; ++r) g[q][r] = xj[q][min((int)threadIdx.x + r * 1024, last)];
#pragma unroll
            for (int r = 0; r < RPT; ++r)
#pragma unroll
                for (int q = 0; q < HDU; ++q) acc[r] += v[q] * g[q][r];
        }
        // 3. the DMA has landed for every wave, then up hops from LDS
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int h = 0; h < nhu; ++h) {
            uint32_t w[RPT];
#pragma unroll
            for (int r = 0; r < RPT; ++r) w[r] = HU[(int64_t)h * dup + min((int)threadIdx.x + r * 1024, last)];
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                T vv = sc[(w[r] >> 1) & 127u];
                vv = (w[r] & 1u) ? -vv : vv;
                acc[r] += vv * col[w[r] >> 8];
            }
        }
        // 4. diagonal last, with the lane's own x from LDS
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int i = threadIdx.x + r * 1024;
            if (i < dup) {
                const int64_t o = (int64_t)c * dup + i;
                const T xi = col[i];
                const uint32_t ib = (imp4[r / 4] >> (8 * (r % 4))) & 0xffu;
                const T hv = acc[r] + (T)(DU[i] + wt[ib]) * xi;
                if (MODE == kPlain) {
                    __builtin_nontemporal_store(hv, &Y[o]);
                } else {
                    const T yy = (T)(sx * (double)hv);
                    __builtin_nontemporal_store(yy, &Y[o]);
                    dot += (double)xi * (double)yy;
                }
            }
        }
        __syncthreads();
    }
    if (MODE == kLanczos) {
        const double s = block_sum(dot, red);
        if (threadIdx.x == 0) PARTIALS[blockIdx.x] = s;
    }
}

// ------------------------------------------------------------------------
// Down-spin hops of a large sector over row panels: the first pass of the
// split H*v.  For every column c and the rows r of one panel,
//     Y(r, c) = sum_h td_h(c) X(r, jd_h(c))        (table order, from 0)
// Every column of X is read by the ~nhd columns that hop to it.  With whole
// columns (D_up rows) those re-reads miss the 4 MB L2 of an XCD; a panel of
// R rows of all D_dw columns fits, so each element of X is fetched once per
// H*v and re-read from L2.  Panel p is worked on by the workgroups of XCD
// p % 8 (workgroup b runs on XCD b % 8), all of them on the same panel at a
// time, panels in turn.  A 256-lane workgroup covers 256/R column segments:
// lane t owns row t % R of column segment t / R.  NH hops are loaded and
// gathered before any is added; a padded hop adds 0 * x, which leaves the sum
// unchanged.  Rows past D_up load row D_up - 1 (in bounds) and are not stored.
template <typename T, int MODE, int R>
__global__ void __launch_bounds__(256) k_hxv_dn(
    const T *__restrict__ X, T *__restrict__ Y, int32_t dup, int32_t ddw, int32_t nhd,
    const int32_t *__restrict__ HDI, const T *__restrict__ HDV, const ChainScalars *__restrict__ CS)
{
    constexpr int SEG = 256 / R;
    constexpr int NH = 8;
    if (MODE == kLanczos && CS->done) return;
    const int nper = gridDim.x / 8;
    const int xcd = blockIdx.x & 7, loc = blockIdx.x >> 3;
    const int seg = threadIdx.x / R, rr = threadIdx.x % R;
    const int npanel = (dup + R - 1) / R;
    const int ncg = (ddw + SEG - 1) / SEG;
    for (int p = xcd; p < npanel; p += 8) {
        const int r0 = p * R + rr;
        const int row = min(r0, dup - 1);
        for (int g = loc; g < ncg; g += nper) {
            const int c = min(g * SEG + seg, ddw - 1);
            const bool st = r0 < dup && g * SEG + seg < ddw;
            T acc = (T)0;
            for (int h0 = 0; h0 < nhd; h0 += NH) {
                int32_t j[NH];
                T v[NH], x[NH];
#pragma unroll
                for (int q = 0; q < NH; ++q) {
                    const int h = min(h0 + q, nhd - 1);
                    j[q] = HDI[(int64_t)h * ddw + c];
                    v[q] = h0 + q < nhd ? HDV[(int64_t)h * ddw + c] : (T)0;
                }
#pragma unroll
                for (int q = 0; q < NH; ++q) x[q] = X[(int64_t)j[q] * dup + row];
#pragma unroll
                for (int q = 0; q < NH; ++q) acc += v[q] * x[q];
            }
            if (st) __builtin_nontemporal_store(acc, &Y[(int64_t)c * dup + row]);
        }
    }
}

// rows per lane the column kernel supports per element type (register budget)
constexpr int kColMaxRptF64 = 13;
constexpr int kColMaxRptF32 = 16;

// dynamic LDS of the column kernel: the column rounded up to whole 1 KB DMA
// chunks (the last 16-byte-granule chunk may be partial)
static size_t col_lds_bytes(int64_t dup, size_t esz)
{
    return ((size_t)dup * esz + 1023) & ~(size_t)1023;
}

static bool use_col_kernel(const dmed_sector *s, int prec)
{
    const Context &c = ctx();
    const size_t esz = prec == DMED_FP32 ? 4 : 8;
    const int64_t dup = s->up.dim;
    const int rpt = (int)((dup + 1023) / 1024);
    const int maxr = prec == DMED_FP32 ? kColMaxRptF32 : kColMaxRptF64;
    return !s->jterms && s->up.packed && s->up.ncoef <= 128 && s->nimp <= 8 && dup >= 512 &&
           rpt <= maxr && col_lds_bytes(dup, esz) + 2048 <= c.lds_max && s->dw.dim >= 64;
}

// A kernel's dynamic-LDS limit is process-wide state: it is raised once per
// kernel, to all the LDS a workgroup may have, so concurrent launches from
// several host threads never change it under one another.
template <auto KERN>
static int lds_attr_once()
{
    static const hipError_t e = [] {
        hipFuncAttributes fa{};
        hipError_t r = hipFuncGetAttributes(&fa, (const void *)KERN);
        if (r != hipSuccess) return r;
        const size_t lim = ctx().lds_max > fa.sharedSizeBytes ? ctx().lds_max - fa.sharedSizeBytes : 0;
        return hipFuncSetAttribute((const void *)KERN, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lim);
    }();
    if (e != hipSuccess)
        return fail(DMED_EHIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
    return 0;
}

template <typename T, int RPT>
static int col_fit_per_cu(size_t shmem)
{
    int nb = 1;
    if (lds_attr_once<k_hxv_col<T, kLanczos, RPT, false>>() != 0 ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, (const void *)k_hxv_col<T, kLanczos, RPT, false>, 1024, shmem) != hipSuccess)
        nb = 1;
    return nb;
}

// ------------------------------------------------------------ autotuner
// Opt-in (DMED_TUNE=1).  The first H*v of a sector shape times the launch
// shapes that give the same y bit for bit: the tile kernel's columns per
// workgroup and workgroup size, and the column kernel's workgroups per CU.
// It keeps the fastest shape for that sector shape and precision.
// DMED_TUNE_FILE=path loads earlier choices on first use and appends new
// ones, so later runs reuse them.  Lanczos-mode partial sums follow the
// grid, so a tuned run's alpha/beta can differ from the fixed rules' in the
// last bits.  By default the fixed rules below apply, and runs are
// reproducible bit for bit.
struct TuneChoice {
    int cb = 0, bs = 0;      // tile kernel (0: fixed rule)
    int wgs = 0;             // column kernel workgroups per CU (0: fixed rule)
};
using TuneKey = std::tuple<int64_t, int64_t, int, int, int, int>;

static TuneKey tune_key(const dmed_sector *s, int prec)
{
    return TuneKey{s->up.dim, s->dw.dim, s->up.nhop, s->dw.nhop, s->nimp,
                   prec == DMED_FP32 ? DMED_FP32 : DMED_FP64};
}

struct TuneTable {
    std::mutex mu;
    std::map<TuneKey, TuneChoice> tab;
    bool loaded = false;
};

static TuneTable &tune_table()
{
    static TuneTable t;
    return t;
}

static bool tune_enabled()
{
    static const bool on = [] {
        const char *e = std::getenv("DMED_TUNE");
        return e && e[0] == '1';
    }();
    return on;
}

// choice recorded for this shape (no tuning here); zeros when none
static TuneChoice tune_get(const dmed_sector *s, int prec)
{
    if (!tune_enabled()) return TuneChoice{};
    TuneTable &t = tune_table();
    std::lock_guard<std::mutex> lk(t.mu);
    auto it = t.tab.find(tune_key(s, prec));
    return it == t.tab.end() ? TuneChoice{} : it->second;
}

// Workgroups of the persistent column kernel: as many per CU as its
// registers and LDS allow, at most two (a third lowered the rate at
// D_up = 3432), in multiples of the 8 XCDs, never more than the columns.
static unsigned col_grid(const dmed_sector *s, int prec)
{
    const int ncu = std::max(8, ctx().num_cu);
    const int64_t dup = s->up.dim;
    const int rpt = (int)((dup + 1023) / 1024);
    const bool f32 = prec == DMED_FP32;
    const size_t shmem = col_lds_bytes(dup, f32 ? 4 : 8);
    int fit = 1;
    switch (rpt) {
#define DMED_FIT(R) \
    case R: fit = f32 ? col_fit_per_cu<float, R>(shmem) : col_fit_per_cu<double, R>(shmem); break;
        DMED_FIT(1) DMED_FIT(2) DMED_FIT(3) DMED_FIT(4) DMED_FIT(5) DMED_FIT(6) DMED_FIT(7)
#undef DMED_FIT
    default: fit = 1; break;
    }
    int per_cu = std::max(1, std::min(fit, 2));
    const TuneChoice tc = tune_get(s, prec);
    if (tc.wgs > 0) per_cu = std::min(tc.wgs, std::max(1, fit));
    int g = (ncu / 8) * 8 * per_cu;
    while (g > 8 && g / 8 > (s->dw.dim + 7) / 8) g -= 8;
    return (unsigned)g;
}

template <typename T, int MODE, int RPT, bool SPLIT>
static int launch_col_t(const dmed_sector *s, const void *x, void *y, ChainScalars *cs,
                        double *partials, unsigned grid, hipStream_t st)
{
    const int64_t dup = s->up.dim;
    const size_t shmem = col_lds_bytes(dup, sizeof(T));
    // 16-byte DMA granules need every column start 16-byte aligned
    const int dma16 = ((uintptr_t)x % 16 == 0) && ((size_t)dup * sizeof(T)) % 16 == 0;
    constexpr auto kern = k_hxv_col<T, MODE, RPT, SPLIT>;
    if (int rc = lds_attr_once<kern>()) return rc;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(1024),
                       shmem, st, static_cast<const T *>(x), static_cast<T *>(y), (int32_t)dup,
                       (int32_t)s->dw.dim, (int32_t)s->up.nhop, (int32_t)s->dw.nhop, (int32_t)s->nimp,
                       (int32_t)s->up.ncoef, s->up.hpk, s->up.coef, s->dw.hidx, s->dw.hval,
                       s->up.diag, s->dw.diag, s->up.imp, s->dw.imp, s->W, (int32_t)dma16,
                       (const ChainScalars *)cs, partials);
    DMED_HIP(hipGetLastError());
    return 0;
}

template <typename T, int MODE, bool SPLIT>
static int launch_col_rpt(int rpt, const dmed_sector *s, const void *x, void *y,
                          ChainScalars *cs, double *partials, unsigned grid, hipStream_t st)
{
    switch (rpt) {
#define DMED_RPT(R) \
    case R: return launch_col_t<T, MODE, R, SPLIT>(s, x, y, cs, partials, grid, st);
        DMED_RPT(1) DMED_RPT(2) DMED_RPT(3) DMED_RPT(4) DMED_RPT(5) DMED_RPT(6) DMED_RPT(7) DMED_RPT(8)
        DMED_RPT(9) DMED_RPT(10) DMED_RPT(11) DMED_RPT(12) DMED_RPT(13)
#undef DMED_RPT
    case 14: case 15: case 16:
        if constexpr (sizeof(T) == 4) {
            if (rpt == 14) return launch_col_t<T, MODE, 14, SPLIT>(s, x, y, cs, partials, grid, st);
            if (rpt == 15) return launch_col_t<T, MODE, 15, SPLIT>(s, x, y, cs, partials, grid, st);
            return launch_col_t<T, MODE, 16, SPLIT>(s, x, y, cs, partials, grid, st);
        }
        break;
    default:
        break;
    }
    return fail(DMED_EINVAL, "column kernel: unsupported rows per lane");
}

// first pass of the split H*v: down hops over row panels of `panel` rows,
// `wgs` 256-lane workgroups per CU
template <typename T, int MODE>
static int launch_dn(const dmed_sector *s, const void *x, void *y, ChainScalars *cs, int panel,
                     int wgs, hipStream_t st)
{
    const int ncu = std::max(8, ctx().num_cu);
    const unsigned grid = (unsigned)((ncu / 8) * 8 * std::max(1, wgs));
    const T *hv;
    if constexpr (sizeof(T) == 8) hv = reinterpret_cast<const T *>(s->dw.hval);
    else hv = reinterpret_cast<const T *>(s->dw.hval32);
    const int32_t dup = (int32_t)s->up.dim, ddw = (int32_t)s->dw.dim, nhd = s->dw.nhop;
    switch (panel) {
#define DMED_PANEL(R)                                                                          \
    case R:                                                                                   \
        hipLaunchKernelGGL((k_hxv_dn<T, MODE, R>), dim3(grid), dim3(256), 0, st,               \
                           static_cast<const T *>(x), static_cast<T *>(y), dup, ddw, nhd,      \
                           s->dw.hidx, hv, (const ChainScalars *)cs);                          \
        break;
        DMED_PANEL(8) DMED_PANEL(16) DMED_PANEL(32) DMED_PANEL(64)
#undef DMED_PANEL
    default:
        return fail(DMED_EINVAL, "down-hop pass: unsupported panel rows");
    }
    DMED_HIP(hipGetLastError());
    return 0;
}

struct LaunchShape {
    int cb, bs;
    bool lds;
    size_t shmem;
    unsigned grid;
};

static LaunchShape pick_shape(const dmed_sector *s, int prec)
{
    const Context &c = ctx();
    const size_t esz = prec == DMED_FP32 ? 4 : 8;
    const int64_t dup = s->up.dim, ddw = s->dw.dim;
    LaunchShape sh{};
    const size_t colb = (size_t)dup * esz;
    const size_t wtb = 8 * 64 * (size_t)s->nimp;
    const size_t lds_budget = c.lds_max - 1024;
    sh.lds = colb + wtb + 16 <= lds_budget;
    // ~2K elements per workgroup, but never fewer than two workgroups per CU
    // while the sector has columns for them; at least ~256 elements each
    const int64_t dupc = std::max<int64_t>(dup, 1);
    int64_t cb = std::max<int64_t>(1, (2048 + dupc - 1) / dupc);
    const int64_t fill = std::max<int64_t>(1, ddw / (2 * std::max(c.num_cu, 1)));
    if (cb > fill) cb = std::max<int64_t>(fill, (256 + dupc - 1) / dupc);
    cb = std::min<int64_t>(std::min<int64_t>(cb, ddw), 64);
    if (sh.lds)
        while (cb > 1 && ((size_t)cb * colb + 16 + 8 * (size_t)cb * s->nimp) > 64 * 1024) --cb;
    sh.cb = (int)std::max<int64_t>(cb, 1);
    const int64_t nel = (int64_t)sh.cb * dup;
    sh.bs = nel >= 4096 ? 1024 : (nel >= 1024 ? 512 : 256);
    const TuneChoice tc = tune_get(s, prec);
    if (tc.cb > 0) {
        sh.cb = (int)std::min<int64_t>(tc.cb, std::max<int64_t>(ddw, 1));
        sh.bs = tc.bs;
    }
    sh.shmem = (sh.lds ? (((size_t)sh.cb * colb + 15) & ~(size_t)15) : 0) +
               8 * (size_t)sh.cb * s->nimp;
    sh.grid = (unsigned)((ddw + sh.cb - 1) / sh.cb);
    return sh;
}

template <typename T, int MODE, bool LDS>
static int launch_t(const dmed_sector *s, const void *x, void *y, ChainScalars *cs,
                    double *partials, const LaunchShape &sh, hipStream_t st)
{
    HxvArgs<T> a;
    a.x = static_cast<const T *>(x);
    a.y = static_cast<T *>(y);
    a.dup = (int32_t)s->up.dim;
    a.ddw = (int32_t)s->dw.dim;
    a.nhu = s->up.nhop;
    a.nhd = s->dw.nhop;
    a.nimp = s->nimp;
    a.cb = sh.cb;
    a.inv_dup = 1.0f / (float)a.dup;
    a.hu_idx = s->up.hidx;
    a.hd_idx = s->dw.hidx;
    if constexpr (sizeof(T) == 8) {
        a.hu_val = reinterpret_cast<const T *>(s->up.hval);
        a.hd_val = reinterpret_cast<const T *>(s->dw.hval);
    } else {
        a.hu_val = reinterpret_cast<const T *>(s->up.hval32);
        a.hd_val = reinterpret_cast<const T *>(s->dw.hval32);
    }
    a.du = s->up.diag;
    a.dd = s->dw.diag;
    a.impu = s->up.imp;
    a.impd = s->dw.imp;
    a.W = s->W;
    a.ju = s->up.ipair;
    a.jd = s->dw.ipair;
    a.jterms = s->jterms;
    a.norb = s->norb;
    a.jx = s->jx;
    a.jp = s->jp;
    a.cs = cs;
    a.partials = partials;
    constexpr auto kern = k_hxv<T, MODE, LDS>;
    if (sh.shmem > 64 * 1024)
        if (int rc = lds_attr_once<kern>()) return rc;
    hipLaunchKernelGGL(kern, dim3(sh.grid), dim3(sh.bs), sh.shmem, st, a);
    DMED_HIP(hipGetLastError());
    return 0;
}

// ms per launch of `fn` on st: the best of three batches of `rep` launches,
// after one untimed launch; -1 on a failure
template <typename F>
static float time_launches(F fn, int rep, hipStream_t st)
{
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return -1.f;
    if (hipEventCreate(&e1) != hipSuccess) { hipEventDestroy(e0); return -1.f; }
    float best = -1.f;
    bool ok = fn() == 0;
    for (int b = 0; b < 3 && ok; ++b) {
        float ms = 0.f;
        ok = hipEventRecord(e0, st) == hipSuccess;
        for (int r = 0; r < rep && ok; ++r) ok = fn() == 0;
        ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
             hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        if (ok && (best < 0 || ms / rep < best)) best = ms / rep;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return ok ? best : -1.f;
}

// keep the fixed rule unless a candidate beats it by more than 3 %
constexpr float kTuneMargin = 0.97f;

static void tune_load_file(TuneTable &t)
{
    t.loaded = true;
    const char *path = std::getenv("DMED_TUNE_FILE");
    if (!path) return;
    FILE *f = std::fopen(path, "r");
    if (!f) return;
    long long dup, ddw;
    int nhu, nhd, nimp, prec, cb, bs, wgs;
    while (std::fscanf(f, "%lld %lld %d %d %d %d %d %d %d", &dup, &ddw, &nhu, &nhd, &nimp, &prec,
                       &cb, &bs, &wgs) == 9)
        t.tab[TuneKey{dup, ddw, nhu, nhd, nimp, prec}] = TuneChoice{cb, bs, wgs};
    std::fclose(f);
}

// With DMED_TUNE=1, make sure this sector shape has a recorded launch shape,
// timing the candidates on st the first time (not while st is capturing a
// graph: then the fixed rules apply until a later eager call).
void tune_ensure(const dmed_sector *s, int prec, hipStream_t st)
{
    if (!tune_enabled() || s->dim == 0) return;
    prec = prec == DMED_FP32 ? DMED_FP32 : DMED_FP64;
    TuneTable &t = tune_table();
    const TuneKey key = tune_key(s, prec);
    {
        std::lock_guard<std::mutex> lk(t.mu);
        if (!t.loaded) tune_load_file(t);
        if (t.tab.count(key)) return;
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return;
    const size_t esz = prec == DMED_FP32 ? 4 : 8;
    void *x = nullptr, *y = nullptr;
    if (hipMalloc(&x, s->dim * esz) != hipSuccess || hipMalloc(&y, s->dim * esz) != hipSuccess) {
        (void)hipGetLastError();
        if (x) hipFree(x);
        std::lock_guard<std::mutex> lk(t.mu);
        t.tab[key] = TuneChoice{};
        return;
    }
    hipMemsetAsync(x, 0, s->dim * esz, st);
    TuneChoice best{};
    float tbest = -1.f;
    const bool f32 = prec == DMED_FP32;
    const int rep = s->dim >= (int64_t)1 << 24 ? 3 : 20;
    if (use_col_kernel(s, prec)) {
        // workgroups per CU the registers and LDS allow: 1, and 2 if it fits
        const int rpt = (int)((s->up.dim + 1023) / 1024);
        const int ncu = std::max(8, ctx().num_cu);
        int fit = 1;
        const size_t shmem = col_lds_bytes(s->up.dim, esz);
        switch (rpt) {
#define DMED_FIT(R) \
    case R: fit = f32 ? col_fit_per_cu<float, R>(shmem) : col_fit_per_cu<double, R>(shmem); break;
            DMED_FIT(1) DMED_FIT(2) DMED_FIT(3) DMED_FIT(4) DMED_FIT(5) DMED_FIT(6) DMED_FIT(7)
#undef DMED_FIT
        default: fit = 1; break;
        }
        // the fixed rule first, then the other count
        const int w0 = std::max(1, std::min(fit, 2));
        for (int w : {w0, 3 - w0}) {
            if (w > std::min(fit, 2)) continue;
            unsigned g = (unsigned)((ncu / 8) * 8 * w);
            while (g > 8 && g / 8 > (s->dw.dim + 7) / 8) g -= 8;
            auto fn = [&]() {
                return f32 ? launch_col_rpt<float, kPlain, false>(rpt, s, x, y, nullptr, nullptr, g, st)
                           : launch_col_rpt<double, kPlain, false>(rpt, s, x, y, nullptr, nullptr, g, st);
            };
            const float ms = time_launches(fn, rep, st);
            if (ms > 0 && (tbest < 0 || ms < kTuneMargin * tbest)) {
                tbest = ms;
                best = TuneChoice{0, 0, w};
            }
        }
    } else {
        LaunchShape base = pick_shape(s, prec);     // no recorded choice yet: the fixed rule
        const int64_t dup = s->up.dim, ddw = s->dw.dim;
        const size_t colb = (size_t)dup * esz;
        std::vector<std::pair<int, int>> cand{{base.cb, base.bs}};
        for (int cb = 1; cb <= 64 && cb <= ddw; cb *= 2)
            for (int bs = 256; bs <= 1024; bs *= 2)
                if (cb != base.cb || bs != base.bs) cand.push_back({cb, bs});
        for (auto [cb, bs] : cand) {
            {
                LaunchShape sh = base;
                sh.cb = cb;
                sh.bs = bs;
                if (sh.lds && (size_t)cb * colb + 16 + 8 * (size_t)cb * s->nimp > 64 * 1024) continue;
                sh.shmem = (sh.lds ? (((size_t)cb * colb + 15) & ~(size_t)15) : 0) + 8 * (size_t)cb * s->nimp;
                sh.grid = (unsigned)((ddw + cb - 1) / cb);
                auto fn = [&]() {
                    if (f32) return sh.lds ? launch_t<float, kPlain, true>(s, x, y, nullptr, nullptr, sh, st)
                                           : launch_t<float, kPlain, false>(s, x, y, nullptr, nullptr, sh, st);
                    return sh.lds ? launch_t<double, kPlain, true>(s, x, y, nullptr, nullptr, sh, st)
                                  : launch_t<double, kPlain, false>(s, x, y, nullptr, nullptr, sh, st);
                };
                const float ms = time_launches(fn, rep, st);
                if (ms > 0 && (tbest < 0 || ms < kTuneMargin * tbest)) {
                    tbest = ms;
                    best = TuneChoice{cb, bs, 0};
                }
            }
        }
    }
    hipStreamSynchronize(st);
    hipFree(x);
    hipFree(y);
    (void)hipGetLastError();
    std::lock_guard<std::mutex> lk(t.mu);
    if (t.tab.count(key)) return;            // another thread got there first
    t.tab[key] = best;
    if (const char *path = std::getenv("DMED_TUNE_FILE")) {
        if (FILE *f = std::fopen(path, "a")) {
            std::fprintf(f, "%lld %lld %d %d %d %d %d %d %d\n", (long long)std::get<0>(key),
                         (long long)std::get<1>(key), std::get<2>(key), std::get<3>(key),
                         std::get<4>(key), std::get<5>(key), best.cb, best.bs, best.wgs);
            std::fclose(f);
        }
    }
}

// Split H*v (down-hop panel pass + column pass) for the column kernel's
// sectors: panel rows and panel-pass workgroups per CU; panel 0 keeps the
// one-pass kernel.
struct SplitShape {
    int panel = 0, wgs = 4;
};

static SplitShape split_shape(const dmed_sector *s, int prec)
{
    SplitShape sp;
    if (const char *e = std::getenv("DMED_HXV_PANEL")) sp.panel = std::atoi(e);
    if (const char *e = std::getenv("DMED_HXV_DNWGS")) sp.wgs = std::max(1, std::atoi(e));
    return sp;
}

int launch_hxv(const dmed_sector *s, const void *x, void *y, int prec, int mode,
               ChainScalars *cs, double *partials, int *nparts, hipStream_t st)
{
    if (s->dim == 0) return 0;
    if (mode == kLanczos && partials == nullptr) return fail(DMED_EINVAL, "no partials");
    tune_ensure(s, prec, st);
    if (use_col_kernel(s, prec)) {
        const unsigned grid = col_grid(s, prec);
        if (nparts) *nparts = (int)grid;
        const int rpt = (int)((s->up.dim + 1023) / 1024);
        const bool f32 = prec == DMED_FP32;
        const SplitShape sp = split_shape(s, prec);
        if (sp.panel > 0) {
            // down hops into y over L2-resident row panels, then the column
            // kernel adds the up hops and the diagonal in place
            int rc = f32 ? (mode == kPlain ? launch_dn<float, kPlain>(s, x, y, cs, sp.panel, sp.wgs, st)
                                           : launch_dn<float, kLanczos>(s, x, y, cs, sp.panel, sp.wgs, st))
                         : (mode == kPlain ? launch_dn<double, kPlain>(s, x, y, cs, sp.panel, sp.wgs, st)
                                           : launch_dn<double, kLanczos>(s, x, y, cs, sp.panel, sp.wgs, st));
            if (rc) return rc;
            if (f32)
                return mode == kPlain ? launch_col_rpt<float, kPlain, true>(rpt, s, x, y, cs, partials, grid, st)
                                      : launch_col_rpt<float, kLanczos, true>(rpt, s, x, y, cs, partials, grid, st);
            return mode == kPlain ? launch_col_rpt<double, kPlain, true>(rpt, s, x, y, cs, partials, grid, st)
                                  : launch_col_rpt<double, kLanczos, true>(rpt, s, x, y, cs, partials, grid, st);
        }
        if (f32)
            return mode == kPlain ? launch_col_rpt<float, kPlain, false>(rpt, s, x, y, cs, partials, grid, st)
                                  : launch_col_rpt<float, kLanczos, false>(rpt, s, x, y, cs, partials, grid, st);
        return mode == kPlain ? launch_col_rpt<double, kPlain, false>(rpt, s, x, y, cs, partials, grid, st)
                              : launch_col_rpt<double, kLanczos, false>(rpt, s, x, y, cs, partials, grid, st);
    }
    const LaunchShape sh = pick_shape(s, prec);
    if (nparts) *nparts = (int)sh.grid;
    if (mode == kLanczos && partials == nullptr) return fail(DMED_EINVAL, "no partials");
    const bool f32 = prec == DMED_FP32;
    if (mode == kPlain) {
        if (f32) return sh.lds ? launch_t<float, kPlain, true>(s, x, y, cs, partials, sh, st)
                               : launch_t<float, kPlain, false>(s, x, y, cs, partials, sh, st);
        return sh.lds ? launch_t<double, kPlain, true>(s, x, y, cs, partials, sh, st)
                      : launch_t<double, kPlain, false>(s, x, y, cs, partials, sh, st);
    }
    if (f32) return sh.lds ? launch_t<float, kLanczos, true>(s, x, y, cs, partials, sh, st)
                           : launch_t<float, kLanczos, false>(s, x, y, cs, partials, sh, st);
    return sh.lds ? launch_t<double, kLanczos, true>(s, x, y, cs, partials, sh, st)
                  : launch_t<double, kLanczos, false>(s, x, y, cs, partials, sh, st);
}

int hxv_nparts(const dmed_sector *s, int prec)
{
    tune_ensure(s, prec, ex().stream);
    if (use_col_kernel(s, prec)) return (int)col_grid(s, prec);
    return (int)pick_shape(s, prec).grid;
}

}  // namespace dmed
