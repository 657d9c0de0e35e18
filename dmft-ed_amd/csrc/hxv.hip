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
// Persistent column kernel for large columns (D_up up to BS*RPT).
// BS-lane workgroups (one or more per CU, as registers and LDS allow) walk
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
// Each lane owns rows tid + BS*r, r < RPT.  Rows past D_up load row D_up-1
// (in bounds) and are never stored: the inner loops carry no per-row branch,
// so all loads of a hop step are in flight before the first is used.
// Summation order per row: down hops in table order, up hops in table order,
// then the diagonal -- the same for every (BS, RPT) and grid, so the launch
// shape never changes y.  In Lanczos mode each column's <x, y> is reduced on
// its own and stored at PARTIALS[c]: alpha does not depend on the grid either.
// HALVES = 2: steps 2-4 run twice, for the first and the second half of the
// lane's rows, one neighbour column at a time.  Half the accumulators, loads
// and hop words are live at once, which brings the kernel within the
// registers of two workgroups per CU (2*BS/256 waves per SIMD); two
// workgroups let one's LDS gathers overlap the other's far loads.  The sums
// are the same, term for term, as with HALVES = 1.
template <typename T, int MODE, int BS, int RPT, int HALVES = 1>
__global__ void __launch_bounds__(BS, HALVES == 2 ? 2 * BS / 256 : 1) k_hxv_col(
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
        const int ic = min((int)threadIdx.x + r * BS, last);
        imp4[r / 4] |= (uint32_t)IMPU[ic] << (8 * (r % 4));
    }
    // Lanczos-mode <x, y> of a column in the order of the 1024-lane form:
    // row i belongs to virtual lane i % 1024, which sums its rows in order;
    // the 64-lane groups of virtual lanes are reduced by the same shuffle
    // tree and their 16 sums by one more.  Thread t of a BS-lane workgroup
    // holds the virtual lanes t + BS*j, j < M, in dot[j].
    constexpr int M = 1024 / BS;
    static_assert(M * BS == 1024, "lanes per workgroup must divide 1024");
    for (int c = xcd * nper + loc; c < ddw; c += gridDim.x) {
        double dot[M];
#pragma unroll
        for (int j = 0; j < M; ++j) dot[j] = 0.0;
        const char *xc = reinterpret_cast<const char *>(X + (int64_t)c * dup);
        // 1. x(:,c) -> LDS: 16-byte granules (1 KB per wave instruction)
        //    when every column start is 16-byte aligned, else 4-byte ones
        if (dma16) {
            for (int k = wave; k * 1024 < colbytes; k += BS / 64) {
                const int off = k * 1024 + lane * 16;
                if (off < colbytes)
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void *)(xc + off),
                        (__attribute__((address_space(3))) void *)(smem + k * 1024), 16, 0, 0);
            }
        } else {
            for (int k = wave; k * 256 < colbytes; k += BS / 64) {
                const int off = k * 256 + lane * 4;
                if (off < colbytes)
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void *)(xc + off),
                        (__attribute__((address_space(3))) void *)(smem + k * 256), 4, 0, 0);
            }
        }
        if ((int)threadIdx.x < nimp)
            wt[threadIdx.x] = DD[c] + WTAB[threadIdx.x * nimp + IMPD[c]];
        if constexpr (HALVES == 2) {
            constexpr int RH = (RPT + 1) / 2;
#pragma unroll 1
            for (int part = 0; part < 2; ++part) {
                const int rb = part * RH;
                T acc[RH];
#pragma unroll
                for (int r = 0; r < RH; ++r) acc[r] = (T)0;
                // 2. down hops for rows rb .. rb+RH-1
                for (int h = 0; h < nhd; ++h) {
                    const T v = (T)HDV[(int64_t)h * ddw + c];
                    const T *xj = X + (int64_t)HDI[(int64_t)h * ddw + c] * dup;
                    T g[RH];
#pragma unroll
                    for (int r = 0; r < RH; ++r)
                        g[r] = xj[min((int)threadIdx.x + (rb + r) * BS, last)];
#pragma unroll
                    for (int r = 0; r < RH; ++r) acc[r] += v * g[r];
                }
                // 3. the DMA has landed (first half), up hops from LDS
                if (part == 0) {
                    __builtin_amdgcn_s_waitcnt(0);
                    __syncthreads();
                }
                for (int h = 0; h < nhu; ++h) {
                    uint32_t w[RH];
#pragma unroll
                    for (int r = 0; r < RH; ++r)
                        w[r] = HU[(int64_t)h * dup + min((int)threadIdx.x + (rb + r) * BS, last)];
#pragma unroll
                    for (int r = 0; r < RH; ++r) {
                        T vv = sc[(w[r] >> 1) & 127u];
                        vv = (w[r] & 1u) ? -vv : vv;
                        acc[r] += vv * col[w[r] >> 8];
                    }
                }
                // 4. diagonal and store
#pragma unroll
                for (int q = 0; q < RH; ++q) {
                    const int r = rb + q;
                    const int i = threadIdx.x + r * BS;
                    if (r < RPT && i < dup) {
                        const int64_t o = (int64_t)c * dup + i;
                        const T xi = col[i];
                        const uint32_t ib = (imp4[r / 4] >> (8 * (r % 4))) & 0xffu;
                        const T hv = acc[q] + (T)(DU[i] + wt[ib]) * xi;
                        if (MODE == kPlain) {
                            __builtin_nontemporal_store(hv, &Y[o]);
                        } else {
                            const T yy = (T)(sx * (double)hv);
                            __builtin_nontemporal_store(yy, &Y[o]);
                            dot[r % M] += (double)xi * (double)yy;
                        }
                    }
                }
            }
        } else {
            T acc[RPT];
#pragma unroll
            for (int r = 0; r < RPT; ++r) acc[r] = (T)0;
            // 2. down hops
            for (int h0 = 0; h0 < nhd; h0 += HDU) {
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
                    for (int r = 0; r < RPT; ++r) g[q][r] = xj[q][min((int)threadIdx.x + r * BS, last)];
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
                for (int r = 0; r < RPT; ++r) w[r] = HU[(int64_t)h * dup + min((int)threadIdx.x + r * BS, last)];
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
                const int i = threadIdx.x + r * BS;
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
                        dot[r % M] += (double)xi * (double)yy;
                    }
                }
            }
        }
        if (MODE == kLanczos) {
#pragma unroll
            for (int j = 0; j < M; ++j) {
                const double ws = wave_sum(dot[j]);
                if (lane == 0) red[wave + (BS / 64) * j] = ws;
            }
            __syncthreads();
            if (wave == 0) {
                const double v = wave_sum(lane < 16 ? red[lane] : 0.0);
                if (lane == 0) PARTIALS[c] = v;
            }
        }
        __syncthreads();
    }
}

// rows per lane the column kernel supports (register budget): per element
// type for 1024-lane workgroups, and for the 512- and 256-lane forms
constexpr int kColMaxRptF64 = 13;
constexpr int kColMaxRptF32 = 16;
constexpr int kColMaxRptSmall = 8;

// dynamic LDS of the column kernel: the column rounded up to whole 1 KB DMA
// chunks (the last 16-byte-granule chunk may be partial)
static size_t col_lds_bytes(int64_t dup, size_t esz)
{
    return ((size_t)dup * esz + 1023) & ~(size_t)1023;
}

static int col_max_rpt(int bs, int prec)
{
    if (bs != 1024) return kColMaxRptSmall;
    return prec == DMED_FP32 ? kColMaxRptF32 : kColMaxRptF64;
}

// rows per lane of a bs-lane column workgroup, 0 when the column does not fit
static int col_rpt(const dmed_sector *s, int prec, int bs)
{
    const int64_t rpt = (s->up.dim + bs - 1) / bs;
    return rpt >= 1 && rpt <= col_max_rpt(bs, prec) ? (int)rpt : 0;
}

// The fixed rule: the 1024-lane column kernel for sectors with D_up >= 512
// whose column fits in LDS and whose hop words pack; the tile kernel otherwise.
static bool use_col_kernel(const dmed_sector *s, int prec)
{
    const Context &c = ctx();
    const size_t esz = prec == DMED_FP32 ? 4 : 8;
    const int64_t dup = s->up.dim;
    return !s->jterms && s->up.packed && s->up.ncoef <= 128 && s->nimp <= 8 && dup >= 512 &&
           col_rpt(s, prec, 1024) > 0 && col_lds_bytes(dup, esz) + 2048 <= c.lds_max &&
           s->dw.dim >= 64;
}

// A kernel's dynamic-LDS limit is process-wide state: it is raised once per
// kernel, to all the LDS a workgroup may have, before the kernel's first
// launch or occupancy query, under a lock.  Every launch site calls this
// first, so concurrent launches from several host threads never see the
// attribute change under them.
static int lds_attr_once(const void *kern)
{
    static std::mutex mu;
    static std::map<const void *, hipError_t> done;
    std::lock_guard<std::mutex> lk(mu);
    auto it = done.find(kern);
    hipError_t e;
    if (it != done.end()) {
        e = it->second;
    } else {
        hipFuncAttributes fa{};
        e = hipFuncGetAttributes(&fa, kern);
        if (e == hipSuccess) {
            const size_t lim =
                ctx().lds_max > fa.sharedSizeBytes ? ctx().lds_max - fa.sharedSizeBytes : 0;
            e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim);
        }
        done[kern] = e;
    }
    if (e != hipSuccess)
        return fail(DMED_EHIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
    return 0;
}

// Column-kernel instantiations: 1024 lanes with 1..13 (FP64) or 1..16 (FP32)
// rows per lane, 512 and 256 lanes with 1..8; the two-half form (HALVES = 2)
// for 1024 lanes and at least 4 rows per lane.  f(kernel pointer) is called
// with the instantiation for (bs, rpt, halves); false if there is none.
template <typename T, int MODE, int BS, int H, typename F>
static bool col_dispatch_bs(int rpt, F &&f)
{
    switch (rpt) {
#define DMED_R(R)                                                                  \
    case R:                                                                        \
        if constexpr (R <= (BS == 1024 ? (sizeof(T) == 4 ? 16 : 13) : 8) &&        \
                      (H == 1 || (BS == 1024 && R >= 4))) {                        \
            f(k_hxv_col<T, MODE, BS, R, H>);                                       \
            return true;                                                           \
        }                                                                          \
        return false;
        DMED_R(1) DMED_R(2) DMED_R(3) DMED_R(4) DMED_R(5) DMED_R(6) DMED_R(7) DMED_R(8)
        DMED_R(9) DMED_R(10) DMED_R(11) DMED_R(12) DMED_R(13) DMED_R(14) DMED_R(15) DMED_R(16)
#undef DMED_R
    default:
        return false;
    }
}

template <typename T, int MODE, typename F>
static bool col_dispatch(int bs, int rpt, bool halves, F &&f)
{
    if (bs == 1024)
        return halves ? col_dispatch_bs<T, MODE, 1024, 2>(rpt, f)
                      : col_dispatch_bs<T, MODE, 1024, 1>(rpt, f);
    if (halves) return false;
    if (bs == 512) return col_dispatch_bs<T, MODE, 512, 1>(rpt, f);
    if (bs == 256) return col_dispatch_bs<T, MODE, 256, 1>(rpt, f);
    return false;
}

// DMED_COL_HALVES=0 keeps the column kernel to its one-pass form
static bool halves_enabled()
{
    static const bool on = [] {
        const char *e = std::getenv("DMED_COL_HALVES");
        return !(e && e[0] == '0');
    }();
    return on;
}

// workgroups of (bs, rpt, halves) per CU that registers and LDS allow; 0
// when there is no such instantiation
template <typename T>
static int col_fit_per_cu(int bs, int rpt, bool halves, size_t shmem)
{
    int nb = 1;
    const bool ok = col_dispatch<T, kLanczos>(bs, rpt, halves, [&](auto kern) {
        if (lds_attr_once((const void *)kern) != 0 ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)kern, bs, shmem) !=
                hipSuccess)
            nb = 1;
    });
    (void)hipGetLastError();
    return ok ? std::max(nb, 1) : 0;
}

// {workgroups per CU of the one-pass form, of the two-half form (0: none)}
struct ColFit {
    int one = 1, two = 0;
    int best() const { return std::max(one, two); }
};

static ColFit col_fit(const dmed_sector *s, int prec, int bs)
{
    // the occupancy query is the same for every sector of this shape: keep it
    static std::mutex mu;
    static std::map<std::tuple<int64_t, int, int>, ColFit> memo;
    const auto key = std::make_tuple(s->up.dim, prec == DMED_FP32 ? 1 : 0, bs);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = memo.find(key);
        if (it != memo.end()) return it->second;
    }
    const int rpt = col_rpt(s, prec, bs);
    const size_t shmem = col_lds_bytes(s->up.dim, prec == DMED_FP32 ? 4 : 8);
    ColFit f;
    if (rpt == 0) {
        f.one = 0;
    } else {
        const bool f32 = prec == DMED_FP32;
        f.one = std::max(1, f32 ? col_fit_per_cu<float>(bs, rpt, false, shmem)
                                : col_fit_per_cu<double>(bs, rpt, false, shmem));
        if (halves_enabled() && bs == 1024)
            f.two = f32 ? col_fit_per_cu<float>(bs, rpt, true, shmem)
                        : col_fit_per_cu<double>(bs, rpt, true, shmem);
    }
    std::lock_guard<std::mutex> lk(mu);
    memo[key] = f;
    return f;
}

// ------------------------------------------------------------ autotuner
// On by default (DMED_TUNE=0 turns it off) for sectors of at least
// DMED_TUNE_MIN_DIM states (2^20 unless set).  The first H*v of a sector
// shape and precision times the launch shapes that leave every element's
// sum unchanged -- it never switches between the column and the tile
// kernel, whose summation orders differ -- and keeps the fastest:
//   column kernel: lanes per workgroup (1024, 512, 256, i.e. rows per lane
//                  ceil(D_up/lanes)) x workgroups per CU (as many as fit, <= 4);
//   tile kernel:   columns per workgroup (the LDS tile, 1..64) x lanes.
// A candidate replaces the fixed rule only when it is more than 3% faster.
// The column kernel's Lanczos-mode partial sums are one per column, so a
// tuned shape changes neither y nor alpha/beta; the tile kernel keeps the
// fixed rule in Lanczos mode (its partial sums follow the grid).
// DMED_TUNE_FILE=path loads earlier choices on first use and appends new ones.
struct TuneChoice {
    int cb = 0, bs = 0;      // tile kernel (0: fixed rule)
    int cbs = 0, wgs = 0;    // column kernel lanes and workgroups per CU (0: fixed rule)
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
        return !(e && e[0] == '0');
    }();
    return on;
}

static int64_t tune_min_dim()
{
    static const int64_t d = [] {
        const char *e = std::getenv("DMED_TUNE_MIN_DIM");
        return e ? (int64_t)std::atoll(e) : (int64_t)1 << 20;
    }();
    return d;
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

// Column-kernel launch shape: lanes, rows per lane, workgroups per CU, grid.
// Fixed rule: 1024 lanes, as many workgroups per CU as registers and LDS
// allow, at most two (a third lowered the rate at D_up = 3432), in multiples
// of the 8 XCDs, never more than the columns.
struct ColShape {
    int bs = 1024, rpt = 0, per_cu = 1, fit = 1;
    unsigned grid = 8;
    bool tuned = false;
    bool halves = false;     // the two-half form (more workgroups per CU)
};

static unsigned col_grid_for(const dmed_sector *s, int per_cu)
{
    const int ncu = std::max(8, ctx().num_cu);
    int g = (ncu / 8) * 8 * per_cu;
    while (g > 8 && g / 8 > (s->dw.dim + 7) / 8) g -= 8;
    return (unsigned)g;
}

// Workgroups per CU: the tuned count, or two where they fit (the fixed rule).
// The one-pass form is kept whenever it fits that many workgroups; the
// two-half form takes over where only it does.
static ColShape col_shape(const dmed_sector *s, int prec, const TuneChoice &tc)
{
    ColShape c;
    ColFit f;
    if (tc.cbs > 0 && col_rpt(s, prec, tc.cbs) > 0) {
        c.bs = tc.cbs;
        f = col_fit(s, prec, c.bs);
        c.per_cu = std::max(1, std::min(tc.wgs, f.best()));
        c.tuned = true;
    } else {
        f = col_fit(s, prec, 1024);
        c.per_cu = std::max(1, std::min(f.best(), 2));
    }
    c.fit = f.best();
    c.halves = c.per_cu > f.one && f.two >= c.per_cu;
    c.rpt = col_rpt(s, prec, c.bs);
    c.grid = col_grid_for(s, c.per_cu);
    return c;
}

template <typename T, int MODE>
static int launch_col(const dmed_sector *s, const void *x, void *y, ChainScalars *cs,
                      double *partials, const ColShape &c, hipStream_t st)
{
    const int64_t dup = s->up.dim;
    const size_t shmem = col_lds_bytes(dup, sizeof(T));
    // 16-byte DMA granules need every column start 16-byte aligned
    const int dma16 = ((uintptr_t)x % 16 == 0) && ((size_t)dup * sizeof(T)) % 16 == 0;
    int rc = 0;
    const bool ok = col_dispatch<T, MODE>(c.bs, c.rpt, c.halves, [&](auto kern) {
        rc = lds_attr_once((const void *)kern);
        if (rc) return;
        hipLaunchKernelGGL(kern, dim3(c.grid), dim3(c.bs), shmem, st, static_cast<const T *>(x),
                           static_cast<T *>(y), (int32_t)dup, (int32_t)s->dw.dim,
                           (int32_t)s->up.nhop, (int32_t)s->dw.nhop, (int32_t)s->nimp,
                           (int32_t)s->up.ncoef, s->up.hpk, s->up.coef, s->dw.hidx, s->dw.hval,
                           s->up.diag, s->dw.diag, s->up.imp, s->dw.imp, s->W, (int32_t)dma16,
                           (const ChainScalars *)cs, partials);
    });
    if (!ok) return fail(DMED_EINVAL, "column kernel: unsupported rows per lane");
    if (rc) return rc;
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
    const auto kern = k_hxv<T, MODE, LDS>;
    if (sh.shmem > 64 * 1024)
        if (int rc = lds_attr_once((const void *)kern)) return rc;
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
    int nhu, nhd, nimp, prec, cb, bs, cbs, wgs;
    while (std::fscanf(f, "%lld %lld %d %d %d %d %d %d %d %d", &dup, &ddw, &nhu, &nhd, &nimp, &prec,
                       &cb, &bs, &cbs, &wgs) == 10)
        t.tab[TuneKey{dup, ddw, nhu, nhd, nimp, prec}] = TuneChoice{cb, bs, cbs, wgs};
    std::fclose(f);
}

static thread_local bool t_no_alloc = false;
bool thread_no_alloc() { return t_no_alloc; }
NoAllocScope::NoAllocScope() : prev(t_no_alloc) { t_no_alloc = true; }
NoAllocScope::~NoAllocScope() { t_no_alloc = prev; }

// Make sure this sector shape has a recorded launch shape, timing the
// candidates on st the first time (not while st is capturing a graph: then
// the fixed rules apply until a later eager call).
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
        if (s->dim < tune_min_dim()) {
            t.tab[key] = TuneChoice{};
            return;
        }
    }
    if (t_no_alloc) return;                   // timing needs scratch vectors
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
    auto consider = [&](const TuneChoice &tc, auto fn) {
        const float ms = time_launches(fn, rep, st);
        if (ms > 0 && (tbest < 0 || ms < kTuneMargin * tbest)) {
            tbest = ms;
            best = tc;
        }
    };
    if (use_col_kernel(s, prec)) {
        // the fixed rule first, then rows per lane x workgroups per CU
        std::vector<TuneChoice> cand{TuneChoice{}};
        for (int bs : {1024, 512, 256}) {
            if (col_rpt(s, prec, bs) == 0) continue;
            const int fit = std::min(col_fit(s, prec, bs).best(), 4);
            for (int w = 1; w <= fit; ++w) cand.push_back(TuneChoice{0, 0, bs, w});
        }
        for (const TuneChoice &tc : cand) {
            const ColShape c = col_shape(s, prec, tc);
            consider(tc, [&]() {
                return f32 ? launch_col<float, kPlain>(s, x, y, nullptr, nullptr, c, st)
                           : launch_col<double, kPlain>(s, x, y, nullptr, nullptr, c, st);
            });
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
            LaunchShape sh = base;
            sh.cb = cb;
            sh.bs = bs;
            if (sh.lds && (size_t)cb * colb + 16 + 8 * (size_t)cb * s->nimp > 64 * 1024) continue;
            sh.shmem = (sh.lds ? (((size_t)cb * colb + 15) & ~(size_t)15) : 0) + 8 * (size_t)cb * s->nimp;
            sh.grid = (unsigned)((ddw + cb - 1) / cb);
            consider(TuneChoice{cb, bs, 0, 0}, [&]() {
                if (f32) return sh.lds ? launch_t<float, kPlain, true>(s, x, y, nullptr, nullptr, sh, st)
                                       : launch_t<float, kPlain, false>(s, x, y, nullptr, nullptr, sh, st);
                return sh.lds ? launch_t<double, kPlain, true>(s, x, y, nullptr, nullptr, sh, st)
                              : launch_t<double, kPlain, false>(s, x, y, nullptr, nullptr, sh, st);
            });
        }
        if (best.cb == base.cb && best.bs == base.bs) best = TuneChoice{};
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
            std::fprintf(f, "%lld %lld %d %d %d %d %d %d %d %d\n", (long long)std::get<0>(key),
                         (long long)std::get<1>(key), std::get<2>(key), std::get<3>(key),
                         std::get<4>(key), std::get<5>(key), best.cb, best.bs, best.cbs, best.wgs);
            std::fclose(f);
        }
    }
}

// Tile-kernel shape: the tuned one in plain mode, the fixed rule in Lanczos
// mode (whose block partial sums follow the grid).
static LaunchShape tile_shape(const dmed_sector *s, int prec, int mode)
{
    LaunchShape sh = pick_shape(s, prec);
    if (mode != kPlain) return sh;
    const TuneChoice tc = tune_get(s, prec);
    if (tc.cb > 0) {
        const size_t esz = prec == DMED_FP32 ? 4 : 8;
        const int64_t ddw = s->dw.dim;
        sh.cb = (int)std::min<int64_t>(tc.cb, std::max<int64_t>(ddw, 1));
        sh.bs = tc.bs;
        sh.shmem = (sh.lds ? (((size_t)sh.cb * s->up.dim * esz + 15) & ~(size_t)15) : 0) +
                   8 * (size_t)sh.cb * s->nimp;
        sh.grid = (unsigned)((ddw + sh.cb - 1) / sh.cb);
    }
    return sh;
}

int launch_hxv(const dmed_sector *s, const void *x, void *y, int prec, int mode,
               ChainScalars *cs, double *partials, int *nparts, hipStream_t st)
{
    if (s->dim == 0) return 0;
    if (mode == kLanczos && partials == nullptr) return fail(DMED_EINVAL, "no partials");
    tune_ensure(s, prec, st);
    const bool f32 = prec == DMED_FP32;
    if (use_col_kernel(s, prec)) {
        const ColShape c = col_shape(s, prec, tune_get(s, prec));
        if (nparts) *nparts = (int)s->dw.dim;         // one partial per column
        if (f32)
            return mode == kPlain ? launch_col<float, kPlain>(s, x, y, cs, partials, c, st)
                                  : launch_col<float, kLanczos>(s, x, y, cs, partials, c, st);
        return mode == kPlain ? launch_col<double, kPlain>(s, x, y, cs, partials, c, st)
                              : launch_col<double, kLanczos>(s, x, y, cs, partials, c, st);
    }
    const LaunchShape sh = tile_shape(s, prec, mode);
    if (nparts) *nparts = (int)sh.grid;
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

// partial sums a Lanczos-mode H*v of s leaves (column kernel: one per column)
int hxv_nparts(const dmed_sector *s, int prec)
{
    if (use_col_kernel(s, prec)) return (int)s->dw.dim;
    return (int)pick_shape(s, prec).grid;
}

// the plain-mode launch shape of s in precision prec (after tuning)
int hxv_shape(const dmed_sector *s, int prec, dmed_hxv_shape *out)
{
    prec = prec == DMED_FP32 ? DMED_FP32 : DMED_FP64;
    tune_ensure(s, prec, ex().stream);
    *out = dmed_hxv_shape{};
    if (use_col_kernel(s, prec)) {
        const ColShape c = col_shape(s, prec, tune_get(s, prec));
        out->kernel = 2;
        out->lanes = c.bs;
        out->rows_per_lane = c.rpt;
        out->columns_per_group = 1;
        out->groups_per_cu = c.per_cu;
        out->grid = (int32_t)c.grid;
        out->tuned = c.tuned ? 1 : 0;
        out->row_passes = c.halves ? 2 : 1;
    } else {
        const LaunchShape sh = tile_shape(s, prec, kPlain);
        out->kernel = 1;
        out->lanes = sh.bs;
        out->rows_per_lane = (int32_t)((sh.cb * s->up.dim + sh.bs - 1) / sh.bs);
        out->columns_per_group = sh.cb;
        out->groups_per_cu = 0;
        out->grid = (int32_t)sh.grid;
        out->tuned = tune_get(s, prec).cb > 0 ? 1 : 0;
        out->row_passes = 1;
    }
    return 0;
}

}  // namespace dmed
