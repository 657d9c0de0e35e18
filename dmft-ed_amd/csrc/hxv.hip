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
// index (coalesced sweeps).  The epilogue can fuse the first half of a
// Lanczos step (w = sx*Hv - b*svp*v_prev and the block partials of <v,w>).
#include "dmed_internal.h"

#include <algorithm>

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
    const T *__restrict__ vp;             // Lanczos epilogue
    const ChainScalars *cs;
    double *__restrict__ partials;
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

    double sx = 1.0, coef_vp = 0.0;
    if (MODE == kLanczos) {
        if (a.cs->done) return;
        sx = a.cs->sx;
        coef_vp = a.cs->b * a.cs->svp;
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
        for (int h = 0; h < a.nhu; ++h)
            acc += a.hu_val[(int64_t)h * dup + i] * xc[a.hu_idx[(int64_t)h * dup + i]];
        const int64_t c = c0 + cc;
        for (int h = 0; h < a.nhd; ++h) {
            const int64_t j = a.hd_idx[(int64_t)h * a.ddw + c];
            acc += a.hd_val[(int64_t)h * a.ddw + c] * a.x[j * dup + i];
        }
        const int64_t o = c0 * dup + e;
        if (MODE == kPlain) {
            a.y[o] = acc;
        } else {
            const double yy = sx * (double)acc - coef_vp * (double)a.vp[o];
            a.y[o] = (T)yy;
            dot += sx * (double)xo * yy;
        }
    }
    if (MODE == kLanczos) {
        const double s = block_sum(dot, red);
        if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
    }
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
    // aim for ~4-8K elements per workgroup, bounded by the LDS budget
    int cb = (int)std::max<int64_t>(1, std::min<int64_t>(ddw, 8192 / std::max<int64_t>(dup, 1)));
    if (sh.lds)
        while (cb > 1 && ((size_t)cb * colb + 16 + 8 * (size_t)cb * s->nimp) > 64 * 1024) --cb;
    sh.cb = std::min<int>(cb, 64);
    const int64_t nel = (int64_t)sh.cb * dup;
    sh.bs = nel >= 8192 ? 1024 : (nel >= 2048 ? 512 : 256);
    sh.shmem = (sh.lds ? (((size_t)sh.cb * colb + 15) & ~(size_t)15) : 0) +
               8 * (size_t)sh.cb * s->nimp;
    sh.grid = (unsigned)((ddw + sh.cb - 1) / sh.cb);
    return sh;
}

template <typename T, int MODE, bool LDS>
static int launch_t(const dmed_sector *s, const void *x, void *y, const void *vp,
                    ChainScalars *cs, double *partials, const LaunchShape &sh, hipStream_t st)
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
    a.vp = static_cast<const T *>(vp);
    a.cs = cs;
    a.partials = partials;
    auto kern = k_hxv<T, MODE, LDS>;
    if (sh.shmem > 64 * 1024)
        DMED_HIP(hipFuncSetAttribute((const void *)kern,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)sh.shmem));
    hipLaunchKernelGGL(kern, dim3(sh.grid), dim3(sh.bs), sh.shmem, st, a);
    DMED_HIP(hipGetLastError());
    return 0;
}

int launch_hxv(const dmed_sector *s, const void *x, void *y, int prec, int mode,
               const void *vp, ChainScalars *cs, double *partials, int *nparts,
               hipStream_t st)
{
    if (s->dim == 0) return 0;
    const LaunchShape sh = pick_shape(s, prec);
    if (nparts) *nparts = (int)sh.grid;
    if (mode == kLanczos && partials == nullptr) return fail(DMED_EINVAL, "no partials");
    const bool f32 = prec == DMED_FP32;
    if (mode == kPlain) {
        if (f32) return sh.lds ? launch_t<float, kPlain, true>(s, x, y, vp, cs, partials, sh, st)
                               : launch_t<float, kPlain, false>(s, x, y, vp, cs, partials, sh, st);
        return sh.lds ? launch_t<double, kPlain, true>(s, x, y, vp, cs, partials, sh, st)
                      : launch_t<double, kPlain, false>(s, x, y, vp, cs, partials, sh, st);
    }
    if (f32) return sh.lds ? launch_t<float, kLanczos, true>(s, x, y, vp, cs, partials, sh, st)
                           : launch_t<float, kLanczos, false>(s, x, y, vp, cs, partials, sh, st);
    return sh.lds ? launch_t<double, kLanczos, true>(s, x, y, vp, cs, partials, sh, st)
                  : launch_t<double, kLanczos, false>(s, x, y, vp, cs, partials, sh, st);
}

int hxv_nparts(const dmed_sector *s, int prec) { return (int)pick_shape(s, prec).grid; }

}  // namespace dmed
