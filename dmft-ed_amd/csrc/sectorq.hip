// sectorq.hip -- persistent work-queue eigensolver for the small sectors of a
// batched solve.
//
// The per-sector work of ed_diag_c (ED_DIAG.f90:115-228) for a sector of
// dimension D is a thick-restart Lanczos with full reorthogonalisation and
// multiplet-completing deflation rounds.  Driven from the host, a sector of a
// few thousand states costs seven launches per Krylov step and a host round
// trip per restart, and the kernels are far too small to fill the chip.  Here one launch
// takes a queue of such sectors (largest first): each 1024-lane workgroup
// pops a sector from an atomic counter and runs the whole solve for it --
// H*v, the three-term step and classical Gram-Schmidt with a second pass on
// the Daniel-Gragg-Kaufman-Stewart test, restarts, the Rayleigh-Ritz eigenproblem of the projected matrix
// (parallel cyclic Jacobi on the matrix in LDS), the deflation rounds and the
// explicit residuals -- with the Krylov basis in its own slice of HBM, which
// stays in the XCD's L2 at these sizes.  Workgroups leave when the queue is
// empty; every loop inside is bounded (Krylov steps by the iteration budget,
// Jacobi sweeps, deflation rounds by neig + 1).
//
// The algorithm is the one of trl_run / trl_deflated (krylov.hip) step for
// step: the same start vectors, three-term steps, Gram-Schmidt passes,
// restart rule, convergence test and keep/evict rules; only the dense
// eigensolver of the projected matrix differs (Jacobi here, Householder + QL on the host), so
// results agree to the solver tolerance rather than bit for bit.
#include "dmed_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace dmed {

namespace {

constexpr int kQBlock = 1024;      // lanes per workgroup
constexpr int kQMaxM = 64;         // Krylov basis size limit (as trl_run)
constexpr int kQMaxNeig = 16;      // eigenpairs per sector
constexpr int kQMaxRows = kQMaxNeig + kQMaxM;
constexpr int kQWaves = kQBlock / 64;

// One sector of the queue.  Device pointers; the workspace slices are
// disjoint per task.
struct SqTask {
    const int32_t *hu_idx, *hd_idx;
    const double *hu_val, *hd_val;
    const double *du, *dd, *W;
    const uint8_t *impu, *impd;
    int32_t dup, ddw, nhu, nhd, nimp;
    int32_t dim;
    int32_t neig, m, maxit, lowest;
    double tol, degen;
    double *V, *Vs, *w, *Y, *X;         // (m+1)D, (m+1)D, D, neig*D, neig*D
    double **evecs;                     // neig output vectors (ascending energy)
    double *evals, *resid;              // [neig]
    int32_t *info;                      // nout, status, H*v count
};

// ---------------------------------------------------------- block helpers
struct Shared {
    double T[kQMaxM * kQMaxM];          // projected matrix, column-major
    double A[kQMaxM * kQMaxM];          // Jacobi work copy
    double S[kQMaxM * kQMaxM];          // Ritz rotation, column q = vector q
    double c[kQMaxRows], c2[kQMaxRows];
    double red[8 * kQWaves];
    double theta[kQMaxM];
    double jc[kQMaxM / 2], js[kQMaxM / 2];
    int order[kQMaxM];
    double th[kQMaxNeig], rr[kQMaxNeig];
    double ke[kQMaxNeig], kr[kQMaxNeig];
    int kslot[kQMaxNeig];
    double bcast;
    int ibcast;
};

__device__ __forceinline__ double wsum64(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// workgroup sum, the same value returned to every lane
__device__ double wg_sum(double v, Shared &sh)
{
    v = wsum64(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh.red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < kQWaves; ++k) s += sh.red[k];
        sh.bcast = s;
    }
    __syncthreads();
    return sh.bcast;
}

// counter-based hash in [-1, 1): the start vectors of krylov.hip
__device__ __forceinline__ double q_hash_uniform(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + 0x9e3779b97f4a7c15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// y = H x over the sector (the tile kernel's per-element sum)
__device__ void wg_hxv(const SqTask &t, const double *__restrict__ x, double *__restrict__ y)
{
    const int dup = t.dup, ddw = t.ddw, n = t.dim;
    for (int e = threadIdx.x; e < n; e += kQBlock) {
        const int id = e / dup, iu = e - id * dup;
        double acc = (t.du[iu] + t.dd[id] + t.W[t.impu[iu] * t.nimp + t.impd[id]]) * x[e];
        const int col = id * dup;
        for (int h = 0; h < t.nhu; ++h)
            acc += t.hu_val[h * dup + iu] * x[col + t.hu_idx[h * dup + iu]];
        for (int h = 0; h < t.nhd; ++h)
            acc += t.hd_val[h * ddw + id] * x[t.hd_idx[h * ddw + id] * dup + iu];
        y[e] = acc;
    }
    __syncthreads();
}

__device__ double wg_norm2(const double *__restrict__ x, int n, Shared &sh)
{
    double s = 0.0;
    for (int e = threadIdx.x; e < n; e += kQBlock) s += x[e] * x[e];
    return wg_sum(s, sh);
}

__device__ void wg_copy(double *__restrict__ d, const double *__restrict__ s, int n)
{
    for (int e = threadIdx.x; e < n; e += kQBlock) d[e] = s[e];
    __syncthreads();
}

// out[r] = <row r, w> for the rows Y[0..ny) then V[0..nv), eight at a time
__device__ void wg_dots(const double *Y, int ny, const double *V, int nv, const double *w, int n,
                        double *out, Shared &sh)
{
    const int nr = ny + nv;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int r0 = 0; r0 < nr; r0 += 8) {
        double s[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) s[q] = 0.0;
        for (int e = threadIdx.x; e < n; e += kQBlock) {
            const double we = w[e];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int r = r0 + q;
                if (r < nr) {
                    const double *b = r < ny ? Y + (size_t)r * n : V + (size_t)(r - ny) * n;
                    s[q] += b[e] * we;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) s[q] = wsum64(s[q]);
        __syncthreads();
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < 8; ++q) sh.red[q * kQWaves + wv] = s[q];
        }
        __syncthreads();
        if ((int)threadIdx.x < 8 && r0 + (int)threadIdx.x < nr) {
            double t = 0.0;
            for (int k = 0; k < kQWaves; ++k) t += sh.red[threadIdx.x * kQWaves + k];
            out[r0 + threadIdx.x] = t;
        }
        __syncthreads();
    }
}

// w -= sum_r c[r] row_r; returns the new ||w||^2
__device__ double wg_subtract(const double *Y, int ny, const double *V, int nv, double *w, int n,
                              const double *c, Shared &sh)
{
    double s2 = 0.0;
    for (int e = threadIdx.x; e < n; e += kQBlock) {
        double s = 0.0;
        for (int j = 0; j < ny; ++j) s += Y[(size_t)j * n + e] * c[j];
        for (int j = 0; j < nv; ++j) s += V[(size_t)j * n + e] * c[ny + j];
        const double x = w[e] - s;
        w[e] = x;
        s2 += x * x;
    }
    return wg_sum(s2, sh);
}

// Classical Gram-Schmidt of w against Y[0..nl) and V[0..jmax), repeated once
// when the first pass cancelled more than half of ||w||^2; the V coefficients
// of both passes are summed into cv[0..jmax).  Returns the final ||w||^2.
__device__ double wg_orth(const double *Y, int nl, const double *V, int jmax, double *w, int n,
                          double *cv, Shared &sh)
{
    const int rows = nl + jmax;
    if (rows <= 0) return wg_norm2(w, n, sh);
    const double n0 = wg_norm2(w, n, sh);
    wg_dots(Y, nl, V, jmax, w, n, sh.c, sh);
    const double n1 = wg_subtract(Y, nl, V, jmax, w, n, sh.c, sh);
    for (int r = threadIdx.x; r < jmax; r += kQBlock) cv[r] = sh.c[nl + r];
    __syncthreads();
    if (!(n1 < 0.5 * n0)) return n1;
    wg_dots(Y, nl, V, jmax, w, n, sh.c2, sh);
    const double n2 = wg_subtract(Y, nl, V, jmax, w, n, sh.c2, sh);
    for (int r = threadIdx.x; r < jmax; r += kQBlock) cv[r] += sh.c2[nl + r];
    __syncthreads();
    return n2;
}

// dst = w / sqrt(nrm2) (0 for a zero norm)
__device__ void wg_normalize(double *__restrict__ dst, const double *__restrict__ w, int n,
                             double nrm2)
{
    const double r = sqrt(nrm2);
    const double a = r > 0.0 ? 1.0 / r : 0.0;
    for (int e = threadIdx.x; e < n; e += kQBlock) dst[e] = a * w[e];
    __syncthreads();
}

// Vs[:, 0..k) = V[:, 0..m) S[0..m, 0..k), sixteen outputs per pass
__device__ void wg_rotate(const double *__restrict__ V, double *__restrict__ Vs, int n, int m,
                          int k, const double *S)
{
    for (int q0 = 0; q0 < k; q0 += 16) {
        for (int e = threadIdx.x; e < n; e += kQBlock) {
            double acc[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[q] = 0.0;
            for (int j = 0; j < m; ++j) {
                const double x = V[(size_t)j * n + e];
#pragma unroll
                for (int q = 0; q < 16; ++q)
                    if (q0 + q < k) acc[q] += x * S[(q0 + q) * m + j];
            }
#pragma unroll
            for (int q = 0; q < 16; ++q)
                if (q0 + q < k) Vs[(size_t)(q0 + q) * n + e] = acc[q];
        }
    }
    __syncthreads();
}

// Eigen-decomposition of the symmetric m x m matrix sh.A (column-major,
// destroyed): eigenvalues ascending in sh.theta, eigenvectors in the columns
// of sh.S.  Parallel cyclic Jacobi: m/2 disjoint rotations per round in the
// round-robin order, m - 1 rounds per sweep, at most 60 sweeps.
__device__ void wg_jacobi(int m, Shared &sh)
{
    double *A = sh.A, *S = sh.S;
    for (int e = threadIdx.x; e < m * m; e += kQBlock) S[e] = (e % m == e / m) ? 1.0 : 0.0;
    __syncthreads();
    const int mm = (m + 1) & ~1;           // players, with a dummy when m is odd
    const int np = mm / 2;
    for (int sweep = 0; sweep < 60 && m > 1; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int e = threadIdx.x; e < m * m; e += kQBlock) {
            const double a = A[e];
            if (e % m != e / m) off += a * a;
            tot += a * a;
        }
        off = wg_sum(off, sh);
        tot = wg_sum(tot, sh);
        if (off <= 1e-32 * tot) break;
        for (int r = 0; r < mm - 1; ++r) {
            if ((int)threadIdx.x < np) {
                const int k = threadIdx.x;
                int p, q;
                if (k == 0) {
                    p = mm - 1;
                    q = r;
                } else {
                    p = (r + k) % (mm - 1);
                    q = (r - k + (mm - 1)) % (mm - 1);
                }
                double c = 1.0, s = 0.0;
                if (p < m && q < m) {
                    const double apq = A[q * m + p];
                    const double app = fabs(A[p * m + p]), aqq = fabs(A[q * m + q]);
                    // a rotation below the diagonals' rounding only zeroes a_pq
                    if (app + 1e3 * fabs(apq) == app && aqq + 1e3 * fabs(apq) == aqq) {
                        A[q * m + p] = 0.0;
                        A[p * m + q] = 0.0;
                    } else if (apq != 0.0) {
                        const double tau = (A[q * m + q] - A[p * m + p]) / (2.0 * apq);
                        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                    }
                }
                sh.jc[k] = c;
                sh.js[k] = s;
                sh.order[2 * k] = p;
                sh.order[2 * k + 1] = q;
            }
            __syncthreads();
            // A <- A J (columns p, q of every pair)
            for (int it = threadIdx.x; it < np * m; it += kQBlock) {
                const int k = it / m, i = it - k * m;
                const int p = sh.order[2 * k], q = sh.order[2 * k + 1];
                if (p >= m || q >= m || sh.js[k] == 0.0) continue;
                const double c = sh.jc[k], s = sh.js[k];
                const double aip = A[p * m + i], aiq = A[q * m + i];
                A[p * m + i] = c * aip - s * aiq;
                A[q * m + i] = s * aip + c * aiq;
                const double vip = S[p * m + i], viq = S[q * m + i];
                S[p * m + i] = c * vip - s * viq;
                S[q * m + i] = s * vip + c * viq;
            }
            __syncthreads();
            // A <- J^T A (rows p, q of every pair)
            for (int it = threadIdx.x; it < np * m; it += kQBlock) {
                const int k = it / m, j = it - k * m;
                const int p = sh.order[2 * k], q = sh.order[2 * k + 1];
                if (p >= m || q >= m || sh.js[k] == 0.0) continue;
                const double c = sh.jc[k], s = sh.js[k];
                const double apj = A[j * m + p], aqj = A[j * m + q];
                A[j * m + p] = c * apj - s * aqj;
                A[j * m + q] = s * apj + c * aqj;
            }
            __syncthreads();
            // the rotated pair's off-diagonal element is zero by construction
            if ((int)threadIdx.x < np) {
                const int k = threadIdx.x;
                const int p = sh.order[2 * k], q = sh.order[2 * k + 1];
                if (p < m && q < m && sh.js[k] != 0.0) {
                    A[q * m + p] = 0.0;
                    A[p * m + q] = 0.0;
                }
            }
            __syncthreads();
        }
    }
    // ascending order: sort the indices, then permute the columns of S
    // through A, which is no longer needed
    if (threadIdx.x == 0) {
        for (int i = 0; i < m; ++i) sh.order[i] = i;
        for (int i = 1; i < m; ++i) {
            const int oi = sh.order[i];
            const double v = A[oi * m + oi];
            int j = i - 1;
            while (j >= 0 && A[sh.order[j] * m + sh.order[j]] > v) {
                sh.order[j + 1] = sh.order[j];
                --j;
            }
            sh.order[j + 1] = oi;
        }
        for (int i = 0; i < m; ++i) sh.theta[i] = A[sh.order[i] * m + sh.order[i]];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < m * m; e += kQBlock) A[e] = S[e];
    __syncthreads();
    for (int e = threadIdx.x; e < m * m; e += kQBlock) {
        const int q = e / m, j = e - q * m;
        S[q * m + j] = A[sh.order[q] * m + j];
    }
    __syncthreads();
}

// One thick-restart Lanczos run (trl_run): the `want` lowest eigenpairs of H
// on the complement of Y[0..nl).  Ritz values to sh.th, residual norms to
// sh.rr, Ritz vectors to X.  V and Vs swap roles at every rotation; the
// caller's pointers are updated.  Returns 0 or DMED_ENOCONV.
__device__ int wg_trl(const SqTask &t, int want, int m, int maxit, double tol, int nl,
                      uint64_t seed, double *&V, double *&Vs, int &used, Shared &sh)
{
    const int n = t.dim;
    const int avail = n - nl;
    m = max(m, want + 2);
    m = min(m, kQMaxM);
    m = min(m, avail);
    const bool full = m == avail;
    double *w = t.w;
    for (int e = threadIdx.x; e < n; e += kQBlock) w[e] = q_hash_uniform(seed, (uint64_t)e);
    __syncthreads();
    {
        const double nn = wg_orth(t.Y, nl, V, 0, w, n, sh.c2, sh);
        wg_normalize(V, w, n, nn);
    }
    for (int e = threadIdx.x; e < m * m; e += kQBlock) sh.T[e] = 0.0;
    __syncthreads();
    int k = 0;
    double beta_last = 0.0;
    bool converged = false;
    for (;;) {
        for (int j = k; j < m; ++j) {
            const double *vj = V + (size_t)j * n;
            wg_hxv(t, vj, w);
            ++used;
            // step j > k: the three-term step first, w -= alpha v_j +
            // beta_{j-1} v_{j-1} (as trl_run does), so that the full pass
            // seldom needs its second pass
            double al = 0.0;
            const double bprev = beta_last;
            if (j > k) {
                const double *vjm = vj - n;
                double s = 0.0;
                for (int e = threadIdx.x; e < n; e += kQBlock) s += vj[e] * w[e];
                al = wg_sum(s, sh);
                for (int e = threadIdx.x; e < n; e += kQBlock)
                    w[e] = w[e] - al * vj[e] - bprev * vjm[e];
                __syncthreads();
            }
            // coefficients on V[0..j] land in sh.A's first column as scratch
            double *cj = sh.A;
            const double nn = wg_orth(t.Y, nl, V, j + 1, w, n, cj, sh);
            if (j > k && threadIdx.x == 0) {
                cj[j] += al;
                cj[j - 1] += bprev;
            }
            __syncthreads();
            wg_normalize(V + (size_t)(j + 1) * n, w, n, nn);
            const double bn = sqrt(nn);
            if (threadIdx.x == 0) {
                for (int q = 0; q <= j; ++q) {
                    const bool arrow = (j == k) && (q < k);
                    if (arrow || q >= j - 1) sh.T[q * m + j] = sh.T[j * m + q] = cj[q];
                }
                if (j + 1 < m) sh.T[(j + 1) * m + j] = sh.T[j * m + j + 1] = bn;
            }
            beta_last = bn;
            __syncthreads();
        }
        for (int e = threadIdx.x; e < m * m; e += kQBlock) sh.A[e] = sh.T[e];
        __syncthreads();
        wg_jacobi(m, sh);
        bool ok = true;
        for (int q = 0; q < want; ++q) {
            const double rn = full ? 0.0 : fabs(beta_last * sh.S[q * m + (m - 1)]);
            if (rn > 0.1 * tol * fmax(1.0, fabs(sh.theta[q]))) ok = false;
        }
        if (ok || used >= maxit || full) {
            converged = ok || full;
            wg_rotate(V, Vs, n, m, want, sh.S);
            double *x = V;
            V = Vs;
            Vs = x;
            break;
        }
        // thick restart: kk Ritz vectors and the residual direction
        const int kk = min(m - 2, want + max(1, (m - want) / 2));
        wg_rotate(V, Vs, n, m, kk, sh.S);
        wg_copy(Vs + (size_t)kk * n, V + (size_t)m * n, n);
        double *x = V;
        V = Vs;
        Vs = x;
        for (int e = threadIdx.x; e < m * m; e += kQBlock) sh.T[e] = 0.0;
        __syncthreads();
        if (threadIdx.x == 0)
            for (int q = 0; q < kk; ++q) {
                sh.T[q * m + q] = sh.theta[q];
                const double arrow = beta_last * sh.S[q * m + (m - 1)];
                sh.T[q * m + kk] = sh.T[kk * m + q] = arrow;
            }
        __syncthreads();
        k = kk;
    }
    for (int q = 0; q < want; ++q) {
        const double *vq = V + (size_t)q * n;
        wg_hxv(t, vq, w);
        const double th = sh.theta[q];
        double s = 0.0;
        for (int e = threadIdx.x; e < n; e += kQBlock) {
            const double r = w[e] - th * vq[e];
            s += r * r;
        }
        const double r = sqrt(wg_sum(s, sh));
        if (threadIdx.x == 0) {
            sh.th[q] = th;
            sh.rr[q] = r;
        }
        if (r > tol * fmax(1.0, fabs(th))) converged = false;
        wg_copy(t.X + (size_t)q * n, vq, n);
    }
    __syncthreads();
    return converged ? 0 : DMED_ENOCONV;
}

// The deflation rounds of trl_deflated, then the kept vectors in ascending
// energy to the task's outputs.
__device__ void wg_solve(const SqTask &t, Shared &sh)
{
    const int n = t.dim, neig = t.neig;
    const int want = t.lowest ? 1 : neig;
    double *V = t.V, *Vs = t.Vs;
    int used = 0, rcfinal = 0, nk = 0;
    constexpr double confirm_tol = 1e-6;
    for (int round = 0; round <= neig; ++round) {
        const int nl = nk;
        if (n - nl < want) break;
        const double tol_r = (nl > 0 && confirm_tol > t.tol) ? confirm_tol : t.tol;
        const uint64_t seed = 1234567ull + 7919ull * (uint64_t)round;
        int rc = wg_trl(t, want, t.m, max(t.maxit - used, 4 * t.m), tol_r, nl, seed, V, Vs, used, sh);
        if (tol_r > t.tol) {
            bool reach = false;
            for (int q = 0; q < want; ++q) {
                const double lo = sh.th[q] - sh.rr[q];
                if (t.lowest ? (lo <= sh.ke[0] + t.degen) : (nk < neig || lo < sh.ke[nk - 1]))
                    reach = true;
            }
            if (!reach) break;
            rc = wg_trl(t, want, t.m, max(t.maxit - used, 4 * t.m), t.tol, nl, seed, V, Vs, used, sh);
        }
        if (rc == DMED_ENOCONV) rcfinal = rc;
        bool added = false;
        for (int q = 0; q < want; ++q) {
            int slot = -1;
            if (threadIdx.x == 0) {
                const double e = sh.th[q], r = sh.rr[q];
                if (t.lowest) {
                    if (!(nk > 0 && !(e <= sh.ke[0] + t.degen)) && nk < neig) slot = nk;
                } else if (nk >= neig) {
                    // nothing new below the top, else evict the highest level
                    if (e < sh.ke[nk - 1]) {
                        slot = sh.kslot[nk - 1];
                        --nk;
                    }
                } else {
                    unsigned busy = 0u;
                    for (int i = 0; i < nk; ++i) busy |= 1u << sh.kslot[i];
                    slot = __ffs(~busy) - 1;
                }
                if (slot >= 0) {
                    // insert in ascending energy (stable after equal energies)
                    int i = nk;
                    while (i > 0 && sh.ke[i - 1] > e) {
                        sh.ke[i] = sh.ke[i - 1];
                        sh.kr[i] = sh.kr[i - 1];
                        sh.kslot[i] = sh.kslot[i - 1];
                        --i;
                    }
                    sh.ke[i] = e;
                    sh.kr[i] = r;
                    sh.kslot[i] = slot;
                    ++nk;
                }
                sh.ibcast = slot;
                sh.bcast = (double)nk;
            }
            __syncthreads();
            slot = sh.ibcast;
            nk = (int)sh.bcast;
            __syncthreads();
            if (slot >= 0) {
                wg_copy(t.Y + (size_t)slot * n, t.X + (size_t)q * n, n);
                added = true;
            }
        }
        if (!added || used >= t.maxit) break;
        if (t.lowest && nk >= neig) break;
    }
    for (int q = 0; q < nk; ++q) wg_copy(t.evecs[q], t.Y + (size_t)sh.kslot[q] * n, n);
    if (threadIdx.x == 0) {
        for (int q = 0; q < nk; ++q) {
            t.evals[q] = sh.ke[q];
            t.resid[q] = sh.kr[q];
        }
        t.info[0] = nk;
        t.info[1] = rcfinal;
        t.info[2] = used;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kQBlock) k_sector_queue(const SqTask *__restrict__ tasks,
                                                          int ntask, int *__restrict__ counter)
{
    extern __shared__ __align__(16) unsigned char smem[];
    Shared &sh = *reinterpret_cast<Shared *>(smem);
    for (;;) {
        if (threadIdx.x == 0) sh.ibcast = atomicAdd(counter, 1);
        __syncthreads();
        const int k = sh.ibcast;
        __syncthreads();
        if (k >= ntask) break;
        const SqTask t = tasks[k];
        wg_solve(t, sh);
    }
}

}  // namespace

// ------------------------------------------------------------------ host
int64_t queue_max_dim()
{
    static const int64_t d = [] {
        const char *e = std::getenv("DMED_SECTOR_QUEUE_MAXDIM");
        return e ? (int64_t)std::atoll(e) : kQueueMaxDim;
    }();
    return d;
}

bool queue_takes(const dmed_sector *s, int neig, int ncv)
{
    return s->jterms == 0 && s->dim <= queue_max_dim() && s->dim < (int64_t)1 << 30 &&
           neig >= 1 && neig <= kQMaxNeig &&
           neig < s->dim && std::max(ncv, neig + 2) <= kQMaxM + 8;
}

int queue_launch(SectorQueue &q, hipStream_t st)
{
    const int nt = (int)q.items.size();
    if (nt == 0) return 0;
    size_t words = 0;
    for (auto &it : q.items) {
        const int64_t n = it.s->dim;
        const int m = std::min(std::max(it.ncv, it.neig + 2), kQMaxM);
        words += (size_t)(2 * (m + 1) + 1 + 2 * it.neig) * n + 64;
    }
    DMED_HIP(hipMalloc(&q.work, words * 8));
    DMED_HIP(hipMalloc(&q.res, (size_t)nt * (2 * kQMaxNeig + 4) * 8));
    DMED_HIP(hipMalloc(&q.ptrs, (size_t)nt * kQMaxNeig * sizeof(double *)));
    DMED_HIP(hipMalloc(&q.counter, sizeof(int)));
    std::vector<SqTask> tasks(nt);
    std::vector<double *> ptrs((size_t)nt * kQMaxNeig, nullptr);
    double *wp = q.work;
    for (int k = 0; k < nt; ++k) {
        auto &it = q.items[k];
        const dmed_sector *s = it.s;
        const int64_t n = s->dim;
        const int m = std::min(std::max(it.ncv, it.neig + 2), kQMaxM);
        SqTask &t = tasks[k];
        t.hu_idx = s->up.hidx;
        t.hd_idx = s->dw.hidx;
        t.hu_val = s->up.hval;
        t.hd_val = s->dw.hval;
        t.du = s->up.diag;
        t.dd = s->dw.diag;
        t.W = s->W;
        t.impu = s->up.imp;
        t.impd = s->dw.imp;
        t.dup = (int32_t)s->up.dim;
        t.ddw = (int32_t)s->dw.dim;
        t.nhu = s->up.nhop;
        t.nhd = s->dw.nhop;
        t.nimp = s->nimp;
        t.dim = (int32_t)n;
        t.neig = it.neig;
        t.m = m;
        t.maxit = it.maxit;
        t.lowest = it.lowest ? 1 : 0;
        t.tol = it.tol;
        t.degen = it.degen;
        t.V = wp;
        wp += (size_t)(m + 1) * n;
        t.Vs = wp;
        wp += (size_t)(m + 1) * n;
        t.w = wp;
        wp += n;
        t.Y = wp;
        wp += (size_t)it.neig * n;
        t.X = wp;
        wp += (size_t)it.neig * n + 64;
        it.vecs.assign(it.neig, nullptr);
        for (int q = 0; q < it.neig; ++q) {
            dmed_vec *v = new dmed_vec{s, DMED_FP64, n, nullptr};
            if (hipMalloc(&v->data, (size_t)std::max<int64_t>(n, 1) * 8) != hipSuccess) {
                delete v;
                return fail(DMED_ENOMEM, "sector queue: eigenvector allocation");
            }
            it.vecs[q] = v;
            ptrs[(size_t)k * kQMaxNeig + q] = (double *)v->data;
        }
        t.evecs = q.ptrs + (size_t)k * kQMaxNeig;
        double *r = q.res + (size_t)k * (2 * kQMaxNeig + 4);
        t.evals = r;
        t.resid = r + kQMaxNeig;
        t.info = reinterpret_cast<int32_t *>(r + 2 * kQMaxNeig);
    }
    DMED_HIP(hipMalloc(&q.tasks, (size_t)nt * sizeof(SqTask)));
    DMED_HIP(hipMemcpyAsync(q.tasks, tasks.data(), (size_t)nt * sizeof(SqTask), hipMemcpyHostToDevice, st));
    DMED_HIP(hipMemcpyAsync(q.ptrs, ptrs.data(), ptrs.size() * sizeof(double *), hipMemcpyHostToDevice, st));
    DMED_HIP(hipMemsetAsync(q.counter, 0, sizeof(int), st));
    DMED_HIP(hipMemsetAsync(q.res, 0, (size_t)nt * (2 * kQMaxNeig + 4) * 8, st));
    DMED_HIP(hipStreamSynchronize(st));     // the host task tables go out of scope
    static_assert(sizeof(Shared) <= 160 * 1024, "queue LDS");
    if (int rc = [&]() -> int {
            static const hipError_t e = hipFuncSetAttribute(
                (const void *)k_sector_queue, hipFuncAttributeMaxDynamicSharedMemorySize,
                (int)sizeof(Shared));
            return e == hipSuccess ? 0 : fail(DMED_EHIP, "sector queue: LDS attribute");
        }())
        return rc;
    const int ncu = std::max(8, ctx().num_cu);
    const unsigned grid = (unsigned)std::min(nt, ncu);
    DMED_HIP(hipEventCreate(&q.ev0));
    DMED_HIP(hipEventCreate(&q.ev1));
    DMED_HIP(hipEventRecord(q.ev0, st));
    hipLaunchKernelGGL(k_sector_queue, dim3(grid), dim3(kQBlock), sizeof(Shared), st,
                       (const SqTask *)q.tasks, nt, q.counter);
    DMED_HIP(hipGetLastError());
    DMED_HIP(hipEventRecord(q.ev1, st));
    q.stream = st;
    return 0;
}

int queue_finish(SectorQueue &q)
{
    const int nt = (int)q.items.size();
    int rc = 0;
    if (nt > 0 && q.stream) {
        std::vector<double> res((size_t)nt * (2 * kQMaxNeig + 4));
        if (hipMemcpyAsync(res.data(), q.res, res.size() * 8, hipMemcpyDeviceToHost, q.stream) !=
                hipSuccess ||
            hipStreamSynchronize(q.stream) != hipSuccess)
            rc = fail(DMED_EHIP, "sector queue: result copy");
        if (!rc && hipEventElapsedTime(&q.kernel_ms, q.ev0, q.ev1) != hipSuccess) q.kernel_ms = -1.f;
        for (int k = 0; k < nt && !rc; ++k) {
            auto &it = q.items[k];
            const double *r = res.data() + (size_t)k * (2 * kQMaxNeig + 4);
            int32_t info[3];
            std::memcpy(info, r + 2 * kQMaxNeig, sizeof info);
            it.nout = info[0];
            it.rc = info[1];
            it.used = info[2];
            it.evals.assign(r, r + it.nout);
            it.resid.assign(r + kQMaxNeig, r + kQMaxNeig + it.nout);
            for (int q2 = it.nout; q2 < it.neig; ++q2) {
                dmed_vec_free(it.vecs[q2]);
                it.vecs[q2] = nullptr;
            }
            it.vecs.resize(it.nout);
        }
    }
    if (nt > 0 && !q.stream)        // never launched: drop the eigenvector buffers
        for (auto &it : q.items) {
            for (dmed_vec *v : it.vecs) dmed_vec_free(v);
            it.vecs.clear();
            it.nout = 0;
            it.rc = DMED_EHIP;
        }
    if (q.ev0) hipEventDestroy(q.ev0);
    if (q.ev1) hipEventDestroy(q.ev1);
    q.ev0 = q.ev1 = nullptr;
    hipFree(q.work);
    hipFree(q.res);
    hipFree(q.ptrs);
    hipFree(q.counter);
    hipFree(q.tasks);
    q.work = q.res = nullptr;
    q.ptrs = nullptr;
    q.counter = nullptr;
    q.tasks = nullptr;
    q.stream = nullptr;
    return rc;
}

}  // namespace dmed
