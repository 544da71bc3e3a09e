// smallmat.hip -- the solvers' s x s algebra on the device (s <= 64).
//
// The block Krylov recurrences need, per iteration, a few s x s FP64
// products, LU solves (PartialPivLU / FullPivLU where the reference calls
// Eigen: bl_cocg_rq_mvp_fft.cpp:56,66, bl_cocg_rq_jacobi_mvp_fft.cpp:63,73,
// bl_bicgstab_jacobi_mvp_fft.cpp:62,93, bl_bicggr_jacobi_mvp_fft.cpp:58,113)
// and, inside the Cholesky QR, a Cholesky factor and its inverse.  Each runs
// here as one work-group on matrices that stay in HBM between steps, so an
// iteration is enqueued without a host round trip; a failure (exactly zero
// LU pivot, Cholesky pivot below its floor) sets a bit of a device flag word
// that the solver reads with its convergence test, once per iteration.
//
// All matrices are complex FP64, column-major with leading dimension equal
// to their row count.  The LU follows the host routine's order: pivot by the
// largest modulus (first in column-major order on ties), rows (and columns
// for full pivoting) swapped, multipliers a_ik / pivot stored below the
// diagonal, right-looking update, then forward and back substitution.
#include "blockops.hpp"
#include <algorithm>

namespace bdda {

namespace {

constexpr int SM_NT = 256;
constexpr int SM_MAX = 64;

__device__ __forceinline__ c64 cmul(c64 a, c64 b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ c64 cdivd(c64 a, c64 b)
{
    const double den = b.x * b.x + b.y * b.y;
    return mk((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}

// X (s x nb) = A^{-1} (bscale B [/ *bdiv]); on an exactly zero pivot X = 0
// and flag |= bit.  Once a solve has flagged an LU breakdown every later
// solve also returns X = 0 (and sets its bit), so the iterates stop changing
// until the solver reads the flag.  A and B are read once into LDS; X may
// alias B.
__global__ __launch_bounds__(SM_NT) void k_lu_solve(const c64* __restrict__ A, const c64* B, c64* X, int s, int nb,
                                                    int full, c64 bscale, const c64* __restrict__ bdiv,
                                                    int* __restrict__ flag, int bit)
{
    if (*flag & BDDA_SM_LU_ANY) {
        for (int e = threadIdx.x; e < s * nb; e += SM_NT) X[e] = mk(0.0, 0.0);
        if (threadIdx.x == 0) atomicOr(flag, bit);
        return;
    }
    if (bdiv) bscale = cdivd(bscale, *bdiv);
    __shared__ c64 a[SM_MAX * SM_MAX];
    __shared__ c64 x[SM_MAX * SM_MAX];
    __shared__ int rp[SM_MAX], cp[SM_MAX];
    __shared__ double bestv[SM_NT / 64];
    __shared__ int besti[SM_NT / 64];
    __shared__ int piv[2];
    __shared__ int fail;
    const int t = threadIdx.x;
    for (int e = t; e < s * s; e += SM_NT) a[e] = A[e];
    for (int e = t; e < s * nb; e += SM_NT) x[e] = cmul(bscale, B[e]);
    if (t < s) rp[t] = cp[t] = t;
    if (t == 0) fail = 0;
    __syncthreads();
    for (int k = 0; k < s; ++k) {
        // pivot: largest modulus over rows i >= k of column k (partial) or the
        // trailing block (full); ties to the first in column-major order
        const int ncol = full ? s - k : 1, nrow = s - k;
        double bv = -1.0;
        int bi = 0x7fffffff;
        for (int e = t; e < ncol * nrow; e += SM_NT) {
            const int jj = e / nrow, ii = e % nrow;
            const int i = k + ii, j = k + jj;
            const c64 v = a[i + j * s];
            const double m = sqrt(v.x * v.x + v.y * v.y);
            const int key = j * s + i;
            if (m > bv || (m == bv && key < bi)) {
                bv = m;
                bi = key;
            }
        }
        // (largest modulus, then smallest key) is a total order, so any
        // reduction order picks the same pivot: lane shuffles within each
        // wave, then the waves' winners in LDS (2 barriers per column instead
        // of a barrier per halving over the group)
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(bv, off);
            const int oi = __shfl_xor(bi, off);
            if (o > bv || (o == bv && oi < bi)) {
                bv = o;
                bi = oi;
            }
        }
        if ((t & 63) == 0) {
            bestv[t >> 6] = bv;
            besti[t >> 6] = bi;
        }
        __syncthreads();
        if (t == 0) {
            for (int w = 1; w < SM_NT / 64; ++w) {
                if (bestv[w] > bv || (bestv[w] == bv && besti[w] < bi)) {
                    bv = bestv[w];
                    bi = besti[w];
                }
            }
            // (no candidate: every modulus NaN; the pivot a_kk then fails below)
            const bool none = bi == 0x7fffffff;
            piv[0] = none ? k : bi % s;  // row
            piv[1] = none ? k : bi / s;  // column
        }
        __syncthreads();
        const int pr = piv[0], pc = piv[1];
        if (pr != k) {
            for (int j = t; j < s; j += SM_NT) {
                const c64 u = a[k + j * s];
                a[k + j * s] = a[pr + j * s];
                a[pr + j * s] = u;
            }
            for (int c = t; c < nb; c += SM_NT) {  // the right-hand side follows the row swap
                const c64 u = x[k + c * s];
                x[k + c * s] = x[pr + c * s];
                x[pr + c * s] = u;
            }
            if (t == 0) {
                const int u = rp[k];
                rp[k] = rp[pr];
                rp[pr] = u;
            }
        }
        __syncthreads();
        if (pc != k) {
            for (int i = t; i < s; i += SM_NT) {
                const c64 u = a[i + k * s];
                a[i + k * s] = a[i + pc * s];
                a[i + pc * s] = u;
            }
            if (t == 0) {
                const int u = cp[k];
                cp[k] = cp[pc];
                cp[pc] = u;
            }
        }
        __syncthreads();
        const c64 p = a[k + k * s];
        if (!(p.x * p.x + p.y * p.y > 0)) {
            if (t == 0) fail = 1;
            break;  // uniform: every thread read the same pivot
        }
        for (int i = k + 1 + t; i < s; i += SM_NT) a[i + k * s] = cdivd(a[i + k * s], p);
        __syncthreads();
        for (int e = t; e < (s - k - 1) * (s - k - 1); e += SM_NT) {
            const int i = k + 1 + e % (s - k - 1), j = k + 1 + e / (s - k - 1);
            const c64 l = a[i + k * s], u = a[k + j * s];
            c64 v = a[i + j * s];
            v.x -= l.x * u.x - l.y * u.y;
            v.y -= l.x * u.y + l.y * u.x;
            a[i + j * s] = v;
        }
        __syncthreads();
    }
    __syncthreads();
    if (fail) {
        for (int e = t; e < s * nb; e += SM_NT) X[e] = mk(0.0, 0.0);
        if (t == 0) atomicOr(flag, bit);
        return;
    }
    // substitution: one thread per right-hand-side column
    for (int c = t; c < nb; c += SM_NT) {
        c64* y = x + c * s;
        for (int j = 0; j < s; ++j) {  // unit lower triangle
            const c64 v = y[j];
            for (int i = j + 1; i < s; ++i) {
                const c64 l = a[i + j * s];
                y[i].x -= l.x * v.x - l.y * v.y;
                y[i].y -= l.x * v.y + l.y * v.x;
            }
        }
        for (int j = s - 1; j >= 0; --j) {  // upper triangle
            const c64 v = cdivd(y[j], a[j + j * s]);
            y[j] = v;
            for (int i = 0; i < j; ++i) {
                const c64 u = a[i + j * s];
                y[i].x -= u.x * v.x - u.y * v.y;
                y[i].y -= u.x * v.y + u.y * v.x;
            }
        }
    }
    __syncthreads();
    // undo the column permutation: X[cp[i]] = y[i]
    for (int e = t; e < s * nb; e += SM_NT) {
        const int i = e % s, c = e / s;
        X[cp[i] + c * s] = x[i + c * s];
    }
}

// C (m x n) = alpha op(A) B (+ beta C0 when C0 is given), op(A) = A (m x k)
// or A^T (A: k x m).  C may alias neither A nor B; C0 may alias C.
__global__ __launch_bounds__(SM_NT) void k_small_mul(const c64* __restrict__ A, int transA, const c64* __restrict__ B,
                                                     c64* C, int m, int k, int n, c64 alpha,
                                                     const c64* __restrict__ alpha_dev, const c64* C0, c64 beta)
{
    c64 al = alpha;
    if (alpha_dev) al = cmul(al, *alpha_dev);
    for (int e = threadIdx.x; e < m * n; e += SM_NT) {
        const int i = e % m, j = e / m;
        double sx = 0, sy = 0;
        for (int l = 0; l < k; ++l) {
            const c64 a = transA ? A[l + i * k] : A[i + l * m];
            const c64 b = B[l + j * k];
            sx += a.x * b.x - a.y * b.y;
            sy += a.x * b.y + a.y * b.x;
        }
        c64 v = cmul(al, mk(sx, sy));
        if (C0) v = v + cmul(beta, C0[e]);
        C[e] = v;
    }
}

// Upper Cholesky factor U (G = U^H U) of the Hermitian s x s G and its
// inverse; a pivot d <= floor * G_jj (or not finite) sets flag |= bit and
// leaves U and Ui undefined
__global__ __launch_bounds__(SM_NT) void k_chol_inv(const c64* __restrict__ G, c64* __restrict__ U,
                                                    c64* __restrict__ Ui, int s, double rel_floor,
                                                    int* __restrict__ flag, int bit)
{
    __shared__ c64 u[SM_MAX * SM_MAX];
    __shared__ c64 ui[SM_MAX * SM_MAX];  // the inverse, built in LDS and then stored
    __shared__ int fail;
    const int t = threadIdx.x;
    for (int e = t; e < s * s; e += SM_NT) u[e] = mk(0.0, 0.0);
    if (t == 0) fail = 0;
    __syncthreads();
    for (int j = 0; j < s; ++j) {
        if (t == 0) {
            double d = G[j + j * s].x;
            const double gjj = d;
            for (int k = 0; k < j; ++k) {
                const c64 v = u[k + j * s];
                d -= v.x * v.x + v.y * v.y;
            }
            if (!(d > rel_floor * gjj) || !isfinite(d)) fail = 1;
            else u[j + j * s] = mk(sqrt(d), 0.0);
        }
        __syncthreads();
        if (fail) break;
        const double ujj = u[j + j * s].x;
        for (int i = j + 1 + t; i < s; i += SM_NT) {
            c64 v = G[j + i * s];
            for (int k = 0; k < j; ++k) {
                const c64 a = u[k + j * s], b = u[k + i * s];
                v.x -= a.x * b.x + a.y * b.y;
                v.y -= a.x * b.y - a.y * b.x;
            }
            u[j + i * s] = mk(v.x / ujj, v.y / ujj);
        }
        __syncthreads();
    }
    if (fail) {
        if (t == 0) atomicOr(flag, bit);
        return;
    }
    for (int e = t; e < s * s; e += SM_NT) U[e] = u[e];
    // inverse by back substitution, one thread per column of Ui; each step
    // reads the column's earlier entries, so the column is built in LDS
    // (through global memory every step would wait for a store to come back)
    for (int j = t; j < s; j += SM_NT) {
        for (int i = j + 1; i < s; ++i) ui[i + j * s] = mk(0.0, 0.0);
        for (int i = j; i >= 0; --i) {
            c64 v = (i == j) ? mk(1.0, 0.0) : mk(0.0, 0.0);
            for (int k = i + 1; k <= j; ++k) {
                const c64 a = u[i + k * s], b = ui[k + j * s];
                v.x -= a.x * b.x - a.y * b.y;
                v.y -= a.x * b.y + a.y * b.x;
            }
            ui[i + j * s] = cdivd(v, u[i + i * s]);
        }
    }
    __syncthreads();
    for (int e = t; e < s * s; e += SM_NT) Ui[e] = ui[e];
}

// out = sqrt(sum |A_e|^2) over n entries, in a fixed order
__global__ __launch_bounds__(SM_NT) void k_small_fro(const c64* __restrict__ A, int n, double* __restrict__ out)
{
    __shared__ double sh[SM_NT];
    double v = 0;
    for (int e = threadIdx.x; e < n; e += SM_NT) v += A[e].x * A[e].x + A[e].y * A[e].y;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = SM_NT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sqrt(sh[0]);
}

// out = mult num / den (complex scalars on the device); zero once an LU
// breakdown is flagged, so the block updates it scales stop
__global__ void k_scalar_div(const c64* __restrict__ num, const c64* __restrict__ den, c64 mult, c64* __restrict__ out,
                             const int* __restrict__ flag)
{
    if (threadIdx.x == 0) *out = (*flag & BDDA_SM_LU_ANY) ? mk(0.0, 0.0) : cmul(mult, cdivd(*num, *den));
}

void check_small(int s, int nb)
{
    if (s < 1 || s > SM_MAX || nb < 0 || nb > SM_MAX) throw HipError{hipErrorInvalidValue, "s x s algebra: s > 64"};
}

}  // namespace

void dm_lu_solve(hipStream_t st, const c64* A, const c64* B, c64* X, int s, int nb, int full, c64 bscale,
                 const c64* bdiv, int* flag, int bit)
{
    check_small(s, nb);
    hipLaunchKernelGGL(k_lu_solve, dim3(1), dim3(SM_NT), 0, st, A, B, X, s, nb, full, bscale, bdiv, flag, bit);
    BDDA_HIP(hipGetLastError());
}

void dm_mul(hipStream_t st, const c64* A, int transA, const c64* B, c64* C, int m, int k, int n, c64 alpha,
            const c64* alpha_dev, const c64* C0, c64 beta)
{
    check_small(std::max(m, std::max(k, n)), 0);
    hipLaunchKernelGGL(k_small_mul, dim3(1), dim3(SM_NT), 0, st, A, transA, B, C, m, k, n, alpha, alpha_dev, C0, beta);
    BDDA_HIP(hipGetLastError());
}

void dm_chol_inv(hipStream_t st, const c64* G, c64* U, c64* Ui, int s, double rel_floor, int* flag, int bit)
{
    check_small(s, 0);
    hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(SM_NT), 0, st, G, U, Ui, s, rel_floor, flag, bit);
    BDDA_HIP(hipGetLastError());
}

void dm_fro(hipStream_t st, const c64* A, int n, double* out)
{
    hipLaunchKernelGGL(k_small_fro, dim3(1), dim3(SM_NT), 0, st, A, n, out);
    BDDA_HIP(hipGetLastError());
}

void dm_scalar_div(hipStream_t st, const c64* num, const c64* den, c64 mult, c64* out, const int* flag)
{
    hipLaunchKernelGGL(k_scalar_div, dim3(1), dim3(64), 0, st, num, den, mult, out, flag);
    BDDA_HIP(hipGetLastError());
}

}  // namespace bdda
