// blockops.hpp -- block-vector kernels for the block Krylov solvers.
//
// A block vector is rows x s complex values, column-major with leading
// dimension ld (the layout of Eigen::MatrixXcd in the reference solvers,
// e.g. bl_cocg_rq_mvp_fft.cpp:23-35).  Reductions accumulate in FP64 for
// either storage precision; small s x s results come back to the host in
// FP64 where the reference's Eigen LU / products run (host side, FP64).
#pragma once
#include "bdda_common.hpp"
#include "devmem.hpp"
#include <vector>

namespace bdda {

using HostMat = std::vector<c64>;  // column-major s x s (or a x b)

struct BlockWork {
    hipStream_t stream = nullptr;
    c64* partial = nullptr;    // device partial sums
    size_t partial_elems = 0;
    c64* red = nullptr;        // device reduced result
    size_t red_elems = 0;
    c64* coef = nullptr;       // device coefficient staging (up to 4 x 64 x 64)
    c64* host_pinned = nullptr;
    size_t host_elems = 0;
    int coef_slot = 0;         // rotating staging slot for combo coefficients
    ~BlockWork();
    void reserve(size_t partial_n, size_t red_n);
    // device buffers back to the allocator (after the stream's work); call
    // before the stream is destroyed
    void release();
};

// C(a x b) = op(A)^ B with op = transpose (conjA = 0) or adjoint (conjA = 1),
// A rows x a (lda), B rows x b (ldb).  Result on the host in FP64.
template <typename R>
HostMat gram(BlockWork& w, const cplx<R>* A, int64_t lda, int a, const cplx<R>* B, int64_t ldb, int b,
             int64_t rows, int conjA);

// sum over all entries of conj(A) .* B (conjA = 1) or A .* B; rows x s
template <typename R>
c64 dot_all(BlockWork& w, const cplx<R>* A, int64_t lda, const cplx<R>* B, int64_t ldb, int64_t rows, int s,
            int conjA);

template <typename R>
double norm_fro(BlockWork& w, const cplx<R>* A, int64_t lda, int64_t rows, int s);

// no NaN or Inf among the n values of A (device, FP64 complex)
bool all_finite(BlockWork& w, const c64* A, int64_t n);

// Z = sum_t a_t V_t + sum_u W_u M_u  (M_u: FP64, k_u x s, column-major, on the
// host (M) or on the device (Md, scaled by ms_u and, when msd_u is given, by the
// device scalar *msd_u); a_t is scaled by the device scalar *ad_t when given).
// Z may alias a V_t but never a W_u.
template <typename R>
struct Combo {
    int nv = 0;
    const cplx<R>* V[3];
    int64_t ldv[3];
    c64 a[3];
    const c64* ad[3] = {nullptr, nullptr, nullptr};
    int nm = 0;
    const cplx<R>* W[2];
    int64_t ldw[2];
    int kw[2];
    const c64* M[2] = {nullptr, nullptr};   // host pointers
    const c64* Md[2] = {nullptr, nullptr};  // device pointers
    c64 ms[2] = {{1.0, 0.0}, {1.0, 0.0}};
    const c64* msd[2] = {nullptr, nullptr};
};
template <typename R>
void combo(BlockWork& w, const Combo<R>& c, cplx<R>* Z, int64_t ldz, int64_t rows, int s);

// elementwise: Z[:, j] = d .* A[:, j]  (d has `rows` entries)
template <typename R>
void scale_rows(hipStream_t st, const cplx<R>* d, const cplx<R>* A, int64_t lda, cplx<R>* Z, int64_t ldz,
                int64_t rows, int s);

// precision conversions (block vectors)
void convert_64_to_32(hipStream_t st, const c64* A, int64_t lda, c32* Z, int64_t ldz, int64_t rows, int s,
                      double scale);
void accumulate_32_to_64(hipStream_t st, const c32* A, int64_t lda, c64* Z, int64_t ldz, int64_t rows, int s,
                         double scale);

template <typename R>
void fill_zero(hipStream_t st, cplx<R>* Z, int64_t ldz, int64_t rows, int s);

// reduced QR by column-sequential classical Gram-Schmidt with one
// re-orthogonalisation pass (a whole column's projections per Gram launch);
// the solvers' fallback when Cholesky QR fails.  Q may alias A.  Returns
// BDDA_OK or BDDA_EBREAKDOWN (a numerically dependent column).
template <typename R>
int qr_cgs2(BlockWork& w, const cplx<R>* A, int64_t lda, cplx<R>* Q, int64_t ldq, int64_t rows, int s,
            HostMat& Rm);

// qr_reduced exactly as the reference orders it: modified Gram-Schmidt with
// one re-orthogonalisation pass, R(i,k) from the first pass
// (qr_reduced.cpp:9-31).  O(s^2) launches.  Q may alias A.
template <typename R>
int qr_mgs2(BlockWork& w, const cplx<R>* A, int64_t lda, cplx<R>* Q, int64_t ldq, int64_t rows, int s,
            HostMat& Rm);

// reduced QR by Cholesky QR applied twice (two Gram + two update sweeps);
// equal in exact arithmetic to the MGS factorisation with positive real
// diag(R).  Falls back to qr_cgs2 when a Cholesky pivot fails.  scratch is
// rows x s (ld = rows) and must not alias A or Q; Q may alias A.
template <typename R>
int qr_fast(BlockWork& w, const cplx<R>* A, int64_t lda, cplx<R>* Q, int64_t ldq, int64_t rows, int s,
            HostMat& Rm, cplx<R>* scratch);

// ---------------- s x s algebra on the device (smallmat.hip): FP64, s <= 64,
// one work-group per call, nothing returns to the host.  Failures set a bit of
// the device flag word: BDDA_SM_LU / BDDA_SM_LU2 (exactly zero LU pivot of the
// solve the caller tags with it; X is then zero) or BDDA_SM_CHOL (Cholesky
// pivot below its floor).  After an LU breakdown every later LU solve and
// scalar quotient returns zero, so the block updates stop.
constexpr int BDDA_SM_LU = 1, BDDA_SM_CHOL = 2, BDDA_SM_LU2 = 4, BDDA_SM_LU_ANY = BDDA_SM_LU | BDDA_SM_LU2;
// X (s x nb) = A^{-1} (bscale B [/ *bdiv]), partial (full = 0) or full
// pivoting; X may alias B
void dm_lu_solve(hipStream_t st, const c64* A, const c64* B, c64* X, int s, int nb, int full, c64 bscale,
                 const c64* bdiv, int* flag, int bit);
// C (m x n) = alpha [*alpha_dev] op(A) B [+ beta C0]; op(A) = A (m x k) or A^T (A: k x m)
void dm_mul(hipStream_t st, const c64* A, int transA, const c64* B, c64* C, int m, int k, int n, c64 alpha,
            const c64* alpha_dev, const c64* C0, c64 beta);
// U = chol_upper(G) (G = U^H U), Ui = U^{-1}; pivot d <= rel_floor G_jj fails
void dm_chol_inv(hipStream_t st, const c64* G, c64* U, c64* Ui, int s, double rel_floor, int* flag, int bit);
// *out = ||A||_F over n entries
void dm_fro(hipStream_t st, const c64* A, int n, double* out);
// *out = mult * (*num) / (*den), or zero after an LU breakdown
void dm_scalar_div(hipStream_t st, const c64* num, const c64* den, c64 mult, c64* out, const int* flag);

// Gram C (a x b) = op(A)^ B and the sum of conj(A) .* B (conjA) or A .* B,
// reduced in FP64 into device memory (stream-ordered, no host wait)
template <typename R>
void gram_dev(BlockWork& w, const cplx<R>* A, int64_t lda, int a, const cplx<R>* B, int64_t ldb, int b, int64_t rows,
              int conjA, c64* out);
template <typename R>
void dot_all_dev(BlockWork& w, const cplx<R>* A, int64_t lda, const cplx<R>* B, int64_t ldb, int64_t rows, int s,
                 int conjA, c64* out);

// qr_fast with its s x s algebra on the device: R (s x s) to device memory,
// tmp holds 4 s x s matrices.  A Cholesky pivot below its floor sets
// BDDA_SM_CHOL in *flag (Q and R are then undefined; the caller redoes the
// factorisation with qr_fast on the host path).
template <typename R>
void qr_fast_dev(BlockWork& w, const cplx<R>* A, int64_t lda, cplx<R>* Q, int64_t ldq, int64_t rows, int s, c64* Rd,
                 cplx<R>* scratch, c64* tmp, int* flag);

// streaming copy, read and write rates (GB/s, best of five) over two device
// buffers of `bytes` each: the HBM bandwidth a plain kernel reaches
void hbm_probe(void* a, void* b, size_t bytes, double* copy_gbs, double* read_gbs, double* write_gbs);

// ---------------- small dense FP64 algebra on the host (s <= 64)
HostMat hm_mul(const HostMat& A, const HostMat& B, int m, int k, int n);       // A (m x k) B (k x n)
HostMat hm_tmul(const HostMat& A, const HostMat& B, int k, int m, int n);      // A^T (A: k x m) B (k x n)
// X = A^{-1} B, partial (full = 0) or full (full = 1) pivoting; returns
// false when a pivot is exactly zero
bool hm_lu_solve(const HostMat& A, const HostMat& B, int s, int nb, int full, HostMat& X);
double hm_fro(const HostMat& A);

}  // namespace bdda
