/*
 * oracle.h -- CPU restatement of block-DDA's solve path, FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library never links or calls it.
 *
 * Every complex array is interleaved (re, im) double, i.e. the memory layout
 * of std::complex<double> / Eigen::VectorXcd::data().  Matrices are
 * column-major with leading dimension = rows.
 *
 * Reference files restated (paths relative to the reference root):
 *   mvp_fft/application_function.cpp:7-78     orc_application_function
 *   mvp_fft/BT_fft.cpp:8-53                   orc_bt_fft / orc_bt_fft_custom
 *   mvp_fft/BT_pad.cpp:6-45                   orc_bt_pad
 *   mvp_fft/BT_reconstruct.cpp:5-36           orc_bt_reconstruct
 *   mvp_fft/MBT_fft_init.cpp:10-42            orc_mbt_fft_init
 *   mvp_fft/MBT_fft_mvp.cpp:8-43              orc_mbt_fft_mvp
 *   bl_krylov_mvp_fft/qr_reduced.cpp:9-31     orc_qr_reduced
 *   bl_krylov_mvp_fft/bl_*_mvp_fft.cpp        orc_solve (6 methods)
 *   Solver.cpp:17-56                          orc_fft_length, DIAG_A convention
 * The FFTW 1-D transforms of the reference (Solver.cpp:23-24) are replaced
 * by this file's own mixed-radix + Bluestein FFT (orc_fft), unnormalised,
 * FFTW sign convention (forward = exp(-2 pi i jk/n)).
 */
#ifndef BDDA_ORACLE_H
#define BDDA_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* element function used by orc_bt_fft_custom: writes an f x f column-major
 * complex block for the 2x3 index matrix n_ind (row 0 = n_ind[0..2],
 * row 1 = n_ind[3..5]), exactly like application_function's contract. */
typedef void (*orc_elem_fn)(const int* n_ind, int f, void* ctx, double* out);

int64_t orc_fft_length(const int* n, int f);
int64_t orc_bt_pad_length(const int* n, int f);

int orc_application_function(const int* n_ind, int f, double lf, double kre,
                             double kim, double* out);
int orc_bt_fft(const int* n, int f, double lf, double kre, double kim,
               double* Au);
int orc_bt_fft_custom(const int* n, int f, orc_elem_fn fn, void* ctx,
                      double* Au);
int64_t orc_bt_pad(const int* n, int f, const double* x, double* xz);
int64_t orc_bt_reconstruct(const int* n, int f, const double* bz, double* b);

/* in-place unnormalised DFT of length len; sign = -1 forward, +1 backward */
int orc_fft(int64_t len, double* data, int sign);
/* transforms of at least fourstep_min points run as four-step FFTs on
 * `threads` OpenMP threads (outside parallel regions); defaults 2^24, 1 */
void orc_fft_config(int64_t fourstep_min, int threads);

int orc_mbt_fft_init(const int* n, int f, double lf, double kre, double kim,
                     double* Au_til);
int orc_mbt_fft_mvp(const int* n, int f, const double* Au_til,
                    const double* p_hat, double* q_hat);

/* the inline block matvec every solver repeats (e.g.
 * bl_cocg_rq_jacobi_mvp_fft.cpp:43-60): Y = DIAG.*X + gather(MBT(scatter X)),
 * then Y = Y ./ DIAG when jacobi != 0.  X, Y are (f*nocc) x L. */
int orc_block_matvec(const int* n, int f, const int* address, int64_t nocc,
                     const double* Au_til, const double* diag, int jacobi,
                     const double* X, double* Y, int L, int nthreads);

int orc_qr_reduced(const double* A, int64_t m, int ncol, double* Q, double* R);
/* X (s x nb) = A^{-1} B, partial (full = 0) or full (full = 1) pivoting: the
 * s x s solves of the block solvers (Eigen PartialPivLU / FullPivLU .solve()). */
void orc_lu_solve(const double* A, const double* B, int s, int nb, int full, double* X);

enum { ORC_COCG_RQ = 0, ORC_BICGSTAB = 1, ORC_BICGGR = 2 };
/* Runs one of the six reference solvers.  hist (length itermax, may be NULL)
 * receives the per-iteration "relative err" the reference prints.  Returns
 * the number of iterations executed through *iters (the index of the
 * iteration that broke the loop + 1, or itermax). */
int orc_solve(int method, int jacobi, const int* n, int f, const int* address,
              int64_t nocc, const double* Au_til, const double* diag,
              const double* B, int L, double tol, int itermax, double* X,
              int* iters, double* hist, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
