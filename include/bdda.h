/*
 * bdda.h -- C ABI of the MI355X (gfx950) block-DDA solve path.
 *
 * The reference exposes this path as free C++ functions over Eigen types and
 * FFTW plans (mvp_fft/MBT_utils.hpp:1-6, DDA_linear_solvers.hpp:1-6,
 * linear_algebra_addon.hpp:1).  This header carries the same operations over
 * plain pointers so that any host language can bind them:
 *   - complex values are interleaved (re, im) doubles: the memory layout of
 *     std::complex<double>, Eigen::VectorXcd::data(), numpy complex128;
 *   - matrices are column-major with an explicit leading dimension (Eigen's
 *     default MatrixXcd layout);
 *   - the cuboid vector layout is the reference's: element (ix, iy, iz, q)
 *     at f*(ny*nz*ix + nz*iy + iz) + q; occupied vectors are (f*N) x L with
 *     dipole m at rows f*m .. f*m+f-1, address[m] = ny*nz*ix + nz*iy + iz
 *     (0-based, Target_manager.cpp:98,167-169);
 *   - every entry point returns a status code (BDDA_OK = 0) and never exits;
 *     bdda_last_error() describes the last failure on the calling thread.
 * One handle = one HIP stream; calls on the same handle must not overlap.
 */
#ifndef BDDA_H
#define BDDA_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double re, im;
} bdda_c64;

typedef struct bdda_mbt bdda_mbt;

enum {
    BDDA_OK = 0,
    BDDA_EINVAL = 1,
    BDDA_ENOMEM = 2,
    BDDA_EHIP = 3,
    BDDA_EBREAKDOWN = 4,
    BDDA_ENOCONV = 5,
    BDDA_ENAN = 6,
    BDDA_EUNSUPPORTED = 7
};

/* precision masks / modes */
enum { BDDA_PREC_FP64 = 1, BDDA_PREC_FP32 = 2 };
enum { BDDA_MODE_FP64 = 0, BDDA_MODE_FP32 = 1, BDDA_MODE_MIXED = 2, BDDA_MODE_MIXED_COLUMNS = 3 };
enum { BDDA_COCG_RQ = 0, BDDA_BICGSTAB = 1, BDDA_BICGGR = 2 };

typedef struct {
    int iters;             /* block iterations (summed over defect-correction steps) */
    int outer_iters;       /* defect-correction steps (1 for FP64 / FP32 modes) */
    int matvecs;           /* block matvecs (each covers all L columns) */
    int converged;         /* 1 when the stopping test was met */
    int status;            /* BDDA_* of the solve */
    double recurrence_res; /* last residual as the reference computes it */
    double true_res;       /* ||B - A X||_F / ||B||_F in FP64 (-1 if not computed) */
    double time_ms;        /* wall time of the solve on the host clock */
} bdda_solve_info;

const char* bdda_last_error(void);
int bdda_version(void);

/* ---------------------------------------------------------------- layout
 * CPU reference-layout helpers (no GPU needed).
 * application_function (mvp_fft/application_function.cpp:7-78): f x f
 *   column-major block for the 2x3 index matrix n_ind (row-major: row 0 =
 *   n_ind[0..2], row 1 = n_ind[3..5]).
 * BT_fft (mvp_fft/BT_fft.cpp:8-53): Au, length f^2 prod(2n-1).
 * BT_pad (mvp_fft/BT_pad.cpp:6-45): xz, length f^2 nx (2ny-1)(2nz-1).
 * BT_reconstruct (mvp_fft/BT_reconstruct.cpp:5-36): b, length f nx ny nz. */
int64_t bdda_fft_length(const int n[3], int f);
int bdda_application_function(const int n_ind[6], int f, double lf, bdda_c64 k, bdda_c64* out);
int bdda_bt_fft(const int n[3], int f, double lf, bdda_c64 k, bdda_c64* Au);
int bdda_bt_pad(const int n[3], int f, const bdda_c64* x, bdda_c64* xz);
int bdda_bt_reconstruct(const int n[3], int f, const bdda_c64* bz, bdda_c64* b);

/* ---------------------------------------------------------------- operator
 * bdda_mbt_create replaces Solver::fft_prep + MBT_fft_init
 * (Solver.cpp:17-25,55; mvp_fft/MBT_fft_init.cpp:10-42): it builds the
 * transformed interaction kernel on the device.  prec_mask selects which
 * precisions (BDDA_PREC_*) the operator can apply.  f is 3 or 6. */
int bdda_mbt_create(const int n[3], int f, double lf, bdda_c64 k, int prec_mask, bdda_mbt** out);
void bdda_mbt_destroy(bdda_mbt* h);
/* padded FFT grid chosen for the operator */
int bdda_mbt_grid(const bdda_mbt* h, int P[3]);

/* MBT_fft_mvp (mvp_fft/MBT_fft_mvp.cpp:8-43) on ncols cuboid vectors
 * (host buffers, f*nx*ny*nz rows each): y = offdiag(A) x.  prec is
 * BDDA_PREC_FP64 or BDDA_PREC_FP32 (inputs/outputs stay complex<double>). */
int bdda_mbt_apply(bdda_mbt* h, int prec, const bdda_c64* x, int64_t ldx, bdda_c64* y, int64_t ldy, int ncols);

/* Occupancy and DIAG_A of the target (Solver.cpp:28-47: DIAG_A = 1/alpha
 * per component).  diag has f*nocc entries. */
int bdda_mbt_set_target(bdda_mbt* h, const int32_t* address, int64_t nocc, const bdda_c64* diag);

/* The inline block matvec every reference solver repeats
 * (bl_cocg_rq_jacobi_mvp_fft.cpp:43-60): Y = DIAG_A .* X + gather(MBT(
 * scatter(X))), then Y = jpre .* Y with jpre = 1/DIAG_A when jacobi != 0. */
int bdda_mbt_apply_occupied(bdda_mbt* h, int prec, int jacobi, const bdda_c64* X, int64_t ldx, bdda_c64* Y,
                            int64_t ldy, int ncols);

/* Block solvers (DDA_linear_solvers.hpp:1-6):
 *   method BDDA_COCG_RQ  + jacobi 0/1 -> bl_cocg_rq_mvp_fft / bl_cocg_rq_jacobi_mvp_fft
 *   method BDDA_BICGSTAB + jacobi 0/1 -> bl_bicgstab_mvp_fft / ..._jacobi_...
 *   method BDDA_BICGGR   + jacobi 0/1 -> bl_bicggr_mvp_fft / ..._jacobi_...
 * mode BDDA_MODE_FP64 keeps the reference's FP64 arithmetic and stopping
 * rule (recurrence residual < tol); BDDA_MODE_MIXED runs the inner block
 * iteration in complex FP32 inside an FP64 defect-correction loop that stops
 * on the true FP64 residual ||B - A X||/||B|| <= tol; BDDA_MODE_MIXED_COLUMNS
 * (no reference counterpart: the reference stops on the block norm, so one
 * column may stay above tol) is the same loop stopping only when every
 * column's true residual ||B_j - A X_j||/||B_j|| <= tol, and info.true_res
 * is then the largest of them.  hist (optional,
 * hist_len entries) receives per-iteration residuals (outer true residuals
 * are stored negated; entries past the last one written are NaN).  Returns
 * BDDA_ENOCONV when itermax was reached. */
int bdda_solve(bdda_mbt* h, int method, int jacobi, int mode, const bdda_c64* B, int64_t ldb, int L, double tol,
               int itermax, bdda_c64* X, int64_t ldx, bdda_solve_info* info, double* hist, int hist_len);

/* Per-column true residuals res[j] = ||B_j - A X_j|| / ||B_j|| (||B_j - A X_j||
 * when B_j = 0), in FP64 with one FP64 block matvec and A without Jacobi
 * scaling.  No reference counterpart: the reference prints only the block
 * recurrence residual ||xi||_F / ||B||_F (bl_cocg_rq_mvp_fft.cpp:58-60), so
 * a block can pass with one column above tol; this reports the worst one. */
int bdda_residual_columns(bdda_mbt* h, const bdda_c64* B, int64_t ldb, const bdda_c64* X, int64_t ldx, int L,
                          double* res);

/* qr_reduced (bl_krylov_mvp_fft/qr_reduced.cpp:9-31) on the GPU: A is m x n
 * column-major (ld = m); Q (m x n) and R (n x n, R(i,k) first-pass
 * coefficients, R(k,k) the norm after re-orthogonalisation). */
int bdda_qr_reduced(const bdda_c64* A, int64_t m, int n, bdda_c64* Q, bdda_c64* R);

/* X (s x nb) = A^{-1} B for the s x s matrices of the block solvers, on the
 * GPU in FP64, by LU with partial (full = 0) or full (full = 1) pivoting:
 * the Eigen PartialPivLU(A).solve(B) / FullPivLU(A).solve(B) calls of the
 * solvers (bl_cocg_rq_jacobi_mvp_fft.cpp:63,73, bl_cocg_rq_mvp_fft.cpp:56,66,
 * bl_bicgstab_jacobi_mvp_fft.cpp:62,93, bl_bicggr_jacobi_mvp_fft.cpp:58,113).
 * The solvers run the same kernel on device-resident matrices.  Host
 * buffers, column-major, 1 <= s <= 64, 0 <= nb <= 64.  An exactly zero pivot
 * returns BDDA_EBREAKDOWN (Eigen would return non-finite values). */
int bdda_lu_solve(const bdda_c64* A, const bdda_c64* B, bdda_c64* X, int s, int nb, int full);

/* ---------------------------------------------------------------- observables
 * Cross sections of a solved block (Result_manager.cpp:65-122), on the GPU
 * in FP64.  B is the incident-field block the solve used (EHinc) and X the
 * dipole block (PM), both (f*nocc) x L; the wavenumber is the operator's and
 * must be real (Result_manager.cpp:71-76).  Per column l:
 *   Cext = 4 pi k Im(sum_i conj(B_il) X_il)
 *   Cabs = 4 pi k sum_i (Im(conj(1/alpha_i)) - 2 k^3 / 3) |X_il|^2,
 *          1/alpha_i = DIAG_A_i of the target
 *   Csca = Cext - Cabs.
 * BDDA_ALPHA_FIRST takes every dipole's alpha from dipole 0, as the
 * reference does (Result_manager.cpp:73,97-98); BDDA_ALPHA_EACH uses each
 * dipole's own DIAG_A (targets of several materials). */
enum { BDDA_ALPHA_FIRST = 0, BDDA_ALPHA_EACH = 1 };
int bdda_cross_sections(bdda_mbt* h, const bdda_c64* B, int64_t ldb, const bdda_c64* X, int64_t ldx, int L,
                        int alpha_mode, double* Cext, double* Cabs, double* Csca);

/* Far-field scattering matrix.  No reference counterpart: the reference
 * writes cross sections only.  Columns 2p and 2p+1 of X are the dipoles
 * excited by plane waves along u_p polarised along e1_p and u_p x e1_p
 * (u, e1: npairs x 3, orthonormal).  In the plane of u_p and e1_p, at
 * theta_a = pi a / (nang - 1), the Bohren-Huffman amplitudes follow from
 *   E_far = k^2 e^{ikr}/r sum_j [(I - n n) P_j - n x M_j] e^{-ik n.r_j},
 *   n = cos(theta) u + sin(theta) e1, r_j = lf (ix, iy, iz):
 *   S2 = -ik theta.E', S4 = -ik phi.E' (polarisation e1),
 *   S3 = -ik theta.E', S1 = -ik phi.E' (polarisation u x e1),
 * with E' = E_far r e^{-ikr}, phi = u x e1 (M_j for f = 6 only), and
 *   S11 = (|S1|^2 + |S2|^2 + |S3|^2 + |S4|^2) / 2,
 *   S12 = (|S2|^2 - |S1|^2 + |S4|^2 - |S3|^2) / 2.
 * S11, S12: npairs x nang (pair-major); S (optional): npairs x nang x 4
 * complex (S1, S2, S3, S4).  The forward amplitude obeys the optical theorem
 * against bdda_cross_sections: Cext(pol e1) = 4 pi / k^2 Re S2(0). */
int bdda_scattering_matrix(bdda_mbt* h, const bdda_c64* X, int64_t ldx, int npairs, const double* u,
                           const double* e1, int nang, double* S11, double* S12, bdda_c64* S);

/* ---------------------------------------------------------------- memory
 * No reference counterpart (the reference sizes nothing ahead of time).
 * HBM estimate for a target of nocc dipoles on the n lattice solved with L
 * right-hand sides in `mode`: init_peak while bdda_mbt_create runs (the full
 * kernel grid is transformed in FP64 before its octant is kept), resident
 * bytes of the operator with its target, the pass workspace of an L-column
 * block matvec (see bdda_mbt_set_workspace_limit) and the Krylov blocks of
 * one solve.  The workspace and the Krylov blocks live for one call: every
 * call that runs block matvecs gives them back to the device allocator
 * before it returns. */
int bdda_estimate_bytes(const int n[3], int f, int64_t nocc, int L, int mode, int64_t* init_peak,
                        int64_t* resident, int64_t* workspace, int64_t* krylov);
/* Cap the pass workspace of a handle: block matvecs whose workspace would
 * exceed `bytes` run in column chunks (0 = no cap, the default). */
int bdda_mbt_set_workspace_limit(bdda_mbt* h, size_t bytes);

/* ---------------------------------------------------------------- device
 * Device-resident variants (pointers from bdda_dev_alloc, complex<double>
 * layout) for callers that keep B and X in HBM. */
/* HBM bandwidth that a plain streaming kernel of this library reaches (no
 * reference counterpart; measurement only): best of five launches each of a
 * copy a -> b (16-byte non-temporal vector accesses, rate counts read plus
 * write bytes), a read of a and a write of b.  a and b are device buffers of
 * `bytes` each (from bdda_dev_alloc); their contents are overwritten. */
int bdda_hbm_probe(void* a, void* b, size_t bytes, double* copy_gbs, double* read_gbs, double* write_gbs);

/* Device memory of the library (no reference counterpart: the reference
 * keeps its blocks in host memory).  Every device buffer, the library's own
 * included, comes from one caching allocator per GPU: freed blocks stay with
 * it and are handed out again (best fit, split, merged with idle neighbours),
 * because memory new to the process is slow to get from the runtime (single
 * hipMalloc calls of 0.3-3 s for the tens of GB a large operator needs).
 * bdda_dev_free waits for the whole device before the block is reused, as
 * hipFree does.  When the runtime runs out of memory, the wholly idle
 * allocations go back to it and the request is tried again; bdda_dev_trim
 * does that on demand (for a caller that shares the GPU with another
 * allocator) and reports the bytes released.  BDDA_DEVICE_CACHE=0 in the
 * environment turns the cache off. */
int bdda_dev_alloc(size_t bytes, void** p);
int bdda_dev_free(void* p);
int bdda_dev_trim(size_t* released);
/* bytes held from the runtime and in live blocks, runtime allocations made
 * and requests served from idle blocks, on the current device */
int bdda_dev_pool_stats(size_t* held, size_t* in_use, int64_t* runtime_allocs, int64_t* reuses);
int bdda_memcpy_h2d(void* dst, const void* src, size_t bytes);
int bdda_memcpy_d2h(void* dst, const void* src, size_t bytes);
/* device-to-device copy (also the known-size reference for HBM counters) */
int bdda_memcpy_d2d(void* dst, const void* src, size_t bytes);
int bdda_device_sync(void);
int bdda_solve_dev(bdda_mbt* h, int method, int jacobi, int mode, const bdda_c64* dB, int L, double tol,
                   int itermax, bdda_c64* dX, bdda_solve_info* info, double* hist, int hist_len);
/* device block matvec in the given precision; dX/dY hold complex<float>
 * (prec FP32) or complex<double> (prec FP64), ld = f*nocc */
int bdda_mbt_apply_occupied_dev(bdda_mbt* h, int prec, int jacobi, const void* dX, void* dY, int ncols);
/* bdda_residual_columns, bdda_cross_sections and bdda_scattering_matrix on
 * device blocks (ld = f*nocc) */
int bdda_residual_columns_dev(bdda_mbt* h, const bdda_c64* dB, const bdda_c64* dX, int L, double* res);
int bdda_cross_sections_dev(bdda_mbt* h, const bdda_c64* dB, const bdda_c64* dX, int L, int alpha_mode, double* Cext,
                            double* Cabs, double* Csca);
int bdda_scattering_matrix_dev(bdda_mbt* h, const bdda_c64* dX, int npairs, const double* u, const double* e1,
                               int nang, double* S11, double* S12, bdda_c64* S);
/* bdda_cross_sections_dev without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  The per-group partial sums
 * live in a file-scope device array of the library (1.5 MiB per GPU, held
 * until the process ends) instead of a per-call hipMalloc; L is limited to
 * 1..256 columns (BDDA_EINVAL beyond).  Results are bit-identical to
 * bdda_cross_sections_dev.  All calls on one GPU share that array, so calls
 * that overlap on one GPU (several threads or streams) race; serialise them. */
int bdda_cross_sections_dev_noalloc(bdda_mbt* h, const bdda_c64* dB, const bdda_c64* dX, int L, int alpha_mode,
                                    double* Cext, double* Cabs, double* Csca);
/* per-pass kernel timing of the following applies (HIP events on the
 * handle's stream); pass_ms receives z-fwd, y-fwd, x-fused, y-inv, z-inv */
int bdda_mbt_set_timing(bdda_mbt* h, int on);
int bdda_mbt_pass_times(const bdda_mbt* h, float pass_ms[5]);
/* algorithmic HBM bytes of one block matvec of ncols columns */
double bdda_mbt_algorithmic_bytes(const bdda_mbt* h, int prec, int ncols);

#ifdef __cplusplus
}
#endif
#endif
