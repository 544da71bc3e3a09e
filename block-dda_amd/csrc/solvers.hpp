// solvers.hpp -- block Krylov solvers over the device MBT operator.
#pragma once
#include "blockops.hpp"
#include "mbt.hpp"
#include <vector>

namespace bdda {

enum Method : int { M_COCG_RQ = 0, M_BICGSTAB = 1, M_BICGGR = 2 };
enum Mode : int { MODE_FP64 = 0, MODE_FP32 = 1, MODE_MIXED = 2, MODE_MIXED_COLUMNS = 3 };

struct SolveInfo {
    int iters = 0;          // inner (block) iterations executed, summed over outer steps
    int outer_iters = 0;    // defect-correction steps (1 for plain modes)
    int matvecs = 0;        // block matvecs (each covers all s columns)
    int converged = 0;
    int status = 0;
    double recurrence_res = 0;  // last reference-style residual (per method)
    double true_res = -1;       // ||B - A X||_F / ||B||_F in FP64 when computed
    double time_ms = 0;
};

// The problem: operator + occupancy + DIAG_A (Solver.cpp:28-47) on device.
struct Problem {
    MbtOperator* op = nullptr;
    int64_t nocc = 0;
    int f = 3;
    c64* diag64 = nullptr;
    c64* jpre64 = nullptr;
    c32* diag32 = nullptr;
    c32* jpre32 = nullptr;
    BlockWork work;
    ~Problem();
    // device buffers back to the allocator; call before work.stream is destroyed
    void release();
    int64_t rows() const { return (int64_t)f * nocc; }
    template <typename R>
    const cplx<R>* diag() const;
    template <typename R>
    const cplx<R>* jpre() const;
};

void problem_set_diag(Problem& P, const c64* diag_host);

// Y = A X (jacobi: D^{-1} A X) on device, s columns, ld = rows
template <typename R>
void problem_matvec(Problem& P, const cplx<R>* X, cplx<R>* Y, int s, int jacobi);

// Solve A X = B (device pointers, ld = rows).  hist (may be null) receives
// the per-iteration residual the reference prints.
int solve_fp(Problem& P, int method, int jacobi, int mode, const c64* B64, c64* X64, int s, double tol,
             int itermax, SolveInfo& info, std::vector<double>* hist, double inner_tol);

// FP64 true relative residual ||B - A X|| / ||B|| (uses one FP64 matvec)
double true_residual(Problem& P, const c64* B64, const c64* X64, int s);

// per column: ||B_j - A X_j|| / ||B_j|| (||B_j - A X_j|| where B_j = 0), one FP64 matvec
void residual_columns(Problem& P, const c64* B64, const c64* X64, int s, double* out);

}  // namespace bdda
