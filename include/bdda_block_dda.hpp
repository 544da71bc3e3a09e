// bdda_block_dda.hpp -- the reference's solve-path C++ interface over the C ABI.
//
// Declares, with the reference's names, parameter lists and meaning:
//   mvp_fft/MBT_utils.hpp:1-6
//     application_function, BT_fft, BT_pad, BT_reconstruct, MBT_fft_init,
//     MBT_fft_mvp
//   DDA_linear_solvers.hpp:1-6
//     bl_cocg_rq[_jacobi]_mvp_fft, bl_bicgstab[_jacobi]_mvp_fft,
//     bl_bicggr[_jacobi]_mvp_fft
//   bl_krylov_mvp_fft/linear_algebra_addon.hpp:1
//     qr_reduced
// so that Solver.cpp (Solver.cpp:55,64-72) compiles and links against the
// MI355X path unchanged.  Header-only: every function forwards to
// libbdda.so through include/bdda.h.
//
// Types.  With BDDA_WITH_EIGEN defined the functions take and return the
// Eigen types of the reference (Eigen::VectorXi, MatrixXi, VectorXcd,
// MatrixXcd; column-major, std::complex<double>).  Without it they use
// bdda_cxx::Dense, a column-major container with the subset of Eigen's
// interface used here (rows, cols, size, data, (i), (i, j)); the memory
// layout is the same.
//
// Plans.  The reference threads two fftw_plan handles through every call.
// Here the plan arguments are template parameters and are never executed:
// the caller may pass its fftw_plan objects unchanged, or bdda_cxx::Plan.
//
// Au_til.  MBT_fft_init builds the transformed interaction kernel on the GPU
// and returns a 4-element token vector in place of the length-F host vector
// (MBT_fft_init.cpp:24-36): element 0 holds a magic number and an operator
// id, elements 1-3 the lattice, f and k.  MBT_fft_mvp and the solvers look
// the device operator up from that token; a vector not made by this
// MBT_fft_init is rejected.  Operators live until the process ends, as the
// reference's plans do (Solver.hpp:30-31).
//
// Precision.  The solvers run in complex FP64 with the reference's stopping
// rule by default.  bdda_cxx::set_solve_mode(BDDA_MODE_MIXED), or the
// environment variable BDDA_SOLVE_MODE=mixed read at the first solve, selects
// the FP32-inner / FP64-defect-correction solve (stopping on the true FP64
// residual).
//
// Errors.  Where the reference prints to cerr and calls exit(EXIT_FAILURE)
// (application_function.cpp:71-76, MBT_fft_init.cpp:29-32), this header does
// the same for every failed library call; with BDDA_CXX_THROW defined it
// throws bdda_cxx::Error instead.  A solve that reaches itermax returns X
// silently, as the reference does.  Each solver prints the reference's
// per-iteration line ("<name>: iter= k relative err= e") after the solve
// unless bdda_cxx::set_verbose(false) was called.
#pragma once
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "bdda.h"

#if defined(BDDA_WITH_EIGEN)
#include <Eigen/Dense>
#endif

namespace bdda_cxx {

#if defined(BDDA_WITH_EIGEN)
using VectorXi = Eigen::VectorXi;
using MatrixXi = Eigen::MatrixXi;
using VectorXcd = Eigen::VectorXcd;
using MatrixXcd = Eigen::MatrixXcd;
#else
// column-major dense container; vectors are n x 1
template <typename T>
class Dense {
public:
    Dense() = default;
    explicit Dense(int64_t rows, int64_t cols = 1) : r_(rows), c_(cols), v_((size_t)(rows * cols)) {}
    int64_t rows() const { return r_; }
    int64_t cols() const { return c_; }
    int64_t size() const { return r_ * c_; }
    T* data() { return v_.data(); }
    const T* data() const { return v_.data(); }
    T& operator()(int64_t i) { return v_[(size_t)i]; }
    const T& operator()(int64_t i) const { return v_[(size_t)i]; }
    T& operator()(int64_t i, int64_t j) { return v_[(size_t)(i + j * r_)]; }
    const T& operator()(int64_t i, int64_t j) const { return v_[(size_t)(i + j * r_)]; }
    T& operator[](int64_t i) { return v_[(size_t)i]; }
    const T& operator[](int64_t i) const { return v_[(size_t)i]; }
    void resize(int64_t rows, int64_t cols = 1)
    {
        r_ = rows;
        c_ = cols;
        v_.assign((size_t)(rows * cols), T());
    }
    void setZero() { std::fill(v_.begin(), v_.end(), T()); }

private:
    int64_t r_ = 0, c_ = 0;
    std::vector<T> v_;
};
using VectorXi = Dense<int>;
using MatrixXi = Dense<int>;
using VectorXcd = Dense<std::complex<double>>;
using MatrixXcd = Dense<std::complex<double>>;
#endif

// stands in for fftw_plan where the caller has none
struct Plan {
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

namespace detail {

[[noreturn]] inline void fail_msg(int code, const std::string& msg)
{
#if defined(BDDA_CXX_THROW)
    throw Error(code, msg);
#else
    (void)code;
    std::cerr << msg << std::endl;
    std::exit(EXIT_FAILURE);
#endif
}

[[noreturn]] inline void fail(int code, const char* where)
{
    fail_msg(code, std::string(where) + ": " + bdda_last_error());
}

inline void check(int rc, const char* where)
{
    if (rc != BDDA_OK) fail(rc, where);
}

inline bdda_c64 c64(const std::complex<double>& z) { return bdda_c64{z.real(), z.imag()}; }

template <class V>
inline void lattice(const V& n, int out[3], const char* where)
{
    if (n.size() != 3) fail_msg(BDDA_EINVAL, std::string(where) + ": the lattice n must have 3 levels");
    for (int i = 0; i < 3; ++i) out[i] = (int)n(i);
}

struct Operator {
    bdda_mbt* h = nullptr;
    int n[3] = {0, 0, 0};
    int f = 0;
    std::mutex mu;  // one call in flight per operator (the reference calls
                    // MBT_fft_mvp from several OpenMP threads at once)
};

struct State {
    std::mutex mu;
    std::vector<std::unique_ptr<Operator>> ops;
    int mode = -1;  // -1: not chosen yet (environment read at the first solve)
    bool verbose = true;
};

inline State& state()
{
    static State s;
    return s;
}

constexpr double kMagic = 6.02214076e23 * 1.6180339887;  // token tag

inline Operator& lookup(const VectorXcd& Au_til, const int n[3], int f, const char* where)
{
    State& s = state();
    bool ok = Au_til.size() == 4 && Au_til(0).real() == kMagic;
    int64_t id = ok ? (int64_t)Au_til(0).imag() : -1;
    Operator* op = nullptr;
    {
        std::lock_guard<std::mutex> g(s.mu);
        if (ok && id >= 0 && id < (int64_t)s.ops.size()) op = s.ops[(size_t)id].get();
    }
    if (!op || op->f != f || op->n[0] != n[0] || op->n[1] != n[1] || op->n[2] != n[2])
        fail_msg(BDDA_EINVAL, std::string(where) + ": Au_til was not made by MBT_fft_init for this n and f");
    return *op;
}

inline int solve_mode()
{
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    if (s.mode < 0) {
        const char* e = std::getenv("BDDA_SOLVE_MODE");
        s.mode = (e && std::strcmp(e, "mixed") == 0) ? BDDA_MODE_MIXED : BDDA_MODE_FP64;
    }
    return s.mode;
}

inline MatrixXcd solve(const char* name, int method, int jacobi, const VectorXi& n_, int f,
                       const std::vector<int>& address, const VectorXcd& Au_til, const VectorXcd& DIAG_A,
                       const MatrixXcd& B, double tol, int itermax)
{
    int n[3];
    lattice(n_, n, name);
    Operator& op = lookup(Au_til, n, f, name);
    const int64_t rows = B.rows();
    const int L = (int)B.cols();
    if ((int64_t)address.size() * f != rows || DIAG_A.size() != rows)
        fail_msg(BDDA_EINVAL, std::string(name) + ": B, DIAG_A and address sizes disagree");
    MatrixXcd X(rows, L);
    std::vector<double> hist((size_t)(itermax > 0 ? itermax : 1) + 64, 0.0);
    bdda_solve_info info{};
    const int mode = solve_mode();
    int rc;
    {
        std::lock_guard<std::mutex> g(op.mu);
        rc = bdda_mbt_set_target(op.h, address.data(), (int64_t)address.size(),
                                 reinterpret_cast<const bdda_c64*>(DIAG_A.data()));
        if (rc == BDDA_OK)
            rc = bdda_solve(op.h, method, jacobi, mode, reinterpret_cast<const bdda_c64*>(B.data()), rows, L, tol,
                            itermax, reinterpret_cast<bdda_c64*>(X.data()), rows, &info, hist.data(),
                            (int)hist.size());
    }
    if (rc != BDDA_OK && rc != BDDA_ENOCONV) fail(rc, name);
    if (state().verbose) {
        // inner recurrence residuals, as the reference prints them; in mixed
        // mode the outer true residuals come negated between them
        int k = 0;
        for (double e : hist) {
            if (e != e) break;  // NaN: past the last entry
            if (e < 0)
                std::cout << name << ": outer true relative err= " << -e << std::endl;
            else
                std::cout << name << ": iter= " << k++ << " relative err= " << e << std::endl;
        }
    }
    return X;
}

// BT_fft over the levels level..3 (1-based) with n_ind fixed on the levels
// above: block j of a level holds displacement n-1-j, taken as the index
// pair (n-j, 1) for j < n and (1, j-n+2) above (BT_fft.cpp:39-50); a
// terminal f x f block M is stored as flipud(M)^T in column-major order
// (BT_fft.cpp:25-31).
inline void bt_fft_level(const int n[3], int f, int ind[6], int level, double lf, bdda_c64 k,
                         std::complex<double>* out, int64_t& pos)
{
    if (level == 4) {
        std::vector<std::complex<double>> M((size_t)f * f);
        check(bdda_application_function(ind, f, lf, k, reinterpret_cast<bdda_c64*>(M.data())), "BT_fft");
        for (int b = 0; b < f; ++b)
            for (int a = 0; a < f; ++a) out[pos + a + (int64_t)b * f] = M[(size_t)((f - 1 - b) + a * f)];
        pos += (int64_t)f * f;
        return;
    }
    const int nl = n[level - 1];
    for (int j = 0; j < 2 * nl - 1; ++j) {
        ind[level - 1] = j < nl ? nl - j : 1;
        ind[3 + level - 1] = j < nl ? 1 : j - nl + 2;
        bt_fft_level(n, f, ind, level + 1, lf, k, out, pos);
    }
}

}  // namespace detail

inline void set_solve_mode(int mode)
{
    std::lock_guard<std::mutex> g(detail::state().mu);
    detail::state().mode = mode;
}

inline void set_verbose(bool on)
{
    std::lock_guard<std::mutex> g(detail::state().mu);
    detail::state().verbose = on;
}

}  // namespace bdda_cxx

// --------------------------------------------------------------- MBT_utils

// mvp_fft/application_function.cpp:7-78 (n and level are unused there too)
inline bdda_cxx::MatrixXcd application_function(const bdda_cxx::VectorXi& n, const int& f, bdda_cxx::MatrixXi& n_ind,
                                                int level, const double& lf, const std::complex<double>& k)
{
    (void)n;
    (void)level;
    int ind[6];
    for (int j = 0; j < 3; ++j) {
        ind[j] = n_ind(0, j);
        ind[3 + j] = n_ind(1, j);
    }
    bdda_cxx::MatrixXcd out(f, f);
    bdda_cxx::detail::check(bdda_application_function(ind, f, lf, bdda_cxx::detail::c64(k),
                                                      reinterpret_cast<bdda_c64*>(out.data())),
                            "application_function");
    return out;
}

// mvp_fft/BT_fft.cpp:8-53: the projection for levels level..3, n_ind fixing
// the levels above (MBT_fft_init calls it with n_ind = 1, level = 1)
inline bdda_cxx::VectorXcd BT_fft(const bdda_cxx::VectorXi& n, const int& f, bdda_cxx::MatrixXi n_ind, const int& level,
                                  const double& lf, const std::complex<double>& k)
{
    int nn[3];
    bdda_cxx::detail::lattice(n, nn, "BT_fft");
    if (f != 3 && f != 6) bdda_cxx::detail::fail_msg(BDDA_EINVAL, "BT_fft: unknown f value");
    int64_t len = (int64_t)f * f;
    for (int l = level; l <= 3; ++l) len *= 2 * nn[l - 1] - 1;
    int ind[6];
    for (int j = 0; j < 3; ++j) {
        ind[j] = n_ind(0, j);
        ind[3 + j] = n_ind(1, j);
    }
    bdda_cxx::VectorXcd Au(len);
    int64_t pos = 0;
    bdda_cxx::detail::bt_fft_level(nn, f, ind, level, lf, bdda_cxx::detail::c64(k), Au.data(), pos);
    return Au;
}

// mvp_fft/BT_pad.cpp:6-45
inline bdda_cxx::VectorXcd BT_pad(const bdda_cxx::VectorXi& n, const int& f, const bdda_cxx::VectorXcd& x)
{
    int nn[3];
    bdda_cxx::detail::lattice(n, nn, "BT_pad");
    bdda_cxx::VectorXcd xz((int64_t)f * f * nn[0] * (2 * nn[1] - 1) * (2 * nn[2] - 1));
    if (x.size() != (int64_t)f * nn[0] * nn[1] * nn[2]) bdda_cxx::detail::fail_msg(BDDA_EINVAL, "BT_pad: wrong x size");
    bdda_cxx::detail::check(bdda_bt_pad(nn, f, reinterpret_cast<const bdda_c64*>(x.data()),
                                        reinterpret_cast<bdda_c64*>(xz.data())),
                            "BT_pad");
    return xz;
}

// mvp_fft/BT_reconstruct.cpp:5-36
inline bdda_cxx::VectorXcd BT_reconstruct(const bdda_cxx::VectorXi& n, const int& f, const bdda_cxx::VectorXcd& bz)
{
    int nn[3];
    bdda_cxx::detail::lattice(n, nn, "BT_reconstruct");
    if (bz.size() != bdda_fft_length(nn, f)) bdda_cxx::detail::fail_msg(BDDA_EINVAL, "BT_reconstruct: wrong bz size");
    bdda_cxx::VectorXcd b((int64_t)f * nn[0] * nn[1] * nn[2]);
    bdda_cxx::detail::check(bdda_bt_reconstruct(nn, f, reinterpret_cast<const bdda_c64*>(bz.data()),
                                                reinterpret_cast<bdda_c64*>(b.data())),
                            "BT_reconstruct");
    return b;
}

// mvp_fft/MBT_fft_init.cpp:10-42 (with Solver::fft_prep, Solver.cpp:17-25):
// builds the device operator in FP64 and FP32 and returns its token
template <class PlanT>
inline bdda_cxx::VectorXcd MBT_fft_init(const bdda_cxx::VectorXi& n, const int& f, const double& lf,
                                        const std::complex<double>& k, PlanT& plan_fwd)
{
    (void)plan_fwd;
    using namespace bdda_cxx::detail;
    auto op = std::make_unique<Operator>();
    lattice(n, op->n, "MBT_fft_init");
    op->f = f;
    check(bdda_mbt_create(op->n, f, lf, c64(k), BDDA_PREC_FP64 | BDDA_PREC_FP32, &op->h), "MBT_fft_init");
    State& s = state();
    int64_t id;
    {
        std::lock_guard<std::mutex> g(s.mu);
        id = (int64_t)s.ops.size();
        s.ops.push_back(std::move(op));
    }
    bdda_cxx::VectorXcd tok(4);
    tok(0) = std::complex<double>(kMagic, (double)id);
    tok(1) = std::complex<double>(n(0), n(1));
    tok(2) = std::complex<double>(n(2), f);
    tok(3) = k;
    return tok;
}

// mvp_fft/MBT_fft_mvp.cpp:8-43: the off-diagonal product on one cuboid
// vector of f*nx*ny*nz entries
template <class PlanT>
inline bdda_cxx::VectorXcd MBT_fft_mvp(const bdda_cxx::VectorXi& n, const int& f, const bdda_cxx::VectorXcd& Au_til,
                                       const bdda_cxx::VectorXcd& p_hat, PlanT& plan_fwd, PlanT& plan_inv)
{
    (void)plan_fwd;
    (void)plan_inv;
    using namespace bdda_cxx::detail;
    int nn[3];
    lattice(n, nn, "MBT_fft_mvp");
    Operator& op = lookup(Au_til, nn, f, "MBT_fft_mvp");
    const int64_t C = (int64_t)f * nn[0] * nn[1] * nn[2];
    if (p_hat.size() != C) fail_msg(BDDA_EINVAL, "MBT_fft_mvp: wrong p_hat size");
    bdda_cxx::VectorXcd y(C);
    std::lock_guard<std::mutex> g(op.mu);
    check(bdda_mbt_apply(op.h, BDDA_PREC_FP64, reinterpret_cast<const bdda_c64*>(p_hat.data()), C,
                         reinterpret_cast<bdda_c64*>(y.data()), C, 1),
          "MBT_fft_mvp");
    return y;
}

// ----------------------------------------------------- DDA_linear_solvers

#define BDDA_CXX_SOLVER(NAME, PRINTED, METHOD, JACOBI)                                                            \
    template <class PlanT>                                                                                        \
    inline bdda_cxx::MatrixXcd NAME(const bdda_cxx::VectorXi& n, const int& f, const std::vector<int>& address,    \
                                    const bdda_cxx::VectorXcd& Au_til, const bdda_cxx::VectorXcd& DIAG_A,         \
                                    const bdda_cxx::MatrixXcd& B, const double& tol, const int& itermax,          \
                                    PlanT& plan_fwd, PlanT& plan_inv)                                             \
    {                                                                                                             \
        (void)plan_fwd;                                                                                           \
        (void)plan_inv;                                                                                           \
        return bdda_cxx::detail::solve(PRINTED, METHOD, JACOBI, n, f, address, Au_til, DIAG_A, B, tol, itermax); \
    }

// bl_krylov_mvp_fft/bl_cocg_rq_mvp_fft.cpp:12-72, bl_cocg_rq_jacobi_mvp_fft.cpp:12-79
BDDA_CXX_SOLVER(bl_cocg_rq_mvp_fft, "bl_cocg_rq", BDDA_COCG_RQ, 0)
BDDA_CXX_SOLVER(bl_cocg_rq_jacobi_mvp_fft, "bl_cocg_rq_jacobi_mvp_fft", BDDA_COCG_RQ, 1)
// bl_krylov_mvp_fft/bl_bicgstab_mvp_fft.cpp:12-101, bl_bicgstab_jacobi_mvp_fft.cpp:12-97
BDDA_CXX_SOLVER(bl_bicgstab_mvp_fft, "bl_bicgstab", BDDA_BICGSTAB, 0)
BDDA_CXX_SOLVER(bl_bicgstab_jacobi_mvp_fft, "bl_bicgstab_jacobi_mvp_fft", BDDA_BICGSTAB, 1)
// bl_krylov_mvp_fft/bl_bicggr_mvp_fft.cpp:10-119, bl_bicggr_jacobi_mvp_fft.cpp:10-120
BDDA_CXX_SOLVER(bl_bicggr_mvp_fft, "bl_bicggr", BDDA_BICGGR, 0)
BDDA_CXX_SOLVER(bl_bicggr_jacobi_mvp_fft, "bl_bicggr_jacobi_mvp_fft", BDDA_BICGGR, 1)
#undef BDDA_CXX_SOLVER

// ------------------------------------------------- linear_algebra_addon

// bl_krylov_mvp_fft/qr_reduced.cpp:9-31
inline std::tuple<bdda_cxx::MatrixXcd, bdda_cxx::MatrixXcd> qr_reduced(const bdda_cxx::MatrixXcd& A)
{
    const int64_t m = A.rows();
    const int nc = (int)A.cols();
    bdda_cxx::MatrixXcd Q(m, nc), R(nc, nc);
    bdda_cxx::detail::check(bdda_qr_reduced(reinterpret_cast<const bdda_c64*>(A.data()), m, nc,
                                            reinterpret_cast<bdda_c64*>(Q.data()),
                                            reinterpret_cast<bdda_c64*>(R.data())),
                            "qr_reduced");
    return std::make_tuple(Q, R);
}
