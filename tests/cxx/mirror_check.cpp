// mirror_check.cpp -- drives the reference-signature C++ interface
// (include/bdda_block_dda.hpp) the way Solver.cpp does, and writes every
// input and output to a record file that tests/test_cxx_mirror.py checks
// against the oracle.
//
//   mirror_check layout <file>   CPU only: application_function, BT_fft (all
//                                levels), BT_pad, BT_reconstruct
//   mirror_check gpu <file>      MBT_fft_init, MBT_fft_mvp, the six solvers,
//                                qr_reduced (needs the GPU)
//   mirror_check badf            application_function with f = 4 (the
//                                reference prints and exits with failure)
//   mirror_check badtoken        MBT_fft_mvp with an Au_til not made by
//                                MBT_fft_init (needs the GPU library only to
//                                load; fails before any device call)
//
// Record file: repeated [int32 name length][name][int32 kind: 0 complex128,
// 1 int32, 2 float64][int64 count][data].
#include <cstdio>
#include <fstream>
#include <string>

#include "bdda_block_dda.hpp"

using bdda_cxx::MatrixXcd;
using bdda_cxx::MatrixXi;
using bdda_cxx::VectorXcd;
using bdda_cxx::VectorXi;
using cd = std::complex<double>;

namespace {

struct Records {
    std::ofstream out;
    explicit Records(const char* path) : out(path, std::ios::binary) {}
    void put(const std::string& name, int kind, int64_t count, const void* data, size_t elem)
    {
        const int32_t nl = (int32_t)name.size();
        out.write((const char*)&nl, 4);
        out.write(name.data(), nl);
        out.write((const char*)&kind, 4);
        out.write((const char*)&count, 8);
        out.write((const char*)data, (std::streamsize)(count * (int64_t)elem));
    }
    void c(const std::string& name, const VectorXcd& v) { put(name, 0, v.size(), v.data(), 16); }
    void i(const std::string& name, const std::vector<int>& v) { put(name, 1, (int64_t)v.size(), v.data(), 4); }
    void d(const std::string& name, const std::vector<double>& v) { put(name, 2, (int64_t)v.size(), v.data(), 8); }
};

// deterministic complex values in [-1, 1)^2 (64-bit LCG, Knuth's constants)
struct Lcg {
    uint64_t s;
    explicit Lcg(uint64_t seed) : s(seed) {}
    double u()
    {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(s >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
    VectorXcd vec(int64_t n, int64_t cols = 1)
    {
        VectorXcd v(n, cols);
        for (int64_t j = 0; j < n * cols; ++j) v.data()[j] = cd(u(), u());
        return v;
    }
};

VectorXi lattice(int a, int b, int c)
{
    VectorXi n(3);
    n(0) = a;
    n(1) = b;
    n(2) = c;
    return n;
}

int run_layout(const char* path)
{
    Records R(path);
    Lcg g(7);
    const double lf = 0.6;
    const cd k(1.2, 0.1);
    for (int f : {3, 6}) {
        const VectorXi n = lattice(2, 3, 4);
        const std::string t = "f" + std::to_string(f) + "_";
        MatrixXi ind(2, 3);
        ind(0, 0) = 2; ind(0, 1) = 1; ind(0, 2) = 3;
        ind(1, 0) = 1; ind(1, 1) = 2; ind(1, 2) = 1;
        R.c(t + "app", application_function(n, f, ind, 4, lf, k));
        MatrixXi ones(2, 3);
        for (int r = 0; r < 2; ++r)
            for (int l = 0; l < 3; ++l) ones(r, l) = 1;
        R.c(t + "bt_fft_l1", BT_fft(n, f, ones, 1, lf, k));
        // a sub-projection: level 1 fixed at index pair (2, 1)
        MatrixXi sub = ones;
        sub(0, 0) = 2;
        R.c(t + "bt_fft_l2", BT_fft(n, f, sub, 2, lf, k));
        // level 1 at (1, 2) and level 2 at (3, 1)
        MatrixXi sub3 = ones;
        sub3(1, 0) = 2;
        sub3(0, 1) = 3;
        R.c(t + "bt_fft_l3", BT_fft(n, f, sub3, 3, lf, k));
        const VectorXcd x = g.vec((int64_t)f * 2 * 3 * 4);
        R.c(t + "x", x);
        R.c(t + "bt_pad", BT_pad(n, f, x));
        const VectorXcd bz = g.vec((int64_t)f * f * 3 * 5 * 7);
        R.c(t + "bz", bz);
        R.c(t + "bt_reconstruct", BT_reconstruct(n, f, bz));
    }
    return 0;
}

int run_gpu(const char* path)
{
    Records R(path);
    Lcg g(11);
    bdda_cxx::set_verbose(true);
    bdda_cxx::Plan plan_fwd, plan_inv;
    const double lf = 0.5;
    const cd k(0.9, 0.0);
    for (int f : {3, 6}) {
        const std::string t = "f" + std::to_string(f) + "_";
        const VectorXi n = lattice(6, 5, 7);
        const int C = 6 * 5 * 7;
        VectorXcd Au_til = MBT_fft_init(n, f, lf, k, plan_fwd);
        // MBT_fft_mvp on one cuboid vector
        const VectorXcd p = g.vec((int64_t)f * C);
        R.c(t + "p_hat", p);
        R.c(t + "mvp", MBT_fft_mvp(n, f, Au_til, p, plan_fwd, plan_inv));
        // target: every third cell occupied, one material; DIAG_A = 1/alpha
        // with the Clausius-Mossotti polarizability of m = 1.5 + 0.1i
        std::vector<int> address;
        for (int c = 0; c < C; c += 3) address.push_back(c);
        R.i(t + "address", address);
        const int N = (int)address.size();
        const cd m2 = cd(1.5, 0.1) * cd(1.5, 0.1);
        const cd alpha = 3.0 * lf * lf * lf / (4.0 * 3.14159265358979323846) * (m2 - 1.0) / (m2 + 2.0);
        VectorXcd DIAG_A((int64_t)f * N);
        for (int m = 0; m < N; ++m)
            for (int q = 0; q < f; ++q) DIAG_A((int64_t)f * m + q) = 1.0 / alpha;
        R.c(t + "diag", DIAG_A);
        const int L = 2;
        MatrixXcd B = g.vec((int64_t)f * N, L);
        R.c(t + "B", B);
        const double tol = 1e-9;
        const int itermax = 300;
        if (f == 3) {
            R.c(t + "cocg_rq", bl_cocg_rq_mvp_fft(n, f, address, Au_til, DIAG_A, B, tol, itermax, plan_fwd, plan_inv));
            R.c(t + "cocg_rq_jacobi",
                bl_cocg_rq_jacobi_mvp_fft(n, f, address, Au_til, DIAG_A, B, tol, itermax, plan_fwd, plan_inv));
        }
        R.c(t + "bicgstab", bl_bicgstab_mvp_fft(n, f, address, Au_til, DIAG_A, B, tol, itermax, plan_fwd, plan_inv));
        R.c(t + "bicgstab_jacobi",
            bl_bicgstab_jacobi_mvp_fft(n, f, address, Au_til, DIAG_A, B, tol, itermax, plan_fwd, plan_inv));
        R.c(t + "bicggr", bl_bicggr_mvp_fft(n, f, address, Au_til, DIAG_A, B, tol, itermax, plan_fwd, plan_inv));
        R.c(t + "bicggr_jacobi",
            bl_bicggr_jacobi_mvp_fft(n, f, address, Au_til, DIAG_A, B, tol, itermax, plan_fwd, plan_inv));
    }
    const MatrixXcd A = g.vec(60, 5);
    R.c("qr_A", A);
    MatrixXcd Q, Rm;
    std::tie(Q, Rm) = qr_reduced(A);
    R.c("qr_Q", Q);
    R.c("qr_R", Rm);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "layout" && argc > 2) return run_layout(argv[2]);
    if (mode == "gpu" && argc > 2) return run_gpu(argv[2]);
    if (mode == "badf") {
        MatrixXi ind(2, 3);
        for (int r = 0; r < 2; ++r)
            for (int l = 0; l < 3; ++l) ind(r, l) = 1 + r * (l == 0);
        application_function(lattice(2, 2, 2), 4, ind, 4, 1.0, cd(1.0, 0.0));
        std::puts("not reached");
        return 0;
    }
    if (mode == "badtoken") {
        bdda_cxx::Plan p;
        const VectorXcd fake(4);
        MBT_fft_mvp(lattice(2, 2, 2), 3, fake, VectorXcd(24), p, p);
        std::puts("not reached");
        return 0;
    }
    std::fprintf(stderr, "usage: mirror_check layout|gpu <file> | badf | badtoken\n");
    return 2;
}
