// solvers.hip -- the reference's six block Krylov solvers on the device
// operator, plus FP32-inner / FP64-outer defect correction.
//
// Recurrences follow, line by line (paths relative to the reference root):
//   bl_cocg_rq_mvp_fft.cpp:22-69, bl_cocg_rq_jacobi_mvp_fft.cpp:22-76
//   bl_bicgstab_mvp_fft.cpp:21-94, bl_bicgstab_jacobi_mvp_fft.cpp:20-95
//   bl_bicggr_mvp_fft.cpp:18-113,  bl_bicggr_jacobi_mvp_fft.cpp:18-116
// Block vectors live on the device (FP32 or FP64); Gram matrices, traces and
// norms are reduced in FP64 and the s x s algebra (LU, products) runs in
// FP64 on the host exactly where the reference calls Eigen.
//
// Documented differences (equal in exact arithmetic):
//   - qr_reduced inside the loops is Cholesky-QR applied twice with a
//     column Gram-Schmidt fallback (same Q, R as MGS with positive diag R);
//   - X += S*alpha*xi is evaluated as X += S*(alpha*xi);
//   - an exactly singular s x s LU pivot or a zero QR column stops the
//     solve with BDDA_EBREAKDOWN instead of producing NaN.
#include "solvers.hpp"
#include <chrono>
#include <cmath>
#include <cstring>
#include <utility>

namespace bdda {

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    explicit DevBuf(size_t count) : n(count)
    {
        if (count) BDDA_HIP(hipMalloc(&p, sizeof(T) * count));
    }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

Problem::~Problem()
{
    for (void* p : {(void*)diag64, (void*)jpre64, (void*)diag32, (void*)jpre32})
        if (p) (void)hipFree(p);
}

template <>
const c64* Problem::diag<double>() const { return diag64; }
template <>
const c32* Problem::diag<float>() const { return diag32; }
template <>
const c64* Problem::jpre<double>() const { return jpre64; }
template <>
const c32* Problem::jpre<float>() const { return jpre32; }

void problem_set_diag(Problem& P, const c64* diag_host)
{
    const int64_t n = P.rows();
    std::vector<c64> jp(n);
    std::vector<c32> d32(n), j32(n);
    for (int64_t i = 0; i < n; ++i) {
        const c64 d = diag_host[i];
        const double den = d.x * d.x + d.y * d.y;
        jp[i] = mk(d.x / den, -d.y / den);  // jpre = 1 / DIAG_A
        d32[i] = cvt<float>(d);
        j32[i] = cvt<float>(jp[i]);
    }
    for (void* p : {(void*)P.diag64, (void*)P.jpre64, (void*)P.diag32, (void*)P.jpre32})
        if (p) BDDA_HIP(hipFree(p));
    P.diag64 = P.jpre64 = nullptr;
    P.diag32 = P.jpre32 = nullptr;
    BDDA_HIP(hipMalloc(&P.diag64, sizeof(c64) * n));
    BDDA_HIP(hipMalloc(&P.jpre64, sizeof(c64) * n));
    BDDA_HIP(hipMalloc(&P.diag32, sizeof(c32) * n));
    BDDA_HIP(hipMalloc(&P.jpre32, sizeof(c32) * n));
    BDDA_HIP(hipMemcpy(P.diag64, diag_host, sizeof(c64) * n, hipMemcpyHostToDevice));
    BDDA_HIP(hipMemcpy(P.jpre64, jp.data(), sizeof(c64) * n, hipMemcpyHostToDevice));
    BDDA_HIP(hipMemcpy(P.diag32, d32.data(), sizeof(c32) * n, hipMemcpyHostToDevice));
    BDDA_HIP(hipMemcpy(P.jpre32, j32.data(), sizeof(c32) * n, hipMemcpyHostToDevice));
}

template <typename R>
void problem_matvec(Problem& P, const cplx<R>* X, cplx<R>* Y, int s, int jacobi)
{
    mbt_apply_occupied<R>(P.op, X, P.rows(), Y, P.rows(), s, P.diag<R>(), jacobi ? P.jpre<R>() : nullptr, P.nocc);
}
template void problem_matvec<float>(Problem&, const c32*, c32*, int, int);
template void problem_matvec<double>(Problem&, const c64*, c64*, int, int);

namespace {

template <typename R>
void copy_block(BlockWork& w, const cplx<R>* A, cplx<R>* Z, int64_t rows, int s)
{
    BDDA_HIP(hipMemcpyAsync(Z, A, sizeof(cplx<R>) * rows * s, hipMemcpyDeviceToDevice, w.stream));
}

inline HostMat neg(const HostMat& A)
{
    HostMat B(A.size());
    for (size_t i = 0; i < A.size(); ++i) B[i] = mk(-A[i].x, -A[i].y);
    return B;
}

inline HostMat scal(const HostMat& A, c64 a)
{
    HostMat B(A.size());
    for (size_t i = 0; i < A.size(); ++i) B[i] = a * A[i];
    return B;
}

inline c64 cdiv(c64 a, c64 b)
{
    const double den = b.x * b.x + b.y * b.y;
    return mk((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}

struct Ctx {
    Problem& P;
    int jacobi;
    int s;
    int64_t rows;
    SolveInfo& info;
    std::vector<double>* hist;
};

// ------------------------------------------------------------ COCG-rq
template <typename R>
int run_cocg_rq(Ctx& c, const cplx<R>* B, cplx<R>* X, double tol, int itermax)
{
    BlockWork& w = c.P.work;
    const int s = c.s;
    const int64_t rows = c.rows, N = rows * s;
    DevBuf<cplx<R>> Bp(c.jacobi ? N : 0), Q(N), Qn(N), S(N), AS(N), T(N), scr(N);
    const cplx<R>* Bu = B;
    if (c.jacobi) {  // B_pre = jpre .* B  (bl_cocg_rq_jacobi_mvp_fft.cpp:22-27)
        scale_rows<R>(w.stream, c.P.jpre<R>(), B, rows, Bp.p, rows, rows, s);
        Bu = Bp.p;
    }
    const double bnorm = norm_fro<R>(w, Bu, rows, rows, s);
    fill_zero<R>(w.stream, X, rows, rows, s);
    HostMat xi, tau;
    int rc = qr_fast<R>(w, Bu, rows, Q.p, rows, rows, s, xi, scr.p);
    if (rc) return rc;
    copy_block<R>(w, Q.p, S.p, rows, s);
    const int full = c.jacobi ? 0 : 1;  // FullPivLU (plain) vs PartialPivLU (jacobi)
    HostMat G2 = gram<R>(w, Q.p, rows, s, Q.p, rows, s, rows, 0);  // Q^T Q
    cplx<R>* pS = S.p;
    cplx<R>* pT = T.p;
    cplx<R>* pQ = Q.p;
    cplx<R>* pQn = Qn.p;
    for (int k = 0; k < itermax; ++k) {
        problem_matvec<R>(c.P, pS, AS.p, s, c.jacobi);
        ++c.info.matvecs;
        HostMat G1 = gram<R>(w, pS, rows, s, AS.p, rows, s, rows, 0);  // S^T AS
        HostMat alpha;
        if (!hm_lu_solve(G1, G2, s, s, full, alpha)) return BDDA_EBREAKDOWN;
        HostMat axi = hm_mul(alpha, xi, s, s, s);
        {   // X = X + S alpha xi
            Combo<R> cb;
            cb.nv = 1; cb.V[0] = X; cb.ldv[0] = rows; cb.a[0] = mk(1.0, 0.0);
            cb.nm = 1; cb.W[0] = pS; cb.ldw[0] = rows; cb.kw[0] = s; cb.M[0] = axi.data();
            combo<R>(w, cb, X, rows, rows, s);
        }
        HostMat nalpha = neg(alpha);
        {   // T = Q - AS alpha
            Combo<R> cb;
            cb.nv = 1; cb.V[0] = pQ; cb.ldv[0] = rows; cb.a[0] = mk(1.0, 0.0);
            cb.nm = 1; cb.W[0] = AS.p; cb.ldw[0] = rows; cb.kw[0] = s; cb.M[0] = nalpha.data();
            combo<R>(w, cb, pT, rows, rows, s);
        }
        rc = qr_fast<R>(w, pT, rows, pQn, rows, rows, s, tau, scr.p);
        if (rc) return rc;
        xi = hm_mul(tau, xi, s, s, s);
        const double err = hm_fro(xi) / bnorm;
        c.info.iters = k + 1;
        c.info.recurrence_res = err;
        if (c.hist) c.hist->push_back(err);
        if (!std::isfinite(err)) return BDDA_ENAN;
        if (err < tol) {
            c.info.converged = 1;
            break;
        }
        HostMat G3 = gram<R>(w, pQn, rows, s, pQn, rows, s, rows, 0);  // Qnew^T Qnew
        HostMat rhs = hm_tmul(tau, G3, s, s, s);                         // tau^T (Qnew^T Qnew)
        HostMat beta;
        if (!hm_lu_solve(G2, rhs, s, s, full, beta)) return BDDA_EBREAKDOWN;
        {   // S = Qnew + S beta
            Combo<R> cb;
            cb.nv = 1; cb.V[0] = pQn; cb.ldv[0] = rows; cb.a[0] = mk(1.0, 0.0);
            cb.nm = 1; cb.W[0] = pS; cb.ldw[0] = rows; cb.kw[0] = s; cb.M[0] = beta.data();
            combo<R>(w, cb, pT, rows, rows, s);
        }
        std::swap(pS, pT);
        std::swap(pQ, pQn);
        G2 = std::move(G3);
    }
    BDDA_HIP(hipStreamSynchronize(w.stream));
    return BDDA_OK;
}

// ------------------------------------------------------------ BiCGSTAB
template <typename R>
int run_bicgstab(Ctx& c, const cplx<R>* B, cplx<R>* X, double tol, int itermax)
{
    BlockWork& w = c.P.work;
    const int s = c.s;
    const int64_t rows = c.rows, N = rows * s;
    DevBuf<cplx<R>> Rb(N), Pb(N), Pn(N), R0(N), V(N), T(N), Z(N);
    if (c.jacobi) scale_rows<R>(w.stream, c.P.jpre<R>(), B, rows, Rb.p, rows, rows, s);
    else copy_block<R>(w, B, Rb.p, rows, s);
    const double bnorm = norm_fro<R>(w, Rb.p, rows, rows, s);
    fill_zero<R>(w.stream, X, rows, rows, s);
    copy_block<R>(w, Rb.p, Pb.p, rows, s);
    copy_block<R>(w, Rb.p, R0.p, rows, s);
    cplx<R>* pP = Pb.p;
    cplx<R>* pPn = Pn.p;
    for (int k = 0; k < itermax; ++k) {
        problem_matvec<R>(c.P, pP, V.p, s, c.jacobi);  // V = A P
        ++c.info.matvecs;
        HostMat M1 = gram<R>(w, R0.p, rows, s, V.p, rows, s, rows, 1);   // R0^H V
        HostMat M2 = gram<R>(w, R0.p, rows, s, Rb.p, rows, s, rows, 1);  // R0^H R
        HostMat alpha;
        if (!hm_lu_solve(M1, M2, s, s, 0, alpha)) return BDDA_EBREAKDOWN;
        HostMat nalpha = neg(alpha);
        {   // T = R - V alpha
            Combo<R> cb;
            cb.nv = 1; cb.V[0] = Rb.p; cb.ldv[0] = rows; cb.a[0] = mk(1.0, 0.0);
            cb.nm = 1; cb.W[0] = V.p; cb.ldw[0] = rows; cb.kw[0] = s; cb.M[0] = nalpha.data();
            combo<R>(w, cb, T.p, rows, rows, s);
        }
        problem_matvec<R>(c.P, T.p, Z.p, s, c.jacobi);  // Z = A T
        ++c.info.matvecs;
        const c64 zt = dot_all<R>(w, Z.p, rows, T.p, rows, rows, s, 1);
        const c64 zz = dot_all<R>(w, Z.p, rows, Z.p, rows, rows, s, 1);
        const c64 qsi = cdiv(zt, zz);
        {   // X = X + P alpha + qsi T
            Combo<R> cb;
            cb.nv = 2; cb.V[0] = X; cb.ldv[0] = rows; cb.a[0] = mk(1.0, 0.0);
            cb.V[1] = T.p; cb.ldv[1] = rows; cb.a[1] = qsi;
            cb.nm = 1; cb.W[0] = pP; cb.ldw[0] = rows; cb.kw[0] = s; cb.M[0] = alpha.data();
            combo<R>(w, cb, X, rows, rows, s);
        }
        {   // R = T - qsi Z
            Combo<R> cb;
            cb.nv = 2; cb.V[0] = T.p; cb.ldv[0] = rows; cb.a[0] = mk(1.0, 0.0);
            cb.V[1] = Z.p; cb.ldv[1] = rows; cb.a[1] = mk(-qsi.x, -qsi.y);
            combo<R>(w, cb, Rb.p, rows, rows, s);
        }
        const double err = norm_fro<R>(w, Rb.p, rows, rows, s) / bnorm;
        c.info.iters = k + 1;
        c.info.recurrence_res = err;
        if (c.hist) c.hist->push_back(err);
        if (!std::isfinite(err)) return BDDA_ENAN;
        if (err < tol) {
            c.info.converged = 1;
            break;
        }
        HostMat M3 = gram<R>(w, R0.p, rows, s, Z.p, rows, s, rows, 1);  // beta = LU(R0^H V)^{-1} (-R0^H Z)
        HostMat beta;
        if (!hm_lu_solve(M1, neg(M3), s, s, 0, beta)) return BDDA_EBREAKDOWN;
        HostMat qb = scal(beta, mk(-qsi.x, -qsi.y));
        {   // P = R + (P - qsi V) beta
            Combo<R> cb;
            cb.nv = 1; cb.V[0] = Rb.p; cb.ldv[0] = rows; cb.a[0] = mk(1.0, 0.0);
            cb.nm = 2;
            cb.W[0] = pP; cb.ldw[0] = rows; cb.kw[0] = s; cb.M[0] = beta.data();
            cb.W[1] = V.p; cb.ldw[1] = rows; cb.kw[1] = s; cb.M[1] = qb.data();
            combo<R>(w, cb, pPn, rows, rows, s);
        }
        std::swap(pP, pPn);
    }
    BDDA_HIP(hipStreamSynchronize(w.stream));
    return BDDA_OK;
}

// ------------------------------------------------------------ BiCGGR
template <typename R>
int run_bicggr(Ctx& c, const cplx<R>* B, cplx<R>* X, double tol, int itermax)
{
    BlockWork& w = c.P.work;
    const int s = c.s;
    const int64_t rows = c.rows, N = rows * s;
    DevBuf<cplx<R>> Rb(N), Rn(N), Pb(N), R0(N), V(N), W(N), U(N), Y(N);
    if (c.jacobi) scale_rows<R>(w.stream, c.P.jpre<R>(), B, rows, Rb.p, rows, rows, s);
    else copy_block<R>(w, B, Rb.p, rows, s);
    const double bnorm = norm_fro<R>(w, Rb.p, rows, rows, s);
    fill_zero<R>(w.stream, X, rows, rows, s);
    copy_block<R>(w, Rb.p, Pb.p, rows, s);
    problem_matvec<R>(c.P, Rb.p, V.p, s, c.jacobi);  // V = A R
    ++c.info.matvecs;
    copy_block<R>(w, V.p, W.p, rows, s);
    copy_block<R>(w, Rb.p, R0.p, rows, s);
    cplx<R>* pR = Rb.p;
    cplx<R>* pRn = Rn.p;
    for (int k = 0; k < itermax; ++k) {
        HostMat M1 = gram<R>(w, R0.p, rows, s, V.p, rows, s, rows, 1);
        HostMat M2 = gram<R>(w, R0.p, rows, s, pR, rows, s, rows, 1);
        HostMat alpha;
        if (!hm_lu_solve(M1, M2, s, s, 0, alpha)) return BDDA_EBREAKDOWN;
        const c64 wr = dot_all<R>(w, W.p, rows, pR, rows, rows, s, 1);
        const c64 ww = dot_all<R>(w, W.p, rows, W.p, rows, rows, s, 1);
        const c64 qsi = cdiv(wr, ww);
        HostMat qa = scal(alpha, mk(-qsi.x, -qsi.y));
        {   // U = (P - qsi V) alpha
            Combo<R> cb;
            cb.nm = 2;
            cb.W[0] = Pb.p; cb.ldw[0] = rows; cb.kw[0] = s; cb.M[0] = alpha.data();
            cb.W[1] = V.p; cb.ldw[1] = rows; cb.kw[1] = s; cb.M[1] = qa.data();
            combo<R>(w, cb, U.p, rows, rows, s);
        }
        problem_matvec<R>(c.P, U.p, Y.p, s, c.jacobi);  // Y = A U
        ++c.info.matvecs;
        {   // X = X + qsi R + U
            Combo<R> cb;
            cb.nv = 3; cb.V[0] = X; cb.ldv[0] = rows; cb.a[0] = mk(1.0, 0.0);
            cb.V[1] = pR; cb.ldv[1] = rows; cb.a[1] = qsi;
            cb.V[2] = U.p; cb.ldv[2] = rows; cb.a[2] = mk(1.0, 0.0);
            combo<R>(w, cb, X, rows, rows, s);
        }
        {   // Rnew = R - qsi W - Y
            Combo<R> cb;
            cb.nv = 3; cb.V[0] = pR; cb.ldv[0] = rows; cb.a[0] = mk(1.0, 0.0);
            cb.V[1] = W.p; cb.ldv[1] = rows; cb.a[1] = mk(-qsi.x, -qsi.y);
            cb.V[2] = Y.p; cb.ldv[2] = rows; cb.a[2] = mk(-1.0, 0.0);
            combo<R>(w, cb, pRn, rows, rows, s);
        }
        const double err = norm_fro<R>(w, pRn, rows, rows, s) / bnorm;
        c.info.iters = k + 1;
        c.info.recurrence_res = err;
        if (c.hist) c.hist->push_back(err);
        if (!std::isfinite(err)) return BDDA_ENAN;
        if (err < tol) {
            c.info.converged = 1;
            break;
        }
        problem_matvec<R>(c.P, pRn, W.p, s, c.jacobi);  // W = A Rnew
        ++c.info.matvecs;
        HostMat M3 = gram<R>(w, R0.p, rows, s, pR, rows, s, rows, 1);   // R0^H R
        HostMat M4 = gram<R>(w, R0.p, rows, s, pRn, rows, s, rows, 1);  // R0^H Rnew / qsi
        const c64 iq = cdiv(mk(1.0, 0.0), qsi);
        (void)iq;
        for (auto& v : M4) v = cdiv(v, qsi);
        HostMat gamma;
        if (!hm_lu_solve(M3, M4, s, s, 0, gamma)) return BDDA_EBREAKDOWN;
        std::swap(pR, pRn);
        {   // P = R + U gamma
            Combo<R> cb;
            cb.nv = 1; cb.V[0] = pR; cb.ldv[0] = rows; cb.a[0] = mk(1.0, 0.0);
            cb.nm = 1; cb.W[0] = U.p; cb.ldw[0] = rows; cb.kw[0] = s; cb.M[0] = gamma.data();
            combo<R>(w, cb, Pb.p, rows, rows, s);
        }
        {   // V = W + Y gamma
            Combo<R> cb;
            cb.nv = 1; cb.V[0] = W.p; cb.ldv[0] = rows; cb.a[0] = mk(1.0, 0.0);
            cb.nm = 1; cb.W[0] = Y.p; cb.ldw[0] = rows; cb.kw[0] = s; cb.M[0] = gamma.data();
            combo<R>(w, cb, V.p, rows, rows, s);
        }
    }
    BDDA_HIP(hipStreamSynchronize(w.stream));
    return BDDA_OK;
}

template <typename R>
int run_method(Ctx& c, int method, const cplx<R>* B, cplx<R>* X, double tol, int itermax)
{
    switch (method) {
    case M_COCG_RQ: return run_cocg_rq<R>(c, B, X, tol, itermax);
    case M_BICGSTAB: return run_bicgstab<R>(c, B, X, tol, itermax);
    case M_BICGGR: return run_bicggr<R>(c, B, X, tol, itermax);
    default: return BDDA_EINVAL;
    }
}

}  // namespace

double true_residual(Problem& P, const c64* B64, const c64* X64, int s)
{
    const int64_t rows = P.rows(), N = rows * s;
    BlockWork& w = P.work;
    DevBuf<c64> AX(N);
    problem_matvec<double>(P, X64, AX.p, s, 0);
    Combo<double> cb;
    cb.nv = 2;
    cb.V[0] = B64; cb.ldv[0] = rows; cb.a[0] = mk(1.0, 0.0);
    cb.V[1] = AX.p; cb.ldv[1] = rows; cb.a[1] = mk(-1.0, 0.0);
    combo<double>(w, cb, AX.p, rows, rows, s);
    const double rn = norm_fro<double>(w, AX.p, rows, rows, s);
    const double bn = norm_fro<double>(w, B64, rows, rows, s);
    return rn / bn;
}

void residual_columns(Problem& P, const c64* B64, const c64* X64, int s, double* out)
{
    const int64_t rows = P.rows(), N = rows * s;
    BlockWork& w = P.work;
    DevBuf<c64> AX(N);
    problem_matvec<double>(P, X64, AX.p, s, 0);
    Combo<double> cb;
    cb.nv = 2;
    cb.V[0] = B64; cb.ldv[0] = rows; cb.a[0] = mk(1.0, 0.0);
    cb.V[1] = AX.p; cb.ldv[1] = rows; cb.a[1] = mk(-1.0, 0.0);
    combo<double>(w, cb, AX.p, rows, rows, s);
    // column norms from the diagonals of R^H R and B^H B
    const HostMat gr = gram<double>(w, AX.p, rows, s, AX.p, rows, s, rows, 1);
    const HostMat gb = gram<double>(w, B64, rows, s, B64, rows, s, rows, 1);
    for (int j = 0; j < s; ++j) {
        const double rn = std::sqrt(std::max(gr[(size_t)j * s + j].x, 0.0));
        const double bn = std::sqrt(std::max(gb[(size_t)j * s + j].x, 0.0));
        out[j] = bn > 0 ? rn / bn : rn;
    }
}

int solve_fp(Problem& P, int method, int jacobi, int mode, const c64* B64, c64* X64, int s, double tol,
             int itermax, SolveInfo& info, std::vector<double>* hist, double inner_tol)
{
    const auto t0 = std::chrono::steady_clock::now();
    info = SolveInfo{};
    const int64_t rows = P.rows(), N = rows * s;
    BlockWork& w = P.work;
    Ctx c{P, jacobi, s, rows, info, hist};
    int rc = BDDA_OK;
    if (mode == MODE_FP64) {
        rc = run_method<double>(c, method, B64, X64, tol, itermax);
        info.outer_iters = 1;
    } else if (mode == MODE_FP32) {
        DevBuf<c32> B32(N), X32(N);
        convert_64_to_32(w.stream, B64, rows, B32.p, rows, rows, s, 1.0);
        rc = run_method<float>(c, method, B32.p, X32.p, tol, itermax);
        fill_zero<double>(w.stream, X64, rows, rows, s);
        accumulate_32_to_64(w.stream, X32.p, rows, X64, rows, rows, s, 1.0);
        info.outer_iters = 1;
    } else if (mode == MODE_MIXED) {
        // FP64 defect correction: r = B - A x (FP64 matvec), A d = r in FP32,
        // x += d, until ||B - A x|| / ||B|| <= tol.
        DevBuf<c64> R64(N), AX(N);
        DevBuf<c32> R32(N), D32(N);
        const double bn = norm_fro<double>(w, B64, rows, rows, s);
        fill_zero<double>(w.stream, X64, rows, rows, s);
        BDDA_HIP(hipMemcpyAsync(R64.p, B64, sizeof(c64) * N, hipMemcpyDeviceToDevice, w.stream));
        double rel = 1.0;
        const int max_outer = 20;
        int total_iters = 0;
        // every pass measures the residual of the current x first, so the
        // reported true residual and convergence always describe the X
        // returned, also after the last correction
        for (int o = 0;; ++o) {
            const double rn = norm_fro<double>(w, R64.p, rows, rows, s);
            rel = rn / bn;
            info.true_res = rel;
            if (hist) hist->push_back(-rel);  // negative marks an outer (true) residual
            if (rel <= tol) {
                info.converged = 1;
                rc = BDDA_OK;
                break;
            }
            if (o == max_outer || (o > 0 && total_iters >= itermax)) {
                rc = BDDA_ENOCONV;
                break;
            }
            if (rc == BDDA_ENOCONV) break;  // the inner solve ran out of iterations: x is its best
            // inner FP32 solve to half the outer target: the FP32 recurrence
            // residual runs below the true one by up to about 2x near 1e-5
            // (measured on the 16384-monomer aggregate: 4.8e-6 recurrence ->
            // 9.2e-6 true), so half the target usually needs no second outer
            // step, while a tighter one costs iterations in the FP32 floor
            double itol = inner_tol > 0 ? inner_tol : 0.5 * tol / rel;
            itol = std::min(std::max(itol, 2e-6), 0.5);
            convert_64_to_32(w.stream, R64.p, rows, R32.p, rows, rows, s, 1.0 / rn);
            SolveInfo inner;
            Ctx ci{P, jacobi, s, rows, inner, hist};
            rc = run_method<float>(ci, method, R32.p, D32.p, itol, std::max(1, itermax - total_iters));
            total_iters += inner.iters;
            info.iters += inner.iters;
            info.matvecs += inner.matvecs;
            info.recurrence_res = inner.recurrence_res;
            ++info.outer_iters;
            if (rc != BDDA_OK && rc != BDDA_ENOCONV) {
                // breakdown or NaN in the FP32 solve: its correction is not
                // usable.  After a first correction x is still the best
                // iterate, so return it with its measured residual.
                if (o > 0) rc = BDDA_ENOCONV;
                break;
            }
            accumulate_32_to_64(w.stream, D32.p, rows, X64, rows, rows, s, rn);
            problem_matvec<double>(P, X64, AX.p, s, 0);  // r = B - A x
            ++info.matvecs;
            Combo<double> cb;
            cb.nv = 2;
            cb.V[0] = B64; cb.ldv[0] = rows; cb.a[0] = mk(1.0, 0.0);
            cb.V[1] = AX.p; cb.ldv[1] = rows; cb.a[1] = mk(-1.0, 0.0);
            combo<double>(w, cb, R64.p, rows, rows, s);
        }
        BDDA_HIP(hipStreamSynchronize(w.stream));
    } else {
        rc = BDDA_EINVAL;
    }
    const auto t1 = std::chrono::steady_clock::now();
    info.time_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    info.status = rc;
    return rc;
}

}  // namespace bdda
