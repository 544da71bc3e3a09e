// solvers.hip -- the reference's six block Krylov solvers on the device
// operator, plus FP32-inner / FP64-outer defect correction.
//
// Recurrences follow, line by line (paths relative to the reference root):
//   bl_cocg_rq_mvp_fft.cpp:22-69, bl_cocg_rq_jacobi_mvp_fft.cpp:22-76
//   bl_bicgstab_mvp_fft.cpp:21-94, bl_bicgstab_jacobi_mvp_fft.cpp:20-95
//   bl_bicggr_mvp_fft.cpp:18-113,  bl_bicggr_jacobi_mvp_fft.cpp:18-116
// Block vectors live on the device (FP32 or FP64); Gram matrices, traces and
// norms are reduced in FP64, and the s x s algebra (LU, products, the
// Cholesky factors of the QR) runs in FP64 on the device (smallmat.hip)
// exactly where the reference calls Eigen.  An iteration is enqueued
// without waiting on the host; the host reads one value per iteration, the
// residual norm of the convergence test, with the breakdown flags.
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
#include <cstdlib>
#include <cstring>
#include <utility>

namespace bdda {

// a solver's device block, freed after the work queued on its stream
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipStream_t st = nullptr;
    DevBuf(size_t count, hipStream_t stream) : n(count), st(stream)
    {
        if (count) p = dev_alloc_n<T>(count);
    }
    ~DevBuf() { dev_free(p, st); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

Problem::~Problem() { release(); }

void Problem::release()
{
    for (void* p : {(void*)diag64, (void*)jpre64, (void*)diag32, (void*)jpre32}) dev_free(p, work.stream);
    diag64 = jpre64 = nullptr;
    diag32 = jpre32 = nullptr;
    work.release();
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
    for (void* p : {(void*)P.diag64, (void*)P.jpre64, (void*)P.diag32, (void*)P.jpre32}) dev_free(p, P.work.stream);
    P.diag64 = P.jpre64 = nullptr;
    P.diag32 = P.jpre32 = nullptr;
    P.diag64 = dev_alloc_n<c64>(n, true);
    P.jpre64 = dev_alloc_n<c64>(n, true);
    P.diag32 = dev_alloc_n<c32>(n, true);
    P.jpre32 = dev_alloc_n<c32>(n, true);
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

struct Ctx {
    Problem& P;
    int jacobi;
    int s;
    int64_t rows;
    SolveInfo& info;
    std::vector<double>* hist;
};




// ------------------------------------------------------------ device s x s state
// The solvers below keep every s x s quantity (Grams, LU solutions, the QR's
// R, traces) in device memory and enqueue a whole iteration without a host
// round trip; the host reads one probe per iteration (the residual norm the
// convergence test needs and the breakdown flags).
struct SmallState {
    BlockWork& w;
    int s;
    DevBuf<c64> buf;
    c64* probe;  // [0].x: the value read per iteration; [1]: flag word
    int* flag;
    SmallState(BlockWork& w_, int s_, int nmat, int64_t rows) : w(w_), s(s_), buf((size_t)nmat * s_ * s_ + 4, w_.stream)
    {
        probe = buf.p + (size_t)nmat * s * s;
        flag = reinterpret_cast<int*>(probe + 1);
        // partial sums for every Gram and dot of the solve, reserved up front:
        // a reallocation would free memory that queued kernels still use
        const size_t np = (size_t)s * s;
        w.reserve(std::max<size_t>(1024, (s <= 16 ? 1024 : 256) * np), std::max<size_t>(np, 1));
        (void)rows;
        fill_zero<double>(w.stream, probe, 2, 2, 1);
    }
    c64* m(int i) { return buf.p + (size_t)i * s * s; }
    struct Read {
        double v;
        int flags;
    };
    Read read()
    {
        BDDA_HIP(hipMemcpyAsync(w.host_pinned + w.host_elems, probe, 2 * sizeof(c64), hipMemcpyDeviceToHost,
                                w.stream));
        BDDA_HIP(hipStreamSynchronize(w.stream));
        Read r;
        r.v = w.host_pinned[w.host_elems].x;
        std::memcpy(&r.flags, &w.host_pinned[w.host_elems + 1], sizeof(int));
        return r;
    }
    // host matrix -> device (stream idle first: the solvers call this only
    // at set-up and on the Cholesky fallback)
    void upload(c64* dst, const HostMat& h)
    {
        BDDA_HIP(hipStreamSynchronize(w.stream));
        BDDA_HIP(hipMemcpy(dst, h.data(), sizeof(c64) * h.size(), hipMemcpyHostToDevice));
    }
    void clear_flag(int bits)
    {
        int f = 0;
        BDDA_HIP(hipStreamSynchronize(w.stream));
        BDDA_HIP(hipMemcpy(&f, flag, sizeof(int), hipMemcpyDeviceToHost));
        f &= ~bits;
        BDDA_HIP(hipMemcpy(flag, &f, sizeof(int), hipMemcpyHostToDevice));
    }
};

// ------------------------------------------------------------ COCG-rq
template <typename R>
int run_cocg_rq(Ctx& c, const cplx<R>* B, cplx<R>* X, double tol, int itermax)
{
    BlockWork& w = c.P.work;
    const int s = c.s;
    const int64_t rows = c.rows, N = rows * s;
    const c64 one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
    DevBuf<cplx<R>> Bp(c.jacobi ? N : 0, w.stream), Q(N, w.stream), Qn(N, w.stream), S(N, w.stream), AS(N, w.stream), T(N, w.stream), scr(N, w.stream);
    const cplx<R>* Bu = B;
    if (c.jacobi) {  // B_pre = jpre .* B  (bl_cocg_rq_jacobi_mvp_fft.cpp:22-27)
        scale_rows<R>(w.stream, c.P.jpre<R>(), B, rows, Bp.p, rows, rows, s);
        Bu = Bp.p;
    }
    const double bnorm = norm_fro<R>(w, Bu, rows, rows, s);
    fill_zero<R>(w.stream, X, rows, rows, s);
    HostMat xi0;
    int rc = qr_fast<R>(w, Bu, rows, Q.p, rows, rows, s, xi0, scr.p);
    if (rc) return rc;
    copy_block<R>(w, Q.p, S.p, rows, s);
    const int full = c.jacobi ? 0 : 1;  // FullPivLU (plain) vs PartialPivLU (jacobi)
    SmallState sm(w, s, 14, rows);
    c64 *G1 = sm.m(0), *G2 = sm.m(1), *G3 = sm.m(2), *alpha = sm.m(3), *axi = sm.m(4), *xi = sm.m(5);
    c64 *xin = sm.m(6), *tau = sm.m(7), *rhs = sm.m(8), *beta = sm.m(9), *qtmp = sm.m(10);  // qtmp: 4 matrices
    sm.upload(xi, xi0);
    gram_dev<R>(w, Q.p, rows, s, Q.p, rows, s, rows, 0, G2);  // Q^T Q
    cplx<R>* pS = S.p;
    cplx<R>* pT = T.p;
    cplx<R>* pQ = Q.p;
    cplx<R>* pQn = Qn.p;
    for (int k = 0; k < itermax; ++k) {
        problem_matvec<R>(c.P, pS, AS.p, s, c.jacobi);
        ++c.info.matvecs;
        gram_dev<R>(w, pS, rows, s, AS.p, rows, s, rows, 0, G1);  // S^T AS
        dm_lu_solve(w.stream, G1, G2, alpha, s, s, full, one, nullptr, sm.flag, BDDA_SM_LU);
        dm_mul(w.stream, alpha, 0, xi, axi, s, s, s, one, nullptr, nullptr, zero);
        {   // X = X + S (alpha xi)
            Combo<R> cb;
            cb.nv = 1; cb.V[0] = X; cb.ldv[0] = rows; cb.a[0] = one;
            cb.nm = 1; cb.W[0] = pS; cb.ldw[0] = rows; cb.kw[0] = s; cb.Md[0] = axi;
            combo<R>(w, cb, X, rows, rows, s);
        }
        {   // T = Q - AS alpha
            Combo<R> cb;
            cb.nv = 1; cb.V[0] = pQ; cb.ldv[0] = rows; cb.a[0] = one;
            cb.nm = 1; cb.W[0] = AS.p; cb.ldw[0] = rows; cb.kw[0] = s; cb.Md[0] = alpha; cb.ms[0] = mk(-1.0, 0.0);
            combo<R>(w, cb, pT, rows, rows, s);
        }
        qr_fast_dev<R>(w, pT, rows, pQn, rows, rows, s, tau, scr.p, qtmp, sm.flag);
        // xi = tau xi, its norm, and (for the next direction) beta = (Q^T Q)^{-1} tau^T (Qnew^T Qnew)
        auto tail = [&]() {
            dm_mul(w.stream, tau, 0, xi, xin, s, s, s, one, nullptr, nullptr, zero);
            dm_fro(w.stream, xin, s * s, &sm.probe[0].x);
            gram_dev<R>(w, pQn, rows, s, pQn, rows, s, rows, 0, G3);
            dm_mul(w.stream, tau, 1, G3, rhs, s, s, s, one, nullptr, nullptr, zero);
            dm_lu_solve(w.stream, G2, rhs, beta, s, s, full, one, nullptr, sm.flag, BDDA_SM_LU2);
        };
        tail();
        SmallState::Read rd = sm.read();
        if (rd.flags & BDDA_SM_CHOL) {
            // a Cholesky pivot of the QR fell below its floor: factor T again
            // on the host path (Cholesky checks, then column Gram-Schmidt)
            HostMat tauh;
            rc = qr_fast<R>(w, pT, rows, pQn, rows, rows, s, tauh, scr.p);
            if (rc) return rc;
            sm.upload(tau, tauh);
            sm.clear_flag(BDDA_SM_CHOL);
            tail();
            rd = sm.read();
        }
        if (rd.flags & BDDA_SM_LU) return BDDA_EBREAKDOWN;
        const double err = rd.v / bnorm;
        c.info.iters = k + 1;
        c.info.recurrence_res = err;
        if (c.hist) c.hist->push_back(err);
        if (!std::isfinite(err)) return BDDA_ENAN;
        if (err < tol) {
            c.info.converged = 1;
            break;
        }
        if (rd.flags & BDDA_SM_LU2) return BDDA_EBREAKDOWN;
        {   // S = Qnew + S beta
            Combo<R> cb;
            cb.nv = 1; cb.V[0] = pQn; cb.ldv[0] = rows; cb.a[0] = one;
            cb.nm = 1; cb.W[0] = pS; cb.ldw[0] = rows; cb.kw[0] = s; cb.Md[0] = beta;
            combo<R>(w, cb, pT, rows, rows, s);
        }
        std::swap(pS, pT);
        std::swap(pQ, pQn);
        std::swap(G2, G3);
        std::swap(xi, xin);
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
    const c64 one = mk(1.0, 0.0);
    DevBuf<cplx<R>> Rb(N, w.stream), Pb(N, w.stream), Pn(N, w.stream), R0(N, w.stream), V(N, w.stream), T(N, w.stream), Z(N, w.stream);
    if (c.jacobi) scale_rows<R>(w.stream, c.P.jpre<R>(), B, rows, Rb.p, rows, rows, s);
    else copy_block<R>(w, B, Rb.p, rows, s);
    const double bnorm = norm_fro<R>(w, Rb.p, rows, rows, s);
    fill_zero<R>(w.stream, X, rows, rows, s);
    copy_block<R>(w, Rb.p, Pb.p, rows, s);
    copy_block<R>(w, Rb.p, R0.p, rows, s);
    SmallState sm(w, s, 6, rows);
    c64 *M1 = sm.m(0), *M2 = sm.m(1), *M3 = sm.m(2), *alpha = sm.m(3), *beta = sm.m(4), *sc = sm.m(5);
    c64 *zt = sc, *zz = sc + 1, *qsi = sc + 2;
    cplx<R>* pP = Pb.p;
    cplx<R>* pPn = Pn.p;
    for (int k = 0; k < itermax; ++k) {
        problem_matvec<R>(c.P, pP, V.p, s, c.jacobi);  // V = A P
        ++c.info.matvecs;
        gram_dev<R>(w, R0.p, rows, s, V.p, rows, s, rows, 1, M1);   // R0^H V
        gram_dev<R>(w, R0.p, rows, s, Rb.p, rows, s, rows, 1, M2);  // R0^H R
        dm_lu_solve(w.stream, M1, M2, alpha, s, s, 0, one, nullptr, sm.flag, BDDA_SM_LU);
        {   // T = R - V alpha
            Combo<R> cb;
            cb.nv = 1; cb.V[0] = Rb.p; cb.ldv[0] = rows; cb.a[0] = one;
            cb.nm = 1; cb.W[0] = V.p; cb.ldw[0] = rows; cb.kw[0] = s; cb.Md[0] = alpha; cb.ms[0] = mk(-1.0, 0.0);
            combo<R>(w, cb, T.p, rows, rows, s);
        }
        problem_matvec<R>(c.P, T.p, Z.p, s, c.jacobi);  // Z = A T
        ++c.info.matvecs;
        dot_all_dev<R>(w, Z.p, rows, T.p, rows, rows, s, 1, zt);
        dot_all_dev<R>(w, Z.p, rows, Z.p, rows, rows, s, 1, zz);
        dm_scalar_div(w.stream, zt, zz, one, qsi, sm.flag);
        {   // X = X + P alpha + qsi T
            Combo<R> cb;
            cb.nv = 2; cb.V[0] = X; cb.ldv[0] = rows; cb.a[0] = one;
            cb.V[1] = T.p; cb.ldv[1] = rows; cb.a[1] = one; cb.ad[1] = qsi;
            cb.nm = 1; cb.W[0] = pP; cb.ldw[0] = rows; cb.kw[0] = s; cb.Md[0] = alpha;
            combo<R>(w, cb, X, rows, rows, s);
        }
        {   // R = T - qsi Z
            Combo<R> cb;
            cb.nv = 2; cb.V[0] = T.p; cb.ldv[0] = rows; cb.a[0] = one;
            cb.V[1] = Z.p; cb.ldv[1] = rows; cb.a[1] = mk(-1.0, 0.0); cb.ad[1] = qsi;
            combo<R>(w, cb, Rb.p, rows, rows, s);
        }
        dot_all_dev<R>(w, Rb.p, rows, Rb.p, rows, rows, s, 1, sm.probe);
        gram_dev<R>(w, R0.p, rows, s, Z.p, rows, s, rows, 1, M3);  // beta = LU(R0^H V)^{-1} (-R0^H Z)
        dm_lu_solve(w.stream, M1, M3, beta, s, s, 0, mk(-1.0, 0.0), nullptr, sm.flag, BDDA_SM_LU2);
        const SmallState::Read rd = sm.read();
        if (rd.flags & BDDA_SM_LU) return BDDA_EBREAKDOWN;
        const double err = std::sqrt(std::max(rd.v, 0.0)) / bnorm;
        c.info.iters = k + 1;
        c.info.recurrence_res = err;
        if (c.hist) c.hist->push_back(err);
        if (!std::isfinite(err)) return BDDA_ENAN;
        if (err < tol) {
            c.info.converged = 1;
            break;
        }
        if (rd.flags & BDDA_SM_LU2) return BDDA_EBREAKDOWN;
        {   // P = R + (P - qsi V) beta
            Combo<R> cb;
            cb.nv = 1; cb.V[0] = Rb.p; cb.ldv[0] = rows; cb.a[0] = one;
            cb.nm = 2;
            cb.W[0] = pP; cb.ldw[0] = rows; cb.kw[0] = s; cb.Md[0] = beta;
            cb.W[1] = V.p; cb.ldw[1] = rows; cb.kw[1] = s; cb.Md[1] = beta; cb.ms[1] = mk(-1.0, 0.0); cb.msd[1] = qsi;
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
    const c64 one = mk(1.0, 0.0);
    DevBuf<cplx<R>> Rb(N, w.stream), Rn(N, w.stream), Pb(N, w.stream), R0(N, w.stream), V(N, w.stream), W(N, w.stream), U(N, w.stream), Y(N, w.stream);
    if (c.jacobi) scale_rows<R>(w.stream, c.P.jpre<R>(), B, rows, Rb.p, rows, rows, s);
    else copy_block<R>(w, B, Rb.p, rows, s);
    const double bnorm = norm_fro<R>(w, Rb.p, rows, rows, s);
    fill_zero<R>(w.stream, X, rows, rows, s);
    copy_block<R>(w, Rb.p, Pb.p, rows, s);
    problem_matvec<R>(c.P, Rb.p, V.p, s, c.jacobi);  // V = A R
    ++c.info.matvecs;
    copy_block<R>(w, V.p, W.p, rows, s);
    copy_block<R>(w, Rb.p, R0.p, rows, s);
    SmallState sm(w, s, 6, rows);
    c64 *M1 = sm.m(0), *M2 = sm.m(1), *M3 = sm.m(2), *alpha = sm.m(3), *gamma = sm.m(4), *sc = sm.m(5);
    c64 *wr = sc, *ww = sc + 1, *qsi = sc + 2;
    cplx<R>* pR = Rb.p;
    cplx<R>* pRn = Rn.p;
    for (int k = 0; k < itermax; ++k) {
        gram_dev<R>(w, R0.p, rows, s, V.p, rows, s, rows, 1, M1);
        gram_dev<R>(w, R0.p, rows, s, pR, rows, s, rows, 1, M2);
        dm_lu_solve(w.stream, M1, M2, alpha, s, s, 0, one, nullptr, sm.flag, BDDA_SM_LU);
        dot_all_dev<R>(w, W.p, rows, pR, rows, rows, s, 1, wr);
        dot_all_dev<R>(w, W.p, rows, W.p, rows, rows, s, 1, ww);
        dm_scalar_div(w.stream, wr, ww, one, qsi, sm.flag);
        {   // U = (P - qsi V) alpha
            Combo<R> cb;
            cb.nm = 2;
            cb.W[0] = Pb.p; cb.ldw[0] = rows; cb.kw[0] = s; cb.Md[0] = alpha;
            cb.W[1] = V.p; cb.ldw[1] = rows; cb.kw[1] = s; cb.Md[1] = alpha; cb.ms[1] = mk(-1.0, 0.0); cb.msd[1] = qsi;
            combo<R>(w, cb, U.p, rows, rows, s);
        }
        problem_matvec<R>(c.P, U.p, Y.p, s, c.jacobi);  // Y = A U
        ++c.info.matvecs;
        {   // X = X + qsi R + U
            Combo<R> cb;
            cb.nv = 3; cb.V[0] = X; cb.ldv[0] = rows; cb.a[0] = one;
            cb.V[1] = pR; cb.ldv[1] = rows; cb.a[1] = one; cb.ad[1] = qsi;
            cb.V[2] = U.p; cb.ldv[2] = rows; cb.a[2] = one;
            combo<R>(w, cb, X, rows, rows, s);
        }
        {   // Rnew = R - qsi W - Y
            Combo<R> cb;
            cb.nv = 3; cb.V[0] = pR; cb.ldv[0] = rows; cb.a[0] = one;
            cb.V[1] = W.p; cb.ldv[1] = rows; cb.a[1] = mk(-1.0, 0.0); cb.ad[1] = qsi;
            cb.V[2] = Y.p; cb.ldv[2] = rows; cb.a[2] = mk(-1.0, 0.0);
            combo<R>(w, cb, pRn, rows, rows, s);
        }
        dot_all_dev<R>(w, pRn, rows, pRn, rows, rows, s, 1, sm.probe);
        const SmallState::Read rd = sm.read();
        // (LU2: the previous iteration's gamma solve broke down)
        if (rd.flags & BDDA_SM_LU_ANY) return BDDA_EBREAKDOWN;
        const double err = std::sqrt(std::max(rd.v, 0.0)) / bnorm;
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
        gram_dev<R>(w, R0.p, rows, s, pR, rows, s, rows, 1, M3);   // R0^H R
        gram_dev<R>(w, R0.p, rows, s, pRn, rows, s, rows, 1, M2);  // R0^H Rnew (divided by qsi in the solve)
        dm_lu_solve(w.stream, M3, M2, gamma, s, s, 0, one, qsi, sm.flag, BDDA_SM_LU2);
        std::swap(pR, pRn);
        {   // P = R + U gamma
            Combo<R> cb;
            cb.nv = 1; cb.V[0] = pR; cb.ldv[0] = rows; cb.a[0] = one;
            cb.nm = 1; cb.W[0] = U.p; cb.ldw[0] = rows; cb.kw[0] = s; cb.Md[0] = gamma;
            combo<R>(w, cb, Pb.p, rows, rows, s);
        }
        {   // V = W + Y gamma
            Combo<R> cb;
            cb.nv = 1; cb.V[0] = W.p; cb.ldv[0] = rows; cb.a[0] = one;
            cb.nm = 1; cb.W[0] = Y.p; cb.ldw[0] = rows; cb.kw[0] = s; cb.Md[0] = gamma;
            combo<R>(w, cb, V.p, rows, rows, s);
        }
    }
    if (!c.info.converged && c.info.iters > 0) {
        // a gamma breakdown in the last iteration that ran
        if (sm.read().flags & BDDA_SM_LU_ANY) return BDDA_EBREAKDOWN;
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
    DevBuf<c64> AX(N, w.stream);
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
    DevBuf<c64> AX(N, w.stream);
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
        DevBuf<c32> B32(N, w.stream), X32(N, w.stream);
        convert_64_to_32(w.stream, B64, rows, B32.p, rows, rows, s, 1.0);
        rc = run_method<float>(c, method, B32.p, X32.p, tol, itermax);
        fill_zero<double>(w.stream, X64, rows, rows, s);
        accumulate_32_to_64(w.stream, X32.p, rows, X64, rows, rows, s, 1.0);
        info.outer_iters = 1;
    } else if (mode == MODE_MIXED || mode == MODE_MIXED_COLUMNS) {
        // FP64 defect correction: r = B - A x (FP64 matvec), A d = r in FP32,
        // x += d, until ||B - A x|| / ||B|| <= tol.
        DevBuf<c64> R64(N, w.stream), AX(N, w.stream);
        DevBuf<c32> R32(N, w.stream), D32(N, w.stream);
        const double bn = norm_fro<double>(w, B64, rows, rows, s);
        fill_zero<double>(w.stream, X64, rows, rows, s);
        BDDA_HIP(hipMemcpyAsync(R64.p, B64, sizeof(c64) * N, hipMemcpyDeviceToDevice, w.stream));
        double rel = 1.0;
        const int max_outer = 20;
        int total_iters = 0;
        // every pass measures the residual of the current x first, so the
        // reported true residual and convergence always describe the X
        // returned, also after the last correction
        // per-column test: ||b_j|| once, ||r_j|| from the diagonal of R^H R
        std::vector<double> bcol;
        if (mode == MODE_MIXED_COLUMNS) {
            const HostMat gb = gram<double>(w, B64, rows, s, B64, rows, s, rows, 1);
            for (int j = 0; j < s; ++j) bcol.push_back(std::sqrt(std::max(gb[(size_t)j * s + j].x, 0.0)));
        }
        for (int o = 0;; ++o) {
            const double rn = norm_fro<double>(w, R64.p, rows, rows, s);
            rel = rn / bn;
            if (mode == MODE_MIXED_COLUMNS) {
                const HostMat gr = gram<double>(w, R64.p, rows, s, R64.p, rows, s, rows, 1);
                double worst = 0;
                for (int j = 0; j < s; ++j) {
                    const double r = std::sqrt(std::max(gr[(size_t)j * s + j].x, 0.0));
                    worst = std::max(worst, bcol[j] > 0 ? r / bcol[j] : r);
                }
                rel = worst;
            }
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
