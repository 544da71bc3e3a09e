// capi.hip -- extern "C" entry points declared in include/bdda.h.
#include "../../include/bdda.h"
#include "observables.hpp"
#include "solvers.hpp"
#include <cmath>
#include <complex>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <vector>

namespace bdda {
static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
const char* last_error() { return g_err.c_str(); }
}  // namespace bdda

using namespace bdda;

struct bdda_mbt {
    MbtOperator* op = nullptr;
    Problem prob;
    hipStream_t stream = nullptr;
    int n[3] = {0, 0, 0};
    int f = 3;
    int have_target = 0;
    ~bdda_mbt()
    {
        // device buffers go back to the allocator after the stream's work,
        // so the stream is destroyed last
        prob.release();
        if (op) mbt_destroy(op);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

#define BDDA_TRY(...)                                                               \
    try {                                                                           \
        __VA_ARGS__                                                                 \
    } catch (const HipError& e) {                                                   \
        set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " at " + e.where); \
        return (e.code == hipErrorOutOfMemory) ? BDDA_ENOMEM : BDDA_EHIP;           \
    } catch (const std::bad_alloc&) {                                               \
        set_error("host allocation failed");                                        \
        return BDDA_ENOMEM;                                                         \
    } catch (const std::exception& e) {                                             \
        set_error(e.what());                                                        \
        return BDDA_EINVAL;                                                         \
    }

static int fail(int code, const char* msg)
{
    set_error(msg);
    return code;
}

// ------------------------------------------------------------------ layout
// Closed forms of the reference recursions:
//   BT_pad        x[f((ix ny + iy) nz + iz) + q] -> xz[((ix My + iy) Mz + iz) f^2 + q]
//   BT_reconstruct b[f((ix ny + iy) nz + iz) + r] <- bz[(((ix+nx-1) My + iy+ny-1) Mz + iz+nz-1) f^2 + r f + f-1]
//   BT_fft        Au[((jx My + jy) Mz + jz) f^2 + r' + c' f] = M(d)[f-1-c'][r'], d = n-1-j
static void app_fn(const int* d, int f, double lf, std::complex<double> k, std::complex<double>* out)
{
    const double X[3] = {d[0] * lf, d[1] * lf, d[2] * lf};
    const double r = std::sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    for (int i = 0; i < f * f; ++i) out[i] = 0.0;
    if (r == 0.0) return;
    const std::complex<double> I(0.0, 1.0);
    const std::complex<double> e = std::exp(I * k * r) / r;
    const std::complex<double> c2 = (1.0 - I * k * r) / (r * r);
    std::complex<double> G[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const double uu = X[a] * X[b] / (r * r);
            const double id = (a == b) ? 1.0 : 0.0;
            G[a][b] = (k * k * (id - uu) - c2 * (id - 3.0 * uu)) * e;
        }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) out[a + f * b] = -G[a][b];
    if (f == 6) {
        const double u[3] = {X[0] / r, X[1] / r, X[2] / r};
        const double ux[3][3] = {{0, -u[2], u[1]}, {u[2], 0, -u[0]}, {-u[1], u[0], 0}};
        const std::complex<double> cf = -k * k * e * (1.0 - 1.0 / (I * k * r));
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                out[a + 6 * (b + 3)] = -cf * ux[a][b];
                out[(a + 3) + 6 * b] = cf * ux[a][b];
                out[(a + 3) + 6 * (b + 3)] = -G[a][b];
            }
    }
}

static int check_prec(const bdda_mbt* h, int prec)
{
    if (prec == BDDA_PREC_FP64 && h->op->have64) return BDDA_OK;
    if (prec == BDDA_PREC_FP32 && h->op->have32) return BDDA_OK;
    return fail(BDDA_EINVAL, "operator was not built for this precision");
}

template <typename R>
static void to_dev(const bdda_c64* src, int64_t ld, int64_t rows, int cols, cplx<R>* dst, hipStream_t st)
{
    if (sizeof(R) == sizeof(double) && ld == rows) {  // same layout: one copy, no staging vector
        BDDA_HIP(hipMemcpyAsync(dst, src, sizeof(cplx<R>) * rows * cols, hipMemcpyHostToDevice, st));
        BDDA_HIP(hipStreamSynchronize(st));
        return;
    }
    std::vector<cplx<R>> tmp((size_t)rows * cols);
    for (int j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i) tmp[(size_t)j * rows + i] = mk((R)src[j * ld + i].re, (R)src[j * ld + i].im);
    BDDA_HIP(hipMemcpyAsync(dst, tmp.data(), sizeof(cplx<R>) * tmp.size(), hipMemcpyHostToDevice, st));
    BDDA_HIP(hipStreamSynchronize(st));
}

template <typename R>
static void from_dev(const cplx<R>* src, int64_t rows, int cols, bdda_c64* dst, int64_t ld, hipStream_t st)
{
    if (sizeof(R) == sizeof(double) && ld == rows) {
        BDDA_HIP(hipMemcpyAsync(dst, src, sizeof(cplx<R>) * rows * cols, hipMemcpyDeviceToHost, st));
        BDDA_HIP(hipStreamSynchronize(st));
        return;
    }
    std::vector<cplx<R>> tmp((size_t)rows * cols);
    BDDA_HIP(hipMemcpyAsync(tmp.data(), src, sizeof(cplx<R>) * tmp.size(), hipMemcpyDeviceToHost, st));
    BDDA_HIP(hipStreamSynchronize(st));
    for (int j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i) dst[j * ld + i] = {(double)tmp[(size_t)j * rows + i].x, (double)tmp[(size_t)j * rows + i].y};
}

// the pass workspace lives for one call of the C ABI: back to the device
// allocator when the call returns, so that idle operators hold no workspace
struct WorkspaceScope {
    MbtOperator* op;
    ~WorkspaceScope()
    {
        if (op) mbt_release_workspace(op);
    }
};

template <typename R>
static void apply_host(bdda_mbt* h, const bdda_c64* x, int64_t ldx, bdda_c64* y, int64_t ldy, int ncols, int occupied,
                       int jacobi)
{
    const int64_t rows = occupied ? h->prob.rows() : (int64_t)h->f * h->n[0] * h->n[1] * h->n[2];
    DevPtr<cplx<R>> bx((size_t)rows * ncols, h->stream), by((size_t)rows * ncols, h->stream);
    cplx<R>* dx = bx.p;
    cplx<R>* dy = by.p;
    to_dev<R>(x, ldx, rows, ncols, dx, h->stream);
    if (occupied)
        mbt_apply_occupied<R>(h->op, dx, rows, dy, rows, ncols, h->prob.diag<R>(), jacobi ? h->prob.jpre<R>() : nullptr,
                              h->prob.nocc);
    else
        mbt_apply_occupied<R>(h->op, dx, rows, dy, rows, ncols, nullptr, nullptr, -1);
    from_dev<R>(dy, rows, ncols, y, ldy, h->stream);
}

static int solve_common(bdda_mbt* h, int method, int jacobi, int mode, const c64* dB, int L, double tol, int itermax,
                        c64* dX, bdda_solve_info* info, double* hist, int hist_len)
{
    if (method < 0 || method > 2) return fail(BDDA_EINVAL, "unknown method");
    if (mode < 0 || mode > 3) return fail(BDDA_EINVAL, "unknown mode");
    if (L < 1 || L > 64) return fail(BDDA_EINVAL, "number of right-hand sides must be in 1..64");
    if (!(tol > 0) || itermax < 1) return fail(BDDA_EINVAL, "tol must be > 0 and itermax >= 1");
    if (!h->have_target) return fail(BDDA_EINVAL, "bdda_mbt_set_target not called");
    if ((mode != MODE_FP64 && !h->op->have32) || (mode != MODE_FP32 && !h->op->have64))
        return fail(BDDA_EINVAL, "operator lacks the precision this mode needs");
    if (!all_finite(h->prob.work, dB, h->prob.rows() * L)) return fail(BDDA_ENAN, "NaN or Inf in B");
    SolveInfo si;
    std::vector<double> hv;
    int rc = solve_fp(h->prob, method, jacobi, mode, dB, dX, L, tol, itermax, si, hist ? &hv : nullptr, 0.0);
    if (rc == BDDA_OK && mode != MODE_MIXED && mode != MODE_MIXED_COLUMNS)
        si.true_res = true_residual(h->prob, dB, dX, L);
    if (rc == BDDA_OK && !si.converged) rc = BDDA_ENOCONV;
    si.status = rc;
    if (info) {
        info->iters = si.iters;
        info->outer_iters = si.outer_iters;
        info->matvecs = si.matvecs;
        info->converged = si.converged;
        info->status = rc;
        info->recurrence_res = si.recurrence_res;
        info->true_res = si.true_res;
        info->time_ms = si.time_ms;
    }
    if (hist)
        for (int i = 0; i < hist_len; ++i) hist[i] = (i < (int)hv.size()) ? hv[i] : std::nan("");
    if (rc == BDDA_EBREAKDOWN) set_error("block breakdown: rank-deficient block (QR or s x s LU)");
    else if (rc == BDDA_ENAN) set_error("NaN or Inf in the iteration");
    else if (rc == BDDA_ENOCONV) set_error("itermax reached before tol");
    return rc;
}

// --------------------------------------------------------------- observables
static int obs_check(const bdda_mbt* h, int L, double* kout)
{
    if (!h->have_target) return fail(BDDA_EINVAL, "bdda_mbt_set_target not called");
    if (!h->op->have64) return fail(BDDA_EINVAL, "observables need the FP64 operator");
    if (h->op->kim != 0.0) return fail(BDDA_EINVAL, "non zero imag(k) is not supported");
    if (L < 1) return fail(BDDA_EINVAL, "need at least one column");
    *kout = h->op->kre;
    return BDDA_OK;
}

static int xs_dev(bdda_mbt* h, const c64* dB, const c64* dX, int L, int alpha_mode, double* Cext, double* Cabs,
                  double* Csca)
{
    double k = 0;
    int rc = obs_check(h, L, &k);
    if (rc) return rc;
    if (alpha_mode != BDDA_ALPHA_FIRST && alpha_mode != BDDA_ALPHA_EACH) return fail(BDDA_EINVAL, "bad alpha_mode");
    if (!Cext || !Cabs || !Csca) return fail(BDDA_EINVAL, "null pointer");
    cross_sections(h->prob, dB, dX, L, k, alpha_mode == BDDA_ALPHA_EACH, Cext, Cabs);
    for (int l = 0; l < L; ++l) Csca[l] = Cext[l] - Cabs[l];
    return BDDA_OK;
}

static int sm_dev(bdda_mbt* h, const c64* dX, int npairs, const double* u, const double* e1, int nang, double* S11,
                  double* S12, bdda_c64* S)
{
    double k = 0;
    int rc = obs_check(h, 2 * npairs, &k);
    if (rc) return rc;
    if (!u || !e1 || !S11 || !S12) return fail(BDDA_EINVAL, "null pointer");
    if (npairs < 1 || nang < 2) return fail(BDDA_EINVAL, "need npairs >= 1 and nang >= 2");
    for (int p = 0; p < npairs; ++p) {
        const double* a = u + 3 * p;
        const double* b = e1 + 3 * p;
        const double na = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        const double nb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
        const double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
        if (std::fabs(na - 1) > 1e-9 || std::fabs(nb - 1) > 1e-9 || std::fabs(ab) > 1e-9)
            return fail(BDDA_EINVAL, "u and e1 must be orthonormal");
    }
    std::vector<c64> amp((size_t)npairs * nang * 4);
    scattering_amplitudes(h->prob, dX, npairs, k, u, e1, nang, amp.data());
    for (int64_t e = 0; e < (int64_t)npairs * nang; ++e) {
        double m2[4];
        for (int q = 0; q < 4; ++q) m2[q] = amp[e * 4 + q].x * amp[e * 4 + q].x + amp[e * 4 + q].y * amp[e * 4 + q].y;
        S11[e] = 0.5 * (m2[0] + m2[1] + m2[2] + m2[3]);
        S12[e] = 0.5 * (m2[1] - m2[0] + m2[3] - m2[2]);
        if (S)
            for (int q = 0; q < 4; ++q) S[e * 4 + q] = {amp[e * 4 + q].x, amp[e * 4 + q].y};
    }
    return BDDA_OK;
}

// a device array of at least one element, freed after the work on stream st
template <typename T>
struct DevMem : DevPtr<T> {
    DevMem(size_t n, hipStream_t st) : DevPtr<T>(std::max<size_t>(n, 1), st) {}
};

extern "C" {

const char* bdda_last_error(void) { return last_error(); }
int bdda_version(void) { return 1; }

int64_t bdda_fft_length(const int n[3], int f)
{
    return (int64_t)f * f * (2 * n[0] - 1) * (int64_t)(2 * n[1] - 1) * (2 * n[2] - 1);
}

int bdda_application_function(const int n_ind[6], int f, double lf, bdda_c64 k, bdda_c64* out)
{
    if (f != 3 && f != 6) return fail(BDDA_EINVAL, "unknown f value (3 or 6)");
    if (!n_ind || !out) return fail(BDDA_EINVAL, "null pointer");
    const int d[3] = {n_ind[0] - n_ind[3], n_ind[1] - n_ind[4], n_ind[2] - n_ind[5]};
    app_fn(d, f, lf, {k.re, k.im}, reinterpret_cast<std::complex<double>*>(out));
    return BDDA_OK;
}

int bdda_bt_fft(const int n[3], int f, double lf, bdda_c64 k, bdda_c64* Au_)
{
    if (f != 3 && f != 6) return fail(BDDA_EINVAL, "unknown f value (3 or 6)");
    auto* Au = reinterpret_cast<std::complex<double>*>(Au_);
    const int My = 2 * n[1] - 1, Mz = 2 * n[2] - 1, Mx = 2 * n[0] - 1;
    std::complex<double> blk[36];
    for (int jx = 0; jx < Mx; ++jx)
        for (int jy = 0; jy < My; ++jy)
            for (int jz = 0; jz < Mz; ++jz) {
                const int d[3] = {n[0] - 1 - jx, n[1] - 1 - jy, n[2] - 1 - jz};
                app_fn(d, f, lf, {k.re, k.im}, blk);
                std::complex<double>* o = Au + (((int64_t)jx * My + jy) * Mz + jz) * f * f;
                for (int c = 0; c < f; ++c)
                    for (int r = 0; r < f; ++r) o[r + c * f] = blk[(f - 1 - c) + f * r];
            }
    return BDDA_OK;
}

int bdda_bt_pad(const int n[3], int f, const bdda_c64* x, bdda_c64* xz)
{
    if (f < 1) return fail(BDDA_EINVAL, "f < 1");
    const int My = 2 * n[1] - 1, Mz = 2 * n[2] - 1;
    const int64_t len = (int64_t)f * f * n[0] * My * Mz;
    std::memset(xz, 0, sizeof(bdda_c64) * len);
    for (int ix = 0; ix < n[0]; ++ix)
        for (int iy = 0; iy < n[1]; ++iy)
            for (int iz = 0; iz < n[2]; ++iz)
                for (int q = 0; q < f; ++q)
                    xz[(((int64_t)ix * My + iy) * Mz + iz) * f * f + q] =
                        x[(int64_t)f * (((int64_t)ix * n[1] + iy) * n[2] + iz) + q];
    return BDDA_OK;
}

int bdda_bt_reconstruct(const int n[3], int f, const bdda_c64* bz, bdda_c64* b)
{
    if (f < 1) return fail(BDDA_EINVAL, "f < 1");
    const int My = 2 * n[1] - 1, Mz = 2 * n[2] - 1;
    for (int ix = 0; ix < n[0]; ++ix)
        for (int iy = 0; iy < n[1]; ++iy)
            for (int iz = 0; iz < n[2]; ++iz)
                for (int r = 0; r < f; ++r)
                    b[(int64_t)f * (((int64_t)ix * n[1] + iy) * n[2] + iz) + r] =
                        bz[((((int64_t)ix + n[0] - 1) * My + iy + n[1] - 1) * Mz + iz + n[2] - 1) * f * f + r * f + f -
                           1];
    return BDDA_OK;
}

// ------------------------------------------------------------------ operator
int bdda_mbt_create(const int n[3], int f, double lf, bdda_c64 k, int prec_mask, bdda_mbt** out)
{
    if (!out || !n) return fail(BDDA_EINVAL, "null pointer");
    *out = nullptr;
    if (f != 3 && f != 6) return fail(BDDA_EINVAL, "unknown f value (3 or 6)");
    if (n[0] < 1 || n[1] < 1 || n[2] < 1) return fail(BDDA_EINVAL, "grid dimensions must be >= 1");
    if (!(lf > 0) || !std::isfinite(lf)) return fail(BDDA_EINVAL, "lf must be positive");
    if (!(prec_mask & 3)) return fail(BDDA_EINVAL, "prec_mask selects no precision");
    int P[3];
    if (!mbt_padded_dims(n, f, P))
        return fail(BDDA_EUNSUPPORTED, "padded grid beyond the passes' limit (P * f <= 3072 on every axis)");
    BDDA_TRY({
        std::unique_ptr<bdda_mbt> h(new bdda_mbt());
        BDDA_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        h->op = mbt_create(n, f, lf, k.re, k.im, (prec_mask & BDDA_PREC_FP64) ? 1 : 0,
                           (prec_mask & BDDA_PREC_FP32) ? 1 : 0, h->stream);
        h->prob.op = h->op;
        h->prob.f = f;
        h->prob.work.stream = h->stream;
        for (int a = 0; a < 3; ++a) h->n[a] = n[a];
        h->f = f;
        *out = h.release();
        return BDDA_OK;
    })
}

void bdda_mbt_destroy(bdda_mbt* h) { delete h; }

int bdda_mbt_grid(const bdda_mbt* h, int P[3])
{
    if (!h || !P) return fail(BDDA_EINVAL, "null pointer");
    P[0] = h->op->d.Px;
    P[1] = h->op->d.Py;
    P[2] = h->op->d.Pz;
    return BDDA_OK;
}

int bdda_mbt_apply(bdda_mbt* h, int prec, const bdda_c64* x, int64_t ldx, bdda_c64* y, int64_t ldy, int ncols)
{
    if (!h || !x || !y) return fail(BDDA_EINVAL, "null pointer");
    if (ncols < 1) return BDDA_OK;
    if (int rc = check_prec(h, prec)) return rc;
    const int64_t rows = (int64_t)h->f * h->n[0] * h->n[1] * h->n[2];
    if (ldx < rows || ldy < rows) return fail(BDDA_EINVAL, "leading dimension smaller than f*nx*ny*nz");
    BDDA_TRY({
        WorkspaceScope wss{h->op};
        if (prec == BDDA_PREC_FP64) apply_host<double>(h, x, ldx, y, ldy, ncols, 0, 0);
        else apply_host<float>(h, x, ldx, y, ldy, ncols, 0, 0);
        return BDDA_OK;
    })
}

int bdda_mbt_set_target(bdda_mbt* h, const int32_t* address, int64_t nocc, const bdda_c64* diag)
{
    if (!h || (!address && nocc > 0) || (!diag && nocc > 0)) return fail(BDDA_EINVAL, "null pointer");
    if (nocc < 1) return fail(BDDA_EINVAL, "no occupied cells");
    const int64_t C = (int64_t)h->n[0] * h->n[1] * h->n[2];
    std::vector<char> seen(C, 0);
    for (int64_t m = 0; m < nocc; ++m) {
        if (address[m] < 0 || address[m] >= C) return fail(BDDA_EINVAL, "address outside the cuboid");
        if (seen[address[m]]) return fail(BDDA_EINVAL, "duplicate address");
        seen[address[m]] = 1;
    }
    for (int64_t i = 0; i < (int64_t)h->f * nocc; ++i)
        if (!std::isfinite(diag[i].re) || !std::isfinite(diag[i].im) || (diag[i].re == 0 && diag[i].im == 0))
            return fail(BDDA_EINVAL, "DIAG_A entries must be finite and nonzero");
    BDDA_TRY({
        mbt_set_address(h->op, address, nocc);
        h->prob.nocc = nocc;
        problem_set_diag(h->prob, reinterpret_cast<const c64*>(diag));
        h->have_target = 1;
        return BDDA_OK;
    })
}

int bdda_mbt_apply_occupied(bdda_mbt* h, int prec, int jacobi, const bdda_c64* X, int64_t ldx, bdda_c64* Y,
                            int64_t ldy, int ncols)
{
    if (!h || !X || !Y) return fail(BDDA_EINVAL, "null pointer");
    if (!h->have_target) return fail(BDDA_EINVAL, "bdda_mbt_set_target not called");
    if (ncols < 1) return BDDA_OK;
    if (int rc = check_prec(h, prec)) return rc;
    if (ldx < h->prob.rows() || ldy < h->prob.rows()) return fail(BDDA_EINVAL, "leading dimension smaller than f*nocc");
    BDDA_TRY({
        WorkspaceScope wss{h->op};
        if (prec == BDDA_PREC_FP64) apply_host<double>(h, X, ldx, Y, ldy, ncols, 1, jacobi);
        else apply_host<float>(h, X, ldx, Y, ldy, ncols, 1, jacobi);
        return BDDA_OK;
    })
}

int bdda_solve(bdda_mbt* h, int method, int jacobi, int mode, const bdda_c64* B, int64_t ldb, int L, double tol,
               int itermax, bdda_c64* X, int64_t ldx, bdda_solve_info* info, double* hist, int hist_len)
{
    if (!h || !B || !X) return fail(BDDA_EINVAL, "null pointer");
    if (!h->have_target) return fail(BDDA_EINVAL, "bdda_mbt_set_target not called");
    const int64_t rows = h->prob.rows();
    if (ldb < rows || ldx < rows) return fail(BDDA_EINVAL, "leading dimension smaller than f*nocc");
    if (L < 1 || L > 64) return fail(BDDA_EINVAL, "number of right-hand sides must be in 1..64");
    BDDA_TRY({
        WorkspaceScope wss{h->op};
        DevMem<c64> bB((size_t)rows * L, h->stream), bX((size_t)rows * L, h->stream);
        c64* dB = bB.p;
        c64* dX = bX.p;
        to_dev<double>(B, ldb, rows, L, dB, h->stream);
        int rc = solve_common(h, method, jacobi, mode, dB, L, tol, itermax, dX, info, hist, hist_len);
        if (rc == BDDA_OK || rc == BDDA_ENOCONV) from_dev<double>(dX, rows, L, X, ldx, h->stream);
        return rc;
    })
}

int bdda_solve_dev(bdda_mbt* h, int method, int jacobi, int mode, const bdda_c64* dB, int L, double tol, int itermax,
                   bdda_c64* dX, bdda_solve_info* info, double* hist, int hist_len)
{
    if (!h || !dB || !dX) return fail(BDDA_EINVAL, "null pointer");
    BDDA_TRY({
        WorkspaceScope wss{h->op};
        return solve_common(h, method, jacobi, mode, reinterpret_cast<const c64*>(dB), L, tol, itermax,
                            reinterpret_cast<c64*>(dX), info, hist, hist_len);
    })
}

static int rescol_dev(bdda_mbt* h, const c64* dB, const c64* dX, int L, double* res)
{
    if (!h->have_target) return fail(BDDA_EINVAL, "bdda_mbt_set_target not called");
    if (L < 1 || L > 64 || !res) return fail(BDDA_EINVAL, "need 1 <= L <= 64 and an output array");
    residual_columns(h->prob, dB, dX, L, res);
    return BDDA_OK;
}

int bdda_residual_columns(bdda_mbt* h, const bdda_c64* B, int64_t ldb, const bdda_c64* X, int64_t ldx, int L,
                          double* res)
{
    if (!h || !B || !X) return fail(BDDA_EINVAL, "null pointer");
    BDDA_TRY({
        WorkspaceScope wss{h->op};
        if (!h->have_target) return fail(BDDA_EINVAL, "bdda_mbt_set_target not called");
        const int64_t rows = h->prob.rows();
        if (L < 1 || ldb < rows || ldx < rows) return fail(BDDA_EINVAL, "bad block shape");
        DevMem<c64> dB((size_t)rows * L, h->stream);
        DevMem<c64> dX((size_t)rows * L, h->stream);
        to_dev<double>(B, ldb, rows, L, dB.p, h->stream);
        to_dev<double>(X, ldx, rows, L, dX.p, h->stream);
        return rescol_dev(h, dB.p, dX.p, L, res);
    })
}

int bdda_residual_columns_dev(bdda_mbt* h, const bdda_c64* dB, const bdda_c64* dX, int L, double* res)
{
    if (!h || !dB || !dX) return fail(BDDA_EINVAL, "null pointer");
    BDDA_TRY({
        WorkspaceScope wss{h->op};
        return rescol_dev(h, reinterpret_cast<const c64*>(dB), reinterpret_cast<const c64*>(dX), L, res);
    })
}

int bdda_cross_sections(bdda_mbt* h, const bdda_c64* B, int64_t ldb, const bdda_c64* X, int64_t ldx, int L,
                        int alpha_mode, double* Cext, double* Cabs, double* Csca)
{
    if (!h || !B || !X) return fail(BDDA_EINVAL, "null pointer");
    BDDA_TRY({
        if (!h->have_target) return fail(BDDA_EINVAL, "bdda_mbt_set_target not called");
        const int64_t rows = h->prob.rows();
        if (L < 1 || ldb < rows || ldx < rows) return fail(BDDA_EINVAL, "bad block shape");
        DevMem<c64> dB((size_t)rows * L, h->stream);
        DevMem<c64> dX((size_t)rows * L, h->stream);
        to_dev<double>(B, ldb, rows, L, dB.p, h->stream);
        to_dev<double>(X, ldx, rows, L, dX.p, h->stream);
        return xs_dev(h, dB.p, dX.p, L, alpha_mode, Cext, Cabs, Csca);
    })
}

int bdda_cross_sections_dev(bdda_mbt* h, const bdda_c64* dB, const bdda_c64* dX, int L, int alpha_mode, double* Cext,
                            double* Cabs, double* Csca)
{
    if (!h || !dB || !dX) return fail(BDDA_EINVAL, "null pointer");
    BDDA_TRY({
        return xs_dev(h, reinterpret_cast<const c64*>(dB), reinterpret_cast<const c64*>(dX), L, alpha_mode, Cext,
                      Cabs, Csca);
    })
}

int bdda_cross_sections_dev_noalloc(bdda_mbt* h, const bdda_c64* dB, const bdda_c64* dX, int L, int alpha_mode,
                                    double* Cext, double* Cabs, double* Csca)
{
    if (!h || !dB || !dX) return fail(BDDA_EINVAL, "null pointer");
    BDDA_TRY({
        double k = 0;
        int rc = obs_check(h, L, &k);
        if (rc) return rc;
        if (alpha_mode != BDDA_ALPHA_FIRST && alpha_mode != BDDA_ALPHA_EACH) return fail(BDDA_EINVAL, "bad alpha_mode");
        if (!Cext || !Cabs || !Csca) return fail(BDDA_EINVAL, "null pointer");
        if (L > XS_MAX_COLS) return fail(BDDA_EINVAL, "bdda_cross_sections_dev_noalloc takes at most 256 columns");
        cross_sections_noalloc(h->prob, reinterpret_cast<const c64*>(dB), reinterpret_cast<const c64*>(dX), L, k,
                               alpha_mode == BDDA_ALPHA_EACH, xs_scratch_device(), Cext, Cabs);
        for (int l = 0; l < L; ++l) Csca[l] = Cext[l] - Cabs[l];
        return BDDA_OK;
    })
}

int bdda_scattering_matrix(bdda_mbt* h, const bdda_c64* X, int64_t ldx, int npairs, const double* u,
                           const double* e1, int nang, double* S11, double* S12, bdda_c64* S)
{
    if (!h || !X) return fail(BDDA_EINVAL, "null pointer");
    BDDA_TRY({
        if (!h->have_target) return fail(BDDA_EINVAL, "bdda_mbt_set_target not called");
        const int64_t rows = h->prob.rows();
        if (npairs < 1 || ldx < rows) return fail(BDDA_EINVAL, "bad block shape");
        DevMem<c64> dX((size_t)rows * 2 * npairs, h->stream);
        to_dev<double>(X, ldx, rows, 2 * npairs, dX.p, h->stream);
        return sm_dev(h, dX.p, npairs, u, e1, nang, S11, S12, S);
    })
}

int bdda_scattering_matrix_dev(bdda_mbt* h, const bdda_c64* dX, int npairs, const double* u, const double* e1,
                               int nang, double* S11, double* S12, bdda_c64* S)
{
    if (!h || !dX) return fail(BDDA_EINVAL, "null pointer");
    BDDA_TRY({ return sm_dev(h, reinterpret_cast<const c64*>(dX), npairs, u, e1, nang, S11, S12, S); })
}

int bdda_estimate_bytes(const int n[3], int f, int64_t nocc, int L, int mode, int64_t* init_peak,
                        int64_t* resident, int64_t* workspace, int64_t* krylov)
{
    if (!n || !init_peak || !resident || !workspace || !krylov) return fail(BDDA_EINVAL, "null pointer");
    if (f != 3 && f != 6) return fail(BDDA_EINVAL, "unknown f value (3 or 6)");
    if (n[0] < 1 || n[1] < 1 || n[2] < 1 || nocc < 1 || L < 1) return fail(BDDA_EINVAL, "bad sizes");
    if (mode < BDDA_MODE_FP64 || mode > BDDA_MODE_MIXED_COLUMNS) return fail(BDDA_EINVAL, "bad mode");
    if (mode == BDDA_MODE_MIXED_COLUMNS) mode = BDDA_MODE_MIXED;  // the same buffers
    int P[3];
    if (!mbt_padded_dims(n, f, P))
        return fail(BDDA_EUNSUPPORTED, "padded grid beyond the passes' limit (P * f <= 3072 on every axis)");
    BDDA_TRY({
        MbtDims d{n[0], n[1], n[2], f, P[0], P[1], P[2]};
        const int Pzs = mbt_z_pitch(d.Pz);
        const int64_t nc = (f == 3) ? 6 : 9;
        const int64_t G = (int64_t)d.Px * d.Py * d.Pz;
        const int64_t gh = (int64_t)(d.Px / 2 + 1) * (d.Py / 2 + 1) * (d.Pz / 2 + 1) * nc * (16 + 8);
        const int64_t C = (int64_t)n[0] * n[1] * n[2];
        const int64_t rows = (int64_t)f * nocc;
        *resident = gh + 4 * C + 4 * nocc + rows * (16 + 16 + 8 + 8) + (int64_t)(d.Px + d.Py + d.Pz) * 24;
        *init_peak = *resident + 2 * G * nc * 16;
        const int64_t ws = (int64_t)mbt_workspace_per_column(d, Pzs, mode == BDDA_MODE_FP32 ? 8 : 16) * L;
        const int64_t kry = rows * L * (mode == BDDA_MODE_FP64 ? 10 * 16 : mode == BDDA_MODE_MIXED ? 4 * 16 + 10 * 8
                                                                                                 : 2 * 16 + 10 * 8);
        *workspace = ws;
        *krylov = kry;
        return BDDA_OK;
    })
}

int bdda_mbt_set_workspace_limit(bdda_mbt* h, size_t bytes)
{
    if (!h) return fail(BDDA_EINVAL, "null pointer");
    h->op->ws_limit = bytes;
    return BDDA_OK;
}

int bdda_qr_reduced(const bdda_c64* A, int64_t m, int n, bdda_c64* Q, bdda_c64* R)
{
    if (!A || !Q || !R) return fail(BDDA_EINVAL, "null pointer");
    if (m < 1 || n < 1 || n > 64) return fail(BDDA_EINVAL, "need m >= 1 and 1 <= n <= 64");
    BDDA_TRY({
        BlockWork w;
        BDDA_HIP(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
        struct S {  // w's buffers go back after its stream's work, then the stream goes
            BlockWork& w;
            ~S()
            {
                w.release();
                (void)hipStreamDestroy(w.stream);
                w.stream = nullptr;
            }
        } sg{w};
        DevPtr<c64> bA((size_t)m * n, w.stream);
        c64* dA = bA.p;
        BDDA_HIP(hipMemcpy(dA, A, sizeof(c64) * m * n, hipMemcpyHostToDevice));
        HostMat Rm;
        int rc = qr_mgs2<double>(w, dA, m, dA, m, m, n, Rm);
        if (rc) return fail(rc, "qr_reduced: rank-deficient input (zero column after orthogonalisation)");
        BDDA_HIP(hipMemcpy(Q, dA, sizeof(c64) * m * n, hipMemcpyDeviceToHost));
        std::memcpy(R, Rm.data(), sizeof(c64) * n * n);
        return BDDA_OK;
    })
}

int bdda_lu_solve(const bdda_c64* A, const bdda_c64* B, bdda_c64* X, int s, int nb, int full)
{
    if (!A || (nb > 0 && (!B || !X))) return fail(BDDA_EINVAL, "null pointer");
    if (s < 1 || s > 64 || nb < 0 || nb > 64) return fail(BDDA_EINVAL, "need 1 <= s <= 64 and 0 <= nb <= 64");
    BDDA_TRY({
        hipStream_t st;
        BDDA_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct S {
            hipStream_t s;
            ~S() { (void)hipStreamDestroy(s); }
        } sg{st};
        const size_t na = (size_t)s * s, nx = (size_t)s * nb;
        DevPtr<c64> bd(na + 2 * nx + 1, st);
        c64* d = bd.p;
        c64 *dA = d, *dB = d + na, *dX = dB + nx;
        int* flag = reinterpret_cast<int*>(dX + nx);
        BDDA_HIP(hipMemcpy(dA, A, sizeof(c64) * na, hipMemcpyHostToDevice));
        if (nx) BDDA_HIP(hipMemcpy(dB, B, sizeof(c64) * nx, hipMemcpyHostToDevice));
        const int zero = 0;
        BDDA_HIP(hipMemcpy(flag, &zero, sizeof(int), hipMemcpyHostToDevice));
        dm_lu_solve(st, dA, dB, dX, s, nb, full ? 1 : 0, mk(1.0, 0.0), nullptr, flag, BDDA_SM_LU);
        int f = 0;
        BDDA_HIP(hipMemcpyAsync(&f, flag, sizeof(int), hipMemcpyDeviceToHost, st));
        BDDA_HIP(hipStreamSynchronize(st));
        if (f) return fail(BDDA_EBREAKDOWN, "lu_solve: exactly zero pivot");
        if (nx) BDDA_HIP(hipMemcpy(X, dX, sizeof(c64) * nx, hipMemcpyDeviceToHost));
        return BDDA_OK;
    })
}

int bdda_hbm_probe(void* a, void* b, size_t bytes, double* copy_gbs, double* read_gbs, double* write_gbs)
{
    if (!a || !b || !copy_gbs || !read_gbs || !write_gbs) return fail(BDDA_EINVAL, "null pointer");
    if (bytes < (1u << 20)) return fail(BDDA_EINVAL, "use buffers of at least 1 MiB");
    BDDA_TRY({
        hbm_probe(a, b, bytes, copy_gbs, read_gbs, write_gbs);
        return BDDA_OK;
    })
}

int bdda_dev_alloc(size_t bytes, void** p)
{
    if (!p) return fail(BDDA_EINVAL, "null pointer");
    *p = nullptr;
    BDDA_TRY({
        *p = dev_malloc(bytes, true);
        return BDDA_OK;
    })
}

int bdda_dev_free(void* p)
{
    // waits for the whole device (a caller's buffer may be in use on any stream)
    dev_free(p, nullptr);
    return BDDA_OK;
}

int bdda_dev_trim(size_t* released)
{
    const size_t n = dev_trim();
    if (released) *released = n;
    return BDDA_OK;
}

int bdda_dev_pool_stats(size_t* held, size_t* in_use, int64_t* runtime_allocs, int64_t* reuses)
{
    const DevStats s = dev_stats();
    if (held) *held = s.held;
    if (in_use) *in_use = s.in_use;
    if (runtime_allocs) *runtime_allocs = s.runtime_allocs;
    if (reuses) *reuses = s.reuses;
    return BDDA_OK;
}

int bdda_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    return e == hipSuccess ? BDDA_OK : fail(BDDA_EHIP, hipGetErrorString(e));
}

int bdda_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    return e == hipSuccess ? BDDA_OK : fail(BDDA_EHIP, hipGetErrorString(e));
}

int bdda_memcpy_d2d(void* dst, const void* src, size_t bytes)
{
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice);
    return e == hipSuccess ? BDDA_OK : fail(BDDA_EHIP, hipGetErrorString(e));
}

int bdda_device_sync(void)
{
    hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? BDDA_OK : fail(BDDA_EHIP, hipGetErrorString(e));
}

int bdda_mbt_apply_occupied_dev(bdda_mbt* h, int prec, int jacobi, const void* dX, void* dY, int ncols)
{
    if (!h || !dX || !dY) return fail(BDDA_EINVAL, "null pointer");
    if (!h->have_target) return fail(BDDA_EINVAL, "bdda_mbt_set_target not called");
    if (int rc = check_prec(h, prec)) return rc;
    BDDA_TRY({
        WorkspaceScope wss{h->op};
        const int64_t rows = h->prob.rows();
        if (prec == BDDA_PREC_FP64)
            mbt_apply_occupied<double>(h->op, (const c64*)dX, rows, (c64*)dY, rows, ncols, h->prob.diag64,
                                       jacobi ? h->prob.jpre64 : nullptr, h->prob.nocc);
        else
            mbt_apply_occupied<float>(h->op, (const c32*)dX, rows, (c32*)dY, rows, ncols, h->prob.diag32,
                                      jacobi ? h->prob.jpre32 : nullptr, h->prob.nocc);
        if (!h->op->timing) BDDA_HIP(hipStreamSynchronize(h->stream));
        return BDDA_OK;
    })
}

int bdda_mbt_set_timing(bdda_mbt* h, int on)
{
    if (!h) return fail(BDDA_EINVAL, "null pointer");
    h->op->timing = on ? 1 : 0;
    return BDDA_OK;
}

int bdda_mbt_pass_times(const bdda_mbt* h, float pass_ms[5])
{
    if (!h || !pass_ms) return fail(BDDA_EINVAL, "null pointer");
    for (int i = 0; i < 5; ++i) pass_ms[i] = h->op->pass_ms[i];
    return BDDA_OK;
}

double bdda_mbt_algorithmic_bytes(const bdda_mbt* h, int prec, int ncols)
{
    if (!h) return 0.0;
    const size_t c = (prec == BDDA_PREC_FP32) ? 8 : 16;
    return mbt_algorithmic_bytes(h->op, ncols, h->have_target ? h->prob.nocc : (int64_t)h->n[0] * h->n[1] * h->n[2], c);
}

}  // extern "C"
