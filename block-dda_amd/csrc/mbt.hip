// mbt.hip -- MBT operator: G^ build and the fused FFT matvec passes (gfx950).
//
// Reference behaviour reproduced (paths relative to the reference root):
//   application_function   mvp_fft/application_function.cpp:21-67
//   MBT_fft_init           mvp_fft/MBT_fft_init.cpp:10-42
//   MBT_fft_mvp            mvp_fft/MBT_fft_mvp.cpp:8-43
//   BT_pad / BT_reconstruct  (index embedding / extraction; here fused into
//                          the first / last pass, no padded vector is formed)
//   inline block matvec    bl_cocg_rq_jacobi_mvp_fft.cpp:43-60 (scatter,
//                          DIAG_A .* x, gather, jpre .*)
// The reference's 1-D FFT of length f^2 prod(2n-1) equals a zero-padded 3-D
// convolution over the cuboid (no index digit borrows); this file computes
// that convolution with 3-D FFTs on a smooth padded grid instead.
#include "mbt_passes.hpp"
#include "devmem.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <stdexcept>

namespace bdda {

// ------------------------------------------------------------------ sizes
static bool smooth7(int v)
{
    for (int p : {2, 3, 5, 7})
        while (v % p == 0) v /= p;
    return v == 1;
}

int choose_padded(int n, int mult, int plan_mult)
{
    const int need = std::max(1, 2 * n - 1);
    int smooth = need;
    while (!(smooth % mult == 0 && smooth7(smooth))) ++smooth;
    // prefer a size with a compile-time plan unless it pads more than 12 %
    int planned = need;
    while (planned <= 1024 && !(has_plan(planned) && planned % plan_mult == 0)) ++planned;
    if (planned <= 1024 && planned <= smooth + smooth / 8) return planned;
    return smooth;
}

bool mbt_padded_dims(const int n[3], int f, int P[3])
{
    // the all-in-LDS fused x kernel (sizes without a plan) needs Px % 8 == 0;
    // for f = 6 the x size is taken as a multiple of 8 so that it has that
    // fallback (the planned x kernel serves both f)
    P[0] = choose_padded(n[0], 8, f == 3 ? 1 : 8);
    P[1] = choose_padded(n[1], 1, 1);
    P[2] = choose_padded(n[2], 1, 1);
    for (int a = 0; a < 3; ++a)
        if ((int64_t)P[a] * f > MBT_MAX_PF) return false;
    return true;
}

AxisPlan make_axis_plan(int P, bool last8)
{
    AxisPlan ap{};
    ap.P = P;
    int a = 0, b = 0, c = 0, e = 0, v = P;
    while (v % 2 == 0) { v /= 2; ++a; }
    while (v % 3 == 0) { v /= 3; ++b; }
    while (v % 5 == 0) { v /= 5; ++c; }
    while (v % 7 == 0) { v /= 7; ++e; }
    std::vector<int> r;
    int n8 = 0;
    if (last8) { n8 = 1; a -= 3; }
    // powers of two: 8s, then fix the remainder with 4s / 2 (or a 6)
    int m8 = a / 3, rem = a % 3;
    std::vector<int> twos;
    for (int i = 0; i < m8; ++i) twos.push_back(8);
    if (rem == 2) twos.push_back(4);
    if (rem == 1) {
        if (b > 0) { r.push_back(6); --b; }
        else if (m8 > 0) { twos.pop_back(); twos.push_back(4); twos.push_back(4); }
        else twos.push_back(2);
    }
    for (int i = 0; i < e; ++i) r.push_back(7);
    for (int i = 0; i < c; ++i) r.push_back(5);
    for (int i = 0; i < b; ++i) r.push_back(3);
    for (int t : twos) r.push_back(t);
    if (n8) r.push_back(8);
    ap.nst = (int)r.size();
    for (int i = 0; i < ap.nst; ++i) ap.rad[i] = r[i];
    return ap;
}

// ---------------------------------------------------------------- kernels

// Interaction kernel on the full padded grid, FP64, pre-scaled by 1/(PxPyPz).
// Circular embedding: axis index x -> displacement d = x (x <= n-1) or
// x - P (x >= P-n+1), zero otherwise.  Component order as in mbt.hpp.
__global__ void k_green_eval(c64* __restrict__ out, MbtDims d, int nc, double lf, double kre,
                             double kim, double scale)
{
    const int64_t total = (int64_t)d.Px * d.Py * d.Pz;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(i % d.Pz);
        const int y = (int)((i / d.Pz) % d.Py);
        const int x = (int)(i / ((int64_t)d.Pz * d.Py));
        auto disp = [](int v, int n, int P, bool& ok) {
            if (v <= n - 1) return v;
            if (v >= P - n + 1) return v - P;
            ok = false;
            return 0;
        };
        bool ok = true;
        const int dx = disp(x, d.nx, d.Px, ok), dy = disp(y, d.ny, d.Py, ok), dz = disp(z, d.nz, d.Pz, ok);
        c64* o = out + i * nc;
        for (int c = 0; c < nc; ++c) o[c] = mk(0.0, 0.0);
        if (!ok || (dx == 0 && dy == 0 && dz == 0)) continue;
        const double X[3] = {dx * lf, dy * lf, dz * lf};
        const double r = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
        const double u[3] = {X[0] / r, X[1] / r, X[2] / r};
        // e^{ikr}/r with complex k
        double sn, cs;
        sincos(kre * r, &sn, &cs);
        const double damp = exp(-kim * r) / r;
        const c64 e = mk(damp * cs, damp * sn);
        const c64 k = mk(kre, kim);
        const c64 k2 = k * k;
        const c64 ikr = mk(-kim * r, kre * r);
        // (1 - ikr)/r^2
        const c64 c2 = (1.0 / (r * r)) * mk(1.0 - ikr.x, -ikr.y);
        const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
        for (int c = 0; c < 6; ++c) {
            const double uu = u[pa[c]] * u[pb[c]];
            const double id = (pa[c] == pb[c]) ? 1.0 : 0.0;
            const c64 g = (k2 * mk(id - uu, 0.0) - c2 * mk(id - 3.0 * uu, 0.0)) * e;
            o[c] = (-scale) * g;  // M00 = -G
        }
        if (nc == 9) {
            // cf = -k^2 e^{ikr}/r (1 - 1/(ikr)); M01 = -F = -cf [u]_x
            const double den = ikr.x * ikr.x + ikr.y * ikr.y;
            const c64 inv_ikr = mk(ikr.x / den, -ikr.y / den);
            const c64 cf = (-1.0) * (k2 * e * mk(1.0 - inv_ikr.x, -inv_ikr.y));
            o[6] = (scale * u[2]) * cf;   // M01_xy = -cf * (-u_z)
            o[7] = (-scale * u[1]) * cf;  // M01_xz = -cf * ( u_y)
            o[8] = (scale * u[0]) * cf;   // M01_yz = -cf * (-u_x)
        }
    }
}

// extract the reduced octant of the transformed kernel (and an FP32 copy)
__global__ void k_reduce_gh(const c64* __restrict__ full, c64* __restrict__ gh64, c32* __restrict__ gh32,
                            MbtDims d, int nc)
{
    const int Hx = d.Px / 2 + 1, Hy = d.Py / 2 + 1, Hz = d.Pz / 2 + 1;
    const int64_t total = (int64_t)Hx * Hy * Hz * nc;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % nc);
        const int64_t f = i / nc;
        const int kz = (int)(f % Hz), ky = (int)((f / Hz) % Hy), kx = (int)(f / ((int64_t)Hz * Hy));
        const c64 v = full[(((int64_t)kx * d.Py + ky) * d.Pz + kz) * nc + c];
        if (gh64) gh64[i] = v;
        if (gh32) gh32[i] = cvt<float>(v);
    }
}

__global__ void k_build_c2o(const int* __restrict__ address, int64_t nocc, int* __restrict__ c2o)
{
    for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < nocc;
         m += (int64_t)gridDim.x * blockDim.x)
        c2o[address[m]] = (int)m;
}

__global__ void k_fill_int(int* p, int64_t n, int v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ------------------------------------------------------------ host side
static int grid_for(int64_t n) { return (int)grid_cap(n); }

template <typename R>
static void upload_twiddles(int P, void** dst)
{
    std::vector<cplx<R>> h(P);
    for (int q = 0; q < P; ++q) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * q / P;
        h[q] = mk((R)std::cos(a), (R)std::sin(a));
    }
    *dst = dev_alloc_n<cplx<R>>(P, true);
    BDDA_HIP(hipMemcpy(*dst, h.data(), sizeof(cplx<R>) * P, hipMemcpyHostToDevice));
}

MbtOperator* mbt_create(const int n[3], int f, double lf, double kre, double kim, int want64, int want32,
                        hipStream_t stream)
{
    auto* op = new MbtOperator();
    op->d.nx = n[0];
    op->d.ny = n[1];
    op->d.nz = n[2];
    op->d.f = f;
    int P[3];
    if (!mbt_padded_dims(n, f, P)) {
        delete op;
        throw std::invalid_argument("padded grid beyond the passes' limit (P * f <= 3072 on every axis)");
    }
    op->d.Px = P[0];
    op->d.Py = P[1];
    op->d.Pz = P[2];
    op->lf = lf;
    op->kre = kre;
    op->kim = kim;
    op->nc = (f == 3) ? 6 : 9;
    op->stream = stream;
    op->ax[0] = make_axis_plan(op->d.Px, op->d.Px % 8 == 0);
    op->ax[1] = make_axis_plan(op->d.Py, false);
    op->ax[2] = make_axis_plan(op->d.Pz, false);
    op->Pzs = mbt_z_pitch(op->d.Pz);
    const MbtDims& d = op->d;
    try {
        for (int a = 0; a < 3; ++a) {
            upload_twiddles<double>(op->ax[a].P, &op->tw64[a]);
            upload_twiddles<float>(op->ax[a].P, &op->tw32[a]);
        }
        // full FP64 kernel grid, forward 3-D FFT, octant extraction
        const int64_t G = (int64_t)d.Px * d.Py * d.Pz;
        const int nc = op->nc;
        // the build transient (two full FP64 kernel grids), freed below
        DevPtr<c64> fullb((size_t)G * nc, stream), tmpb((size_t)G * nc, stream);
        c64* full = fullb.p;
        c64* tmp = tmpb.p;
        const double scale = 1.0 / (double)G;
        hipLaunchKernelGGL(k_green_eval, dim3(grid_for(G)), dim3(256), 0, stream, full, d, nc, lf, kre, kim, scale);
        BDDA_HIP(hipGetLastError());
        const c64* tw[3] = {(const c64*)op->tw64[0], (const c64*)op->tw64[1], (const c64*)op->tw64[2]};
        // z: lines along z, stride nc
        launch_lines<double, false>(full, tmp, d.Pz, d.Pz, nc, nc, (int64_t)d.Px * d.Py, op->ax[2], tw[2], stream);
        // y: stride Pz nc
        launch_lines<double, false>(tmp, full, d.Py, d.Py, (int64_t)d.Pz * nc, (int64_t)d.Pz * nc, d.Px, op->ax[1], tw[1],
                                    stream);
        // x: stride Py Pz nc
        launch_lines<double, false>(full, tmp, d.Px, d.Px, (int64_t)d.Py * d.Pz * nc, (int64_t)d.Py * d.Pz * nc, 1,
                                    op->ax[0], tw[0], stream);
        op->gh_elems = (size_t)(d.Px / 2 + 1) * (d.Py / 2 + 1) * (d.Pz / 2 + 1);
        if (want64) op->gh64 = dev_alloc_n<c64>(op->gh_elems * nc, true);
        if (want32) op->gh32 = dev_alloc_n<c32>(op->gh_elems * nc, true);
        hipLaunchKernelGGL(k_reduce_gh, dim3(grid_for((int64_t)op->gh_elems * nc)), dim3(256), 0, stream, tmp,
                           (c64*)op->gh64, (c32*)op->gh32, d, nc);
        BDDA_HIP(hipGetLastError());
        BDDA_HIP(hipStreamSynchronize(stream));
        op->have64 = want64 ? 1 : 0;
        op->have32 = want32 ? 1 : 0;
        for (auto& e : op->ev) BDDA_HIP(hipEventCreate(&e));
    } catch (...) {
        mbt_destroy(op);
        throw;
    }
    return op;
}

void mbt_destroy(MbtOperator* op)
{
    if (!op) return;
    // (the stream outlives the operator; its handle owns both)
    hipStream_t st = op->stream;
    for (int a = 0; a < 3; ++a) {
        dev_free(op->tw64[a], st);
        dev_free(op->tw32[a], st);
    }
    dev_free(op->gh64, st);
    dev_free(op->gh32, st);
    dev_free(op->ws, st);
    dev_free(op->c2o, st);
    dev_free(op->addr_dev, st);
    dev_free(op->lines, st);
    dev_free(op->lmask, st);
    for (auto& e : op->ev)
        if (e) (void)hipEventDestroy(e);
    delete op;
}

void mbt_set_address(MbtOperator* op, const int* address_host, int64_t nocc)
{
    const int64_t C = (int64_t)op->d.nx * op->d.ny * op->d.nz;
    if (!op->c2o) {
        op->c2o = dev_alloc_n<int>(C, true);
        op->c2o_cells = C;
    }
    if (op->addr_n < nocc) {
        dev_free(op->addr_dev, op->stream);
        op->addr_dev = nullptr;
        op->addr_n = 0;
        op->addr_dev = dev_alloc_n<int>(std::max<int64_t>(nocc, 1), true);
        op->addr_n = nocc;
    }
    BDDA_HIP(hipMemcpyAsync(op->addr_dev, address_host, sizeof(int) * nocc, hipMemcpyHostToDevice, op->stream));
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(C)), dim3(256), 0, op->stream, op->c2o, C, -1);
    hipLaunchKernelGGL(k_build_c2o, dim3(grid_for(nocc)), dim3(256), 0, op->stream, (const int*)op->addr_dev, nocc,
                       op->c2o);
    BDDA_HIP(hipGetLastError());
    // occupied (ix, iy) lines: flags and the ascending list of their ids
    const int64_t nxny = (int64_t)op->d.nx * op->d.ny;
    std::vector<unsigned char> mask(nxny, 0);
    for (int64_t m = 0; m < nocc; ++m) mask[address_host[m] / op->d.nz] = 1;
    std::vector<int> ids;
    for (int64_t l = 0; l < nxny; ++l)
        if (mask[l]) ids.push_back((int)l);
    if (!op->lmask) op->lmask = dev_alloc_n<unsigned char>(nxny, true);
    dev_free(op->lines, op->stream);
    op->lines = nullptr;
    op->nlines = 0;
    op->lines = dev_alloc_n<int>(std::max<size_t>(ids.size(), 1), true);
    BDDA_HIP(hipMemcpyAsync(op->lmask, mask.data(), nxny, hipMemcpyHostToDevice, op->stream));
    if (!ids.empty())
        BDDA_HIP(hipMemcpyAsync(op->lines, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice, op->stream));
    BDDA_HIP(hipStreamSynchronize(op->stream));
    op->nlines = (int64_t)ids.size();
    op->c2o_host_addr.assign(address_host, address_host + nocc);
}

void mbt_ensure_workspace(MbtOperator* op, size_t bytes)
{
    if (op->ws_bytes >= bytes) return;
    if (op->ws) {
        dev_free(op->ws, op->stream);
        op->ws = nullptr;
        op->ws_bytes = 0;
    }
    op->ws = dev_malloc(bytes);
    op->ws_bytes = bytes;
}

void mbt_release_workspace(MbtOperator* op)
{
    dev_free(op->ws, op->stream);
    op->ws = nullptr;
    op->ws_bytes = 0;
}

template <typename R>
void mbt_apply_occupied(MbtOperator* op, const cplx<R>* X, int64_t ldx, cplx<R>* Y, int64_t ldy, int s,
                        const cplx<R>* diag, const cplx<R>* jpre, int64_t nocc)
{
    // nocc < 0 selects the cuboid layout (MBT_fft_mvp: every cell present)
    const int* c2o = (nocc >= 0) ? op->c2o : nullptr;
    if (nocc >= 0 && !op->c2o) throw HipError{hipErrorInvalidValue, "address not set"};
    mbt_apply_precision<R>(op, X, ldx, Y, ldy, s, diag, jpre, c2o);
}

template void mbt_apply_occupied<float>(MbtOperator*, const c32*, int64_t, c32*, int64_t, int, const c32*,
                                        const c32*, int64_t);
template void mbt_apply_occupied<double>(MbtOperator*, const c64*, int64_t, c64*, int64_t, int, const c64*,
                                         const c64*, int64_t);

// Pz rounded up to 16 elements, and 16 more when that is a multiple of 64:
// the x pass reads 64-byte row segments one plane (Py Pzs elements) apart,
// and planes of power-of-two multiples of 64 elements measured slower (x pass
// at 512^3, s = 16: FP32 33.1 -> 28.0 ms, FP64 61.8 -> 52.7 ms with the pitch
// 528 instead of 512; 448^3: 23.5 -> 20.9 ms).  Padding 352 (343^3) to 368
// measured no faster, and 32 more at 512 no better than 16.
int mbt_z_pitch(int Pz)
{
    const int p = (int)ceil_div(Pz, 16) * 16;
    return p % 64 == 0 ? p + 16 : p;
}

size_t mbt_workspace_per_column(const MbtDims& d, int Pzs, size_t elem_bytes)
{
    return elem_bytes * (size_t)d.f * ((size_t)d.nx * d.ny + (size_t)d.nx * d.Py) * (size_t)Pzs;
}

double mbt_algorithmic_bytes(const MbtOperator* op, int s, int64_t nocc, size_t c)
{
    const MbtDims& d = op->d;
    const double f = d.f;
    // the planned f = 3 passes move only the rows of occupied (ix, iy) lines
    const bool lined = d.f == 3 && op->lines && nocc >= 0 && has_plan(d.Py) && has_plan(d.Pz);
    const double nxnyPz = (lined ? (double)op->nlines : (double)d.nx * d.ny) * d.Pz;
    const double nxPyPz = (double)d.nx * d.Py * d.Pz;
    const double per = s * f * c * (3.0 * nocc + 4.0 * nxnyPz + 4.0 * nxPyPz);
    return per + (double)c * op->nc * op->gh_elems;
}

}  // namespace bdda
