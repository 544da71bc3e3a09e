// observables.hip -- cross sections (Result_manager.cpp:65-122) and the
// far-field amplitude matrix of a solved dipole block, on the device in FP64.
#include "observables.hpp"
#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace bdda {

constexpr int NT_OBS = 256;

// per (column l, group g): sum conj(B) X (2 doubles) and sum w |X|^2 with
// w_i = -Im(DIAG_sel(i)) - 2k^3/3, DIAG_sel(i) = DIAG_i or DIAG_(i mod f)
__global__ __launch_bounds__(NT_OBS) void k_xs_partial(const c64* __restrict__ B, const c64* __restrict__ X,
                                                       const c64* __restrict__ diag, int64_t rows, int f,
                                                       int alpha_each, double rad, double* __restrict__ partial)
{
    __shared__ double sh[3][NT_OBS];
    const int l = blockIdx.y;
    const c64* Bl = B + (int64_t)l * rows;
    const c64* Xl = X + (int64_t)l * rows;
    double e0 = 0, e1 = 0, a = 0;
    for (int64_t i = blockIdx.x * (int64_t)NT_OBS + threadIdx.x; i < rows; i += (int64_t)gridDim.x * NT_OBS) {
        const c64 b = Bl[i], x = Xl[i];
        e0 += b.x * x.x + b.y * x.y;  // conj(b) x
        e1 += b.x * x.y - b.y * x.x;
        const c64 d = diag[alpha_each ? i : i % f];
        a += (-d.y - rad) * (x.x * x.x + x.y * x.y);
    }
    sh[0][threadIdx.x] = e0;
    sh[1][threadIdx.x] = e1;
    sh[2][threadIdx.x] = a;
    __syncthreads();
    for (int w = NT_OBS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) partial[((int64_t)l * gridDim.x + blockIdx.x) * 3 + q] = sh[q][0];
}

void cross_sections(Problem& P, const c64* dB, const c64* dX, int L, double k, int alpha_each, double* Cext,
                    double* Cabs)
{
    const int64_t rows = P.rows();
    const int nblk = (int)std::min<int64_t>(256, std::max<int64_t>(1, ceil_div(rows, NT_OBS)));
    DevPtr<double> dpb((size_t)3 * nblk * L, P.work.stream);
    double* dpart = dpb.p;
    std::vector<double> part((size_t)3 * nblk * L);
    {
        hipLaunchKernelGGL(k_xs_partial, dim3(nblk, L), dim3(NT_OBS), 0, P.work.stream, dB, dX, P.diag64, rows, P.f,
                           alpha_each, 2.0 * k * k * k / 3.0, dpart);
        BDDA_HIP(hipGetLastError());
        BDDA_HIP(hipMemcpyAsync(part.data(), dpart, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                                P.work.stream));
        BDDA_HIP(hipStreamSynchronize(P.work.stream));
    }
    for (int l = 0; l < L; ++l) {
        double im = 0, ab = 0;
        for (int g = 0; g < nblk; ++g) {
            im += part[((size_t)l * nblk + g) * 3 + 1];
            ab += part[((size_t)l * nblk + g) * 3 + 2];
        }
        Cext[l] = 4.0 * M_PI * k * im;
        Cabs[l] = 4.0 * M_PI * k * ab;
    }
}

// ------------------------------------------------ cross sections, no allocation
// Partial sums of k_xs_partial for callers that cannot allocate device
// memory: a file-scope device array holds them (XS_MAX_COLS columns of at
// most 256 groups, 3 sums each: 1.5 MiB), followed by the per-column totals.
// Every GPU the library runs on holds its own copy until the process ends.
constexpr int XS_MAX_GROUPS = 256;
static __device__ __align__(256) double scratch[(size_t)XS_MAX_COLS * (XS_MAX_GROUPS * 3 + 2)];

// out[l] and out[L + l] = sums over groups g (in order) of partial[(l G + g) 3 + 1]
// and of partial[(l G + g) 3 + 2]: the same order as the host sum of cross_sections
__global__ void k_xs_total(const double* __restrict__ partial, int ngroups, int L, double* __restrict__ out)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    double im = 0, ab = 0;
    for (int g = 0; g < ngroups; ++g) {
        im += partial[((int64_t)l * ngroups + g) * 3 + 1];
        ab += partial[((int64_t)l * ngroups + g) * 3 + 2];
    }
    out[l] = im;
    out[L + l] = ab;
}

double* xs_scratch_device()
{
    // hipGetSymbolAddress once per device, with that device current; the
    // array's host-side name is not a device address and is never passed on
    static std::mutex mu;
    static std::vector<double*> ptr;
    int dev = 0;
    BDDA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)ptr.size() <= dev) ptr.resize(dev + 1, nullptr);
    if (!ptr[dev]) {
        void* p = nullptr;
        BDDA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptr[dev] = static_cast<double*>(p);
    }
    return ptr[dev];
}

void cross_sections_noalloc(Problem& P, const c64* dB, const c64* dX, int L, double k, int alpha_each,
                            double* dscratch, double* Cext, double* Cabs)
{
    if (L < 1 || L > XS_MAX_COLS) throw std::invalid_argument("cross sections without allocation take 1.." +
                                                              std::to_string(XS_MAX_COLS) + " columns");
    const int64_t rows = P.rows();
    const int nblk = (int)std::min<int64_t>(XS_MAX_GROUPS, std::max<int64_t>(1, ceil_div(rows, NT_OBS)));
    double* dpart = dscratch;
    double* dsum = dscratch + (size_t)XS_MAX_COLS * XS_MAX_GROUPS * 3;
    hipStream_t st = P.work.stream;
    hipLaunchKernelGGL(k_xs_partial, dim3(nblk, L), dim3(NT_OBS), 0, st, dB, dX, P.diag64, rows, P.f, alpha_each,
                       2.0 * k * k * k / 3.0, dpart);
    BDDA_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_xs_total, dim3((unsigned)ceil_div(L, 64)), dim3(64), 0, st, dpart, nblk, L, dsum);
    BDDA_HIP(hipGetLastError());
    // the two planes of totals go straight into the caller's arrays (no
    // staging buffer) and are scaled there as cross_sections scales them
    BDDA_HIP(hipMemcpyAsync(Cext, dsum, sizeof(double) * L, hipMemcpyDeviceToHost, st));
    BDDA_HIP(hipMemcpyAsync(Cabs, dsum + L, sizeof(double) * L, hipMemcpyDeviceToHost, st));
    BDDA_HIP(hipStreamSynchronize(st));
    for (int l = 0; l < L; ++l) {
        Cext[l] = 4.0 * M_PI * k * Cext[l];
        Cabs[l] = 4.0 * M_PI * k * Cabs[l];
    }
}

// ------------------------------------------------------------- far field
constexpr int TJ_FF = 128;  // dipoles per LDS tile

struct FfGeom {
    int nx, ny, nz;
    double lf, k;
    int nang;
};

// group (chunk c, pair p, angle block): sums over the chunk's dipoles of
// P_j e^{-ik n.r_j} (and M_j for F = 6) for both columns of the pair, then
// the theta/phi projections of the far field -> partial[((c npairs + p)
// nang + a) 4 + q], q = theta1, phi1, theta2, phi2
template <int F>
__global__ __launch_bounds__(NT_OBS) void k_ff_partial(const c64* __restrict__ X, int64_t rows, int64_t nocc,
                                                       const int* __restrict__ addr, FfGeom g,
                                                       const double* __restrict__ u3, const double* __restrict__ e13,
                                                       int npairs, int64_t chunk, c64* __restrict__ partial)
{
    __shared__ double pos[TJ_FF][3];
    __shared__ c64 Pj[TJ_FF][2][F];
    const int p = blockIdx.y;
    const int a = blockIdx.z * NT_OBS + threadIdx.x;
    const bool live = a < g.nang;
    const double th = live ? M_PI * a / (g.nang - 1) : 0.0;
    const double ct = cos(th), st = sin(th);
    const double u[3] = {u3[3 * p], u3[3 * p + 1], u3[3 * p + 2]};
    const double e1[3] = {e13[3 * p], e13[3 * p + 1], e13[3 * p + 2]};
    const double nv[3] = {ct * u[0] + st * e1[0], ct * u[1] + st * e1[1], ct * u[2] + st * e1[2]};
    double acc[2][F][2];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int q = 0; q < F; ++q) acc[c][q][0] = acc[c][q][1] = 0.0;
    const int64_t j0 = blockIdx.x * chunk;
    const int64_t j1 = min(nocc, j0 + chunk);
    const int64_t nyz = (int64_t)g.ny * g.nz;
    for (int64_t t0 = j0; t0 < j1; t0 += TJ_FF) {
        const int nt = (int)min((int64_t)TJ_FF, j1 - t0);
        for (int e = threadIdx.x; e < nt; e += NT_OBS) {
            const int64_t cell = addr[t0 + e];
            const int64_t ix = cell / nyz, rem = cell - ix * nyz;
            const int64_t iy = rem / g.nz, iz = rem - iy * g.nz;
            pos[e][0] = g.lf * ix;
            pos[e][1] = g.lf * iy;
            pos[e][2] = g.lf * iz;
        }
        for (int e = threadIdx.x; e < nt * 2 * F; e += NT_OBS) {
            const int j = e / (2 * F), rest = e % (2 * F), c = rest / F, q = rest % F;
            Pj[j][c][q] = X[(int64_t)(2 * p + c) * rows + (t0 + j) * F + q];
        }
        __syncthreads();
        if (live)
            for (int j = 0; j < nt; ++j) {
                const double ph = -g.k * (nv[0] * pos[j][0] + nv[1] * pos[j][1] + nv[2] * pos[j][2]);
                double s, c;
                sincos(ph, &s, &c);
#pragma unroll
                for (int cc = 0; cc < 2; ++cc)
#pragma unroll
                    for (int q = 0; q < F; ++q) {
                        const c64 v = Pj[j][cc][q];
                        acc[cc][q][0] += c * v.x - s * v.y;
                        acc[cc][q][1] += c * v.y + s * v.x;
                    }
            }
        __syncthreads();
    }
    if (!live) return;
    const double tv[3] = {-st * u[0] + ct * e1[0], -st * u[1] + ct * e1[1], -st * u[2] + ct * e1[2]};
    const double pv[3] = {u[1] * e1[2] - u[2] * e1[1], u[2] * e1[0] - u[0] * e1[2], u[0] * e1[1] - u[1] * e1[0]};
    c64* out = partial + (((int64_t)blockIdx.x * npairs + p) * g.nang + a) * 4;
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
        double eth[2] = {0, 0}, eph[2] = {0, 0};
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                eth[r] += tv[q] * acc[cc][q][r];
                eph[r] += pv[q] * acc[cc][q][r];
                if constexpr (F == 6) {  // -n x M: theta.(-n x M) = phi.M, phi.(-n x M) = -theta.M
                    eth[r] += pv[q] * acc[cc][3 + q][r];
                    eph[r] -= tv[q] * acc[cc][3 + q][r];
                }
            }
        out[2 * cc] = mk(eth[0], eth[1]);
        out[2 * cc + 1] = mk(eph[0], eph[1]);
    }
}

// out[e] = sum_c partial[c n + e] in chunk order
__global__ void k_ff_reduce(const c64* __restrict__ partial, int nchunk, int64_t n, c64* __restrict__ out)
{
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    double sx = 0, sy = 0;
    for (int c = 0; c < nchunk; ++c) {
        const c64 v = partial[(int64_t)c * n + e];
        sx += v.x;
        sy += v.y;
    }
    out[e] = mk(sx, sy);
}

void scattering_amplitudes(Problem& P, const c64* dX, int npairs, double k, const double* u, const double* e1,
                           int nang, c64* S)
{
    const MbtOperator* op = P.op;
    const int64_t nocc = P.nocc, rows = P.rows();
    const int nchunk = (int)std::min<int64_t>(std::max<int64_t>(1, ceil_div(nocc, 4 * TJ_FF)), 256);
    const int64_t chunk = ceil_div(ceil_div(nocc, nchunk), TJ_FF) * TJ_FF;
    const int64_t n = (int64_t)npairs * nang * 4;
    FfGeom g{op->d.nx, op->d.ny, op->d.nz, op->lf, k, nang};
    std::vector<double> geo((size_t)6 * npairs);
    for (int p = 0; p < npairs; ++p)
        for (int q = 0; q < 3; ++q) {
            geo[3 * p + q] = u[3 * p + q];
            geo[3 * npairs + 3 * p + q] = e1[3 * p + q];
        }
    std::vector<c64> red(n);
    {
        hipStream_t st = P.work.stream;
        DevPtr<double> dgeob(geo.size(), st);
        DevPtr<c64> dpartb((size_t)n * nchunk, st), doutb((size_t)n, st);
        double* dgeo = dgeob.p;
        c64* dpart = dpartb.p;
        c64* dout = doutb.p;
        BDDA_HIP(hipMemcpyAsync(dgeo, geo.data(), sizeof(double) * geo.size(), hipMemcpyHostToDevice, st));
        const dim3 grid(nchunk, npairs, (unsigned)ceil_div(nang, NT_OBS));
        if (P.f == 3)
            hipLaunchKernelGGL(k_ff_partial<3>, grid, dim3(NT_OBS), 0, st, dX, rows, nocc, op->addr_dev, g, dgeo,
                               dgeo + 3 * npairs, npairs, chunk, dpart);
        else
            hipLaunchKernelGGL(k_ff_partial<6>, grid, dim3(NT_OBS), 0, st, dX, rows, nocc, op->addr_dev, g, dgeo,
                               dgeo + 3 * npairs, npairs, chunk, dpart);
        BDDA_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_ff_reduce, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, dpart, nchunk, n, dout);
        BDDA_HIP(hipGetLastError());
        BDDA_HIP(hipMemcpyAsync(red.data(), dout, sizeof(c64) * n, hipMemcpyDeviceToHost, st));
        BDDA_HIP(hipStreamSynchronize(st));
    }
    // S = -ik * k^2 * projection: S2 = theta1, S4 = phi1, S3 = theta2, S1 = phi2
    const double k3 = k * k * k;
    for (int64_t e = 0; e < (int64_t)npairs * nang; ++e) {
        const c64* r = &red[e * 4];
        auto f = [&](c64 v) { return mk(k3 * v.y, -k3 * v.x); };  // -i k^3 v
        S[e * 4 + 0] = f(r[3]);
        S[e * 4 + 1] = f(r[0]);
        S[e * 4 + 2] = f(r[2]);
        S[e * 4 + 3] = f(r[1]);
    }
}

}  // namespace bdda
