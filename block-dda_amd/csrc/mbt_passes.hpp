// mbt_passes.hpp -- the five fused FFT passes of one MBT block matvec and
// the per-precision driver; included by mbt_f32.hip and mbt_f64.hip, each
// of which instantiates one precision (separate translation units so the
// two compile in parallel).
//
// Pass sequence of one block matvec (s RHS, f components, all batched):
//   z forward + scatter (X -> U), y forward (U -> V), x forward * G^ * x
//   inverse (V in place), y inverse (V -> U), z inverse + gather + diag
//   (+ Jacobi) (U -> Y).  Layouts are described in mbt.hpp.
#pragma once
#include <algorithm>
#include <cstdlib>
#include "mbt.hpp"

namespace bdda {

// element budget per work-group (elements held in LDS and spread over
// NT_PASS threads): 24 per thread in FP32, 12 per thread in FP64, so one
// group uses 48 KB of LDS in either precision and three fit on a CU.
constexpr int NT_PASS = 256;
template <typename R>
struct Budget;
template <>
struct Budget<float> {
    static constexpr int MAXE = 12;
    static constexpr int SEG = 16;  // elements per 128-byte segment
};
template <>
struct Budget<double> {
    static constexpr int MAXE = 12;
    static constexpr int SEG = 8;
};
static_assert(Budget<float>::MAXE * NT_PASS == MBT_MAX_PF && Budget<double>::MAXE * NT_PASS == MBT_MAX_PF,
              "mbt_padded_dims must refuse what the all-in-LDS passes cannot hold");



inline int64_t grid_cap(int64_t n, int bs = 256)
{
    return std::min<int64_t>(std::max<int64_t>(ceil_div(n, bs), 1), 65536);
}

template <typename R>
inline int tile_len(int P, int64_t Zdim)
{
    const int cap = Budget<R>::MAXE * NT_PASS / P;
    int t = cap;
    if (t >= Budget<R>::SEG) t = (t / Budget<R>::SEG) * Budget<R>::SEG;
    t = (int)std::min<int64_t>(t, Zdim);
    return std::max(t, 1);
}

// Batched FFT along the middle axis of in[outer][nlin][pitch] into
// out[outer][nlout][pitch] over the first Zdim columns (zero padding
// nlin -> P, keeping nlout outputs).  A group owns Tz consecutive columns.
template <typename R, bool INV>
__global__ __launch_bounds__(NT_PASS) void k_lines_strided(const cplx<R>* __restrict__ in,
                                                          cplx<R>* __restrict__ out, int nlin,
                                                          int nlout, int64_t Zdim, int64_t pitch, int64_t ntiles,
                                                          AxisPlan ap, const cplx<R>* __restrict__ twg,
                                                          int Tz)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<R>* tw = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* lds = tw + ap.P;
    const int P = ap.P;
    const int64_t outer = blockIdx.x / ntiles;
    const int64_t z0 = (blockIdx.x % ntiles) * (int64_t)Tz;
    const int tz = (int)min((int64_t)Tz, Zdim - z0);
    for (int i = threadIdx.x; i < P; i += NT_PASS) tw[i] = twg[i];
    const cplx<R>* src = in + (outer * nlin) * pitch + z0;
    for (int e = threadIdx.x; e < P * Tz; e += NT_PASS) {
        const int p = e / Tz, t = e % Tz;
        cplx<R> v = czero<R>();
        if (p < nlin && t < tz) v = src[(int64_t)p * pitch + t];
        lds[p * Tz + t] = v;
    }
    __syncthreads();
    const StageGeom g{P, Tz, Tz, 1, true};
    run_stages<Budget<R>::MAXE, NT_PASS, R, INV>(lds, tw, g, ap, 0, ap.nst);
    cplx<R>* dst = out + (outer * nlout) * pitch + z0;
    for (int e = threadIdx.x; e < nlout * Tz; e += NT_PASS) {
        const int p = e / Tz, t = e % Tz;
        if (t < tz) dst[(int64_t)p * pitch + t] = lds[p * Tz + t];
    }
}

template <typename R, bool INV>
inline void launch_lines(const cplx<R>* in, cplx<R>* out, int nlin, int nlout, int64_t Zdim, int64_t pitch,
                         int64_t outer, const AxisPlan& ap, const cplx<R>* tw, hipStream_t st)
{
    if (ap.P * 1LL > (int64_t)Budget<R>::MAXE * NT_PASS) throw HipError{hipErrorInvalidValue, "axis too long"};
    const int Tz = tile_len<R>(ap.P, Zdim);
    const int64_t ntiles = ceil_div(Zdim, Tz);
    const size_t shm = sizeof(cplx<R>) * (ap.P + (size_t)ap.P * Tz);
    hipLaunchKernelGGL((k_lines_strided<R, INV>), dim3((unsigned)(ntiles * outer)), dim3(NT_PASS), shm, st, in, out,
                       nlin, nlout, Zdim, pitch, ntiles, ap, tw, Tz);
    BDDA_HIP(hipGetLastError());
}

// Sign of a reduced G^ component when frequency axes are reflected.
// flips: bit0 = x, bit1 = y, bit2 = z reflected.  Odd-axis masks:
//   M00: xx,yy,zz even; xy odd in x,y; xz odd in x,z; yz odd in y,z
//   M01 (antisymmetric): xy ~ u_z odd in z; xz ~ u_y odd in y; yz ~ u_x odd in x
__device__ __forceinline__ int comp_odd_mask(int c)
{
    constexpr int masks[9] = {0, 0, 0, 3, 5, 6, 4, 2, 1};
    return masks[c];
}


}  // namespace bdda
#include "mbt_plans.hpp"
namespace bdda {

// ------------------------------------------------------------------
// all-in-LDS kernels for padded sizes without a compile-time plan.  Layouts
// as in mbt.hpp: U[((rhs F + c) nx + ix) ny + iy][Pzs],
// V[((rhs F + c) nx + ix) Py + ky][Pzs].

// z forward with the scatter of the occupied vector fused in.  A group owns
// TL consecutive (ix, iy) lines with all F components.
template <typename R, int F>
__global__ __launch_bounds__(NT_PASS) void k_zfwd_scatter(const cplx<R>* __restrict__ X, int64_t ldx,
                                                         const int* __restrict__ c2o,
                                                         cplx<R>* __restrict__ U, MbtDims d, int Pzs,
                                                         AxisPlan ap, const cplx<R>* __restrict__ twg,
                                                         int TL, int64_t ntiles)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<R>* tw = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* lds = tw + ap.P;
    const int Pz = ap.P;
    const int Tp = TL * F;
    const int64_t nxny = (int64_t)d.nx * d.ny;
    const int rhs = (int)(blockIdx.x / ntiles);
    const int64_t l0 = (blockIdx.x % ntiles) * (int64_t)TL;
    const int nl = (int)min((int64_t)TL, nxny - l0);
    for (int i = threadIdx.x; i < Pz; i += NT_PASS) tw[i] = twg[i];
    const cplx<R>* Xr = X + rhs * ldx;
    for (int e = threadIdx.x; e < TL * Pz * F; e += NT_PASS) {
        const int c = e % F;
        const int rest = e / F;
        const int p = rest % Pz, tl = rest / Pz;
        cplx<R> v = czero<R>();
        if (tl < nl && p < d.nz) {
            const int64_t cell = (l0 + tl) * d.nz + p;
            const int64_t m = c2o ? (int64_t)c2o[cell] : cell;
            if (m >= 0) v = Xr[m * F + c];
        }
        lds[p * Tp + tl * F + c] = v;
    }
    __syncthreads();
    const StageGeom g{Pz, Tp, Tp, 1, true};
    run_stages<Budget<R>::MAXE, NT_PASS, R, false>(lds, tw, g, ap, 0, ap.nst);
    for (int e = threadIdx.x; e < F * nl * Pz; e += NT_PASS) {
        const int p = e % Pz;
        const int rest = e / Pz;
        const int tl = rest % nl, c = rest / nl;
        U[((((int64_t)rhs * F + c) * nxny) + l0 + tl) * Pzs + p] = lds[p * Tp + tl * F + c];
    }
}

// z inverse with the gather fused in:
// Y[f m + c] = diag[f m + c] X[f m + c] + IFFT_z(U)[cell][c]   (then * jpre)
template <typename R, int F>
__global__ __launch_bounds__(NT_PASS) void k_zinv_gather(const cplx<R>* __restrict__ U,
                                                        const int* __restrict__ c2o,
                                                        const cplx<R>* __restrict__ X, int64_t ldx,
                                                        const cplx<R>* __restrict__ diag,
                                                        const cplx<R>* __restrict__ jpre,
                                                        cplx<R>* __restrict__ Y, int64_t ldy, MbtDims d, int Pzs,
                                                        AxisPlan ap, const cplx<R>* __restrict__ twg,
                                                        int TL, int64_t ntiles)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<R>* tw = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* lds = tw + ap.P;
    const int Pz = ap.P;
    const int Tp = TL * F;
    const int64_t nxny = (int64_t)d.nx * d.ny;
    const int rhs = (int)(blockIdx.x / ntiles);
    const int64_t l0 = (blockIdx.x % ntiles) * (int64_t)TL;
    const int nl = (int)min((int64_t)TL, nxny - l0);
    for (int i = threadIdx.x; i < Pz; i += NT_PASS) tw[i] = twg[i];
    for (int e = threadIdx.x; e < F * TL * Pz; e += NT_PASS) {
        const int p = e % Pz;
        const int rest = e / Pz;
        const int tl = rest % TL, c = rest / TL;
        cplx<R> v = czero<R>();
        if (tl < nl) v = U[((((int64_t)rhs * F + c) * nxny) + l0 + tl) * Pzs + p];
        lds[p * Tp + tl * F + c] = v;
    }
    __syncthreads();
    const StageGeom g{Pz, Tp, Tp, 1, true};
    run_stages<Budget<R>::MAXE, NT_PASS, R, true>(lds, tw, g, ap, 0, ap.nst);
    const cplx<R>* Xr = X + rhs * ldx;
    cplx<R>* Yr = Y + rhs * ldy;
    for (int e = threadIdx.x; e < nl * d.nz * F; e += NT_PASS) {
        const int c = e % F;
        const int rest = e / F;
        const int p = rest % d.nz, tl = rest / d.nz;
        const int64_t cell = (l0 + tl) * d.nz + p;
        const int64_t m = c2o ? (int64_t)c2o[cell] : cell;
        if (m < 0) continue;
        const int64_t i = m * F + c;
        cplx<R> v = lds[p * Tp + tl * F + c];
        if (diag) v = diag[i] * Xr[i] + v;
        if (jpre) v = v * jpre[i];
        Yr[i] = v;
    }
}

// x pass: forward FFT along x, per-frequency f x f multiply by G^, inverse
// FFT along x, in place on V.  A group owns Tq consecutive positions q of
// the padded (ky, kz) plane with all F components.  The forward's last
// stage (radix 8) keeps its outputs in registers, the multiply runs there,
// and the inverse's first stage (radix 8, no twiddles) starts from them.
template <typename R, int F>
__global__ __launch_bounds__(NT_PASS) void k_xfused(cplx<R>* __restrict__ V, const cplx<R>* __restrict__ Gh,
                                                   MbtDims d, int Pzs, AxisPlan ap, const cplx<R>* __restrict__ twg,
                                                   int Tq, int64_t ntiles)
{
    constexpr int RL = 8;
    constexpr int NC = (F == 3) ? 6 : 9;
    extern __shared__ unsigned char smem_raw[];
    cplx<R>* tw = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* lds = tw + ap.P;
    const int Px = ap.P;
    const int64_t plane = (int64_t)d.Py * Pzs;
    const int Tp = Tq * F;
    const int rhs = (int)(blockIdx.x / ntiles);
    const int64_t q0 = (blockIdx.x % ntiles) * (int64_t)Tq;
    const int tqn = (int)min((int64_t)Tq, plane - q0);
    for (int i = threadIdx.x; i < Px; i += NT_PASS) tw[i] = twg[i];
    cplx<R>* base = V + (int64_t)rhs * F * d.nx * plane + q0;
    const int64_t cstride = (int64_t)d.nx * plane;
    for (int e = threadIdx.x; e < Px * Tp; e += NT_PASS) {
        const int p = e / Tp, t = e % Tp;
        const int tq = t / F, c = t % F;
        cplx<R> v = czero<R>();
        if (p < d.nx && tq < tqn) v = base[c * cstride + (int64_t)p * plane + tq];
        lds[p * Tp + t] = v;
    }
    __syncthreads();
    const StageGeom g{Px, Tp, Tp, 1, true};
    run_stages<Budget<R>::MAXE, NT_PASS, R, false>(lds, tw, g, ap, 0, ap.nst - 1);

    const int nb = Px / RL;  // butterflies per line in the last stage; L = nb
    const int total = nb * Tq;
    const int b = threadIdx.x;  // one frequency-line butterfly per thread (host checks total <= NT)
    cplx<R> h[F][RL];
    int tq = 0, j = 0;
    if (b < total) {
        tq = b % Tq;
        j = b / Tq;
#pragma unroll
        for (int c = 0; c < F; ++c)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                cplx<R> x = lds[(j + r * nb) * Tp + tq * F + c];
                if (r > 0 && j > 0) x = x * tw[r * j];
                h[c][r] = x;
            }
    }
    __syncthreads();
    if (b < total) {
#pragma unroll
        for (int c = 0; c < F; ++c) Dft<RL, R, false>::run(h[c]);
        const int64_t q = q0 + tq;
        const int ky = (int)(q / Pzs), kz = (int)(q % Pzs);
        if (tq < tqn && kz < d.Pz) {
            const int Hy = d.Py / 2 + 1, Hz = d.Pz / 2 + 1;
            int fl = 0;
            int kyr = ky, kzr = kz;
            if (ky > d.Py / 2) { kyr = d.Py - ky; fl |= 2; }
            if (kz > d.Pz / 2) { kzr = d.Pz - kz; fl |= 4; }
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const int kx = j + r * nb;
                int flr = fl, kxr = kx;
                if (kx > Px / 2) { kxr = Px - kx; flr |= 1; }
                const cplx<R>* gp = Gh + (((int64_t)kxr * Hy + kyr) * Hz + kzr) * NC;
                cplx<R> gv[NC];
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    cplx<R> v = gp[c];
                    if (__popc(comp_odd_mask(c) & flr) & 1) v = -v;
                    gv[c] = v;
                }
                // symmetric S = [[xx,xy,xz],[xy,yy,yz],[xz,yz,zz]]
                const cplx<R> ex = h[0][r], ey = h[1][r], ez = h[2][r];
                cplx<R> ox = gv[0] * ex + gv[3] * ey + gv[4] * ez;
                cplx<R> oy = gv[3] * ex + gv[1] * ey + gv[5] * ez;
                cplx<R> oz = gv[4] * ex + gv[5] * ey + gv[2] * ez;
                if constexpr (F == 6) {
                    // A = [[0,a,b],[-a,0,c],[-b,-c,0]] with a=xy, b=xz, c=yz of M01
                    const cplx<R> hx = h[3][r], hy = h[4][r], hz = h[5][r];
                    const cplx<R> a = gv[6], bb = gv[7], cc = gv[8];
                    ox = ox + a * hy + bb * hz;
                    oy = oy - a * hx + cc * hz;
                    oz = oz - bb * hx - cc * hy;
                    // H' = -A E + S H
                    const cplx<R> px = gv[0] * hx + gv[3] * hy + gv[4] * hz - (a * ey + bb * ez);
                    const cplx<R> py = gv[3] * hx + gv[1] * hy + gv[5] * hz - (cc * ez - a * ex);
                    const cplx<R> pz = gv[4] * hx + gv[5] * hy + gv[2] * hz - (-bb * ex - cc * ey);
                    h[3][r] = px;
                    h[4][r] = py;
                    h[5][r] = pz;
                }
                h[0][r] = ox;
                h[1][r] = oy;
                h[2][r] = oz;
            }
        }
#pragma unroll
        for (int c = 0; c < F; ++c) {
            Dft<RL, R, true>::run(h[c]);
#pragma unroll
            for (int r = 0; r < RL; ++r) lds[(j * RL + r) * Tp + tq * F + c] = h[c][r];
        }
    }
    __syncthreads();
    run_stages<Budget<R>::MAXE, NT_PASS, R, true>(lds, tw, g, ap, 1, ap.nst);
    for (int e = threadIdx.x; e < d.nx * Tp; e += NT_PASS) {
        const int p = e / Tp, t = e % Tp;
        const int tq2 = t / F, c = t % F;
        if (tq2 < tqn) base[c * cstride + (int64_t)p * plane + tq2] = lds[p * Tp + t];
    }
}

inline int ap_last(const AxisPlan& ap) { return ap.nst ? ap.rad[ap.nst - 1] : 0; }

// BDDA_PASSES=lds forces the all-in-LDS kernels (cross-checks)
inline bool force_lds_passes()
{
    const char* v = getenv("BDDA_PASSES");
    return v && v[0] == 'l';
}

template <typename R, int F>
void apply_impl(MbtOperator* op, const cplx<R>* X, int64_t ldx, cplx<R>* Y, int64_t ldy, int s, const cplx<R>* diag,
                const cplx<R>* jpre, const int* c2o)
{
    const MbtDims& d = op->d;
    hipStream_t st = op->stream;
    const bool is64 = sizeof(R) == 8;
    const cplx<R>* gh = (const cplx<R>*)(is64 ? op->gh64 : op->gh32);
    if (!gh) throw HipError{hipErrorInvalidValue, "operator built without this precision"};
    const cplx<R>* tw[3];
    for (int a = 0; a < 3; ++a) tw[a] = (const cplx<R>*)(is64 ? op->tw64[a] : op->tw32[a]);
    const int Pzs = op->Pzs;
    const int64_t nxny = (int64_t)d.nx * d.ny;
    const int64_t nU = (int64_t)s * F * nxny * Pzs;
    const int64_t nV = (int64_t)s * F * d.nx * d.Py * Pzs;
    // (mbt_workspace_per_column is this per column)
    mbt_ensure_workspace(op, sizeof(cplx<R>) * (nU + nV));
    cplx<R>* U = (cplx<R>*)op->ws;
    cplx<R>* V = U + nU;
    // The planned kernels serve f = 3 and f = 6.  With six components between
    // the transforms the fused x kernel's f = 6 instances take 140-320 VGPRs
    // (one to three waves per SIMD), and still measured 3x faster than the
    // all-in-LDS x kernel: the reference's f = 6 workload on 192x192x225
    // (L = 4) solved in 0.184 s against 0.543 s in FP64.
    const bool planned = !force_lds_passes();
    // Occupied vectors: only the (ix, iy) lines that hold an occupied cell
    // are transformed along z.  The y inverse pass skips the U rows of the
    // other lines and the z inverse pass skips those lines, since nothing is
    // gathered from them.  The z forward pass skips their rows only when the
    // planned y forward pass runs with the line mask, which reads them as zero.
    const bool lined = c2o && op->lines;
    const bool ymask = planned && lined && lines_plan_ok<R>(d.Py, Pzs);
    const int* zl = ymask ? op->lines : nullptr;
    const bool tm = op->timing != 0;
    if (tm) BDDA_HIP(hipEventRecord(op->ev[0], st));
    // z forward + scatter: X -> U
    bool done = planned && launch_zfwd_plan<R, F>(X, ldx, c2o, U, d, Pzs, tw[2], s, st, zl, op->nlines);
    if (!done) {
        const int P = d.Pz;
        if ((int64_t)P * F > (int64_t)Budget<R>::MAXE * NT_PASS) throw HipError{hipErrorInvalidValue, "Pz too long"};
        const int TL = std::max(1, Budget<R>::MAXE * NT_PASS / (P * F));
        const int64_t ntiles = ceil_div(nxny, TL);
        const size_t shm = sizeof(cplx<R>) * (P + (size_t)P * TL * F);
        hipLaunchKernelGGL((k_zfwd_scatter<R, F>), dim3((unsigned)(ntiles * s)), dim3(NT_PASS), shm, st, X, ldx, c2o,
                           U, d, Pzs, op->ax[2], tw[2], TL, ntiles);
        BDDA_HIP(hipGetLastError());
    }
    if (tm) BDDA_HIP(hipEventRecord(op->ev[1], st));
    // y forward: U[outer = (rhs F + c) nx + ix][ny][Pzs] -> V[outer][Py][Pzs]
    const int64_t outer = (int64_t)s * F * d.nx;
    done = planned && launch_lines_plan<R, false>(U, V, d.ny, d.Py, d.Pz, Pzs, outer, d.Py, tw[1], st,
                                                      ymask ? op->lmask : nullptr, d.nx);
    if (!done) launch_lines<R, false>(U, V, d.ny, d.Py, d.Pz, Pzs, outer, op->ax[1], tw[1], st);
    if (tm) BDDA_HIP(hipEventRecord(op->ev[2], st));
    // x fused, in place on V
    done = false;
    done = planned && launch_xfused_plan<R, F>(V, gh, d, Pzs, tw[0], s, st);
    if (!done) {
        const int P = d.Px;
        const int Tq = Budget<R>::MAXE * NT_PASS / (P * F);
        if (Tq < 1 || (P / 8) * Tq > NT_PASS || P % 8 != 0 || ap_last(op->ax[0]) != 8)
            throw HipError{hipErrorInvalidValue, "no fused x kernel for this padded size"};
        const int64_t plane = (int64_t)d.Py * Pzs;
        const int64_t ntiles = ceil_div(plane, Tq);
        const size_t shm = sizeof(cplx<R>) * (P + (size_t)P * Tq * F);
        hipLaunchKernelGGL((k_xfused<R, F>), dim3((unsigned)(ntiles * s)), dim3(NT_PASS), shm, st, V, gh, d, Pzs,
                           op->ax[0], tw[0], Tq, ntiles);
        BDDA_HIP(hipGetLastError());
    }
    if (tm) BDDA_HIP(hipEventRecord(op->ev[3], st));
    // y inverse: V -> U keeping ny
    done = planned && launch_lines_plan<R, true>(V, U, d.Py, d.ny, d.Pz, Pzs, outer, d.Py, tw[1], st,
                                                     lined ? op->lmask : nullptr, d.nx);
    if (!done) launch_lines<R, true>(V, U, d.Py, d.ny, d.Pz, Pzs, outer, op->ax[1], tw[1], st);
    if (tm) BDDA_HIP(hipEventRecord(op->ev[4], st));
    // z inverse + gather: U -> Y
    done = planned && launch_zinv_plan<R, F>(U, c2o, X, ldx, diag, jpre, Y, ldy, d, Pzs, tw[2], s, st,
                                                 lined ? op->lines : nullptr, op->nlines);
    if (!done) {
        const int P = d.Pz;
        const int TL = std::max(1, Budget<R>::MAXE * NT_PASS / (P * F));
        const int64_t ntiles = ceil_div(nxny, TL);
        const size_t shm = sizeof(cplx<R>) * (P + (size_t)P * TL * F);
        hipLaunchKernelGGL((k_zinv_gather<R, F>), dim3((unsigned)(ntiles * s)), dim3(NT_PASS), shm, st, U, c2o, X,
                           ldx, diag, jpre, Y, ldy, d, Pzs, op->ax[2], tw[2], TL, ntiles);
        BDDA_HIP(hipGetLastError());
    }
    if (tm) {
        BDDA_HIP(hipEventRecord(op->ev[5], st));
        BDDA_HIP(hipEventSynchronize(op->ev[5]));
        for (int i = 0; i < 5; ++i) BDDA_HIP(hipEventElapsedTime(&op->pass_ms[i], op->ev[i], op->ev[i + 1]));
    }
}

template <typename R>
void mbt_apply_precision(MbtOperator* op, const cplx<R>* X, int64_t ldx, cplx<R>* Y, int64_t ldy, int s,
                         const cplx<R>* diag, const cplx<R>* jpre, const int* c2o)
{
    // columns in chunks whose pass workspace fits the operator's limit
    const size_t per_col = mbt_workspace_per_column(op->d, op->Pzs, sizeof(cplx<R>));
    int chunk = s;
    if (op->ws_limit > 0) chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)s, op->ws_limit / per_col));
    for (int j0 = 0; j0 < s; j0 += chunk) {
        const int sc = std::min(chunk, s - j0);
        if (op->d.f == 3) apply_impl<R, 3>(op, X + j0 * ldx, ldx, Y + j0 * ldy, ldy, sc, diag, jpre, c2o);
        else apply_impl<R, 6>(op, X + j0 * ldx, ldx, Y + j0 * ldy, ldy, sc, diag, jpre, c2o);
    }
}

}  // namespace bdda
