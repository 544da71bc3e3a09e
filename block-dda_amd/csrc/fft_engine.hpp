// fft_engine.hpp -- LDS-staged Stockham FFT building blocks for gfx950.
//
// A work-group holds T lines of one axis (length P) in LDS; element (p, t)
// lives at lds[p * sP + t * sT].  Each Stockham stage of radix RAD reads
// RAD values per butterfly into registers (stride P/RAD), applies the stage
// twiddles, runs a register DFT, and writes the results back in autosort
// order, so the last stage leaves the transform in natural order:
//
//   stage with L = product of the radices already applied,
//   butterfly j (0 <= j < P/RAD), k = j mod L:
//     in  : x[j + r P/RAD] * w_{L RAD}^{r k}          r = 0..RAD-1
//     out : y[(j / L) L RAD + k + r L]
//
// Butterflies of all T lines are spread over the NT threads of the group;
// every thread keeps at most ceil(MAXE / RAD) butterflies in registers,
// where MAXE >= ceil(P T / NT) is a compile-time bound.  One LDS buffer is
// enough: all reads of a stage finish before the barrier, all writes after.
//
// Twiddle table: tw[q] = exp(-2 pi i q / P), q < P, held in LDS.
#pragma once
#include "bdda_common.hpp"

namespace bdda {

// ------------------------------------------------------------------ DFTs
template <typename R, bool INV>
__device__ __forceinline__ cplx<R> rot_mi(cplx<R> a)  // a * (-i) forward, a * (+i) inverse
{
    return INV ? mul_pi(a) : mul_mi(a);
}

template <int RAD, typename R, bool INV>
struct Dft;

template <typename R, bool INV>
struct Dft<2, R, INV> {
    __device__ __forceinline__ static void run(cplx<R>* v)
    {
        const cplx<R> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};

template <typename R, bool INV>
struct Dft<4, R, INV> {
    __device__ __forceinline__ static void run(cplx<R>* v)
    {
        const cplx<R> t0 = v[0] + v[2], t1 = v[0] - v[2];
        const cplx<R> t2 = v[1] + v[3], t3 = rot_mi<R, INV>(v[1] - v[3]);
        v[0] = t0 + t2;
        v[2] = t0 - t2;
        v[1] = t1 + t3;
        v[3] = t1 - t3;
    }
};

template <typename R, bool INV>
struct Dft<8, R, INV> {
    __device__ __forceinline__ static void run(cplx<R>* v)
    {
        cplx<R> e[4] = {v[0], v[2], v[4], v[6]};
        cplx<R> o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, R, INV>::run(e);
        Dft<4, R, INV>::run(o);
        constexpr R h = R(0.70710678118654752440084436210484904);
        // w^1 = (1 -/+ i)/sqrt2, w^2 = -/+ i, w^3 = (-1 -/+ i)/sqrt2
        const cplx<R> o1 = INV ? mk(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                               : mk(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
        const cplx<R> o2 = rot_mi<R, INV>(o[2]);
        const cplx<R> o3 = INV ? mk(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                               : mk(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
        v[0] = e[0] + o[0];
        v[4] = e[0] - o[0];
        v[1] = e[1] + o1;
        v[5] = e[1] - o1;
        v[2] = e[2] + o2;
        v[6] = e[2] - o2;
        v[3] = e[3] + o3;
        v[7] = e[3] - o3;
    }
};

// cos/sin(2 pi i / P), i < P, evaluated at compile time
template <int P>
struct TrigTab {
    double c[P], s[P];
    constexpr TrigTab() : c(), s()
    {
        for (int i = 0; i < P; ++i) {
            c[i] = ct_cos(2.0 * kPi * i / P);
            s[i] = ct_sin(2.0 * kPi * i / P);
        }
    }
};

// odd prime radix p: pairs a_j = x_j + x_{p-j}, b_j = x_j - x_{p-j}
template <int P, typename R, bool INV>
struct DftOdd {
    __device__ __forceinline__ static void run(cplx<R>* v)
    {
        static constexpr TrigTab<P> tab{};
        constexpr int H = (P - 1) / 2;
        cplx<R> a[H], b[H];
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            a[j - 1] = v[j] + v[P - j];
            b[j - 1] = v[j] - v[P - j];
        }
        const cplx<R> x0 = v[0];
        cplx<R> y0 = x0;
#pragma unroll
        for (int j = 0; j < H; ++j) y0 = y0 + a[j];
#pragma unroll
        for (int k = 1; k <= H; ++k) {
            cplx<R> A = x0, B = czero<R>();
#pragma unroll
            for (int j = 1; j <= H; ++j) {
                const R c = R(tab.c[(j * k) % P]);
                const R s = R(tab.s[(j * k) % P]);
                A = A + c * a[j - 1];
                B = B + s * b[j - 1];
            }
            // forward: y_k = A - iB, y_{p-k} = A + iB
            const cplx<R> iB = mul_pi(B);
            v[k] = INV ? (A + iB) : (A - iB);
            v[P - k] = INV ? (A - iB) : (A + iB);
        }
        v[0] = y0;
    }
};
template <typename R, bool INV>
struct Dft<3, R, INV> : DftOdd<3, R, INV> {};
template <typename R, bool INV>
struct Dft<5, R, INV> : DftOdd<5, R, INV> {};
template <typename R, bool INV>
struct Dft<7, R, INV> : DftOdd<7, R, INV> {};

// radix 6 = 2 x 3 (Good-Thomas free, with the w_6 twiddles inside)
template <typename R, bool INV>
struct Dft<6, R, INV> {
    __device__ __forceinline__ static void run(cplx<R>* v)
    {
        // y_k = E_k' + w6^k O_k' where E = DFT3(x0,x2,x4), O = DFT3(x1,x3,x5)
        cplx<R> e[3] = {v[0], v[2], v[4]};
        cplx<R> o[3] = {v[1], v[3], v[5]};
        Dft<3, R, INV>::run(e);
        Dft<3, R, INV>::run(o);
        constexpr R c1 = R(0.5), s1 = R(0.86602540378443864676372317075293618);
        // w6^1 = (1/2 -/+ i sqrt3/2), w6^2 = (-1/2 -/+ i sqrt3/2)
        const cplx<R> w1 = INV ? mk(c1, s1) : mk(c1, -s1);
        const cplx<R> w2 = INV ? mk(-c1, s1) : mk(-c1, -s1);
        const cplx<R> o1 = w1 * o[1];
        const cplx<R> o2 = w2 * o[2];
        v[0] = e[0] + o[0];
        v[3] = e[0] - o[0];
        v[1] = e[1] + o1;
        v[4] = e[1] - o1;
        v[2] = e[2] + o2;
        v[5] = e[2] - o2;
    }
};

// ------------------------------------------------------------ stage geometry
struct StageGeom {
    int P;      // line length
    int T;      // lines in the group
    int sP, sT; // LDS strides
    bool tfast; // butterfly index: t fastest (for [p][t] layouts) or j fastest
};

template <int RAD>
__device__ __forceinline__ void bfly_coords(const StageGeom& g, int b, int& t, int& j)
{
    constexpr int dummy = RAD;
    (void)dummy;
    const int nb = g.P / RAD;
    if (g.tfast) {
        t = b % g.T;
        j = b / g.T;
    } else {
        j = b % nb;
        t = b / nb;
    }
}

// load + twiddle one stage's butterflies into v
template <int RAD, int MB, int NT, typename R, bool INV>
__device__ __forceinline__ void stage_load(const cplx<R>* lds, const cplx<R>* tw, const StageGeom& g,
                                           int L, cplx<R> (&v)[MB][RAD])
{
    const int nb = g.P / RAD;
    const int total = nb * g.T;
    const int tstep = g.P / (L * RAD);
#pragma unroll
    for (int i = 0; i < MB; ++i) {
        const int b = threadIdx.x + i * NT;
        if (b < total) {
            int t, j;
            bfly_coords<RAD>(g, b, t, j);
            const int k = j % L;
#pragma unroll
            for (int r = 0; r < RAD; ++r) {
                cplx<R> x = lds[(j + r * nb) * g.sP + t * g.sT];
                if (r > 0 && k > 0) {
                    cplx<R> w = tw[r * k * tstep];
                    if (INV) w.y = -w.y;
                    x = x * w;
                }
                v[i][r] = x;
            }
        }
    }
}

template <int RAD, int MB, int NT, typename R, bool INV>
__device__ __forceinline__ void stage_compute(cplx<R> (&v)[MB][RAD], const StageGeom& g)
{
    const int total = (g.P / RAD) * g.T;
#pragma unroll
    for (int i = 0; i < MB; ++i) {
        const int b = threadIdx.x + i * NT;
        if (b < total) Dft<RAD, R, INV>::run(v[i]);
    }
}

template <int RAD, int MB, int NT, typename R>
__device__ __forceinline__ void stage_store(cplx<R>* lds, const StageGeom& g, int L,
                                            const cplx<R> (&v)[MB][RAD])
{
    const int nb = g.P / RAD;
    const int total = nb * g.T;
#pragma unroll
    for (int i = 0; i < MB; ++i) {
        const int b = threadIdx.x + i * NT;
        if (b < total) {
            int t, j;
            bfly_coords<RAD>(g, b, t, j);
            const int k = j % L;
            const int base = (j / L) * L * RAD + k;
#pragma unroll
            for (int r = 0; r < RAD; ++r) lds[(base + r * L) * g.sP + t * g.sT] = v[i][r];
        }
    }
}

// one full stage in LDS (load, barrier, compute, store, barrier)
template <int RAD, int MAXE, int NT, typename R, bool INV>
__device__ __forceinline__ void stage_lds(cplx<R>* lds, const cplx<R>* tw, const StageGeom& g, int L)
{
    constexpr int MB = (MAXE + RAD - 1) / RAD;
    cplx<R> v[MB][RAD];
    stage_load<RAD, MB, NT, R, INV>(lds, tw, g, L, v);
    __syncthreads();
    stage_compute<RAD, MB, NT, R, INV>(v, g);
    stage_store<RAD, MB, NT, R>(lds, g, L, v);
    __syncthreads();
}

// Radix plan of one axis, passed by value to kernels.
struct AxisPlan {
    int P;
    int nst;
    int rad[8];
};

// run stages [s0, s1) of the plan (forward order rad[0..]) or, for the
// inverse, of the reversed plan; L accumulates from the first stage.
template <int MAXE, int NT, typename R, bool INV>
__device__ __forceinline__ void run_stages(cplx<R>* lds, const cplx<R>* tw, const StageGeom& g,
                                           const AxisPlan& ap, int s0, int s1)
{
    int L = 1;
    for (int s = 0; s < s0; ++s) L *= INV ? ap.rad[ap.nst - 1 - s] : ap.rad[s];
    for (int s = s0; s < s1; ++s) {
        const int rad = INV ? ap.rad[ap.nst - 1 - s] : ap.rad[s];
        switch (rad) {
        case 2: stage_lds<2, MAXE, NT, R, INV>(lds, tw, g, L); break;
        case 3: stage_lds<3, MAXE, NT, R, INV>(lds, tw, g, L); break;
        case 4: stage_lds<4, MAXE, NT, R, INV>(lds, tw, g, L); break;
        case 5: stage_lds<5, MAXE, NT, R, INV>(lds, tw, g, L); break;
        case 6: stage_lds<6, MAXE, NT, R, INV>(lds, tw, g, L); break;
        case 7: stage_lds<7, MAXE, NT, R, INV>(lds, tw, g, L); break;
        case 8: stage_lds<8, MAXE, NT, R, INV>(lds, tw, g, L); break;
        default: break;
        }
        L *= rad;
    }
}

}  // namespace bdda
