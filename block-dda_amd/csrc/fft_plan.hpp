// fft_plan.hpp -- compile-time-planned register-resident Stockham FFT.
//
// A work-group of NT threads transforms NQ lines (length P) of NC
// components; NQ indexes the contiguous axis in HBM, so a
// wave reading butterfly inputs x[j + r P/RAD][q][c] touches consecutive
// addresses.  Data stays in registers: stage-0 inputs come straight from
// the caller's loader, twiddles and the radix-RAD DFT run in registers, and
// only the exchange between two stages goes through LDS (NCX components per
// round).  The last stage's outputs sit at positions j + r P/RAD, the same
// pattern as stage 0's inputs, so they go straight to the caller's storer.
//
// Stockham stage (L = product of earlier radices), butterfly j, k = j mod L:
//   in x[j + r P/RAD] * w_{L RAD}^{r k},   out y[(j - k) RAD + k + r L].
// Butterfly b = j NQ + q (q fastest); thread t owns b = t + i NT.
//
// The radix sequence is a template parameter: every stage's radix, L, the
// number of butterflies per thread and all strides are compile-time
// constants, so the register arrays have exact sizes, every index folds and
// nothing is placed in scratch memory.  Kernels for the padded sizes in the
// plan table (mbt_plans.hpp) are built on this engine; other sizes use the
// all-in-LDS engine of fft_engine.hpp.
#pragma once
#include <type_traits>
#include "fft_engine.hpp"

namespace bdda {

template <int... RS>
struct Radices {
    static constexpr int N = sizeof...(RS);
    static constexpr int P = (RS * ... * 1);
    static constexpr int r[N] = {RS...};
};

template <int I>
using ic = std::integral_constant<int, I>;

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (B < E) {
        f(ic<B>{});
        static_for<B + 1, E>(f);
    }
}

// Odd radix RAD on inputs of which only v[0..NZ) can be nonzero (the first
// stage of a forward transform of a zero-padded line): the pair sums and
// differences of DftOdd lose their zero halves at compile time instead of
// adding +0.0, which the compiler may not drop (x + 0.0 differs from x for
// x = -0.0).
template <int RAD, typename R, bool INV, int NZ>
__device__ __forceinline__ void dft_odd_nz(cplx<R>* v)
{
    static constexpr TrigTab<RAD> tab{};
    constexpr int H = (RAD - 1) / 2;
    cplx<R> a[H], b[H];
    static_for<1, H + 1>([&](auto JJ) {
        constexpr int j = decltype(JJ)::value;
        if constexpr (j < NZ && RAD - j < NZ) {
            a[j - 1] = v[j] + v[RAD - j];
            b[j - 1] = v[j] - v[RAD - j];
        } else if constexpr (j < NZ) {
            a[j - 1] = v[j];
            b[j - 1] = v[j];
        } else if constexpr (RAD - j < NZ) {
            a[j - 1] = v[RAD - j];
            b[j - 1] = -v[RAD - j];
        }
    });
    const cplx<R> x0 = v[0];
    cplx<R> y0 = x0;
    static_for<1, H + 1>([&](auto JJ) {
        constexpr int j = decltype(JJ)::value;
        if constexpr (j < NZ || RAD - j < NZ) y0 = y0 + a[j - 1];
    });
    static_for<1, H + 1>([&](auto KK) {
        constexpr int k = decltype(KK)::value;
        cplx<R> A = x0, B = czero<R>();
        bool first = true;
        static_for<1, H + 1>([&](auto JJ) {
            constexpr int j = decltype(JJ)::value;
            if constexpr (j < NZ || RAD - j < NZ) {
                const R c = R(tab.c[(j * k) % RAD]);
                const R sn = R(tab.s[(j * k) % RAD]);
                A = A + c * a[j - 1];
                B = first ? sn * b[j - 1] : B + sn * b[j - 1];
                first = false;
            }
        });
        const cplx<R> iB = mul_pi(B);
        v[k] = INV ? (A + iB) : (A - iB);
        v[RAD - k] = INV ? (A - iB) : (A + iB);
    });
    v[0] = y0;
}

// R: real type, PL: Radices<...>, NT threads, NQ lines (power of two), NC
// components, NCX components per LDS exchange round.  Methods take INV: the
// inverse transform uses the radices in reverse order, so the forward's last
// stage and the inverse's first stage share radix and butterfly layout.  DB:
// the exchange uses two LDS buffers of NCX * P * NQ elements (lds points at
// the first, the second follows it) and one barrier per round.  An instance
// used for several transforms in a row needs an even number of rounds per
// forward + inverse pair, which 2 (NS - 1) rounds of each component set are.
template <typename R, class PL, int NT, int NQ, int NC, int NCX, bool DB = false>
struct PlanFFT {
    static constexpr int P = PL::P;
    static constexpr int NS = PL::N;
    template <int S, bool INV>
    static constexpr int rad() { return INV ? PL::r[NS - 1 - S] : PL::r[S]; }
    template <int S, bool INV>
    static constexpr int Lof()
    {
        int L = 1;
        for (int s = 0; s < S; ++s) L *= INV ? PL::r[NS - 1 - s] : PL::r[s];
        return L;
    }
    template <int S, bool INV>
    static constexpr int mb() { return (P / rad<S, INV>() * NQ + NT - 1) / NT; }
    static constexpr int emax()
    {
        int m = 0;
        for (int s = 0; s < NS; ++s) {
            const int rr = PL::r[s];
            const int v = ((P / rr * NQ + NT - 1) / NT) * rr;
            m = v > m ? v : m;
        }
        return m;
    }
    static constexpr int E = emax();
    // butterfly b -> (line q, butterfly j); NQ need not be a power of two (a
    // line count that fills whole waves may be odd), and the split is by a
    // compile-time constant
    __device__ __forceinline__ static int qof(int b) { return (int)((unsigned)b % (unsigned)NQ); }
    __device__ __forceinline__ static int jof(int b) { return (int)((unsigned)b / (unsigned)NQ); }

    cplx<R> v[NC][E];
    const cplx<R>* tw;  // LDS twiddles, tw[i] = exp(-2 pi i i / P)
    cplx<R>* lds;       // NCX * P * NQ exchange elements

    __device__ __forceinline__ PlanFFT(const cplx<R>* tw_, cplx<R>* lds_) : tw(tw_), lds(lds_) {}

    // stage-0 inputs: ld(j, RB, q, c) returns element p = j + RB of line q,
    // component c, where RB (an integral_constant, r P/RAD) is the
    // compile-time part of p, so that callers can fold it into a uniform
    // base address.  HALF: the input is a zero-padded line of
    // at most (P+1)/2 values (P >= 2n - 1), so butterfly inputs at
    // p = j + r P/RAD with r P/RAD >= (P+1)/2 are zero whatever j is; they are
    // set at compile time and the arithmetic on them folds away.
    template <bool INV, bool HALF = false, class LD>
    __device__ __forceinline__ void load(LD&& ld)
    {
        constexpr int RAD = rad<0, INV>(), nb = P / RAD, total = nb * NQ;
        static_for<0, mb<0, INV>()>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const int b = threadIdx.x + i * NT;
            if (total % NT == 0 || b < total) {
                const int q = qof(b), j = jof(b);
                static_for<0, RAD>([&](auto RR) {
                    constexpr int r = decltype(RR)::value;
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        if constexpr (HALF && r * nb >= (P + 1) / 2) v[c][i * RAD + r] = czero<R>();
                        else v[c][i * RAD + r] = ld(j, ic<r * nb>{}, q, c);
                    }
                });
            }
        });
    }

    // last-stage outputs: st(j, RB, q, c, value) for p = j + RB.  HALF: only outputs p < (P+1)/2
    // are wanted (the inverse transform keeps n <= (P+1)/2 of them); the
    // others are never stored and their arithmetic is dead code.
    template <bool INV, bool HALF = false, class ST>
    __device__ __forceinline__ void store(ST&& st)
    {
        constexpr int S = NS - 1, RAD = rad<S, INV>(), nb = P / RAD, total = nb * NQ;
        static_for<0, mb<S, INV>()>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const int b = threadIdx.x + i * NT;
            if (total % NT == 0 || b < total) {
                const int q = qof(b), j = jof(b);
                static_for<0, RAD>([&](auto RR) {
                    constexpr int r = decltype(RR)::value;
                    if constexpr (!(HALF && r * nb >= (P + 1) / 2)) {
#pragma unroll
                        for (int c = 0; c < NC; ++c) st(j, ic<r * nb>{}, q, c, v[c][i * RAD + r]);
                    }
                });
            }
        });
    }

    // last-stage outputs in two phases: first key(j, RB, q) for every output
    // position p = j + RB of the thread (a negative key skips it), then
    // st(key, c, value).  A scatter whose addresses come from a lookup table
    // issues all of its lookups before any dependent access, so their
    // latencies overlap instead of adding up.  HALF as in store().
    template <bool INV, bool HALF, class KEY, class ST>
    __device__ __forceinline__ void store_keyed(KEY&& key, ST&& st)
    {
        constexpr int S = NS - 1, RAD = rad<S, INV>(), nb = P / RAD, total = nb * NQ, M = mb<S, INV>();
        int kv[M][RAD];
        static_for<0, M>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const int b = threadIdx.x + i * NT;
            const bool ok = total % NT == 0 || b < total;
            const int q = qof(b), j = jof(b);
            static_for<0, RAD>([&](auto RR) {
                constexpr int r = decltype(RR)::value;
                if constexpr (!(HALF && r * nb >= (P + 1) / 2)) kv[i][r] = ok ? key(j, ic<r * nb>{}, q) : -1;
            });
        });
        static_for<0, M>([&](auto I) {
            constexpr int i = decltype(I)::value;
            static_for<0, RAD>([&](auto RR) {
                constexpr int r = decltype(RR)::value;
                if constexpr (!(HALF && r * nb >= (P + 1) / 2)) {
                    if (kv[i][r] >= 0) {
#pragma unroll
                        for (int c = 0; c < NC; ++c) st(kv[i][r], c, v[c][i * RAD + r]);
                    }
                }
            });
        });
    }

    // forward last-stage slots: fn(p, q, cplx<R>* vals[NC]) may rewrite vals
    template <class FN>
    __device__ __forceinline__ void visit(FN&& fn)
    {
        constexpr int S = NS - 1, RAD = rad<S, false>(), nb = P / RAD, total = nb * NQ;
        static_for<0, mb<S, false>()>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const int b = threadIdx.x + i * NT;
            if (total % NT == 0 || b < total) {
                const int q = qof(b), j = jof(b);
#pragma unroll
                for (int r = 0; r < RAD; ++r) {
                    cplx<R> x[NC];
#pragma unroll
                    for (int c = 0; c < NC; ++c) x[c] = v[c][i * RAD + r];
                    fn(j + r * nb, q, x);
#pragma unroll
                    for (int c = 0; c < NC; ++c) v[c][i * RAD + r] = x[c];
                }
            }
        });
    }

    // NZ: stage-0 butterfly inputs r >= NZ are known zero (HALF loads)
    template <int S, bool INV, int NZ = 64>
    __device__ __forceinline__ void compute()
    {
        constexpr int RAD = rad<S, INV>(), L = Lof<S, INV>(), nb = P / RAD, total = nb * NQ;
        constexpr int tstep = P / (L * RAD);
        static_for<0, mb<S, INV>()>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const int b = threadIdx.x + i * NT;
            if (total % NT == 0 || b < total) {
                if constexpr (L > 1) {
                    const int j = jof(b);
                    const int kt = (j % L) * tstep;
#pragma unroll
                    for (int r = 1; r < RAD; ++r) {
                        cplx<R> w = tw[r * kt];
                        if (INV) w.y = -w.y;
#pragma unroll
                        for (int c = 0; c < NC; ++c) v[c][i * RAD + r] = v[c][i * RAD + r] * w;
                    }
                }
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    if constexpr (NZ < RAD && RAD % 2 == 1) dft_odd_nz<RAD, R, INV, NZ>(&v[c][i * RAD]);
                    else Dft<RAD, R, INV>::run(&v[c][i * RAD]);
                }
            }
        });
    }

    // stage S outputs -> LDS -> stage S+1 inputs, NCX components per round
    template <int S, bool INV>
    __device__ __forceinline__ void exchange()
    {
        constexpr int RO = rad<S, INV>(), L = Lof<S, INV>(), nbo = P / RO, to = nbo * NQ;
        constexpr int RI = rad<S + 1, INV>(), nbi = P / RI, ti = nbi * NQ;
        constexpr int PQ = P * NQ;
        constexpr int NR = (NC + NCX - 1) / NCX;
        static_for<0, NR>([&](auto CB) {
            constexpr int c0 = decltype(CB)::value * NCX;
            // DB: rounds alternate between two buffers (round number counted
            // over the forward's exchanges, then the inverse's), so a round's
            // stores cannot meet the previous round's loads and one barrier
            // per round is enough
            constexpr int rnd = ((INV ? NS - 1 : 0) + S) * NR + decltype(CB)::value;
            cplx<R>* lds = this->lds + (DB && (rnd & 1) ? NCX * PQ : 0);
            static_for<0, mb<S, INV>()>([&](auto I) {
                constexpr int i = decltype(I)::value;
                const int b = threadIdx.x + i * NT;
                if (to % NT == 0 || b < to) {
                    const int q = qof(b), j = jof(b);
                    const int k = j % L;
                    const int base = ((j - k) * RO + k) * NQ + q;
#pragma unroll
                    for (int r = 0; r < RO; ++r)
#pragma unroll
                        for (int cc = 0; cc < NCX; ++cc)
                            if (c0 + cc < NC) lds[cc * PQ + base + r * L * NQ] = v[c0 + cc][i * RO + r];
                }
            });
            __syncthreads();
            static_for<0, mb<S + 1, INV>()>([&](auto I) {
                constexpr int i = decltype(I)::value;
                const int b = threadIdx.x + i * NT;
                if (ti % NT == 0 || b < ti) {
#pragma unroll
                    for (int r = 0; r < RI; ++r)
#pragma unroll
                        for (int cc = 0; cc < NCX; ++cc)
                            if (c0 + cc < NC) v[c0 + cc][i * RI + r] = lds[cc * PQ + b + r * nbi * NQ];
                }
            });
            if constexpr (!DB) __syncthreads();
        });
    }

    // HALF: the stage-0 inputs were loaded by load<INV, true>
    template <bool INV, bool HALF = false>
    __device__ __forceinline__ void run()
    {
        constexpr int RAD0 = rad<0, INV>(), nb0 = P / RAD0;
        constexpr int NZ0 = HALF ? ((P + 1) / 2 + nb0 - 1) / nb0 : 64;  // r nb0 < (P+1)/2
        static_for<0, NS>([&](auto SI) {
            constexpr int S = decltype(SI)::value;
            compute<S, INV, S == 0 ? NZ0 : 64>();
            if constexpr (S + 1 < NS) exchange<S, INV>();
        });
    }
};

}  // namespace bdda
