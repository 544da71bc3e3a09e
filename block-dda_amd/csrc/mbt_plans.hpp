// mbt_plans.hpp -- pass kernels on the compile-time-planned FFT engine and
// the table of padded sizes that have them.
//
// Every padded size P in BDDA_PLAN_TABLE has a fixed radix sequence; the
// four passes of an MBT block matvec (z forward + scatter, y forward/inverse,
// fused x, z inverse + gather) are instantiated for it in each precision.
// Sizes outside the table run on the all-in-LDS kernels (mbt_passes.hpp).
#pragma once
#include "fft_plan.hpp"
#include "mbt.hpp"

namespace bdda {

// P, then its radices (forward order)
// (every 2n - 1 up to 1023 finds a planned size at most 12.5 % above the
// smallest 2/3/5/7-smooth one: sizes without a plan run the all-in-LDS
// kernels, measured 3.4x slower at 280^3)
#define BDDA_PLAN_TABLE_0(X) \
    X(32, 8, 4)              \
    X(40, 8, 5)              \
    X(48, 8, 6)              \
    X(56, 8, 7)              \
    X(64, 8, 8)              \
    X(72, 8, 3, 3)           \
    X(80, 8, 5, 2)           \
    X(96, 8, 4, 3)           \
    X(112, 8, 7, 2)          \
    X(120, 8, 5, 3)          \
    X(128, 8, 4, 4)          \
    X(160, 8, 5, 4)
#define BDDA_PLAN_TABLE_1(X) \
    X(168, 8, 7, 3)          \
    X(192, 8, 6, 4)          \
    X(200, 8, 5, 5)          \
    X(224, 8, 7, 4)          \
    X(240, 8, 6, 5)          \
    X(256, 8, 8, 4)          \
    X(280, 8, 7, 5)          \
    X(288, 8, 6, 6)          \
    X(320, 8, 8, 5)
#define BDDA_PLAN_TABLE_2(X) \
    X(336, 8, 7, 6)          \
    X(343, 7, 7, 7)          \
    X(384, 8, 8, 6)          \
    X(392, 8, 7, 7)          \
    X(448, 8, 8, 7)          \
    X(512, 8, 8, 8)
#define BDDA_PLAN_TABLE_3(X) \
    X(576, 8, 8, 3, 3)       \
    X(640, 8, 8, 5, 2)       \
    X(720, 8, 6, 5, 3)       \
    X(784, 8, 7, 7, 2)       \
    X(896, 8, 8, 7, 2)       \
    X(960, 8, 8, 5, 3)       \
    X(1024, 8, 8, 8, 2)
// (the four groups are compiled in separate translation units, mbt_k_*.hip)
#define BDDA_PLAN_TABLE(X) BDDA_PLAN_TABLE_0(X) BDDA_PLAN_TABLE_1(X) BDDA_PLAN_TABLE_2(X) BDDA_PLAN_TABLE_3(X)

// calls f(Radices<...>{}) for the plan of size P; false when P has none
template <class F>
inline bool with_plan(int P, F&& f)
{
    switch (P) {
#define BDDA_PLAN_CASE(p, ...) \
    case p: f(Radices<__VA_ARGS__>{}); return true;
        BDDA_PLAN_TABLE(BDDA_PLAN_CASE)
#undef BDDA_PLAN_CASE
    default: return false;
    }
}

inline bool has_plan(int P)
{
    return with_plan(P, [](auto) {});
}

// The fused x pass's plan: the table's, except that FP32 takes 192 as
// 8 x 8 x 3.  With eight lines per group that gives every stage one
// butterfly round of at most 8 values (256 threads) instead of two rounds of
// radix 6 (12 values, 192 threads), 141 VGPRs instead of 186, and measured
// 22 % less time for the pass.  FP64 keeps 8 x 6 x 4: its 8 x 8 x 3 kernel
// takes 256 VGPRs and measured 33 % slower.
template <typename R, class F>
inline bool with_xplan(int P, F&& f)
{
    if constexpr (sizeof(R) == 4) {
        if (P == 192) {
            f(Radices<8, 8, 3>{});
            return true;
        }
        if (P == 384) {  // (8 8 6: 12.8 ms at 384^3, s = 16; 6 4 4 4 on 384 threads: 11.2 ms)
            f(Radices<6, 4, 4, 4>{});
            return true;
        }
    }
    return with_plan(P, f);
}

// Threads per group for NQ lines of a plan.  A stage of radix r has
// P / r * NQ butterflies, spread over the group in rounds; the choice is the
// multiple of 64 that leaves the fewest lanes idle over all stages, with at
// most MB rounds per stage (each round's values live in registers), ties
// going to the size nearest TARGET.  343 = 7^3 with 16 lines (784
// butterflies per stage) gets 832 threads, one butterfly each, instead of
// four rounds of 256 with the last one a quarter full.
template <class PL, int NQ, int MB, int TARGET>
constexpr int plan_threads()
{
    int best = 0;
    long best_used = 0, best_slots = 1;
    for (int nt = 64; nt <= 1024; nt += 64) {
        long used = 0, slots = 0;
        bool fits = true;
        for (int s = 0; s < PL::N; ++s) {
            const long b = (long)(PL::P / PL::r[s]) * NQ;
            const long rounds = (b + nt - 1) / nt;
            fits = fits && rounds <= MB;
            used += b;
            slots += rounds * nt;
        }
        if (!fits) continue;
        const long lhs = used * best_slots, rhs = best_used * slots;
        const int dn = nt > TARGET ? nt - TARGET : TARGET - nt;
        const int db = best > TARGET ? best - TARGET : TARGET - best;
        if (best == 0 || lhs > rhs || (lhs == rhs && dn < db)) {
            best = nt;
            best_used = used;
            best_slots = slots;
        }
    }
    return best == 0 ? 1024 : best;
}

// Fraction of issued butterfly lanes that do work, at plan_threads' choice.
template <class PL, int NQ, int MB, int TARGET>
constexpr double plan_fill()
{
    const int nt = plan_threads<PL, NQ, MB, TARGET>();
    long used = 0, slots = 0;
    for (int s = 0; s < PL::N; ++s) {
        const long b = (long)(PL::P / PL::r[s]) * NQ;
        used += b;
        slots += ((b + nt - 1) / nt) * nt;
    }
    return (double)used / (double)slots;
}

// Lines per group: the base count, or one more when that fills the waves
// better by over 5 % and every stage still takes one round (so a thread keeps
// one line throughout).  343 = 7^3 has 49 butterflies per line and stage:
// 4 lines fill 3.06 of 4 waves (76.6 % of the lanes work), 5 lines 3.83
// (95.7 %).  Used by the z inverse pass, whose lines are whole rows.
template <class PL, int BASE, int MB, int TARGET>
constexpr int plan_lines()
{
    constexpr int alt = BASE + 1;
    constexpr int nt = plan_threads<PL, alt, MB, TARGET>();
    bool one_round = true;
    for (int s = 0; s < PL::N; ++s) one_round = one_round && (long)(PL::P / PL::r[s]) * alt <= nt;
    return (one_round && plan_fill<PL, alt, MB, TARGET>() > plan_fill<PL, BASE, MB, TARGET>() + 0.05) ? alt : BASE;
}

// Work-group -> (tile, right-hand side).  Work-groups are dispatched round-
// robin over the 8 XCDs (group b runs on XCD b mod 8), each with its own L2.
// Runs of PAIR consecutive tiles are dealt to XCDs in the same round-robin,
// and on each XCD the s right-hand sides of a run go back to back, the PAIR
// tiles of one right-hand side next to each other.  So the per-tile data the
// right-hand sides share (G^, the occupancy map, DIAG) comes from HBM once
// per tile, and with PAIR = 2 two half-line row segments of adjacent tiles
// (64-byte runs of a 128-byte L2 line) are read through the same L2.  A
// bijection on [0, ntiles * s) for ntiles a multiple of PAIR.
template <int PAIR = 1>
__device__ __forceinline__ void xcd_tile_rhs(int64_t b, int64_t ntiles, int s, int64_t& tile, int& rhs)
{
    constexpr int NXCD = 8;
    const int64_t nrun = ntiles / PAIR;
    const int64_t per = (int64_t)PAIR * s;  // groups per run
    const int64_t full = (nrun / NXCD) * NXCD * per;
    int64_t run, w;
    if (b < full) {
        const int64_t local = b / NXCD;
        run = (local / per) * NXCD + b % NXCD;
        w = local % per;
    } else {
        const int64_t r = b - full;
        run = (nrun / NXCD) * NXCD + r / per;
        w = r % per;
    }
    rhs = (int)(w / PAIR);
    tile = run * PAIR + w % PAIR;
}

// base + a 32-bit byte offset: with a uniform base the compiler issues the
// access as one scalar-base load or store with a 32-bit lane offset, instead
// of 64-bit address arithmetic per element (the launchers check that the
// offsets fit)
template <typename T>
__device__ __forceinline__ T* at_b(T* base, uint32_t off_bytes)
{
    using C = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return reinterpret_cast<T*>(reinterpret_cast<C*>(base) + off_bytes);
}

// tile widths of the planned passes (row segments of 32 to 128 bytes)
template <typename R>
struct PlanTiles {
    static constexpr int NQ_Y = sizeof(R) == 4 ? 16 : 8;  // y pass: columns per group
    // x pass: kz per group (all F planes).  Eight FP32 / four FP64 lines
    // (64-byte segments) keep three or more groups per CU; sixteen lines
    // measured 11-15 % slower at 343^3 for lack of resident waves.
    static constexpr int NQ_X = sizeof(R) == 4 ? 8 : 4;
    // z pass: lines per group (the group's lines x F components x P sit in
    // LDS for the gather / scatter transposition)
    static constexpr int TL_Z = sizeof(R) == 4 ? 4 : 2;
    // z pass: components per LDS exchange round (fewer rounds with more,
    // more resident groups with fewer)
    static constexpr int NCX_Z = 1;
};

// LDS elements of the z passes beyond the twiddle table
template <class PL, int F, int TL, int NCX>
constexpr size_t zpass_lds_elems()
{
    const size_t gather = (size_t)TL * F * ((PL::P + 1) / 2 + 1);  // forward gather rows (pitch nz | 1)
    const size_t xch = (size_t)NCX * TL * PL::P;                   // exchange / store staging
    return gather > xch ? gather : xch;
}

// ---------------------------------------------------------------- y pass
// in[outer][nlin][pitch] -> out[outer][nlout][pitch] over the first Zdim
// columns; a group owns NQ consecutive columns.
// lmask (optional): lmask[ix ny + iy] != 0 for the (ix, iy) lines that hold
// an occupied cell, where ix = outer mod nxm and iy runs over the short side
// (the forward pass's input, the inverse pass's output).  The forward pass
// takes the rows of the other lines as zero (they hold stale values); the
// inverse pass does not write them (nothing is gathered from them).
// FP64 forward passes skip reading those rows, which needs the flags, and so
// a barrier, before the loads; FP32 forward passes read every row and zero the
// stale inputs afterwards (measured at 343: FP64 4.27 against 4.57 ms,
// FP32 2.75 ms either way).
template <typename R, class PL, bool INV, int NT, int NQ>
__global__ __launch_bounds__(NT) void k_lines_plan(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out,
                                                   int nlin, int nlout, int64_t Zdim, int64_t pitch, int64_t ntiles,
                                                   const cplx<R>* __restrict__ twg,
                                                   const unsigned char* __restrict__ lmask, int nxm)
{
    constexpr int P = PL::P;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<R>* tw = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* lds = tw + P;
    const int64_t outer = blockIdx.x / ntiles;
    const int64_t z0 = (blockIdx.x % ntiles) * (int64_t)NQ;
    const int tz = (int)min((int64_t)NQ, Zdim - z0);
    if constexpr (INV) for (int i = threadIdx.x; i < P; i += NT) tw[i] = twg[i];
    const cplx<R>* src = in + (outer * nlin) * pitch + z0;
    cplx<R>* dst = out + (outer * nlout) * pitch + z0;
    const int nshort = INV ? nlout : nlin;
    const unsigned char* mrow = lmask ? lmask + (outer % nxm) * nshort : nullptr;
    // the mask row in LDS (the short side is at most (P + 1) / 2 lines)
    constexpr bool SKIP = !INV && sizeof(R) == 8;
    constexpr int MSKN = ((P + 1) / 2 + 15) / 16 * 16;
    __shared__ __align__(16) unsigned char msk[MSKN];
    auto stage_flags = [&]() {
        if (mrow)
            for (int i = threadIdx.x; i < nshort; i += NT) msk[i] = mrow[i];
    };
    if constexpr (INV) stage_flags();
    if constexpr (SKIP) {
        stage_flags();
        if (mrow) __syncthreads();
    }
    using FFT = PlanFFT<R, PL, NT, NQ, 1, 1>;
    FFT F(tw, lds);
    const uint32_t pitch32 = (uint32_t)pitch;
    auto lane_b = [&](int j, int q) { return ((uint32_t)j * pitch32 + (uint32_t)q) * (uint32_t)sizeof(cplx<R>); };
    // The flags of the positions p = j + r P / RAD a thread loads (skipping
    // forward pass) or stores (inverse pass) are read once into bit r of
    // `keep`, so that the row accesses do not each wait for an LDS read.  With
    // more than one butterfly per thread in that stage the flags are read per
    // access instead.
    constexpr int SK = INV ? PL::N - 1 : 0;
    constexpr int RADK = FFT::template rad<SK, INV>(), NBK = P / RADK;
    constexpr bool ONE = FFT::template mb<SK, INV>() == 1;
    unsigned keep = ~0u;
    auto read_flags = [&]() {  // (after a barrier that follows the staging)
        if constexpr (ONE) {
            if (mrow) {
                keep = 0u;
                const int j = FFT::jof(threadIdx.x);
                static_for<0, RADK>([&](auto RR) {
                    constexpr int r = decltype(RR)::value;
                    const int p = j + r * NBK;
                    // (read unconditionally, within the array, so the reads
                    // overlap; entries at p >= nshort are not used)
                    const unsigned f = msk[p < MSKN ? p : MSKN - 1];
                    keep |= (p < nshort && f) ? 1u << r : 0u;
                });
            }
        }
    };
    auto flagged = [&](int p, auto RB) {
        if constexpr (ONE) return ((keep >> (decltype(RB)::value / NBK)) & 1u) != 0u;
        else return !mrow || msk[p] != 0;
    };
    if constexpr (SKIP) read_flags();
    F.template load<INV, !INV>([&](int j, auto RB, int q, int) {
        const int p = j + decltype(RB)::value;
        return (p < nlin && q < tz && (!SKIP || flagged(p, RB)))
                   ? *at_b(src + (int64_t)decltype(RB)::value * pitch, lane_b(j, q))
                   : czero<R>();
    });
    if constexpr (!INV) {
        // forward: the twiddle table (and the FP32 flag row) staged behind the
        // row loads already issued, so that the loads do not wait for it
        // (y forward 2.78 -> 2.72 ms in FP32; the same order measured slower
        // in the y inverse, x and z passes)
        for (int i = threadIdx.x; i < P; i += NT) tw[i] = twg[i];
        if constexpr (!SKIP) stage_flags();
    }
    __syncthreads();
    if constexpr (INV) {
        read_flags();
    } else if (!SKIP && mrow) {
        // FP32 forward: the inputs from the rows of unlisted lines become zero
        constexpr int NB0 = P / RADK;
        static_for<0, FFT::template mb<0, false>()>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const int j = FFT::jof(threadIdx.x + i * NT);
            static_for<0, RADK>([&](auto RR) {
                constexpr int r = decltype(RR)::value;
                if constexpr (r * NB0 < (P + 1) / 2) {
                    const int p = j + r * NB0;
                    const unsigned f = msk[p < MSKN ? p : MSKN - 1];
                    if (p < nlin && !f) F.v[0][i * RADK + r] = czero<R>();
                }
            });
        });
    }
    F.template run<INV, !INV>();
    F.template store<INV, INV>([&](int j, auto RB, int q, int, cplx<R> v) {
        const int p = j + decltype(RB)::value;
        if (p < nlout && q < tz && (!INV || flagged(p, RB)))
            *at_b(dst + (int64_t)decltype(RB)::value * pitch, lane_b(j, q)) = v;
    });
}

// ---------------------------------------------------------------- x pass
// Minimum waves per SIMD asked of the compiler.
// - FP32 plans whose butterfly registers take at most 42 VGPRs (343 = 7^3:
//   one butterfly of 7 values x 3 components per thread) fit in 72 VGPRs, so
//   seven waves fit per SIMD (four groups of seven waves per CU instead of
//   three); larger plans would spill.
// - FP32 plans of one radix-8 butterfly (48 VGPRs of values) took 138-148
//   VGPRs left alone (336 = 8 7 6, 384 = 8 8 6, 512 = 8 8 7): three waves
//   per SIMD, one 8-wave group per CU.  Four waves (128 VGPRs, a few spilled
//   registers) give two groups: x pass 10.6 -> 8.0 ms at 336^3, 15.8 -> 13.3
//   ms at 384^3, 24.9 -> 23.5 ms at 448^3 (s = 16); five or six waves spill
//   tens of registers.
template <typename R, class PL, int F, int NT, int NQ>
constexpr int xfused_min_waves()
{
    constexpr int regs = PlanFFT<R, PL, NT, NQ, F, 1>::E * F * 2;
    return sizeof(R) != 4 ? 1 : regs <= 42 ? 7 : regs <= 48 ? 4 : 1;
}

// Right-hand sides per x group, one after the other.  A group reads the G^
// rows of its (ky, kz) tile for each right-hand side; at the large planned
// sizes the right-hand sides' groups of one tile no longer find those rows in
// L2 (512^3: 41.7 GB fetched against 26.6 GB algorithmic).  Two per group
// read them while they are still there: x pass at s = 16, FP32 512^3 28.0 ->
// 20.5 ms, FP64 512^3 56.4 -> 43.4 ms and 448^3 58.7 -> 50.0 ms.  Four were
// no better; FP32 at 448 spills with two (57 registers); 343 and below gain
// nothing.
template <typename R, class PL>
struct XRhsPerGroup {
    static constexpr int value = (sizeof(R) == 4 ? PL::P == 512 : PL::P >= 448) ? 2 : 1;
};

// Two LDS exchange buffers for the FP32 x groups (PlanFFT's DB): one barrier
// per exchange round instead of two (13 per right-hand side instead of 25 for
// three stages), for twice the exchange LDS.  Measured x pass at P^3, s = 16,
// one buffer -> two: 128 0.347 -> 0.327 ms, 160 0.99 -> 0.95, 168 0.83 ->
// 0.80, 200 1.81 -> 1.53, 224 1.82 -> 1.76, 240 2.90 -> 2.77, 288 4.68 ->
// 4.61, 320 5.47 -> 5.42, 384 11.02 -> 10.94, 448 20.0 -> 19.1, 512 21.7 ->
// 21.3, 576 52.5 -> 49.2; slower at 192 (+2 %), 256 (+1 %), 336 (+4 %) and
// 343 (+0.5 %), where the larger LDS leaves fewer groups per CU; sizes below
// 128 within noise.  Plans whose two buffers would pass 80 KB per group (two
// groups in a CU's 160 KB) keep one.
template <typename R, class PL>
struct XDoubleBuffer {
    static constexpr int P = PL::P;
    static constexpr bool value = sizeof(R) == 4 && P * (1 + 2 * PlanTiles<R>::NQ_X) * 8 <= 80 * 1024 && P != 192 &&
                                  P != 256 && P != 336 && P != 343;
};

// A group owns (RPG right-hand sides, ky, NQ consecutive kz) across all F
// component planes.
template <typename R, class PL, int F, int NT, int NQ>
__global__ __launch_bounds__(NT, (xfused_min_waves<R, PL, F, NT, NQ>())) void k_xfused_plan(
    cplx<R>* __restrict__ V, const cplx<R>* __restrict__ Gh, MbtDims d, int Pzs, const cplx<R>* __restrict__ twg,
    int ntz, int nrhs)
{
    constexpr int P = PL::P;
    constexpr int NC = (F == 3) ? 6 : 9;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<R>* tw = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* lds = tw + P;
    const int64_t plane = (int64_t)d.Py * Pzs;
    const int64_t cstride = (int64_t)d.nx * plane;
    int64_t tile;
    int rg;
    // kz tiles in pairs (ntz rounded up to even; the padding groups exit); a
    // group takes RPG right-hand sides one after the other
    constexpr int RPG = XRhsPerGroup<R, PL>::value;
    const int ntzp = (ntz + 1) & ~1;
    xcd_tile_rhs<2>(blockIdx.x, (int64_t)d.Py * ntzp, (nrhs + RPG - 1) / RPG, tile, rg);
    const int ky = (int)(tile / ntzp);
    const int kz0 = (int)(tile % ntzp) * NQ;
    if (kz0 >= d.Pz) return;  // whole group, before any barrier
    const int tqn = min(NQ, d.Pz - kz0);
    for (int i = threadIdx.x; i < P; i += NT) tw[i] = twg[i];
    auto one_rhs = [&](const int rhs) {
    cplx<R>* base = V + (int64_t)rhs * F * cstride + (int64_t)ky * Pzs + kz0;
    PlanFFT<R, PL, NT, NQ, F, 1, XDoubleBuffer<R, PL>::value> X(tw, lds);
    // V addresses split into a uniform row base (component, compile-time row
    // offset RB) and a 32-bit per-lane byte offset (j plane + q, j < P/r0;
    // the launcher checks that it fits), so each access is a scalar-base
    // load or store
    const uint32_t plane32 = (uint32_t)plane;
    auto vaddr = [&](int j, auto RB, int q, int c) {
        char* row = reinterpret_cast<char*>(base + (c * cstride + (int64_t)decltype(RB)::value * plane));
        return reinterpret_cast<cplx<R>*>(row + ((uint32_t)j * plane32 + (uint32_t)q) * (uint32_t)sizeof(cplx<R>));
    };
    X.template load<false, true>([&](int j, auto RB, int q, int c) {
        const int p = j + decltype(RB)::value;
        return (p < d.nx && q < tqn) ? *vaddr(j, RB, q, c) : czero<R>();
    });
    __syncthreads();
    X.template run<false, true>();
    // G^ is stored for one octant; the (kx, ky, kz) image of a reduced element
    // carries the signs of the axis reflections: G'_ab = s_a s_b G_ab for the
    // symmetric part, and det(s) s_a s_b A_ab for the antisymmetric part of f = 6
    const int Hy = d.Py / 2 + 1, Hz = d.Pz / 2 + 1;
    static_assert(NT % NQ == 0 || PlanFFT<R, PL, NT, NQ, F, 1>::template mb<PL::N - 1, false>() == 1,
                  "a thread keeps one kz for all its butterflies");
    const bool fy = ky > d.Py / 2;
    const int kyr = fy ? d.Py - ky : ky;
    // (clamped: lanes of a partial kz tile compute on a valid G^ row and never store)
    const int kz = min(kz0 + (int)(threadIdx.x % NQ), d.Pz - 1);
    const bool fz = kz > d.Pz / 2;
    const int kzr = fz ? d.Pz - kz : kz;
    const int off_yz = (kyr * Hz + kzr) * NC;  // < 2^31 for every planned size
    const int stride_x = Hy * Hz * NC;
    const R sy = fy ? R(-1) : R(1), sz = fz ? R(-1) : R(1), syz = sy * sz;
    X.visit([&](int kx, int, cplx<R>* h) {
        const bool fx = kx > P / 2;
        const int kxr = fx ? P - kx : kx;
        const cplx<R>* gp = reinterpret_cast<const cplx<R>*>(
            reinterpret_cast<const char*>(Gh) + (uint32_t)(kxr * stride_x + off_yz) * (uint32_t)sizeof(cplx<R>));
        const R sx = fx ? R(-1) : R(1), sxy = sx * sy, sxz = sx * sz;
        cplx<R> gv[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) gv[c] = gp[c];
        gv[3] = sxy * gv[3];
        gv[4] = sxz * gv[4];
        gv[5] = syz * gv[5];
        const cplx<R> ex = h[0], ey = h[1], ez = h[2];
        cplx<R> ox = gv[0] * ex + gv[3] * ey + gv[4] * ez;
        cplx<R> oy = gv[3] * ex + gv[1] * ey + gv[5] * ez;
        cplx<R> oz = gv[4] * ex + gv[5] * ey + gv[2] * ez;
        if constexpr (F == 6) {
            const R sxyz = sxy * sz;  // det of the reflection
            const cplx<R> hx = h[3], hy = h[4], hz = h[5];
            const cplx<R> a = (sxyz * sxy) * gv[6], bb = (sxyz * sxz) * gv[7], cc = (sxyz * syz) * gv[8];
            ox = ox + a * hy + bb * hz;
            oy = oy - a * hx + cc * hz;
            oz = oz - bb * hx - cc * hy;
            h[3] = gv[0] * hx + gv[3] * hy + gv[4] * hz - (a * ey + bb * ez);
            h[4] = gv[3] * hx + gv[1] * hy + gv[5] * hz - (cc * ez - a * ex);
            h[5] = gv[4] * hx + gv[5] * hy + gv[2] * hz - (-bb * ex - cc * ey);
        }
        h[0] = ox;
        h[1] = oy;
        h[2] = oz;
    });
    X.template run<true>();
    X.template store<true, true>([&](int j, auto RB, int q, int c, cplx<R> v) {
        const int p = j + decltype(RB)::value;
        if (p < d.nx && q < tqn) *vaddr(j, RB, q, c) = v;
    });
    };
    if constexpr (RPG == 1) {
        one_rhs(rg);
    } else {
        for (int rr = 0; rr < RPG; ++rr)
            if (rg * RPG + rr < nrhs) one_rhs(rg * RPG + rr);  // (whole group)
    }
}

// ---------------------------------------------------------------- z passes
// A group owns TL consecutive (ix, iy) lines with all F components.  The
// FFT engine's lanes run over (line, butterfly) with the line fastest, so
// each of its loads and stores touches TL runs of 64/TL consecutive z
// positions per wave.  The inverse pass reads U and scatters into the
// occupied vector straight from the engine's registers.  The forward pass
// stages the gather through LDS (one contiguous sweep over the cells of the
// group's lines), which measured faster than gathering per line run.

// forward + scatter: U[((rhs F + c) nxny + l) Pzs + kz] = FFT_z(X[f c2o[cell] + c])
// lines (optional, nlines entries): the (ix, iy) lines that hold an occupied
// cell; the groups then take TL of them each and the rows of the others are
// not written (the y pass reads them as zero through its line mask)
template <typename R, class PL, int F, int NT, int TL, int NCX>
__global__ __launch_bounds__(NT) void k_zfwd_plan(const cplx<R>* __restrict__ X, int64_t ldx,
                                                  const int* __restrict__ c2o, cplx<R>* __restrict__ U, MbtDims d,
                                                  int Pzs, const cplx<R>* __restrict__ twg, int64_t ntiles, int nrhs,
                                                  const int* __restrict__ lines, int64_t nlines)
{
    constexpr int P = PL::P;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ __align__(16) int lid[(TL + 3) / 4 * 4];
    cplx<R>* tw = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* lds = tw + P;
    const int64_t nxny = (int64_t)d.nx * d.ny;
    int64_t tile;
    int rhs;
    xcd_tile_rhs(blockIdx.x, ntiles, nrhs, tile, rhs);
    const int64_t l0 = tile * TL;
    const int nl = (int)min((int64_t)TL, (lines ? nlines : nxny) - l0);
    if (threadIdx.x < TL)
        lid[threadIdx.x] = (int)threadIdx.x < nl ? (lines ? lines[l0 + threadIdx.x] : (int)(l0 + threadIdx.x)) : 0;
    for (int i = threadIdx.x; i < P; i += NT) tw[i] = twg[i];
    __syncthreads();  // line ids
    const cplx<R>* Xr = X + rhs * ldx;
    // gather of the occupied values into LDS [(line F + c) PZ + p], one
    // occupancy lookup per cell; rows of odd pitch PZ >= nz keep both the
    // z-contiguous gather and the line-strided FFT loads free of bank conflicts
    // (nz <= H = (P+1)/2 since P >= 2 nz - 1: the sweep runs over TL rows of
    // H so that its index split is by a compile-time constant)
    const int PZ = d.nz | 1;
    constexpr int H = (P + 1) / 2;
    // All occupancy lookups of a thread are issued before any of its X loads,
    // and all X loads before the LDS stores, so the dependent loads overlap
    // across the thread's cells instead of running one cell at a time.
    constexpr int NIT = (TL * H + NT - 1) / NT;
    int mi[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int e = threadIdx.x + it * NT;
        const int p = e % H, tl = e / H;
        mi[it] = -1;
        if (e < TL * H && p < d.nz && tl < nl) {
            const int cell = lid[tl] * d.nz + p;
            mi[it] = c2o ? c2o[cell] : cell;
        }
    }
    cplx<R> xv[NIT][F];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const uint32_t ob = (uint32_t)mi[it] * (uint32_t)(F * sizeof(cplx<R>));
#pragma unroll
        for (int c = 0; c < F; ++c) xv[it][c] = (mi[it] >= 0) ? *at_b(Xr + c, ob) : czero<R>();
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int e = threadIdx.x + it * NT;
        const int p = e % H, tl = e / H;
        if (e < TL * H && p < d.nz) {
#pragma unroll
            for (int c = 0; c < F; ++c) lds[(tl * F + c) * PZ + p] = xv[it][c];
        }
    }
    __syncthreads();
    PlanFFT<R, PL, NT, TL, F, NCX> Z(tw, lds);
    Z.template load<false, true>([&](int j, auto RB, int q, int c) {
        const int p = j + decltype(RB)::value;
        return p < d.nz ? lds[(q * F + c) * PZ + p] : czero<R>();
    });
    __syncthreads();
    Z.template run<false, true>();
    // through LDS, NCX components at a time, so that each wave writes 64
    // consecutive kz of one row (direct stores from the butterfly registers
    // write 16-element runs and measured slower in FP32)
    // (row offsets lid Pzs are 32-bit: the launcher checks nxny Pzs elements fit)
    cplx<R>* Ur = U + (int64_t)rhs * F * nxny * Pzs;
    static_for<0, (F + NCX - 1) / NCX>([&](auto CB) {
        constexpr int c0 = decltype(CB)::value * NCX;
        if constexpr (c0 > 0) __syncthreads();  // previous group's copy-out has read LDS
        Z.template store<false>([&](int j, auto RB, int q, int c, cplx<R> v) {
            if (c >= c0 && c < c0 + NCX) lds[((c - c0) * TL + q) * P + j + decltype(RB)::value] = v;
        });
        __syncthreads();
        for (int e = threadIdx.x; e < NCX * TL * P; e += NT) {
            const int p = e % P, tl = (e / P) % TL, cc = e / (P * TL);
            if (tl < nl && c0 + cc < F)
                *at_b(Ur + (int64_t)(c0 + cc) * nxny * Pzs, ((uint32_t)lid[tl] * (uint32_t)Pzs + (uint32_t)p) *
                                                                (uint32_t)sizeof(cplx<R>)) = lds[e];
        }
    });
}

// lines (optional, nlines entries): only the lines that hold an occupied
// cell are transformed (the others have nothing to gather)
template <typename R, class PL, int F, int NT, int TL, int NCX>
__global__ __launch_bounds__(NT) void k_zinv_plan(const cplx<R>* __restrict__ U, const int* __restrict__ c2o,
                                                  const cplx<R>* __restrict__ X, int64_t ldx,
                                                  const cplx<R>* __restrict__ diag, const cplx<R>* __restrict__ jpre,
                                                  cplx<R>* __restrict__ Y, int64_t ldy, MbtDims d, int Pzs,
                                                  const cplx<R>* __restrict__ twg, int64_t ntiles, int nrhs,
                                                  const int* __restrict__ lines, int64_t nlines)
{
    constexpr int P = PL::P;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ __align__(16) int lid[(TL + 3) / 4 * 4];
    cplx<R>* tw = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* lds = tw + P;
    const int64_t nxny = (int64_t)d.nx * d.ny;
    int64_t tile;
    int rhs;
    xcd_tile_rhs(blockIdx.x, ntiles, nrhs, tile, rhs);
    const int64_t l0 = tile * TL;
    const int nl = (int)min((int64_t)TL, (lines ? nlines : nxny) - l0);
    if (threadIdx.x < TL)
        lid[threadIdx.x] = (int)threadIdx.x < nl ? (lines ? lines[l0 + threadIdx.x] : (int)(l0 + threadIdx.x)) : 0;
    for (int i = threadIdx.x; i < P; i += NT) tw[i] = twg[i];
    __syncthreads();  // line ids
    // (row offsets lid Pzs are 32-bit: the launcher checks nxny Pzs elements fit)
    const cplx<R>* Ur = U + (int64_t)rhs * F * nxny * Pzs;
    PlanFFT<R, PL, NT, TL, F, NCX> Z(tw, lds);
    Z.template load<true>([&](int j, auto RB, int q, int c) {
        const cplx<R>* row = Ur + ((int64_t)c * nxny * Pzs + decltype(RB)::value);
        return q < nl ? *at_b(row, ((uint32_t)lid[q] * (uint32_t)Pzs + (uint32_t)j) * (uint32_t)sizeof(cplx<R>))
                      : czero<R>();
    });
    __syncthreads();  // twiddles
    Z.template run<true>();
    const cplx<R>* Xr = X + rhs * ldx;
    cplx<R>* Yr = Y + rhs * ldy;
    // occupancy lookups of all the thread's cells first, then the updates
    Z.template store_keyed<true, true>(
        [&](int j, auto RB, int q) {
            const int p = j + decltype(RB)::value;
            if (p >= d.nz || q >= nl) return -1;
            const int cell = lid[q] * d.nz + p;
            return c2o ? c2o[cell] : cell;
        },
        [&](int m, int c, cplx<R> v) {
            const uint32_t ob = (uint32_t)m * (uint32_t)(F * sizeof(cplx<R>));
            if (diag) v = *at_b(diag + c, ob) * *at_b(Xr + c, ob) + v;
            if (jpre) v = v * *at_b(jpre + c, ob);
            *at_b(Yr + c, ob) = v;
        });
}

// ---------------------------------------------------------------- launchers
// Each planned pass is launched through a per-plan function template
// (*_plan_launch<R, PL, ...>).  Their explicit instantiations, with the
// kernels they launch, live in the translation units mbt_k_<precision>_<g>.hip,
// one per group of the plan table, so the kernels compile in parallel; every
// other unit sees extern declarations (end of this file) and only dispatches.

// threads of a y group for a plan (0: plan_threads' choice).  FP32 343 (7 7
// 7, 784 butterflies per stage over 16 columns) gets 448 threads with two
// butterflies each instead of 832 with one: y forward 2.74 -> 2.52 ms, y
// inverse 2.35 -> 2.13 ms at the 16384-monomer workload (s = 16); in FP64
// 256 threads were no faster and 128 slower.
template <typename R, class PL>
struct YThreads {
    static constexpr int value = 0;
};

template <typename R, class PL, bool INV>
void lines_plan_launch(const cplx<R>* in, cplx<R>* out, int nlin, int nlout, int64_t Zdim, int64_t pitch,
                       int64_t outer, const cplx<R>* tw, hipStream_t st, const unsigned char* lmask, int nxm)
{
    using T = PlanTiles<R>;
    const int64_t ntiles = ceil_div(Zdim, T::NQ_Y);
    const size_t shm = sizeof(cplx<R>) * (PL::P + (size_t)PL::P * T::NQ_Y);
    constexpr int NT = YThreads<R, PL>::value ? YThreads<R, PL>::value : plan_threads<PL, T::NQ_Y, 4, 256>();
    hipLaunchKernelGGL((k_lines_plan<R, PL, INV, NT, T::NQ_Y>), dim3((unsigned)(ntiles * outer)), dim3(NT), shm, st,
                       in, out, nlin, nlout, Zdim, pitch, ntiles, tw, lmask, nxm);
    BDDA_HIP(hipGetLastError());
}

// threads of an x group for a plan (0: plan_threads' choice, one round of
// butterflies per stage where that fills the lanes best).  Measured FP32 x
// pass, s = 16, at P^3: two butterflies per thread on 256 threads at 280
// (8 7 5: 4.10 -> 3.70 ms) and 336 (8 7 6: 8.06 -> 6.33 ms), against 8.06
// ms for 336 on 448 threads; at 320 (8 8 5: 5.48 -> 6.64 ms), 448 (8 8 7:
// 21.0 -> 29.0 ms) and 512 (8 8 8: 27.5 -> 30.1 ms) the same was slower.
template <typename R, class PL>
struct XThreads {
    static constexpr int value = 0;
};

// false: the grid's lane offsets do not fit 32 bits (the all-in-LDS kernel takes it)
template <typename R, class PL, int F>
bool xfused_plan_launch(cplx<R>* V, const cplx<R>* gh, const MbtDims& d, int Pzs, const cplx<R>* tw, int s,
                        hipStream_t st)
{
    using T = PlanTiles<R>;
    if constexpr (F * PL::P > MBT_MAX_PF) {
        return false;  // refused earlier by mbt_padded_dims; never launched
    } else {
        // per-lane byte offsets into V (within a row block) and G^ are 32-bit
        const uint64_t gh_bytes = (uint64_t)(d.Px / 2 + 1) * (d.Py / 2 + 1) * (d.Pz / 2 + 1) * (F == 3 ? 6 : 9) *
                                  sizeof(cplx<R>);
        if ((uint64_t)(PL::P / PL::r[0]) * d.Py * Pzs * sizeof(cplx<R>) >= (1ull << 32) || gh_bytes >= (1ull << 32))
            return false;
        // (odd line counts that fill the waves, 9 or 5 at 343, measured 19 % and
        // 14 % slower here: their row segments straddle 64-byte boundaries)
        constexpr int NQ = T::NQ_X;
        const int ntz = (int)ceil_div(d.Pz, NQ);
        const size_t shm = sizeof(cplx<R>) * (PL::P + (size_t)PL::P * NQ * (XDoubleBuffer<R, PL>::value ? 2 : 1));
        constexpr int NT = XThreads<R, PL>::value ? XThreads<R, PL>::value : plan_threads<PL, NQ, 2, 512>();
        constexpr int RPG = XRhsPerGroup<R, PL>::value;
        hipLaunchKernelGGL((k_xfused_plan<R, PL, F, NT, NQ>),
                           dim3((unsigned)((int64_t)((s + RPG - 1) / RPG) * d.Py * ((ntz + 1) & ~1))),
                           dim3(NT), shm, st, V, gh, d, Pzs, tw, ntz, s);
        BDDA_HIP(hipGetLastError());
        return true;
    }
}

template <typename R, class PL, int F>
bool zfwd_plan_launch(const cplx<R>* X, int64_t ldx, const int* c2o, cplx<R>* U, const MbtDims& d, int Pzs,
                      const cplx<R>* tw, int s, hipStream_t st, const int* lines, int64_t nlines)
{
    using T = PlanTiles<R>;
    if constexpr (F * PL::P > MBT_MAX_PF) {
        return false;  // refused earlier by mbt_padded_dims; never launched
    } else {
        constexpr int TL = T::TL_Z;  // more lines would cut the resident groups (LDS gather rows)
        const int64_t nxny = (int64_t)d.nx * d.ny;
        const int64_t ntiles = ceil_div(lines ? nlines : nxny, TL);
        if (ntiles == 0) return true;
        const size_t shm = sizeof(cplx<R>) * (PL::P + zpass_lds_elems<PL, F, TL, T::NCX_Z>());
        constexpr int NT = plan_threads<PL, TL, 2, 256>();
        hipLaunchKernelGGL((k_zfwd_plan<R, PL, F, NT, TL, T::NCX_Z>), dim3((unsigned)(ntiles * s)), dim3(NT), shm, st,
                           X, ldx, c2o, U, d, Pzs, tw, ntiles, s, lines, nlines);
        BDDA_HIP(hipGetLastError());
        return true;
    }
}

template <typename R, class PL, int F>
bool zinv_plan_launch(const cplx<R>* U, const int* c2o, const cplx<R>* X, int64_t ldx, const cplx<R>* diag,
                      const cplx<R>* jpre, cplx<R>* Y, int64_t ldy, const MbtDims& d, int Pzs, const cplx<R>* tw,
                      int s, hipStream_t st, const int* lines, int64_t nlines)
{
    using T = PlanTiles<R>;
    if constexpr (F * PL::P > MBT_MAX_PF) {
        return false;  // refused earlier by mbt_padded_dims; never launched
    } else {
        constexpr int TL = plan_lines<PL, T::TL_Z, 2, 256>();
        const int64_t nxny = (int64_t)d.nx * d.ny;
        const int64_t ntiles = ceil_div(lines ? nlines : nxny, TL);
        if (ntiles == 0) return true;
        const size_t shm = sizeof(cplx<R>) * (PL::P + (size_t)T::NCX_Z * TL * PL::P);
        constexpr int NT = plan_threads<PL, TL, 2, 256>();
        hipLaunchKernelGGL((k_zinv_plan<R, PL, F, NT, TL, T::NCX_Z>), dim3((unsigned)(ntiles * s)), dim3(NT), shm, st,
                           U, c2o, X, ldx, diag, jpre, Y, ldy, d, Pzs, tw, ntiles, s, lines, nlines);
        BDDA_HIP(hipGetLastError());
        return true;
    }
}

// the planned y kernel serves (P, pitch): a plan exists and its lane offsets fit 32 bits
template <typename R>
inline bool lines_plan_ok(int P, int64_t pitch)
{
    return has_plan(P) && (uint64_t)P * pitch * sizeof(cplx<R>) < (1ull << 32);
}

template <typename R, bool INV>
inline bool launch_lines_plan(const cplx<R>* in, cplx<R>* out, int nlin, int nlout, int64_t Zdim, int64_t pitch,
                              int64_t outer, int P, const cplx<R>* tw, hipStream_t st,
                              const unsigned char* lmask = nullptr, int nxm = 1)
{
    if (!lines_plan_ok<R>(P, pitch)) return false;
    return with_plan(P, [&](auto pl) {
        lines_plan_launch<R, decltype(pl), INV>(in, out, nlin, nlout, Zdim, pitch, outer, tw, st, lmask, nxm);
    });
}

template <typename R, int F>
inline bool launch_xfused_plan(cplx<R>* V, const cplx<R>* gh, const MbtDims& d, int Pzs, const cplx<R>* tw, int s,
                               hipStream_t st)
{
    bool fits = true;
    const bool planned = with_xplan<R>(d.Px, [&](auto pl) {
        fits = xfused_plan_launch<R, decltype(pl), F>(V, gh, d, Pzs, tw, s, st);
    });
    return planned && fits;
}

// the planned z kernels serve the grid: 32-bit cell indices, lane offsets into
// the occupied rows and row offsets into one U plane
template <typename R, int F>
inline bool zplan_ok(const MbtDims& d, int Pzs)
{
    const uint64_t cells = (uint64_t)d.nx * d.ny * d.nz;
    return has_plan(d.Pz) && cells < (1ull << 31) && cells * F * sizeof(cplx<R>) < (1ull << 32) &&
           (uint64_t)d.nx * d.ny * Pzs * sizeof(cplx<R>) < (1ull << 32);
}

// lines: null, or the nlines (ix, iy) lines that hold an occupied cell
template <typename R, int F>
inline bool launch_zfwd_plan(const cplx<R>* X, int64_t ldx, const int* c2o, cplx<R>* U, const MbtDims& d, int Pzs,
                             const cplx<R>* tw, int s, hipStream_t st, const int* lines = nullptr,
                             int64_t nlines = 0)
{
    if (!zplan_ok<R, F>(d, Pzs)) return false;
    bool fits = true;
    const bool planned = with_plan(d.Pz, [&](auto pl) {
        fits = zfwd_plan_launch<R, decltype(pl), F>(X, ldx, c2o, U, d, Pzs, tw, s, st, lines, nlines);
    });
    return planned && fits;
}

template <typename R, int F>
inline bool launch_zinv_plan(const cplx<R>* U, const int* c2o, const cplx<R>* X, int64_t ldx, const cplx<R>* diag,
                             const cplx<R>* jpre, cplx<R>* Y, int64_t ldy, const MbtDims& d, int Pzs,
                             const cplx<R>* tw, int s, hipStream_t st, const int* lines = nullptr,
                             int64_t nlines = 0)
{
    if (!zplan_ok<R, F>(d, Pzs)) return false;
    bool fits = true;
    const bool planned = with_plan(d.Pz, [&](auto pl) {
        fits = zinv_plan_launch<R, decltype(pl), F>(U, c2o, X, ldx, diag, jpre, Y, ldy, d, Pzs, tw, s, st, lines,
                                                    nlines);
    });
    return planned && fits;
}

// ---------------------------------------------------------------- instantiations
// BDDA_PLAN_INST(EXT, R, PL): the per-plan launch functions of one plan in
// one precision, declared (EXT = extern) or defined (EXT empty).
#define BDDA_PLAN_INST_X(EXT, R, PL, F)                                                                         \
    EXT template bool xfused_plan_launch<R, PL, F>(cplx<R>*, const cplx<R>*, const MbtDims&, int, const cplx<R>*, \
                                                   int, hipStream_t);
#define BDDA_PLAN_INST_XZ(EXT, R, PL, F)                                                                        \
    BDDA_PLAN_INST_X(EXT, R, PL, F)                                                                            \
    EXT template bool zfwd_plan_launch<R, PL, F>(const cplx<R>*, int64_t, const int*, cplx<R>*, const MbtDims&,   \
                                                 int, const cplx<R>*, int, hipStream_t, const int*, int64_t);   \
    EXT template bool zinv_plan_launch<R, PL, F>(const cplx<R>*, const int*, const cplx<R>*, int64_t,            \
                                                 const cplx<R>*, const cplx<R>*, cplx<R>*, int64_t,             \
                                                 const MbtDims&, int, const cplx<R>*, int, hipStream_t,          \
                                                 const int*, int64_t);
#define BDDA_PLAN_INST(EXT, R, PL)                                                                               \
    EXT template void lines_plan_launch<R, PL, false>(const cplx<R>*, cplx<R>*, int, int, int64_t, int64_t,      \
                                                      int64_t, const cplx<R>*, hipStream_t,                    \
                                                      const unsigned char*, int);                              \
    EXT template void lines_plan_launch<R, PL, true>(const cplx<R>*, cplx<R>*, int, int, int64_t, int64_t,       \
                                                     int64_t, const cplx<R>*, hipStream_t, const unsigned char*, \
                                                     int);                                                      \
    BDDA_PLAN_INST_XZ(EXT, R, PL, 3)                                                                           \
    BDDA_PLAN_INST_XZ(EXT, R, PL, 6)

// named plan types (the radix lists carry commas, which macro arguments cannot)
#define BDDA_PLAN_TYPE(p, ...) using PlanOf##p = Radices<__VA_ARGS__>;
BDDA_PLAN_TABLE(BDDA_PLAN_TYPE)
#undef BDDA_PLAN_TYPE
using PlanX192F32 = Radices<8, 8, 3>;  // with_xplan's FP32 x plans at 192 and 384
using PlanX384F32 = Radices<6, 4, 4, 4>;
template <>
struct XThreads<float, PlanX384F32> {
    static constexpr int value = 384;
};
template <>
struct XThreads<float, PlanOf336> {
    static constexpr int value = 256;
};
template <>
struct XThreads<float, PlanOf280> {
    static constexpr int value = 256;
};
template <>
struct YThreads<float, PlanOf343> {
    static constexpr int value = 448;
};

#ifndef BDDA_PLAN_INSTANTIATE
#define BDDA_PLAN_EXTERN(p, ...) BDDA_PLAN_INST(extern, float, PlanOf##p) BDDA_PLAN_INST(extern, double, PlanOf##p)
BDDA_PLAN_TABLE(BDDA_PLAN_EXTERN)
#undef BDDA_PLAN_EXTERN
BDDA_PLAN_INST_X(extern, float, PlanX192F32, 3)
BDDA_PLAN_INST_X(extern, float, PlanX192F32, 6)
BDDA_PLAN_INST_X(extern, float, PlanX384F32, 3)
BDDA_PLAN_INST_X(extern, float, PlanX384F32, 6)
#endif

}  // namespace bdda
