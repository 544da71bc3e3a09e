// mbt.hpp -- the MBT (multilevel block-Toeplitz) operator on one GPU.
//
// The reference evaluates y_i = sum_j M(lf (i - j)) x_j over the nx*ny*nz
// cuboid through one 1-D FFT of length f^2 (2nx-1)(2ny-1)(2nz-1)
// (mvp_fft/MBT_fft_mvp.cpp:8-43).  Here the same linear convolution is a
// zero-padded 3-D FFT convolution on a Px*Py*Pz grid (P >= 2n-1 per axis,
// 2/3/5/7-smooth) with a per-frequency f x f multiply by the transformed
// interaction kernel G^ (symmetry-reduced to one octant).
//
// Device layouts (complex T interleaved):
//   occupied vectors   X[rhs * ld + f m + c]                (reference layout)
//   U  (after z pass)  U[(((rhs f + c) nx + ix) ny + iy) Pzs + kz]
//   V  (after y pass)  V[(((rhs f + c) nx + ix) Py + ky) Pzs + kz]
// One plane per component; rows are padded to Pzs (a multiple of 16
// elements, 128 or 256 bytes) so that every tile a pass moves starts on a
// cache-line boundary.
//   G^ reduced         Gh[((kx' (Py/2+1) + ky') (Pz/2+1) + kz')][nc]
//                      nc = 6 (f=3: xx yy zz xy xz yz of -G)
//                      nc = 9 (f=6: the same + xy xz yz of -F)
#pragma once
#include "bdda_common.hpp"
#include "fft_engine.hpp"
#include <vector>

namespace bdda {

struct MbtDims {
    int nx, ny, nz, f;
    int Px, Py, Pz;
};

// choose the padded FFT size for one axis: the smallest size >= 2n-1 with a
// compile-time plan (a multiple of plan_mult) unless that pads more than
// 12 % beyond the smallest 2/3/5/7-smooth multiple of `mult`
int choose_padded(int n, int mult, int plan_mult);
AxisPlan make_axis_plan(int P, bool last8);

// The all-in-LDS passes (sizes without a compile-time plan, and f = 6) hold
// one padded line of every component in a work-group: P f <= MBT_MAX_PF.
constexpr int MBT_MAX_PF = 3072;
// padded sizes of the operator for lattice n; false when an axis exceeds
// what the passes can transform
bool mbt_padded_dims(const int n[3], int f, int P[3]);

enum Precision : int { PREC_FP64 = 0, PREC_FP32 = 1 };

struct MbtOperator {
    MbtDims d;
    double lf;
    double kre, kim;
    AxisPlan ax[3];             // x plan ends with radix 8
    int Pzs = 0;                // row pitch of U and V (mbt_z_pitch): Pz rounded up to 16, padded
    void* tw64[3] = {nullptr, nullptr, nullptr};  // per-axis twiddle tables
    void* tw32[3] = {nullptr, nullptr, nullptr};
    void* gh64 = nullptr;       // reduced G^, FP64 (scaled by 1/(PxPyPz))
    void* gh32 = nullptr;       // reduced G^, FP32 copy
    int have64 = 0, have32 = 0;
    int nc = 6;
    size_t gh_elems = 0;        // reduced elements per component set
    hipStream_t stream = nullptr;
    // workspace for the passes (grown on demand)
    void* ws = nullptr;
    size_t ws_bytes = 0;
    size_t ws_limit = 0;        // > 0: block matvecs run in column chunks within it
    // occupancy map for the occupied-vector path
    int* c2o = nullptr;         // cell -> occupied index or -1
    int64_t c2o_cells = 0;
    int* addr_dev = nullptr;    // occupied index -> cell (the address array)
    int64_t addr_n = 0;
    std::vector<int> c2o_host_addr;  // address array the map was built from
    // (ix, iy) lines holding at least one occupied cell: their ids
    // (ny ix + iy, ascending) and a per-line flag; the occupied-vector passes
    // transform only these lines along z
    int* lines = nullptr;
    int64_t nlines = 0;
    unsigned char* lmask = nullptr;
    // launch timing of the last apply (ms per pass), filled when timing on
    int timing = 0;
    hipEvent_t ev[8] = {};
    float pass_ms[6] = {};
};

// Create the operator: padded sizes, plans, G^ build (FP64 always, FP32
// copy when want32).  Throws HipError on runtime failure.
MbtOperator* mbt_create(const int n[3], int f, double lf, double kre, double kim,
                        int want64, int want32, hipStream_t stream);
void mbt_destroy(MbtOperator* op);

// set the occupancy (address[m] = ny nz ix + nz iy + iz, 0-based)
void mbt_set_address(MbtOperator* op, const int* address_host, int64_t nocc);

// Y[:, r] = diag .* X[:, r] + gather(conv(scatter(X[:, r])))  (then .* jpre)
// for r < s, all pointers on the device.  diag/jpre may be null.
template <typename R>
void mbt_apply_occupied(MbtOperator* op, const cplx<R>* X, int64_t ldx, cplx<R>* Y, int64_t ldy,
                        int s, const cplx<R>* diag, const cplx<R>* jpre, int64_t nocc);

// grow the pass workspace to at least `bytes`
void mbt_ensure_workspace(MbtOperator* op, size_t bytes);
// give the pass workspace back to the device allocator (after the stream's
// work); the C ABI's calls that run block matvecs do so before they return
void mbt_release_workspace(MbtOperator* op);

// one precision's pass sequence (mbt_f32.hip / mbt_f64.hip); c2o == null
// selects the cuboid layout
template <typename R>
void mbt_apply_precision(MbtOperator* op, const cplx<R>* X, int64_t ldx, cplx<R>* Y, int64_t ldy, int s,
                         const cplx<R>* diag, const cplx<R>* jpre, const int* c2o);

extern template void mbt_apply_precision<float>(MbtOperator*, const cplx<float>*, int64_t, cplx<float>*, int64_t,
                                                int, const cplx<float>*, const cplx<float>*, const int*);
extern template void mbt_apply_precision<double>(MbtOperator*, const cplx<double>*, int64_t, cplx<double>*, int64_t,
                                                 int, const cplx<double>*, const cplx<double>*, const int*);

// pass workspace of one column (U and V grids) in precision R
size_t mbt_workspace_per_column(const MbtDims& d, int Pzs, size_t elem_bytes);
// row pitch Pzs of U and V for a padded z length Pz
int mbt_z_pitch(int Pz);

// bytes per block matvec under the minimal fused pipeline (the algorithmic
// model used for roofline numbers)
double mbt_algorithmic_bytes(const MbtOperator* op, int s, int64_t nocc, size_t elem_bytes);

}  // namespace bdda
