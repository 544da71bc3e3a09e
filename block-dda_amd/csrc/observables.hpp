// observables.hpp -- cross sections and the far-field scattering matrix of a
// solved block of dipole moments, computed on the device.
#pragma once
#include "solvers.hpp"

namespace bdda {

// Cext[l], Cabs[l] for the L columns of the device blocks dB (incident field
// EHinc) and dX (dipoles PM), both (f*nocc) x L FP64 with ld = f*nocc
// (Result_manager.cpp:65-122):
//   Cext = 4 pi k Im(sum_i conj(B_i) X_i)
//   Cabs = 4 pi k sum_i (Im(conj(1/alpha_i)) - 2 k^3 / 3) |X_i|^2,
// 1/alpha_i = DIAG_i of component i; alpha_each = 0 uses the DIAG of dipole 0
// for every dipole, as the reference does (Result_manager.cpp:73,97-98).
void cross_sections(Problem& P, const c64* dB, const c64* dX, int L, double k, int alpha_each, double* Cext,
                    double* Cabs);

// cross_sections without any device allocation: the partial sums go to
// dscratch (from xs_scratch_device(), the device address of a file-scope
// array on the current device) instead of a per-call hipMalloc.  L is
// 1..XS_MAX_COLS; results are bit-identical to cross_sections.  Calls on one
// GPU share that array, so calls that overlap on one GPU race.
constexpr int XS_MAX_COLS = 256;
double* xs_scratch_device();
void cross_sections_noalloc(Problem& P, const c64* dB, const c64* dX, int L, double k, int alpha_each,
                            double* dscratch, double* Cext, double* Cabs);

// Far-field amplitudes of npairs column pairs (2p: incident polarisation
// e1_p, 2p+1: e2_p = u_p x e1_p, both along u_p) in the plane spanned by u_p
// and e1_p at theta_a = pi a / (nang - 1): the Bohren-Huffman amplitude
// matrix S[(p nang + a) 4 + {0,1,2,3}] = S1, S2, S3, S4 with
//   E_far = k^2 e^{ikr}/r sum_j [(I - n n) P_j - n x M_j] e^{-ik n.r_j},
//   S2 = -ik theta.E', S4 = -ik phi.E' (pol e1), S3 = -ik theta.E',
//   S1 = -ik phi.E' (pol e2), E' = E_far r e^{-ikr},
// n = cos(theta) u + sin(theta) e1, theta-hat = -sin(theta) u + cos(theta)
// e1, phi-hat = u x e1, r_j = lf (ix, iy, iz).  M_j exists for f = 6.
void scattering_amplitudes(Problem& P, const c64* dX, int npairs, double k, const double* u, const double* e1,
                           int nang, c64* S);

}  // namespace bdda
