// C wrapper over the reference's Mie routines (gmie/gmie.cpp, gmie/gmie_coeff.cpp),
// compiled from the reference sources by `make -C oracle ref` into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: used by tests/golden/make_mie_golden.py to write
// Mie fixtures; nothing in the product links it.
#include <complex>
#include <tuple>
#include <vector>
#include "gmie.hpp"

extern "C" {

// Qsca, Qabs, Qext -> q[0..2]; S1, S2 -> interleaved complex arrays of nang
int ref_gmie(double wl0, double rp, int nang, const double* eper_p, const double* mper_p, const double* eper_m,
             const double* mper_m, double* q, double* S1, double* S2)
{
    const std::complex<double> ep(eper_p[0], eper_p[1]), mp(mper_p[0], mper_p[1]);
    const std::complex<double> em(eper_m[0], eper_m[1]), mm(mper_m[0], mper_m[1]);
    double qsca, qabs, qext;
    std::vector<std::complex<double>> s1, s2;
    std::tie(qsca, qabs, qext, s1, s2) = gmie(wl0, rp, nang, ep, mp, em, mm);
    q[0] = qsca;
    q[1] = qabs;
    q[2] = qext;
    for (int j = 0; j < nang; ++j) {
        S1[2 * j] = s1[j].real();
        S1[2 * j + 1] = s1[j].imag();
        S2[2 * j] = s2[j].real();
        S2[2 * j + 1] = s2[j].imag();
    }
    return 0;
}

// first nmax Mie coefficients a_n, b_n (interleaved complex); returns the count
int ref_gmie_coeff(double wl0, double rp, const double* eper_p, const double* mper_p, const double* eper_m,
                   const double* mper_m, int nmax, double* a, double* b)
{
    const std::complex<double> ep(eper_p[0], eper_p[1]), mp(mper_p[0], mper_p[1]);
    const std::complex<double> em(eper_m[0], eper_m[1]), mm(mper_m[0], mper_m[1]);
    std::vector<std::complex<double>> va, vb, vc, vd;
    std::tie(va, vb, vc, vd) = gmie_coeff(wl0, rp, ep, mp, em, mm);
    const int n = (int)va.size() < nmax ? (int)va.size() : nmax;
    for (int j = 0; j < n; ++j) {
        a[2 * j] = va[j].real();
        a[2 * j + 1] = va[j].imag();
        b[2 * j] = vb[j].real();
        b[2 * j + 1] = vb[j].imag();
    }
    return n;
}
}
