// O(N) host-side stages of the RCWA path: wave vectors, gap-medium modes,
// the half-space S-matrices in closed form, and the source/field/efficiency
// tail.  These are the reference functions whose every matrix is a 2x2 block
// of diagonals, so they cost O(N) and stay on the host.
#pragma once
#include <complex>
#include <vector>

namespace rcwa_host {

using cplx = std::complex<double>;

inline cplx usqrt(cplx z) {  // rcwa.hpp:24-28
  if (z.imag() == 0.0) z = cplx(z.real(), 0.0);
  return std::sqrt(z);
}

struct Incidence {
  double lambda0, theta, phi;
};

struct Media {
  cplx er_ref, ur_ref, er_trn, ur_trn;
};

struct WaveVectors {
  cplx k_inc[3];
  std::vector<cplx> Kx, Ky, Kz_ref, Kz_trn;
};

// rcwa.cpp:85-160
WaveVectors wave_vectors(double Lx, double Ly, const Incidence& inc, int hx, int hy,
                         const Media& m);

// Gap medium V0 (rcwa.cpp:163-234) as its four diagonals d[2*bi+bj][m]; W0 = I.
std::vector<cplx> gap_v0_diagonals(const std::vector<cplx>& Kx, const std::vector<cplx>& Ky);

// SMatrixReflection (which = 0, rcwa.cpp:354-411) and SMatrixTransmission
// (which = 1, rcwa.cpp:414-466) in closed form: with W0 = W_end = I every
// block is a 2x2 block of diagonals; out[blk][2*bi+bj][m], blk = S11, S12,
// S21, S22.  The reference's dense LU solves reduce to one 2x2 system per
// harmonic.
std::vector<cplx> end_smatrix_diagonals(int which, const std::vector<cplx>& Kx,
                                        const std::vector<cplx>& Ky, const std::vector<cplx>& Kz,
                                        cplx er, cplx ur);

struct Efficiencies {
  std::vector<double> R, T;  // PQ each, harmonic index order (= Q x P column-major)
  double R_tot = 0, T_tot = 0;
  std::vector<cplx> r, t;    // 3PQ each: the reflected / transmitted fields (x, y, z)
};

// ComputeSourceModeCoeff + Compute{Reflected,Transmitted}Field +
// ComputeDiffractionEfficiencies (rcwa.cpp:469-585) with W_ref = W_trn = I.
// s11c / s21c hold the S11 / S21 columns at the source entries PQ/2 and
// PQ/2+PQ, each 2PQ long, back to back.
Efficiencies efficiencies(double theta, double pte, double ptm, const WaveVectors& wv,
                          const Media& m, const cplx* s11c, const cplx* s21c);

}  // namespace rcwa_host
