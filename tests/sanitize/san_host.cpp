// The host-side code of the RCWA path built with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/sanitize/Makefile; run by
// tests/test_sanitize.py): the O(N) host stages the product runs on the host
// (rcwa-cuda_amd/csrc/host_math.cpp: wave vectors, gap-medium diagonals, the
// half-space S-matrices, the source/field/efficiency tail of
// rcwa.cpp:469-585) and the CPU FP64 oracle (oracle/rcwa_oracle.cpp), over
// seeded random problems and the edge shapes the path accepts (one harmonic,
// no layers, Nx != Ny, hx != hy, odd grids).  Any out-of-bounds access, use
// after free, leak or undefined behaviour aborts the run with a report; the
// numerical checks tie the host tail to the oracle:
//   - rcwa_host::wave_vectors = the oracle's wave vectors (1e-13 relative);
//   - rcwa_host::efficiencies on the oracle's own global S11/S21 columns =
//     the oracle's per-order R, T (1e-12);
//   - lossless stacks: R + T = 1 (1e-10);
//   - the oracle's eigenpairs: |A v - lambda v| <= 1e-12 |A| n;
//   - ConvMat of a constant grid = that constant times I.
// Prints "ok" and exits 0 when every check holds.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "host_math.hpp"

extern "C" {
const char* oracle_last_error();
int oracle_convmat(const double* grid, int Nx, int Ny, int hx, int hy, double* out);
int oracle_eig(int n, const double* A, double* values, double* vectors, int* iters);
int oracle_wave_vectors(double Lx, double Ly, double lam0, double theta, double phi, int hx,
                        int hy, const double* media, double* k_inc, double* Kx, double* Ky,
                        double* Kz_ref, double* Kz_trn);
int oracle_solve(int Nx, int Ny, int L, double Lx, double Ly, const double* er, const double* ur,
                 const double* thick, double lam0, double theta, double phi, int hx, int hy,
                 const double* media, int npol, const double* pols, double* R, double* T,
                 double* tot, double* Sg, int layer_limit);
}

using cplx = std::complex<double>;

static int failures = 0;

static void expect(bool ok, const char* what, double got, double bound) {
  if (!ok) {
    std::printf("FAIL %s: %.3e (bound %.1e)\n", what, got, bound);
    ++failures;
  }
}

struct Case {
  int hx, hy, L, Nx, Ny;
  bool lossless;
};

static void solve_case(std::mt19937_64& rng, const Case& c) {
  std::uniform_real_distribution<double> U(0.0, 1.0);
  const double Lx = 300 + 400 * U(rng), Ly = 300 + 400 * U(rng);
  const double lam0 = 400 + 500 * U(rng), theta = U(rng), phi = U(rng);
  const size_t G = (size_t)c.Nx * c.Ny;
  std::vector<cplx> er((size_t)c.L * G), ur((size_t)c.L * G, cplx(1, 0));
  for (int l = 0; l < c.L; ++l) {
    const cplx base(1 + 11 * U(rng), c.lossless ? 0.0 : U(rng) - 0.5);
    const cplx incl(1 + 11 * U(rng), c.lossless ? 0.0 : U(rng) - 0.5);
    for (size_t g = 0; g < G; ++g) er[l * G + g] = U(rng) < 0.4 ? incl : base;
  }
  std::vector<double> thick(c.L);
  for (double& t : thick) t = 20 + 280 * U(rng);
  const cplx er_trn(1 + 2 * U(rng), c.lossless ? 0.0 : -0.2 * U(rng));
  const double media[8] = {1, 0, 1, 0, er_trn.real(), er_trn.imag(), 1, 0};
  const double pols[6] = {1, 0, 0, 1, 0.6, 0.8};
  const int npol = 3;
  const int P = 2 * c.hx + 1, Q = 2 * c.hy + 1, N = P * Q, n = 2 * N;
  std::vector<double> R(npol * N), T(npol * N), tot(2 * npol), Sg(4 * 2 * (size_t)n * n);
  if (oracle_solve(c.Nx, c.Ny, c.L, Lx, Ly, reinterpret_cast<const double*>(er.data()),
                   reinterpret_cast<const double*>(ur.data()), thick.data(), lam0, theta, phi,
                   c.hx, c.hy, media, npol, pols, R.data(), T.data(), tot.data(), Sg.data(),
                   -1)) {
    std::printf("FAIL oracle_solve: %s\n", oracle_last_error());
    ++failures;
    return;
  }
  // the product's host stages on the oracle's global S-matrix
  rcwa_host::Media m{cplx(1, 0), cplx(1, 0), er_trn, cplx(1, 0)};
  rcwa_host::WaveVectors wv =
      rcwa_host::wave_vectors(Lx, Ly, rcwa_host::Incidence{lam0, theta, phi}, c.hx, c.hy, m);
  std::vector<cplx> ki(3), kx(N), ky(N), kr(N), kt(N);
  oracle_wave_vectors(Lx, Ly, lam0, theta, phi, c.hx, c.hy, media,
                      reinterpret_cast<double*>(ki.data()), reinterpret_cast<double*>(kx.data()),
                      reinterpret_cast<double*>(ky.data()), reinterpret_cast<double*>(kr.data()),
                      reinterpret_cast<double*>(kt.data()));
  double dk = 0, scale = 0;
  for (int i = 0; i < N; ++i) {
    dk = std::max({dk, std::abs(wv.Kx[i] - kx[i]), std::abs(wv.Ky[i] - ky[i]),
                   std::abs(wv.Kz_ref[i] - kr[i]), std::abs(wv.Kz_trn[i] - kt[i])});
    scale = std::max({scale, std::abs(kx[i]), std::abs(ky[i]), std::abs(kr[i]), std::abs(kt[i])});
  }
  expect(dk <= 1e-13 * std::max(1.0, scale), "wave vectors vs oracle", dk, 1e-13);
  std::vector<cplx> v0 = rcwa_host::gap_v0_diagonals(wv.Kx, wv.Ky);
  std::vector<cplx> e0 = rcwa_host::end_smatrix_diagonals(0, wv.Kx, wv.Ky, wv.Kz_ref, m.er_ref,
                                                          m.ur_ref);
  std::vector<cplx> e1 = rcwa_host::end_smatrix_diagonals(1, wv.Kx, wv.Ky, wv.Kz_trn, m.er_trn,
                                                          m.ur_trn);
  bool finite = v0.size() == 4 * (size_t)N && e0.size() == 16 * (size_t)N &&
                e1.size() == 16 * (size_t)N;
  for (const cplx& z : v0) finite = finite && std::isfinite(z.real()) && std::isfinite(z.imag());
  for (const cplx& z : e0) finite = finite && std::isfinite(z.real()) && std::isfinite(z.imag());
  for (const cplx& z : e1) finite = finite && std::isfinite(z.real()) && std::isfinite(z.imag());
  expect(finite, "gap-medium and half-space diagonals finite", 0, 0);
  // the S11 / S21 columns at the source entries N/2 and N/2 + N, back to back
  const cplx* S11 = reinterpret_cast<const cplx*>(Sg.data());
  const cplx* S21 = S11 + 2 * (size_t)n * n;
  std::vector<cplx> s11c(2 * (size_t)n), s21c(2 * (size_t)n);
  const int cols[2] = {N / 2, N / 2 + N};
  for (int k = 0; k < 2; ++k)
    for (int i = 0; i < n; ++i) {
      s11c[(size_t)k * n + i] = S11[i + (size_t)cols[k] * n];
      s21c[(size_t)k * n + i] = S21[i + (size_t)cols[k] * n];
    }
  for (int ip = 0; ip < npol; ++ip) {
    rcwa_host::Efficiencies e =
        rcwa_host::efficiencies(theta, pols[2 * ip], pols[2 * ip + 1], wv, m, s11c.data(),
                                s21c.data());
    double d = std::max(std::abs(e.R_tot - tot[2 * ip]), std::abs(e.T_tot - tot[2 * ip + 1]));
    for (int i = 0; i < N; ++i)
      d = std::max({d, std::abs(e.R[i] - R[(size_t)ip * N + i]),
                    std::abs(e.T[i] - T[(size_t)ip * N + i])});
    expect(d <= 1e-12, "host efficiencies vs oracle", d, 1e-12);
    expect(e.r.size() == 3 * (size_t)N && e.t.size() == 3 * (size_t)N, "field sizes", 0, 0);
    if (c.lossless) {
      const double dev = std::abs(e.R_tot + e.T_tot - 1.0);
      expect(dev <= 1e-10, "lossless R + T = 1", dev, 1e-10);
    }
  }
}

static void eig_case(std::mt19937_64& rng, int n) {
  std::normal_distribution<double> Nd(0.0, 1.0);
  std::vector<cplx> A((size_t)n * n), V((size_t)n * n), lam(n);
  double an = 0;
  for (cplx& z : A) {
    z = cplx(Nd(rng), Nd(rng));
    an = std::max(an, std::abs(z));
  }
  int iters = 0;
  if (oracle_eig(n, reinterpret_cast<const double*>(A.data()), reinterpret_cast<double*>(lam.data()),
                 reinterpret_cast<double*>(V.data()), &iters)) {
    std::printf("FAIL oracle_eig: %s\n", oracle_last_error());
    ++failures;
    return;
  }
  double res = 0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      cplx s = -lam[j] * V[i + (size_t)j * n];
      for (int k = 0; k < n; ++k) s += A[i + (size_t)k * n] * V[k + (size_t)j * n];
      res = std::max(res, std::abs(s));
    }
  expect(res <= 1e-12 * an * n, "eigen residual", res, 1e-12 * an * n);
}

static void convmat_case(int Nx, int Ny, int hx, int hy) {
  const cplx c0(3.25, -0.5);
  std::vector<cplx> g((size_t)Nx * Ny, c0);
  const int N = (2 * hx + 1) * (2 * hy + 1);
  std::vector<cplx> C((size_t)N * N);
  if (oracle_convmat(reinterpret_cast<const double*>(g.data()), Nx, Ny, hx, hy,
                     reinterpret_cast<double*>(C.data()))) {
    std::printf("FAIL oracle_convmat: %s\n", oracle_last_error());
    ++failures;
    return;
  }
  double d = 0;
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) d = std::max(d, std::abs(C[i + (size_t)j * N] - (i == j ? c0 : 0.0)));
  expect(d <= 1e-13, "ConvMat of a constant grid", d, 1e-13);
}

int main() {
  std::mt19937_64 rng(20261204);
  const Case cases[] = {
      {1, 1, 2, 8, 8, false}, {2, 1, 3, 12, 9, false}, {0, 0, 1, 4, 4, false},
      {1, 2, 0, 8, 8, false}, {2, 2, 2, 16, 16, true}, {3, 1, 1, 15, 7, true},
  };
  for (const Case& c : cases) solve_case(rng, c);
  for (int n : {1, 2, 7, 24}) eig_case(rng, n);
  convmat_case(7, 5, 1, 2);
  convmat_case(8, 8, 2, 2);
  convmat_case(1, 1, 0, 0);
  if (failures) {
    std::printf("%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
