// A caller written like rcwa3d/src/main.cpp:112-198 (same call sequence and
// names), compiled against include/rcwa_amd.hpp.  It solves the structure
// given on stdin and prints R, T, R_tot, T_tot (stage by stage) and the same
// numbers from the one-call Solve(), for the parity test.
//
// stdin: Nx Ny L Lx Ly hx hy lambda theta phi pte ptm
//        er_ref(re im) ur_ref er_trn ur_trn, t[L], then er[L*Ny*Nx] (re im), ur[...]
#include <cstdio>
#include <iostream>
#include "rcwa_amd.hpp"

static int run();
int main() {
  try {
    return run();
  } catch (const rcwa_amd::Error& e) {
    std::fprintf(stderr, "rcwa_amd::Error status %d: %s\n", (int)e.status, e.what());
    return 3;
  }
}

static int run() {
  int Nx, Ny, L, hx, hy;
  double Lx, Ly, lam, th, ph, pte, ptm;
  double m[8];
  if (!(std::cin >> Nx >> Ny >> L >> Lx >> Ly >> hx >> hy >> lam >> th >> ph >> pte >> ptm)) return 2;
  for (double& v : m) std::cin >> v;
  std::vector<Real> thickness(L);
  for (auto& v : thickness) std::cin >> v;
  std::vector<Complex> er((size_t)L * Nx * Ny), ur((size_t)L * Nx * Ny);
  for (auto& v : er) { double a, b; std::cin >> a >> b; v = Complex(a, b); }
  for (auto& v : ur) { double a, b; std::cin >> a >> b; v = Complex(a, b); }
  if (!std::cin) {
    std::fprintf(stderr, "malformed input\n");
    return 2;
  }

  Device device(Nx, Ny, L, Lx, Ly, er, ur, thickness, hx, hy);
  Source source(lam, th, ph, pte, ptm);
  RCWAParams params(hx, hy, Complex(m[0], m[1]), Complex(m[2], m[3]), Complex(m[4], m[5]),
                    Complex(m[6], m[7]));
  int P = 2 * params.Nx_harmonics + 1;
  int Q = 2 * params.Ny_harmonics + 1;
  int PQ = P * Q;
  std::vector<Complex> k_inc(3, Complex(0.0, 0.0));
  Vector Kx, Ky, Kz_ref, Kz_trn;
  Matrix W0(2 * PQ, 2 * PQ);
  Matrix V0(2 * PQ, 2 * PQ);
  Matrix W_ref(2 * PQ, 2 * PQ);
  Matrix W_trn(2 * PQ, 2 * PQ);
  for (int layer = 0; layer < device.num_layers; ++layer) {
    device.erc.at(layer) = ConvMat(device.er, layer, device.Nx, device.Ny, params.Nx_harmonics,
                                   params.Ny_harmonics);
    device.urc.at(layer) = ConvMat(device.ur, layer, device.Nx, device.Ny, params.Nx_harmonics,
                                   params.Ny_harmonics);
  }
  ComputeWaveVectors(device, source, params, k_inc, Kx, Ky, Kz_ref, Kz_trn);
  GapMedium(Kx, Ky, W0, V0);
  ScatteringMatrix S_device = SMatrixInit(params.Nx_harmonics, params.Ny_harmonics);
  ScatteringMatrix S_layer = ScatteringMatrix(Matrix::Zero(2 * PQ, 2 * PQ), Matrix::Zero(2 * PQ, 2 * PQ),
                                              Matrix::Zero(2 * PQ, 2 * PQ), Matrix::Zero(2 * PQ, 2 * PQ));
  for (int layer = 0; layer < device.num_layers; ++layer) {
    S_layer = SMatrixLayer(layer, device, source, params, Kx, Ky, W0, V0);
    S_device = RedhefferProduct(S_device, S_layer);
  }
  std::fprintf(stderr, "layers folded\n");
  ScatteringMatrix S_ref = SMatrixReflection(params, Kx, Ky, Kz_ref, W0, V0, W_ref);
  ScatteringMatrix S_trn = SMatrixTransmission(params, Kx, Ky, Kz_trn, W0, V0, W_trn);
  ScatteringMatrix S_global = RedhefferProduct(S_ref, S_device);
  S_global = RedhefferProduct(S_global, S_trn);
  Vector csrc(2 * PQ);
  ComputeSourceModeCoeff(source, params, k_inc, W_ref, csrc);
  Vector r = Vector::Zero(3 * PQ);
  ComputeReflectedField(params, S_global, csrc, Kx, Ky, Kz_ref, W_ref, r);
  Vector t = Vector::Zero(3 * PQ);
  ComputeTransmittedField(params, S_global, csrc, Kx, Ky, Kz_trn, W_trn, t);
  Results results = ComputeDiffractionEfficiencies(params, r, t, k_inc, Kz_ref, Kz_trn);
  std::fprintf(stderr, "staged path done\n");
  Results one = Solve(device, source, params);
  std::printf("%d\n", PQ);
  for (int i = 0; i < PQ; ++i) std::printf("%.17g %.17g\n", results.R(i), results.T(i));
  std::printf("%.17g %.17g\n", results.R_tot, results.T_tot);
  for (int i = 0; i < PQ; ++i) std::printf("%.17g %.17g\n", one.R(i), one.T(i));
  std::printf("%.17g %.17g\n", one.R_tot, one.T_tot);
  return 0;
}
