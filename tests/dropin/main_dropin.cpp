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

  // the rcwa.hpp helpers, each checked against what it must equal
  // Vec_3d: the source polarisation as rcwa.cpp:481-498 forms it (W_ref = I,
  // so csrc holds its x and y components at the zero order)
  Vec_3d kinc{};
  for (int i = 0; i < 3; ++i) kinc(i) = k_inc[i];
  Vec_3d a_TE{};
  if (std::abs(source.theta) < 1e-6) a_TE(1) = 1.0;
  else { a_TE(0) = -kinc(1); a_TE(1) = kinc(0); }
  a_TE = a_TE.normalized();
  Vec_3d a_TM = kinc.cross(a_TE).normalized();
  Vec_3d Pol = (source.pte * a_TE + source.ptm * a_TM).normalized();
  std::printf("helper vec3d %.3e\n", std::max(std::abs(Pol(0) - csrc(PQ / 2)),
                                              std::abs(Pol(1) - csrc(PQ / 2 + PQ))));
  // unsigned_sqrt over a matrix: -1 - 0i -> +i, and the scalar form elsewhere
  Matrix Zs(2, 2);
  Zs(0, 0) = Complex(-1.0, -0.0);
  Zs(1, 0) = Complex(2.0, 1.0);
  Zs(0, 1) = Kz_ref(0);
  Zs(1, 1) = Complex(-4.0, 0.0);
  Matrix Zr = unsigned_sqrt(Zs);
  double usq = std::abs(Zr(0, 0) - Complex(0, 1)) + std::abs(Zr(1, 1) - Complex(0, 2));
  for (long j = 0; j < 2; ++j)
    for (long i = 0; i < 2; ++i) usq += std::abs(Zr(i, j) - unsigned_sqrt(Zs(i, j)));
  std::printf("helper usqrt %.3e\n", usq);
  // MeshGrid of the harmonic orders (main.cpp:204-213)
  Vector kx_t = Vector::LinSpaced(P, -params.Nx_harmonics, params.Nx_harmonics);
  Vector ky_t = Vector::LinSpaced(Q, -params.Ny_harmonics, params.Ny_harmonics);
  Matrix KX, KY;
  MeshGrid(kx_t, ky_t, KX, KY);
  double mg = (KX.rows() == Q && KX.cols() == P && KY.rows() == Q && KY.cols() == P) ? 0.0 : 1.0;
  for (int i = 0; i < Q; ++i)
    for (int j = 0; j < P; ++j)
      mg += std::abs(KX(i, j) - Complex(j - params.Nx_harmonics, 0)) +
            std::abs(KY(i, j) - Complex(i - params.Ny_harmonics, 0));
  std::printf("helper meshgrid %.3e\n", mg);
  // ComplexEigenSolver + extraction on the global S11: residual of A V = V diag(w)
  const Matrix& A = S_global.S11;
  EigenvalSolverResults ev = extraction(rcwa_amd::ComplexEigenSolver<Matrix>(A));
  const long n = A.rows();
  Matrix AV = A * ev.eigenvectors;
  double res = 0, an = A.norm();
  for (long j = 0; j < n; ++j) {
    double cn = 0;
    for (long i = 0; i < n; ++i) cn += std::norm(ev.eigenvectors(i, j));
    if (std::abs(std::sqrt(cn) - 1.0) > 1e-12) res += 1.0;
    for (long i = 0; i < n; ++i)
      res = std::max(res, std::abs(AV(i, j) - ev.eigenvectors(i, j) * ev.eigenvalues(j)) / an);
  }
  std::printf("helper eig %.3e\n", res);
  // PartialPivLU solve, and the matrix product on the GPU (70^3 >= 64^3)
  Matrix B = S_device.S12 + Matrix::Identity(n, n);
  Matrix rhs = S_device.S11.block(0, 0, n, 3);
  Matrix x = rcwa_amd::lu(B).solve(rhs);
  std::printf("helper lu %.3e\n", (B * x - rhs).norm() / (B.norm() * x.norm()));
  Matrix G1(70, 70), G2(70, 70), Gh(70, 70);
  for (int j = 0; j < 70; ++j)
    for (int i = 0; i < 70; ++i) {
      G1(i, j) = Complex(std::sin(i + 2.0 * j), std::cos(3.0 * i - j));
      G2(i, j) = Complex(std::cos(i * 0.5 + j), std::sin(i - 0.25 * j));
    }
  Matrix Gd = G1 * G2;
  for (int j = 0; j < 70; ++j)
    for (int k = 0; k < 70; ++k)
      for (int i = 0; i < 70; ++i) Gh(i, j) += G1(i, k) * G2(k, j);
  std::printf("helper gemm %.3e\n", (Gd - Gh).norm() / Gh.norm());
  return 0;
}
