// C++ drop-in for rcwa3d's interface (rcwa3d/include/rcwa.hpp), on top of the
// C ABI of librcwa_amd.so (include/rcwa_amd.h).
//
// Same type names, field names, constructor signatures and free functions as
// rcwa.hpp, so a caller written against rcwa3d (rcwa3d/src/main.cpp) compiles
// against this header after changing the include.  The O(N^3) stages
// (ConvMat, SMatrixLayer, RedhefferProduct) run on the GPU; the O(N) ones on
// the host.  Matrix is a small owned column-major complex matrix when Eigen is
// absent (the reference's Matrix is Eigen's column-major dynamic matrix,
// rcwa.hpp:13); only the members the reference's driver uses are provided.
//
// Differences by design: errors throw rcwa_amd::Error (the reference aborts
// at rcwa.cpp:300) and nothing is printed.  SMatrixLayer takes W0/V0 for
// signature compatibility; they must be GapMedium's outputs for the same
// Kx, Ky, as at the reference's only call site (main.cpp:141, 151).
#pragma once

#include <cmath>
#include <complex>
#include <cstring>
#include <ostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "rcwa_amd.h"

using Real = double;
using Complex = std::complex<Real>;

namespace rcwa_amd {

struct Error : std::runtime_error {
  rcwa_status status;
  Error(rcwa_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check(rcwa_status s) {
  if (s != RCWA_OK) throw Error(s, rcwa_last_error());
}

// Column-major dense matrix (rows x cols).
template <typename T>
class Dense {
 public:
  Dense() = default;
  Dense(long r, long c) : r_(r), c_(c), a_((size_t)(r * c)) {}
  explicit Dense(long n) : r_(n), c_(1), a_((size_t)n) {}
  static Dense Zero(long r, long c) { return Dense(r, c); }
  static Dense Zero(long n) { return Dense(n); }
  static Dense Identity(long r, long c) {
    Dense m(r, c);
    for (long i = 0; i < std::min(r, c); ++i) m(i, i) = T(1);
    return m;
  }
  long rows() const { return r_; }
  long cols() const { return c_; }
  long size() const { return r_ * c_; }
  T& operator()(long i, long j) { return a_[(size_t)(i + j * r_)]; }
  const T& operator()(long i, long j) const { return a_[(size_t)(i + j * r_)]; }
  T& operator()(long i) { return a_[(size_t)i]; }
  const T& operator()(long i) const { return a_[(size_t)i]; }
  T* data() { return a_.data(); }
  const T* data() const { return a_.data(); }
  void resize(long r, long c) {
    r_ = r;
    c_ = c;
    a_.assign((size_t)(r * c), T(0));
  }
  T sum() const {
    T s(0);
    for (const T& v : a_) s += v;
    return s;
  }
  // Eigen's LinSpaced(size, low, high) for vectors
  static Dense LinSpaced(long n, T lo, T hi) {
    Dense v(n);
    for (long i = 0; i < n; ++i) v(i) = n > 1 ? lo + (hi - lo) * T((double)i / (n - 1)) : hi;
    return v;
  }
  Dense transpose() const {
    Dense t(c_, r_);
    for (long j = 0; j < c_; ++j)
      for (long i = 0; i < r_; ++i) t(j, i) = (*this)(i, j);
    return t;
  }
  Dense conjugate() const {
    Dense t(*this);
    for (T& v : t.a_) v = conj_(v);
    return t;
  }
  Dense adjoint() const { return transpose().conjugate(); }
  // copies of a block, a column and a row (reads; write blocks with set_block)
  Dense block(long i, long j, long r, long c) const {
    Dense b(r, c);
    for (long q = 0; q < c; ++q)
      for (long p = 0; p < r; ++p) b(p, q) = (*this)(i + p, j + q);
    return b;
  }
  void set_block(long i, long j, const Dense& b) {
    for (long q = 0; q < b.cols(); ++q)
      for (long p = 0; p < b.rows(); ++p) (*this)(i + p, j + q) = b(p, q);
  }
  Dense col(long j) const { return block(0, j, r_, 1); }
  Dense row(long i) const { return block(i, 0, 1, c_); }
  // Eigen's replicate(rows_factor, cols_factor)
  Dense replicate(long fr, long fc) const {
    Dense m(r_ * fr, c_ * fc);
    for (long bj = 0; bj < fc; ++bj)
      for (long bi = 0; bi < fr; ++bi) m.set_block(bi * r_, bj * c_, *this);
    return m;
  }
  double norm() const {
    double s = 0;
    for (const T& v : a_) s += std::norm(v);
    return std::sqrt(s);
  }
  Dense& operator+=(const Dense& b) {
    for (size_t i = 0; i < a_.size(); ++i) a_[i] += b.a_[i];
    return *this;
  }
  Dense& operator-=(const Dense& b) {
    for (size_t i = 0; i < a_.size(); ++i) a_[i] -= b.a_[i];
    return *this;
  }
  Dense& operator*=(T s) {
    for (T& v : a_) v *= s;
    return *this;
  }

 private:
  static T conj_(const T& v) {
    if constexpr (std::is_same<T, Complex>::value) return std::conj(v);
    else return v;
  }
  long r_ = 0, c_ = 0;
  std::vector<T> a_;
};

template <typename T>
std::ostream& operator<<(std::ostream& os, const Dense<T>& m) {
  for (long i = 0; i < m.rows(); ++i) {
    for (long j = 0; j < m.cols(); ++j) os << (j ? " " : "") << m(i, j);
    os << '\n';
  }
  return os;
}

// Process-wide default context, created on first use.
inline rcwa_ctx*& default_ctx() {
  static rcwa_ctx* c = nullptr;
  return c;
}
inline rcwa_ctx* ctx() {
  if (!default_ctx()) check(rcwa_ctx_create(0, 0, &default_ctx()));
  return default_ctx();
}
inline void set_context(rcwa_ctx* c) { default_ctx() = c; }

inline const rcwa_c128* C(const Complex* p) { return reinterpret_cast<const rcwa_c128*>(p); }
inline rcwa_c128* C(Complex* p) { return reinterpret_cast<rcwa_c128*>(p); }

template <typename T>
Dense<T> operator+(Dense<T> a, const Dense<T>& b) { return a += b; }
template <typename T>
Dense<T> operator-(Dense<T> a, const Dense<T>& b) { return a -= b; }
template <typename T>
Dense<T> operator-(Dense<T> a) { return a *= T(-1); }
template <typename T>
Dense<T> operator*(Dense<T> a, T s) { return a *= s; }
template <typename T>
Dense<T> operator*(T s, Dense<T> a) { return a *= s; }

// Matrix product (Eigen's operator*): complex products from 64 x 64 x 64
// upwards run on the GPU (rcwa_zgemm), smaller ones on the host.
template <typename T>
Dense<T> operator*(const Dense<T>& a, const Dense<T>& b) {
  if (a.cols() != b.rows()) throw Error(RCWA_E_ARG, "operator*: inner dimensions differ");
  Dense<T> c(a.rows(), b.cols());
  if constexpr (std::is_same<T, Complex>::value) {
    if (a.rows() * b.cols() * a.cols() >= 64L * 64 * 64) {
      rcwa_ctx* x = nullptr;
      if (!default_ctx()) check(rcwa_ctx_create(0, 0, &default_ctx()));
      x = default_ctx();
      check(rcwa_zgemm(x, (int)a.rows(), (int)b.cols(), (int)a.cols(), 1, C(a.data()),
                       C(b.data()), C(c.data())));
      return c;
    }
  }
  for (long j = 0; j < b.cols(); ++j)
    for (long k = 0; k < a.cols(); ++k) {
      const T bk = b(k, j);
      for (long i = 0; i < a.rows(); ++i) c(i, j) += a(i, k) * bk;
    }
  return c;
}

// Eigen's PartialPivLU (rcwa.cpp takes .lu() of every matrix it solves
// with): row pivoting on the largest magnitude per column, on the host.
template <typename T>
class PartialPivLU {
 public:
  explicit PartialPivLU(const Dense<T>& a) : f_(a), piv_(a.rows()) {
    if (a.rows() != a.cols()) throw Error(RCWA_E_ARG, "lu: matrix is not square");
    const long n = a.rows();
    for (long k = 0; k < n; ++k) {
      long p = k;
      for (long i = k + 1; i < n; ++i)
        if (std::abs(f_(i, k)) > std::abs(f_(p, k))) p = i;
      piv_[k] = p;
      if (p != k)
        for (long j = 0; j < n; ++j) std::swap(f_(k, j), f_(p, j));
      if (f_(k, k) == T(0)) continue;
      for (long i = k + 1; i < n; ++i) f_(i, k) /= f_(k, k);
      for (long j = k + 1; j < n; ++j)
        for (long i = k + 1; i < n; ++i) f_(i, j) -= f_(i, k) * f_(k, j);
    }
  }
  Dense<T> solve(const Dense<T>& b) const {
    const long n = f_.rows();
    if (b.rows() != n) throw Error(RCWA_E_ARG, "lu solve: shape");
    Dense<T> x(b);
    for (long j = 0; j < x.cols(); ++j) {
      for (long k = 0; k < n; ++k)
        if (piv_[k] != k) std::swap(x(k, j), x(piv_[k], j));
      for (long k = 0; k < n; ++k)
        for (long i = k + 1; i < n; ++i) x(i, j) -= f_(i, k) * x(k, j);
      for (long k = n - 1; k >= 0; --k) {
        x(k, j) /= f_(k, k);
        for (long i = 0; i < k; ++i) x(i, j) -= f_(i, k) * x(k, j);
      }
    }
    return x;
  }
  Dense<T> inverse() const { return solve(Dense<T>::Identity(f_.rows(), f_.rows())); }

 private:
  Dense<T> f_;
  std::vector<long> piv_;
};

template <typename T>
PartialPivLU<T> lu(const Dense<T>& a) { return PartialPivLU<T>(a); }

}  // namespace rcwa_amd

using Matrix = rcwa_amd::Dense<Complex>;
using Vector = rcwa_amd::Dense<Complex>;
using Real_Matrix = rcwa_amd::Dense<Real>;
using Real_Vector = rcwa_amd::Dense<Real>;

// rcwa.hpp:17: 3-vector of complex numbers (brace-initialised to zeros).
struct Vec_3d {
  Complex v[3]{};
  Complex& operator()(int i) { return v[i]; }
  const Complex& operator()(int i) const { return v[i]; }
  Complex& operator[](int i) { return v[i]; }
  const Complex& operator[](int i) const { return v[i]; }
  static constexpr long size() { return 3; }
  Complex* data() { return v; }
  const Complex* data() const { return v; }
  double norm() const { return std::sqrt(std::norm(v[0]) + std::norm(v[1]) + std::norm(v[2])); }
  Vec_3d normalized() const {
    const double n = norm();
    return Vec_3d{{v[0] / n, v[1] / n, v[2] / n}};
  }
  // Eigen's cross(): this x b
  Vec_3d cross(const Vec_3d& b) const {
    return Vec_3d{{v[1] * b.v[2] - v[2] * b.v[1], v[2] * b.v[0] - v[0] * b.v[2],
                   v[0] * b.v[1] - v[1] * b.v[0]}};
  }
};
inline Vec_3d operator+(const Vec_3d& a, const Vec_3d& b) {
  return Vec_3d{{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]}};
}
inline Vec_3d operator*(Real s, const Vec_3d& a) { return Vec_3d{{s * a.v[0], s * a.v[1], s * a.v[2]}}; }
inline Vec_3d operator*(Complex s, const Vec_3d& a) {
  return Vec_3d{{s * a.v[0], s * a.v[1], s * a.v[2]}};
}

// rcwa.hpp:24-28
inline Complex unsigned_sqrt(Complex z) {
  if (z.imag() == 0.0) z = Complex(z.real(), 0.0);
  return std::sqrt(z);
}

// rcwa.hpp:37-44: the elementwise form, for vectors and matrices.
inline Matrix unsigned_sqrt(const Matrix& m) {
  Matrix r(m.rows(), m.cols());
  for (long j = 0; j < m.cols(); ++j)
    for (long i = 0; i < m.rows(); ++i) r(i, j) = unsigned_sqrt(m(i, j));
  return r;
}

// rcwa.hpp:218-224: X = x^T repeated over y.size() rows, Y = y repeated over
// x.size() columns.
template <typename Scalar>
void MeshGrid(const rcwa_amd::Dense<Scalar>& x, const rcwa_amd::Dense<Scalar>& y,
              rcwa_amd::Dense<Scalar>& X, rcwa_amd::Dense<Scalar>& Y) {
  X = x.transpose().replicate(y.size(), 1);
  Y = y.replicate(1, x.size());
}

// rcwa.hpp:227-230
struct EigenvalSolverResults {
  Vector eigenvalues{};
  Matrix eigenvectors{};
};

namespace rcwa_amd {
// The eigensolver the reference takes from Eigen (ComplexEigenSolver,
// rcwa.cpp:298-306), on the GPU: eigenvalues and unit 2-norm eigenvectors of
// a general complex matrix.  Their order and phases differ from Eigen's
// (which sorts them); the S-matrices built from them do not depend on either.
template <typename MatrixType>
class ComplexEigenSolver {
 public:
  ComplexEigenSolver() = default;
  explicit ComplexEigenSolver(const MatrixType& a, bool computeEigenvectors = true) {
    compute(a, computeEigenvectors);
  }
  ComplexEigenSolver& compute(const MatrixType& a, bool computeEigenvectors = true) {
    (void)computeEigenvectors;
    if (a.rows() != a.cols()) throw Error(RCWA_E_ARG, "ComplexEigenSolver: not square");
    const long n = a.rows();
    w_ = Vector(n);
    v_ = Matrix(n, n);
    const rcwa_status st =
        rcwa_eig(ctx(), (int)n, 1, C(a.data()), C(w_.data()), C(v_.data()));
    ok_ = st == RCWA_OK;
    if (st != RCWA_OK && st != RCWA_E_EIG_NOCONV) check(st);
    return *this;
  }
  const Vector& eigenvalues() const { return w_; }
  const Matrix& eigenvectors() const { return v_; }
  bool converged() const { return ok_; }

 private:
  Vector w_;
  Matrix v_;
  bool ok_ = false;
};
}  // namespace rcwa_amd

// rcwa.hpp:233 / rcwa.cpp:617-621.  The reference aborts when the solver did
// not converge (rcwa.cpp:300); this throws rcwa_amd::Error(RCWA_E_EIG_NOCONV).
inline EigenvalSolverResults extraction(rcwa_amd::ComplexEigenSolver<Matrix>&& solver) {
  if (!solver.converged())
    throw rcwa_amd::Error(RCWA_E_EIG_NOCONV, "eigensolver QR did not converge");
  return EigenvalSolverResults{solver.eigenvalues(), solver.eigenvectors()};
}

// rcwa.hpp:48-78
struct Device {
  int Nx{}, Ny{}, num_layers{};
  Real Lx{}, Ly{};
  std::vector<Complex> er{}, ur{};  // [layer][iy][ix]
  std::vector<Real> t{};
  std::vector<Matrix> erc{}, urc{};
  Device() : Nx(1), Ny(1), num_layers(1), Lx(1.0), Ly(1.0) {}
  Device(int Nx_, int Ny_, int num_layers_, Real Lx_, Real Ly_, const std::vector<Complex>& er_,
         const std::vector<Complex>& ur_, const std::vector<Real>& t_, int Nx_harmonics_,
         int Ny_harmonics_)
      : Nx(Nx_), Ny(Ny_), num_layers(num_layers_), Lx(Lx_), Ly(Ly_), er(er_), ur(ur_), t(t_),
        erc(num_layers_, Matrix::Zero((2 * Ny_harmonics_ + 1) * (2 * Nx_harmonics_ + 1),
                                      (2 * Ny_harmonics_ + 1) * (2 * Nx_harmonics_ + 1))),
        urc(num_layers_, Matrix::Zero((2 * Ny_harmonics_ + 1) * (2 * Nx_harmonics_ + 1),
                                      (2 * Ny_harmonics_ + 1) * (2 * Nx_harmonics_ + 1))) {}
};

// rcwa.hpp:81-95
struct Source {
  Real lambda0{}, k0{}, theta{}, phi{}, pte{}, ptm{};
  Source(Real lambda0_, Real theta_, Real phi_, Real pte_, Real ptm_)
      : lambda0(lambda0_), k0(2 * M_PI / lambda0_), theta(theta_), phi(phi_), pte(pte_),
        ptm(ptm_) {}
};

// rcwa.hpp:98-112
struct RCWAParams {
  int Nx_harmonics{}, Ny_harmonics{};
  Complex er_ref{}, ur_ref{}, er_trn{}, ur_trn{};
  RCWAParams(int Nx_harmonics_, int Ny_harmonics_, Complex er_ref_, Complex ur_ref_,
             Complex er_trn_, Complex ur_trn_)
      : Nx_harmonics(Nx_harmonics_), Ny_harmonics(Ny_harmonics_), er_ref(er_ref_),
        ur_ref(ur_ref_), er_trn(er_trn_), ur_trn(ur_trn_) {}
};

// rcwa.hpp:116-128
struct Results {
  Real_Matrix R{}, T{};
  Real R_tot{}, T_tot{};
  Results(Real_Matrix R_, Real_Matrix T_, Real R_tot_, Real T_tot_)
      : R(R_), T(T_), R_tot(R_tot_), T_tot(T_tot_) {}
};

// rcwa.hpp:132-162
struct ScatteringMatrix {
  Matrix S11{}, S12{}, S21{}, S22{};
  ScatteringMatrix(Matrix S11_, Matrix S12_, Matrix S21_, Matrix S22_)
      : S11(S11_), S12(S12_), S21(S21_), S22(S22_) {}
};

namespace rcwa_amd {
inline rcwa_params to_params(const RCWAParams& p) {
  return rcwa_params{p.Nx_harmonics, p.Ny_harmonics, {p.er_ref.real(), p.er_ref.imag()},
                     {p.ur_ref.real(), p.ur_ref.imag()}, {p.er_trn.real(), p.er_trn.imag()},
                     {p.ur_trn.real(), p.ur_trn.imag()}};
}
inline bool is_identity(const Matrix& M) {
  for (long j = 0; j < M.cols(); ++j)
    for (long i = 0; i < M.rows(); ++i)
      if (M(i, j) != Complex(i == j ? 1.0 : 0.0, 0.0)) return false;
  return true;
}
}  // namespace rcwa_amd

// rcwa.hpp:172 (rcwa.cpp:6-82) -- GPU
inline Matrix ConvMat(const std::vector<Complex>& field, int layer, int Nx, int Ny,
                      int Nx_harmonics, int Ny_harmonics) {
  const long PQ = (2L * Nx_harmonics + 1) * (2L * Ny_harmonics + 1);
  Matrix C(PQ, PQ);
  rcwa_amd::check(rcwa_convmat(rcwa_amd::ctx(), rcwa_amd::C(field.data()), layer, Nx, Ny,
                               Nx_harmonics, Ny_harmonics, rcwa_amd::C(C.data())));
  return C;
}

// rcwa.hpp:175 (rcwa.cpp:85-160) -- host
inline void ComputeWaveVectors(const Device& device, const Source& source,
                               const RCWAParams& params, std::vector<Complex>& k_inc, Vector& Kx,
                               Vector& Ky, Vector& Kz_ref, Vector& Kz_trn) {
  const long PQ = (2L * params.Nx_harmonics + 1) * (2L * params.Ny_harmonics + 1);
  k_inc.resize(3);
  Kx = Vector(PQ);
  Ky = Vector(PQ);
  Kz_ref = Vector(PQ);
  Kz_trn = Vector(PQ);
  const rcwa_source s{source.lambda0, source.theta, source.phi, source.pte, source.ptm};
  const rcwa_params p = rcwa_amd::to_params(params);
  rcwa_amd::check(rcwa_wave_vectors(device.Lx, device.Ly, &s, &p, rcwa_amd::C(k_inc.data()),
                                    rcwa_amd::C(Kx.data()), rcwa_amd::C(Ky.data()),
                                    rcwa_amd::C(Kz_ref.data()), rcwa_amd::C(Kz_trn.data())));
}

// rcwa.hpp:178 (rcwa.cpp:163-234) -- host
inline void GapMedium(const Vector& Kx, const Vector& Ky, Matrix& W0, Matrix& V0) {
  const long n = 2 * Kx.size();
  W0 = Matrix(n, n);
  V0 = Matrix(n, n);
  rcwa_amd::check(rcwa_gap_medium((int)Kx.size(), rcwa_amd::C(Kx.data()), rcwa_amd::C(Ky.data()),
                                  rcwa_amd::C(W0.data()), rcwa_amd::C(V0.data())));
}

// rcwa.hpp:181 (rcwa.cpp:237-250)
inline ScatteringMatrix SMatrixInit(int Nx_harmonics, int Ny_harmonics) {
  const long n = 2 * (2L * Nx_harmonics + 1) * (2L * Ny_harmonics + 1);
  return ScatteringMatrix(Matrix::Zero(n, n), Matrix::Identity(n, n), Matrix::Identity(n, n),
                          Matrix::Zero(n, n));
}

// rcwa.hpp:184 (rcwa.cpp:254-352) -- GPU
inline ScatteringMatrix SMatrixLayer(int layer, const Device& device, const Source& source,
                                     const RCWAParams& params, const Vector& Kx, const Vector& Ky,
                                     const Matrix& W0, const Matrix& V0) {
  (void)params;
  (void)W0;
  (void)V0;
  const Matrix& erc = device.erc.at(layer);
  const Matrix& urc = device.urc.at(layer);
  const long n = 2 * erc.rows();
  Matrix S11(n, n), S12(n, n);
  const double k0t = source.k0 * device.t.at(layer);
  rcwa_amd::check(rcwa_layer_smatrix(rcwa_amd::ctx(), (int)erc.rows(), 1, rcwa_amd::C(erc.data()),
                                     rcwa_amd::C(urc.data()), rcwa_amd::C(Kx.data()),
                                     rcwa_amd::C(Ky.data()), &k0t, rcwa_amd::C(S11.data()),
                                     rcwa_amd::C(S12.data())));
  return ScatteringMatrix(S11, S12, S12, S11);
}

namespace rcwa_amd {
inline ScatteringMatrix end_smatrix(int which, const RCWAParams& params, const Vector& Kx,
                                    const Vector& Ky, const Vector& Kz, Matrix& W) {
  const long n = 2 * Kx.size();
  Matrix a(n, n), b(n, n), c(n, n), d(n, n);
  const rcwa_params p = to_params(params);
  check(rcwa_end_smatrix(which, (int)Kx.size(), &p, C(Kx.data()), C(Ky.data()), C(Kz.data()),
                         C(a.data()), C(b.data()), C(c.data()), C(d.data())));
  W = Matrix::Identity(n, n);
  return ScatteringMatrix(a, b, c, d);
}
}  // namespace rcwa_amd

// rcwa.hpp:190 (rcwa.cpp:354-411) -- host, closed form
inline ScatteringMatrix SMatrixReflection(const RCWAParams& params, const Vector& Kx,
                                          const Vector& Ky, Vector& Kz_ref, const Matrix& W0,
                                          const Matrix& V0, Matrix& W_ref) {
  (void)W0;
  (void)V0;
  return rcwa_amd::end_smatrix(0, params, Kx, Ky, Kz_ref, W_ref);
}

// rcwa.hpp:193 (rcwa.cpp:414-466) -- host, closed form
inline ScatteringMatrix SMatrixTransmission(const RCWAParams& params, const Vector& Kx,
                                            const Vector& Ky, Vector& Kz_trn, const Matrix& W0,
                                            const Matrix& V0, Matrix& W_trn) {
  (void)W0;
  (void)V0;
  return rcwa_amd::end_smatrix(1, params, Kx, Ky, Kz_trn, W_trn);
}

// rcwa.hpp:214 (rcwa.cpp:588-615) -- GPU
inline ScatteringMatrix RedhefferProduct(const ScatteringMatrix& A, const ScatteringMatrix& B) {
  const long n = A.S11.rows();
  ScatteringMatrix S(Matrix(n, n), Matrix(n, n), Matrix(n, n), Matrix(n, n));
  const rcwa_c128* a[4] = {rcwa_amd::C(A.S11.data()), rcwa_amd::C(A.S12.data()),
                           rcwa_amd::C(A.S21.data()), rcwa_amd::C(A.S22.data())};
  const rcwa_c128* b[4] = {rcwa_amd::C(B.S11.data()), rcwa_amd::C(B.S12.data()),
                           rcwa_amd::C(B.S21.data()), rcwa_amd::C(B.S22.data())};
  rcwa_c128* s[4] = {rcwa_amd::C(S.S11.data()), rcwa_amd::C(S.S12.data()),
                     rcwa_amd::C(S.S21.data()), rcwa_amd::C(S.S22.data())};
  rcwa_amd::check(rcwa_star_product(rcwa_amd::ctx(), (int)n, 1, a, b, s));
  return S;
}

// rcwa.hpp:200 (rcwa.cpp:469-523)
inline void ComputeSourceModeCoeff(const Source& source, const RCWAParams& params,
                                   const std::vector<Complex>& k_inc, const Matrix& W_ref,
                                   Vector& csrc) {
  Complex aTE[3];
  if (std::abs(source.theta) < 1e-6) {
    aTE[0] = 0.0; aTE[1] = 1.0; aTE[2] = 0.0;
  } else {
    aTE[0] = -k_inc.at(1); aTE[1] = k_inc.at(0); aTE[2] = 0.0;
  }
  auto normalize = [](Complex* v) {
    const double nn = std::sqrt(std::norm(v[0]) + std::norm(v[1]) + std::norm(v[2]));
    for (int i = 0; i < 3; ++i) v[i] /= nn;
  };
  normalize(aTE);
  Complex aTM[3] = {k_inc[1] * aTE[2] - k_inc[2] * aTE[1], k_inc[2] * aTE[0] - k_inc[0] * aTE[2],
                    k_inc[0] * aTE[1] - k_inc[1] * aTE[0]};
  normalize(aTM);
  Complex pol[3];
  for (int i = 0; i < 3; ++i) pol[i] = source.pte * aTE[i] + source.ptm * aTM[i];
  normalize(pol);
  const long PQ = (2L * params.Nx_harmonics + 1) * (2L * params.Ny_harmonics + 1);
  Vector esrc = Vector::Zero(2 * PQ);
  esrc(PQ / 2) = pol[0];
  esrc(PQ / 2 + PQ) = pol[1];
  if (rcwa_amd::is_identity(W_ref)) {
    csrc = esrc;
    return;
  }
  const long n = W_ref.rows();
  Matrix Wi(n, n);
  rcwa_amd::check(rcwa_inverse(rcwa_amd::ctx(), (int)n, 1, rcwa_amd::C(W_ref.data()),
                               rcwa_amd::C(Wi.data())));
  csrc = Vector(n);
  for (long j = 0; j < n; ++j)
    for (long i = 0; i < n; ++i) csrc(i) += Wi(i, j) * esrc(j);
}

namespace rcwa_amd {
inline void field(const Matrix& S, const Vector& csrc, const Vector& Kx, const Vector& Ky,
                  const Vector& Kz, const Matrix& W, Vector& r) {
  const long n = S.rows(), PQ = Kx.size();
  Vector e(n), f(n);
  for (long j = 0; j < n; ++j)
    if (csrc(j) != Complex(0, 0))
      for (long i = 0; i < n; ++i) e(i) += S(i, j) * csrc(j);
  for (long j = 0; j < n; ++j)
    for (long i = 0; i < n; ++i) f(i) += W(i, j) * e(j);
  for (long i = 0; i < n; ++i) r(i) = f(i);
  for (long i = 0; i < PQ; ++i) r(n + i) = -(Kx(i) * r(i) + Ky(i) * r(PQ + i)) / Kz(i);
}
}  // namespace rcwa_amd

// rcwa.hpp:204 (rcwa.cpp:526-542); r must be pre-sized to 3PQ
inline void ComputeReflectedField(const RCWAParams& params, const ScatteringMatrix& S_global,
                                  const Vector& csrc, const Vector& Kx, const Vector& Ky,
                                  const Vector& Kz_ref, const Matrix& W_ref, Vector& r) {
  (void)params;
  rcwa_amd::field(S_global.S11, csrc, Kx, Ky, Kz_ref, W_ref, r);
}

// rcwa.hpp:206 (rcwa.cpp:544-561); t must be pre-sized to 3PQ
inline void ComputeTransmittedField(const RCWAParams& params, const ScatteringMatrix& S_global,
                                    const Vector& csrc, const Vector& Kx, const Vector& Ky,
                                    const Vector& Kz_trn, const Matrix& W_trn, Vector& t) {
  (void)params;
  rcwa_amd::field(S_global.S21, csrc, Kx, Ky, Kz_trn, W_trn, t);
}

// rcwa.hpp:209 (rcwa.cpp:564-585)
inline Results ComputeDiffractionEfficiencies(const RCWAParams& params, const Vector& r,
                                              const Vector& t, const std::vector<Complex>& k_inc,
                                              const Vector& Kz_ref, const Vector& Kz_trn) {
  const long P = 2L * params.Nx_harmonics + 1, Q = 2L * params.Ny_harmonics + 1, PQ = P * Q;
  Real_Matrix R(Q, P), T(Q, P);
  const double den = (k_inc.at(2) / params.ur_ref).real();
  for (long m = 0; m < PQ; ++m) {
    const double rr = std::norm(r(m)) + std::norm(r(PQ + m)) + std::norm(r(2 * PQ + m));
    const double tt = std::norm(t(m)) + std::norm(t(PQ + m)) + std::norm(t(2 * PQ + m));
    R(m) = (-Kz_ref(m) / params.ur_ref).real() / den * rr;
    T(m) = (Kz_trn(m) / params.ur_trn).real() / den * tt;
  }
  return Results(R, T, R.sum(), T.sum());
}

// The whole of main.cpp:126-191 on the GPU in one call.
inline Results Solve(const Device& device, const Source& source, const RCWAParams& params) {
  const rcwa_device d{device.Nx, device.Ny, device.num_layers, device.Lx, device.Ly,
                      rcwa_amd::C(device.er.data()), rcwa_amd::C(device.ur.data()),
                      device.t.data(), 0};
  const rcwa_source s{source.lambda0, source.theta, source.phi, source.pte, source.ptm};
  const rcwa_params p = rcwa_amd::to_params(params);
  const long P = 2L * params.Nx_harmonics + 1, Q = 2L * params.Ny_harmonics + 1;
  Real_Matrix R(Q, P), T(Q, P);
  rcwa_results out{R.data(), T.data(), 0.0, 0.0};
  rcwa_amd::check(rcwa_solve(rcwa_amd::ctx(), &d, &p, &s, 1, RCWA_FP64, &out));
  return Results(R, T, out.R_tot, out.T_tot);
}
