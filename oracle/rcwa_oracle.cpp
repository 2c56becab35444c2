// CPU FP64 oracle for the RCWA solve path.
//
// TEST INFRASTRUCTURE ONLY: this file is the checker.  Only the test suite,
// the smoke check and the CPU-baseline leg of the benchmark load it (through
// ctypes, as oracle/liboracle.so).  The product library never links it.
//
// It is an op-for-op restatement, written without Eigen, of rcwa3d:
//   rcwa3d/include/rcwa.hpp:11-17   scalar/matrix types (column-major complex128)
//   rcwa3d/include/rcwa.hpp:24-44   unsigned_sqrt
//   rcwa3d/src/rcwa.cpp:6-621       every stage function
//   rcwa3d/src/main.cpp:126-191     the call order of one solve
// The linear algebra the reference takes from Eigen (PartialPivLU, GEMM,
// ComplexEigenSolver, kissfft) is replaced by textbook algorithms written
// here: partial-pivot LU, Householder Hessenberg reduction, single-shift
// complex QR with Wilkinson shifts (complex Schur form), triangular
// eigenvectors by back-substitution, and a radix-2 / direct DFT.
//
// How each part is pinned (closed forms, invariants, LAPACK cross-check) is
// stated in tests/test_oracle.py.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <vector>

namespace oracle {

using Real = double;
using Complex = std::complex<double>;

// Column-major dense complex matrix (rcwa.hpp:13 uses Eigen's default
// column-major storage; this keeps the same element order).
struct Matrix {
  int rows = 0, cols = 0;
  std::vector<Complex> a;
  Matrix() {}
  Matrix(int r, int c) : rows(r), cols(c), a((size_t)r * c, Complex(0, 0)) {}
  Complex& operator()(int i, int j) { return a[(size_t)j * rows + i]; }
  const Complex& operator()(int i, int j) const { return a[(size_t)j * rows + i]; }
  Complex* col(int j) { return a.data() + (size_t)j * rows; }
  const Complex* col(int j) const { return a.data() + (size_t)j * rows; }
  static Matrix Identity(int n) {
    Matrix m(n, n);
    for (int i = 0; i < n; ++i) m(i, i) = 1.0;
    return m;
  }
};
using Vector = std::vector<Complex>;

struct OracleError : std::runtime_error {
  int code;
  OracleError(int c, const char* m) : std::runtime_error(m), code(c) {}
};

// rcwa.hpp:24-28: an imaginary part of -0 is replaced by +0 before the
// principal square root, so sqrt(-x - 0i) = +i sqrt(x).
static inline Complex unsigned_sqrt(Complex z) {
  if (z.imag() == 0.0) z = Complex(z.real(), 0.0);
  return std::sqrt(z);
}

// ---------------------------------------------------------------- BLAS-like
static Matrix matmul(const Matrix& A, const Matrix& B) {
  if (A.cols != B.rows) throw OracleError(1, "matmul shape");
  Matrix C(A.rows, B.cols);
  const int m = A.rows, kk = A.cols, n = B.cols;
  const int KB = 64;
  for (int k0 = 0; k0 < kk; k0 += KB) {
    int k1 = std::min(kk, k0 + KB);
    for (int j = 0; j < n; ++j) {
      Complex* c = C.col(j);
      for (int k = k0; k < k1; ++k) {
        const Complex b = B(k, j);
        if (b == Complex(0, 0)) continue;
        const Complex* a = A.col(k);
        for (int i = 0; i < m; ++i) c[i] += a[i] * b;
      }
    }
  }
  return C;
}

static Matrix add(const Matrix& A, const Matrix& B, double sb = 1.0) {
  Matrix C = A;
  for (size_t i = 0; i < C.a.size(); ++i) C.a[i] += sb * B.a[i];
  return C;
}

static Matrix scale(const Matrix& A, Complex s) {
  Matrix C = A;
  for (auto& v : C.a) v *= s;
  return C;
}

static Matrix transpose(const Matrix& A) {
  Matrix T(A.cols, A.rows);
  for (int j = 0; j < A.cols; ++j)
    for (int i = 0; i < A.rows; ++i) T(j, i) = A(i, j);
  return T;
}

// diag(d) * M  and  M * diag(d)
static Matrix diag_left(const Vector& d, const Matrix& M) {
  Matrix C = M;
  for (int j = 0; j < M.cols; ++j)
    for (int i = 0; i < M.rows; ++i) C(i, j) *= d[i];
  return C;
}
static Matrix diag_right(const Matrix& M, const Vector& d) {
  Matrix C = M;
  for (int j = 0; j < M.cols; ++j)
    for (int i = 0; i < M.rows; ++i) C(i, j) *= d[j];
  return C;
}
static Matrix diag_dense(const Vector& d) {
  Matrix C((int)d.size(), (int)d.size());
  for (size_t i = 0; i < d.size(); ++i) C((int)i, (int)i) = d[i];
  return C;
}

// --------------------------------------------------------------- LU (partial pivoting)
// The reference's `.lu()` is Eigen's PartialPivLU (row pivoting on the
// largest-magnitude entry of each column).
struct LU {
  Matrix f;
  std::vector<int> piv;
  explicit LU(const Matrix& A) : f(A), piv(A.rows) {
    if (A.rows != A.cols) throw OracleError(1, "lu: not square");
    const int n = A.rows;
    for (int k = 0; k < n; ++k) {
      int p = k;
      double best = std::norm(f(k, k));
      for (int i = k + 1; i < n; ++i) {
        double v = std::norm(f(i, k));
        if (v > best) { best = v; p = i; }
      }
      piv[k] = p;
      if (p != k)
        for (int j = 0; j < n; ++j) std::swap(f(k, j), f(p, j));
      Complex d = f(k, k);
      if (d == Complex(0, 0)) continue;  // singular column: leave zeros (inf/nan follow, as in Eigen)
      Complex inv = 1.0 / d;
      Complex* ck = f.col(k);
      for (int i = k + 1; i < n; ++i) ck[i] *= inv;
      for (int j = k + 1; j < n; ++j) {
        Complex* cj = f.col(j);
        const Complex u = cj[k];
        if (u == Complex(0, 0)) continue;
        for (int i = k + 1; i < n; ++i) cj[i] -= ck[i] * u;
      }
    }
  }
  Matrix solve(const Matrix& B) const {
    const int n = f.rows;
    if (B.rows != n) throw OracleError(1, "lu solve shape");
    Matrix X = B;
    for (int j = 0; j < X.cols; ++j) {
      Complex* x = X.col(j);
      for (int k = 0; k < n; ++k)
        if (piv[k] != k) std::swap(x[k], x[piv[k]]);
      for (int k = 0; k < n; ++k) {          // L (unit lower)
        const Complex v = x[k];
        if (v == Complex(0, 0)) continue;
        const Complex* c = f.col(k);
        for (int i = k + 1; i < n; ++i) x[i] -= c[i] * v;
      }
      for (int k = n - 1; k >= 0; --k) {     // U
        x[k] /= f(k, k);
        const Complex v = x[k];
        if (v == Complex(0, 0)) continue;
        const Complex* c = f.col(k);
        for (int i = 0; i < k; ++i) x[i] -= c[i] * v;
      }
    }
    return X;
  }
  Vector solve(const Vector& b) const {
    Matrix B((int)b.size(), 1);
    for (size_t i = 0; i < b.size(); ++i) B.a[i] = b[i];
    Matrix X = solve(B);
    return X.a;
  }
};

// --------------------------------------------------------------- FFT
// Forward transform sum_n x[n] e^{-2 pi i k n / N}, unscaled (the Eigen::FFT
// convention used at rcwa.cpp:24-38).  Radix-2 when N is a power of two,
// direct summation otherwise.
static void dft_inplace(std::vector<Complex>& x) {
  const size_t N = x.size();
  if (N <= 1) return;
  if ((N & (N - 1)) == 0) {
    for (size_t i = 1, j = 0; i < N; ++i) {
      size_t bit = N >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) std::swap(x[i], x[j]);
    }
    for (size_t len = 2; len <= N; len <<= 1) {
      const double ang = -2.0 * M_PI / (double)len;
      for (size_t i = 0; i < N; i += len)
        for (size_t k = 0; k < len / 2; ++k) {
          Complex w(std::cos(ang * k), std::sin(ang * k));
          Complex u = x[i + k], v = x[i + k + len / 2] * w;
          x[i + k] = u + v;
          x[i + k + len / 2] = u - v;
        }
    }
  } else {
    std::vector<Complex> y(N);
    for (size_t k = 0; k < N; ++k) {
      Complex s(0, 0);
      for (size_t n = 0; n < N; ++n) {
        const double ang = -2.0 * M_PI * (double)((k * n) % N) / (double)N;
        s += x[n] * Complex(std::cos(ang), std::sin(ang));
      }
      y[k] = s;
    }
    x = y;
  }
}

// --------------------------------------------------------------- eigensolver
// Replaces Eigen::ComplexEigenSolver (rcwa.cpp:298): Hessenberg reduction,
// complex Schur form by implicitly shifted QR, eigenvectors of the triangular
// factor by back-substitution, back-transformation, unit 2-norm columns.
// S-matrices depend on W only through span/scale-invariant combinations, so
// eigenvector order and normalisation need not match Eigen's.
static inline double abs1(Complex z) { return std::fabs(z.real()) + std::fabs(z.imag()); }

struct EigResult {
  Vector values;
  Matrix vectors;
  int iterations = 0;
};

// Householder reflector H = I - tau v v^H, v[0] = 1, with H^H x = beta e1.
static void make_reflector(Complex* x, int len, Complex& tau, Complex& beta) {
  Complex alpha = x[0];
  double xn2 = 0;
  for (int i = 1; i < len; ++i) xn2 += std::norm(x[i]);
  if (xn2 == 0.0 && alpha.imag() == 0.0) {
    tau = 0.0;
    beta = alpha;
    return;
  }
  double nrm = std::sqrt(std::norm(alpha) + xn2);
  double b = (alpha.real() >= 0) ? -nrm : nrm;
  tau = Complex((b - alpha.real()) / b, -alpha.imag() / b);
  Complex s = 1.0 / (alpha - b);
  for (int i = 1; i < len; ++i) x[i] *= s;
  x[0] = 1.0;
  beta = b;
}

static void hessenberg(Matrix& A, Matrix& Qm) {
  const int n = A.rows;
  Qm = Matrix::Identity(n);
  std::vector<Complex> v(n), w(n);
  for (int k = 0; k + 2 < n; ++k) {
    const int len = n - k - 1;
    for (int i = 0; i < len; ++i) v[i] = A(k + 1 + i, k);
    Complex tau, beta;
    make_reflector(v.data(), len, tau, beta);
    if (tau == Complex(0, 0)) continue;
    // right: A(:, k+1:) -= tau (A(:,k+1:) v) v^H
    for (int i = 0; i < n; ++i) w[i] = 0;
    for (int j = 0; j < len; ++j) {
      const Complex vj = v[j];
      const Complex* c = A.col(k + 1 + j);
      for (int i = 0; i < n; ++i) w[i] += c[i] * vj;
    }
    for (int j = 0; j < len; ++j) {
      const Complex f = tau * std::conj(v[j]);
      Complex* c = A.col(k + 1 + j);
      for (int i = 0; i < n; ++i) c[i] -= w[i] * f;
    }
    // left: A(k+1:, k:) -= conj(tau) v (v^H A(k+1:, k:))
    for (int j = k; j < n; ++j) {
      Complex* c = A.col(j) + k + 1;
      Complex d(0, 0);
      for (int i = 0; i < len; ++i) d += std::conj(v[i]) * c[i];
      d *= std::conj(tau);
      for (int i = 0; i < len; ++i) c[i] -= v[i] * d;
    }
    A(k + 1, k) = beta;
    for (int i = k + 2; i < n; ++i) A(i, k) = 0;
    // Q := Q H
    for (int i = 0; i < n; ++i) w[i] = 0;
    for (int j = 0; j < len; ++j) {
      const Complex* c = Qm.col(k + 1 + j);
      for (int i = 0; i < n; ++i) w[i] += c[i] * v[j];
    }
    for (int j = 0; j < len; ++j) {
      const Complex f = tau * std::conj(v[j]);
      Complex* c = Qm.col(k + 1 + j);
      for (int i = 0; i < n; ++i) c[i] -= w[i] * f;
    }
  }
}

// Givens G = [c s; -conj(s) c] with G [f; g] = [r; 0], c real.
static inline void givens(Complex f, Complex g, double& c, Complex& s, Complex& r) {
  const double af = std::abs(f), ag = std::abs(g);
  if (ag == 0.0) { c = 1.0; s = 0.0; r = f; return; }
  if (af == 0.0) { c = 0.0; s = std::conj(g) / ag; r = ag; return; }
  const double nr = std::hypot(af, ag);
  const Complex ph = f / af;
  c = af / nr;
  s = ph * std::conj(g) / nr;
  r = ph * nr;
}

// Wilkinson shift: the eigenvalue of [[a b][c d]] nearer to d.
static Complex wilkinson(Complex a, Complex b, Complex c, Complex d) {
  Complex x = 0.5 * (a - d);
  Complex bc = b * c;
  Complex y = std::sqrt(x * x + bc);
  if ((x.real() * y.real() + x.imag() * y.imag()) < 0) y = -y;
  Complex den = x + y;
  if (den == Complex(0, 0)) return d;
  return d - bc / den;
}

static EigResult eig(const Matrix& A0, int max_iter_per_row = 30) {
  const int n = A0.rows;
  EigResult res;
  Matrix T = A0, Z;
  hessenberg(T, Z);
  const double eps = std::numeric_limits<double>::epsilon();
  int iu = n - 1, iter = 0, total = 0;
  const int maxit = max_iter_per_row * std::max(n, 1);
  while (iu > 0) {
    int il = iu;
    while (il > 0) {
      const double sub = abs1(T(il, il - 1));
      const double tst = abs1(T(il - 1, il - 1)) + abs1(T(il, il));
      if (sub <= eps * tst || sub < std::numeric_limits<double>::min()) {
        T(il, il - 1) = 0;
        break;
      }
      --il;
    }
    if (il == iu) { --iu; iter = 0; continue; }
    ++iter; ++total;
    if (total > maxit) throw OracleError(3, "eig: QR did not converge");
    Complex mu;
    if (iter == 10) mu = T(il, il) + 0.75 * std::fabs(T(il + 1, il).real());
    else if (iter == 20) mu = T(iu, iu) + 0.75 * std::fabs(T(iu, iu - 1).real());
    else mu = wilkinson(T(iu - 1, iu - 1), T(iu - 1, iu), T(iu, iu - 1), T(iu, iu));
    Complex f = T(il, il) - mu, g = T(il + 1, il);
    for (int k = il; k < iu; ++k) {
      double c; Complex s, r;
      if (k > il) { f = T(k, k - 1); g = T(k + 1, k - 1); }
      givens(f, g, c, s, r);
      if (k > il) { T(k, k - 1) = r; T(k + 1, k - 1) = 0; }
      for (int j = k; j < n; ++j) {           // rows k, k+1
        const Complex x = T(k, j), y = T(k + 1, j);
        T(k, j) = c * x + s * y;
        T(k + 1, j) = -std::conj(s) * x + c * y;
      }
      const int ie = std::min(k + 2, iu);
      for (int i = 0; i <= ie; ++i) {         // cols k, k+1
        const Complex x = T(i, k), y = T(i, k + 1);
        T(i, k) = c * x + std::conj(s) * y;
        T(i, k + 1) = -s * x + c * y;
      }
      for (int i = 0; i < n; ++i) {
        const Complex x = Z(i, k), y = Z(i, k + 1);
        Z(i, k) = c * x + std::conj(s) * y;
        Z(i, k + 1) = -s * x + c * y;
      }
    }
  }
  res.iterations = total;
  res.values.resize(n);
  for (int i = 0; i < n; ++i) res.values[i] = T(i, i);
  // Eigenvectors of T: T X = X diag(T), X unit upper triangular.
  double tnorm = 0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= j; ++i) tnorm = std::max(tnorm, abs1(T(i, j)));
  const double smlnum = std::numeric_limits<double>::min() * (n / eps);
  Matrix X(n, n);
  for (int k = n - 1; k >= 0; --k) {
    const Complex lk = T(k, k);
    const double smin = std::max(eps * abs1(lk), smlnum);
    Complex* x = X.col(k);
    x[k] = 1.0;
    for (int i = 0; i < k; ++i) x[i] = -T(i, k);
    for (int i = k - 1; i >= 0; --i) {
      Complex d = T(i, i) - lk;
      if (abs1(d) < smin) d = smin;
      x[i] /= d;
      const Complex xi = x[i];
      if (xi == Complex(0, 0)) continue;
      const Complex* tc = T.col(i);
      for (int m = 0; m < i; ++m) x[m] -= tc[m] * xi;
    }
  }
  (void)tnorm;
  res.vectors = matmul(Z, X);
  for (int k = 0; k < n; ++k) {
    double s = 0;
    Complex* c = res.vectors.col(k);
    for (int i = 0; i < n; ++i) s += std::norm(c[i]);
    s = 1.0 / std::sqrt(s);
    for (int i = 0; i < n; ++i) c[i] *= s;
  }
  return res;
}

// --------------------------------------------------------------- RCWA stages
struct ScatteringMatrix {
  Matrix S11, S12, S21, S22;
};

// rcwa.cpp:6-82
static Matrix ConvMat(const Complex* layer_grid, int Nx, int Ny, int hx, int hy) {
  // F is Ny x Nx; the layer slice is row-major [iy][ix] (rcwa.cpp:18).
  std::vector<Complex> F((size_t)Nx * Ny);
  for (int iy = 0; iy < Ny; ++iy)
    for (int ix = 0; ix < Nx; ++ix) F[(size_t)iy * Nx + ix] = layer_grid[(size_t)iy * Nx + ix];
  std::vector<Complex> line;
  line.resize(Nx);
  for (int iy = 0; iy < Ny; ++iy) {                      // rows (rcwa.cpp:27-31)
    for (int ix = 0; ix < Nx; ++ix) line[ix] = F[(size_t)iy * Nx + ix];
    dft_inplace(line);
    for (int ix = 0; ix < Nx; ++ix) F[(size_t)iy * Nx + ix] = line[ix];
  }
  line.resize(Ny);
  for (int ix = 0; ix < Nx; ++ix) {                      // columns (rcwa.cpp:34-38)
    for (int iy = 0; iy < Ny; ++iy) line[iy] = F[(size_t)iy * Nx + ix];
    dft_inplace(line);
    for (int iy = 0; iy < Ny; ++iy) F[(size_t)iy * Nx + ix] = line[iy];
  }
  const double inv = 1.0 / (double)((long long)Nx * Ny); // rcwa.cpp:41
  const int hX = Nx / 2, hY = Ny / 2;
  // fftshift (rcwa.cpp:44-53): shifted index j holds frequency j - N/2,
  // i.e. unshifted index (j - N/2) mod N.
  auto fshift = [&](int ky, int kx) {
    int uy = ((ky - hY) % Ny + Ny) % Ny;
    int ux = ((kx - hX) % Nx + Nx) % Nx;
    return F[(size_t)uy * Nx + ux] * inv;
  };
  const int P = 2 * hx + 1, Q = 2 * hy + 1, PQ = P * Q;
  Matrix C(PQ, PQ);
  for (int m = 0; m < PQ; ++m) {                         // rcwa.cpp:61-78
    const int pxm = m / Q - hx, pym = m % Q - hy;
    for (int nn = 0; nn < PQ; ++nn) {
      const int pxn = nn / Q - hx, pyn = nn % Q - hy;
      const int kx = (pxm - pxn) + hX, ky = (pym - pyn) + hY;
      if (kx >= 0 && kx < Nx && ky >= 0 && ky < Ny) C(m, nn) = fshift(ky, kx);
    }
  }
  return C;
}

struct Params {
  int hx, hy;
  Complex er_ref, ur_ref, er_trn, ur_trn;
};
struct Source {
  double lambda0, k0, theta, phi, pte, ptm;
};

// rcwa.cpp:85-160
static void ComputeWaveVectors(double Lx, double Ly, const Source& s, const Params& p,
                               Vector& k_inc, Vector& Kx, Vector& Ky, Vector& Kz_ref,
                               Vector& Kz_trn) {
  const Complex n_inc = std::sqrt(p.er_ref * p.ur_ref);  // plain sqrt, rcwa.cpp:91
  const double st = std::sin(s.theta), ct = std::cos(s.theta);
  const double sp = std::sin(s.phi), cp = std::cos(s.phi);
  k_inc.assign(3, 0);
  k_inc[0] = n_inc * st * cp;
  k_inc[1] = n_inc * st * sp;
  k_inc[2] = n_inc * ct;
  const int P = 2 * p.hx + 1, Q = 2 * p.hy + 1;
  Vector kx(P), ky(Q);
  for (int i = 0; i < P; ++i) kx[i] = k_inc[0] - (double)(i - p.hx) * (s.lambda0 / Lx);
  for (int i = 0; i < Q; ++i) ky[i] = k_inc[1] - (double)(i - p.hy) * (s.lambda0 / Ly);
  Kx.resize(P * Q); Ky.resize(P * Q); Kz_ref.resize(P * Q); Kz_trn.resize(P * Q);
  const Complex er_ = std::conj(p.ur_ref) * std::conj(p.er_ref);
  const Complex et_ = std::conj(p.ur_trn) * std::conj(p.er_trn);
  for (int ix = 0; ix < P; ++ix)                         // MeshGrid + column-major flatten
    for (int iy = 0; iy < Q; ++iy) {
      const int m = ix * Q + iy;
      Kx[m] = kx[ix];
      Ky[m] = ky[iy];
      const Complex k2 = Kx[m] * Kx[m] + Ky[m] * Ky[m];
      Kz_ref[m] = -std::conj(unsigned_sqrt(er_ - k2));   // rcwa.cpp:134
      Kz_trn[m] = std::conj(unsigned_sqrt(et_ - k2));    // rcwa.cpp:135
    }
}

static Matrix blocks(const Matrix& a, const Matrix& b, const Matrix& c, const Matrix& d) {
  const int N = a.rows;
  Matrix M(2 * N, 2 * N);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) {
      M(i, j) = a(i, j);
      M(i, j + N) = b(i, j);
      M(i + N, j) = c(i, j);
      M(i + N, j + N) = d(i, j);
    }
  return M;
}

static Matrix q_half(const Vector& Kx, const Vector& Ky, Complex er_ur) {
  const int N = (int)Kx.size();
  Matrix Q(2 * N, 2 * N);
  for (int i = 0; i < N; ++i) {
    Q(i, i) = Kx[i] * Ky[i];
    Q(i, i + N) = er_ur - Kx[i] * Kx[i];
    Q(i + N, i) = Ky[i] * Ky[i] - er_ur;
    Q(i + N, i + N) = -(Kx[i] * Ky[i]);
  }
  return Q;
}

// rcwa.cpp:163-234
static void GapMedium(const Vector& Kx, const Vector& Ky, Matrix& W0, Matrix& V0) {
  const int N = (int)Kx.size();
  Vector lam(2 * N);
  for (int i = 0; i < N; ++i) {
    const Complex kz = std::conj(unsigned_sqrt(1.0 - Kx[i] * Kx[i] - Ky[i] * Ky[i]));
    lam[i] = lam[i + N] = Complex(0, 1) * kz;
  }
  Matrix Q = q_half(Kx, Ky, 1.0);
  W0 = Matrix::Identity(2 * N);
  V0 = Matrix(2 * N, 2 * N);
  for (int j = 0; j < 2 * N; ++j)
    for (int i = 0; i < 2 * N; ++i) V0(i, j) = Q(i, j) / lam[j];
}

// rcwa.cpp:237-250
static ScatteringMatrix SMatrixInit(int n) {
  return {Matrix(n, n), Matrix::Identity(n), Matrix::Identity(n), Matrix(n, n)};
}

// rcwa.cpp:254-352
static ScatteringMatrix SMatrixLayer(const Matrix& erc, const Matrix& urc, double k0, double t,
                                     const Vector& Kx, const Vector& Ky, const Matrix& W0,
                                     const Matrix& V0, int* eig_iters = nullptr) {
  const Matrix KX = diag_dense(Kx), KY = diag_dense(Ky);
  LU le(erc);
  const Matrix erc_inv_Kx = le.solve(KX);                 // rcwa.cpp:267
  const Matrix erc_inv_Ky = le.solve(KY);                 // rcwa.cpp:270
  Vector mKy(Ky.size());
  for (size_t i = 0; i < Ky.size(); ++i) mKy[i] = -Ky[i];
  const Matrix Pm = blocks(diag_left(Kx, erc_inv_Ky), add(urc, diag_left(Kx, erc_inv_Kx), -1.0),
                           add(diag_left(Ky, erc_inv_Ky), urc, -1.0), diag_left(mKy, erc_inv_Kx));
  LU lu(urc);
  const Matrix urc_inv_Kx = lu.solve(KX);                 // rcwa.cpp:280
  const Matrix urc_inv_Ky = lu.solve(KY);                 // rcwa.cpp:281
  const Matrix Qm = blocks(diag_left(Kx, urc_inv_Ky), add(erc, diag_left(Kx, urc_inv_Kx), -1.0),
                           add(diag_left(Ky, urc_inv_Ky), erc, -1.0), diag_left(mKy, urc_inv_Kx));
  const Matrix Omega2 = matmul(Pm, Qm);                   // rcwa.cpp:288
  EigResult er = eig(Omega2);                              // rcwa.cpp:298-306
  if (eig_iters) *eig_iters = er.iterations;
  const int n = Omega2.rows;
  Vector lam(n), X(n);
  for (int i = 0; i < n; ++i) {
    lam[i] = unsigned_sqrt(er.values[i]);                  // rcwa.cpp:317
    X[i] = std::exp(-k0 * t * lam[i]);                     // rcwa.cpp:339
  }
  Matrix Vm = matmul(Qm, er.vectors);                      // rcwa.cpp:326-330
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) Vm(i, j) /= lam[j];
  const Matrix WiW0 = LU(er.vectors).solve(W0);            // rcwa.cpp:335
  const Matrix ViV0 = LU(Vm).solve(V0);                    // rcwa.cpp:336
  const Matrix A = add(WiW0, ViV0), B = add(WiW0, ViV0, -1.0);
  const Matrix Xd = diag_dense(X);
  LU Alu(A);                                               // rcwa.cpp:345
  const Matrix AiX = Alu.solve(Xd);
  const Matrix XBAiX = matmul(diag_left(X, B), AiX);       // X*B*A^-1*X
  LU D(add(A, matmul(XBAiX, B), -1.0));                    // rcwa.cpp:346
  const Matrix S11 = D.solve(add(matmul(XBAiX, A), B, -1.0));            // rcwa.cpp:347
  const Matrix S12 = matmul(D.solve(Xd), add(A, matmul(B, Alu.solve(B)), -1.0));  // rcwa.cpp:348
  return {S11, S12, S12, S11};                             // rcwa.cpp:351
}

// rcwa.cpp:354-411 (which = 0) and rcwa.cpp:414-466 (which = 1)
static ScatteringMatrix SMatrixEnd(int which, const Params& p, const Vector& Kx, const Vector& Ky,
                                   const Vector& Kz, const Matrix& W0, const Matrix& V0,
                                   Matrix& Wend) {
  const int N = (int)Kx.size();
  const Complex er = which == 0 ? p.er_ref : p.er_trn;
  const Complex ur = which == 0 ? p.ur_ref : p.ur_trn;
  Matrix Q = scale(q_half(Kx, Ky, ur * er), 1.0 / ur);
  Wend = Matrix::Identity(2 * N);
  const Complex ii = which == 0 ? Complex(0, -1) : Complex(0, 1);
  Matrix V(2 * N, 2 * N);
  for (int j = 0; j < 2 * N; ++j) {
    const Complex l = ii * Kz[j % N];
    for (int i = 0; i < 2 * N; ++i) V(i, j) = Q(i, j) / l;
  }
  const Matrix W0iW = LU(W0).solve(Wend);
  const Matrix V0iV = LU(V0).solve(V);
  const Matrix A = add(W0iW, V0iV), B = add(W0iW, V0iV, -1.0);
  LU Alu(A);
  const Matrix AiB = Alu.solve(B);
  const Matrix S_a = scale(AiB, -1.0);
  const Matrix S_b = Alu.solve(scale(Matrix::Identity(2 * N), 2.0));
  const Matrix S_c = scale(add(A, matmul(B, AiB), -1.0), 0.5);
  const Matrix S_d = transpose(LU(transpose(A)).solve(transpose(B)));
  if (which == 0) return {S_a, S_b, S_c, S_d};             // rcwa.cpp:405-410
  return {S_d, S_c, S_b, S_a};                             // rcwa.cpp:460-465
}

// rcwa.cpp:588-615
static ScatteringMatrix RedhefferProduct(const ScatteringMatrix& A, const ScatteringMatrix& B) {
  const int n = A.S11.rows;
  const Matrix I = Matrix::Identity(n);
  LU l1(add(I, matmul(B.S11, A.S22), -1.0));
  LU l2(add(I, matmul(A.S22, B.S11), -1.0));
  ScatteringMatrix S;
  S.S11 = add(A.S11, matmul(matmul(A.S12, l1.solve(B.S11)), A.S21));
  S.S12 = matmul(A.S12, l1.solve(B.S12));
  S.S21 = matmul(B.S21, l2.solve(A.S21));
  S.S22 = add(B.S22, matmul(matmul(B.S21, l2.solve(A.S22)), B.S12));
  return S;
}

// rcwa.cpp:469-523
static Vector ComputeSourceModeCoeff(const Source& s, const Params& p, const Vector& k_inc,
                                     const Matrix& W_ref) {
  Complex aTE[3];
  if (std::fabs(s.theta) < 1e-6) { aTE[0] = 0; aTE[1] = 1; aTE[2] = 0; }
  else { aTE[0] = -k_inc[1]; aTE[1] = k_inc[0]; aTE[2] = 0; }
  auto normalize = [](Complex* v) {
    double nn = std::sqrt(std::norm(v[0]) + std::norm(v[1]) + std::norm(v[2]));
    for (int i = 0; i < 3; ++i) v[i] /= nn;
  };
  normalize(aTE);
  Complex aTM[3] = {k_inc[1] * aTE[2] - k_inc[2] * aTE[1], k_inc[2] * aTE[0] - k_inc[0] * aTE[2],
                    k_inc[0] * aTE[1] - k_inc[1] * aTE[0]};
  normalize(aTM);
  Complex pol[3];
  for (int i = 0; i < 3; ++i) pol[i] = s.pte * aTE[i] + s.ptm * aTM[i];
  normalize(pol);
  const int PQ = (2 * p.hx + 1) * (2 * p.hy + 1);
  Vector esrc(2 * PQ, 0);
  esrc[PQ / 2] = pol[0];
  esrc[PQ / 2 + PQ] = pol[1];
  return LU(W_ref).solve(esrc);
}

// rcwa.cpp:526-561
static Vector ComputeField(const Matrix& Sblk, const Vector& csrc, const Vector& Kx,
                           const Vector& Ky, const Vector& Kz, const Matrix& W) {
  const int PQ = (int)Kx.size();
  Vector e(2 * PQ, 0), out(3 * PQ, 0);
  for (int j = 0; j < 2 * PQ; ++j)
    for (int i = 0; i < 2 * PQ; ++i) e[i] += Sblk(i, j) * csrc[j];
  for (int i = 0; i < 2 * PQ; ++i) {
    Complex s(0, 0);
    for (int j = 0; j < 2 * PQ; ++j) s += W(i, j) * e[j];
    out[i] = s;
  }
  for (int i = 0; i < PQ; ++i) out[2 * PQ + i] = -(Kx[i] * out[i] + Ky[i] * out[PQ + i]) / Kz[i];
  return out;
}

}  // namespace oracle

// ====================================================================== C API
using namespace oracle;

namespace {
thread_local char g_err[256];
int fail(int code, const char* m) {
  std::snprintf(g_err, sizeof g_err, "%s", m);
  return code;
}
Vector vec_from(const double* p, int n) {
  Vector v(n);
  for (int i = 0; i < n; ++i) v[i] = Complex(p[2 * i], p[2 * i + 1]);
  return v;
}
Matrix mat_from(const double* p, int r, int c) {
  Matrix m(r, c);
  std::memcpy((void*)m.a.data(), p, sizeof(Complex) * (size_t)r * c);
  return m;
}
void mat_to(const Matrix& m, double* p) {
  if (p) std::memcpy(p, m.a.data(), sizeof(Complex) * m.a.size());
}
}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err; }

int oracle_convmat(const double* grid, int Nx, int Ny, int hx, int hy, double* out) {
  try {
    mat_to(ConvMat(reinterpret_cast<const Complex*>(grid), Nx, Ny, hx, hy), out);
    return 0;
  } catch (const OracleError& e) { return fail(e.code, e.what()); }
}

int oracle_eig(int n, const double* A, double* values, double* vectors, int* iters) {
  try {
    EigResult r = eig(mat_from(A, n, n));
    std::memcpy(values, r.values.data(), sizeof(Complex) * n);
    mat_to(r.vectors, vectors);
    if (iters) *iters = r.iterations;
    return 0;
  } catch (const OracleError& e) { return fail(e.code, e.what()); }
}

// Schur factor only (for tests of the QR phase): T upper triangular, Z unitary.
int oracle_hessenberg(int n, double* A, double* Q) {
  Matrix a = mat_from(A, n, n), q;
  hessenberg(a, q);
  mat_to(a, A);
  mat_to(q, Q);
  return 0;
}

int oracle_lu_solve(int n, int nrhs, const double* A, const double* B, double* X) {
  mat_to(LU(mat_from(A, n, n)).solve(mat_from(B, n, nrhs)), X);
  return 0;
}

int oracle_matmul(int m, int k, int nn, const double* A, const double* B, double* C) {
  mat_to(matmul(mat_from(A, m, k), mat_from(B, k, nn)), C);
  return 0;
}

// Wave vectors; outputs are PQ complex each (k_inc is 3 complex).
int oracle_wave_vectors(double Lx, double Ly, double lam0, double theta, double phi, int hx,
                        int hy, const double* media /*er_ref,ur_ref,er_trn,ur_trn*/,
                        double* k_inc, double* Kx, double* Ky, double* Kz_ref, double* Kz_trn) {
  Params p{hx, hy, {media[0], media[1]}, {media[2], media[3]}, {media[4], media[5]},
           {media[6], media[7]}};
  Source s{lam0, 2 * M_PI / lam0, theta, phi, 1, 0};
  Vector ki, kx, ky, kr, kt;
  ComputeWaveVectors(Lx, Ly, s, p, ki, kx, ky, kr, kt);
  std::memcpy(k_inc, ki.data(), 3 * sizeof(Complex));
  std::memcpy(Kx, kx.data(), kx.size() * sizeof(Complex));
  std::memcpy(Ky, ky.data(), ky.size() * sizeof(Complex));
  std::memcpy(Kz_ref, kr.data(), kr.size() * sizeof(Complex));
  std::memcpy(Kz_trn, kt.data(), kt.size() * sizeof(Complex));
  return 0;
}

// Layer S-matrix for given convolution matrices and wave vectors (N = PQ).
int oracle_smatrix_layer(int N, const double* erc, const double* urc, const double* Kx,
                         const double* Ky, double k0, double t, double* S11, double* S12) {
  try {
    Vector kx = vec_from(Kx, N), ky = vec_from(Ky, N);
    Matrix W0, V0;
    GapMedium(kx, ky, W0, V0);
    ScatteringMatrix S = SMatrixLayer(mat_from(erc, N, N), mat_from(urc, N, N), k0, t, kx, ky,
                                      W0, V0);
    mat_to(S.S11, S11);
    mat_to(S.S12, S12);
    return 0;
  } catch (const OracleError& e) { return fail(e.code, e.what()); }
}

int oracle_redheffer(int n, const double* const* A, const double* const* B, double* const* S) {
  ScatteringMatrix a{mat_from(A[0], n, n), mat_from(A[1], n, n), mat_from(A[2], n, n),
                     mat_from(A[3], n, n)};
  ScatteringMatrix b{mat_from(B[0], n, n), mat_from(B[1], n, n), mat_from(B[2], n, n),
                     mat_from(B[3], n, n)};
  ScatteringMatrix s = RedhefferProduct(a, b);
  mat_to(s.S11, S[0]);
  mat_to(s.S12, S[1]);
  mat_to(s.S21, S[2]);
  mat_to(s.S22, S[3]);
  return 0;
}

// Full solve in the order of main.cpp:126-191.
//  er, ur: [L][Ny][Nx] complex; thick: [L]; media: er_ref, ur_ref, er_trn, ur_trn
//  pols: npol pairs (pte, ptm); R, T: npol * Q*P doubles (Q rows, column-major);
//  tot: npol * 2 doubles (R_tot, T_tot); Sg (optional): 4 n*n complex blocks.
int oracle_solve(int Nx, int Ny, int L, double Lx, double Ly, const double* er, const double* ur,
                 const double* thick, double lam0, double theta, double phi, int hx, int hy,
                 const double* media, int npol, const double* pols, double* R, double* T,
                 double* tot, double* Sg, int layer_limit) {
  try {
    Params p{hx, hy, {media[0], media[1]}, {media[2], media[3]}, {media[4], media[5]},
             {media[6], media[7]}};
    Source s{lam0, 2 * M_PI / lam0, theta, phi, 1.0, 0.0};
    const Complex* E = reinterpret_cast<const Complex*>(er);
    const Complex* U = reinterpret_cast<const Complex*>(ur);
    const size_t G = (size_t)Nx * Ny;
    const int nl = (layer_limit >= 0 && layer_limit < L) ? layer_limit : L;
    std::vector<Matrix> erc(nl), urc(nl);
    for (int l = 0; l < nl; ++l) {                       // main.cpp:126-138
      erc[l] = ConvMat(E + l * G, Nx, Ny, hx, hy);
      urc[l] = ConvMat(U + l * G, Nx, Ny, hx, hy);
    }
    Vector k_inc, Kx, Ky, Kz_ref, Kz_trn;
    ComputeWaveVectors(Lx, Ly, s, p, k_inc, Kx, Ky, Kz_ref, Kz_trn);   // main.cpp:140
    Matrix W0, V0;
    GapMedium(Kx, Ky, W0, V0);                            // main.cpp:141
    const int n = 2 * (int)Kx.size();
    ScatteringMatrix Sd = SMatrixInit(n);                 // main.cpp:143
    for (int l = 0; l < nl; ++l)                          // main.cpp:146-164
      Sd = RedhefferProduct(Sd, SMatrixLayer(erc[l], urc[l], s.k0, thick[l], Kx, Ky, W0, V0));
    Matrix Wr, Wt;
    ScatteringMatrix Sr = SMatrixEnd(0, p, Kx, Ky, Kz_ref, W0, V0, Wr);   // main.cpp:166
    ScatteringMatrix St = SMatrixEnd(1, p, Kx, Ky, Kz_trn, W0, V0, Wt);   // main.cpp:167
    ScatteringMatrix Sgl = RedhefferProduct(RedhefferProduct(Sr, Sd), St); // main.cpp:169-170
    if (Sg) {
      const size_t nn = (size_t)n * n * 2;
      mat_to(Sgl.S11, Sg);
      mat_to(Sgl.S12, Sg + nn);
      mat_to(Sgl.S21, Sg + 2 * nn);
      mat_to(Sgl.S22, Sg + 3 * nn);
    }
    const int PQ = n / 2, P = 2 * hx + 1, Q = 2 * hy + 1;
    (void)P; (void)Q;
    for (int ip = 0; ip < npol; ++ip) {                   // main.cpp:182-191
      s.pte = pols[2 * ip];
      s.ptm = pols[2 * ip + 1];
      Vector csrc = ComputeSourceModeCoeff(s, p, k_inc, Wr);
      Vector r = ComputeField(Sgl.S11, csrc, Kx, Ky, Kz_ref, Wr);
      Vector t = ComputeField(Sgl.S21, csrc, Kx, Ky, Kz_trn, Wt);
      // rcwa.cpp:564-585; reshaped(Q, P) column-major = harmonic index order.
      const double den = (k_inc[2] / p.ur_ref).real();
      double rt = 0, tt = 0;
      for (int m = 0; m < PQ; ++m) {
        const double rr = std::norm(r[m]) + std::norm(r[PQ + m]) + std::norm(r[2 * PQ + m]);
        const double tq = std::norm(t[m]) + std::norm(t[PQ + m]) + std::norm(t[2 * PQ + m]);
        const double Rm = (-Kz_ref[m] / p.ur_ref).real() / den * rr;
        const double Tm = (Kz_trn[m] / p.ur_trn).real() / den * tq;
        R[(size_t)ip * PQ + m] = Rm;
        T[(size_t)ip * PQ + m] = Tm;
        rt += Rm;
        tt += Tm;
      }
      tot[2 * ip] = rt;
      tot[2 * ip + 1] = tt;
    }
    return 0;
  } catch (const OracleError& e) {
    return fail(e.code, e.what());
  }
}

}  // extern "C"
