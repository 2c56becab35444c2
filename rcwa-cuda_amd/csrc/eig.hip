// Batched eigendecomposition of general complex matrices on gfx950: the
// replacement of Eigen::ComplexEigenSolver at rcwa.cpp:298-306.
//
// Three phases, each one batched launch:
//   1. Householder reduction to upper Hessenberg form H = Q^H A Q, Q
//      accumulated (one workgroup per matrix);
//   2. complex Schur form T = Z^H A Z by implicitly shifted single-shift QR
//      sweeps with Wilkinson shifts (exceptional shifts at the 10th and 20th
//      sweep on one block) and deflation on negligible subdiagonals (one
//      workgroup per matrix; every Givens rotation is applied by all threads,
//      rows then columns);
//   3. eigenvectors: the unit upper-triangular X with T X = X diag(T) by
//      back-substitution, one wave per eigenvector (tiny pivots are replaced
//      by max(eps |t_kk|, safe minimum), as in LAPACK's xTREVC), then
//      W = Z X on the matrix cores and unit 2-norm columns.
// The S-matrices the path builds from W depend only on the spans of
// eigenvectors and not on their scale or order, so W need not match Eigen's
// normalisation or sort order.
#include <cfloat>
#include <cstdlib>
#include "common.hpp"
#include "kernels.hpp"

namespace {

constexpr int HESS_THREADS = 1024;
constexpr int QR_THREADS = 512;

__device__ __forceinline__ zc shfl_down_z(zc v, int off) {
  return zmake(__shfl_down(v.x, off), __shfl_down(v.y, off));
}

__global__ void set_identity_kernel(int n, ZMat Q) {
  zc* q = Q.p + (long long)blockIdx.y * Q.stride;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < (long long)n * n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    q[i + j * Q.ld] = zmake(i == j ? 1.0 : 0.0, 0.0);
  }
}

__global__ __launch_bounds__(HESS_THREADS) void hess_kernel(int n, ZMat Am, ZMat Qm) {
  extern __shared__ zc v[];
  __shared__ double red[HESS_THREADS / 64];
  __shared__ zc s_tau, s_beta, s_scale;
  __shared__ int s_skip;
  const int b = blockIdx.x;
  zc* A = Am.p + (long long)b * Am.stride;
  zc* Q = Qm.p + (long long)b * Qm.stride;
  const long long lda = Am.ld, ldq = Qm.ld;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nwaves = HESS_THREADS / 64;
  for (int k = 0; k + 2 < n; ++k) {
    const int len = n - k - 1;
    double part = 0.0;
    for (int i = tid; i < len; i += HESS_THREADS) {
      const zc x = A[(k + 1 + i) + k * lda];
      v[i] = x;
      if (i > 0) part += zabs2(x);
    }
    for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off);
    if (lane == 0) red[wave] = part;
    __syncthreads();
    if (tid == 0) {
      double xn2 = 0.0;
      for (int w = 0; w < nwaves; ++w) xn2 += red[w];
      const zc alpha = v[0];
      if (xn2 == 0.0 && alpha.y == 0.0) {
        s_skip = 1;
      } else {
        s_skip = 0;
        const double nrm = sqrt(zabs2(alpha) + xn2);
        const double beta = (alpha.x >= 0.0) ? -nrm : nrm;
        s_tau = zmake((beta - alpha.x) / beta, -alpha.y / beta);
        s_beta = zmake(beta, 0.0);
        s_scale = zinv(zsub(alpha, zmake(beta, 0.0)));
      }
    }
    __syncthreads();
    if (s_skip) continue;
    const zc tau = s_tau, scale = s_scale;
    for (int i = tid; i < len; i += HESS_THREADS) v[i] = (i == 0) ? zmake(1, 0) : zmul(v[i], scale);
    __syncthreads();
    // A(:, k+1:) -= tau (A(:, k+1:) v) v^H ; Q(:, k+1:) likewise
    for (int i = tid; i < n; i += HESS_THREADS) {
      zc w = zmake(0, 0), wq = zmake(0, 0);
      for (int j = 0; j < len; ++j) {
        const zc vj = v[j];
        w = zfma(A[i + (k + 1 + j) * lda], vj, w);
        wq = zfma(Q[i + (k + 1 + j) * ldq], vj, wq);
      }
      w = zmul(w, tau);
      wq = zmul(wq, tau);
      for (int j = 0; j < len; ++j) {
        const zc cv = zconj(v[j]);
        A[i + (k + 1 + j) * lda] = zsub(A[i + (k + 1 + j) * lda], zmul(w, cv));
        Q[i + (k + 1 + j) * ldq] = zsub(Q[i + (k + 1 + j) * ldq], zmul(wq, cv));
      }
    }
    __syncthreads();
    // A(k+1:, j) -= conj(tau) v (v^H A(k+1:, j)),  j = k+1..n-1
    const zc ctau = zconj(tau);
    for (int j = k + 1 + wave; j < n; j += nwaves) {
      zc* col = A + (k + 1) + j * lda;
      zc d = zmake(0, 0);
      for (int i = lane; i < len; i += 64) d = zadd(d, zmul(zconj(v[i]), col[i]));
      for (int off = 32; off > 0; off >>= 1) d = zadd(d, shfl_down_z(d, off));
      d = zmake(__shfl(d.x, 0), __shfl(d.y, 0));
      d = zmul(d, ctau);
      for (int i = lane; i < len; i += 64) col[i] = zsub(col[i], zmul(v[i], d));
    }
    for (int i = tid; i < len; i += HESS_THREADS)
      A[(k + 1 + i) + k * lda] = (i == 0) ? s_beta : zmake(0, 0);
    __syncthreads();
  }
}

__device__ __forceinline__ void givens(zc f, zc g, double& c, zc& s, zc& r) {
  const double af = zabs(f), ag = zabs(g);
  if (ag == 0.0) { c = 1.0; s = zmake(0, 0); r = f; return; }
  if (af == 0.0) { c = 0.0; s = zscale(zconj(g), 1.0 / ag); r = zmake(ag, 0); return; }
  const double nr = hypot(af, ag);
  const zc ph = zscale(f, 1.0 / af);
  c = af / nr;
  s = zscale(zmul(ph, zconj(g)), 1.0 / nr);
  r = zscale(ph, nr);
}

__device__ zc wilkinson(zc a, zc b, zc c, zc d) {
  const zc x = zscale(zsub(a, d), 0.5);
  const zc bc = zmul(b, c);
  zc y = zsqrt(zadd(zmul(x, x), bc));
  if (x.x * y.x + x.y * y.y < 0) y = zneg(y);
  const zc den = zadd(x, y);
  if (den.x == 0.0 && den.y == 0.0) return d;
  return zsub(d, zdiv(bc, den));
}

__global__ __launch_bounds__(QR_THREADS) void schur_kernel(int n, ZMat Tm, ZMat Zm, int* info,
                                                           int maxit_per_row) {
  __shared__ int s_il;
  __shared__ double s_c;
  __shared__ zc s_s, s_mu;
  const int b = blockIdx.x;
  zc* T = Tm.p + (long long)b * Tm.stride;
  zc* Z = Zm.p + (long long)b * Zm.stride;
  const long long ld = Tm.ld, ldz = Zm.ld;
  const int tid = threadIdx.x;
  const double eps = DBL_EPSILON;
  const long long maxit = (long long)maxit_per_row * (n > 1 ? n : 1);
  int iu = n - 1, iter = 0;
  long long total = 0;
  while (iu > 0) {
    if (tid == 0) s_il = 0;
    __syncthreads();
    for (int k = iu - tid; k >= 1; k -= QR_THREADS) {
      const double sub = zabs1(T[k + (k - 1) * ld]);
      const double tst = zabs1(T[(k - 1) + (k - 1) * ld]) + zabs1(T[k + k * ld]);
      if (sub <= eps * tst || sub < DBL_MIN) atomicMax(&s_il, k);
    }
    __syncthreads();
    const int il = s_il;
    if (il == iu) {
      if (tid == 0) T[iu + (iu - 1) * ld] = zmake(0, 0);
      --iu;
      iter = 0;
      __syncthreads();
      continue;
    }
    ++iter;
    ++total;
    if (total > maxit) {
      if (tid == 0 && info) info[b] = 1;
      break;
    }
    if (tid == 0) {
      if (il > 0) T[il + (il - 1) * ld] = zmake(0, 0);
      zc mu;
      if (iter == 10)
        mu = zadd(T[il + il * ld], zmake(0.75 * fabs(T[(il + 1) + il * ld].x), 0));
      else if (iter == 20)
        mu = zadd(T[iu + iu * ld], zmake(0.75 * fabs(T[iu + (iu - 1) * ld].x), 0));
      else
        mu = wilkinson(T[(iu - 1) + (iu - 1) * ld], T[(iu - 1) + iu * ld],
                       T[iu + (iu - 1) * ld], T[iu + iu * ld]);
      s_mu = mu;
    }
    __syncthreads();
    for (int k = il; k < iu; ++k) {
      if (tid == 0) {
        zc f, g;
        if (k == il) {
          f = zsub(T[il + il * ld], s_mu);
          g = T[(il + 1) + il * ld];
        } else {
          f = T[k + (k - 1) * ld];
          g = T[(k + 1) + (k - 1) * ld];
        }
        double c;
        zc s, r;
        givens(f, g, c, s, r);
        if (k > il) {
          T[k + (k - 1) * ld] = r;
          T[(k + 1) + (k - 1) * ld] = zmake(0, 0);
        }
        s_c = c;
        s_s = s;
      }
      __syncthreads();
      const double c = s_c;
      const zc s = s_s, sc = zconj(s);
      for (int j = k + tid; j < n; j += QR_THREADS) {
        const zc x = T[k + j * ld], y = T[(k + 1) + j * ld];
        T[k + j * ld] = zadd(zscale(x, c), zmul(s, y));
        T[(k + 1) + j * ld] = zsub(zscale(y, c), zmul(sc, x));
      }
      for (int i = tid; i < n; i += QR_THREADS) {
        const zc x = Z[i + k * ldz], y = Z[i + (k + 1) * ldz];
        Z[i + k * ldz] = zadd(zscale(x, c), zmul(sc, y));
        Z[i + (k + 1) * ldz] = zsub(zscale(y, c), zmul(s, x));
      }
      __syncthreads();
      const int ie = min(k + 2, iu);
      for (int i = tid; i <= ie; i += QR_THREADS) {
        const zc x = T[i + k * ld], y = T[i + (k + 1) * ld];
        T[i + k * ld] = zadd(zscale(x, c), zmul(sc, y));
        T[i + (k + 1) * ld] = zsub(zscale(y, c), zmul(s, x));
      }
      __syncthreads();
    }
  }
}

// Windowed variant of the QR sweep.  A sweep over the active block [il, iu]
// is cut into windows of R consecutive Givens rotations.  Per window:
//   1. the diagonal band rows/cols [k0-1, k0+R+1] of T is staged in LDS;
//   2. wave 0 chases the bulge through the band: it computes the R
//      rotations and applies each one to the band only (rows, then columns,
//      in the sequential order of the unwindowed sweep);
//   3. all threads write the band back and apply the R rotations, held in
//      LDS, to everything outside the band: the rows [k0, k0+R] right of the
//      band (one thread per column, its R+1 entries in registers), the
//      columns [k0, k0+R] above the band and the same columns of Z (one
//      thread per row).  Each element outside the band is read and written
//      once per window instead of once per rotation.
// The arithmetic per element and the rotation order are those of the
// unwindowed sweep.
template <int R, int NT>
__global__ __launch_bounds__(NT) void schur_win_kernel(int n, ZMat Tm, ZMat Zm, int* info,
                                                       int maxit_per_row) {
  constexpr int LDB = R + 3;
  __shared__ zc band[LDB * LDB];
  __shared__ double rc[R];
  __shared__ zc rs[R];
  __shared__ int s_il;
  __shared__ zc s_mu;
  const int b = blockIdx.x;
  zc* T = Tm.p + (long long)b * Tm.stride;
  zc* Z = Zm.p + (long long)b * Zm.stride;
  const long long ld = Tm.ld, ldz = Zm.ld;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double eps = DBL_EPSILON;
  const long long maxit = (long long)maxit_per_row * (n > 1 ? n : 1);
  int iu = n - 1, iter = 0;
  long long total = 0;
  while (iu > 0) {
    if (tid == 0) s_il = 0;
    __syncthreads();
    for (int k = iu - tid; k >= 1; k -= NT) {
      const double sub = zabs1(T[k + (k - 1) * ld]);
      const double tst = zabs1(T[(k - 1) + (k - 1) * ld]) + zabs1(T[k + k * ld]);
      if (sub <= eps * tst || sub < DBL_MIN) atomicMax(&s_il, k);
    }
    __syncthreads();
    const int il = s_il;
    if (il == iu) {
      if (tid == 0) T[iu + (iu - 1) * ld] = zmake(0, 0);
      --iu;
      iter = 0;
      __syncthreads();
      continue;
    }
    ++iter;
    ++total;
    if (total > maxit) {
      if (tid == 0 && info) info[b] = 1;
      break;
    }
    if (tid == 0) {
      if (il > 0) T[il + (il - 1) * ld] = zmake(0, 0);
      zc mu;
      if (iter == 10)
        mu = zadd(T[il + il * ld], zmake(0.75 * fabs(T[(il + 1) + il * ld].x), 0));
      else if (iter == 20)
        mu = zadd(T[iu + iu * ld], zmake(0.75 * fabs(T[iu + (iu - 1) * ld].x), 0));
      else
        mu = wilkinson(T[(iu - 1) + (iu - 1) * ld], T[(iu - 1) + iu * ld],
                       T[iu + (iu - 1) * ld], T[iu + iu * ld]);
      s_mu = mu;
    }
    __syncthreads();
    for (int k0 = il; k0 < iu;) {
      const int k1 = min(k0 + R, iu);          // rotations k0 .. k1-1
      const int nr = k1 - k0;
      const int b0 = k0 > 0 ? k0 - 1 : 0;      // band rows/cols [b0, be]
      const int be = min(k1 + 1, n - 1);
      const int nb = be - b0 + 1;
      for (int idx = tid; idx < nb * nb; idx += NT) {
        const int i = idx % nb, j = idx / nb;
        band[i + j * LDB] = T[(b0 + i) + (long long)(b0 + j) * ld];
      }
      __syncthreads();
      if (wave == 0) {
        for (int k = k0; k < k1; ++k) {
          double c = 0.0;
          zc sr = zmake(0, 0);
          if (lane == 0) {
            zc f, g, r;
            if (k == il) {
              f = zsub(band[(k - b0) + (k - b0) * LDB], s_mu);
              g = band[(k + 1 - b0) + (k - b0) * LDB];
            } else {
              f = band[(k - b0) + (k - 1 - b0) * LDB];
              g = band[(k + 1 - b0) + (k - 1 - b0) * LDB];
            }
            givens(f, g, c, sr, r);
            if (k > il) {
              band[(k - b0) + (k - 1 - b0) * LDB] = r;
              band[(k + 1 - b0) + (k - 1 - b0) * LDB] = zmake(0, 0);
            }
            rc[k - k0] = c;
            rs[k - k0] = sr;
          }
          c = __shfl(c, 0);
          sr = zmake(__shfl(sr.x, 0), __shfl(sr.y, 0));
          const zc sc = zconj(sr);
          // rows k, k+1 over band columns k .. be
          for (int j = k + lane; j <= be; j += 64) {
            zc* p0 = &band[(k - b0) + (j - b0) * LDB];
            const zc x = p0[0], y = p0[1];
            p0[0] = zadd(zscale(x, c), zmul(sr, y));
            p0[1] = zsub(zscale(y, c), zmul(sc, x));
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
          // columns k, k+1 over band rows b0 .. min(k+2, iu)
          const int ie = min(k + 2, iu);
          for (int i = b0 + lane; i <= ie; i += 64) {
            zc* q0 = &band[(i - b0) + (k - b0) * LDB];
            const zc x = q0[0], y = q0[LDB];
            q0[0] = zadd(zscale(x, c), zmul(sc, y));
            q0[LDB] = zsub(zscale(y, c), zmul(sr, x));
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
        }
      }
      __syncthreads();
      // band back to T
      for (int idx = tid; idx < nb * nb; idx += NT) {
        const int i = idx % nb, j = idx / nb;
        T[(b0 + i) + (long long)(b0 + j) * ld] = band[i + j * LDB];
      }
      // rows k0..k1 right of the band: one column per thread
      for (int j = be + 1 + tid; j < n; j += NT) {
        zc v[R + 1];
        zc* col = T + k0 + (long long)j * ld;
#pragma unroll
        for (int q = 0; q <= R; ++q)
          if (q <= nr) v[q] = col[q];
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (q < nr) {
            const double c = rc[q];
            const zc sr = rs[q], sc = zconj(sr);
            const zc x = v[q], y = v[q + 1];
            v[q] = zadd(zscale(x, c), zmul(sr, y));
            v[q + 1] = zsub(zscale(y, c), zmul(sc, x));
          }
#pragma unroll
        for (int q = 0; q <= R; ++q)
          if (q <= nr) col[q] = v[q];
      }
      // columns k0..k1 above the band (rows 0..b0-1 of T) and of Z (all rows)
      for (int i = tid; i < b0 + n; i += NT) {
        zc* base;
        long long lds;
        if (i < b0) { base = T + i + (long long)k0 * ld; lds = ld; }
        else { base = Z + (i - b0) + (long long)k0 * ldz; lds = ldz; }
        zc v[R + 1];
#pragma unroll
        for (int q = 0; q <= R; ++q)
          if (q <= nr) v[q] = base[q * lds];
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (q < nr) {
            const double c = rc[q];
            const zc sr = rs[q], sc = zconj(sr);
            const zc x = v[q], y = v[q + 1];
            v[q] = zadd(zscale(x, c), zmul(sc, y));
            v[q + 1] = zsub(zscale(y, c), zmul(sr, x));
          }
#pragma unroll
        for (int q = 0; q <= R; ++q)
          if (q <= nr) base[q * lds] = v[q];
      }
      __syncthreads();
      k0 = k1;
    }
  }
}

__global__ __launch_bounds__(64) void trevc_kernel(int n, ZCMat Tm, ZMat Xm) {
  extern __shared__ zc rhs[];
  const int k = blockIdx.x, b = blockIdx.y, lane = threadIdx.x;
  const zc* T = Tm.p + (long long)b * Tm.stride;
  zc* X = Xm.p + (long long)b * Xm.stride + (long long)k * Xm.ld;
  const long long ld = Tm.ld;
  const zc lk = T[k + k * ld];
  const double smlnum = DBL_MIN * ((double)n / DBL_EPSILON);
  const double smin = fmax(DBL_EPSILON * zabs1(lk), smlnum);
  for (int i = lane; i < n; i += 64) {
    if (i < k) rhs[i] = zneg(T[i + k * ld]);
    else X[i] = zmake(i == k ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int i = k - 1; i >= 0; --i) {
    zc d = zsub(T[i + i * ld], lk);
    if (zabs1(d) < smin) d = zmake(smin, 0.0);
    const zc xi = zdiv(rhs[i], d);
    if (lane == 0) X[i] = xi;
    const zc* tc = T + i * ld;
    for (int m = lane; m < i; m += 64) rhs[m] = zsub(rhs[m], zmul(tc[m], xi));
    __syncthreads();
  }
}

__global__ void diag_kernel(int n, ZCMat Tm, zc* w) {
  const int b = blockIdx.y;
  const zc* T = Tm.p + (long long)b * Tm.stride;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    w[(long long)b * n + i] = T[i + i * Tm.ld];
}

}  // namespace

int hessenberg_batched(hipStream_t s, int n, ZMat A, ZMat Q, int batch) {
  if (n <= 0 || batch <= 0) return 0;
  long long blocks = ((long long)n * n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(set_identity_kernel, dim3((unsigned)blocks, batch), dim3(256), 0, s, n, Q);
  hipLaunchKernelGGL(hess_kernel, dim3(batch), dim3(HESS_THREADS), (size_t)n * sizeof(zc), s, n,
                     A, Q);
  return (int)hipGetLastError();
}

int schur_batched_unwindowed(hipStream_t s, int n, ZMat T, ZMat Z, int batch, int* info,
                             int maxit_per_row);

int schur_batched(hipStream_t s, int n, ZMat T, ZMat Z, int batch, int* info,
                  int maxit_per_row) {
  if (n <= 0 || batch <= 0) return 0;
  // RCWA_QR_SWEEP=direct selects the unwindowed sweep (same arithmetic).
  static const bool direct = [] {
    const char* e = getenv("RCWA_QR_SWEEP");
    return e && e[0] == 'd';
  }();
  if (direct) return schur_batched_unwindowed(s, n, T, Z, batch, info, maxit_per_row);
  hipLaunchKernelGGL((schur_win_kernel<12, 256>), dim3(batch), dim3(256), 0, s, n, T, Z, info,
                     maxit_per_row);
  return (int)hipGetLastError();
}

int schur_batched_unwindowed(hipStream_t s, int n, ZMat T, ZMat Z, int batch, int* info,
                             int maxit_per_row) {
  if (n <= 0 || batch <= 0) return 0;
  hipLaunchKernelGGL(schur_kernel, dim3(batch), dim3(QR_THREADS), 0, s, n, T, Z, info,
                     maxit_per_row);
  return (int)hipGetLastError();
}

int trevc_batched(hipStream_t s, int n, ZCMat T, ZMat X, int batch) {
  if (n <= 0 || batch <= 0) return 0;
  hipLaunchKernelGGL(trevc_kernel, dim3(n, batch), dim3(64), (size_t)n * sizeof(zc), s, n, T, X);
  return (int)hipGetLastError();
}

size_t eig_workspace_bytes(int n, int batch) {
  return 2 * (size_t)batch * n * n * sizeof(zc) + hess_blocked_ws_bytes(n, batch) + 512;
}

int eig_batched(hipStream_t s, int n, ZMat A, zc* w, ZMat W, int batch, void* ws, int* info,
                int maxit_per_row, hipEvent_t* marks) {
  if (n <= 0 || batch <= 0) return 0;
  zc* Zp = (zc*)ws;
  zc* Xp = Zp + (size_t)batch * n * n;
  ZMat Z{Zp, n, (long long)n * n}, X{Xp, n, (long long)n * n};
  int e;
  if (marks) (void)hipEventRecord(marks[0], s);
  // RCWA_HESS=unblocked selects the one-workgroup unblocked reduction.
  static const bool unblocked = [] {
    const char* v = getenv("RCWA_HESS");
    return v && v[0] == 'u';
  }();
  if (unblocked || n > 4096) {
    if ((e = hessenberg_batched(s, n, A, Z, batch))) return e;
  } else {
    void* hws = (void*)(((uintptr_t)(Xp + (size_t)batch * n * n) + 255) & ~(uintptr_t)255);
    if ((e = hessenberg_blocked_batched(s, n, A, Z, batch, hws))) return e;
  }
  if (marks) (void)hipEventRecord(marks[1], s);
  if ((e = schur_batched(s, n, A, Z, batch, info, maxit_per_row))) return e;
  if (marks) (void)hipEventRecord(marks[2], s);
  if ((e = trevc_batched(s, n, ZCMat{A.p, A.ld, A.stride}, X, batch))) return e;
  hipLaunchKernelGGL(diag_kernel, dim3((n + 255) / 256, batch), dim3(256), 0, s, n,
                     ZCMat{A.p, A.ld, A.stride}, w);
  if ((e = zgemm_batched(s, n, n, n, zmake(1, 0), ZCMat{Zp, n, (long long)n * n},
                         ZCMat{Xp, n, (long long)n * n}, zmake(0, 0), ZCMat{nullptr, 0, 0}, W,
                         batch)))
    return e;
  if ((e = znormalize_cols(s, n, W, batch))) return e;
  if (marks) (void)hipEventRecord(marks[3], s);
  return 0;
}
