// Blocked Householder reduction to upper Hessenberg form, H = Q^H A Q, for a
// batch of complex matrices (first phase of the eigensolver that replaces
// Eigen::ComplexEigenSolver, rcwa.cpp:298).
//
// Panels of NB columns.  For the panel starting at column p the reflectors
// H_j = I - tau_j v_j v_j^H (v_j zero above row j+1, one at row j+1) are
// combined as H_p ... H_{p+nb-1} = I - V T V^H (compact WY, T upper
// triangular).  The panel kernel (one workgroup per matrix) produces V, T and
// the rows p+1.. of Y = A0 V T, where A0 is the matrix at the start of the
// panel:
//   rows p+1.. of column j are brought up to date with the earlier
//   reflectors of the panel, a_j <- (I - V T^H V^H)(A0 e_j - Y V^H e_j);
//   its reflector is generated from rows j+1..n-1;
//   y_j = tau_j (A0[:, j+1:] v_j - Y (V^H v_j)), t_j = -tau_j T (V^H v_j).
// Rows 0..p see only the right transforms, so they are left to the matrix
// cores after the panel (as LAPACK's zlahr2/zgehrd do):
//   Y[:p+1, :]      = (A0[:p+1, p+1:] V[p+1:, :]) T
//   A[:p+1, p:p+nb] -= Y[:p+1, :] V[p:p+nb, :]^H
// The columns right of the panel then receive the whole block at once:
//   A[:, p+nb:]    -= Y V[p+nb:, :]^H                  (A <- A H)
//   A[p+1:, p+nb:] -= V (T^H (V^H A[p+1:, p+nb:]))      (A <- H^H A)
//   Q[:, p+1:]     -= (Q[:, p+1:] V T) V^H              (Q <- Q H)
// The reduction's only memory-bound part is the product A0[p+1:, j+1:] v_j,
// one read of the trailing block per column (n^3/3 entries over the
// reduction, against n^3/2 when every row is formed in the panel); everything
// else is ZGEMM.
#include <algorithm>
#include "common.hpp"
#include "kernels.hpp"

namespace {

constexpr int NB = 64;
constexpr int PT = 1024;  // panel threads

__device__ __forceinline__ zc shfl_down_z(zc v, int off) {
  return zmake(__shfl_down(v.x, off), __shfl_down(v.y, off));
}
__device__ __forceinline__ zc wave_sum(zc v) {
  for (int off = 32; off > 0; off >>= 1) v = zadd(v, shfl_down_z(v, off));
  return zmake(__shfl(v.x, 0), __shfl(v.y, 0));
}

// Workspace per matrix: V (n x NB), Y (n x NB), T (NB x NB), column ld = n.
// RN rows of the panel's product per thread (n <= RN * PT), CB columns per
// iteration of it (their loads issued together; the sums run in column order
// whatever CB is).
template <int RN, int CB>
__global__ __launch_bounds__(PT) void hess_panel_kernel(int n, int p, int nbe, ZMat Am, zc* Vw,
                                                        zc* Yw, zc* Tw) {
  extern __shared__ zc sm[];
  zc* bcol = sm;            // n: current column / reflector
  __shared__ zc wq[NB], w2[NB], zq[NB];
  __shared__ double red[PT / 64];
  __shared__ zc s_tau, s_beta, s_scale;
  const int b = blockIdx.x;
  zc* A = Am.p + (long long)b * Am.stride;
  const long long lda = Am.ld;
  zc* V = Vw + (size_t)b * n * NB;
  zc* Y = Yw + (size_t)b * n * NB;
  zc* T = Tw + (size_t)b * NB * NB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nw = PT / 64;
  for (int jj = 0; jj < nbe; ++jj) {
    const int j = p + jj;
    // (a) right transforms of the panel so far: b = A0 e_j - Y V(j, :)^H
    // (rows p+1.. only: rows 0..p are updated after the panel)
    for (int i = p + 1 + tid; i < n; i += PT) {
      zc v = A[i + (long long)j * lda];
      for (int q = 0; q < jj; ++q) v = zsub(v, zmul(Y[i + (size_t)q * n], zconj(V[j + (size_t)q * n])));
      bcol[i] = v;
    }
    __syncthreads();
    if (jj > 0) {
      // left transforms: w = V^H b (rows p+1..), w2 = T^H w, b -= V w2
      for (int q = wave; q < jj; q += nw) {
        zc acc = zmake(0, 0);
        for (int i = p + 1 + lane; i < n; i += 64)
          acc = zadd(acc, zmul(zconj(V[i + (size_t)q * n]), bcol[i]));
        acc = wave_sum(acc);
        if (lane == 0) wq[q] = acc;
      }
      __syncthreads();
      if (tid < jj) {
        zc acc = zmake(0, 0);
        for (int r = 0; r <= tid; ++r) acc = zadd(acc, zmul(zconj(T[r + tid * NB]), wq[r]));
        w2[tid] = acc;
      }
      __syncthreads();
      for (int i = p + 1 + tid; i < n; i += PT) {
        zc v = bcol[i];
        for (int q = 0; q < jj; ++q) v = zsub(v, zmul(V[i + (size_t)q * n], w2[q]));
        bcol[i] = v;
      }
      __syncthreads();
    }
    // (b) reflector from b[j+1 .. n-1]
    double part = 0.0;
    for (int i = j + 2 + tid; i < n; i += PT) part += zabs2(bcol[i]);
    for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off);
    if (lane == 0) red[wave] = part;
    __syncthreads();
    if (tid == 0) {
      double xn2 = 0.0;
      for (int w = 0; w < nw; ++w) xn2 += red[w];
      const zc alpha = bcol[j + 1];
      if (xn2 == 0.0 && alpha.y == 0.0) {
        s_tau = zmake(0, 0);
        s_beta = alpha;
        s_scale = zmake(0, 0);
      } else {
        const double nrm = sqrt(zabs2(alpha) + xn2);
        const double beta = (alpha.x >= 0.0) ? -nrm : nrm;
        s_tau = zmake((beta - alpha.x) / beta, -alpha.y / beta);
        s_beta = zmake(beta, 0.0);
        s_scale = zinv(zsub(alpha, zmake(beta, 0.0)));
      }
    }
    __syncthreads();
    const zc tau = s_tau, scale = s_scale;
    // store rows p+1.. of the finished column j and the reflector v (in bcol
    // and V(:, jj), zero above row j+1)
    for (int i = tid; i < n; i += PT) {
      zc a, v;
      if (i <= j) { a = bcol[i]; v = zmake(0, 0); }
      else if (i == j + 1) { a = s_beta; v = zmake(1, 0); }
      else { a = zmake(0, 0); v = zmul(bcol[i], scale); }
      if (i > p) A[i + (long long)j * lda] = a;
      V[i + (size_t)jj * n] = v;
      bcol[i] = v;
    }
    __syncthreads();
    // (c) y = A0[p+1:, j+1:] v ; z = V(:, :jj)^H v
    for (int q = wave; q < jj; q += nw) {
      zc acc = zmake(0, 0);
      for (int i = j + 1 + lane; i < n; i += 64)
        acc = zadd(acc, zmul(zconj(V[i + (size_t)q * n]), bcol[i]));
      acc = wave_sum(acc);
      if (lane == 0) zq[q] = acc;
    }
    zc yacc[RN];
#pragma unroll
    for (int r = 0; r < RN; ++r) yacc[r] = zmake(0, 0);
    // CB columns per iteration, their loads issued together
    int c = j + 1;
    for (; c + CB - 1 < n; c += CB) {
      const zc* col = A + (long long)c * lda;
#pragma unroll
      for (int r = 0; r < RN; ++r) {
        const int i = p + 1 + tid + r * PT;
        if (i < n) {
          zc a[CB];
#pragma unroll
          for (int q = 0; q < CB; ++q) a[q] = col[i + q * lda];
#pragma unroll
          for (int q = 0; q < CB; ++q) yacc[r] = zfma(a[q], bcol[c + q], yacc[r]);
        }
      }
    }
    for (; c < n; ++c) {
      const zc vc = bcol[c];
      const zc* col = A + (long long)c * lda;
#pragma unroll
      for (int r = 0; r < RN; ++r) {
        const int i = p + 1 + tid + r * PT;
        if (i < n) yacc[r] = zfma(col[i], vc, yacc[r]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RN; ++r) {
      const int i = p + 1 + tid + r * PT;
      if (i < n) {
        zc y = yacc[r];
        for (int q = 0; q < jj; ++q) y = zsub(y, zmul(Y[i + (size_t)q * n], zq[q]));
        Y[i + (size_t)jj * n] = zmul(y, tau);
      }
    }
    if (tid <= jj) {
      zc t;
      if (tid == jj) {
        t = tau;
      } else {
        zc acc = zmake(0, 0);
        for (int r = tid; r < jj; ++r) acc = zadd(acc, zmul(T[tid + r * NB], zq[r]));
        t = zneg(zmul(tau, acc));
      }
      T[tid + jj * NB] = t;
    }
    for (int q = jj + 1 + tid; q < NB; q += PT) T[q + jj * NB] = zmake(0, 0);
    __syncthreads();
  }
}

// The panel split by column for small batches (a few matrices of large n,
// where one workgroup per matrix leaves most of the chip idle): per column
// j = p + jj of the panel, hess_col_prep_kernel (one workgroup per matrix)
// does steps (a) and (b) of hess_panel_kernel, stores the finished column and
// the reflector v = V(:, jj), and forms zq = V(:, :jj)^H v (global, NB per
// matrix) and column jj of T; hess_col_y_kernel then forms
// Y(i, jj) = tau (A0[i, j+1:] v - Y(i, :jj) zq) for the rows p+1.. in blocks
// of YR rows over the whole chip.  The same arithmetic per entry as the
// panel kernel.
constexpr int YR = 256;

__global__ __launch_bounds__(PT) void hess_col_prep_kernel(int n, int p, int jj, ZMat Am, zc* Vw,
                                                           zc* Yw, zc* Tw, zc* Zq) {
  extern __shared__ zc sm[];
  zc* bcol = sm;  // n
  __shared__ zc wq[NB], w2[NB], zq[NB];
  __shared__ double red[PT / 64];
  __shared__ zc s_tau, s_beta, s_scale;
  const int b = blockIdx.x;
  zc* A = Am.p + (long long)b * Am.stride;
  const long long lda = Am.ld;
  zc* V = Vw + (size_t)b * n * NB;
  zc* Y = Yw + (size_t)b * n * NB;
  zc* T = Tw + (size_t)b * NB * NB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nw = PT / 64;
  const int j = p + jj;
  for (int i = p + 1 + tid; i < n; i += PT) {
    zc v = A[i + (long long)j * lda];
    for (int q = 0; q < jj; ++q) v = zsub(v, zmul(Y[i + (size_t)q * n], zconj(V[j + (size_t)q * n])));
    bcol[i] = v;
  }
  __syncthreads();
  if (jj > 0) {
    for (int q = wave; q < jj; q += nw) {
      zc acc = zmake(0, 0);
      for (int i = p + 1 + lane; i < n; i += 64)
        acc = zadd(acc, zmul(zconj(V[i + (size_t)q * n]), bcol[i]));
      acc = wave_sum(acc);
      if (lane == 0) wq[q] = acc;
    }
    __syncthreads();
    if (tid < jj) {
      zc acc = zmake(0, 0);
      for (int r = 0; r <= tid; ++r) acc = zadd(acc, zmul(zconj(T[r + tid * NB]), wq[r]));
      w2[tid] = acc;
    }
    __syncthreads();
    for (int i = p + 1 + tid; i < n; i += PT) {
      zc v = bcol[i];
      for (int q = 0; q < jj; ++q) v = zsub(v, zmul(V[i + (size_t)q * n], w2[q]));
      bcol[i] = v;
    }
    __syncthreads();
  }
  double part = 0.0;
  for (int i = j + 2 + tid; i < n; i += PT) part += zabs2(bcol[i]);
  for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off);
  if (lane == 0) red[wave] = part;
  __syncthreads();
  if (tid == 0) {
    double xn2 = 0.0;
    for (int w = 0; w < nw; ++w) xn2 += red[w];
    const zc alpha = bcol[j + 1];
    if (xn2 == 0.0 && alpha.y == 0.0) {
      s_tau = zmake(0, 0);
      s_beta = alpha;
      s_scale = zmake(0, 0);
    } else {
      const double nrm = sqrt(zabs2(alpha) + xn2);
      const double beta = (alpha.x >= 0.0) ? -nrm : nrm;
      s_tau = zmake((beta - alpha.x) / beta, -alpha.y / beta);
      s_beta = zmake(beta, 0.0);
      s_scale = zinv(zsub(alpha, zmake(beta, 0.0)));
    }
  }
  __syncthreads();
  const zc tau = s_tau, scale = s_scale;
  for (int i = tid; i < n; i += PT) {
    zc a, v;
    if (i <= j) { a = bcol[i]; v = zmake(0, 0); }
    else if (i == j + 1) { a = s_beta; v = zmake(1, 0); }
    else { a = zmake(0, 0); v = zmul(bcol[i], scale); }
    if (i > p) A[i + (long long)j * lda] = a;
    V[i + (size_t)jj * n] = v;
    bcol[i] = v;
  }
  __syncthreads();
  for (int q = wave; q < jj; q += nw) {
    zc acc = zmake(0, 0);
    for (int i = j + 1 + lane; i < n; i += 64)
      acc = zadd(acc, zmul(zconj(V[i + (size_t)q * n]), bcol[i]));
    acc = wave_sum(acc);
    if (lane == 0) zq[q] = acc;
  }
  __syncthreads();
  zc* zg = Zq + (size_t)b * NB;
  if (tid < jj) zg[tid] = zq[tid];
  if (tid == jj) zg[NB - 1] = tau;  // the column's tau, for hess_col_y_kernel
  if (tid <= jj) {
    zc t;
    if (tid == jj) {
      t = tau;
    } else {
      zc acc = zmake(0, 0);
      for (int r = tid; r < jj; ++r) acc = zadd(acc, zmul(T[tid + r * NB], zq[r]));
      t = zneg(zmul(tau, acc));
    }
    T[tid + jj * NB] = t;
  }
  for (int q = jj + 1 + tid; q < NB; q += PT) T[q + jj * NB] = zmake(0, 0);
}

// Y(i, jj) for the rows i = p+1 + YR blockIdx.x + threadIdx.x: the product
// with the trailing columns j+1.. (eight columns' loads in flight together,
// v staged in LDS), less Y(i, :jj) zq, times tau.
__global__ __launch_bounds__(YR) void hess_col_y_kernel(int n, int p, int jj, ZCMat Am,
                                                        const zc* Vw, zc* Yw, const zc* Zq) {
  extern __shared__ zc sv[];  // v[j+1 .. n-1]
  const int b = blockIdx.y;
  const zc* A = Am.p + (long long)b * Am.stride;
  const long long lda = Am.ld;
  const zc* V = Vw + (size_t)b * n * NB;
  zc* Y = Yw + (size_t)b * n * NB;
  const zc* zg = Zq + (size_t)b * NB;
  const int j = p + jj;
  for (int c = j + 1 + threadIdx.x; c < n; c += YR) sv[c - j - 1] = V[c + (size_t)jj * n];
  __syncthreads();
  const int i = p + 1 + blockIdx.x * YR + threadIdx.x;
  if (i >= n) return;
  zc y = zmake(0, 0);
  int c = j + 1;
  for (; c + 7 < n; c += 8) {
    const zc* col = A + i + (long long)c * lda;
    zc a[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] = col[q * lda];
#pragma unroll
    for (int q = 0; q < 8; ++q) y = zfma(a[q], sv[c - j - 1 + q], y);
  }
  for (; c < n; ++c) y = zfma(A[i + (long long)c * lda], sv[c - j - 1], y);
  for (int q = 0; q < jj; ++q) y = zsub(y, zmul(Y[i + (size_t)q * n], zg[q]));
  Y[i + (size_t)jj * n] = zmul(y, zg[NB - 1]);
}

// dst (cols x rows) = src(rows x cols)^H, src column-major with ld; dst ld = cols.
__global__ void conj_transpose_kernel(int rows, int cols, const zc* src, long long sld,
                                      long long sstride, zc* dst, long long dstride) {
  const int b = blockIdx.y;
  const zc* s = src + (long long)b * sstride;
  zc* d = dst + (long long)b * dstride;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       idx < (long long)rows * cols; idx += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(idx % rows), c = (int)(idx / rows);
    d[c + (long long)r * cols] = zconj(s[r + (long long)c * sld]);
  }
}

__global__ void identity_q_kernel(int n, ZMat Q) {
  zc* q = Q.p + (long long)blockIdx.y * Q.stride;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < (long long)n * n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    q[i + j * Q.ld] = zmake(i == j ? 1.0 : 0.0, 0.0);
  }
}

inline dim3 g2(long long elems, int batch) {
  long long blocks = (elems + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  return dim3((unsigned)blocks, (unsigned)batch);
}

}  // namespace

int zconj_transpose_batched(hipStream_t s, int rows, int cols, ZCMat src, ZMat dst, int batch) {
  if (rows <= 0 || cols <= 0 || batch <= 0) return 0;
  hipLaunchKernelGGL(conj_transpose_kernel, g2((long long)rows * cols, batch), dim3(256), 0, s,
                     rows, cols, src.p, src.ld, src.stride, dst.p, dst.stride);
  return (int)hipGetLastError();
}

size_t hess_blocked_ws_bytes(int n, int batch) {
  // V, Y (n x NB), T (NB x NB), VH (NB x n), W (NB x n), W2 (NB x n), QV (n x NB), QVT (n x NB),
  // and zq (NB) of the column-split panel
  const size_t per = (size_t)n * NB * 7 + (size_t)NB * NB * 2 + NB;
  return (size_t)batch * per * sizeof(zc) + 1024;
}

int hessenberg_blocked_batched(hipStream_t s, int n, ZMat A, ZMat Q, int batch, void* ws) {
  if (n <= 0 || batch <= 0) return 0;
  if (n > 4096) return (int)hipErrorInvalidValue;
  const size_t nNB = (size_t)n * NB;
  zc* V = (zc*)ws;
  zc* Y = V + batch * nNB;
  zc* T = Y + batch * nNB;
  zc* T2 = T + (size_t)batch * NB * NB;
  zc* VH = T2 + (size_t)batch * NB * NB;
  zc* W = VH + batch * nNB;
  zc* W2 = W + batch * nNB;
  zc* QV = W2 + batch * nNB;
  zc* QVT = QV + batch * nNB;
  zc* ZQ = QVT + batch * nNB;
  // a small batch takes the panel column by column over the whole chip
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const bool by_column = (long long)batch * 4 <= cus && n >= 512;
  const zc one = zmake(1, 0), mone = zmake(-1, 0), zero = zmake(0, 0);
  const ZCMat none{nullptr, 0, 0};
  hipLaunchKernelGGL(identity_q_kernel, g2((long long)n * n, batch), dim3(256), 0, s, n, Q);
  for (int p = 0; p + 2 < n; p += NB) {
    const int nbe = std::min(NB, n - 2 - p);  // reflectors in this panel
    if (by_column) {
      for (int jj = 0; jj < nbe; ++jj) {
        hipLaunchKernelGGL(hess_col_prep_kernel, dim3(batch), dim3(PT), (size_t)n * sizeof(zc), s,
                           n, p, jj, A, V, Y, T, ZQ);
        const int rows = n - p - 1;
        hipLaunchKernelGGL(hess_col_y_kernel, dim3((rows + YR - 1) / YR, batch), dim3(YR),
                           (size_t)(n - p - jj - 1) * sizeof(zc), s, n, p, jj,
                           ZCMat{A.p, A.ld, A.stride}, V, Y, ZQ);
      }
    } else {
      // one product row per thread up to n = PT, else up to four (sized
      // registers for the product; eight columns in flight either way)
      if (n <= PT)
        hipLaunchKernelGGL((hess_panel_kernel<1, 8>), dim3(batch), dim3(PT), (size_t)n * sizeof(zc),
                           s, n, p, nbe, A, V, Y, T);
      else
        hipLaunchKernelGGL((hess_panel_kernel<(4096 + PT - 1) / PT, 8>), dim3(batch), dim3(PT),
                           (size_t)n * sizeof(zc), s, n, p, nbe, A, V, Y, T);
    }
    const int c0 = p + nbe;                   // first column right of the panel
    const int m = n - c0;                     // trailing columns
    const int r1 = p + 1, mr = n - r1;        // rows of V that are nonzero
    int e;
    // rows 0..p: Y[:r1, :] = (A0[:r1, r1:] V[r1:, :]) T, then
    // A[:r1, p:c0] -= Y[:r1, :] V[p:c0, :]^H (row p of V is zero)
    if ((e = zgemm_batched(s, r1, nbe, mr, one, ZCMat{A.p + (long long)r1 * A.ld, A.ld, A.stride},
                           ZCMat{V + r1, n, (long long)nNB}, zero, none,
                           ZMat{QV, r1, (long long)nNB}, batch)))
      return e;
    if ((e = zgemm_batched(s, r1, nbe, nbe, one, ZCMat{QV, r1, (long long)nNB},
                           ZCMat{T, NB, (long long)NB * NB}, zero, none,
                           ZMat{Y, n, (long long)nNB}, batch)))
      return e;
    hipLaunchKernelGGL(conj_transpose_kernel, g2((long long)nbe * nbe, batch), dim3(256), 0, s,
                       nbe, nbe, V + p, (long long)n, (long long)nNB, T2, (long long)NB * NB);
    if ((e = zgemm_batched(s, r1, nbe, nbe, mone, ZCMat{Y, n, (long long)nNB},
                           ZCMat{T2, nbe, (long long)NB * NB}, one,
                           ZCMat{A.p + (long long)p * A.ld, A.ld, A.stride},
                           ZMat{A.p + (long long)p * A.ld, A.ld, A.stride}, batch)))
      return e;
    if (m > 0) {
      // A[:, c0:] -= Y V[c0:, :]^H
      hipLaunchKernelGGL(conj_transpose_kernel, g2((long long)m * nbe, batch), dim3(256), 0, s, m,
                         nbe, V + c0, (long long)n, (long long)nNB, VH, (long long)nNB);
      if ((e = zgemm_batched(s, n, m, nbe, mone, ZCMat{Y, n, (long long)nNB},
                             ZCMat{VH, nbe, (long long)nNB}, one,
                             ZCMat{A.p + (long long)c0 * A.ld, A.ld, A.stride},
                             ZMat{A.p + (long long)c0 * A.ld, A.ld, A.stride}, batch)))
        return e;
      // A[r1:, c0:] -= V (T^H (V^H A[r1:, c0:]))
      hipLaunchKernelGGL(conj_transpose_kernel, g2((long long)mr * nbe, batch), dim3(256), 0, s,
                         mr, nbe, V + r1, (long long)n, (long long)nNB, VH, (long long)nNB);
      if ((e = zgemm_batched(s, nbe, m, mr, one, ZCMat{VH, nbe, (long long)nNB},
                             ZCMat{A.p + r1 + (long long)c0 * A.ld, A.ld, A.stride}, zero, none,
                             ZMat{W, nbe, (long long)nNB}, batch)))
        return e;
      hipLaunchKernelGGL(conj_transpose_kernel, g2((long long)NB * NB, batch), dim3(256), 0, s,
                         nbe, nbe, T, (long long)NB, (long long)NB * NB, T2,
                         (long long)NB * NB);
      if ((e = zgemm_batched(s, nbe, m, nbe, one, ZCMat{T2, nbe, (long long)NB * NB},
                             ZCMat{W, nbe, (long long)nNB}, zero, none,
                             ZMat{W2, nbe, (long long)nNB}, batch)))
        return e;
      if ((e = zgemm_batched(s, mr, m, nbe, mone, ZCMat{V + r1, n, (long long)nNB},
                             ZCMat{W2, nbe, (long long)nNB}, one,
                             ZCMat{A.p + r1 + (long long)c0 * A.ld, A.ld, A.stride},
                             ZMat{A.p + r1 + (long long)c0 * A.ld, A.ld, A.stride}, batch)))
        return e;
    }
    // Q[:, r1:] -= (Q[:, r1:] V[r1:, :] T) V[r1:, :]^H
    if ((e = zgemm_batched(s, n, nbe, mr, one, ZCMat{Q.p + (long long)r1 * Q.ld, Q.ld, Q.stride},
                           ZCMat{V + r1, n, (long long)nNB}, zero, none,
                           ZMat{QV, n, (long long)nNB}, batch)))
      return e;
    if ((e = zgemm_batched(s, n, nbe, nbe, one, ZCMat{QV, n, (long long)nNB},
                           ZCMat{T, NB, (long long)NB * NB}, zero, none,
                           ZMat{QVT, n, (long long)nNB}, batch)))
      return e;
    hipLaunchKernelGGL(conj_transpose_kernel, g2((long long)mr * nbe, batch), dim3(256), 0, s, mr,
                       nbe, V + r1, (long long)n, (long long)nNB, VH, (long long)nNB);
    if ((e = zgemm_batched(s, n, mr, nbe, mone, ZCMat{QVT, n, (long long)nNB},
                           ZCMat{VH, nbe, (long long)nNB}, one,
                           ZCMat{Q.p + (long long)r1 * Q.ld, Q.ld, Q.stride},
                           ZMat{Q.p + (long long)r1 * Q.ld, Q.ld, Q.stride}, batch)))
      return e;
  }
  return (int)hipGetLastError();
}
