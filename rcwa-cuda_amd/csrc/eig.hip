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
constexpr int SCHUR_G = 8;  // sweeps per deferred-update pass

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

// ---------------------------------------------------------------------------
// QR sweeps with deferred updates.  The iteration itself only ever reads the
// active block T[il:iu+1, il:iu+1], so each sweep updates just that block
// (windowed as in schur_win_kernel) and logs its rotations.  Everything else
// the sweep changes -- Z, the rows of T above the block and the columns of T
// right of it -- is updated every G sweeps by a wavefront pass: rotation k of
// logged sweep a acts at time tau = k + 2a, so one pass over a row (or
// column) applies all G sweeps in the order of the unwindowed algorithm while
// holding only 2G consecutive entries, in a circular register window whose
// slot indices are static inside an unrolled block of 2G time steps.  Each
// element outside the active block is read and written once per G sweeps.
// Range of logged sweeps: the time window [tmin, tmax] and the columns (rows)
// lo..hi that any rotation touches, plus the largest il and smallest iu.
struct FlushRange {
  int tmin, tmax, lo, hi, ilmax, iumin;
};

template <int G>
__device__ FlushRange flush_range(int m, const int* s_il, const int* s_iu) {
  int tmin = 1 << 30, tmax = -1, cmin = 1 << 30, cmax = -1, ilmax = -1, iumin = 1 << 30;
  for (int a = 0; a < m; ++a) {
    tmin = min(tmin, s_il[a] + 2 * a);
    tmax = max(tmax, s_iu[a] - 1 + 2 * a);
    cmin = min(cmin, s_il[a]);
    cmax = max(cmax, s_iu[a]);
    ilmax = max(ilmax, s_il[a]);
    iumin = min(iumin, s_iu[a]);
  }
  FlushRange f;
  f.tmin = tmin;
  f.tmax = tmax;
  f.lo = max(cmin, tmin - 2 * (G - 1));
  f.hi = min(cmax, tmax + 1);
  f.ilmax = ilmax;
  f.iumin = iumin;
  return f;
}

// Applies the m logged sweeps to one strip: entries base[c * step], c = lo..hi.
// kind 0: a row of Z (every sweep), kind 1: row idx of T (sweeps with
// idx < il), both by column rotations; kind 2: column idx of T (sweeps with
// idx > iu) by row rotations.
template <int G>
__device__ void flush_strip(const FlushRange& f, int kind, int idx, zc* base, long long step,
                            int n, int m, const int* s_il, const int* s_iu, const double* lc,
                            const zc* ls) {
  constexpr int W = 2 * G;  // power of two: slot of column c is c & (W - 1)
  const int lo = f.lo, hi = f.hi, tmin = f.tmin, tmax = f.tmax;
  // per-sweep bounds in scalar registers, the applicability test as a mask
  int ila[G], iua[G];
  unsigned mask = 0;
#pragma unroll
  for (int a = 0; a < G; ++a) {
    ila[a] = a < m ? __builtin_amdgcn_readfirstlane(s_il[a]) : 0;
    iua[a] = a < m ? __builtin_amdgcn_readfirstlane(s_iu[a]) : 0;
    const bool applies = kind == 0 || (kind == 1 ? idx < ila[a] : idx > iua[a]);
    if (a < m && applies) mask |= 1u << a;
  }
  zc w[W];
  for (int c = lo; c <= min(tmin, hi); ++c) {
    const zc v = base[(long long)c * step];
#pragma unroll
    for (int q = 0; q < W; ++q)
      if ((c & (W - 1)) == q) w[q] = v;
  }
  for (int tb = (tmin / W) * W; tb <= tmax; tb += W) {
#pragma unroll
    for (int u = 0; u < W; ++u) {
      const int tau = tb + u;
      if (tau < tmin || tau > tmax) continue;
      const int cn = tau + 1;
      if (cn > tmin && cn <= hi) w[(u + 1) & (W - 1)] = base[(long long)cn * step];
#pragma unroll
      for (int a = 0; a < G; ++a) {
        const int k = tau - 2 * a;
        if (!((mask >> a) & 1u) || k < ila[a] || k >= iua[a]) continue;
        const double c = lc[(size_t)a * n + (k - ila[a])];
        const zc sr = ls[(size_t)a * n + (k - ila[a])];
        zc& x = w[(u - 2 * a) & (W - 1)];
        zc& y = w[(u - 2 * a + 1) & (W - 1)];
        const zc xv = x, yv = y;
        if (kind == 2) {  // row rotation on a column of T
          x = zadd(zscale(xv, c), zmul(sr, yv));
          y = zsub(zscale(yv, c), zmul(zconj(sr), xv));
        } else {          // column rotation on a row of T or Z
          x = zadd(zscale(xv, c), zmul(zconj(sr), yv));
          y = zsub(zscale(yv, c), zmul(sr, xv));
        }
      }
      const int done = tau - 2 * (G - 1);  // no rotation reaches this column any more
      if (done >= lo && done <= hi) base[(long long)done * step] = w[(u - 2 * (G - 1)) & (W - 1)];
    }
  }
  for (int c = max(lo, tmax - 2 * (G - 1) + 1); c <= hi; ++c) {
    zc v = w[0];
#pragma unroll
    for (int q = 0; q < W; ++q)
      if ((c & (W - 1)) == q) v = w[q];
    base[(long long)c * step] = v;
  }
}

// Rows of Z and T first, then columns of T.  A row strip and a column strip
// of T can share entries; left and right rotations commute, but each strip
// needs its whole row or column, so the two phases are separated by a barrier.
// Called by all NT threads of the workgroup.
template <int G>
__device__ void defer_flush(int n, int m, const int* s_il, const int* s_iu, const double* lc,
                            const zc* ls, zc* T, long long ld, zc* Z, long long ldz, int tid,
                            int nt) {
  const FlushRange f = flush_range<G>(m, s_il, s_iu);
  const int nrow = n + max(f.ilmax, 0), ncol = max(n - 1 - f.iumin, 0);
#pragma unroll 1
  for (int ph = 0; ph < 2; ++ph) {
    const int cnt = ph == 0 ? nrow : ncol;
    for (int st = tid; st < cnt; st += nt) {
      int kind, idx;
      zc* base;
      long long step;
      if (ph == 1) {
        kind = 2;
        idx = f.iumin + 1 + st;
        base = T + (long long)idx * ld;
        step = 1;
      } else if (st < n) {
        kind = 0;
        idx = st;
        base = Z + st;
        step = ldz;
      } else {
        kind = 1;
        idx = st - n;
        base = T + idx;
        step = ld;
      }
      flush_strip<G>(f, kind, idx, base, step, n, m, s_il, s_iu, lc, ls);
    }
    __syncthreads();
  }
}

template <int R, int NT, int G>
__global__ __launch_bounds__(NT) void schur_defer_kernel(int n, ZMat Tm, ZMat Zm, int* info,
                                                         int maxit_per_row, double* logc,
                                                         zc* logs) {
  constexpr int LDB = R + 3;
  __shared__ zc band[LDB * LDB];
  __shared__ double rc[R];
  __shared__ zc rs[R];
  __shared__ int s_il, s_lil[G], s_liu[G];
  __shared__ zc s_mu;
  const int b = blockIdx.x;
  zc* T = Tm.p + (long long)b * Tm.stride;
  zc* Z = Zm.p + (long long)b * Zm.stride;
  double* lc = logc + (size_t)b * G * n;
  zc* ls = logs + (size_t)b * G * n;
  const long long ld = Tm.ld, ldz = Zm.ld;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double eps = DBL_EPSILON;
  const long long maxit = (long long)maxit_per_row * (n > 1 ? n : 1);
  int iu = n - 1, iter = 0, nlog = 0;
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
      s_lil[nlog] = il;
      s_liu[nlog] = iu;
    }
    __syncthreads();
    for (int k0 = il; k0 < iu;) {
      const int k1 = min(k0 + R, iu);
      const int nr = k1 - k0;
      const int b0 = k0 > 0 ? k0 - 1 : 0;
      const int be = min(k1 + 1, n - 1);
      const int nb = be - b0 + 1;
      for (int idx = tid; idx < nb * nb; idx += NT) {
        const int i = idx % nb, j = idx / nb;
        band[i + j * LDB] = T[(b0 + i) + (long long)(b0 + j) * ld];
      }
      __syncthreads();
      if (wave == 0) {
        const int je = min(be, iu);
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
          for (int j = k + lane; j <= je; j += 64) {
            zc* p0 = &band[(k - b0) + (j - b0) * LDB];
            const zc x = p0[0], y = p0[1];
            p0[0] = zadd(zscale(x, c), zmul(sr, y));
            p0[1] = zsub(zscale(y, c), zmul(sc, x));
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
          const int ie = min(k + 2, iu);
          for (int i = max(b0, il) + lane; i <= ie; i += 64) {
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
      for (int idx = tid; idx < nb * nb; idx += NT) {
        const int i = idx % nb, j = idx / nb;
        T[(b0 + i) + (long long)(b0 + j) * ld] = band[i + j * LDB];
      }
      if (tid < nr) {
        lc[(size_t)nlog * n + (k0 - il) + tid] = rc[tid];
        ls[(size_t)nlog * n + (k0 - il) + tid] = rs[tid];
      }
      // rows k0..k1 of the active block right of the band
      for (int j = be + 1 + tid; j <= iu; j += NT) {
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
      // columns k0..k1 of the active block above the band
      for (int i = il + tid; i < b0; i += NT) {
        zc* base = T + i + (long long)k0 * ld;
        zc v[R + 1];
#pragma unroll
        for (int q = 0; q <= R; ++q)
          if (q <= nr) v[q] = base[q * ld];
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
          if (q <= nr) base[q * ld] = v[q];
      }
      __syncthreads();
      k0 = k1;
    }
    ++nlog;
    if (nlog == G) {
      __threadfence_block();
      defer_flush<G>(n, nlog, s_lil, s_liu, lc, ls, T, ld, Z, ldz, tid, NT);
      __syncthreads();
      nlog = 0;
    }
  }
  __syncthreads();
  if (nlog > 0) defer_flush<G>(n, nlog, s_lil, s_liu, lc, ls, T, ld, Z, ldz, tid, NT);
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

size_t schur_ws_bytes(int n, int batch) {
  return (size_t)batch * SCHUR_G * n * (sizeof(double) + sizeof(zc)) + 512;
}

int schur_batched(hipStream_t s, int n, ZMat T, ZMat Z, int batch, int* info,
                  int maxit_per_row, void* ws) {
  if (n <= 0 || batch <= 0) return 0;
  // RCWA_QR_SWEEP=direct selects the unwindowed sweep, =window the windowed
  // sweep that updates all of T and Z at once (same arithmetic, same result).
  static const char mode = [] {
    const char* e = getenv("RCWA_QR_SWEEP");
    return e ? e[0] : 0;
  }();
  if (mode == 'd') return schur_batched_unwindowed(s, n, T, Z, batch, info, maxit_per_row);
  if (ws && mode != 'w') {
    double* lc = (double*)ws;
    zc* ls = (zc*)(((uintptr_t)(lc + (size_t)batch * SCHUR_G * n) + 255) & ~(uintptr_t)255);
    hipLaunchKernelGGL((schur_defer_kernel<12, 256, SCHUR_G>), dim3(batch), dim3(256), 0, s, n,
                       T, Z, info, maxit_per_row, lc, ls);
    return (int)hipGetLastError();
  }
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
  return 2 * (size_t)batch * n * n * sizeof(zc) + hess_blocked_ws_bytes(n, batch) +
         schur_ws_bytes(n, batch) + 1024;
}

int eig_batched(hipStream_t s, int n, ZMat A, zc* w, ZMat W, int batch, void* ws, int* info,
                int maxit_per_row, hipEvent_t* marks) {
  if (n <= 0 || batch <= 0) return 0;
  zc* Zp = (zc*)ws;
  zc* Xp = Zp + (size_t)batch * n * n;
  ZMat Z{Zp, n, (long long)n * n}, X{Xp, n, (long long)n * n};
  void* hws = (void*)(((uintptr_t)(Xp + (size_t)batch * n * n) + 255) & ~(uintptr_t)255);
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
    if ((e = hessenberg_blocked_batched(s, n, A, Z, batch, hws))) return e;
  }
  if (marks) (void)hipEventRecord(marks[1], s);
  void* sws = (void*)(((uintptr_t)hws + hess_blocked_ws_bytes(n, batch) + 255) & ~(uintptr_t)255);
  if ((e = schur_batched(s, n, A, Z, batch, info, maxit_per_row, sws))) return e;
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
