// Batched eigendecomposition of general complex matrices on gfx950: the
// replacement of Eigen::ComplexEigenSolver at rcwa.cpp:298-306.
//
// Three phases, each batched over the matrices:
//   1. Householder reduction to upper Hessenberg form H = Q^H A Q, Q
//      accumulated (blocked: compact-WY panels in hess.hip, updates on the
//      matrix cores; the unblocked one-workgroup form for n > 4096);
//   2. complex Schur form T = Z^H A Z by implicitly shifted single-shift QR
//      sweeps with Wilkinson shifts (exceptional shifts at the 10th and 20th
//      sweep on one block) and deflation on negligible subdiagonals: one
//      workgroup per matrix sweeps the active block and logs its rotations,
//      qr_apply_kernel applies the log to the rows of Z, and T is formed
//      afterwards as Z^H (A Z) on the matrix cores;
//   3. eigenvectors: the unit upper-triangular X with T X = X diag(T) by
//      blocked back-substitution (tiny pivots are replaced by
//      max(eps |t_kk|, safe minimum), as in LAPACK's xTREVC), then W = Z X on
//      the matrix cores and unit 2-norm columns, or, for the layer stage,
//      W^-1 = D X^-1 Z^H directly.
// The S-matrices the path builds from W depend only on the spans of
// eigenvectors and not on their scale or order, so W need not match Eigen's
// normalisation or sort order.
#include <algorithm>
#include <cfloat>
#include <cstdlib>
#include "common.hpp"
#include "kernels.hpp"

namespace {

constexpr int HESS_THREADS = 1024;
// Rotation logs of a batch of QR iterations (per matrix: room for cap
// rotations and scap sweeps, of which at most limit rotations are used before
// the sweep kernel applies the log itself; nsw[b] = sweeps logged and not yet
// applied).
struct QrLog {
  double* c;
  zc* s;
  int* il;
  int* iu;
  int* off;
  int* nsw;
  int* state;    // per matrix: iu, sweeps on the current block, total sweeps, status
  int* pending;  // matrices that stopped on a full log and must be resumed
  long long cap;
  long long limit;
  int scap;
};

// status values of QrLog::state[4b + 3]
constexpr int QR_RUNNING = 0, QR_DONE = 1, QR_FAILED = 2;

constexpr int APPLY_G = 8;     // sweeps per pass of qr_apply_kernel
constexpr int APPLY_TC = 128;  // time steps per LDS-staged chunk of qr_apply_kernel
constexpr int TREVC_TB = 32;   // rows per block of the blocked eigenvector solve
// threads (rows of Z) per workgroup of qr_apply_kernel: at one wave per SIMD,
// 128 leaves fewer idle waves in a matrix's last workgroup than 256 (882
// rows: 7 workgroups, 98% of the lanes live, against 86%)
constexpr int APPLY_NT = 128;

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

// The rotation of givens() from one reciprocal square root t = (|f|^2 h)^-1/2,
// h = |f|^2 + |g|^2: c = |f|^2 t, s = f conj(g) t, r = f h t.  Used when
// 2^-500 <= |f|^2 and h <= 2^500 (every intermediate then stays normal);
// otherwise givens().  v_rsq_f64 is refined by two Newton steps.
__device__ __forceinline__ void givens_rsq(zc f, zc g, double& c, zc& s, zc& r) {
  const double f2 = f.x * f.x + f.y * f.y, g2 = g.x * g.x + g.y * g.y, h = f2 + g2;
  if (f2 >= 0x1p-500 && h <= 0x1p500) {
    const double p = f2 * h;
    double t = __builtin_amdgcn_rsq(p);
    double e = fma(-p * t, t, 1.0);
    t = fma(0.5 * t, e, t);
    e = fma(-p * t, t, 1.0);
    t = fma(0.5 * t, e, t);
    c = f2 * t;
    s = zscale(zmul(f, zconj(g)), t);
    r = zscale(f, h * t);
  } else {
    givens(f, g, c, s, r);
  }
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

// ---------------------------------------------------------------------------
// QR sweeps with deferred updates.  The iteration itself only ever reads the
// active block T[il:iu+1, il:iu+1] (the block only shrinks, or moves to rows
// and columns that no sweep since its last visit has touched), so each sweep
// updates just that block (in windows of R rotations staged in LDS) and logs
// its rotations.  Everything else the sweeps change -- Z, the rows of T above the
// block and the columns of T right of it -- is applied after the iteration by
// qr_apply_kernel, which spreads the rows and columns of every matrix over the
// whole chip (or in the sweep kernel itself when its log fills up).  A pass
// applies G logged sweeps at once in wavefront order: rotation k of sweep a
// acts at time tau = k + 2a, so one pass over a row (or column) applies the G
// sweeps in the order of the one-sweep-at-a-time algorithm while holding only 2G
// consecutive entries, in a circular register window whose slot indices are
// static inside an unrolled block of 2G time steps.  Each element outside the
// active block is read and written once per G sweeps.
// The rotation log of one matrix: sweep q acted on rows/columns il[q]..iu[q]
// with rotations (c, s)[off[q] + k - il[q]], k = il[q]..iu[q]-1.
struct QrLogM {
  double* c;
  zc* s;
  int* il;
  int* iu;
  int* off;
};

__device__ __forceinline__ QrLogM qr_log_of(const QrLog& L, int b) {
  QrLogM m;
  m.c = L.c + (size_t)b * L.cap;
  m.s = L.s + (size_t)b * L.cap;
  m.il = L.il + (size_t)b * L.scap;
  m.iu = L.iu + (size_t)b * L.scap;
  m.off = L.off + (size_t)b * L.scap;
  return m;
}

// Read-only views of one matrix's rotation log in the constant address
// space, for qr_apply_kernel, a launch in which the log is never written:
// uniform indices make these scalar loads into SGPRs.  (The sweep kernel,
// which writes the log, reads it through QrLogM.)
typedef const __attribute__((address_space(4))) double* KDouble;
typedef const __attribute__((address_space(4))) int* KInt;

struct QrLogK {
  KDouble c;
  KDouble s;  // (re, im) pairs
  KInt il;
  KInt iu;
  KInt off;
};

__device__ __forceinline__ QrLogK qr_log_k(const QrLog& L, int b) {
  QrLogK k;
  k.c = (KDouble)(L.c + (size_t)b * L.cap);
  k.s = (KDouble)(L.s + (size_t)b * L.cap);
  k.il = (KInt)(L.il + (size_t)b * L.scap);
  k.iu = (KInt)(L.iu + (size_t)b * L.scap);
  k.off = (KInt)(L.off + (size_t)b * L.scap);
  return k;
}

// One group of m <= G consecutive logged sweeps (first sweep g0).
template <int G>
struct SweepGroup {
  int m, il[G], iu[G], off[G];
  int tmin, tmax, lo, hi, ilmax, iumin;
};

// LogView: QrLogK (scalar loads; only where the log is read-only for the
// whole kernel) or QrLogM (ordinary loads; in the kernel that writes the log).
template <int G, typename LogView>
__device__ __forceinline__ SweepGroup<G> load_group(const LogView& L, int g0, int m) {
  SweepGroup<G> g;
  g.m = m;
  int tmin = 1 << 30, tmax = -1, cmin = 1 << 30, cmax = -1, ilmax = -1, iumin = 1 << 30;
#pragma unroll
  for (int a = 0; a < G; ++a) {
    // sweeps a >= m get an empty range, so they never rotate anything
    g.il[a] = a < m ? L.il[g0 + a] : 0;
    g.iu[a] = a < m ? L.iu[g0 + a] : 0;
    g.off[a] = a < m ? L.off[g0 + a] : 0;
    if (a < m) {
      tmin = min(tmin, g.il[a] + 2 * a);
      tmax = max(tmax, g.iu[a] - 1 + 2 * a);
      cmin = min(cmin, g.il[a]);
      cmax = max(cmax, g.iu[a]);
      ilmax = max(ilmax, g.il[a]);
      iumin = min(iumin, g.iu[a]);
    }
  }
  g.tmin = tmin;
  g.tmax = tmax;
  g.lo = max(cmin, tmin - 2 * (G - 1));  // first column (row) any rotation touches
  g.hi = min(cmax, tmax + 1);            // last one
  g.ilmax = ilmax;
  g.iumin = iumin;
  return g;
}

// LDS table of one chunk of time steps T0 .. T0+TC-1: entry a * TC + u holds
// the rotation of sweep a of the group at time T0 + u, the identity where that
// sweep has none.
template <int G, int TC>
__device__ void fill_table(const SweepGroup<G>& g, const QrLogM& L, int T0, double* tc, zc* ts,
                           int tid, int nt) {
  for (int e = tid; e < G * TC; e += nt) {
    const int a = e / TC, k = T0 + (e % TC) - 2 * a;
    double c = 1.0;
    zc sr = zmake(0, 0);
    int il = 0, iu = 0, off = 0;
#pragma unroll
    for (int q = 0; q < G; ++q)
      if (q == a) {
        il = g.il[q];
        iu = g.iu[q];
        off = g.off[q];
      }
    if (k >= il && k < iu) {
      c = L.c[off + (k - il)];
      sr = L.s[off + (k - il)];
    }
    tc[e] = c;
    ts[e] = sr;
  }
}

// Applies a group of logged sweeps to one row of Z, entries base[c * step],
// c = lo..hi, for the chunk T0 .. T1 of the group's time range, as column
// rotations read from the LDS table.  Rotation k of sweep a acts at time
// tau = k + 2a, so one
// pass applies the group in the order of the one-sweep-at-a-time algorithm while
// holding only the entries in flight, in a circular register window (slot of
// column c: c mod W) whose slot indices are static inside an unrolled block of
// W time steps.  Every step of a block runs unconditionally -- a time with no
// rotation for a sweep meets the identity -- so that the window stays in fixed
// registers.  Entries still in flight at the chunk end are stored back and
// reloaded by the next chunk, so chunks compose exactly.
template <int G, int TC>
__device__ void apply_chunk_strip(const SweepGroup<G>& g, int T0, int T1, zc* base,
                                  long long step, const double* tc, const zc* ts) {
  constexpr int W = 2 * G;
  const int lo = g.lo, hi = g.hi, first = g.tmin - 2 * (G - 1);
  // Columns in flight at the chunk start: T0 - 2(G-1) .. T0.  T0 is a
  // multiple of W, so column T0 - 2(G-1) + q sits in the static slot
  // (q - 2(G-1)) mod W and the loads issue together.  Slots that hold no
  // column of the strip stay zero: identity rotations still combine them, and
  // 1 * x + 0 * y must not meet garbage.
  zc w[W];
  const int cfirst = max(lo, max(T0, g.tmin) - 2 * (G - 1));
#pragma unroll
  for (int q = 0; q <= 2 * (G - 1); ++q) {
    const int c = T0 - 2 * (G - 1) + q;
    w[(q - 2 * (G - 1)) & (W - 1)] =
        (c >= cfirst && c <= min(T0, hi)) ? base[(long long)c * step] : zmake(0, 0);
  }
  w[1] = zmake(0, 0);  // the slot of column T0 + 1, which the first step loads
  // A block's incoming columns tb+1 .. tb+W are loaded one block ahead, so
  // the loads run under the previous block's rotations.  No rotation reaches
  // column cn before step cn - 1 and no earlier step stores it, so the early
  // load reads its final value; its slot is free when step cn - 1 fills it
  // (the slot's previous column, cn - W, was stored at step cn - 2).  A
  // column the preload already holds is reloaded unchanged.
  // Loads are unconditional (no branch, so the compiler keeps the next
  // block's loads in flight): a column outside lo .. hi is read at the
  // nearest end of the range instead; such a slot only ever meets identity
  // rotations (1 * x + 0 * y with finite y) and is never stored.
  zc pre[W];
#pragma unroll
  for (int u = 0; u < W; ++u) {
    const int cn = min(max(T0 + u + 1, lo), hi);
    pre[u] = base[(long long)cn * step];
  }
  int tb = T0;
  for (; tb <= T1; tb += W) {
    zc nxt[W];
#pragma unroll
    for (int u = 0; u < W; ++u) {
      const int cn = min(max(tb + W + u + 1, lo), hi);
      nxt[u] = base[(long long)cn * step];
    }
#pragma unroll
    for (int u = 0; u < W; ++u) {
      const int tau = tb + u;
      w[(u + 1) & (W - 1)] = pre[u];
#pragma unroll
      for (int a = 0; a < G; ++a) {
        const double c = tc[a * TC + u + (tb - T0)];
        const zc sr = ts[a * TC + u + (tb - T0)];
        zc& x = w[(u - 2 * a) & (W - 1)];
        zc& y = w[(u - 2 * a + 1) & (W - 1)];
        const zc xv = x, yv = y;
        x = zadd(zscale(xv, c), zmul(zconj(sr), yv));
        y = zsub(zscale(yv, c), zmul(sr, xv));
      }
      // No rotation reaches column `done` any more: store it before step
      // done + W - 1 reuses its slot (also past T1, where every rotation is
      // the identity and the loads are zero).
      const int done = tau - 2 * (G - 1);
      if (done >= lo && done <= hi && done >= first)
        base[(long long)done * step] = w[(u - 2 * (G - 1)) & (W - 1)];
    }
#pragma unroll
    for (int u = 0; u < W; ++u) pre[u] = nxt[u];
  }
  // Still in flight after the last executed step tend = tb - 1 (tb is a
  // multiple of W): columns tb - 2(G-1) .. tb, again in static slots.
#pragma unroll
  for (int q = 0; q <= 2 * (G - 1); ++q) {
    const int c = tb - 2 * (G - 1) + q;
    if (c >= lo && c <= hi) base[(long long)c * step] = w[(q - 2 * (G - 1)) & (W - 1)];
  }
}

// Applies the rotation log left by the sweep kernel to the Schur vectors:
// blockIdx.y = matrix, each thread one row of Z (the rows are independent,
// so the threads synchronise only to stage the rotation tables).
template <int G, int TC, int NT>
__global__ __launch_bounds__(NT) void qr_apply_kernel(int n, ZMat Zm, QrLog log) {
  __shared__ double tc[G * TC];
  __shared__ zc ts[G * TC];
  const int b = blockIdx.y, tid = threadIdx.x;
  const int row = blockIdx.x * NT + tid;
  zc* zrow = Zm.p + (long long)b * Zm.stride + row;  // entries zrow[c * ld]
  const QrLogK LK = qr_log_k(log, b);
  const QrLogM L = qr_log_of(log, b);
  const int nsw = ((KInt)log.nsw)[b];
  for (int g0 = 0; g0 < nsw; g0 += G) {
    const SweepGroup<G> g = load_group<G>(LK, g0, min(G, nsw - g0));
    const int t0 = (g.tmin / (2 * G)) * (2 * G);
    for (int T0 = t0; T0 <= g.tmax; T0 += TC) {
      const int T1 = min(T0 + TC - 1, g.tmax);
      __syncthreads();
      fill_table<G, TC>(g, L, T0, tc, ts, tid, NT);
      __syncthreads();
      if (row < n) apply_chunk_strip<G, TC>(g, T0, T1, zrow, Zm.ld, tc, ts);
    }
  }
}

// Applies a window's row rotations (rows k0..k0+nr) to one column: v points
// at T[k0, j].
template <int R>
__device__ __forceinline__ void rotate_rows(zc* v, int nr, const double* rc, const zc* rs) {
  zc x[R + 1];
#pragma unroll
  for (int q = 0; q <= R; ++q)
    if (q <= nr) x[q] = v[q];
#pragma unroll
  for (int q = 0; q < R; ++q)
    if (q < nr) {
      const double c = rc[q];
      const zc sr = rs[q], sc = zconj(sr);
      const zc a = x[q], b = x[q + 1];
      x[q] = zadd(zscale(a, c), zmul(sr, b));
      x[q + 1] = zsub(zscale(b, c), zmul(sc, a));
    }
#pragma unroll
  for (int q = 0; q <= R; ++q)
    if (q <= nr) v[q] = x[q];
}

// Applies a window's column rotations (columns k0..k0+nr) to one row: v points
// at T[i, k0], consecutive columns ld apart.
template <int R>
__device__ __forceinline__ void rotate_cols(zc* v, long long ld, int nr, const double* rc,
                                            const zc* rs) {
  zc x[R + 1];
#pragma unroll
  for (int q = 0; q <= R; ++q)
    if (q <= nr) x[q] = v[q * ld];
#pragma unroll
  for (int q = 0; q < R; ++q)
    if (q < nr) {
      const double c = rc[q];
      const zc sr = rs[q], sc = zconj(sr);
      const zc a = x[q], b = x[q + 1];
      x[q] = zadd(zscale(a, c), zmul(sc, b));
      x[q + 1] = zsub(zscale(b, c), zmul(sr, a));
    }
#pragma unroll
  for (int q = 0; q <= R; ++q)
    if (q <= nr) v[q * ld] = x[q];
}

// One workgroup per matrix.  Resumes from the saved state and runs until
// the matrix is in Schur form, the sweep limit is exceeded, or its rotation
// log is full (then it saves its state, counts itself in *pending and stops;
// the log is applied and the kernel launched again).
template <int R, int NT>
__global__ __launch_bounds__(NT) void schur_defer_kernel(int n, ZMat Tm, ZMat Zm, int* info,
                                                         int maxit_per_row, QrLog log) {
  constexpr int LDB = R + 3;
  __shared__ zc band[LDB * LDB];
  __shared__ double rc[2][R];  // rotations of the current and the previous window
  __shared__ zc rs[2][R];
  __shared__ int s_il;
  __shared__ zc s_mu;
  const int b = blockIdx.x;
  zc* T = Tm.p + (long long)b * Tm.stride;
  const QrLogM L = qr_log_of(log, b);
  const long long ld = Tm.ld;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double eps = DBL_EPSILON;
  const long long maxit = (long long)maxit_per_row * (n > 1 ? n : 1);
  int* st = log.state + 4 * b;
  int iu = st[0], iter = st[1], nsw = 0, used = 0;
  long long total = st[2];
  if (st[3] != QR_RUNNING) {
    if (tid == 0) log.nsw[b] = 0;
    return;
  }
  bool failed = false, stopped = false;
  for (;;) {
    if (iu <= 0 || failed) break;
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
    if (nsw == log.scap || (long long)used + (iu - il) > log.limit) {
      // log full (rare: it holds about 2n sweeps): stop before this sweep;
      // the resumed kernel repeats the deflation test and carries on
      stopped = true;
      break;
    }
    ++iter;
    ++total;
    if (total > maxit) {
      if (tid == 0 && info) info[b] = 1;
      failed = true;
      continue;
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
      L.il[nsw] = il;
      L.iu[nsw] = iu;
      L.off[nsw] = used;
    }
    __syncthreads();
    // Windows of R rotations, software-pipelined: wave 0 chases the bulge
    // through window w (band in LDS) while waves 1.. apply the rotations of
    // window w-1 to the rest of the active block.  Window w-1's rows in the
    // columns that band w adds (pbe+1 .. be) are updated by wave 0 itself
    // before it loads band w, so the two groups never touch the same entries:
    // wave 0 works on rows/columns b0..be, the other waves on columns right of
    // be and on rows above pb0.
    int pk0 = 0, pb0 = 0, pbe = 0, pnr = 0;
    bool prev = false;
    for (int k0 = il, w = 0;; ++w) {
      const bool cur = k0 < iu;
      const int k1 = cur ? min(k0 + R, iu) : k0;
      const int nr = k1 - k0;
      const int b0 = k0 > 0 ? k0 - 1 : 0;
      const int be = min(k1 + 1, n - 1);
      const int nb = be - b0 + 1;
      const double* prc = rc[(w + 1) & 1];  // window w-1's rotations
      const zc* prs = rs[(w + 1) & 1];
      const int near_end = cur ? min(be, iu) : pbe;
      if (wave == 0) {
        if (prev) {
          for (int j = pbe + 1 + lane; j <= near_end; j += 64)
            rotate_rows<R>(T + pk0 + (long long)j * ld, pnr, prc, prs);
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
          __builtin_amdgcn_wave_barrier();
        }
        if (cur) {
          for (int idx = lane; idx < nb * nb; idx += 64) {
            const int i = idx % nb, j = idx / nb;
            band[i + j * LDB] = T[(b0 + i) + (long long)(b0 + j) * ld];
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
          double* crc = rc[w & 1];
          zc* crs = rs[w & 1];
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
              crc[k - k0] = c;
              crs[k - k0] = sr;
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
          for (int idx = lane; idx < nb * nb; idx += 64) {
            const int i = idx % nb, j = idx / nb;
            T[(b0 + i) + (long long)(b0 + j) * ld] = band[i + j * LDB];
          }
          if (lane < nr) {
            L.c[used + (k0 - il) + lane] = crc[lane];
            L.s[used + (k0 - il) + lane] = crs[lane];
          }
        }
      } else if (prev) {
        const int t = tid - 64, nt = NT - 64;
        // window w-1's rows right of what wave 0 covers
        for (int j = near_end + 1 + t; j <= iu; j += nt)
          rotate_rows<R>(T + pk0 + (long long)j * ld, pnr, prc, prs);
        // window w-1's columns above its band
        for (int i = il + t; i < pb0; i += nt)
          rotate_cols<R>(T + i + (long long)pk0 * ld, ld, pnr, prc, prs);
      }
      __syncthreads();
      if (!cur) break;
      prev = true;
      pk0 = k0;
      pb0 = b0;
      pbe = be;
      pnr = nr;
      k0 = k1;
    }
    used += iu - il;
    ++nsw;
  }
  if (tid == 0) {
    log.nsw[b] = nsw;
    st[0] = iu;
    st[1] = iter;
    st[2] = (int)total;
    st[3] = stopped ? QR_RUNNING : failed ? QR_FAILED : QR_DONE;
    if (stopped) atomicAdd(log.pending, 1);
  }
}

// The lane index through an opaque move.  Index arithmetic derived from it
// is then recomputed inside the loop that uses it; derived from the plain lane
// index it is loop-invariant, and the compiler hoists it out of every loop
// and holds it in registers for the whole kernel (about 150 registers in the
// strips of the multishift kernel alone).
__device__ __forceinline__ int lane_here() {
  int l;
  asm volatile("v_mov_b32 %0, %1" : "=v"(l) : "v"((int)(threadIdx.x & 63)));
  return l;
}

// Raw buffer access to one matrix: offsets at or past the end read zero and
// drop stores, so masked-off lanes need no branch (and the loads of a batch
// stay in flight together).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t matrix_rsrc(const zc* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<zc*>(p), 0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ zc buf_load(__amdgpu_buffer_rsrc_t r, unsigned off) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  return zmake(__builtin_bit_cast(double, ((unsigned long long)v.y << 32) | v.x),
               __builtin_bit_cast(double, ((unsigned long long)v.w << 32) | v.z));
}

__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, unsigned off, zc z) {
  const unsigned long long a = __builtin_bit_cast(unsigned long long, z.x);
  const unsigned long long b = __builtin_bit_cast(unsigned long long, z.y);
  u32x4 v;
  v.x = (unsigned)a;
  v.y = (unsigned)(a >> 32);
  v.z = (unsigned)b;
  v.w = (unsigned)(b >> 32);
  __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, 0);
}

// The rotations of one window applied in order to RS = R + GAP (NC - 1) + 1
// consecutive entries held in registers: a column's rows (colform false, row
// rotations) or a row's columns (colform true, column rotations, which use
// the conjugate sine).  Entry NC t + c is chain c's rotation at step t and
// acts on entries GAP (NC - 1 - c) + t and the next; entries with no rotation
// hold the identity.
template <int R, int NC, int GAP>
__device__ __forceinline__ void rotate_window(zc* x, bool colform, const double* rc,
                                              const zc* rs) {
#pragma unroll
  for (int t = 0; t < R; ++t)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int e = NC * t + c, o = GAP * (NC - 1 - c) + t;
      const double cs = rc[e];
      zc s = rs[e];
      if (colform) s.y = -s.y;
      const zc a = x[o], b = x[o + 1];
      x[o] = zadd(zscale(a, cs), zmul(s, b));
      x[o + 1] = zsub(zscale(b, cs), zmul(zconj(s), a));
    }
}

// rotate_window for many chains (the multishift kernel): the same
// rotations in the same order, with a scheduling barrier between steps so
// that the NC rotations of a step are loaded from LDS when that step runs
// instead of all R NC of them ahead (which would hold 3 R NC registers).
template <int R, int NC, int GAP>
__device__ __forceinline__ void rotate_window_staged(zc* x, bool colform, const double* rc,
                                                     const zc* rs) {
#pragma unroll
  for (int t = 0; t < R; ++t) {
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int e = NC * t + c, o = GAP * (NC - 1 - c) + t;
      const double cs = rc[e];
      zc s = rs[e];
      if (colform) s.y = -s.y;
      const zc a = x[o], b = x[o + 1];
      x[o] = zadd(zscale(a, cs), zmul(s, b));
      x[o + 1] = zsub(zscale(b, cs), zmul(zconj(s), a));
    }
  }
  __builtin_amdgcn_sched_barrier(0);
}

// rotate_window_staged with the rotations of step t + 1 read from LDS while
// step t computes (the strips of the multishift kernel, whose 64 lanes apply
// every rotation: the LDS latency of a step's NC rotations was exposed once
// per step).
template <int R, int NC, int GAP>
__device__ __forceinline__ void rotate_window_pipe(zc* x, bool colform, const double* rc,
                                                   const zc* rs) {
  double cc[NC];
  zc ss[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    cc[c] = rc[c];
    ss[c] = rs[c];
  }
#pragma unroll
  for (int t = 0; t < R; ++t) {
    __builtin_amdgcn_sched_barrier(0);
    double cn[NC];
    zc sn[NC];
    if (t + 1 < R) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        cn[c] = rc[NC * (t + 1) + c];
        sn[c] = rs[NC * (t + 1) + c];
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int o = GAP * (NC - 1 - c) + t;
      const double cs = cc[c];
      zc s = ss[c];
      if (colform) s.y = -s.y;
      const zc a = x[o], b = x[o + 1];
      x[o] = zadd(zscale(a, cs), zmul(s, b));
      x[o + 1] = zsub(zscale(b, cs), zmul(zconj(s), a));
    }
    if (t + 1 < R) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        cc[c] = cn[c];
        ss[c] = sn[c];
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
}

// One workgroup per matrix: the iteration of schur_defer_kernel (same shifts,
// deflation and rotation log), with the bulge chase pipelined and, for
// NC = 2, two sweeps chased together.
//
// Two chains.  A pass over the active block il..iu runs NC sweeps with the
// same Wilkinson shift: chain 0, and chain 1 (NC = 2, blocks of at least 2R
// rows, not on the sweeps that take an exceptional shift) GAP = 2 positions
// behind.  At step t chain 0
// rotates at position k and chain 1 at k - GAP: their row rotations act on
// different rows, their column rotations on different columns, and a row
// rotation of one commutes with a column rotation of the other, so the pass
// equals the two sweeps one after the other -- which is how the log records
// them, and how the log is applied.
//
// Windows.  Window x holds steps il + xR .. il + xR + R - 1 of chain 0
// (k0(x) = il + xR); its rotations span rows k0 - GP .. k0 + R (GP = GAP
// (NC - 1)); its band is rows/columns b0(x) = k0 - GP - 1 (il for x = 0) ..
// ce(x) = min(k0 + R + 1, iu).  J(x) = (ce(x-1), ce(x)] are the columns band
// x adds.  In step w wave 0
//   - issues the loads for window w+1: rows k0(w) - R - GP .. k0(w) + R of its
//     column of J(w+1) (element e: column e / JR, row e % JR), and band w+1's
//     fresh block (rows/columns ce(w) .. ce(w+1), less the corner);
//   - stores band w-1 and logs window w-1's rotations (their first readers in
//     memory come after the next barrier);
//   - applies window w-1's rotations to the columns of J(w) (loaded a step
//     earlier and held in LDS), stores the rows above band w and puts the
//     GP + 2 rows inside it into LDS; copies the overlap from band w-1 and
//     zeros the entries below the subdiagonal that no other source fills;
//   - chases window w through band w, chain c in lanes c * 64/NC ..;
//   - applies window w-1's rotations to the top rows of J(w+1) and writes
//     band w+1's prefetched block into LDS.
// The other waves apply window w-1's rotations to the columns right of
// ce(w+1) (chunks of 64 columns, loaded row-fastest and turned into one
// column per lane through the wave's LDS tile) and to the rows above b0(w-1)
// (chunks of 64 rows).  Each region has one writer per step.  Memory access
// is through buffer loads/stores whose masked lanes use an offset past the
// matrix (read as zero, stores dropped), so that no branch splits a batch of
// loads; rows outside il .. iu are masked.
template <int R, int NT, int NC>
__global__ __launch_bounds__(NT) void schur_pipe_kernel(int n, ZMat Tm, ZMat Zm, int* info,
                                                        int maxit_per_row, QrLog log) {
  constexpr int GAP = 2, GP = GAP * (NC - 1);
  constexpr int RS = R + GP + 1;   // rows spanned by a window's rotations
  constexpr int NR = NC * R;       // rotation entries per window
  constexpr int LDB = R + GP + 3;  // band side
  constexpr int FS = R + 1;        // side of a band's fresh block, at most
  constexpr int PF = (FS * FS + 63) / 64;
  constexpr int SW = NT / 64 - 1;  // strip waves
  constexpr int JR = 2 * R + GP + 1, JL = (R * JR + 63) / 64;  // J column rows; loads per lane
  constexpr int JS = 2 * R - 1, JSL = (R * JS + 63) / 64;      // J rows stored per step
  constexpr int BL = (LDB * LDB + 63) / 64;
  constexpr int HALF = 64 / NC;  // lanes per chain in the chase
  __shared__ zc band[2][LDB * LDB];
  __shared__ zc stile[SW * 64 * RS];  // each strip wave's transpose tile
  __shared__ zc jcol[R * JR];         // wave 0: column l of J(w) at l * JR
  __shared__ double rc[2][NR];        // rotations of the current and the previous window
  __shared__ zc rs[2][NR];
  __shared__ int s_il;
  __shared__ zc s_mu[2];
  const int b = blockIdx.x;
  zc* T = Tm.p + (long long)b * Tm.stride;
  const QrLogM L = qr_log_of(log, b);
  const long long ld = Tm.ld;
  const unsigned tbytes = (unsigned)(ld * n * 16);  // below 2^31 (checked by the host)
  const __amdgpu_buffer_rsrc_t trs = matrix_rsrc(T, tbytes);
  auto off_of = [&](int i, int j) { return (unsigned)(i + (long long)j * ld) * 16u; };
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double eps = DBL_EPSILON;
  const long long maxit = (long long)maxit_per_row * (n > 1 ? n : 1);
  int* st = log.state + 4 * b;
  int iu = st[0], iter = st[1], nsw = 0, used = 0;
  long long total = st[2];
  if (st[3] != QR_RUNNING) {
    if (tid == 0) log.nsw[b] = 0;
    return;
  }
  bool failed = false, stopped = false;
  for (;;) {
    if (iu <= 0 || failed) break;
    if (tid == 0) s_il = 0;
    __syncthreads();
    for (int k = iu - tid; k >= 1; k -= NT) {
      const double sub = zabs1(T[k + (k - 1) * ld]);
      const double tst = zabs1(T[(k - 1) + (k - 1) * ld]) + zabs1(T[k + k * ld]);
      if (sub <= eps * tst || sub < DBL_MIN) atomicMax(&s_il, k);
    }
    __syncthreads();
    const int il = __builtin_amdgcn_readfirstlane(s_il);
    if (il == iu) {
      if (tid == 0) T[iu + (iu - 1) * ld] = zmake(0, 0);
      --iu;
      iter = 0;
      __syncthreads();
      continue;
    }
    if (nsw + 1 > log.scap || (long long)used + (iu - il) > log.limit) {
      // log full: stop before this pass; the resumed kernel repeats the
      // deflation test and carries on (a fresh launch always has room for
      // one sweep, so every round makes progress)
      stopped = true;
      break;
    }
    // two chained sweeps where the block is large, neither sweep takes an
    // exceptional shift and the log has room for both
    const bool dbl = NC == 2 && iu - il >= 2 * R && iter + 1 != 10 && iter + 2 != 10 &&
                     iter + 1 != 20 && iter + 2 != 20 && nsw + 2 <= log.scap &&
                     (long long)used + 2LL * (iu - il) <= log.limit;
    const int nch = dbl ? 2 : 1;
    iter += nch;
    total += nch;
    if (total > maxit) {
      if (tid == 0 && info) info[b] = 1;
      failed = true;
      continue;
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
      // both chains take the Wilkinson shift: two sweeps with it need about
      // 13% more sweeps than sweeps with the shift updated in between, and
      // far fewer than with the trailing 2x2 block's other eigenvalue for
      // the second sweep (36% more)
      s_mu[0] = mu;
      s_mu[1] = mu;
      for (int c = 0; c < nch; ++c) {
        L.il[nsw + c] = il;
        L.iu[nsw + c] = iu;
        L.off[nsw + c] = used + c * (iu - il);
      }
    }
    __syncthreads();
    const int nw = (iu - il + (dbl ? GAP : 0) + R - 1) / R;
    auto wk0 = [&](int x) { return il + x * R; };
    auto wce = [&](int x) { return x < nw ? min(il + x * R + R + 1, iu) : iu; };
    auto wb0 = [&](int x) { return x > 0 ? il + x * R - GP - 1 : il; };
    // chain c's rotation at step t of the window starting at k0 is active
    auto active = [&](int k0, int t, int c) {
      const int kk = k0 + t - c * GAP;
      return (c == 0 || dbl) && kk >= il && kk <= iu - 1;
    };
    for (int w = 0; w <= nw; ++w) {
      const bool cur = w < nw, prev = w > 0;
      const int k0 = wk0(w), b0 = wb0(w), ce = wce(w), nb = ce - b0 + 1;
      const int pk0 = prev ? wk0(w - 1) : 0, pb0 = prev ? wb0(w - 1) : 0;
      const int pce = prev ? wce(w - 1) : 0;
      const double* prc = rc[(w + 1) & 1];  // window w-1's rotations
      const zc* prs = rs[(w + 1) & 1];
      zc* cb = band[w & 1];
      const zc* pbnd = band[(w + 1) & 1];
      if (wave == 0) {
        // loads for window w+1: J(w+1), rows k0 - R - GP + q
        const int ce1 = wce(w + 1);
        const int nj1 = cur ? ce1 - ce : 0;
        const int jr0 = k0 - R - GP;
        zc yv[JL];
#pragma unroll
        for (int i = 0; i < JL; ++i) {
          const int e = lane + 64 * i, col = e / JR, row = jr0 + e % JR;
          const bool ok = col < nj1 && row >= il && row <= iu;
          yv[i] = buf_load(trs, ok ? off_of(row, ce + 1 + col) : tbytes);
        }
        // band w+1's fresh block: rows/columns ce .. ce1, less the corner
        zc pf[PF];
        const int m1 = ce1 - ce + 1;
        if (w + 1 < nw) {
#pragma unroll
          for (int p = 0; p < PF; ++p) {
            const int idx = lane + 64 * p, i = idx % m1, j = idx / m1;
            pf[p] = buf_load(trs, idx < m1 * m1 && idx != 0 && i <= j + 1
                                      ? off_of(ce + i, ce + j)
                                      : tbytes);
          }
        }
        if (prev) {
          // band w-1 (upper Hessenberg part) and window w-1's rotations
          const int pnb = pce - pb0 + 1;
#pragma unroll
          for (int p = 0; p < BL; ++p) {
            const int idx = lane + 64 * p, i = idx % pnb, j = idx / pnb;
            const bool ok = idx < pnb * pnb && i <= j + 1;
            buf_store(trs, ok ? off_of(pb0 + i, pb0 + j) : tbytes, pbnd[ok ? i + j * LDB : 0]);
          }
          if (lane < NR) {
            const int t = lane / NC, c = lane % NC;
            if (active(pk0, t, c)) {
              const int kk = pk0 + t - c * GAP;
              L.c[used + c * (iu - il) + (kk - il)] = prc[lane];
              L.s[used + c * (iu - il) + (kk - il)] = prs[lane];
            }
          }
        }
        if (cur && prev) {
          // J(w): jcol row q is row k0 - 2R - GP + q; window w-1's rotations
          // on q = R .. 2R + GP, rows q < 2R - 1 to memory, the rest (rows
          // b0 .. k0) into band w
          const int nj = ce - pce;
          if (lane < nj) {
            zc xr[RS];
#pragma unroll
            for (int i = 0; i < RS; ++i) xr[i] = jcol[lane * JR + R + i];
            rotate_window<R, NC, GAP>(xr, false, prc, prs);
#pragma unroll
            for (int i = 0; i < R - 1; ++i) jcol[lane * JR + R + i] = xr[i];
#pragma unroll
            for (int i = R - 1; i < RS; ++i)
              cb[(i - (R - 1)) + (pce + 1 + lane - b0) * LDB] = xr[i];
          }
#pragma unroll
          for (int i = 0; i < JSL; ++i) {
            const int e = lane + 64 * i, col = e / JS, q = e % JS, row = k0 - 2 * R - GP + q;
            const bool ok = col < nj && row >= il;
            buf_store(trs, ok ? off_of(row, pce + 1 + col) : tbytes, jcol[ok ? col * JR + q : 0]);
          }
          // overlap with band w-1: rows/columns b0 .. pce
          const int no = pce - b0 + 1, o = b0 - pb0;
          if (lane < no * no) {
            const int i = lane % no, j = lane / no;
            cb[i + j * LDB] = pbnd[(o + i) + (o + j) * LDB];
          }
          // rows pce+1 .. ce of columns b0 .. pce-1 (below the subdiagonal)
          const int mzr = ce - pce, mzc = pce - b0;
          for (int idx = lane; idx < mzr * mzc; idx += 64)
            cb[(pce + 1 - b0 + idx % mzr) + (idx / mzr) * LDB] = zmake(0, 0);
        } else if (cur) {
          for (int idx = lane; idx < nb * nb; idx += 64) {
            const int i = idx % nb, j = idx / nb;
            cb[i + j * LDB] =
                i <= j + 1 ? T[(b0 + i) + (long long)(b0 + j) * ld] : zmake(0, 0);
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (cur) {
          double* crc = rc[w & 1];
          zc* crs = rs[w & 1];
          const int ch = lane / HALF, hl = lane % HALF;
          for (int t = 0; t < R; ++t) {
            // every lane of a chain computes that chain's rotation
            const int kk = k0 + t - ch * GAP;
            const bool act = active(k0, t, ch);
            double c = 1.0;
            zc sr = zmake(0, 0), r = zmake(0, 0);
            if (act) {
              zc f, g;
              if (kk == il) {
                f = zsub(cb[(kk - b0) + (kk - b0) * LDB], s_mu[ch]);
                g = cb[(kk + 1 - b0) + (kk - b0) * LDB];
              } else {
                f = cb[(kk - b0) + (kk - 1 - b0) * LDB];
                g = cb[(kk + 1 - b0) + (kk - 1 - b0) * LDB];
              }
              givens_rsq(f, g, c, sr, r);
            }
            if (hl == 0) {
              if (act && kk > il) {
                cb[(kk - b0) + (kk - 1 - b0) * LDB] = r;
                cb[(kk + 1 - b0) + (kk - 1 - b0) * LDB] = zmake(0, 0);
              }
              crc[NC * t + ch] = c;
              crs[NC * t + ch] = sr;
            }
            if (act) {
              const zc sc = zconj(sr);
              for (int j = kk + hl; j <= ce; j += HALF) {
                zc* p0 = &cb[(kk - b0) + (j - b0) * LDB];
                const zc x = p0[0], y = p0[1];
                p0[0] = zadd(zscale(x, c), zmul(sr, y));
                p0[1] = zsub(zscale(y, c), zmul(sc, x));
              }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (act) {
              const zc sc = zconj(sr);
              const int ie = min(kk + 2, iu);
              for (int i = b0 + hl; i <= ie; i += HALF) {
                zc* q0 = &cb[(i - b0) + (kk - b0) * LDB];
                const zc x = q0[0], y = q0[LDB];
                q0[0] = zadd(zscale(x, c), zmul(sc, y));
                q0[LDB] = zsub(zscale(y, c), zmul(sr, x));
              }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
          }
        }
        if (nj1 > 0) {
          // J(w+1) into jcol; window w-1's rotations on its rows q = 0 .. RS-1
#pragma unroll
          for (int i = 0; i < JL; ++i)
            if (lane + 64 * i < R * JR) jcol[lane + 64 * i] = yv[i];
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
          if (prev && lane < nj1) {
            zc xr[RS];
#pragma unroll
            for (int i = 0; i < RS; ++i) xr[i] = jcol[lane * JR + i];
            rotate_window<R, NC, GAP>(xr, false, prc, prs);
#pragma unroll
            for (int i = 0; i < RS; ++i) jcol[lane * JR + i] = xr[i];
          }
        }
        if (w + 1 < nw) {
          // band w+1's fresh block (into band w-1's buffer, read for the last
          // time above); its corner row/column ce sits at ce - b0(w+1)
          zc* nbnd = band[(w + 1) & 1];
          const int f0 = ce - wb0(w + 1);
#pragma unroll
          for (int p = 0; p < PF; ++p) {
            const int idx = lane + 64 * p;
            if (idx < m1 * m1 && idx != 0)
              nbnd[(f0 + idx % m1) + (f0 + idx / m1) * LDB] = pf[p];
          }
        }
      } else if (prev) {
        // window w-1's rotations: chunks of 64 columns of the right strip
        // (columns r0 .. iu) and of 64 rows of the top strip (rows il ..
        // pb0-1); a segment covers rows/columns pk0 - GP .. pk0 + R
        const int sw = wave - 1;
        const int r0 = wce(w + 1) + 1;
        const int nright = max(iu - r0 + 1, 0);
        const int ntop = max(pb0 - il, 0);
        const int nrc = (nright + 63) / 64, nch2 = nrc + (ntop + 63) / 64;
        const int s0 = pk0 - GP;
        zc* tile = stile + sw * (64 * RS);
        for (int ch = sw; ch < nch2; ch += SW) {
          zc v[RS];
          unsigned off[RS];
          const bool right = ch < nrc;
          if (right) {
            const int cb0 = r0 + ch * 64, ncol = min(64, iu - cb0 + 1);
#pragma unroll
            for (int i = 0; i < RS; ++i) {
              const int e = lane + 64 * i, col = e / RS, row = s0 + e % RS;
              off[i] = col < ncol && row >= il && row <= iu ? off_of(row, cb0 + col) : tbytes;
            }
          } else {
            const int row = il + (ch - nrc) * 64 + lane;
#pragma unroll
            for (int q = 0; q < RS; ++q) {
              const int col = s0 + q;
              off[q] = row < pb0 && col >= il && col <= iu ? off_of(row, col) : tbytes;
            }
          }
#pragma unroll
          for (int i = 0; i < RS; ++i) v[i] = buf_load(trs, off[i]);
          if (right) {
#pragma unroll
            for (int i = 0; i < RS; ++i) tile[lane + 64 * i] = v[i];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int q = 0; q < RS; ++q) v[q] = tile[lane * RS + q];
          }
          rotate_window<R, NC, GAP>(v, !right, prc, prs);
          if (right) {
#pragma unroll
            for (int q = 0; q < RS; ++q) tile[lane * RS + q] = v[q];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int i = 0; i < RS; ++i) v[i] = tile[lane + 64 * i];
          }
#pragma unroll
          for (int i = 0; i < RS; ++i) buf_store(trs, off[i], v[i]);
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
        }
      }
      __syncthreads();
    }
    used += nch * (iu - il);
    nsw += nch;
  }
  if (tid == 0) {
    log.nsw[b] = nsw;
    st[0] = iu;
    st[1] = iter;
    st[2] = (int)total;
    st[3] = stopped ? QR_RUNNING : failed ? QR_FAILED : QR_DONE;
    if (stopped) atomicAdd(log.pending, 1);
  }
}


constexpr int GAP_MS = 2;  // positions between the chains of a multishift pass

// ---------------------------------------------------------------------------
// Multishift QR with aggressive early deflation (the sweeps of
// schur_ms_kernel).  Each pass of the iteration first looks at the trailing
// NW x NW window of the active block: its complex Schur form T_w = V^H H_w V
// (single-shift QR inside LDS), the spike s V(0, :)^H that the transform
// leaves in the column left of the window (s = H(kwtop, kwtop-1)), and the
// eigenvalues whose spike entry is negligible deflate at once; the others
// are moved to the top of the window (swaps of adjacent diagonal entries), the
// window is reduced back to Hessenberg form, and the bottom NC of them become
// the shifts of the pass: NC single-shift sweeps chained two positions apart.
// This is LAPACK's zlaqr0/zlaqr3 scheme (Braman, Byers and Mathias, 2002)
// with single-shift bulges.  V reaches the Schur vectors as a dense block of
// the rotation log (qr_apply_ms_kernel) and the rows of the active block above
// the window as a product on the host of the workgroup.

// Dense NW x NW unitary blocks of the AED steps, per matrix: block q acts on
// rows/columns k0[q] .. k0[q] + nw[q] - 1 after the first pos[q] logged
// sweeps, V[q] with leading dimension NW.
struct BlockLog {
  zc* V;
  int* pos;
  int* k0;
  int* nw;
  int* nblk;  // blocks logged per matrix in this launch
  int bcap;   // blocks per matrix
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// A double of lane l (uniform) in every lane.
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ zc readlane_z(zc v, int l) {
  return zmake(readlane_f64(v.x, l), readlane_f64(v.y, l));
}

// One lane's share of a rotation in the AED window: the pair (x, y) =
// (p[0], p[stride]) becomes (c x + s1 y, c y - conj(s1) x) -- a row rotation
// (rows q, q+1 of one column: stride 1, s1 = s) or a column rotation (columns
// q, q+1 of one row: stride LDW, s1 = conj(s)).  Every lane runs the same
// instructions; lanes without a pair (act false) read a valid address and do
// not store.  Returns the new x.
__device__ __forceinline__ zc lane_rotate(zc* p, int stride, bool rowform, bool act, double c,
                                          zc s) {
  const zc x = p[0], y = p[stride];
  const zc s1 = zmake(s.x, rowform ? s.y : -s.y);
  const zc xn = zadd(zscale(x, c), zmul(s1, y));
  const zc yn = zsub(zscale(y, c), zmul(zconj(s1), x));
  if (act) {
    p[0] = xn;
    p[stride] = yn;
  }
  return xn;
}

// Complex Schur form of the nw x nw upper Hessenberg Tw (leading dimension
// NW + 1, nw <= NW <= 32) by single-shift QR sweeps, each rotation applied to
// the whole window and accumulated into Vw (starts as the identity).  One
// wave; every lane follows the same control flow.  Wilkinson shifts with the
// exceptional shifts of schur_pipe_kernel at the 10th and 20th sweep on one
// block.  Returns false when the sweep budget runs out; *sweeps counts the
// sweeps done.
//
// Rotation q of a sweep: rows q, q+1 of columns q .. nw-1 (lane l: column
// q+l) together with columns q, q+1 of Vw (lane 32+i: row i), then columns
// q, q+1 of rows 0 .. min(q+2, hi) of Tw (lane i: row i).  The next
// rotation's inputs T(q+1, q), T(q+2, q) are the results of lanes q+1, q+2
// of the second pass, taken by readlane.
template <int NW>
__device__ bool small_schur(zc* Tw, zc* Vw, int nw, int lane, long long budget, int& sweeps) {
  static_assert(NW <= 32, "one lane per window row in each half-wave");
  constexpr int LDW = NW + 1;
  int hi = nw - 1, it = 0;
  const double eps = DBL_EPSILON;
  while (hi > 0) {
    const int k = hi - lane;
    bool neg = false;
    if (k >= 1) {
      const double sub = zabs1(Tw[k + (k - 1) * LDW]);
      const double tst = zabs1(Tw[(k - 1) + (k - 1) * LDW]) + zabs1(Tw[k + k * LDW]);
      neg = sub <= eps * tst || sub < DBL_MIN;
    }
    const unsigned long long msk = __ballot(neg);
    const int lo = msk ? hi - (__ffsll((long long)msk) - 1) : 0;
    wave_sync();
    if (lo == hi) {
      if (lane == 0) Tw[hi + (hi - 1) * LDW] = zmake(0, 0);
      wave_sync();
      --hi;
      it = 0;
      continue;
    }
    if (sweeps >= budget) return false;
    ++sweeps;
    ++it;
    if (lo > 0 && lane == 0) Tw[lo + (lo - 1) * LDW] = zmake(0, 0);
    wave_sync();
    zc mu;
    if (it == 10)
      mu = zadd(Tw[lo + lo * LDW], zmake(0.75 * fabs(Tw[(lo + 1) + lo * LDW].x), 0));
    else if (it == 20)
      mu = zadd(Tw[hi + hi * LDW], zmake(0.75 * fabs(Tw[hi + (hi - 1) * LDW].x), 0));
    else
      mu = wilkinson(Tw[(hi - 1) + (hi - 1) * LDW], Tw[(hi - 1) + hi * LDW], Tw[hi + (hi - 1) * LDW],
                     Tw[hi + hi * LDW]);
    zc f = zsub(Tw[lo + lo * LDW], mu), g = Tw[(lo + 1) + lo * LDW];
    const bool vlane = lane >= 32;
    const int vi = lane - 32;
    for (int q = lo; q < hi; ++q) {
      double c;
      zc s, r;
      givens_rsq(f, g, c, s, r);
      if (q > lo && lane == 0) {
        Tw[q + (q - 1) * LDW] = r;
        Tw[(q + 1) + (q - 1) * LDW] = zmake(0, 0);
      }
      // rows q, q+1 of column q + lane; columns q, q+1 of row vi of Vw
      const bool a1 = vlane ? vi < nw : q + lane < nw;
      zc* const p1 = vlane ? (a1 ? &Vw[vi + q * LDW] : Vw) : (a1 ? &Tw[q + (q + lane) * LDW] : Tw);
      lane_rotate(p1, vlane ? LDW : 1, !vlane, a1, c, s);
      // columns q, q+1 of row `lane` (rows 0 .. min(q + 2, hi))
      const bool a2 = !vlane && lane <= min(q + 2, hi);
      const zc xn = lane_rotate(a2 ? &Tw[lane + q * LDW] : Tw, LDW, false, a2, c, s);
      f = readlane_z(xn, q + 1);
      g = readlane_z(xn, min(q + 2, 31));
    }
    wave_sync();
  }
  return true;
}

// The deflation test of the AED window (zlaqr3): from the bottom, an
// eigenvalue whose spike entry |s| |V(0, j)| is negligible against its own
// size deflates; one that is not is moved up to the top of the undeflated
// entries.  Returns ns, the undeflated count: they occupy 0 .. ns-1, the
// deflated ones ns .. nw-1.
//
// Moving the eigenvalue t at j up to ilst is a chain of swaps of adjacent
// diagonal entries (LAPACK's ztrexc step): swap q takes the rotation of
// (T(q, q+1), t - T(q, q)) to rows q, q+1 right of the pair, to columns q,
// q+1 above it and to Vw; the diagonal entries trade places and T(q, q+1)
// keeps its value.  One rotate per lane per swap, as in small_schur (lane
// l < nw-q-2: column q+2+l; lanes nw-q-2 .. nw-3: rows 0 .. q-1; lanes 32
// ..: the rows of Vw; lane nw-1 stores the diagonal); the next swap's
// T(q-1, q) is formed in every lane from T(q-1, q:q+1), read one swap ahead.
template <int NW>
__device__ int aed_deflate(zc* Tw, zc* Vw, int nw, zc spike, double smlnum, int lane) {
  static_assert(NW <= 32, "one lane per window row in each half-wave");
  constexpr int LDW = NW + 1;
  const double s1 = zabs1(spike);
  int ns = nw, ilst = 0;
  while (ns > ilst) {
    const int j = ns - 1;
    const zc tm = Tw[j + j * LDW];
    double foo = zabs1(tm);
    if (foo == 0.0) foo = s1;
    if (s1 * zabs1(Vw[j * LDW]) <= fmax(smlnum, DBL_EPSILON * foo)) {
      --ns;
      continue;
    }
    zc tq = Tw[(j - 1) + j * LDW];
    for (int q = j - 1; q >= ilst; --q) {
      const zc t11 = Tw[q + q * LDW];
      const int qm = q > 0 ? q - 1 : 0;
      const zc u0 = Tw[qm + q * LDW], u1 = Tw[qm + (q + 1) * LDW];
      const int nA = nw - q - 2;
      const bool rowform = lane < nA;
      const bool act = rowform || (lane >= nA && lane < nw - 2) || (lane >= 32 && lane - 32 < nw);
      const int off = rowform ? q + (q + 2 + lane) * LDW
                      : lane < 32 ? (act ? (lane - nA) + q * LDW : 0)
                                  : (act ? (lane - 32) + q * LDW : 0);
      zc* const p = (lane < 32 ? Tw : Vw) + off;
      double c;
      zc s, r;
      givens_rsq(tq, zsub(tm, t11), c, s, r);
      lane_rotate(p, rowform ? 1 : LDW, rowform, act, c, s);
      if (lane == nw - 1) {
        Tw[q + q * LDW] = tm;
        Tw[(q + 1) + (q + 1) * LDW] = t11;
        Tw[(q + 1) + q * LDW] = zmake(0, 0);
      }
      tq = zadd(zscale(u0, c), zmul(zconj(s), u1));
    }
    wave_sync();
    ++ilst;
  }
  return ns;
}

// Householder reflector (LAPACK's zlarfg convention) of the len entries
// x[0 .. len-1] (LDS, one per lane): H = I - tau v v^H, v[0] = 1, with
// H^H x = beta e1.  v goes to hv[0 .. len-1]; returns tau, beta in *beta.
__device__ zc house(const zc* x, int len, zc* hv, zc* beta, int lane) {
  const zc alpha = x[0];
  double xn = (lane >= 1 && lane < len) ? zabs2(x[lane]) : 0.0;
  for (int off = 32; off > 0; off >>= 1) xn += __shfl_xor(xn, off);
  wave_sync();
  if (xn == 0.0 && alpha.y == 0.0) {
    if (lane < len) hv[lane] = lane == 0 ? zmake(1, 0) : zmake(0, 0);
    *beta = alpha;
    wave_sync();
    return zmake(0, 0);
  }
  const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn);
  const double bt = alpha.x >= 0 ? -nrm : nrm;
  const zc tau = zmake((bt - alpha.x) / bt, -alpha.y / bt);
  const zc scl = zinv(zsub(alpha, zmake(bt, 0)));
  if (lane < len) hv[lane] = lane == 0 ? zmake(1, 0) : zmul(x[lane], scl);
  *beta = zmake(bt, 0);
  wave_sync();
  return tau;
}

// Rows r0 .. r0+len-1 of Tw, columns c0 .. nw-1: Tw <- H^H Tw.  The loops
// run over the static bound NW (terms past len are masked to zero) so that
// their LDS reads issue together; every address stays inside the window's
// LDS area.
template <int NW>
__device__ void house_left(zc* Tw, int nw, int r0, int len, int c0, const zc* hv, zc tau,
                           int lane) {
  constexpr int LDW = NW + 1;
  const int j = c0 + lane;
  if (lane < NW && j < nw) {
    zc w = zmake(0, 0);
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      const zc h = hv[i], t = Tw[(r0 + i) + j * LDW];
      if (i < len) w = zadd(w, zmul(zconj(h), t));
    }
    w = zmul(w, zconj(tau));
#pragma unroll
    for (int i = 0; i < NW; ++i)
      if (i < len) Tw[(r0 + i) + j * LDW] = zsub(Tw[(r0 + i) + j * LDW], zmul(hv[i], w));
  }
  wave_sync();
}

// Columns c0 .. c0+len-1 of Tw (lanes 0 .. nw-1, its rows) and of Vw (lanes
// 32 .. 32+nw-1): M <- M H.  Static bounds as in house_left.
template <int NW>
__device__ void house_right2(zc* Tw, zc* Vw, int nw, int c0, int len, const zc* hv, zc tau,
                             int lane) {
  constexpr int LDW = NW + 1;
  const int i = lane & 31;
  zc* M = lane < 32 ? Tw : Vw;
  if (i < nw) {
    zc w = zmake(0, 0);
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      const zc h = hv[j], m = M[i + (c0 + j) * LDW];
      if (j < len) w = zadd(w, zmul(m, h));
    }
    w = zmul(w, tau);
#pragma unroll
    for (int j = 0; j < NW; ++j)
      if (j < len) M[i + (c0 + j) * LDW] = zsub(M[i + (c0 + j) * LDW], zmul(w, zconj(hv[j])));
  }
  wave_sync();
}

// Back to Hessenberg form after the AED: the spike sv[0 .. ns-1] (the column
// left of the window) and the undeflated block Tw[0:ns, 0:ns].  A reflector
// folds the spike into its first entry (returned: the new H(kwtop,
// kwtop-1)), then reflectors reduce Tw[0:ns, 0:ns] to Hessenberg form; all of
// them act on the whole window and accumulate into Vw.
template <int NW>
__device__ zc aed_rehess(zc* Tw, zc* Vw, zc* sv, zc* hv, zc* xv, int nw, int ns, int lane) {
  constexpr int LDW = NW + 1;
  zc beta;
  zc tau = house(sv, ns, hv, &beta, lane);
  house_left<NW>(Tw, nw, 0, ns, 0, hv, tau, lane);
  house_right2<NW>(Tw, Vw, nw, 0, ns, hv, tau, lane);
  const zc spike = beta;
  for (int k = 0; k + 2 < ns; ++k) {
    const int len = ns - k - 1;
    if (lane < len) xv[lane] = Tw[(k + 1 + lane) + k * LDW];
    wave_sync();
    zc bk;
    tau = house(xv, len, hv, &bk, lane);
    if (lane < len) Tw[(k + 1 + lane) + k * LDW] = lane == 0 ? bk : zmake(0, 0);
    wave_sync();
    house_left<NW>(Tw, nw, k + 1, len, k + 1, hv, tau, lane);
    house_right2<NW>(Tw, Vw, nw, k + 1, len, hv, tau, lane);
  }
  return spike;
}

// One workgroup per matrix: multishift QR with aggressive early deflation.
// Each pass: the AED step on the trailing aw x aw window of the active block
// (aw <= NW; or the whole block once it has at most NW rows: its Schur form
// is computed in LDS and the block is done), then, unless it deflated more
// than nib% of the window, a pass of nch <= NC single-shift sweeps with the bottom undeflated window
// eigenvalues as shifts, chained GAP = 2 positions apart (chain c rotates at
// k - c GAP at step t while chain 0 is at k), chased through the LDS band and
// applied to the strips exactly as in schur_pipe_kernel, whose pipelining,
// band layout and strip geometry this kernel shares (with nch chains in
// place of two).  Chains that trail by GAP positions act on disjoint rows and
// columns, and a row rotation commutes with a column rotation of another
// chain, so a pass equals its sweeps one after the other, which is how the
// rotation log records them.  The dense V of every AED step goes to the
// block log, ordered with the sweeps (pos = sweeps logged before it).
template <int R, int NT, int NC, int NW>
__global__ __launch_bounds__(NT) void schur_ms_kernel(int n, ZMat Tm, ZMat Zm, int* info,
                                                      int maxit_per_row, QrLog log, BlockLog blog,
                                                      int aedw, int nib) {
  static_assert(GAP_MS * (NC - 1) + 1 <= R, "chains must fit behind one window step");
  static_assert(NW <= 64 && 2 * R <= NW + 1, "AED window: one lane per row, passes beyond it");
  // the window supplies the shifts and fits the LDS area
  const int AW = max(NC, min(aedw, NW));
  constexpr int GAP = GAP_MS, GP = GAP * (NC - 1);
  constexpr int RS = R + GP + 1;   // rows spanned by a window's rotations
  constexpr int NR = NC * R;       // rotation entries per window
  constexpr int LDB = R + GP + 3;  // band side
  constexpr int FS = R + 1;        // side of a band's fresh block, at most
  constexpr int PF = (FS * FS + 63) / 64;
  constexpr int SW = NT / 64 - 1;  // strip waves
  constexpr int JR = 2 * R + GP + 1, JL = (R * JR + 63) / 64;  // J column rows; loads per lane
  constexpr int JS = 2 * R - 1, JSL = (R * JS + 63) / 64;      // J rows stored per step
  constexpr int BL = (LDB * LDB + 63) / 64;
  constexpr int HALF = 64 / NC;  // lanes per chain in the chase
  __shared__ zc band[2][LDB * LDB];
  // each strip wave's transpose tile; between passes the AED window, its
  // Schur vectors and two reflector vectors
  constexpr int LDW = NW + 1;  // leading dimension of the AED window in LDS
  constexpr int TILE = SW * 64 * RS, AEDL = 2 * NW * LDW + 3 * NW;
  __shared__ zc stile[TILE > AEDL ? TILE : AEDL];
  __shared__ zc jcol[R * JR];         // wave 0: column l of J(w) at l * JR
  __shared__ double rc[2][NR];        // rotations of the current and the previous window
  __shared__ zc rs[2][NR];
  __shared__ int s_il;
  __shared__ zc s_mu[NC];
  __shared__ int s_aed[3];
  __shared__ zc s_spike;
  const int b = blockIdx.x;
  zc* T = Tm.p + (long long)b * Tm.stride;
  const QrLogM L = qr_log_of(log, b);
  const long long ld = Tm.ld;
  const unsigned tbytes = (unsigned)(ld * n * 16);  // below 2^31 (checked by the host)
  const __amdgpu_buffer_rsrc_t trs = matrix_rsrc(T, tbytes);
  auto off_of = [&](int i, int j) { return (unsigned)(i + (long long)j * ld) * 16u; };
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double eps = DBL_EPSILON;
  const long long maxit = (long long)maxit_per_row * (n > 1 ? n : 1);
  int* st = log.state + 4 * b;
  int iu = st[0], iter = st[1], nsw = 0, used = 0, nblk = 0;
  const double smlnum = DBL_MIN * ((double)n / DBL_EPSILON);
  zc* const Tw = stile;
  zc* const Vw = stile + NW * LDW;
  zc* const svw = Vw + NW * LDW;
  zc* const hvw = svw + NW;
  zc* const xvw = hvw + NW;
  zc* const BV = blog.V + (long long)b * blog.bcap * NW * NW;
  long long total = st[2];
  if (st[3] != QR_RUNNING) {
    if (tid == 0) {
      log.nsw[b] = 0;
      blog.nblk[b] = 0;
    }
    return;
  }
  bool failed = false, stopped = false;
  for (;;) {
    if (iu <= 0 || failed) break;
    if (tid == 0) s_il = 0;
    __syncthreads();
    for (int k = iu - tid; k >= 1; k -= NT) {
      const double sub = zabs1(T[k + (k - 1) * ld]);
      const double tst = zabs1(T[(k - 1) + (k - 1) * ld]) + zabs1(T[k + k * ld]);
      if (sub <= eps * tst || sub < DBL_MIN) atomicMax(&s_il, k);
    }
    __syncthreads();
    const int il = __builtin_amdgcn_readfirstlane(s_il);
    if (il == iu) {
      if (tid == 0) T[iu + (iu - 1) * ld] = zmake(0, 0);
      --iu;
      iter = 0;
      __syncthreads();
      continue;
    }
    if (nblk + 1 > blog.bcap || nsw + 1 > log.scap ||
        (long long)used + (iu - il) > log.limit) {
      // a log is full (no room for the AED block and one sweep after it):
      // stop here; the logs are applied and the resumed kernel carries on.
      // A fresh launch always has room for both, so every launch makes
      // progress.
      stopped = true;
      break;
    }
    // ---- aggressive early deflation on the trailing window (the whole
    // active block once it fits one)
    const int mblk = iu - il + 1;
    const bool whole = mblk <= NW;
    const int aw = whole ? mblk : AW;
    const int kwtop = iu - aw + 1;
    if (tid == 0 && il > 0) T[il + (il - 1) * ld] = zmake(0, 0);
    for (int idx = tid; idx < aw * aw; idx += NT) {
      const int i = idx % aw, j = idx / aw;
      Tw[i + j * LDW] = i <= j + 1 ? T[(kwtop + i) + (long long)(kwtop + j) * ld] : zmake(0, 0);
      Vw[i + j * LDW] = zmake(i == j ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    if (wave == 0) {
      // The window's own QR gets LAPACK zlahqr's budget, 30 max(10, nw)
      // sweeps; a finished block's sweeps also count against the matrix's
      // limit (maxit_per_row * n), an AED window's do not: when the window
      // does not converge it deflates nothing, and its diagonal still serves
      // as shifts
      int sw = 0;
      const long long own = 30LL * max(10, aw);
      const bool conv =
          small_schur<NW>(Tw, Vw, aw, lane, whole ? min(own, maxit - total) : own, sw);
      int ns = aw;
      zc spike = zmake(0, 0);
      if (conv && whole) {
        ns = 0;
      } else if (!whole) {
        spike = T[kwtop + (long long)(kwtop - 1) * ld];
        if (conv) ns = aed_deflate<NW>(Tw, Vw, aw, spike, smlnum, lane);
        // the bottom NC undeflated eigenvalues of the window are the shifts
        const int take = min(NC, ns);
        if (lane < take) s_mu[lane] = Tw[(ns - take + lane) * (LDW + 1)];
        if (ns > 0 && lane < ns) svw[lane] = zmul(spike, zconj(Vw[lane * LDW]));
        wave_sync();
        if (ns > 1)
          spike = aed_rehess<NW>(Tw, Vw, svw, hvw, xvw, aw, ns, lane);
        else
          spike = ns == 1 ? svw[0] : zmake(0, 0);
      }
      if (lane == 0) {
        s_aed[0] = ns;
        s_aed[1] = whole ? sw : 0;
        s_aed[2] = conv || !whole ? 1 : 0;
        s_spike = spike;
      }
    }
    __syncthreads();
    const int ns = __builtin_amdgcn_readfirstlane(s_aed[0]);
    total += __builtin_amdgcn_readfirstlane(s_aed[1]);
    if (!s_aed[2]) {
      // the sweep budget ran out on a finished block
      if (tid == 0 && info) info[b] = 1;
      failed = true;
      continue;
    }
    if (!whole) {
      // the window and the spike back into the active block; the rows of the
      // active block above the window times V
      for (int idx = tid; idx < aw * aw; idx += NT) {
        const int i = idx % aw, j = idx / aw;
        if (i <= j + 1) T[(kwtop + i) + (long long)(kwtop + j) * ld] = Tw[i + j * LDW];
      }
      if (tid == 0) T[kwtop + (long long)(kwtop - 1) * ld] = s_spike;
      for (int r = il + tid; r < kwtop; r += NT) {
        zc* trow = T + r + (long long)kwtop * ld;
        // aw == AW here (the window is not the whole block); entries past it
        // are zero in x and never read from V
        zc x[NW];
#pragma unroll
        for (int q = 0; q < NW; ++q) x[q] = q < aw ? trow[(long long)min(q, aw - 1) * ld] : zmake(0, 0);
#pragma unroll 1
        for (int j = 0; j < aw; ++j) {
          zc acc = zmake(0, 0);
#pragma unroll
          for (int q = 0; q < NW; ++q)
            acc = zadd(acc, zmul(x[q], q < aw ? Vw[q + j * LDW] : zmake(0, 0)));
          trow[j * ld] = acc;
        }
      }
    }
    // log V for the Schur vectors
    for (int idx = tid; idx < NW * NW; idx += NT)
      BV[(long long)nblk * NW * NW + idx] = Vw[(idx % NW) + (idx / NW) * LDW];
    if (tid == 0) {
      blog.pos[(long long)b * blog.bcap + nblk] = nsw;
      blog.k0[(long long)b * blog.bcap + nblk] = kwtop;
      blog.nw[(long long)b * blog.bcap + nblk] = aw;
    }
    ++nblk;
    __syncthreads();
    if (whole) {
      iu = il - 1;
      iter = 0;
      continue;
    }
    const int nd = aw - ns;
    iu = kwtop + ns - 1;
    iter = nd > 0 ? 0 : iter + 1;
    // many deflations: look at the next window before sweeping (zlaqr0's
    // nibble point, there 14%); a block that now fits NW rows is finished in
    // the window
    if ((nd > 0 && 100 * nd > nib * aw) || ns == 0 || iu - il + 1 <= NW) continue;
    // as many chains as shifts, and as the rotation log still holds (at least
    // one: the check above left room for a sweep over the larger block)
    const int nch = (int)min((long long)min(NC, ns),
                             min((long long)(log.scap - nsw), (log.limit - used) / (iu - il)));
    total += nch;
    if (total > maxit) {
      if (tid == 0 && info) info[b] = 1;
      failed = true;
      continue;
    }
    if (tid == 0) {
      if (iter > 0 && iter % 6 == 0) {
        // no deflation for six passes: exceptional shifts off the bottom
        // diagonal entries
        for (int c = 0; c < nch; ++c) {
          const int k = iu - c;
          s_mu[c] = zadd(T[k + (long long)k * ld],
                         zmake(0.75 * fabs(T[k + (long long)(k - 1) * ld].x), 0));
        }
      }
      for (int c = 0; c < nch; ++c) {
        L.il[nsw + c] = il;
        L.iu[nsw + c] = iu;
        L.off[nsw + c] = used + c * (iu - il);
      }
    }
    __syncthreads();
    const int nw = (iu - il + GAP * (nch - 1) + R - 1) / R;
    auto wk0 = [&](int x) { return il + x * R; };
    auto wce = [&](int x) { return x < nw ? min(il + x * R + R + 1, iu) : iu; };
    auto wb0 = [&](int x) { return x > 0 ? il + x * R - GP - 1 : il; };
    // chain c's rotation at step t of the window starting at k0 is active
    auto active = [&](int k0, int t, int c) {
      const int kk = k0 + t - c * GAP;
      return c < nch && kk >= il && kk <= iu - 1;
    };
    for (int w = 0; w <= nw; ++w) {
      const bool cur = w < nw, prev = w > 0;
      const int k0 = wk0(w), b0 = wb0(w), ce = wce(w), nb = ce - b0 + 1;
      const int pk0 = prev ? wk0(w - 1) : 0, pb0 = prev ? wb0(w - 1) : 0;
      const int pce = prev ? wce(w - 1) : 0;
      const double* prc = rc[(w + 1) & 1];  // window w-1's rotations
      const zc* prs = rs[(w + 1) & 1];
      zc* cb = band[w & 1];
      const zc* pbnd = band[(w + 1) & 1];
      if (wave == 0) {
        const int lane = lane_here();
        // loads for window w+1: J(w+1), rows k0 - R - GP + q
        const int ce1 = wce(w + 1);
        const int nj1 = cur ? ce1 - ce : 0;
        const int jr0 = k0 - R - GP;
        zc yv[JL];
#pragma unroll
        for (int i = 0; i < JL; ++i) {
          const int e = lane + 64 * i, col = e / JR, row = jr0 + e % JR;
          const bool ok = col < nj1 && row >= il && row <= iu;
          yv[i] = buf_load(trs, ok ? off_of(row, ce + 1 + col) : tbytes);
        }
        // band w+1's fresh block: rows/columns ce .. ce1, less the corner
        zc pf[PF];
        const int m1 = ce1 - ce + 1;
        if (w + 1 < nw) {
#pragma unroll
          for (int p = 0; p < PF; ++p) {
            const int idx = lane + 64 * p, i = idx % m1, j = idx / m1;
            pf[p] = buf_load(trs, idx < m1 * m1 && idx != 0 && i <= j + 1
                                      ? off_of(ce + i, ce + j)
                                      : tbytes);
          }
        }
        if (prev) {
          // band w-1 (upper Hessenberg part) and window w-1's rotations
          const int pnb = pce - pb0 + 1;
#pragma unroll
          for (int p = 0; p < BL; ++p) {
            const int idx = lane + 64 * p, i = idx % pnb, j = idx / pnb;
            const bool ok = idx < pnb * pnb && i <= j + 1;
            buf_store(trs, ok ? off_of(pb0 + i, pb0 + j) : tbytes, pbnd[ok ? i + j * LDB : 0]);
          }
          for (int e = lane; e < NR; e += 64) {
            const int t = e / NC, c = e % NC;
            if (active(pk0, t, c)) {
              const int kk = pk0 + t - c * GAP;
              L.c[used + c * (iu - il) + (kk - il)] = prc[e];
              L.s[used + c * (iu - il) + (kk - il)] = prs[e];
            }
          }
        }
        if (cur && prev) {
          // J(w): jcol row q is row k0 - 2R - GP + q; window w-1's rotations
          // on q = R .. 2R + GP, rows q < 2R - 1 to memory, the rest (rows
          // b0 .. k0) into band w
          const int nj = ce - pce;
          if (lane < nj) {
            zc xr[RS];
#pragma unroll
            for (int i = 0; i < RS; ++i) xr[i] = jcol[lane * JR + R + i];
            rotate_window_staged<R, NC, GAP>(xr, false, prc, prs);
#pragma unroll
            for (int i = 0; i < R - 1; ++i) jcol[lane * JR + R + i] = xr[i];
#pragma unroll
            for (int i = R - 1; i < RS; ++i)
              cb[(i - (R - 1)) + (pce + 1 + lane - b0) * LDB] = xr[i];
          }
#pragma unroll
          for (int i = 0; i < JSL; ++i) {
            const int e = lane + 64 * i, col = e / JS, q = e % JS, row = k0 - 2 * R - GP + q;
            const bool ok = col < nj && row >= il;
            buf_store(trs, ok ? off_of(row, pce + 1 + col) : tbytes, jcol[ok ? col * JR + q : 0]);
          }
          // overlap with band w-1: rows/columns b0 .. pce
          const int no = pce - b0 + 1, o = b0 - pb0;
          for (int idx = lane; idx < no * no; idx += 64) {
            const int i = idx % no, j = idx / no;
            cb[i + j * LDB] = pbnd[(o + i) + (o + j) * LDB];
          }
          // rows pce+1 .. ce of columns b0 .. pce-1 (below the subdiagonal)
          const int mzr = ce - pce, mzc = pce - b0;
          for (int idx = lane; idx < mzr * mzc; idx += 64)
            cb[(pce + 1 - b0 + idx % mzr) + (idx / mzr) * LDB] = zmake(0, 0);
        } else if (cur) {
          for (int idx = lane; idx < nb * nb; idx += 64) {
            const int i = idx % nb, j = idx / nb;
            cb[i + j * LDB] =
                i <= j + 1 ? T[(b0 + i) + (long long)(b0 + j) * ld] : zmake(0, 0);
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (cur) {
          double* crc = rc[w & 1];
          zc* crs = rs[w & 1];
          const int ch = lane / HALF, hl = lane % HALF;
          for (int t = 0; t < R; ++t) {
            // every lane of a chain computes that chain's rotation
            const int kk = k0 + t - ch * GAP;
            const bool act = active(k0, t, ch);
            double c = 1.0;
            zc sr = zmake(0, 0), r = zmake(0, 0);
            if (act) {
              zc f, g;
              if (kk == il) {
                f = zsub(cb[(kk - b0) + (kk - b0) * LDB], s_mu[ch]);
                g = cb[(kk + 1 - b0) + (kk - b0) * LDB];
              } else {
                f = cb[(kk - b0) + (kk - 1 - b0) * LDB];
                g = cb[(kk + 1 - b0) + (kk - 1 - b0) * LDB];
              }
              givens_rsq(f, g, c, sr, r);
            }
            if (hl == 0) {
              if (act && kk > il) {
                cb[(kk - b0) + (kk - 1 - b0) * LDB] = r;
                cb[(kk + 1 - b0) + (kk - 1 - b0) * LDB] = zmake(0, 0);
              }
              crc[NC * t + ch] = c;
              crs[NC * t + ch] = sr;
            }
            if (act) {
              const zc sc = zconj(sr);
              for (int j = kk + hl; j <= ce; j += HALF) {
                zc* p0 = &cb[(kk - b0) + (j - b0) * LDB];
                const zc x = p0[0], y = p0[1];
                p0[0] = zadd(zscale(x, c), zmul(sr, y));
                p0[1] = zsub(zscale(y, c), zmul(sc, x));
              }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (act) {
              const zc sc = zconj(sr);
              const int ie = min(kk + 2, iu);
              for (int i = b0 + hl; i <= ie; i += HALF) {
                zc* q0 = &cb[(i - b0) + (kk - b0) * LDB];
                const zc x = q0[0], y = q0[LDB];
                q0[0] = zadd(zscale(x, c), zmul(sc, y));
                q0[LDB] = zsub(zscale(y, c), zmul(sr, x));
              }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
          }
        }
        if (nj1 > 0) {
          // J(w+1) into jcol; window w-1's rotations on its rows q = 0 .. RS-1
#pragma unroll
          for (int i = 0; i < JL; ++i)
            if (lane + 64 * i < R * JR) jcol[lane + 64 * i] = yv[i];
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
          if (prev && lane < nj1) {
            zc xr[RS];
#pragma unroll
            for (int i = 0; i < RS; ++i) xr[i] = jcol[lane * JR + i];
            rotate_window_staged<R, NC, GAP>(xr, false, prc, prs);
#pragma unroll
            for (int i = 0; i < RS; ++i) jcol[lane * JR + i] = xr[i];
          }
        }
        if (w + 1 < nw) {
          // band w+1's fresh block (into band w-1's buffer, read for the last
          // time above); its corner row/column ce sits at ce - b0(w+1)
          zc* nbnd = band[(w + 1) & 1];
          const int f0 = ce - wb0(w + 1);
#pragma unroll
          for (int p = 0; p < PF; ++p) {
            const int idx = lane + 64 * p;
            if (idx < m1 * m1 && idx != 0)
              nbnd[(f0 + idx % m1) + (f0 + idx / m1) * LDB] = pf[p];
          }
        }
      } else if (prev) {
        // window w-1's rotations: chunks of 64 columns of the right strip
        // (columns r0 .. iu) and of 64 rows of the top strip (rows il ..
        // pb0-1); a segment covers rows/columns pk0 - GP .. pk0 + R
        const int sw = wave - 1;
        const int r0 = wce(w + 1) + 1;
        const int nright = max(iu - r0 + 1, 0);
        const int ntop = max(pb0 - il, 0);
        const int nrc = (nright + 63) / 64, nch2 = nrc + (ntop + 63) / 64;
        const int s0 = pk0 - GP;
        zc* tile = stile + sw * (64 * RS);
        for (int ch = sw; ch < nch2; ch += SW) {
          const int lane = lane_here();
          zc v[RS];
          unsigned off[RS];
          const bool right = ch < nrc;
          if (right) {
            const int cb0 = r0 + ch * 64, ncol = min(64, iu - cb0 + 1);
#pragma unroll
            for (int i = 0; i < RS; ++i) {
              const int e = lane + 64 * i, col = e / RS, row = s0 + e % RS;
              off[i] = col < ncol && row >= il && row <= iu ? off_of(row, cb0 + col) : tbytes;
            }
          } else {
            const int row = il + (ch - nrc) * 64 + lane;
#pragma unroll
            for (int q = 0; q < RS; ++q) {
              const int col = s0 + q;
              off[q] = row < pb0 && col >= il && col <= iu ? off_of(row, col) : tbytes;
            }
          }
#pragma unroll
          for (int i = 0; i < RS; ++i) v[i] = buf_load(trs, off[i]);
          if (right) {
#pragma unroll
            for (int i = 0; i < RS; ++i) tile[lane + 64 * i] = v[i];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int q = 0; q < RS; ++q) v[q] = tile[lane * RS + q];
          }
          rotate_window_pipe<R, NC, GAP>(v, !right, prc, prs);
          if (right) {
#pragma unroll
            for (int q = 0; q < RS; ++q) tile[lane * RS + q] = v[q];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int i = 0; i < RS; ++i) v[i] = tile[lane + 64 * i];
          }
#pragma unroll
          for (int i = 0; i < RS; ++i) buf_store(trs, off[i], v[i]);
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
        }
      }
      __syncthreads();
    }
    used += nch * (iu - il);
    nsw += nch;
  }
  if (tid == 0) {
    log.nsw[b] = nsw;
    blog.nblk[b] = nblk;
    st[0] = iu;
    st[1] = iter;
    st[2] = (int)total;
    st[3] = stopped ? QR_RUNNING : failed ? QR_FAILED : QR_DONE;
    if (stopped) atomicAdd(log.pending, 1);
  }
}


// The logs of schur_ms_kernel applied to the Schur vectors: the logged
// sweeps in groups of at most G (never across a block), and before sweep
// pos[q] the dense block q (Z[:, k0 .. k0+nw-1] <- Z[:, k0 .. k0+nw-1] V),
// each thread one row of Z.
template <int G, int TC, int NT, int NW>
__global__ __launch_bounds__(NT) void qr_apply_ms_kernel(int n, ZMat Zm, QrLog log,
                                                         BlockLog blog) {
  __shared__ double tc[G * TC];
  __shared__ zc ts[G * TC];
  __shared__ zc vb[NW * NW];
  const int b = blockIdx.y, tid = threadIdx.x;
  const int row = blockIdx.x * NT + tid;
  zc* zrow = Zm.p + (long long)b * Zm.stride + row;  // entries zrow[c * ld]
  const QrLogK LK = qr_log_k(log, b);
  const QrLogM L = qr_log_of(log, b);
  const int nsw = ((KInt)log.nsw)[b];
  const int nblk = ((KInt)blog.nblk)[b];
  const KInt bpos = (KInt)(blog.pos + (long long)b * blog.bcap);
  const KInt bk0 = (KInt)(blog.k0 + (long long)b * blog.bcap);
  const KInt bnw = (KInt)(blog.nw + (long long)b * blog.bcap);
  const zc* BV = blog.V + (long long)b * blog.bcap * NW * NW;
  int g0 = 0;
  for (int q = 0; q <= nblk; ++q) {
    const int end = q < nblk ? bpos[q] : nsw;
    for (; g0 < end; g0 += G) {
      const SweepGroup<G> g = load_group<G>(LK, g0, min(G, end - g0));
      const int t0 = (g.tmin / (2 * G)) * (2 * G);
      for (int T0 = t0; T0 <= g.tmax; T0 += TC) {
        const int T1 = min(T0 + TC - 1, g.tmax);
        __syncthreads();
        fill_table<G, TC>(g, L, T0, tc, ts, tid, NT);
        __syncthreads();
        if (row < n) apply_chunk_strip<G, TC>(g, T0, T1, zrow, Zm.ld, tc, ts);
      }
    }
    g0 = end;
    if (q == nblk) break;
    const int k0 = bk0[q], aw = bnw[q];
    __syncthreads();
    for (int e = tid; e < NW * NW; e += NT) {
      const int i = e % NW, j = e / NW;
      vb[e] = i < aw && j < aw ? BV[(long long)q * NW * NW + e] : zmake(0, 0);
    }
    __syncthreads();
    if (row < n) {
      zc x[NW];
#pragma unroll
      for (int i = 0; i < NW; ++i) x[i] = i < aw ? zrow[(long long)(k0 + i) * Zm.ld] : zmake(0, 0);
      for (int j = 0; j < aw; ++j) {
        zc acc = zmake(0, 0);
#pragma unroll
        for (int i = 0; i < NW; ++i) acc = zadd(acc, zmul(x[i], vb[i + j * NW]));
        zrow[(long long)(k0 + j) * Zm.ld] = acc;
      }
    }
  }
}

__global__ void qr_state_init_kernel(int n, int batch, int* state, int* pending) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b == 0) *pending = 0;
  if (b < batch) {
    state[4 * b + 0] = n - 1;
    state[4 * b + 1] = 0;
    state[4 * b + 2] = 0;
    state[4 * b + 3] = QR_RUNNING;
  }
}

__global__ void qr_clear_pending_kernel(int* pending) { *pending = 0; }

// Eigenvectors of the upper triangular T, blocked.  X is upper triangular
// with X[k][k] = 1 and, for i < k, (T[i][i] - T[k][k]) X[i][k] =
// -sum_{i < m <= k} T[i][m] X[m][k] (back-substitution, the divisor raised to
// smin where it is tiny, as LAPACK's ztrevc does).  Rows are taken in blocks
// of TB from the bottom: trevc_diag_kernel solves block I for every column
// right of its start (the columns are independent), then one ZGEMM subtracts
// T[0:I0, I] X[I, I0:n] from the rows above, so the O(n^3) part of the work
// runs on the matrix cores.  trevc_init_kernel sets X[k][k] = 1, X[i][k] =
// -T[i][k] for the rows i < k of k's own block (the term m = k for the rows
// of other blocks arrives with the update, through X[k][k] = 1) and 0
// elsewhere.
template <int TB>
__global__ void trevc_init_kernel(int n, ZCMat Tm, ZMat Xm) {
  const int b = blockIdx.y;
  const zc* T = Tm.p + (long long)b * Tm.stride;
  zc* X = Xm.p + (long long)b * Xm.stride;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < (long long)n * n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % n), k = (int)(idx / n);
    zc v = zmake(0, 0);
    if (i == k)
      v = zmake(1, 0);
    else if (i < k && i >= (k / TB) * TB)
      v = zneg(T[i + (long long)k * Tm.ld]);
    X[i + (long long)k * Xm.ld] = v;
  }
}

// Rows I0 .. I0+TB-1 of every column k > I0: one thread per column, its TB
// unknowns in registers, the diagonal block of T in LDS (read by all lanes
// at one address at a time).  UNIT: T has a unit diagonal and no shift is
// taken (the columns of T^-1 by the same substitution: T u_k = e_k).
template <int TB, bool UNIT>
__global__ __launch_bounds__(256) void trevc_diag_kernel(int n, int I0, ZCMat Tm, ZMat Xm) {
  __shared__ zc tb[TB * TB];
  const int b = blockIdx.y, tid = threadIdx.x;
  const zc* T = Tm.p + (long long)b * Tm.stride;
  zc* X = Xm.p + (long long)b * Xm.stride;
  const long long ld = Tm.ld, ldx = Xm.ld;
  const int nbk = min(TB, n - I0);
  for (int e = tid; e < TB * TB; e += 256) {
    const int i = e % TB, j = e / TB;
    tb[e] = (i < nbk && j < nbk) ? T[(I0 + i) + (long long)(I0 + j) * ld] : zmake(0, 0);
  }
  __syncthreads();
  const int k = I0 + 1 + blockIdx.x * 256 + tid;
  if (k >= n) return;
  const zc lk = UNIT ? zmake(0, 0) : T[k + (long long)k * ld];
  const double smlnum = DBL_MIN * ((double)n / DBL_EPSILON);
  const double smin = fmax(DBL_EPSILON * zabs1(lk), smlnum);
  const int iend = min(k, I0 + nbk) - I0;  // unknown rows I0 .. I0+iend-1
  zc* xc = X + (long long)k * ldx + I0;
  zc x[TB];
#pragma unroll
  for (int i = 0; i < TB; ++i) x[i] = i < iend ? xc[i] : zmake(0, 0);
#pragma unroll
  for (int i = TB - 1; i >= 0; --i) {
    if (i < iend) {
      zc xi = x[i];
      if (!UNIT) {
        zc d = zsub(tb[i + i * TB], lk);
        if (zabs1(d) < smin) d = zmake(smin, 0.0);
        xi = zdiv(xi, d);
      }
      x[i] = xi;
#pragma unroll
      for (int m = 0; m < i; ++m) x[m] = zsub(x[m], zmul(tb[m + i * TB], xi));
    }
  }
#pragma unroll
  for (int i = 0; i < TB; ++i)
    if (i < iend) xc[i] = x[i];
}

// out[b n + j] = (|X[:, j]|_2, 0)
__global__ __launch_bounds__(256) void col_norms_kernel(int n, ZCMat Xm, zc* out) {
  __shared__ double red[4];
  const int j = blockIdx.x, b = blockIdx.y;
  const zc* col = Xm.p + (long long)b * Xm.stride + (long long)j * Xm.ld;
  double s = 0;
  for (int i = threadIdx.x; i < n; i += 256) s += zabs2(col[i]);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[(long long)b * n + j] = zmake(sqrt(red[0] + red[1] + red[2] + red[3]), 0);
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

namespace {
// n^2 rotations: the 21x21-harmonic layers need about 0.8 n^2 (1.6 sweeps per
// eigenvalue over an average active block of n/2); a matrix that needs more
// stops and resumes after its log is applied.
long long qr_log_default(int n) { return (long long)n * n + 64; }
int qr_log_sweeps(int n) { return 8 * n + 64; }
// The multishift iteration: windows of MS_R steps, MS_NC chains, blocks of
// up to MS_NW rows finished in the deflation window, room for n/3 + 32 AED
// blocks per matrix and launch.
constexpr int MS_R = 16, MS_NC = 8, MS_NW = 32;
// The AED window's rows: the window's QR costs about the same for every
// matrix size while a pass over the active block grows with n, so larger
// matrices take larger windows (more deflation per pass).  Measured, ms per
// QR phase at n = 882 (21x21 harmonics, 510 matrices per launch): windows of
// 32, 24, 20, 16 rows with the nibble at 14%: 1077, 1007, 982, 954; with the
// nibble at 40%: 24, 20, 16, 14, 12, 10, 8 rows: 949, 936, 939, 936, 918,
// 942, 1011; 12 rows with the nibble at 25 and 60%: 925, 960.  At n = 3362
// (41x41 harmonics, 50 matrices) 32 rows took 8.0 s, 12 rows 8.8 s.  12
// rows (LAPACK's zlaqr0 takes 1.5 rows per shift) up to n = 1024, 32 from
// n = 2048, linear between.
int ms_aed_rows(int n) {
  if (n <= 1024) return 12;
  if (n >= 2048) return MS_NW;
  return 12 + (MS_NW - 12) * (n - 1024) / 1024;
}
// After an AED that deflated more than this share of its window (percent)
// the pass is skipped and the next AED runs at once (zlaqr0's nibble point,
// 14% there).
constexpr int MS_NIB = 40;
int qr_block_cap(int n) { return n / 3 + 32; }
}  // namespace

size_t schur_ws_bytes(int n, int batch) {
  const size_t cap = (size_t)qr_log_default(n), scap = (size_t)qr_log_sweeps(n);
  const size_t bcap = (size_t)qr_block_cap(n);
  return (size_t)batch * (cap * (sizeof(double) + sizeof(zc)) + scap * 3 * sizeof(int) +
                          5 * sizeof(int) + bcap * (MS_NW * MS_NW * sizeof(zc) + 3 * sizeof(int)) +
                          sizeof(int)) +
         16 * 256;
}

int schur_batched(hipStream_t s, int n, ZMat T, ZMat Z, int batch, int* info,
                  int maxit_per_row, void* ws, long long log_cap, int phase) {
  if (n <= 0 || batch <= 0) return 0;
  // The default is the multishift iteration with aggressive early deflation
  // (schur_ms_kernel).  RCWA_QR_SWEEP (read on every call) selects another
  // form of the QR iteration, for tests and comparisons: =pipe the pipelined
  // sweep with two chained sweeps and one Wilkinson shift per pass, =single
  // the same with one sweep, =oneshot the logged sweep without the pipelined
  // wave 0 (also used for matrices of 2^31 bytes or more, which the pipelined
  // kernels' 32-bit buffer offsets cannot address).
  const char* mode_env = getenv("RCWA_QR_SWEEP");
  const char mode = mode_env ? mode_env[0] : 0;
  if (!ws) return (int)hipErrorInvalidValue;
  {
    auto align = [](uintptr_t p) { return (p + 255) & ~(uintptr_t)255; };
    const long long cap_full = qr_log_default(n);
    const int scap = qr_log_sweeps(n);
    QrLog log;
    log.c = (double*)ws;
    log.s = (zc*)align((uintptr_t)(log.c + (size_t)batch * cap_full));
    log.il = (int*)align((uintptr_t)(log.s + (size_t)batch * cap_full));
    log.iu = (int*)align((uintptr_t)(log.il + (size_t)batch * scap));
    log.off = (int*)align((uintptr_t)(log.iu + (size_t)batch * scap));
    log.nsw = (int*)align((uintptr_t)(log.off + (size_t)batch * scap));
    log.state = (int*)align((uintptr_t)(log.nsw + batch));
    log.pending = (int*)align((uintptr_t)(log.state + 4 * (size_t)batch));
    log.cap = cap_full;
    log.scap = scap;
    BlockLog blog;
    blog.bcap = qr_block_cap(n);
    blog.V = (zc*)align((uintptr_t)(log.pending + 1));
    blog.pos = (int*)align((uintptr_t)(blog.V + (size_t)batch * blog.bcap * MS_NW * MS_NW));
    blog.k0 = (int*)align((uintptr_t)(blog.pos + (size_t)batch * blog.bcap));
    blog.nw = (int*)align((uintptr_t)(blog.k0 + (size_t)batch * blog.bcap));
    blog.nblk = (int*)align((uintptr_t)(blog.nw + (size_t)batch * blog.bcap));
    // a smaller limit (at least one sweep's worth) only makes the sweep kernel
    // stop for its log to be applied more often
    log.limit = log_cap > 0 ? std::max<long long>(std::min(log_cap, cap_full), n) : cap_full;
    int e = 0;
    if (phase != 2) {
      hipLaunchKernelGGL(qr_state_init_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, n,
                         batch, log.state, log.pending);
      if ((e = (int)hipGetLastError())) return e;
    }
    // Each round: sweeps until done or log full, then the log applied.  A
    // round makes progress on every pending matrix (at least one sweep), and
    // the sweep limit bounds the sweeps, so the loop ends.  Phase 1 queues
    // the first round and returns; phase 2 starts at its read-back.
    const bool huge = (long long)T.ld * n * 16 >= (1LL << 31);
    const bool ms = !huge && mode != 'o' && mode != 's' && mode != 'p';
    for (bool queued = phase == 2;; queued = false) {
      if (queued) {
        // the first round, queued by phase 1
      } else if (mode == 'o' || huge)
        hipLaunchKernelGGL((schur_defer_kernel<12, 256>), dim3(batch), dim3(256), 0, s, n, T, Z,
                           info, maxit_per_row, log);
      else if (mode == 's')
        hipLaunchKernelGGL((schur_pipe_kernel<12, 256, 1>), dim3(batch), dim3(256), 0, s, n, T,
                           Z, info, maxit_per_row, log);
      else if (mode == 'p')
        hipLaunchKernelGGL((schur_pipe_kernel<12, 256, 2>), dim3(batch), dim3(256), 0, s, n, T,
                           Z, info, maxit_per_row, log);
      else
        hipLaunchKernelGGL((schur_ms_kernel<MS_R, 256, MS_NC, MS_NW>), dim3(batch), dim3(256), 0,
                           s, n, T, Z, info, maxit_per_row, log, blog, ms_aed_rows(n), MS_NIB);
      if ((e = (int)hipGetLastError())) return e;
      // the log reaches the Schur vectors only: the sweeps keep the active
      // block of T up to date, and eig_batched forms the whole Schur form
      // afterwards as Z^H A Z
      if (queued)
        ;
      else if (ms)
        hipLaunchKernelGGL((qr_apply_ms_kernel<APPLY_G, APPLY_TC, APPLY_NT, MS_NW>),
                           dim3((n + APPLY_NT - 1) / APPLY_NT, batch), dim3(APPLY_NT), 0, s, n, Z,
                           log, blog);
      else
        hipLaunchKernelGGL((qr_apply_kernel<APPLY_G, APPLY_TC, APPLY_NT>),
                           dim3((n + APPLY_NT - 1) / APPLY_NT, batch), dim3(APPLY_NT), 0, s, n, Z,
                           log);
      if ((e = (int)hipGetLastError())) return e;
      if (phase == 1) return 0;
      int pending = 0;
      if ((e = (int)hipMemcpyAsync(&pending, log.pending, sizeof(int), hipMemcpyDeviceToHost, s)))
        return e;
      if ((e = (int)hipStreamSynchronize(s))) return e;
      if (pending == 0) return 0;
      hipLaunchKernelGGL(qr_clear_pending_kernel, dim3(1), dim3(1), 0, s, log.pending);
      if ((e = (int)hipGetLastError())) return e;
    }
  }
}

int trevc_batched(hipStream_t s, int n, ZCMat T, ZMat X, int batch, bool unit) {
  if (n <= 0 || batch <= 0) return 0;
  constexpr int TB = TREVC_TB;
  long long blocks = ((long long)n * n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL((trevc_init_kernel<TB>), dim3((unsigned)blocks, batch), dim3(256), 0, s, n,
                     T, X);
  if (int e = (int)hipGetLastError()) return e;
  const zc mone = zmake(-1, 0), one = zmake(1, 0);
  for (int I0 = ((n - 1) / TB) * TB; I0 >= 0; I0 -= TB) {
    if (n - I0 - 1 > 0) {
      if (unit)
        hipLaunchKernelGGL((trevc_diag_kernel<TB, true>), dim3((n - I0 - 1 + 255) / 256, batch),
                           dim3(256), 0, s, n, I0, T, X);
      else
        hipLaunchKernelGGL((trevc_diag_kernel<TB, false>), dim3((n - I0 - 1 + 255) / 256, batch),
                           dim3(256), 0, s, n, I0, T, X);
      if (int e = (int)hipGetLastError()) return e;
    }
    const int nbk = std::min(TB, n - I0);
    if (I0 > 0) {
      // X[0:I0, I0:n] -= T[0:I0, I0:I0+nbk] X[I0:I0+nbk, I0:n]
      const ZCMat Ab{T.p + (long long)I0 * T.ld, T.ld, T.stride};
      const ZCMat Bb{X.p + I0 + (long long)I0 * X.ld, X.ld, X.stride};
      const ZMat Cb{X.p + (long long)I0 * X.ld, X.ld, X.stride};
      if (int e = zgemm_batched(s, I0, n - I0, nbk, mone, Ab, Bb, one,
                                ZCMat{Cb.p, Cb.ld, Cb.stride}, Cb, batch))
        return e;
    }
  }
  return 0;
}

size_t eig_workspace_bytes(int n, int batch) {
  return 3 * (size_t)batch * n * n * sizeof(zc) + hess_blocked_ws_bytes(n, batch) +
         schur_ws_bytes(n, batch) + 1024;
}

static int eig_finish(hipStream_t s, int n, ZMat A, zc* w, ZMat W, int batch, zc* Zp, zc* Xp,
                      zc* Ap, void* hws, hipEvent_t* marks, bool inverse);

int eig_batched(hipStream_t s, int n, ZMat A, zc* w, ZMat W, int batch, void* ws, int* info,
                int maxit_per_row, hipEvent_t* marks, long long log_cap, bool inverse,
                int phase) {
  if (n <= 0 || batch <= 0) return 0;
  zc* Zp = (zc*)ws;
  zc* Xp = Zp + (size_t)batch * n * n;
  zc* Ap = Xp + (size_t)batch * n * n;  // a copy of A, later Z^H
  const long long nn = (long long)n * n;
  ZMat Z{Zp, n, nn};
  void* hws = (void*)(((uintptr_t)(Ap + (size_t)batch * n * n) + 255) & ~(uintptr_t)255);
  void* sws = (void*)(((uintptr_t)hws + hess_blocked_ws_bytes(n, batch) + 255) & ~(uintptr_t)255);
  int e;
  if (phase == 2) {
    if ((e = schur_batched(s, n, A, Z, batch, info, maxit_per_row, sws, log_cap, 2))) return e;
    return eig_finish(s, n, A, w, W, batch, Zp, Xp, Ap, hws, marks, inverse);
  }
  if (marks) (void)hipEventRecord(marks[0], s);
  if ((e = zcopy_batched(s, n, n, ZCMat{A.p, A.ld, A.stride}, ZMat{Ap, n, nn}, batch))) return e;
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
  if ((e = schur_batched(s, n, A, Z, batch, info, maxit_per_row, sws, log_cap, phase))) return e;
  if (phase == 1) return 0;
  return eig_finish(s, n, A, w, W, batch, Zp, Xp, Ap, hws, marks, inverse);
}

// After the QR iteration: the Schur form, its eigenvectors and W or W^-1
// (Zp, Xp, Ap and hws: eig_batched's workspace parts).
static int eig_finish(hipStream_t s, int n, ZMat A, zc* w, ZMat W, int batch, zc* Zp, zc* Xp,
                      zc* Ap, void* hws, hipEvent_t* marks, bool inverse) {
  const long long nn = (long long)n * n;
  ZMat Z{Zp, n, nn}, X{Xp, n, nn};
  int e;
  // The Schur form T = Z^H A Z on the matrix cores (the sweeps left the
  // entries of T outside each active block stale): X <- A Z, Ap <- Z^H,
  // A <- Z^H (A Z).
  const zc one = zmake(1, 0), zero = zmake(0, 0);
  if ((e = zgemm_batched(s, n, n, n, one, ZCMat{Ap, n, nn}, ZCMat{Zp, n, nn}, zero,
                         ZCMat{nullptr, 0, 0}, X, batch)))
    return e;
  // (only the tiles meeting T's upper triangle: nothing reads T below it)
  if ((e = zconj_transpose_batched(s, n, n, ZCMat{Zp, n, nn}, ZMat{Ap, n, nn}, batch))) return e;
  if ((e = zgemm_batched_tri(s, n, n, n, one, ZCMat{Ap, n, nn}, ZCMat{Xp, n, nn}, zero,
                             ZCMat{nullptr, 0, 0}, A, batch, ZGEMM_TRI_C)))
    return e;
  if (marks) (void)hipEventRecord(marks[2], s);
  if ((e = trevc_batched(s, n, ZCMat{A.p, A.ld, A.stride}, X, batch, false))) return e;
  hipLaunchKernelGGL(diag_kernel, dim3((n + 255) / 256, batch), dim3(256), 0, s, n,
                     ZCMat{A.p, A.ld, A.stride}, w);
  if (inverse) {
    // W^-1 of the unit-norm eigenvectors W = Z X D^-1 (D = diag |X e_k|, Z
    // unitary): W^-1 = D X^-1 Z^H.  X is unit upper triangular, so X^-1 comes
    // from the same blocked substitution (into Z's buffer, free now), and Ap
    // still holds Z^H.
    zc* nrm = (zc*)hws;  // n per matrix (the Hessenberg workspace is free)
    hipLaunchKernelGGL(col_norms_kernel, dim3(n, batch), dim3(256), 0, s, n,
                       ZCMat{Xp, n, nn}, nrm);
    if ((e = (int)hipGetLastError())) return e;
    if ((e = trevc_batched(s, n, ZCMat{Xp, n, nn}, Z, batch, true))) return e;
    if ((e = zgemm_batched_tri(s, n, n, n, one, ZCMat{Zp, n, nn}, ZCMat{Ap, n, nn}, zero,
                               ZCMat{nullptr, 0, 0}, W, batch, ZGEMM_TRI_A)))
      return e;
    if ((e = zscale_rc(s, n, n, W, nrm, n, nullptr, 0, batch))) return e;
    if (marks) (void)hipEventRecord(marks[3], s);
    return 0;
  }
  // W = Z X, X upper triangular (the eigenvectors of T)
  if ((e = zgemm_batched_tri(s, n, n, n, zmake(1, 0), ZCMat{Zp, n, (long long)n * n},
                             ZCMat{Xp, n, (long long)n * n}, zmake(0, 0), ZCMat{nullptr, 0, 0}, W,
                             batch, ZGEMM_TRI_B)))
    return e;
  if ((e = znormalize_cols(s, n, W, batch))) return e;
  if (marks) (void)hipEventRecord(marks[3], s);
  return 0;
}
