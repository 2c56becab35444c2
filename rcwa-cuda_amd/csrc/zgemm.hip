// Batched complex128 GEMM on the gfx950 FP64 matrix cores.
//
//   C[b] = alpha * A[b] * B[b] + beta * D[b]      (column-major)
//
// Every dense product of the path goes through here: Omega^2 = P Q
// (rcwa.cpp:288), V = Q W (rcwa.cpp:326), the layer S-matrix products
// (rcwa.cpp:345-348), the Redheffer products (rcwa.cpp:602-611), the
// eigenvector back-transformation and the rank-nb updates of the blocked
// Gauss-Jordan inverse.
//
// A complex product is three real FP64 MFMAs per 16x16x4 step (the 3M form
// of the complex product, Higham, "Stability of a method for multiplying
// complex matrices with three real matrix multiplications", SIAM J. Matrix
// Anal. Appl. 13 (1992)): P1 += Ar Br, P2 += Ai Bi, P3 += (Ar + Ai)(Br + Bi),
// then Cr = P1 - P2 and Ci = P3 - (P1 + P2).  Each lane loads its complex
// element once and forms the sum of its parts in registers.  The method is
// normwise stable: the error of Ci is bounded by a small multiple of
// u k max|A| max|B| like that of the four-product form; only the componentwise
// bound of a tiny imaginary part is lost, which no consumer on the path uses.
//
// Lane layout of v_mfma_f64_16x16x4f64 (measured on gfx950 with an
// in-bounds probe): lane l supplies A[l%16][l/16] and B[l/16][l%16]; result
// register r of lane l holds D[4r + l/16][l%16].
//
// Tiling: WM x WN wave64s, each owning 32x32 of C = 2x2 MFMA tiles of 16x16;
// the workgroup tile is 32 WM x 32 WN.  K advances in slabs of 16 staged
// through LDS (double-buffered: the next slab's global loads are issued
// before the MFMAs on the current one).  The tile shapes below differ only in
// how C is cut into workgroups: every element is summed over the same slabs
// in the same order, so all shapes give the same bits.  The shape of a
// product comes from the autotuner (zgemm_tune): the fastest on scratch
// operands of the product's shape class, measured on first use where tuning
// is allowed, else 2x2 (64 x 64, 256 threads).
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int BK = 16;
constexpr unsigned XCDS = 8;  // accelerator complex dies of the MI355X, each with its own L2
constexpr int PADA = 1, PADB = 1;

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Shapes two or more waves wide along N run two workgroups per CU (their LDS fits
// twice): registers capped at 256 per lane, which the three accumulator sets
// still fit without spilling; the one-wave-wide shapes hold more slab entries
// per lane and keep one wave per SIMD.
// tri: ZGEMM_TRI_C (only the tiles of C that meet its upper triangle are
// formed; the others are left as they are), ZGEMM_TRI_B (B is upper
// triangular: the tile of columns n0..n0+BN-1 needs rows k < n0 + BN only)
// and ZGEMM_TRI_A (A is upper triangular: the tile of rows m0.. needs
// columns k >= m0 only).
template <int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(WN >= 2 ? 2 : 1))) void zgemm_kernel(
    int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta, ZCMat D, ZMat C, int tiles_n,
    int tri, const int* sel) {
  constexpr int BM = 32 * WM, BN = 32 * WN, NT = 64 * WM * WN;
  constexpr int LA = BM * BK / NT, LB = BK * BN / NT;  // slab entries per thread
  __shared__ zc As[2][BK][BM + PADA];
  __shared__ zc Bs[2][BK][BN + PADB];
  // One-dimensional grid of tiles x batch.  Workgroups are dispatched to the
  // XCDs round-robin by their index; the index is remapped so that each XCD
  // receives a contiguous range of tiles (the tiles of one matrix together),
  // whose A rows and B columns then meet in that XCD's L2.
  const unsigned total = gridDim.x, L = blockIdx.x;
  const unsigned per = total / XCDS, rem = total % XCDS, xcd = L % XCDS, idx = L / XCDS;
  const unsigned v = xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
  const int tiles = ((M + BM - 1) / BM) * tiles_n;  // tiles per matrix
  const int b = (int)(v / (unsigned)tiles);
  const int t = (int)(v % (unsigned)tiles);
  const int tn = t % tiles_n;
  const int tm = t / tiles_n;
  const int m0 = tm * BM, n0 = tn * BN;
  if (sel && sel[b] == 0) return;  // a product not selected: C keeps its contents
  if ((tri & ZGEMM_TRI_C) && m0 >= n0 + BN) return;  // uniform over the workgroup
  if (tri & ZGEMM_TRI_B) K = min(K, n0 + BN);
  const int kbeg = (tri & ZGEMM_TRI_A) ? min(m0, K) : 0;  // a multiple of BK
  const zc* Ap = A.p + (long long)b * A.stride;
  const zc* Bp = B.p + (long long)b * B.stride;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int li = lane & 15, lk = lane >> 4;

  v4d acc1[2][2], acc2[2][2], acc3[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc1[i][j] = v4d{0, 0, 0, 0};
      acc2[i][j] = v4d{0, 0, 0, 0};
      acc3[i][j] = v4d{0, 0, 0, 0};
    }

  // The addend D is loaded before the K loop, so that its loads are in flight
  // under the products: for the rank-16 to rank-64 updates (Gauss-Jordan,
  // Hessenberg, eigenvector blocks) reading and writing C is most of the work.
  const zc* Dp = D.p ? D.p + (long long)b * D.stride : nullptr;
  const bool use_d = (beta.x != 0.0 || beta.y != 0.0) && Dp != nullptr;
  zc dv[2][2][4];
  if (use_d) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int col = n0 + wn * 32 + j * 16 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm * 32 + i * 16 + 4 * r + lk;
          dv[i][j][r] = (row < M && col < N) ? Dp[row + (long long)col * D.ld] : zmake(0, 0);
        }
      }
  }

  zc ra[LA], rb[LB];
  auto gload = [&](int k0) {
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      const int idx = tid + NT * r;
      const int i = idx % BM, k = idx / BM;  // A: rows fastest
      const int gi = m0 + i, gk = k0 + k;
      ra[r] = (gi < M && gk < K) ? Ap[gi + (long long)gk * A.ld] : zmake(0, 0);
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      const int idx = tid + NT * r;
      const int kb = idx % BK, j = idx / BK;  // B: k fastest
      const int gkb = k0 + kb, gj = n0 + j;
      rb[r] = (gkb < K && gj < N) ? Bp[gkb + (long long)gj * B.ld] : zmake(0, 0);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      const int idx = tid + NT * r;
      As[buf][idx / BM][idx % BM] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      const int idx = tid + NT * r;
      Bs[buf][idx % BK][idx / BK] = rb[r];
    }
  };

  const int nk = (K - kbeg + BK - 1) / BK;
  gload(kbeg);
  sstore(0);
  __syncthreads();
  for (int t = 0; t < nk; ++t) {
    const int cur = t & 1;
    if (t + 1 < nk) gload(kbeg + (t + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      zc a[2], bb[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[cur][kk + lk][wm * 32 + i * 16 + li];
#pragma unroll
      for (int j = 0; j < 2; ++j) bb[j] = Bs[cur][kk + lk][wn * 32 + j * 16 + li];
      double as[2], bs[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) as[i] = a[i].x + a[i].y;
#pragma unroll
      for (int j = 0; j < 2; ++j) bs[j] = bb[j].x + bb[j].y;
      // twelve independent accumulators between two updates of the same one
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc1[i][j] = mfma(a[i].x, bb[j].x, acc1[i][j]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc2[i][j] = mfma(a[i].y, bb[j].y, acc2[i][j]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc3[i][j] = mfma(as[i], bs[j], acc3[i][j]);
    }
    if (t + 1 < nk) {
      sstore(cur ^ 1);
    }
    __syncthreads();
  }

  zc* Cp = C.p + (long long)b * C.stride;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn * 32 + j * 16 + li;
      if (col >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 32 + i * 16 + 4 * r + lk;  // D[4r + l/16][l%16]
        if (row >= M) continue;
        const double p1 = acc1[i][j][r], p2 = acc2[i][j][r];
        zc v = zmul(alpha, zmake(p1 - p2, acc3[i][j][r] - (p1 + p2)));
        if (use_d) v = zfma(beta, dv[i][j][r], v);
        Cp[row + (long long)col * C.ld] = v;
      }
    }
}

}  // namespace

namespace {

// The tile shapes (waves along M x waves along N).  0 is the default.
struct Shape {
  int wm, wn;
};
constexpr Shape kShapes[] = {{2, 2}, {1, 1}, {4, 1}, {2, 4}, {4, 2}};
constexpr int kNumShapes = (int)(sizeof(kShapes) / sizeof(kShapes[0]));

int launch_shape(int v, hipStream_t s, int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta,
                 ZCMat D, ZMat C, int batch, int tri, const int* sel = nullptr) {
  const int bm = 32 * kShapes[v].wm, bn = 32 * kShapes[v].wn;
  const int tiles_m = (M + bm - 1) / bm, tiles_n = (N + bn - 1) / bn;
  const long long total = (long long)tiles_m * tiles_n * batch;
  if (total > 0x7fffffffLL) return (int)hipErrorInvalidValue;
  const dim3 grid((unsigned)total), block(64 * kShapes[v].wm * kShapes[v].wn);
  switch (v) {
    case 1:
      hipLaunchKernelGGL((zgemm_kernel<1, 1>), grid, block, 0, s, M, N, K, alpha, A, B, beta, D, C,
                         tiles_n, tri, sel);
      break;
    case 2:
      hipLaunchKernelGGL((zgemm_kernel<4, 1>), grid, block, 0, s, M, N, K, alpha, A, B, beta, D, C,
                         tiles_n, tri, sel);
      break;
    case 3:
      hipLaunchKernelGGL((zgemm_kernel<2, 4>), grid, block, 0, s, M, N, K, alpha, A, B, beta, D, C,
                         tiles_n, tri, sel);
      break;
    case 4:
      hipLaunchKernelGGL((zgemm_kernel<4, 2>), grid, block, 0, s, M, N, K, alpha, A, B, beta, D, C,
                         tiles_n, tri, sel);
      break;
    default:
      hipLaunchKernelGGL((zgemm_kernel<2, 2>), grid, block, 0, s, M, N, K, alpha, A, B, beta, D, C,
                         tiles_n, tri, sel);
  }
  return (int)hipGetLastError();
}

// Shape class of a product: M, N, K exact, the batch by its power of two.
struct Key {
  int M, N, K, bclass, dev;
  bool operator<(const Key& o) const {
    return std::tie(M, N, K, bclass, dev) < std::tie(o.M, o.N, o.K, o.bclass, o.dev);
  }
};

int batch_class(int batch) {
  int c = 0;
  while ((1 << c) < batch) ++c;
  return c;
}

struct Tuner {
  std::mutex mu;
  std::map<Key, int> table;
  bool loaded = false;
  std::atomic<int> forced{-1};
  // RCWA_TUNE_FILE: the table is read from it on first use and every tuned
  // entry is appended ("M N K bclass device shape" per line).
  const char* file() const { return getenv("RCWA_TUNE_FILE"); }
  void load_locked() {
    if (loaded) return;
    loaded = true;
    const char* f = file();
    if (!f) return;
    if (FILE* fp = fopen(f, "r")) {
      Key k;
      int v;
      while (fscanf(fp, "%d %d %d %d %d %d", &k.M, &k.N, &k.K, &k.bclass, &k.dev, &v) == 6)
        if (v >= 0 && v < kNumShapes) table[k] = v;
      fclose(fp);
    }
  }
  void save_locked(const Key& k, int v) {
    const char* f = file();
    if (!f) return;
    if (FILE* fp = fopen(f, "a")) {
      fprintf(fp, "%d %d %d %d %d %d\n", k.M, k.N, k.K, k.bclass, k.dev, v);
      fclose(fp);
    }
  }
};

Tuner& tuner() {
  static Tuner t;
  return t;
}

thread_local bool tl_tune = false;

// Times every shape on scratch operands of the key's shape (a batch of up to
// 2^bclass products, capped at 1 GiB of operands) on stream s; returns the
// fastest, or -1 when the scratch cannot be had or a launch fails.
int tune_key(hipStream_t s, const Key& k) {
  const size_t per = ((size_t)k.M * k.K + (size_t)k.K * k.N + (size_t)k.M * k.N) * sizeof(zc);
  int tb = 1 << k.bclass;
  while (tb > 1 && (size_t)tb * per > (size_t)1 << 30) tb >>= 1;
  zc* buf = nullptr;
  if (hipMalloc(&buf, (size_t)tb * per) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  int best = -1;
  if (hipMemsetAsync(buf, 0, (size_t)tb * per, s) == hipSuccess) {
    const long long sa = (long long)k.M * k.K, sb = (long long)k.K * k.N, sc = (long long)k.M * k.N;
    zc* Ap = buf;
    zc* Bp = Ap + (size_t)tb * sa;
    zc* Cp = Bp + (size_t)tb * sb;
    const ZCMat A{Ap, k.M, sa}, B{Bp, k.K, sb};
    const ZMat C{Cp, k.M, sc};
    const ZCMat D{Cp, k.M, sc};
    const zc one = zmake(1, 0);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best_ms = 0;
    for (int v = 0; v < kNumShapes; ++v) {
      if (launch_shape(v, s, k.M, k.N, k.K, one, A, B, one, D, C, tb, 0)) break;  // warm-up
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < 3; ++r) (void)launch_shape(v, s, k.M, k.N, k.K, one, A, B, one, D, C, tb, 0);
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) break;
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (best < 0 || ms < best_ms) {
        best = v;
        best_ms = ms;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(buf);
  (void)hipGetLastError();
  return best;
}

// The shape for a product: forced (tests), tuned, or tuned now when this
// thread allows tuning (never during stream capture), else the default.
int shape_for(hipStream_t s, int M, int N, int K, int batch, int tri) {
  Tuner& t = tuner();
  const int f = t.forced.load();
  if (f >= 0) return f;
  if (tri) return 0;  // the structured forms skip tiles by the default tile
  int dev = 0;
  (void)hipGetDevice(&dev);
  const Key k{M, N, K, batch_class(batch), dev};
  {
    std::lock_guard<std::mutex> g(t.mu);
    t.load_locked();
    auto it = t.table.find(k);
    if (it != t.table.end()) return it->second;
  }
  if (!tl_tune) return 0;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
    (void)hipGetLastError();
    return 0;
  }
  const int v = tune_key(s, k);
  if (v < 0) return 0;
  std::lock_guard<std::mutex> g(t.mu);
  t.table[k] = v;
  t.save_locked(k, v);
  return v;
}

}  // namespace

ZgemmTuneScope::ZgemmTuneScope(bool on) : prev(tl_tune) { tl_tune = on; }
ZgemmTuneScope::~ZgemmTuneScope() { tl_tune = prev; }

int zgemm_shape_count() { return kNumShapes; }

void zgemm_force_shape(int v) { tuner().forced.store(v >= 0 && v < kNumShapes ? v : -1); }

int zgemm_tuned_entries(int* out, int max_entries) {
  Tuner& t = tuner();
  std::lock_guard<std::mutex> g(t.mu);
  t.load_locked();
  int i = 0;
  for (const auto& e : t.table) {
    if (out && i < max_entries) {
      int* o = out + 6 * i;
      o[0] = e.first.M;
      o[1] = e.first.N;
      o[2] = e.first.K;
      o[3] = e.first.bclass;
      o[4] = e.first.dev;
      o[5] = e.second;
    }
    ++i;
  }
  return i;
}

int zgemm_batched_tri(hipStream_t s, int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta,
                      ZCMat D, ZMat C, int batch, int tri) {
  if (M <= 0 || N <= 0 || batch <= 0) return 0;
  return launch_shape(shape_for(s, M, N, K, batch, tri), s, M, N, K, alpha, A, B, beta, D, C,
                      batch, tri);
}

int zgemm_batched(hipStream_t s, int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta,
                  ZCMat D, ZMat C, int batch) {
  return zgemm_batched_tri(s, M, N, K, alpha, A, B, beta, D, C, batch, 0);
}

int zgemm_batched_sel(hipStream_t s, int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta,
                      ZCMat D, ZMat C, int batch, const int* sel) {
  if (M <= 0 || N <= 0 || batch <= 0) return 0;
  return launch_shape(shape_for(s, M, N, K, batch, 0), s, M, N, K, alpha, A, B, beta, D, C, batch,
                      0, sel);
}
