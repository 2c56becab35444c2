// Split-precision complex GEMM on the gfx950 matrix cores, the product used
// by the split form of the path (PQ, W/V and the S-matrix / Redheffer
// multiplies); the eigensolver and every inverse stay FP64.
//
//   C = alpha * A * B + beta * D     (complex128 in and out, column-major)
//
// Ozaki-style splitting: every row of A and every column of B is scaled by a
// power of two so that its largest real or imaginary part lies in [0.5, 1),
// then each real number x is split into two FP16 slices
//   x = x1 + 2^-11 x2 + O(2^-22),   x1 = fp16(x),  x2 = fp16((x - x1) 2^11).
// A real product keeps the three significant slice products
//   X Y ~= X1 Y1 + 2^-11 (X1 Y2 + X2 Y1)
// on v_mfma_f32_16x16x16_f16 with FP32 accumulation (separate "hi" and "lo"
// accumulators), and a complex product is four real ones, 12 MFMAs per
// 16x16x16 step.  The FP32 accumulators cover SLAB = 64 consecutive k; each
// slab's hi + 2^-11 lo is then added into FP64 accumulators, which the
// epilogue unscales and combines with alpha/beta in FP64.  The normwise error
// is about 2^-22 from the split plus sqrt(SLAB) 2^-24 from accumulation.
//
// Operand layout after the split pass (per batch entry and slice): groups of
// four consecutive k for one row (A) or one column (B), i.e.
// S[k/4][row][k%4], so that each lane's four-element MFMA operand is one
// 8-byte load and 16 lanes read 128 contiguous bytes.
#include "common.hpp"
#include "runtime.hpp"

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int TILE = 64;
// FP32 accumulation runs over SLAB consecutive k; slab sums are added in FP64.
constexpr int SLAB = 64;

inline int rup(int x, int m) { return (x + m - 1) / m * m; }

// rows = number of vectors to split (rows of A or columns of B), K = their
// length.  Element (v, k) of batch b is at src + b*stride + v*vs + k*ks.
// Output: 4 slices (re1, re2, im1, im2) of Vp x Kp halfs, and the power-of-two
// scale 2^e of every vector.
__global__ void split_kernel(int rows, int K, const zc* src, long long stride, long long vs,
                             long long ks, int Vp, int Kp, _Float16* out, double* scale) {
  const int b = blockIdx.y;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= Vp) return;
  const zc* p = src + (long long)b * stride + (long long)v * vs;
  const size_t slice = (size_t)Vp * Kp;
  _Float16* o = out + (size_t)b * 4 * slice;
  double m = 0.0;
  if (v < rows)
    for (int k = 0; k < K; ++k) {
      const zc z = p[k * ks];
      m = fmax(m, fmax(fabs(z.x), fabs(z.y)));
    }
  int e = 0;
  if (m > 0.0) frexp(m, &e);
  const double s = ldexp(1.0, -e);
  scale[(size_t)b * Vp + v] = ldexp(1.0, e);
  for (int q = 0; q < Kp / 4; ++q) {
    h4 r1, r2, i1, i2;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = 4 * q + t;
      zc z = zmake(0, 0);
      if (v < rows && k < K) z = p[k * ks];
      const double xr = z.x * s, xi = z.y * s;
      const _Float16 hr = (_Float16)xr, hi = (_Float16)xi;
      r1[t] = hr;
      i1[t] = hi;
      r2[t] = (_Float16)((xr - (double)hr) * 2048.0);
      i2[t] = (_Float16)((xi - (double)hi) * 2048.0);
    }
    const size_t at = ((size_t)q * Vp + v) * 4;
    *(h4*)(o + at) = r1;
    *(h4*)(o + slice + at) = r2;
    *(h4*)(o + 2 * slice + at) = i1;
    *(h4*)(o + 3 * slice + at) = i2;
  }
}

__device__ __forceinline__ f4 mf(h4 a, h4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

__global__ __launch_bounds__(256) void split_mm_kernel(int M, int N, int Kp, int Mp, int Np,
                                                       const _Float16* As, const double* sA,
                                                       const _Float16* Bs, const double* sB,
                                                       zc alpha, zc beta, ZCMat D, ZMat C,
                                                       int tiles_n) {
  const int b = blockIdx.z;
  const int tn = blockIdx.x % tiles_n, tm = blockIdx.x / tiles_n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lk = lane >> 4;
  const size_t sa = (size_t)Mp * Kp, sb = (size_t)Np * Kp;
  const _Float16* A0 = As + (size_t)b * 4 * sa;
  const _Float16* B0 = Bs + (size_t)b * 4 * sb;
  const int r0 = tm * TILE + wm * 32, c0 = tn * TILE + wn * 32;
  f4 hr[2][2], lr[2][2], hi_[2][2], lo_[2][2];
  double accR[2][2][4], accI[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      hr[i][j] = lr[i][j] = hi_[i][j] = lo_[i][j] = f4{0, 0, 0, 0};
#pragma unroll
      for (int r = 0; r < 4; ++r) accR[i][j][r] = accI[i][j][r] = 0.0;
    }
  const double lo_s = 1.0 / 2048.0;
  for (int k0 = 0; k0 < Kp; k0 += 16) {
    const size_t q = (size_t)(k0 / 4 + lk);
    h4 a[2][4], bb[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const size_t at = (q * Mp + r0 + i * 16 + li) * 4;
#pragma unroll
      for (int s = 0; s < 4; ++s) a[i][s] = *(const h4*)(A0 + s * sa + at);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const size_t at = (q * Np + c0 + j * 16 + li) * 4;
#pragma unroll
      for (int s = 0; s < 4; ++s) bb[j][s] = *(const h4*)(B0 + s * sb + at);
    }
    // 12 products per tile, ordered so that consecutive MFMAs update
    // different accumulators (16 independent chains)
    h4 nai[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      nai[i][0] = -a[i][2];
      nai[i][1] = -a[i][3];
    }
#define SPLIT_STEP(ACC, X, Y)                 \
  _Pragma("unroll") for (int i = 0; i < 2; ++i) \
  _Pragma("unroll") for (int j = 0; j < 2; ++j) ACC[i][j] = mf(X, Y, ACC[i][j]);
    SPLIT_STEP(hr, a[i][0], bb[j][0])
    SPLIT_STEP(hi_, a[i][0], bb[j][2])
    SPLIT_STEP(lr, a[i][0], bb[j][1])
    SPLIT_STEP(lo_, a[i][0], bb[j][3])
    SPLIT_STEP(hr, nai[i][0], bb[j][2])
    SPLIT_STEP(hi_, a[i][2], bb[j][0])
    SPLIT_STEP(lr, a[i][1], bb[j][0])
    SPLIT_STEP(lo_, a[i][1], bb[j][2])
    SPLIT_STEP(lr, nai[i][0], bb[j][3])
    SPLIT_STEP(lo_, a[i][2], bb[j][1])
    SPLIT_STEP(lr, nai[i][1], bb[j][2])
    SPLIT_STEP(lo_, a[i][3], bb[j][0])
#undef SPLIT_STEP
    // promote the FP32 partial sums to FP64 every SLAB k-values
    if (((k0 + 16) % SLAB) == 0 || k0 + 16 >= Kp) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            accR[i][j][r] += (double)hr[i][j][r] + lo_s * (double)lr[i][j][r];
            accI[i][j][r] += (double)hi_[i][j][r] + lo_s * (double)lo_[i][j][r];
          }
          hr[i][j] = lr[i][j] = hi_[i][j] = lo_[i][j] = f4{0, 0, 0, 0};
        }
    }
  }
  zc* Cp = C.p + (long long)b * C.stride;
  const zc* Dp = D.p ? D.p + (long long)b * D.stride : nullptr;
  const bool use_d = (beta.x != 0.0 || beta.y != 0.0) && Dp != nullptr;
  const double* sAb = sA + (size_t)b * Mp;
  const double* sBb = sB + (size_t)b * Np;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = c0 + j * 16 + li;
      if (col >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = r0 + i * 16 + 4 * lk + r;
        if (row >= M) continue;
        const double sc = sAb[row] * sBb[col];
        const zc prod = zmake(accR[i][j][r] * sc, accI[i][j][r] * sc);
        zc v = zmul(alpha, prod);
        if (use_d) v = zfma(beta, Dp[row + (long long)col * D.ld], v);
        Cp[row + (long long)col * C.ld] = v;
      }
    }
}

}  // namespace

size_t zgemm_split_ws_bytes(int M, int N, int K, int batch) {
  const size_t Mp = rup(M, TILE), Np = rup(N, TILE), Kp = rup(K, 16);
  return (size_t)batch * (4 * (Mp + Np) * Kp * sizeof(_Float16) + (Mp + Np) * sizeof(double)) +
         1024;
}

int zgemm_split_batched_ws(hipStream_t s, int M, int N, int K, zc alpha, ZCMat A, ZCMat B,
                           zc beta, ZCMat D, ZMat C, int batch, void* ws) {
  if (M <= 0 || N <= 0 || batch <= 0) return 0;
  const int Mp = rup(M, TILE), Np = rup(N, TILE), Kp = rup(K > 0 ? K : 1, 16);
  _Float16* As = (_Float16*)ws;
  _Float16* Bs = As + (size_t)batch * 4 * Mp * Kp;
  double* sA = (double*)(((uintptr_t)(Bs + (size_t)batch * 4 * Np * Kp) + 255) & ~(uintptr_t)255);
  double* sB = sA + (size_t)batch * Mp;
  hipLaunchKernelGGL(split_kernel, dim3((Mp + 255) / 256, batch), dim3(256), 0, s, M, K, A.p,
                     A.stride, 1LL, A.ld, Mp, Kp, As, sA);
  hipLaunchKernelGGL(split_kernel, dim3((Np + 255) / 256, batch), dim3(256), 0, s, N, K, B.p,
                     B.stride, B.ld, 1LL, Np, Kp, Bs, sB);
  const int tiles_n = Np / TILE, tiles_m = Mp / TILE;
  hipLaunchKernelGGL(split_mm_kernel, dim3(tiles_m * tiles_n, 1, batch), dim3(256), 0, s, M, N,
                     Kp, Mp, Np, As, sA, Bs, sB, alpha, beta, D, C, tiles_n);
  return (int)hipGetLastError();
}
