// Split-precision complex GEMM on the gfx950 FP16 matrix cores, the product
// used by the split form of the path (the layer S-matrix and Redheffer
// multiplies); Omega^2 = P Q, V = Q W, the eigensolver and every inverse stay
// FP64.
//
//   C = alpha * A * B + beta * D     (complex128 in and out, column-major)
//
// Ozaki scheme with fixed-point slices.  Every row of A and every column of B
// is scaled by a power of two s so that its real and imaginary parts lie in
// (-1, 1); each scaled real number u is then cut into four 9-bit integer
// slices
//   u = 2^-9 (a1 + 2^-9 a2 + 2^-18 a3 + 2^-27 a4) + r,
//   |a1| <= 512, |a2|, |a3|, |a4| <= 256, |r| <= 2^-37,
// all exact in FP16.  A real product keeps the slice products of order
// 2^-36 and above:
//   u v ~= 2^-18 [ a1 b1 + 2^-9 (a1 b2 + a2 b1) + 2^-18 (a1 b3 + a2 b2 + a3 b1)
//                  + 2^-27 (a1 b4 + a2 b3 + a3 b2 + a4 b1) ].
// The first two groups are sums of integers below 2^18 and 2^17; within a
// slab of 32 consecutive k the complex combination (two real products)
// keeps each below 2^24, so v_mfma_f32_16x16x16_f16 accumulates both
// EXACTLY in FP32, each in its own accumulator.  The last two groups share a
// third FP32 accumulator (the 2^-27 group fed with slices pre-scaled by
// 2^-9); its rounding is 2^-24 of a term 2^-18 down, i.e. 2^-42 of the
// result.  After every slab the accumulators are added into FP64
// accumulators; the epilogue unscales and combines with alpha/beta in FP64.
// Normwise error: about 2^-36 (the dropped slice products and the slicing
// remainder).  Three slices (2^-27) were not enough: the layer S-matrix
// products amplify their error about a hundredfold at 21x21 harmonics.
//
// A complex 16x16x16 step takes 40 MFMAs: 4 for the leading group, 8 for the
// second and 28 for the last two.
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
// The exact FP32 accumulations run over SLAB consecutive k.
constexpr int SLAB = 32;
// Slices per real part: a1, a2, a3, 2^-9 a3, 2^-9 a4 (the same for B).
constexpr int NS = 5;

inline int rup(int x, int m) { return (x + m - 1) / m * m; }

// rows = number of vectors to split (rows of A or columns of B), K = their
// length.  Element (v, k) of batch b is at src + b*stride + v*vs + k*ks.
// Output: 2*ns slices (real part slices, then imaginary part slices) of
// Vp x Kp halfs, and the power-of-two scale 2^e of every vector.
__global__ void split_kernel(int rows, int K, const zc* src, long long stride, long long vs,
                             long long ks, int Vp, int Kp, _Float16* out, double* scale) {
  const int b = blockIdx.y;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= Vp) return;
  const zc* p = src + (long long)b * stride + (long long)v * vs;
  const size_t slice = (size_t)Vp * Kp;
  _Float16* o = out + (size_t)b * 2 * NS * slice;
  double m = 0.0;
  if (v < rows)
    for (int k = 0; k < K; ++k) {
      const zc z = p[k * ks];
      m = fmax(m, fmax(fabs(z.x), fabs(z.y)));
    }
  // 2^e > m, so the scaled parts lie in (-1, 1)
  int e = 0;
  if (m > 0.0) frexp(m, &e);
  scale[(size_t)b * Vp + v] = ldexp(1.0, e);
  const double s = ldexp(1.0, -e);
  for (int q = 0; q < Kp / 4; ++q) {
    h4 out_s[2][NS];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = 4 * q + t;
      zc z = zmake(0, 0);
      if (v < rows && k < K) z = p[k * ks];
      const double part[2] = {z.x * s, z.y * s};
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        // exact in FP64: scaling by powers of two and subtracting the rounded
        // leading bits
        const double t1 = part[c] * 512.0;
        const double a1 = rint(t1);
        const double t2 = (t1 - a1) * 512.0;
        const double a2 = rint(t2);
        const double t3 = (t2 - a2) * 512.0;
        const double a3 = rint(t3);
        const double a4 = rint((t3 - a3) * 512.0);
        out_s[c][0][t] = (_Float16)a1;
        out_s[c][1][t] = (_Float16)a2;
        out_s[c][2][t] = (_Float16)a3;
        out_s[c][3][t] = (_Float16)(a3 * (1.0 / 512.0));
        out_s[c][4][t] = (_Float16)(a4 * (1.0 / 512.0));
      }
    }
    const size_t at = ((size_t)q * Vp + v) * 4;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int j = 0; j < NS; ++j) *(h4*)(o + (size_t)(c * NS + j) * slice + at) = out_s[c][j];
  }
}

__device__ __forceinline__ f4 mf(h4 a, h4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// Slices: index c*NS + j (c = 0 real, 1 imaginary; j = a1, a2, a3, a3/512,
// a4/512), for A and for B.
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
  const _Float16* A0 = As + (size_t)b * 2 * NS * sa;
  const _Float16* B0 = Bs + (size_t)b * 2 * NS * sb;
  const int r0 = tm * TILE + wm * 32, c0 = tn * TILE + wn * 32;
  // g0: leading group, g1: the 2^-9 group (both exact), g2: the 2^-18 and
  // 2^-27 groups; r/i: real/imaginary part
  f4 g0r[2][2], g0i[2][2], g1r[2][2], g1i[2][2], g2r[2][2], g2i[2][2];
  double accR[2][2][4], accI[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      g0r[i][j] = g0i[i][j] = g1r[i][j] = g1i[i][j] = g2r[i][j] = g2i[i][j] = f4{0, 0, 0, 0};
#pragma unroll
      for (int r = 0; r < 4; ++r) accR[i][j][r] = accI[i][j][r] = 0.0;
    }
  const double s1 = 1.0 / 512.0, s2 = 1.0 / (512.0 * 512.0);
  for (int k0 = 0; k0 < Kp; k0 += 16) {
    const size_t q = (size_t)(k0 / 4 + lk);
    // ar[i][j]: real-part slice j of row tile i; ai: imaginary; nai: -ai
    h4 ar[2][NS], ai[2][NS], nai[2][NS], br[2][NS], bi[2][NS];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const size_t at = (q * Mp + r0 + i * 16 + li) * 4;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        ar[i][s] = *(const h4*)(A0 + s * sa + at);
        ai[i][s] = *(const h4*)(A0 + (NS + s) * sa + at);
        nai[i][s] = -ai[i][s];
      }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const size_t at = (q * Np + c0 + j * 16 + li) * 4;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        br[j][s] = *(const h4*)(B0 + s * sb + at);
        bi[j][s] = *(const h4*)(B0 + (NS + s) * sb + at);
      }
    }
    // one complex slice product x * y into accumulators (re, im)
#define SPLIT_CMUL(RE, IM, XA, YB)                          \
  RE[i][j] = mf(ar[i][XA], br[j][YB], RE[i][j]);            \
  IM[i][j] = mf(ar[i][XA], bi[j][YB], IM[i][j]);            \
  RE[i][j] = mf(nai[i][XA], bi[j][YB], RE[i][j]);           \
  IM[i][j] = mf(ai[i][XA], br[j][YB], IM[i][j]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        SPLIT_CMUL(g0r, g0i, 0, 0)   // a1 b1
        SPLIT_CMUL(g1r, g1i, 0, 1)   // a1 b2
        SPLIT_CMUL(g1r, g1i, 1, 0)   // a2 b1
        SPLIT_CMUL(g2r, g2i, 0, 2)   // a1 b3
        SPLIT_CMUL(g2r, g2i, 1, 1)   // a2 b2
        SPLIT_CMUL(g2r, g2i, 2, 0)   // a3 b1
        SPLIT_CMUL(g2r, g2i, 0, 4)   // a1 (b4/512)
        SPLIT_CMUL(g2r, g2i, 1, 3)   // a2 (b3/512)
        SPLIT_CMUL(g2r, g2i, 3, 1)   // (a3/512) b2
        SPLIT_CMUL(g2r, g2i, 4, 0)   // (a4/512) b1
      }
#undef SPLIT_CMUL
    // promote to FP64 every SLAB k-values (keeps the first two groups exact)
    if (((k0 + 16) % SLAB) == 0 || k0 + 16 >= Kp) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            accR[i][j][r] += (double)g0r[i][j][r] +
                             (s1 * (double)g1r[i][j][r] + s2 * (double)g2r[i][j][r]);
            accI[i][j][r] += (double)g0i[i][j][r] +
                             (s1 * (double)g1i[i][j][r] + s2 * (double)g2i[i][j][r]);
          }
          g0r[i][j] = g0i[i][j] = g1r[i][j] = g1i[i][j] = g2r[i][j] = g2i[i][j] =
              f4{0, 0, 0, 0};
        }
    }
  }
  zc* Cp = C.p + (long long)b * C.stride;
  const zc* Dp = D.p ? D.p + (long long)b * D.stride : nullptr;
  const bool use_d = (beta.x != 0.0 || beta.y != 0.0) && Dp != nullptr;
  const double* sAb = sA + (size_t)b * Mp;
  const double* sBb = sB + (size_t)b * Np;
  const double unit = 1.0 / (512.0 * 512.0);  // the 2^-9 of each leading slice
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
        const double sc = sAb[row] * sBb[col] * unit;
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
  return (size_t)batch *
             (2 * NS * (Mp + Np) * Kp * sizeof(_Float16) + (Mp + Np) * sizeof(double)) +
         1024;
}

int zgemm_split_batched_ws(hipStream_t s, int M, int N, int K, zc alpha, ZCMat A, ZCMat B,
                           zc beta, ZCMat D, ZMat C, int batch, void* ws) {
  if (M <= 0 || N <= 0 || batch <= 0) return 0;
  const int Mp = rup(M, TILE), Np = rup(N, TILE), Kp = rup(K > 0 ? K : 1, 16);
  _Float16* As = (_Float16*)ws;
  _Float16* Bs = As + (size_t)batch * 2 * NS * Mp * Kp;
  double* sA =
      (double*)(((uintptr_t)(Bs + (size_t)batch * 2 * NS * Np * Kp) + 255) & ~(uintptr_t)255);
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
