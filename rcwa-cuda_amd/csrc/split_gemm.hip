// Split-precision complex GEMM on the gfx950 int8 matrix cores, the product
// used by the split form of the path (the layer S-matrix and Redheffer
// multiplies); Omega^2 = P Q, V = Q W, the eigensolver and every inverse stay
// FP64.
//
//   C = alpha * A * B + beta * D     (complex128 in and out, column-major)
//
// Ozaki scheme with fixed-point slices.  Every row of A and every column of B
// is scaled by a power of two s so that its real and imaginary parts lie in
// (-1/2, 1/2); each scaled real number u is then cut into six 7-bit integer
// slices
//   u = 2^-7 a1 + 2^-14 a2 + ... + 2^-42 a6 + r,   |a_i| <= 64, |r| <= 2^-43,
// all exact in int8 (rint of the scaled remainder, exact in FP64).  A real
// product keeps the slice products a_i b_j with i + j <= 7 (21 of 36), grouped
// by weight 2^-7(i+j): the dropped ones and the remainders are below 2^-41
// of the product's scale, the normwise error of the product.
//
// v_mfma_i32_16x16x64_i8 accumulates each weight group EXACTLY in int32: a
// group holds at most six slice pairs, each a complex combination of two
// products below 2^12, so over K <= 2^15 every sum stays below 2^31.  The
// product is therefore exact up to the dropped terms however the instruction
// orders its k, and no partial sum is promoted during the K loop: the
// epilogue combines the six groups in FP64, unscales and applies alpha/beta.
// Longer K is cut into chunks of 2^15 accumulated through C.
//
// A complex 16x16x64 step takes 84 MFMAs (21 slice pairs x 4 real products,
// the real part's -ai bi from stored negated slices of A).  Every lane of a
// 16x16x64 MFMA supplies sixteen consecutive k of one row of A and the same
// sixteen k of one column of B.
//
// Three passes per product:
//   split_scale_kernel  the power-of-two scale of every row of A and column
//                       of B (one read, coalesced: a thread per row where
//                       rows are contiguous, a wave per column where k is);
//   split_slice_kernel  the slices, elementwise (a thread per vector and
//                       group of sixteen k), in the MFMA operand layout
//                       S[k/16][vector][k%16] per slice plane: a lane's
//                       sixteen bytes are one load and 16 lanes read 256
//                       contiguous bytes;
//   split_mm_kernel     the products.
#include "common.hpp"
#include "runtime.hpp"

typedef int v4i __attribute__((ext_vector_type(4)));

namespace {

constexpr int TILE = 64;
constexpr int KS = 64;               // k per MFMA step
constexpr int NSL = 6;               // slices per real number
constexpr int NG = 6;                // weight groups i + j = 2 .. 7
constexpr int PA = 3 * NSL;          // planes of A: re, im, -im
constexpr int PB = 2 * NSL;          // planes of B: re, im
constexpr int KCHUNK = 1 << 15;      // K of one exact int32 accumulation

inline int rup(int x, int m) { return (x + m - 1) / m * m; }

// Vector v of batch entry b: element k at src + b*stride + v*vs + k*ks.
// scale[b*Vp + v] = 2^e with 2^e >= 2 |re|, 2 |im| of every element (1 for a
// zero vector and for the padding v >= rows).  kfast: ks == 1, one wave per
// vector with the lanes along k; otherwise a thread per vector (vs == 1
// keeps the waves' loads contiguous).
__global__ __launch_bounds__(256) void split_scale_kernel(int rows, int K, const zc* src,
                                                          long long stride, long long vs,
                                                          long long ks, int Vp, double* scale,
                                                          int kfast) {
  const int b = blockIdx.y;
  const zc* base = src + (long long)b * stride;
  double m = 0.0;
  int v;
  if (kfast) {
    const int lane = threadIdx.x & 63;
    v = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (v < rows)
      for (int k = lane; k < K; k += 64) {
        const zc z = base[(long long)v * vs + k];
        m = fmax(m, fmax(fabs(z.x), fabs(z.y)));
      }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if (lane != 0 || v >= Vp) return;
  } else {
    v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= Vp) return;
    if (v < rows)
      for (int k = 0; k < K; ++k) {
        const zc z = base[(long long)v * vs + (long long)k * ks];
        m = fmax(m, fmax(fabs(z.x), fabs(z.y)));
      }
  }
  int e = 0;
  if (m > 0.0) frexp(m, &e);  // m < 2^e
  scale[(size_t)b * Vp + v] = ldexp(1.0, e + 1);
}

// The slice planes of Vp x Kp bytes each (re slices, im slices and, with
// neg, the negated im slices): thread t takes vector v = t % Vp and k = 16
// (t / Vp) .. +15.
__global__ __launch_bounds__(256) void split_slice_kernel(int rows, int K, const zc* src,
                                                          long long stride, long long vs,
                                                          long long ks, int Vp, int Kp,
                                                          const double* scale, signed char* out,
                                                          int neg) {
  const int b = blockIdx.y;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int q16 = Kp / 16;
  if (t >= (long long)Vp * q16) return;
  const int v = (int)(t % Vp), q = (int)(t / Vp);
  const zc* p = src + (long long)b * stride + (long long)v * vs;
  const size_t plane = (size_t)Vp * Kp;
  signed char* o = out + (size_t)b * (neg ? PA : PB) * plane;
  const double s = 1.0 / scale[(size_t)b * Vp + v];  // a power of two: exact
  // byte u of word w of slice c*NSL + i: slice i of part c at k = 16 q + 4 w + u
  unsigned w[2][NSL][4];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int i = 0; i < NSL; ++i)
#pragma unroll
      for (int x = 0; x < 4; ++x) w[c][i][x] = 0;
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int k = 16 * q + u;
    zc z = zmake(0, 0);
    if (v < rows && k < K) z = p[(long long)k * ks];
    const double part[2] = {z.x * s, z.y * s};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      // exact in FP64: scaling by powers of two and subtracting the rounded
      // leading bits; every slice lies in [-64, 64]
      double tt = part[c] * 128.0;
#pragma unroll
      for (int i = 0; i < NSL; ++i) {
        const double a = rint(tt);
        tt = (tt - a) * 128.0;
        w[c][i][u >> 2] |= ((unsigned)(int)a & 0xffu) << (8 * (u & 3));
      }
    }
  }
  const size_t at = ((size_t)q * Vp + v) * 16;
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int i = 0; i < NSL; ++i)
      *(v4i*)(o + (size_t)(c * NSL + i) * plane + at) =
          v4i{(int)w[c][i][0], (int)w[c][i][1], (int)w[c][i][2], (int)w[c][i][3]};
  if (neg) {
#pragma unroll
    for (int i = 0; i < NSL; ++i) {
      unsigned nw[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        // bytewise negation of the imaginary slices (each in [-64, 64])
        nw[x] = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int byte = (int)(signed char)((w[1][i][x] >> (8 * u)) & 0xffu);
          nw[x] |= ((unsigned)(-byte) & 0xffu) << (8 * u);
        }
      }
      *(v4i*)(o + (size_t)(2 * NSL + i) * plane + at) =
          v4i{(int)nw[0], (int)nw[1], (int)nw[2], (int)nw[3]};
    }
  }
}

__device__ __forceinline__ v4i mfi(v4i a, v4i b, v4i c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// Planes of A: ar_i (0 .. 5), ai_i (6 .. 11), -ai_i (12 .. 17); of B: br_i,
// bi_i.  256 threads = 4 waves in 2x2, a 32x32 block of C per wave (2x2
// MFMA tiles of 16x16).  K <= KCHUNK (the host splits longer products).
__global__ __launch_bounds__(256) void split_mm_kernel(int M, int N, int Kp, int Mp, int Np,
                                                       const signed char* As, const double* sA,
                                                       const signed char* Bs, const double* sB,
                                                       zc alpha, zc beta, ZCMat D, ZMat C,
                                                       int tiles_n) {
  const int b = blockIdx.z;
  const int tn = blockIdx.x % tiles_n, tm = blockIdx.x / tiles_n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lk = lane >> 4;
  const size_t pa = (size_t)Mp * Kp, pb = (size_t)Np * Kp;
  const signed char* A0 = As + (size_t)b * PA * pa;
  const signed char* B0 = Bs + (size_t)b * PB * pb;
  const int r0 = tm * TILE + wm * 32, c0 = tn * TILE + wn * 32;
  // weight group g (i + j = g + 2) of tile (i, j): real and imaginary parts
  v4i gr[2][2][NG], gi[2][2][NG];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int g = 0; g < NG; ++g) gr[i][j][g] = gi[i][j][g] = v4i{0, 0, 0, 0};
  for (int k0 = 0; k0 < Kp; k0 += KS) {
    const size_t q = (size_t)(k0 / 16 + lk);
    v4i br[2][NSL], bi[2][NSL];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const size_t at = (q * Np + c0 + j * 16 + li) * 16;
#pragma unroll
      for (int s = 0; s < NSL; ++s) {
        br[j][s] = *(const v4i*)(B0 + s * pb + at);
        bi[j][s] = *(const v4i*)(B0 + (NSL + s) * pb + at);
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const size_t at = (q * Mp + r0 + i * 16 + li) * 16;
      v4i ar[NSL], ai[NSL], nai[NSL];
#pragma unroll
      for (int s = 0; s < NSL; ++s) {
        ar[s] = *(const v4i*)(A0 + s * pa + at);
        ai[s] = *(const v4i*)(A0 + (NSL + s) * pa + at);
        nai[s] = *(const v4i*)(A0 + (2 * NSL + s) * pa + at);
      }
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int x = 0; x < NSL; ++x)
#pragma unroll
          for (int y = 0; y + x < NG; ++y) {
            gr[i][j][x + y] = mfi(ar[x], br[j][y], gr[i][j][x + y]);
            gr[i][j][x + y] = mfi(nai[x], bi[j][y], gr[i][j][x + y]);
            gi[i][j][x + y] = mfi(ar[x], bi[j][y], gi[i][j][x + y]);
            gi[i][j][x + y] = mfi(ai[x], br[j][y], gi[i][j][x + y]);
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
        // sum of the groups from the smallest weight up: 2^-7(g+2)
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int g = NG - 1; g >= 0; --g) {
          re = re * (1.0 / 128.0) + (double)gr[i][j][g][r];
          im = im * (1.0 / 128.0) + (double)gi[i][j][g][r];
        }
        const double sc = sAb[row] * sBb[col] * (1.0 / 16384.0);
        zc v = zmul(alpha, zmake(re * sc, im * sc));
        if (use_d) v = zfma(beta, Dp[row + (long long)col * D.ld], v);
        Cp[row + (long long)col * C.ld] = v;
      }
    }
}

// Scales and slices of Vp vectors of length K (Kp padded) of every batch
// entry into out / scale.
int split_operand(hipStream_t s, int rows, int K, const zc* src, long long stride, long long vs,
                  long long ks, int Vp, int Kp, signed char* out, double* scale, int batch,
                  int neg) {
  if (ks == 1)
    hipLaunchKernelGGL(split_scale_kernel, dim3((Vp + 3) / 4, batch), dim3(256), 0, s, rows, K,
                       src, stride, vs, ks, Vp, scale, 1);
  else
    hipLaunchKernelGGL(split_scale_kernel, dim3((Vp + 255) / 256, batch), dim3(256), 0, s, rows,
                       K, src, stride, vs, ks, Vp, scale, 0);
  if (int e = (int)hipGetLastError()) return e;
  const long long threads = (long long)Vp * (Kp / 16);
  hipLaunchKernelGGL(split_slice_kernel, dim3((unsigned)((threads + 255) / 256), batch),
                     dim3(256), 0, s, rows, K, src, stride, vs, ks, Vp, Kp, scale, out, neg);
  return (int)hipGetLastError();
}

}  // namespace

size_t zgemm_split_ws_bytes(int M, int N, int K, int batch) {
  const size_t Mp = rup(M, TILE), Np = rup(N, TILE);
  const size_t Kp = rup(std::min(K > 0 ? K : 1, KCHUNK), KS);
  return (size_t)batch * ((PA * Mp + PB * Np) * Kp + (Mp + Np) * sizeof(double)) + 1024;
}

int zgemm_split_batched_ws(hipStream_t s, int M, int N, int K, zc alpha, ZCMat A, ZCMat B,
                           zc beta, ZCMat D, ZMat C, int batch, void* ws) {
  if (M <= 0 || N <= 0 || batch <= 0) return 0;
  const int Mp = rup(M, TILE), Np = rup(N, TILE);
  const int tiles_n = Np / TILE, tiles_m = Mp / TILE;
  // chunks of at most KCHUNK k, accumulated through C (beta = 1, D = C after
  // the first)
  for (int k0 = 0; k0 == 0 || k0 < K; k0 += KCHUNK) {
    const int Kc = std::min(K - k0, KCHUNK);
    const int Kp = rup(Kc > 0 ? Kc : 1, KS);
    signed char* As = (signed char*)ws;
    signed char* Bs = As + (size_t)batch * PA * Mp * Kp;
    double* sA =
        (double*)(((uintptr_t)(Bs + (size_t)batch * PB * Np * Kp) + 255) & ~(uintptr_t)255);
    double* sB = sA + (size_t)batch * Mp;
    // rows of A: element k of row v at A + v + k lda; columns of B: element
    // k of column v at B + v ldb + k
    if (int e = split_operand(s, M, Kc, A.p + (long long)k0 * A.ld, A.stride, 1LL, A.ld, Mp, Kp,
                              As, sA, batch, 1))
      return e;
    if (int e = split_operand(s, N, Kc, B.p + k0, B.stride, B.ld, 1LL, Np, Kp, Bs, sB, batch, 0))
      return e;
    const zc b_ = k0 == 0 ? beta : zmake(1, 0);
    const ZCMat D_ = k0 == 0 ? D : ZCMat{C.p, C.ld, C.stride};
    hipLaunchKernelGGL(split_mm_kernel, dim3(tiles_m * tiles_n, 1, batch), dim3(256), 0, s, M, N,
                       Kp, Mp, Np, As, sA, Bs, sB, alpha, b_, D_, C, tiles_n);
    if (int e = (int)hipGetLastError()) return e;
    if (K <= 0) break;
  }
  return 0;
}
