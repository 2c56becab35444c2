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
// Tiling: 256 threads = 4 wave64s in a 2x2 arrangement; the workgroup tile
// is 64x64 of C, each wave owns 32x32 = 2x2 MFMA tiles of 16x16.  K advances
// in slabs of 16 staged through LDS (double-buffered: the next slab's global
// loads are issued before the MFMAs on the current one).
#include "common.hpp"
#include "kernels.hpp"

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int BM = 64, BN = 64, BK = 16;
constexpr unsigned XCDS = 8;  // accelerator complex dies of the MI355X, each with its own L2
constexpr int PADA = 1, PADB = 1;

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Two workgroups per CU (their LDS fits twice): registers capped at 256 per
// lane, which the three accumulator sets still fit without spilling.
// tri: ZGEMM_TRI_C (only the tiles of C that meet its upper triangle are
// formed; the others are left as they are), ZGEMM_TRI_B (B is upper
// triangular: the tile of columns n0..n0+BN-1 needs rows k < n0 + BN only)
// and ZGEMM_TRI_A (A is upper triangular: the tile of rows m0.. needs
// columns k >= m0 only).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void zgemm_kernel(
    int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta, ZCMat D, ZMat C, int tiles_n,
    int tri) {
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
  if ((tri & ZGEMM_TRI_C) && m0 >= n0 + BN) return;  // uniform over the workgroup
  if (tri & ZGEMM_TRI_B) K = min(K, n0 + BN);
  const int kbeg = (tri & ZGEMM_TRI_A) ? min(m0, K) : 0;  // a multiple of BK
  const zc* Ap = A.p + (long long)b * A.stride;
  const zc* Bp = B.p + (long long)b * B.stride;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
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

  zc ra[4], rb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + 256 * r;
      const int i = idx & (BM - 1), k = idx >> 6;       // A: rows fastest
      const int gi = m0 + i, gk = k0 + k;
      ra[r] = (gi < M && gk < K) ? Ap[gi + (long long)gk * A.ld] : zmake(0, 0);
      const int kb = idx & (BK - 1), j = idx >> 4;        // B: k fastest
      const int gkb = k0 + kb, gj = n0 + j;
      rb[r] = (gkb < K && gj < N) ? Bp[gkb + (long long)gj * B.ld] : zmake(0, 0);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + 256 * r;
      As[buf][idx >> 6][idx & (BM - 1)] = ra[r];
      Bs[buf][idx & (BK - 1)][idx >> 4] = rb[r];
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

int zgemm_batched_tri(hipStream_t s, int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta,
                      ZCMat D, ZMat C, int batch, int tri) {
  if (M <= 0 || N <= 0 || batch <= 0) return 0;
  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  const long long total = (long long)tiles_m * tiles_n * batch;
  if (total > 0x7fffffffLL) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(zgemm_kernel, dim3((unsigned)total, 1, 1), dim3(256), 0, s, M, N, K, alpha, A,
                     B, beta, D, C, tiles_n, tri);
  return (int)hipGetLastError();
}

int zgemm_batched(hipStream_t s, int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta,
                  ZCMat D, ZMat C, int batch) {
  return zgemm_batched_tri(s, M, N, K, alpha, A, B, beta, D, C, batch, 0);
}
