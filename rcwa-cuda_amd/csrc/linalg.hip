// Batched dense complex linear algebra besides GEMM: the blocked
// Gauss-Jordan inverse that replaces the reference's PartialPivLU solves
// (rcwa.cpp:267-281, 335-348, 605-611) and small elementwise kernels.
//
// Gauss-Jordan in place, blocked by nb columns.  One panel step:
//   1. the panel kernel runs nb unblocked Gauss-Jordan steps on the n x nb
//      column panel: pivot search in the not-yet-pivoted rows, a row swap
//      applied to all n columns, and elimination restricted to the panel
//      (for n <= 4096 the panel lives in registers, one to four rows per
//      thread, and the swaps reach the other columns in a separate pass);
//      afterwards the panel holds the nb columns G of the accumulated
//      elementary transform E = I + (G - I_K) e_K^T;
//   2. the other columns receive E as one rank-nb update,
//      M(:, j) += (G - I_K) M(K, j)  for j outside the panel,
//      which is a ZGEMM on the matrix cores.
// For n <= 4096 the scheme has two levels: blocks of 64 columns are factored
// as register panels (16, 8 or 4 columns) whose updates stay inside the
// block, and the rest of the matrix receives one rank-64 update per block
// (the block then holds the columns of the composite transform of its
// panels).
// Row swaps are undone at the end as column swaps in reverse order, so the
// result is A^-1 (not (PA)^-1).
#include <algorithm>
#include "common.hpp"
#include "kernels.hpp"

namespace {

constexpr int GJ_NB = 32;      // panel width of the global-memory panel kernel (n > 4096)
constexpr int GJ_THREADS = 512;
constexpr int GJ_BLOCK = 64;   // outer block of the two-level scheme (n <= 4096)

struct ArgMax {
  double v;
  int i;
};

__device__ __forceinline__ ArgMax better(ArgMax a, ArgMax b) {
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

__global__ __launch_bounds__(GJ_THREADS) void gj_panel_kernel(int n, int k0, int nb, ZMat Mm,
                                                              int* perm, int* info) {
  __shared__ zc rowc[GJ_NB];
  __shared__ ArgMax red[GJ_THREADS / 64];
  __shared__ int s_piv;
  __shared__ zc s_inv;
  const int b = blockIdx.x;
  zc* A = Mm.p + (long long)b * Mm.stride;
  const long long ld = Mm.ld;
  int* pv = perm + (long long)b * n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int kend = min(n, k0 + nb);
  for (int c = k0; c < kend; ++c) {
    // 1. pivot: largest |A(i,c)|^2 over rows c..n-1 (first index on ties)
    ArgMax best{-1.0, n};
    for (int i = c + tid; i < n; i += GJ_THREADS) {
      ArgMax cand{zabs2(A[i + c * ld]), i};
      best = better(best, cand);
    }
    for (int off = 32; off > 0; off >>= 1) {
      ArgMax o{__shfl_down(best.v, off), __shfl_down(best.i, off)};
      best = better(best, o);
    }
    if (lane == 0) red[wave] = best;
    __syncthreads();
    if (tid == 0) {
      ArgMax r = red[0];
      for (int w = 1; w < GJ_THREADS / 64; ++w) r = better(r, red[w]);
      int p = (r.i < n) ? r.i : c;
      if (!(r.v > 0.0)) {
        if (info) info[b] = 1;
        p = c;
      }
      s_piv = p;
      pv[c] = p;
    }
    __syncthreads();
    const int p = s_piv;
    // 2. swap rows c and p across all columns
    if (p != c) {
      for (int j = tid; j < n; j += GJ_THREADS) {
        zc t = A[c + j * ld];
        A[c + j * ld] = A[p + j * ld];
        A[p + j * ld] = t;
      }
    }
    __syncthreads();
    // 3. scaled pivot row within the panel
    if (tid == 0) s_inv = zinv(A[c + c * ld]);
    __syncthreads();
    const zc inv = s_inv;
    if (tid < kend - k0) rowc[tid] = zmul(A[c + (k0 + tid) * ld], inv);
    __syncthreads();
    // 4. eliminate column c from every other row of the panel
    for (int i = tid; i < n; i += GJ_THREADS) {
      const zc aic = A[i + c * ld];
      if (i == c) {
        for (int j = k0; j < kend; ++j)
          A[c + j * ld] = (j == c) ? inv : rowc[j - k0];
      } else {
        for (int j = k0; j < kend; ++j) {
          if (j == c) continue;
          A[i + j * ld] = zsub(A[i + j * ld], zmul(aic, rowc[j - k0]));
        }
        A[i + c * ld] = zneg(zmul(aic, inv));
      }
    }
    __syncthreads();
  }
}

// The same panel step with the panel in registers: thread t of REG_THREADS
// owns rows t, t + REG_THREADS, ... (RPT of them) of the n x NBP panel, so
// n <= RPT * REG_THREADS.  Pivot rows are exchanged between the owning
// threads through LDS and the swap is applied to the panel only;
// gj_swap_kernel applies the panel's swaps to the other columns afterwards.
// Same pivot choice and the same arithmetic per entry as gj_panel_kernel.
constexpr int REG_THREADS = 1024;

template <int RPT, int NBP>
__global__ __launch_bounds__(REG_THREADS) void gj_panel_reg_kernel(int n, int k0, int nb,
                                                                   ZMat Mm, int* perm,
                                                                   int* info) {
  __shared__ zc xrow[2][NBP];  // rows exchanged by a swap
  __shared__ zc rowc[NBP];     // scaled pivot row
  __shared__ ArgMax red[REG_THREADS / 64];
  __shared__ int s_piv;
  __shared__ zc s_inv;
  const int b = blockIdx.x;
  zc* A = Mm.p + (long long)b * Mm.stride;
  const long long ld = Mm.ld;
  int* pv = perm + (long long)b * n;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  zc p[RPT][NBP];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int row = t + r * REG_THREADS;
#pragma unroll
    for (int j = 0; j < NBP; ++j)
      p[r][j] = (row < n && j < nb) ? A[row + (long long)(k0 + j) * ld] : zmake(0, 0);
  }
#pragma unroll
  for (int q = 0; q < NBP; ++q) {
    if (q >= nb) continue;  // uniform: the last panel may be narrower
    const int c = k0 + q;
    // 1. pivot: largest |A(i,c)|^2 over rows c..n-1 (first index on ties)
    ArgMax best{-1.0, n};
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int row = t + r * REG_THREADS;
      if (row < n && row >= c) best = better(best, ArgMax{zabs2(p[r][q]), row});
    }
    for (int off = 32; off > 0; off >>= 1) {
      ArgMax o{__shfl_down(best.v, off), __shfl_down(best.i, off)};
      best = better(best, o);
    }
    if (lane == 0) red[wave] = best;
    __syncthreads();
    if (t == 0) {
      ArgMax r = red[0];
      for (int w = 1; w < REG_THREADS / 64; ++w) r = better(r, red[w]);
      int pr = (r.i < n) ? r.i : c;
      if (!(r.v > 0.0)) {
        if (info) info[b] = 1;
        pr = c;
      }
      s_piv = pr;
      pv[c] = pr;
    }
    __syncthreads();
    const int pr = s_piv;
    // 2. swap panel rows c and pr (owned by threads c % REG_THREADS and
    // pr % REG_THREADS, in their register rows c / REG_THREADS, pr / ...)
    if (pr != c) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int row = t + r * REG_THREADS;
        if (row == c || row == pr) {
#pragma unroll
          for (int j = 0; j < NBP; ++j) xrow[row == c ? 0 : 1][j] = p[r][j];
        }
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int row = t + r * REG_THREADS;
        if (row == c || row == pr) {
#pragma unroll
          for (int j = 0; j < NBP; ++j) p[r][j] = xrow[row == c ? 1 : 0][j];
        }
      }
    }
    // 3. scaled pivot row
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      if (t + r * REG_THREADS == c) {
        const zc inv = zinv(p[r][q]);
        s_inv = inv;
#pragma unroll
        for (int j = 0; j < NBP; ++j) rowc[j] = zmul(p[r][j], inv);
      }
    }
    __syncthreads();
    const zc inv = s_inv;
    // 4. eliminate column c from every other row of the panel
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int row = t + r * REG_THREADS;
      if (row == c) {
#pragma unroll
        for (int j = 0; j < NBP; ++j) p[r][j] = (j == q) ? inv : rowc[j];
      } else if (row < n) {
        const zc aic = p[r][q];
#pragma unroll
        for (int j = 0; j < NBP; ++j)
          if (j != q) p[r][j] = zsub(p[r][j], zmul(aic, rowc[j]));
        p[r][q] = zneg(zmul(aic, inv));
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int row = t + r * REG_THREADS;
    if (row < n) {
#pragma unroll
      for (int j = 0; j < NBP; ++j)
        if (j < nb) A[row + (long long)(k0 + j) * ld] = p[r][j];
    }
  }
}

// Applies the row swaps of panel k0..k0+nb-1 (in order) to the columns
// outside the panel; one thread per column.  The nb swaps touch at most 2 nb
// rows; their net permutation (row dst[k] receives the old row src[k]) is
// formed once per workgroup, so that every load of a column is independent of
// its stores and the loads issue together.
constexpr int GJ_SWAP_MAX = 32;  // rows touched by the swaps of one panel (nb <= 16)
__global__ __launch_bounds__(256) void gj_swap_kernel(int n, int k0, int nb, ZMat Mm,
                                                      const int* perm) {
  __shared__ int s_dst[GJ_SWAP_MAX], s_src[GJ_SWAP_MAX];
  __shared__ int s_m;
  const int b = blockIdx.y;
  const int* pv = perm + (long long)b * n;
  if (threadIdx.x == 0) {
    int m = 0;
    for (int q = 0; q < nb; ++q) {
      const int c = k0 + q, pr = pv[c];
      if (pr == c) continue;
      int ic = -1, ip = -1;
      for (int k = 0; k < m; ++k) {
        if (s_dst[k] == c) ic = k;
        if (s_dst[k] == pr) ip = k;
      }
      if (ic < 0) { ic = m; s_dst[m] = c; s_src[m] = c; ++m; }
      if (ip < 0) { ip = m; s_dst[m] = pr; s_src[m] = pr; ++m; }
      const int t = s_src[ic];
      s_src[ic] = s_src[ip];
      s_src[ip] = t;
    }
    s_m = m;
  }
  __syncthreads();
  const int m = s_m;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (m == 0 || j >= n || (j >= k0 && j < k0 + nb)) return;
  zc* col = Mm.p + (long long)b * Mm.stride + (long long)j * Mm.ld;
  // entries past m repeat the last pair (no branch, so the array stays in
  // registers and the loads issue together): they store the value the last
  // pair stores, to the same row
  double xr[GJ_SWAP_MAX], xi[GJ_SWAP_MAX];
#pragma unroll
  for (int k = 0; k < GJ_SWAP_MAX; ++k) {
    const zc v = col[s_src[min(k, m - 1)]];
    xr[k] = v.x;
    xi[k] = v.y;
  }
#pragma unroll
  for (int k = 0; k < GJ_SWAP_MAX; ++k) col[s_dst[min(k, m - 1)]] = zmake(xr[k], xi[k]);
}

// Gp(i, r) = A(i, k0+r) - delta(i, k0+r) (n x nb);
// Tm(r, j) = A(k0+r, j0+j) (nb x ncols, leading dimension nb), 0 where j0+j
// lies in the panel.
__global__ void gj_prep_kernel(int n, int k0, int nb, int j0, int ncols, ZCMat Mm, zc* Gp,
                               zc* Tm) {
  const int b = blockIdx.y;
  const zc* A = Mm.p + (long long)b * Mm.stride;
  const long long ld = Mm.ld;
  zc* G = Gp + (long long)b * n * nb;
  zc* T = Tm + (long long)b * ncols * nb;
  const int kend = min(n, k0 + nb), w = kend - k0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long first = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  // Gp: rows fastest (consecutive lanes read consecutive rows of a column)
  for (long long idx = first; idx < (long long)n * w; idx += stride) {
    const int i = (int)(idx % n), r = (int)(idx / n);
    zc g = A[i + (long long)(k0 + r) * ld];
    if (i == k0 + r) g.x -= 1.0;
    G[i + (long long)r * n] = g;
  }
  // Tm: the panel's rows, r fastest (the w rows of one column are contiguous)
  for (long long idx = first; idx < (long long)ncols * w; idx += stride) {
    const int r = (int)(idx % w), i = (int)(idx / w);
    const int j = j0 + i;
    T[r + (long long)i * w] = (j >= k0 && j < kend) ? zmake(0, 0) : A[(k0 + r) + (long long)j * ld];
  }
}

__global__ __launch_bounds__(256) void gj_unswap_kernel(int n, ZMat Mm, const int* perm) {
  const int b = blockIdx.x;
  zc* A = Mm.p + (long long)b * Mm.stride;
  const long long ld = Mm.ld;
  const int* pv = perm + (long long)b * n;
  for (int c = n - 1; c >= 0; --c) {
    const int p = pv[c];
    if (p != c) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        zc t = A[i + c * ld];
        A[i + c * ld] = A[i + p * ld];
        A[i + p * ld] = t;
      }
    }
    __syncthreads();
  }
}

__global__ void identity_kernel(int n, ZMat M, zc diag) {
  const int b = blockIdx.y;
  zc* A = M.p + (long long)b * M.stride;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < (long long)n * n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    A[i + j * M.ld] = (i == j) ? diag : zmake(0, 0);
  }
}

__global__ void copy_kernel(int rows, int cols, ZCMat S, ZMat D) {
  const int b = blockIdx.y;
  const zc* s = S.p + (long long)b * S.stride;
  zc* d = D.p + (long long)b * D.stride;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       idx < (long long)rows * cols; idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % rows), j = (int)(idx / rows);
    d[i + j * D.ld] = s[i + j * S.ld];
  }
}

__global__ void scale_rc_kernel(int rows, int cols, ZMat M, const zc* r, long long rs,
                                const zc* c, long long cs) {
  const int b = blockIdx.y;
  zc* A = M.p + (long long)b * M.stride;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       idx < (long long)rows * cols; idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % rows), j = (int)(idx / rows);
    zc v = A[i + j * M.ld];
    if (r) v = zmul(r[b * rs + i], v);
    if (c) v = zmul(v, c[b * cs + j]);
    A[i + j * M.ld] = v;
  }
}

__global__ void scale_cols_to_kernel(int rows, int cols, ZCMat S, const zc* c, long long cs,
                                     ZMat D) {
  const int b = blockIdx.y;
  const zc* s = S.p + (long long)b * S.stride;
  zc* d = D.p + (long long)b * D.stride;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       idx < (long long)rows * cols; idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % rows), j = (int)(idx / rows);
    d[i + j * D.ld] = zmul(s[i + j * S.ld], c[b * cs + j]);
  }
}

__global__ __launch_bounds__(256) void normalize_cols_kernel(int n, ZMat M) {
  __shared__ double red[4];
  const int j = blockIdx.x, b = blockIdx.y;
  zc* col = M.p + (long long)b * M.stride + (long long)j * M.ld;
  double s = 0;
  for (int i = threadIdx.x; i < n; i += 256) s += zabs2(col[i]);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  const double tot = red[0] + red[1] + red[2] + red[3];
  const double inv = tot > 0 ? 1.0 / sqrt(tot) : 1.0;
  for (int i = threadIdx.x; i < n; i += 256) col[i] = zscale(col[i], inv);
}

__global__ void nonfinite_kernel(const zc* p, long long count, int* flag) {
  int bad = 0;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < count;
       idx += (long long)gridDim.x * blockDim.x) {
    const zc v = p[idx];
    if (!isfinite(v.x) || !isfinite(v.y)) bad = 1;
  }
  if (bad) atomicAdd(flag, 1);
}

// out[b] = scale * ||M_b||_F for a batch of rows x cols matrices: one
// workgroup per matrix, each thread a strided share of the columns' entries.
__global__ __launch_bounds__(1024) void frob_kernel(int rows, int cols, ZCMat M, double scale,
                                                    double* out) {
  __shared__ double red[16];
  const zc* p = M.p + (long long)blockIdx.x * M.stride;
  double acc = 0;
  for (int j = 0; j < cols; ++j) {
    const zc* col = p + (long long)j * M.ld;
    for (int i = threadIdx.x; i < rows; i += 1024) acc += zabs2(col[i]);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0;
    for (int w = 0; w < 16; ++w) tot += red[w];
    out[blockIdx.x] = scale * sqrt(tot);
  }
}

inline dim3 grid_for(long long elems, int batch) {
  long long blocks = (elems + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  return dim3((unsigned)blocks, (unsigned)batch);
}

}  // namespace

size_t gj_workspace_bytes(int n, int batch) {
  const size_t nb = std::max(GJ_NB, GJ_BLOCK);
  return (size_t)batch * (2 * n * nb * sizeof(zc) + n * sizeof(int)) + 256;
}

int gj_inverse_batched(hipStream_t s, int n, ZMat M, int batch, void* ws, int* info) {
  if (n <= 0 || batch <= 0) return 0;
  zc* Gp = (zc*)ws;
  zc* Tm = Gp + (size_t)batch * n * GJ_BLOCK;
  int* perm = (int*)(Tm + (size_t)batch * n * GJ_BLOCK);
  // M(:, j0:j0+ncols) += Gp * Tm for the panel k0..k0+w-1
  auto update = [&](int k0, int w, int j0, int ncols) {
    hipLaunchKernelGGL(gj_prep_kernel, grid_for((long long)std::max(n, ncols) * w, batch),
                       dim3(256), 0, s, n, k0, w, j0, ncols, ZCMat{M.p, M.ld, M.stride}, Gp, Tm);
    ZMat C{M.p + (long long)j0 * M.ld, M.ld, M.stride};
    return zgemm_batched(s, n, ncols, w, zmake(1, 0), ZCMat{Gp, n, (long long)n * w},
                         ZCMat{Tm, w, (long long)ncols * w}, zmake(1, 0),
                         ZCMat{C.p, C.ld, C.stride}, C, batch);
  };
  // Register panels: one row per thread up to n = 1024 (16-column panels),
  // two up to 2048 (8 columns), four up to 4096 (4 columns).
  int rpt = 0, nbp = 0;
  if (n <= REG_THREADS) { rpt = 1; nbp = 16; }
  else if (n <= 2 * REG_THREADS) { rpt = 2; nbp = 8; }
  else if (n <= 4 * REG_THREADS) { rpt = 4; nbp = 4; }
  if (rpt) {
    // Two levels: blocks of GJ_BLOCK columns, each factored as panels of
    // nbp columns in registers with the updates kept inside the block; then
    // one rank-GJ_BLOCK update of the rest of the matrix per block.
    for (int K0 = 0; K0 < n; K0 += GJ_BLOCK) {
      const int W = std::min(GJ_BLOCK, n - K0);
      for (int k0 = K0; k0 < K0 + W; k0 += nbp) {
        const int w = std::min(nbp, K0 + W - k0);
        if (rpt == 1)
          hipLaunchKernelGGL((gj_panel_reg_kernel<1, 16>), dim3(batch), dim3(REG_THREADS), 0, s,
                             n, k0, w, M, perm, info);
        else if (rpt == 2)
          hipLaunchKernelGGL((gj_panel_reg_kernel<2, 8>), dim3(batch), dim3(REG_THREADS), 0, s,
                             n, k0, w, M, perm, info);
        else
          hipLaunchKernelGGL((gj_panel_reg_kernel<4, 4>), dim3(batch), dim3(REG_THREADS), 0, s,
                             n, k0, w, M, perm, info);
        if (2 * w > GJ_SWAP_MAX) return (int)hipErrorInvalidValue;
        hipLaunchKernelGGL(gj_swap_kernel, dim3((n + 255) / 256, batch), dim3(256), 0, s, n, k0,
                           w, M, perm);
        if (W > w) {
          int e = update(k0, w, K0, W);
          if (e) return e;
        }
      }
      if (W < n) {
        int e = update(K0, W, 0, n);
        if (e) return e;
      }
    }
  } else {
    for (int k0 = 0; k0 < n; k0 += GJ_NB) {
      const int w = std::min(GJ_NB, n - k0);
      hipLaunchKernelGGL(gj_panel_kernel, dim3(batch), dim3(GJ_THREADS), 0, s, n, k0, w, M, perm,
                         info);
      int e = update(k0, w, 0, n);
      if (e) return e;
    }
  }
  hipLaunchKernelGGL(gj_unswap_kernel, dim3(batch), dim3(256), 0, s, n, M, perm);
  return (int)hipGetLastError();
}

namespace {

constexpr int BLK_INV_BASE = 512;  // blocks this size or smaller: Gauss-Jordan
constexpr int BLK_INV_MAXB = 64;   // larger batches keep Gauss-Jordan busy enough

bool use_block_inverse(int n, int batch) { return n > BLK_INV_BASE && batch <= BLK_INV_MAXB; }

// entries per matrix of the recursion's work buffers: per level n1 x n2 for
// A^-1 B and n2 x n1 for C A^-1 (n1 = m / 2), the deeper levels after them
size_t block_inverse_rec_entries(int m) {
  size_t z = 0;
  while (m > BLK_INV_BASE) {
    const int n1 = m / 2, n2 = m - n1;
    z += 2 * (size_t)n1 * n2;
    m = n2;  // the larger half: the inverse of A (n1) runs in the same space first
  }
  return z;
}

// M = [A B; C D] (A n1 x n1) in place:
//   A <- A^-1 (recursively), W = A^-1 B, D <- S = D - C W, D <- S^-1
//   (recursively), V = C A^-1, B <- -W S^-1, A <- A^-1 - B V, C <- -S^-1 V,
// which is M^-1 = [A^-1 + A^-1 B S^-1 C A^-1, -A^-1 B S^-1; -S^-1 C A^-1,
// S^-1].
int block_inverse_rec(hipStream_t s, int m, ZMat M, int batch, void* gjws, zc* rws, int* info) {
  if (m <= BLK_INV_BASE) return gj_inverse_batched(s, m, M, batch, gjws, info);
  const int n1 = m / 2, n2 = m - n1;
  const long long ld = M.ld, st = M.stride;
  const zc one = zmake(1, 0), mone = zmake(-1, 0), zero = zmake(0, 0);
  const ZCMat none{nullptr, 0, 0};
  const ZMat A{M.p, ld, st}, B{M.p + (long long)n1 * ld, ld, st}, C{M.p + n1, ld, st},
      D{M.p + n1 + (long long)n1 * ld, ld, st};
  auto cm = [](const ZMat& x) { return ZCMat{x.p, x.ld, x.stride}; };
  const long long blk = (long long)n1 * n2;
  const ZMat W{rws, n1, blk}, V{rws + (size_t)batch * blk, n2, blk};
  zc* sub = rws + 2 * (size_t)batch * blk;
  int e;
  if ((e = block_inverse_rec(s, n1, A, batch, gjws, sub, info))) return e;
  if ((e = zgemm_batched(s, n1, n2, n1, one, cm(A), cm(B), zero, none, W, batch))) return e;
  if ((e = zgemm_batched(s, n2, n2, n1, mone, cm(C), cm(W), one, cm(D), D, batch))) return e;
  if ((e = block_inverse_rec(s, n2, D, batch, gjws, sub, info))) return e;
  if ((e = zgemm_batched(s, n2, n1, n1, one, cm(C), cm(A), zero, none, V, batch))) return e;
  if ((e = zgemm_batched(s, n1, n2, n2, mone, cm(W), cm(D), zero, none, B, batch))) return e;
  if ((e = zgemm_batched(s, n1, n1, n2, mone, cm(B), cm(V), one, cm(A), A, batch))) return e;
  return zgemm_batched(s, n2, n1, n2, mone, cm(D), cm(V), zero, none, C, batch);
}

}  // namespace

size_t block_inverse_ws_bytes(int n, int cap) {
  size_t b = gj_workspace_bytes(n, cap);
  if (n > BLK_INV_BASE)
    b += 256 + (size_t)std::min(cap, BLK_INV_MAXB) * block_inverse_rec_entries(n) * sizeof(zc);
  return b;
}

int block_inverse_batched(hipStream_t s, int n, ZMat M, int batch, void* ws, int* info) {
  if (!use_block_inverse(n, batch)) return gj_inverse_batched(s, n, M, batch, ws, info);
  // the Gauss-Jordan part for this batch first, then the recursion's buffers:
  // both fit block_inverse_ws_bytes(n, cap) for batch <= cap
  zc* rws = (zc*)(((uintptr_t)ws + gj_workspace_bytes(n, batch) + 255) & ~(uintptr_t)255);
  return block_inverse_rec(s, n, M, batch, ws, rws, info);
}

namespace {

// Column sums of |re| + |im| of M and F, COND_COLS columns per workgroup of
// four waves (a wave per column at a time, lanes along the rows); each
// workgroup writes the largest sums of its columns to part[b][chunk][2].
constexpr int COND_COLS = 16;
__global__ __launch_bounds__(256) void cond1_part_kernel(int n, ZCMat Mm, ZCMat Fm, double* part) {
  __shared__ double red[2][4];
  const int b = blockIdx.y, chunk = blockIdx.x, nch = gridDim.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const zc* Mp = Mm.p + (long long)b * Mm.stride;
  const zc* Fp = Fm.p + (long long)b * Fm.stride;
  double mm = 0.0, mf = 0.0;
  const int j1 = min(n, (chunk + 1) * COND_COLS);
  for (int j = chunk * COND_COLS + wave; j < j1; j += 4) {
    double cm = 0.0, cf = 0.0;
    for (int i = lane; i < n; i += 64) {
      const zc a = Mp[i + (long long)j * Mm.ld], f = Fp[i + (long long)j * Fm.ld];
      cm += fabs(a.x) + fabs(a.y);
      cf += fabs(f.x) + fabs(f.y);
    }
    for (int o = 32; o > 0; o >>= 1) {
      cm += __shfl_xor(cm, o);
      cf += __shfl_xor(cf, o);
    }
    mm = fmax(mm, cm);
    mf = fmax(mf, cf);
  }
  if (lane == 0) {
    red[0][wave] = mm;
    red[1][wave] = mf;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double* o = part + ((long long)b * nch + chunk) * 2;
    o[0] = fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3]));
    o[1] = fmax(fmax(red[1][0], red[1][1]), fmax(red[1][2], red[1][3]));
  }
}

// One wave per matrix: the largest column sums over the chunks, then the flag.
__global__ __launch_bounds__(64) void cond1_flag_kernel(int nch, const double* part, double kappa,
                                                        int* flag) {
  const int b = blockIdx.x, lane = threadIdx.x;
  const double* p = part + (long long)b * nch * 2;
  double a = 0.0, f = 0.0;
  for (int c = lane; c < nch; c += 64) {
    a = fmax(a, p[2 * c]);
    f = fmax(f, p[2 * c + 1]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    a = fmax(a, __shfl_xor(a, o));
    f = fmax(f, __shfl_xor(f, o));
  }
  if (lane == 0) flag[b] = !(a * f <= kappa);  // a NaN estimate counts as ill-conditioned
}

}  // namespace

size_t cond1_scratch_bytes(int n, int batch) {
  return (size_t)batch * ((n + COND_COLS - 1) / COND_COLS) * 2 * sizeof(double);
}

int cond1_flags(hipStream_t s, int n, ZCMat M, ZCMat F, double kappa, int* flag, int batch,
                double* scratch) {
  if (n <= 0 || batch <= 0) return 0;
  const int nch = (n + COND_COLS - 1) / COND_COLS;
  hipLaunchKernelGGL(cond1_part_kernel, dim3(nch, batch), dim3(256), 0, s, n, M, F, scratch);
  if (int e = (int)hipGetLastError()) return e;
  hipLaunchKernelGGL(cond1_flag_kernel, dim3(batch), dim3(64), 0, s, nch, scratch, kappa, flag);
  return (int)hipGetLastError();
}

int zset_identity(hipStream_t s, int n, ZMat M, int batch, zc diag) {
  hipLaunchKernelGGL(identity_kernel, grid_for((long long)n * n, batch), dim3(256), 0, s, n, M,
                     diag);
  return (int)hipGetLastError();
}

int zcopy_batched(hipStream_t s, int rows, int cols, ZCMat src, ZMat dst, int batch) {
  hipLaunchKernelGGL(copy_kernel, grid_for((long long)rows * cols, batch), dim3(256), 0, s, rows,
                     cols, src, dst);
  return (int)hipGetLastError();
}

int zscale_rc(hipStream_t s, int rows, int cols, ZMat M, const zc* r, long long rstride,
              const zc* c, long long cstride, int batch) {
  hipLaunchKernelGGL(scale_rc_kernel, grid_for((long long)rows * cols, batch), dim3(256), 0, s,
                     rows, cols, M, r, rstride, c, cstride);
  return (int)hipGetLastError();
}

int zscale_cols_to(hipStream_t s, int rows, int cols, ZCMat src, const zc* c, long long cstride,
                   ZMat dst, int batch) {
  hipLaunchKernelGGL(scale_cols_to_kernel, grid_for((long long)rows * cols, batch), dim3(256), 0,
                     s, rows, cols, src, c, cstride, dst);
  return (int)hipGetLastError();
}

int znormalize_cols(hipStream_t s, int n, ZMat M, int batch) {
  hipLaunchKernelGGL(normalize_cols_kernel, dim3(n, batch), dim3(256), 0, s, n, M);
  return (int)hipGetLastError();
}

int zcount_nonfinite(hipStream_t s, const zc* p, long long count, int* flag) {
  hipLaunchKernelGGL(nonfinite_kernel, grid_for(count, 1), dim3(256), 0, s, p, count, flag);
  return (int)hipGetLastError();
}

int zfrob_norms(hipStream_t s, int rows, int cols, ZCMat M, double scale, double* out,
                int batch) {
  if (batch <= 0) return 0;
  hipLaunchKernelGGL(frob_kernel, dim3(batch), dim3(1024), 0, s, rows, cols, M, scale, out);
  return (int)hipGetLastError();
}
