// Stage kernels of the RCWA layer pipeline that are not dense products:
// convolution-matrix assembly (rcwa.cpp:6-82), P/Q assembly
// (rcwa.cpp:264-285), mode constants (rcwa.cpp:317, 339) and the A/B
// matrices (rcwa.cpp:335-338).
#include <algorithm>
#include "common.hpp"
#include "kernels.hpp"
#include "rcwa_amd.h"

namespace {

// ---------------------------------------------------------------- ConvMat
// Only the (4hx+1) x (4hy+1) Fourier coefficients that the Toeplitz gather
// reads (rcwa.cpp:71-76) are formed, as a truncated DFT instead of a full
// 2-D FFT: pass 1 reduces each grid row to the 4hx+1 x-frequencies, pass 2
// reduces the columns of that result to the 4hy+1 y-frequencies.  Twiddles
// are exact table lookups e^{-2 pi i (f*x mod N)/N}, staged in LDS.
constexpr int DFT_THREADS = 256;

__global__ __launch_bounds__(DFT_THREADS) void dft_rows_kernel(const zc* grids, int Nx, int Ny,
                                                               int nfx, int hxx, zc* R) {
  extern __shared__ zc sm[];
  zc* row = sm;        // Nx
  zc* tw = sm + Nx;    // Nx
  const int iy = blockIdx.x, b = blockIdx.y;
  const zc* g = grids + ((long long)b * Ny + iy) * Nx;
  for (int i = threadIdx.x; i < Nx; i += DFT_THREADS) {
    row[i] = g[i];
    double s, c;
    sincospi(-2.0 * (double)i / (double)Nx, &s, &c);
    tw[i] = zmake(c, s);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int f = wave; f < nfx; f += DFT_THREADS / 64) {
    const int fx = f - hxx;                       // -2hx .. 2hx
    const long long fm = ((fx % Nx) + Nx) % Nx;
    zc acc = zmake(0, 0);
    for (int ix = lane; ix < Nx; ix += 64) acc = zfma(row[ix], tw[(fm * ix) % Nx], acc);
    for (int off = 32; off > 0; off >>= 1)
      acc = zadd(acc, zmake(__shfl_down(acc.x, off), __shfl_down(acc.y, off)));
    if (lane == 0) R[((long long)b * Ny + iy) * nfx + f] = acc;
  }
}

__global__ __launch_bounds__(DFT_THREADS) void dft_cols_kernel(const zc* R, int Nx, int Ny,
                                                               int nfx, int nfy, int hyy, zc* F) {
  extern __shared__ zc sm[];
  zc* col = sm;        // Ny
  zc* tw = sm + Ny;    // Ny
  const int f = blockIdx.x, b = blockIdx.y;
  for (int i = threadIdx.x; i < Ny; i += DFT_THREADS) {
    col[i] = R[((long long)b * Ny + i) * nfx + f];
    double s, c;
    sincospi(-2.0 * (double)i / (double)Ny, &s, &c);
    tw[i] = zmake(c, s);
  }
  __syncthreads();
  const double inv = 1.0 / ((double)Nx * (double)Ny);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int q = wave; q < nfy; q += DFT_THREADS / 64) {
    const int fy = q - hyy;
    const long long fm = ((fy % Ny) + Ny) % Ny;
    zc acc = zmake(0, 0);
    for (int iy = lane; iy < Ny; iy += 64) acc = zfma(col[iy], tw[(fm * iy) % Ny], acc);
    for (int off = 32; off > 0; off >>= 1)
      acc = zadd(acc, zmake(__shfl_down(acc.x, off), __shfl_down(acc.y, off)));
    if (lane == 0) F[((long long)b * nfy + q) * nfx + f] = zscale(acc, inv);
  }
}

__global__ void gather_kernel(const zc* F, int Nx, int Ny, int hx, int hy, ZMat C,
                              const int* map) {
  const int P = 2 * hx + 1, Q = 2 * hy + 1, PQ = P * Q;
  const int nfx = 4 * hx + 1, nfy = 4 * hy + 1;
  const int b = blockIdx.y;
  const zc* Fb = F + (long long)b * nfx * nfy;
  zc* Cb = C.p + (long long)(map ? map[b] : b) * C.stride;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < (long long)PQ * PQ;
       idx += (long long)gridDim.x * blockDim.x) {
    const int m = (int)(idx % PQ), n = (int)(idx / PQ);
    const int dp = (m / Q) - (n / Q), dq = (m % Q) - (n % Q);
    const int kx = dp + Nx / 2, ky = dq + Ny / 2;
    zc v = zmake(0, 0);
    if (kx >= 0 && kx < Nx && ky >= 0 && ky < Ny) v = Fb[(dq + 2 * hy) * nfx + (dp + 2 * hx)];
    Cb[m + (long long)n * C.ld] = v;
  }
}

// ---------------------------------------------------------------- P, Q
__global__ void build_pq_kernel(int N, ZCMat erc, ZCMat urc, ZCMat erci, ZCMat urci,
                                const zc* Kx, const zc* Ky, ZMat P, ZMat Q) {
  const int b = blockIdx.y, n = 2 * N;
  const zc* kx = Kx + (long long)b * N;
  const zc* ky = Ky + (long long)b * N;
  const zc* E = erc.p + (long long)b * erc.stride;
  const zc* U = urc.p + (long long)b * urc.stride;
  const zc* Ei = erci.p + (long long)b * erci.stride;
  const zc* Ui = urci.p + (long long)b * urci.stride;
  zc* Pb = P.p + (long long)b * P.stride;
  zc* Qb = Q.p + (long long)b * Q.stride;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < (long long)n * n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    const int bi = i / N, bj = j / N, ii = i - bi * N, jj = j - bj * N;
    const zc ki = bi == 0 ? kx[ii] : ky[ii];            // row factor: Kx (top) / Ky (bottom)
    const zc kj = bj == 0 ? ky[jj] : kx[jj];            // column factor: Ky (left) / Kx (right)
    const zc e = Ei[ii + (long long)jj * erci.ld];
    const zc ui = Ui[ii + (long long)jj * urci.ld];
    const zc u = U[ii + (long long)jj * urc.ld];
    const zc ee = E[ii + (long long)jj * erc.ld];
    zc p = zmul(ki, zmul(e, kj));
    zc q = zmul(ki, zmul(ui, kj));
    // (0,0): +K e K ; (0,1): u - K e K ; (1,0): K e K - u ; (1,1): -K e K
    if (bi == 0 && bj == 1) { p = zsub(u, p); q = zsub(ee, q); }
    else if (bi == 1 && bj == 0) { p = zsub(p, u); q = zsub(q, ee); }
    else if (bi == 1 && bj == 1) { p = zneg(p); q = zneg(q); }
    Pb[i + (long long)j * P.ld] = p;
    Qb[i + (long long)j * Q.ld] = q;
  }
}

__global__ void modes_kernel(int n, const zc* lam2, zc* lam, zc* ilam, zc* X, const double* k0t,
                             int batch) {
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < (long long)n * batch;
       idx += (long long)gridDim.x * blockDim.x) {
    const int b = (int)(idx / n);
    const zc l = zusqrt(lam2[idx]);
    lam[idx] = l;
    ilam[idx] = zinv(l);
    X[idx] = zexp(zscale(l, -k0t[b]));
  }
}

// out = M V0 for the gap medium's V0, given by its four N-diagonals
// v0[b][(2 bi + bj) N + m] (block (bi, bj), rcwa.cpp:225): column j = bj N + m
// of the product is M(:, m) V0_{0,bj}(m) + M(:, N + m) V0_{1,bj}(m).
__global__ void mul_v0_kernel(int N, ZCMat M, const zc* v0, ZMat out) {
  const int b = blockIdx.y, n = 2 * N;
  const zc* m = M.p + (long long)b * M.stride;
  const zc* d = v0 + (long long)b * 4 * N;
  zc* o = out.p + (long long)b * out.stride;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < (long long)n * n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    const int bj = j / N, jj = j - bj * N;
    o[i + (long long)j * out.ld] = zadd(zmul(m[i + (long long)jj * M.ld], d[(0 * 2 + bj) * N + jj]),
                                        zmul(m[i + (long long)(N + jj) * M.ld],
                                             d[(1 * 2 + bj) * N + jj]));
  }
}

// A = Wi + diag(ilam) Y and B = Wi - diag(ilam) Y; A overwrites Y.
__global__ void build_ab_kernel(int n, ZCMat Wi, ZMat Y, const zc* ilam, ZMat B) {
  const int b = blockIdx.y;
  const zc* w = Wi.p + (long long)b * Wi.stride;
  zc* y = Y.p + (long long)b * Y.stride;
  zc* Bb = B.p + (long long)b * B.stride;
  const zc* il = ilam + (long long)b * n;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < (long long)n * n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    const zc t = zmul(il[i], y[i + (long long)j * Y.ld]);
    const zc wv = w[i + (long long)j * Wi.ld];
    y[i + (long long)j * Y.ld] = zadd(wv, t);
    Bb[i + (long long)j * B.ld] = zsub(wv, t);
  }
}

__global__ void expand_kernel(int N, const zc* d, long long dstride, ZMat M) {
  const int b = blockIdx.y, n = 2 * N;
  const zc* db = d + (long long)b * dstride;
  zc* Mb = M.p + (long long)b * M.stride;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < (long long)n * n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    const int bi = i / N, bj = j / N, ii = i - bi * N, jj = j - bj * N;
    Mb[i + (long long)j * M.ld] = (ii == jj) ? db[(bi * 2 + bj) * N + ii] : zmake(0, 0);
  }
}

__global__ void gather_cols_kernel(int n, ZCMat M, const int* cols, int ncol, zc* out) {
  const int b = blockIdx.y;
  const zc* Mb = M.p + (long long)b * M.stride;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n * ncol;
       idx += gridDim.x * blockDim.x) {
    const int i = idx % n, c = idx / n;
    out[((long long)b * ncol + c) * n + i] = Mb[i + (long long)cols[c] * M.ld];
  }
}

inline dim3 grid2(long long elems, int batch) {
  long long blocks = (elems + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  return dim3((unsigned)blocks, (unsigned)batch);
}

}  // namespace

size_t convmat_rowbuf_elems(int Nx, int Ny, int hx, int batch) {
  (void)Nx;
  return (size_t)batch * Ny * (4 * hx + 1);
}

namespace {

// Cell-centre sampling of the shapes of each layer (first[l] .. first[l+1]-1
// in the layer-sorted list, painter's order).  Contraction is off so that
// every inside test rounds as the plain expression does.
__global__ void rasterize_kernel(int Nx, int Ny, double Lx, double Ly, int L, const zc* bg,
                                 const rcwa_shape* shapes, const int* first, zc* grid) {
#pragma clang fp contract(off)
  const long long total = (long long)L * Ny * Nx;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(idx % Nx);
    const long long r = idx / Nx;
    const int iy = (int)(r % Ny), l = (int)(r / Ny);
    const double x = ((double)ix + 0.5) * Lx / Nx;
    const double y = ((double)iy + 0.5) * Ly / Ny;
    zc v = bg[l];
    for (int k = first[l]; k < first[l + 1]; ++k) {
      const rcwa_shape sh = shapes[k];
      bool in;
      if (sh.kind == RCWA_SHAPE_ELLIPSE) {
        const double u = (x - sh.cx) / sh.ax, w = (y - sh.cy) / sh.ay;
        in = u * u + w * w <= 1.0;
      } else if (sh.kind == RCWA_SHAPE_RECT) {
        in = fabs(x - sh.cx) <= sh.ax && fabs(y - sh.cy) <= sh.ay;
      } else {
        const double u = x - sh.cx, w = y - sh.cy;
        in = u * u + w * w <= sh.ax * sh.ax;
      }
      if (in) v = zmake(sh.value.re, sh.value.im);
    }
    grid[idx] = v;
  }
}

}  // namespace

int rasterize(hipStream_t s, int Nx, int Ny, double Lx, double Ly, int L, const zc* bg,
              const rcwa_shape* shapes, const int* first, zc* grid) {
  const long long total = (long long)L * Ny * Nx;
  const long long blocks = std::min<long long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(rasterize_kernel, dim3((unsigned)blocks), dim3(256), 0, s, Nx, Ny, Lx, Ly, L,
                     bg, shapes, first, grid);
  return (int)hipGetLastError();
}

int convmat_dft(hipStream_t s, const zc* grids, int Nx, int Ny, int hx, int hy, zc* F, int batch,
                zc* rowbuf) {
  const int nfx = 4 * hx + 1, nfy = 4 * hy + 1;
  hipLaunchKernelGGL(dft_rows_kernel, dim3(Ny, batch), dim3(DFT_THREADS),
                     2 * (size_t)Nx * sizeof(zc), s, grids, Nx, Ny, nfx, 2 * hx, rowbuf);
  hipLaunchKernelGGL(dft_cols_kernel, dim3(nfx, batch), dim3(DFT_THREADS),
                     2 * (size_t)Ny * sizeof(zc), s, rowbuf, Nx, Ny, nfx, nfy, 2 * hy, F);
  return (int)hipGetLastError();
}

int convmat_gather(hipStream_t s, const zc* F, int Nx, int Ny, int hx, int hy, ZMat C,
                   int batch, const int* map) {
  const long long PQ = (long long)(2 * hx + 1) * (2 * hy + 1);
  hipLaunchKernelGGL(gather_kernel, grid2(PQ * PQ, batch), dim3(256), 0, s, F, Nx, Ny, hx, hy, C, map);
  return (int)hipGetLastError();
}

int build_pq(hipStream_t s, int N, ZCMat erc, ZCMat urc, ZCMat erci, ZCMat urci, const zc* Kx,
             const zc* Ky, ZMat P, ZMat Q, int batch) {
  hipLaunchKernelGGL(build_pq_kernel, grid2(4LL * N * N, batch), dim3(256), 0, s, N, erc, urc,
                     erci, urci, Kx, Ky, P, Q);
  return (int)hipGetLastError();
}

int layer_modes(hipStream_t s, int n, const zc* lam2, zc* lam, zc* ilam, zc* X, const double* k0t,
                int batch) {
  hipLaunchKernelGGL(modes_kernel, grid2((long long)n * batch, 1), dim3(256), 0, s, n, lam2, lam,
                     ilam, X, k0t, batch);
  return (int)hipGetLastError();
}

int mul_v0(hipStream_t s, int N, ZCMat M, const zc* v0, ZMat out, int batch) {
  hipLaunchKernelGGL(mul_v0_kernel, grid2(4LL * N * N, batch), dim3(256), 0, s, N, M, v0, out);
  return (int)hipGetLastError();
}

int build_ab(hipStream_t s, int n, ZCMat Wi, ZMat Y, const zc* ilam, ZMat B, int batch) {
  hipLaunchKernelGGL(build_ab_kernel, grid2((long long)n * n, batch), dim3(256), 0, s, n, Wi, Y,
                     ilam, B);
  return (int)hipGetLastError();
}

int expand_block_diag(hipStream_t s, int N, const zc* d, long long dstride, ZMat M, int batch) {
  hipLaunchKernelGGL(expand_kernel, grid2(4LL * N * N, batch), dim3(256), 0, s, N, d, dstride, M);
  return (int)hipGetLastError();
}

int gather_cols(hipStream_t s, int n, ZCMat M, const int* cols, int ncol, zc* out, int batch) {
  hipLaunchKernelGGL(gather_cols_kernel, grid2((long long)n * ncol, batch), dim3(256), 0, s, n, M,
                     cols, ncol, out);
  return (int)hipGetLastError();
}

namespace {

// Mask grids of (layer, material) pairs: out[k][iy][ix] = 1 where
// material[layer[k]][iy][ix] == mat[k], else 0.
__global__ void material_mask_kernel(long long cells, const unsigned char* material,
                                     const int* layer, const int* mat, zc* out) {
  const int k = blockIdx.y;
  const unsigned char* m = material + (long long)layer[k] * cells;
  const unsigned char want = (unsigned char)mat[k];
  zc* o = out + (long long)k * cells;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < cells;
       idx += (long long)gridDim.x * blockDim.x)
    o[idx] = zmake(m[idx] == want ? 1.0 : 0.0, 0.0);
}

// out[map[g * L + l]] = sum over the pairs k = first[l] .. first[l+1]-1 of
// coef[g * M + mat[k]] * C[k]: blockIdx.y = layer, blockIdx.z = structure.
__global__ void mix_kernel(long long N2, int L, int M, const int* first, const int* mat,
                           const zc* C, const zc* coef, const int* map, zc* out) {
  const int l = blockIdx.y, g = blockIdx.z;
  const int k0 = first[l], k1 = first[l + 1];
  const zc* cf = coef + (long long)g * M;
  zc* o = out + (long long)map[(long long)g * L + l] * N2;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < N2;
       idx += (long long)gridDim.x * blockDim.x) {
    zc acc = zmake(0, 0);
    for (int k = k0; k < k1; ++k) acc = zfma(cf[mat[k]], C[(long long)k * N2 + idx], acc);
    o[idx] = acc;
  }
}

__global__ void gather_mats_kernel(long long elems, const zc* src, const int* smap, zc* dst,
                                   const int* dmap) {
  const int k = blockIdx.y;
  const zc* s = src + (long long)smap[k] * elems;
  zc* d = dst + (long long)dmap[k] * elems;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < elems;
       idx += (long long)gridDim.x * blockDim.x)
    d[idx] = s[idx];
}

}  // namespace

int material_masks(hipStream_t s, int Nx, int Ny, const unsigned char* material,
                   const int* layer, const int* mat, int npairs, zc* out) {
  if (npairs <= 0) return 0;
  hipLaunchKernelGGL(material_mask_kernel, grid2((long long)Nx * Ny, npairs), dim3(256), 0, s,
                     (long long)Nx * Ny, material, layer, mat, out);
  return (int)hipGetLastError();
}

int mix_convmat(hipStream_t s, int N, int L, int M, const int* first, const int* mat,
                const zc* C, const zc* coef, const int* map, zc* out, int G) {
  if (L <= 0 || G <= 0) return 0;
  if (L > 65535 || G > 65535) return (int)hipErrorInvalidValue;
  const long long N2 = (long long)N * N;
  dim3 grid = grid2(N2, L);
  grid.z = (unsigned)G;
  hipLaunchKernelGGL(mix_kernel, grid, dim3(256), 0, s, N2, L, M, first, mat, C, coef, map,
                     out);
  return (int)hipGetLastError();
}

int gather_mats(hipStream_t s, long long elems, const zc* src, const int* smap, zc* dst,
                const int* dmap, int count) {
  // at most 65535 matrices per launch (the grid's y extent)
  for (int c0 = 0; c0 < count; c0 += 65535) {
    hipLaunchKernelGGL(gather_mats_kernel, grid2(elems, std::min(65535, count - c0)), dim3(256),
                       0, s, elems, src, smap + c0, dst, dmap + c0);
    if (int e = (int)hipGetLastError()) return e;
  }
  return 0;
}
