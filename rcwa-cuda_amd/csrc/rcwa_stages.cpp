// Stage entry points of the C ABI: one per function of rcwa3d/include/rcwa.hpp
// (host buffers in and out), plus the HBM-resident batched ConvMat.  Each
// O(N^3) stage runs on the GPU through the same kernels as the full solve;
// the O(N) stages are the host functions of host_math.cpp.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "rcwa_amd.h"
#include "common.hpp"
#include "kernels.hpp"
#include "host_math.hpp"
#include "runtime.hpp"

using rcwa_host::cplx;

#define CK(expr)                                                                   \
  do {                                                                             \
    int _e = (int)(expr);                                                          \
    if (_e) {                                                                      \
      if (_e < 100) return (rcwa_status)_e;                                        \
      return (rcwa_status)rcwa_internal_hip_fail((hipError_t)_e, #expr, __FILE__, __LINE__); \
    }                                                                              \
  } while (0)
#define CKH(expr)                                                                  \
  do {                                                                             \
    hipError_t _h = (expr);                                                        \
    if (_h != hipSuccess)                                                          \
      return (rcwa_status)rcwa_internal_hip_fail(_h, #expr, __FILE__, __LINE__);   \
  } while (0)

namespace {

Ctx* C(rcwa_ctx* c) { return reinterpret_cast<Ctx*>(c); }

std::vector<cplx> vec(const rcwa_c128* p, size_t n) {
  std::vector<cplx> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = cplx(p[i].re, p[i].im);
  return v;
}

rcwa_status h2d(Ctx& c, void* d, const void* h, size_t bytes) {
  CKH(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c.stream));
  return RCWA_OK;
}
rcwa_status d2h(Ctx& c, void* h, const void* d, size_t bytes) {
  CKH(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c.stream));
  CKH(hipStreamSynchronize(c.stream));
  return RCWA_OK;
}

rcwa_status check_info(Ctx& c, int* dinfo, int count) {
  std::vector<int> h(count);
  rcwa_status st = d2h(c, h.data(), dinfo, count * sizeof(int));
  if (st) return st;
  for (int i = 0; i < count; ++i)
    if (h[i]) return rcwa_set_error(RCWA_E_SINGULAR, "exactly singular matrix inverted");
  return RCWA_OK;
}

}  // namespace

// S = A * B for device blocks (the work of rcwa_star_product_batched and its
// allocation-free form): A is copied into S, the product forms S in place.
// t[5] are batch x n x n work slots, gj the Gauss-Jordan workspace, info
// batch ints already cleared (set where an inverse met an exactly zero
// pivot).  Enqueued on the context stream; the caller reads info.
rcwa_status star_product_device(Ctx& c, int n, int batch, const rcwa_c128* const* A_dev,
                                const rcwa_c128* const* B_dev, rcwa_c128* const* S_dev, zc* t[5],
                                void* gj, int* dinfo) {
  const long long n2 = (long long)n * n;
  for (int k = 0; k < 4; ++k)
    for (int q = 0; q < 4; ++q)
      if ((const void*)S_dev[k] == (const void*)B_dev[q])
        return rcwa_set_error(RCWA_E_ARG, "S must not alias B");
  zc* a[4];
  ZCMat Bm[4];
  for (int k = 0; k < 4; ++k) {
    a[k] = (zc*)S_dev[k];
    Bm[k] = ZCMat{(const zc*)B_dev[k], n, n2};
    if ((const void*)A_dev[k] != (const void*)S_dev[k])
      CK(zcopy_batched(c.stream, n, n, ZCMat{(const zc*)A_dev[k], n, n2}, ZMat{a[k], n, n2},
                       batch));
  }
  return star_product(c, n, batch, a, Bm, t, gj, dinfo, true);
}

extern "C" {

rcwa_status rcwa_convmat(rcwa_ctx* ctx, const rcwa_c128* field, int layer, int Nx, int Ny,
                         int hx, int hy, rcwa_c128* out) {
  if (!ctx || !field || !out || Nx <= 0 || Ny <= 0 || hx < 0 || hy < 0 || layer < 0)
    return rcwa_set_error(RCWA_E_ARG, "bad argument");
  Ctx& c = *C(ctx);
  CKH(hipSetDevice(c.device));
  const int PQ = (2 * hx + 1) * (2 * hy + 1);
  const size_t G = (size_t)Nx * Ny, nf = (size_t)(4 * hx + 1) * (4 * hy + 1);
  rcwa_status st = c.reserve((G + Ny * (4 * hx + 1) + nf + (size_t)PQ * PQ) * sizeof(zc) + 4096);
  if (st) return st;
  zc* g = (zc*)c.take(G * sizeof(zc));
  zc* rb = (zc*)c.take((size_t)Ny * (4 * hx + 1) * sizeof(zc));
  zc* F = (zc*)c.take(nf * sizeof(zc));
  zc* Cm = (zc*)c.take((size_t)PQ * PQ * sizeof(zc));
  if ((st = c.arena_check()) != RCWA_OK) return st;
  if ((st = h2d(c, g, field + (size_t)layer * G, G * sizeof(zc)))) return st;
  CK(convmat_dft(c.stream, g, Nx, Ny, hx, hy, F, 1, rb));
  CK(convmat_gather(c.stream, F, Nx, Ny, hx, hy, ZMat{Cm, PQ, (long long)PQ * PQ}, 1));
  return d2h(c, out, Cm, (size_t)PQ * PQ * sizeof(zc));
}

rcwa_status rcwa_convmat_batched(rcwa_ctx* ctx, const rcwa_c128* grids_dev, int Nx, int Ny,
                                 int hx, int hy, int batch, rcwa_c128* C_dev) {
  if (!ctx || !grids_dev || !C_dev || Nx <= 0 || Ny <= 0 || hx < 0 || hy < 0 || batch <= 0)
    return rcwa_set_error(RCWA_E_ARG, "bad argument");
  Ctx& c = *C(ctx);
  CKH(hipSetDevice(c.device));
  const int PQ = (2 * hx + 1) * (2 * hy + 1);
  const size_t nf = (size_t)(4 * hx + 1) * (4 * hy + 1);
  rcwa_status st =
      c.reserve(((size_t)batch * (Ny * (4 * hx + 1) + nf)) * sizeof(zc) + 4096);
  if (st) return st;
  zc* rb = (zc*)c.take((size_t)batch * Ny * (4 * hx + 1) * sizeof(zc));
  zc* F = (zc*)c.take((size_t)batch * nf * sizeof(zc));
  if ((st = c.arena_check()) != RCWA_OK) return st;
  CK(convmat_dft(c.stream, (const zc*)grids_dev, Nx, Ny, hx, hy, F, batch, rb));
  CK(convmat_gather(c.stream, F, Nx, Ny, hx, hy, ZMat{(zc*)C_dev, PQ, (long long)PQ * PQ}, batch));
  CKH(hipStreamSynchronize(c.stream));
  return RCWA_OK;
}

rcwa_status rcwa_rasterize(rcwa_ctx* ctx, int Nx, int Ny, double Lx, double Ly, int num_layers,
                           const rcwa_c128* background, const rcwa_shape* shapes, int nshapes,
                           rcwa_c128* grid_dev) {
  if (!ctx || !background || !grid_dev || Nx <= 0 || Ny <= 0 || num_layers <= 0 || nshapes < 0 ||
      (nshapes > 0 && !shapes) || !(Lx > 0) || !(Ly > 0))
    return rcwa_set_error(RCWA_E_ARG, "bad argument");
  for (int k = 0; k < nshapes; ++k) {
    const rcwa_shape& sh = shapes[k];
    if (sh.layer < 0 || sh.layer >= num_layers || sh.kind < RCWA_SHAPE_ELLIPSE ||
        sh.kind > RCWA_SHAPE_CIRCLE)
      return rcwa_set_error(RCWA_E_ARG, "shape with a bad layer or kind");
  }
  Ctx& c = *C(ctx);
  CKH(hipSetDevice(c.device));
  // layer-sorted copy (stable: painter's order within a layer kept)
  std::vector<rcwa_shape> sorted(shapes, shapes + nshapes);
  std::stable_sort(sorted.begin(), sorted.end(),
                   [](const rcwa_shape& a, const rcwa_shape& b) { return a.layer < b.layer; });
  std::vector<int> first(num_layers + 1, 0);
  for (const rcwa_shape& sh : sorted) ++first[sh.layer + 1];
  for (int l = 0; l < num_layers; ++l) first[l + 1] += first[l];
  const size_t nb = (size_t)num_layers * sizeof(zc), ns = sorted.size() * sizeof(rcwa_shape),
               nf = first.size() * sizeof(int);
  rcwa_status st = c.reserve(nb + ns + nf + 3 * 256);
  if (st) return st;
  zc* bg = (zc*)c.take(nb);
  rcwa_shape* sd = (rcwa_shape*)c.take(ns > 0 ? ns : 1);
  int* fd = (int*)c.take(nf);
  if ((st = c.arena_check()) != RCWA_OK) return st;
  CKH(hipMemcpyAsync(bg, background, nb, hipMemcpyHostToDevice, c.stream));
  if (ns) CKH(hipMemcpyAsync(sd, sorted.data(), ns, hipMemcpyHostToDevice, c.stream));
  CKH(hipMemcpyAsync(fd, first.data(), nf, hipMemcpyHostToDevice, c.stream));
  CK(rasterize(c.stream, Nx, Ny, Lx, Ly, num_layers, bg, sd, fd, (zc*)grid_dev));
  CKH(hipStreamSynchronize(c.stream));
  return RCWA_OK;
}

rcwa_status rcwa_wave_vectors(double Lx, double Ly, const rcwa_source* src,
                              const rcwa_params* p, rcwa_c128* k_inc, rcwa_c128* Kx,
                              rcwa_c128* Ky, rcwa_c128* Kz_ref, rcwa_c128* Kz_trn) {
  if (!src || !p || !k_inc || !Kx || !Ky || !Kz_ref || !Kz_trn)
    return rcwa_set_error(RCWA_E_ARG, "null argument");
  rcwa_host::Media m{cplx(p->er_ref.re, p->er_ref.im), cplx(p->ur_ref.re, p->ur_ref.im),
                     cplx(p->er_trn.re, p->er_trn.im), cplx(p->ur_trn.re, p->ur_trn.im)};
  rcwa_host::WaveVectors w = rcwa_host::wave_vectors(
      Lx, Ly, {src->lambda0, src->theta, src->phi}, p->Nx_harmonics, p->Ny_harmonics, m);
  std::memcpy((void*)k_inc, w.k_inc, 3 * sizeof(cplx));
  const size_t N = w.Kx.size();
  std::memcpy((void*)Kx, w.Kx.data(), N * sizeof(cplx));
  std::memcpy((void*)Ky, w.Ky.data(), N * sizeof(cplx));
  std::memcpy((void*)Kz_ref, w.Kz_ref.data(), N * sizeof(cplx));
  std::memcpy((void*)Kz_trn, w.Kz_trn.data(), N * sizeof(cplx));
  return RCWA_OK;
}

rcwa_status rcwa_gap_medium(int PQ, const rcwa_c128* Kx, const rcwa_c128* Ky, rcwa_c128* W0,
                            rcwa_c128* V0) {
  if (PQ <= 0 || !Kx || !Ky || !W0 || !V0) return rcwa_set_error(RCWA_E_ARG, "bad argument");
  std::vector<cplx> d = rcwa_host::gap_v0_diagonals(vec(Kx, PQ), vec(Ky, PQ));
  const size_t n = 2 * (size_t)PQ;
  std::memset((void*)W0, 0, n * n * sizeof(rcwa_c128));
  std::memset((void*)V0, 0, n * n * sizeof(rcwa_c128));
  for (size_t i = 0; i < n; ++i) W0[i + i * n] = {1.0, 0.0};
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj)
      for (int m = 0; m < PQ; ++m) {
        const cplx v = d[(size_t)(2 * bi + bj) * PQ + m];
        V0[(bi * PQ + m) + (bj * PQ + m) * n] = {v.real(), v.imag()};
      }
  return RCWA_OK;
}

rcwa_status rcwa_end_smatrix(int which, int PQ, const rcwa_params* p, const rcwa_c128* Kx,
                             const rcwa_c128* Ky, const rcwa_c128* Kz, rcwa_c128* S11,
                             rcwa_c128* S12, rcwa_c128* S21, rcwa_c128* S22) {
  if ((which != 0 && which != 1) || PQ <= 0 || !p || !Kx || !Ky || !Kz)
    return rcwa_set_error(RCWA_E_ARG, "bad argument");
  const cplx er = which == 0 ? cplx(p->er_ref.re, p->er_ref.im) : cplx(p->er_trn.re, p->er_trn.im);
  const cplx ur = which == 0 ? cplx(p->ur_ref.re, p->ur_ref.im) : cplx(p->ur_trn.re, p->ur_trn.im);
  std::vector<cplx> d =
      rcwa_host::end_smatrix_diagonals(which, vec(Kx, PQ), vec(Ky, PQ), vec(Kz, PQ), er, ur);
  rcwa_c128* outs[4] = {S11, S12, S21, S22};
  const size_t n = 2 * (size_t)PQ;
  for (int k = 0; k < 4; ++k) {
    if (!outs[k]) continue;
    std::memset((void*)outs[k], 0, n * n * sizeof(rcwa_c128));
    for (int bi = 0; bi < 2; ++bi)
      for (int bj = 0; bj < 2; ++bj)
        for (int m = 0; m < PQ; ++m) {
          const cplx v = d[(size_t)(4 * k + 2 * bi + bj) * PQ + m];
          outs[k][(bi * PQ + m) + (bj * PQ + m) * n] = {v.real(), v.imag()};
        }
  }
  return RCWA_OK;
}

rcwa_status rcwa_layer_smatrix(rcwa_ctx* ctx, int PQ, int batch, const rcwa_c128* erc,
                               const rcwa_c128* urc, const rcwa_c128* Kx, const rcwa_c128* Ky,
                               const double* k0t, rcwa_c128* S11, rcwa_c128* S12) {
  if (!ctx || PQ <= 0 || batch <= 0 || !erc || !urc || !Kx || !Ky || !k0t || !S11 || !S12)
    return rcwa_set_error(RCWA_E_ARG, "bad argument");
  Ctx& c = *C(ctx);
  CKH(hipSetDevice(c.device));
  c.precision = RCWA_FP64;
  const int N = PQ, n = 2 * N;
  const size_t n2 = (size_t)n * n, N2 = (size_t)N * N;
  const size_t bytes = (2 * batch * N2 + 2 * batch * n2 + batch * (6 * N + 4 * n)) * sizeof(zc) +
                       batch * (sizeof(double) + 2 * sizeof(int)) + layer_work_bytes(N, batch) +
                       64 * 256;
  rcwa_status st = c.reserve(bytes);
  if (st) return st;
  zc* derc = (zc*)c.take(batch * N2 * sizeof(zc));
  zc* durc = (zc*)c.take(batch * N2 * sizeof(zc));
  zc* dS11 = (zc*)c.take(batch * n2 * sizeof(zc));
  zc* dS12 = (zc*)c.take(batch * n2 * sizeof(zc));
  zc* dKx = (zc*)c.take((size_t)batch * N * sizeof(zc));
  zc* dKy = (zc*)c.take((size_t)batch * N * sizeof(zc));
  zc* dv0 = (zc*)c.take((size_t)batch * 4 * N * sizeof(zc));
  zc* lam2 = (zc*)c.take((size_t)batch * n * sizeof(zc));
  zc* lam = (zc*)c.take((size_t)batch * n * sizeof(zc));
  zc* ilam = (zc*)c.take((size_t)batch * n * sizeof(zc));
  zc* X = (zc*)c.take((size_t)batch * n * sizeof(zc));
  double* dk = (double*)c.take(batch * sizeof(double));
  int* dinfo = (int*)c.take(2 * batch * sizeof(int));
  LayerWork w;
  w.carve(c, N, batch);
  if ((st = c.arena_check()) != RCWA_OK) return st;
  std::vector<cplx> kx = vec(Kx, N), ky = vec(Ky, N);
  std::vector<cplx> v0 = rcwa_host::gap_v0_diagonals(kx, ky);
  std::vector<cplx> hKx((size_t)batch * N), hKy((size_t)batch * N), hv0((size_t)batch * 4 * N);
  for (int b = 0; b < batch; ++b) {
    std::copy(kx.begin(), kx.end(), hKx.begin() + (size_t)b * N);
    std::copy(ky.begin(), ky.end(), hKy.begin() + (size_t)b * N);
    std::copy(v0.begin(), v0.end(), hv0.begin() + (size_t)b * 4 * N);
  }
  if ((st = h2d(c, derc, erc, batch * N2 * sizeof(zc)))) return st;
  if ((st = h2d(c, durc, urc, batch * N2 * sizeof(zc)))) return st;
  if ((st = h2d(c, dKx, hKx.data(), hKx.size() * sizeof(zc)))) return st;
  if ((st = h2d(c, dKy, hKy.data(), hKy.size() * sizeof(zc)))) return st;
  if ((st = h2d(c, dv0, hv0.data(), hv0.size() * sizeof(zc)))) return st;
  if ((st = h2d(c, dk, k0t, batch * sizeof(double)))) return st;
  CKH(hipMemsetAsync(dinfo, 0, 2 * batch * sizeof(int), c.stream));
  st = layer_stage(c, N, batch, ZCMat{derc, N, (long long)N2}, ZCMat{durc, N, (long long)N2}, dKx,
                   dKy, dv0, dk, lam2, lam, ilam, X, ZMat{dS11, n, (long long)n2},
                   ZMat{dS12, n, (long long)n2}, w, dinfo, dinfo + batch);
  if (st) return st;
  std::vector<int> hinfo(2 * batch);
  if ((st = d2h(c, hinfo.data(), dinfo, 2 * batch * sizeof(int)))) return st;
  for (int b = 0; b < batch; ++b) {
    if (hinfo[b]) return rcwa_set_error(RCWA_E_EIG_NOCONV, "eigensolver QR did not converge");
    if (hinfo[batch + b]) return rcwa_set_error(RCWA_E_SINGULAR, "exactly singular matrix inverted");
  }
  if ((st = d2h(c, S11, dS11, batch * n2 * sizeof(zc)))) return st;
  return d2h(c, S12, dS12, batch * n2 * sizeof(zc));
}

rcwa_status rcwa_star_product(rcwa_ctx* ctx, int n, int batch, const rcwa_c128* const* A,
                              const rcwa_c128* const* B, rcwa_c128* const* S) {
  if (!ctx || n <= 0 || batch <= 0 || !A || !B || !S)
    return rcwa_set_error(RCWA_E_ARG, "bad argument");
  Ctx& c = *C(ctx);
  CKH(hipSetDevice(c.device));
  c.precision = RCWA_FP64;
  const size_t n2 = (size_t)n * n, blk = batch * n2 * sizeof(zc);
  rcwa_status st = c.reserve(13 * blk + gj_workspace_bytes(n, batch) + batch * sizeof(int) + 64 * 256);
  if (st) return st;
  zc* a[4];
  zc* b[4];
  zc* t[5];
  for (auto& p : a) p = (zc*)c.take(blk);
  for (auto& p : b) p = (zc*)c.take(blk);
  for (auto& p : t) p = (zc*)c.take(blk);
  void* gj = c.take(gj_workspace_bytes(n, batch));
  int* dinfo = (int*)c.take(batch * sizeof(int));
  if ((st = c.arena_check()) != RCWA_OK) return st;
  for (int k = 0; k < 4; ++k) {
    if (!A[k] || !B[k] || !S[k]) return rcwa_set_error(RCWA_E_ARG, "null block");
    if ((st = h2d(c, a[k], A[k], blk))) return st;
    if ((st = h2d(c, b[k], B[k], blk))) return st;
  }
  CKH(hipMemsetAsync(dinfo, 0, batch * sizeof(int), c.stream));
  const ZCMat Bm[4] = {{b[0], n, (long long)n2}, {b[1], n, (long long)n2},
                       {b[2], n, (long long)n2}, {b[3], n, (long long)n2}};
  if ((st = star_product(c, n, batch, a, Bm, t, gj, dinfo, false))) return st;
  if ((st = check_info(c, dinfo, batch))) return st;
  for (int k = 0; k < 4; ++k)
    if ((st = d2h(c, S[k], a[k], blk))) return st;
  return RCWA_OK;
}

rcwa_status rcwa_layer_smatrix_batched(rcwa_ctx* ctx, int PQ, int batch, const rcwa_c128* erc_dev,
                                       const rcwa_c128* urc_dev, const rcwa_c128* Kx,
                                       const rcwa_c128* Ky, const double* k0t,
                                       rcwa_c128* S11_dev, rcwa_c128* S12_dev) {
  if (!ctx || PQ <= 0 || batch <= 0 || !erc_dev || !urc_dev || !Kx || !Ky || !k0t || !S11_dev ||
      !S12_dev)
    return rcwa_set_error(RCWA_E_ARG, "bad argument");
  Ctx& c = *C(ctx);
  CKH(hipSetDevice(c.device));
  c.precision = RCWA_FP64;
  const int N = PQ, n = 2 * N;
  const size_t N2 = (size_t)N * N;
  for (int b = 0; b < batch; ++b)
    if (!std::isfinite(k0t[b]) || k0t[b] < 0)
      return rcwa_set_error(RCWA_E_ARG, "k0 t must be finite and >= 0");
  const size_t bytes = (batch * (6 * (size_t)N + 4 * (size_t)n)) * sizeof(zc) +
                       batch * (sizeof(double) + 2 * sizeof(int)) + layer_work_bytes(N, batch) +
                       64 * 256;
  rcwa_status st = c.reserve(bytes);
  if (st) return st;
  zc* dKx = (zc*)c.take((size_t)batch * N * sizeof(zc));
  zc* dKy = (zc*)c.take((size_t)batch * N * sizeof(zc));
  zc* dv0 = (zc*)c.take((size_t)batch * 4 * N * sizeof(zc));
  zc* lam2 = (zc*)c.take((size_t)batch * n * sizeof(zc));
  zc* lam = (zc*)c.take((size_t)batch * n * sizeof(zc));
  zc* ilam = (zc*)c.take((size_t)batch * n * sizeof(zc));
  zc* X = (zc*)c.take((size_t)batch * n * sizeof(zc));
  double* dk = (double*)c.take(batch * sizeof(double));
  int* dinfo = (int*)c.take(2 * batch * sizeof(int));
  LayerWork w;
  w.carve(c, N, batch);
  if ((st = c.arena_check()) != RCWA_OK) return st;
  std::vector<cplx> hv0((size_t)batch * 4 * N);
  for (int b = 0; b < batch; ++b) {
    std::vector<cplx> v0 =
        rcwa_host::gap_v0_diagonals(vec(Kx + (size_t)b * N, N), vec(Ky + (size_t)b * N, N));
    std::copy(v0.begin(), v0.end(), hv0.begin() + (size_t)b * 4 * N);
  }
  for (const cplx& z : hv0)
    if (!std::isfinite(z.real()) || !std::isfinite(z.imag()))
      return rcwa_set_error(RCWA_E_SINGULAR, "k_z = 0 in the gap medium (Wood anomaly)");
  if ((st = h2d(c, dKx, Kx, (size_t)batch * N * sizeof(zc)))) return st;
  if ((st = h2d(c, dKy, Ky, (size_t)batch * N * sizeof(zc)))) return st;
  if ((st = h2d(c, dv0, hv0.data(), hv0.size() * sizeof(zc)))) return st;
  if ((st = h2d(c, dk, k0t, batch * sizeof(double)))) return st;
  CKH(hipMemsetAsync(dinfo, 0, 2 * batch * sizeof(int), c.stream));
  const long long n2 = (long long)n * n;
  st = layer_stage(c, N, batch, ZCMat{(const zc*)erc_dev, N, (long long)N2},
                   ZCMat{(const zc*)urc_dev, N, (long long)N2}, dKx, dKy, dv0, dk, lam2, lam, ilam,
                   X, ZMat{(zc*)S11_dev, n, n2}, ZMat{(zc*)S12_dev, n, n2}, w, dinfo,
                   dinfo + batch);
  if (st) return st;
  std::vector<int> hinfo(2 * batch);
  if ((st = d2h(c, hinfo.data(), dinfo, 2 * batch * sizeof(int)))) return st;
  for (int b = 0; b < batch; ++b) {
    if (hinfo[b]) return rcwa_set_error(RCWA_E_EIG_NOCONV, "eigensolver QR did not converge");
    if (hinfo[batch + b]) return rcwa_set_error(RCWA_E_SINGULAR, "exactly singular matrix inverted");
  }
  return RCWA_OK;
}

rcwa_status rcwa_star_product_batched(rcwa_ctx* ctx, int n, int batch,
                                      const rcwa_c128* const* A_dev,
                                      const rcwa_c128* const* B_dev, rcwa_c128* const* S_dev) {
  if (!ctx || n <= 0 || batch <= 0 || !A_dev || !B_dev || !S_dev)
    return rcwa_set_error(RCWA_E_ARG, "bad argument");
  for (int k = 0; k < 4; ++k)
    if (!A_dev[k] || !B_dev[k] || !S_dev[k]) return rcwa_set_error(RCWA_E_ARG, "null block");
  Ctx& c = *C(ctx);
  CKH(hipSetDevice(c.device));
  c.precision = RCWA_FP64;
  const size_t n2 = (size_t)n * n, blk = batch * n2 * sizeof(zc);
  rcwa_status st = c.reserve(5 * blk + gj_workspace_bytes(n, batch) + batch * sizeof(int) +
                             64 * 256);
  if (st) return st;
  zc* t[5];
  for (auto& p : t) p = (zc*)c.take(blk);
  void* gj = c.take(gj_workspace_bytes(n, batch));
  int* dinfo = (int*)c.take(batch * sizeof(int));
  if ((st = c.arena_check()) != RCWA_OK) return st;
  CKH(hipMemsetAsync(dinfo, 0, batch * sizeof(int), c.stream));
  if ((st = star_product_device(c, n, batch, A_dev, B_dev, S_dev, t, gj, dinfo))) return st;
  return check_info(c, dinfo, batch);
}

rcwa_status rcwa_eig(rcwa_ctx* ctx, int n, int batch, const rcwa_c128* A, rcwa_c128* w,
                     rcwa_c128* V) {
  if (!ctx || n <= 0 || batch <= 0 || !A || !w || !V)
    return rcwa_set_error(RCWA_E_ARG, "bad argument");
  Ctx& c = *C(ctx);
  CKH(hipSetDevice(c.device));
  const size_t n2 = (size_t)n * n, blk = batch * n2 * sizeof(zc);
  rcwa_status st = c.reserve(2 * blk + eig_workspace_bytes(n, batch) +
                             batch * (n * sizeof(zc) + sizeof(int)) + 64 * 256);
  if (st) return st;
  zc* dA = (zc*)c.take(blk);
  zc* dV = (zc*)c.take(blk);
  zc* dw = (zc*)c.take(batch * n * sizeof(zc));
  void* ws = c.take(eig_workspace_bytes(n, batch));
  int* dinfo = (int*)c.take(batch * sizeof(int));
  if ((st = c.arena_check()) != RCWA_OK) return st;
  if ((st = h2d(c, dA, A, blk))) return st;
  CKH(hipMemsetAsync(dinfo, 0, batch * sizeof(int), c.stream));
  CK(eig_batched(c.stream, n, ZMat{dA, n, (long long)n2}, dw, ZMat{dV, n, (long long)n2}, batch,
                 ws, dinfo, c.eig_maxit, nullptr, c.qr_log));
  std::vector<int> h(batch);
  if ((st = d2h(c, h.data(), dinfo, batch * sizeof(int)))) return st;
  for (int b = 0; b < batch; ++b)
    if (h[b]) return rcwa_set_error(RCWA_E_EIG_NOCONV, "eigensolver QR did not converge");
  if ((st = d2h(c, w, dw, batch * n * sizeof(zc)))) return st;
  return d2h(c, V, dV, blk);
}

rcwa_status rcwa_inverse(rcwa_ctx* ctx, int n, int batch, const rcwa_c128* A, rcwa_c128* Ainv) {
  if (!ctx || n <= 0 || batch <= 0 || !A || !Ainv) return rcwa_set_error(RCWA_E_ARG, "bad argument");
  Ctx& c = *C(ctx);
  CKH(hipSetDevice(c.device));
  const size_t n2 = (size_t)n * n, blk = batch * n2 * sizeof(zc);
  rcwa_status st = c.reserve(blk + gj_workspace_bytes(n, batch) + batch * sizeof(int) + 4096);
  if (st) return st;
  zc* dA = (zc*)c.take(blk);
  void* ws = c.take(gj_workspace_bytes(n, batch));
  int* dinfo = (int*)c.take(batch * sizeof(int));
  if ((st = c.arena_check()) != RCWA_OK) return st;
  if ((st = h2d(c, dA, A, blk))) return st;
  CKH(hipMemsetAsync(dinfo, 0, batch * sizeof(int), c.stream));
  CK(gj_inverse_batched(c.stream, n, ZMat{dA, n, (long long)n2}, batch, ws, dinfo));
  if ((st = check_info(c, dinfo, batch))) return st;
  return d2h(c, Ainv, dA, blk);
}

static rcwa_status zgemm_impl(rcwa_ctx* ctx, int m, int n, int k, int batch, const rcwa_c128* A,
                              const rcwa_c128* B, rcwa_c128* Cm, rcwa_precision prec) {
  if (!ctx || m <= 0 || n <= 0 || k <= 0 || batch <= 0 || !A || !B || !Cm)
    return rcwa_set_error(RCWA_E_ARG, "bad argument");
  Ctx& c = *C(ctx);
  CKH(hipSetDevice(c.device));
  const size_t sa = (size_t)m * k, sb = (size_t)k * n, sc = (size_t)m * n;
  rcwa_status st = c.reserve(batch * (sa + sb + sc) * sizeof(zc) + 4096);
  if (st) return st;
  zc* dA = (zc*)c.take(batch * sa * sizeof(zc));
  zc* dB = (zc*)c.take(batch * sb * sizeof(zc));
  zc* dC = (zc*)c.take(batch * sc * sizeof(zc));
  if ((st = c.arena_check()) != RCWA_OK) return st;
  if ((st = h2d(c, dA, A, batch * sa * sizeof(zc)))) return st;
  if ((st = h2d(c, dB, B, batch * sb * sizeof(zc)))) return st;
  c.precision = prec;
  CK(c.gemm(m, n, k, zmake(1, 0), ZCMat{dA, m, (long long)sa}, ZCMat{dB, k, (long long)sb},
            zmake(0, 0), ZCMat{nullptr, 0, 0}, ZMat{dC, m, (long long)sc}, batch));
  c.precision = RCWA_FP64;
  return d2h(c, Cm, dC, batch * sc * sizeof(zc));
}

rcwa_status rcwa_zgemm(rcwa_ctx* ctx, int m, int n, int k, int batch, const rcwa_c128* A,
                       const rcwa_c128* B, rcwa_c128* Cm) {
  return zgemm_impl(ctx, m, n, k, batch, A, B, Cm, RCWA_FP64);
}

rcwa_status rcwa_zgemm_split(rcwa_ctx* ctx, int m, int n, int k, int batch, const rcwa_c128* A,
                             const rcwa_c128* B, rcwa_c128* Cm) {
  return zgemm_impl(ctx, m, n, k, batch, A, B, Cm, RCWA_SPLIT);
}

rcwa_status rcwa_efficiencies(const rcwa_source* src, const rcwa_params* p, const rcwa_c128* k_inc,
                              const rcwa_c128* Kx, const rcwa_c128* Ky, const rcwa_c128* Kz_ref,
                              const rcwa_c128* Kz_trn, const rcwa_c128* s11c,
                              const rcwa_c128* s21c, rcwa_results* out) {
  if (!src || !p || !k_inc || !Kx || !Ky || !Kz_ref || !Kz_trn || !s11c || !s21c || !out)
    return rcwa_set_error(RCWA_E_ARG, "null argument");
  const int N = (2 * p->Nx_harmonics + 1) * (2 * p->Ny_harmonics + 1);
  rcwa_host::WaveVectors w;
  for (int i = 0; i < 3; ++i) w.k_inc[i] = cplx(k_inc[i].re, k_inc[i].im);
  w.Kx = vec(Kx, N);
  w.Ky = vec(Ky, N);
  w.Kz_ref = vec(Kz_ref, N);
  w.Kz_trn = vec(Kz_trn, N);
  rcwa_host::Media m{cplx(p->er_ref.re, p->er_ref.im), cplx(p->ur_ref.re, p->ur_ref.im),
                     cplx(p->er_trn.re, p->er_trn.im), cplx(p->ur_trn.re, p->ur_trn.im)};
  std::vector<cplx> a = vec(s11c, 4 * (size_t)N), b = vec(s21c, 4 * (size_t)N);
  rcwa_host::Efficiencies e =
      rcwa_host::efficiencies(src->theta, src->pte, src->ptm, w, m, a.data(), b.data());
  if (out->R) std::memcpy(out->R, e.R.data(), N * sizeof(double));
  if (out->T) std::memcpy(out->T, e.T.data(), N * sizeof(double));
  if (out->r) std::memcpy(out->r, e.r.data(), 3 * (size_t)N * sizeof(rcwa_c128));
  if (out->t) std::memcpy(out->t, e.t.data(), 3 * (size_t)N * sizeof(rcwa_c128));
  out->R_tot = e.R_tot;
  out->T_tot = e.T_tot;
  out->cond = 0;  // no layer matrices at this stage
  return RCWA_OK;
}

}  // extern "C"
