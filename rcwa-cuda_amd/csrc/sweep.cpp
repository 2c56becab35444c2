// Sweep plans (rcwa_sweep_*): a structure given as material masks, whose
// convolution matrices are computed once on the device; each sweep point
// then mixes them with its material values instead of building and
// transforming grids (main.cpp:126-138 does ConvMat of every layer per
// call).  The points run through the batched solve (solve_sweep).
#include <algorithm>
#include <cstring>
#include <vector>

#include "rcwa_amd.h"
#include "common.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

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

void free_plan(SweepPlan* p) {
  if (!p) return;
  if (p->C) (void)hipFree(p->C);
  if (p->dfirst) (void)hipFree(p->dfirst);
  if (p->dmat) (void)hipFree(p->dmat);
  delete p;
}

// Device temporaries of plan creation, freed on every exit.
struct Temps {
  std::vector<void*> p;
  ~Temps() {
    for (void* q : p) (void)hipFree(q);
  }
  hipError_t alloc(void** out, size_t bytes) {
    hipError_t e = hipMalloc(out, std::max(bytes, (size_t)1));
    if (e == hipSuccess) p.push_back(*out);
    return e;
  }
};

rcwa_status build_plan(Ctx& c, const rcwa_sweep_structure* st, const rcwa_params* par,
                       SweepPlan* plan) {
  const int L = st->num_layers, M = st->num_materials, Nx = st->Nx, Ny = st->Ny;
  const int hx = par->Nx_harmonics, hy = par->Ny_harmonics;
  const int N = (2 * hx + 1) * (2 * hy + 1);
  const size_t cells = (size_t)Nx * Ny;
  plan->ctx = &c;
  plan->M = M;
  plan->N = N;
  plan->par = *par;
  plan->t.assign(st->t, st->t + L);
  plan->dev.Nx = Nx;
  plan->dev.Ny = Ny;
  plan->dev.num_layers = L;
  plan->dev.Lx = st->Lx;
  plan->dev.Ly = st->Ly;
  plan->dev.er = nullptr;
  plan->dev.ur = nullptr;
  plan->dev.t = plan->t.data();
  plan->dev.grids_on_device = 1;
  // (layer, material) pairs present, layer by layer
  std::vector<int> play;
  plan->first.assign(L + 1, 0);
  plan->mat.clear();
  for (int l = 0; l < L; ++l) {
    std::vector<char> has(M, 0);
    const unsigned char* g = st->material + (size_t)l * cells;
    for (size_t i = 0; i < cells; ++i) {
      if (g[i] >= M) return rcwa_set_error(RCWA_E_ARG, "material index out of range");
      has[g[i]] = 1;
    }
    for (int m = 0; m < M; ++m)
      if (has[m]) {
        plan->mat.push_back(m);
        play.push_back(l);
      }
    plan->first[l + 1] = (int)plan->mat.size();
  }
  const int K = (int)plan->mat.size();
  const long long N2 = (long long)N * N;
  CKH(hipSetDevice(c.device));
  CKH(hipMalloc((void**)&plan->C, std::max((size_t)K * (size_t)N2 * sizeof(zc), (size_t)1)));
  CKH(hipMalloc((void**)&plan->dfirst, (size_t)(L + 1) * sizeof(int)));
  CKH(hipMalloc((void**)&plan->dmat, std::max((size_t)K * sizeof(int), (size_t)1)));
  hipStream_t s = c.stream;
  CKH(hipMemcpyAsync(plan->dfirst, plan->first.data(), (size_t)(L + 1) * sizeof(int),
                     hipMemcpyHostToDevice, s));
  if (K == 0) return hipStreamSynchronize(s) == hipSuccess ? RCWA_OK : RCWA_E_HIP;
  CKH(hipMemcpyAsync(plan->dmat, plan->mat.data(), (size_t)K * sizeof(int),
                     hipMemcpyHostToDevice, s));
  // masks -> truncated DFT -> Toeplitz gather, a few pairs at a time
  const int KB = (int)std::max<size_t>(1, std::min<size_t>(K, (size_t)(1u << 30) / (cells * sizeof(zc))));
  Temps tmp;
  unsigned char* dmatl = nullptr;
  int* dlay = nullptr;
  zc* masks = nullptr;
  zc* F = nullptr;
  zc* rowbuf = nullptr;
  CKH(tmp.alloc((void**)&dmatl, (size_t)L * cells));
  CKH(tmp.alloc((void**)&dlay, (size_t)K * sizeof(int)));
  CKH(tmp.alloc((void**)&masks, (size_t)KB * cells * sizeof(zc)));
  CKH(tmp.alloc((void**)&F, (size_t)KB * (4 * hx + 1) * (4 * hy + 1) * sizeof(zc)));
  CKH(tmp.alloc((void**)&rowbuf, convmat_rowbuf_elems(Nx, Ny, hx, KB) * sizeof(zc)));
  CKH(hipMemcpyAsync(dmatl, st->material, (size_t)L * cells, hipMemcpyHostToDevice, s));
  CKH(hipMemcpyAsync(dlay, play.data(), (size_t)K * sizeof(int), hipMemcpyHostToDevice, s));
  for (int k0 = 0; k0 < K; k0 += KB) {
    const int kb = std::min(KB, K - k0);
    CK(material_masks(s, Nx, Ny, dmatl, dlay + k0, plan->dmat + k0, kb, masks));
    CK(convmat_dft(s, masks, Nx, Ny, hx, hy, F, kb, rowbuf));
    CK(convmat_gather(s, F, Nx, Ny, hx, hy, ZMat{plan->C + (size_t)k0 * N2, N, N2}, kb));
  }
  CKH(hipStreamSynchronize(s));
  return RCWA_OK;
}

}  // namespace

extern "C" {

rcwa_status rcwa_sweep_create(rcwa_ctx* ctx, const rcwa_sweep_structure* st,
                              const rcwa_params* par, rcwa_sweep** out) {
  if (!ctx || !st || !par || !out) return rcwa_set_error(RCWA_E_ARG, "null argument");
  *out = nullptr;
  if (st->Nx <= 0 || st->Ny <= 0 || st->num_layers < 0 || st->num_materials <= 0 ||
      st->num_materials > 256 || (st->num_layers > 0 && (!st->material || !st->t)))
    return rcwa_set_error(RCWA_E_ARG, "bad sweep structure");
  if (!(st->Lx > 0) || !(st->Ly > 0)) return rcwa_set_error(RCWA_E_ARG, "periods must be > 0");
  if (par->Nx_harmonics < 0 || par->Ny_harmonics < 0)
    return rcwa_set_error(RCWA_E_ARG, "negative harmonic count");
  SweepPlan* plan = new SweepPlan();
  rcwa_status s = build_plan(*reinterpret_cast<Ctx*>(ctx), st, par, plan);
  if (s != RCWA_OK) {
    free_plan(plan);
    return s;
  }
  *out = reinterpret_cast<rcwa_sweep*>(plan);
  return RCWA_OK;
}

rcwa_status rcwa_sweep_run(rcwa_sweep* sweep, int npoints, const rcwa_source* src, int nsrc,
                           const rcwa_c128* eps, const rcwa_c128* mu, rcwa_precision precision,
                           rcwa_results* out) {
  if (!sweep) return rcwa_set_error(RCWA_E_ARG, "null sweep");
  SweepPlan& plan = *reinterpret_cast<SweepPlan*>(sweep);
  CKH(hipSetDevice(plan.ctx->device));
  return solve_sweep(*plan.ctx, plan, npoints, src, nsrc, eps, mu, precision, out);
}

rcwa_status rcwa_sweep_destroy(rcwa_sweep* sweep) {
  free_plan(reinterpret_cast<SweepPlan*>(sweep));
  return RCWA_OK;
}

}  // extern "C"
