// Host runtime of the MI355X RCWA solve path and the C ABI of
// include/rcwa_amd.h: contexts, the HBM arena, the batched layer pipeline,
// the batched Redheffer fold and the stage entry points.
//
// A solve runs the reference's sequence (rcwa3d/src/main.cpp:126-191) with
// every O(N^3) stage on the GPU, batched over all layers of all solves in a
// chunk:
//   ConvMat (truncated DFT + Toeplitz gather) for eps and mu of every layer;
//   SMatrixLayer for every layer: [[eps]]^-1, [[mu]]^-1, P, Q, Omega^2 = PQ,
//     eigendecomposition, V, W^-1, V^-1, A, B, A^-1 and the S11/S12 blocks;
//   the left fold S_dev = ((S_0 * S_1) * ...) per solve, batched over solves;
//   S_glob = (S_ref * S_dev) * S_trn;
//   the S11/S21 columns at the two source entries go back to the host, which
//   forms fields and efficiencies (O(N)).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "rcwa_amd.h"
#include "common.hpp"
#include "kernels.hpp"
#include "host_math.hpp"
#include "runtime.hpp"

using rcwa_host::cplx;

// ------------------------------------------------------------------ errors
namespace {
thread_local std::string g_last_error;
}

rcwa_status rcwa_set_error(rcwa_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}

int rcwa_internal_hip_fail(hipError_t e, const char* what, const char* file, int line) {
  g_last_error = std::string("HIP error ") + hipGetErrorString(e) + " in " + what + " (" + file +
                 ":" + std::to_string(line) + ")";
  return RCWA_E_HIP;
}

#define CK(expr)                                                                   \
  do {                                                                             \
    int _e = (int)(expr);                                                          \
    if (_e) {                                                                      \
      if (_e != RCWA_E_HIP && _e < 100) return (rcwa_status)_e;                    \
      return (rcwa_status)rcwa_internal_hip_fail((hipError_t)_e, #expr, __FILE__, __LINE__); \
    }                                                                              \
  } while (0)
#define CKH(expr)                                                                  \
  do {                                                                             \
    hipError_t _h = (expr);                                                        \
    if (_h != hipSuccess)                                                          \
      return (rcwa_status)rcwa_internal_hip_fail(_h, #expr, __FILE__, __LINE__);   \
  } while (0)

// ------------------------------------------------------------------ context
rcwa_status Ctx::init(int dev, size_t bud) {
  device = dev;
  CKH(hipSetDevice(dev));
  CKH(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  CKH(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
  size_t fr = 0, tot = 0;
  CKH(hipMemGetInfo(&fr, &tot));
  budget = bud ? bud : (size_t)(0.8 * (double)fr);
  CKH(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  for (auto& e : ev) CKH(hipEventCreate(&e));
  return RCWA_OK;
}

void Ctx::release() {
  if (split_ws) (void)hipFree(split_ws);
  split_ws = nullptr;
  split_cap = 0;
  if (split_ws2) (void)hipFree(split_ws2);
  split_ws2 = nullptr;
  split_cap2 = 0;
  if (arena) (void)hipFree(arena);
  arena = nullptr;
  arena_cap = 0;
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (stream4) (void)hipStreamDestroy(stream4);
  stream4 = nullptr;
  if (stream3) (void)hipStreamDestroy(stream3);
  stream3 = nullptr;
  if (stream2) (void)hipStreamDestroy(stream2);
  stream2 = nullptr;
  if (stream) (void)hipStreamDestroy(stream);
  stream = nullptr;
}

rcwa_status Ctx::reserve(size_t bytes) {
  bytes = (bytes + 4095) & ~(size_t)4095;
  if (bytes > budget)
    return rcwa_set_error(RCWA_E_NOMEM, "request needs " + std::to_string(bytes >> 20) +
                                            " MiB of HBM, budget is " +
                                            std::to_string(budget >> 20) + " MiB");
  if (arena_cap < bytes) {
    CKH(hipSetDevice(device));
    CKH(hipStreamSynchronize(stream));
    if (stream2) CKH(hipStreamSynchronize(stream2));
    if (stream3) CKH(hipStreamSynchronize(stream3));
    if (stream4) CKH(hipStreamSynchronize(stream4));
    if (arena) CKH(hipFree(arena));
    arena = nullptr;
    arena_cap = 0;
    CKH(hipMalloc((void**)&arena, bytes));
    arena_cap = bytes;
  }
  arena_off = 0;
  return RCWA_OK;
}

rcwa_status Ctx::set_precision(rcwa_precision p) {
  precision = p;
  // The split form's slice workspace lies outside the arena and its budget:
  // an FP64 solve releases it, so that the arena can grow into that HBM.
  if (p == RCWA_FP64 && (split_ws || split_ws2)) {
    CKH(hipSetDevice(device));
    CKH(hipStreamSynchronize(stream));
    if (stream2) CKH(hipStreamSynchronize(stream2));
    if (stream3) CKH(hipStreamSynchronize(stream3));
    if (stream4) CKH(hipStreamSynchronize(stream4));
    if (split_ws) CKH(hipFree(split_ws));
    if (split_ws2) CKH(hipFree(split_ws2));
    split_ws = split_ws2 = nullptr;
    split_cap = split_cap2 = 0;
  }
  return RCWA_OK;
}

void* Ctx::take(size_t bytes) {
  arena_off = (arena_off + 255) & ~(size_t)255;
  void* p = arena + arena_off;
  arena_off += bytes;
  return p;
}

rcwa_status Ctx::arena_check() const {
  if (arena_off > arena_cap)
    return rcwa_set_error(RCWA_E_NOMEM, "internal: arena plan exceeded its reservation");
  return RCWA_OK;
}

// ------------------------------------------------------------------ GEMM dispatch
namespace {
// A product in the context's precision on stream st; the split form's slices
// go to ws (grown on demand, after st has drained).
int gemm_on(rcwa_precision prec, hipStream_t st, void*& ws, size_t& cap, int M, int N, int K,
            zc alpha, ZCMat A, ZCMat B, zc beta, ZCMat D, ZMat C, int batch) {
  if (prec == RCWA_SPLIT) {
    const size_t need = zgemm_split_ws_bytes(M, N, K, batch);
    if (need > cap) {
      hipError_t e = hipStreamSynchronize(st);
      if (e == hipSuccess && ws) e = hipFree(ws);
      ws = nullptr;
      cap = 0;
      if (e == hipSuccess) e = hipMalloc(&ws, need);
      if (e != hipSuccess) return (int)e;
      cap = need;
    }
    return zgemm_split_batched_ws(st, M, N, K, alpha, A, B, beta, D, C, batch, ws);
  }
  return zgemm_batched(st, M, N, K, alpha, A, B, beta, D, C, batch);
}
}  // namespace

int Ctx::gemm(int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta, ZCMat D, ZMat C,
              int batch) {
  return gemm_on(precision, stream, split_ws, split_cap, M, N, K, alpha, A, B, beta, D, C, batch);
}

int Ctx::gemm2(int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta, ZCMat D, ZMat C,
               int batch) {
  return gemm_on(precision, stream2, split_ws2, split_cap2, M, N, K, alpha, A, B, beta, D, C,
                 batch);
}

// ------------------------------------------------------------------ layer stage
size_t layer_work_bytes(int N, int Bl) {
  const size_t n = 2 * (size_t)N, n2 = n * n;
  const size_t z = (size_t)Bl * (6 * n2 + 2 * (size_t)N * N);
  return z * sizeof(zc) + gj_workspace_bytes((int)n, Bl) + eig_workspace_bytes((int)n, Bl) +
         10 * 256 + 4096;
}

static rcwa_status layer_stage_finish(Ctx& c, int N, int Bl, const zc* v0, const double* k0t,
                                      zc* lam2, zc* lam, zc* ilam, zc* X, ZMat S11, ZMat S12,
                                      LayerWork& w, int* info_sing, double* cond, int parts,
                                      const std::function<rcwa_status(int)>& on_part,
                                      hipEvent_t* marks);

rcwa_status layer_stage(Ctx& c, int N, int Bl, ZCMat erc, ZCMat urc, const zc* Kx, const zc* Ky,
                        const zc* v0, const double* k0t, zc* lam2, zc* lam, zc* ilam, zc* X,
                        ZMat S11, ZMat S12, LayerWork& w, int* info_eig, int* info_sing,
                        ZCMat erci_pre, ZCMat urci_pre, double* cond, int parts,
                        const std::function<rcwa_status(int)>& on_part, int phase, int ev_group) {
  const int n = 2 * N;
  const long long n2 = (long long)n * n, N2 = (long long)N * N;
  hipStream_t s = c.stream;
  auto M = [&](zc* p) { return ZMat{p, n, n2}; };
  auto CM = [&](const zc* p) { return ZCMat{p, n, n2}; };
  const zc one = zmake(1, 0), zero = zmake(0, 0);
  const ZCMat none{nullptr, 0, 0};
  zc *b0 = w.b[0], *b1 = w.b[1], *b2 = w.b[2], *b3 = w.b[3];
  hipEvent_t* marks = c.ev + 8 + 4 * ev_group;
  if (phase == 2) {
    CK(eig_batched(s, n, M(b2), lam2, M(b3), Bl, w.eig, info_eig, c.eig_maxit,
                   marks, c.qr_log, true,2));
    return layer_stage_finish(c, N, Bl, v0, k0t, lam2, lam, ilam, X, S11, S12, w, info_sing, cond,
                              parts, on_part, marks);
  }
  // [[eps]]^-1 and [[mu]]^-1 (rcwa.cpp:267-281: two LU solves each), unless
  // the caller formed them
  ZCMat erci{w.erci, N, N2}, urci{w.urci, N, N2};
  if (erci_pre.p) {
    erci = erci_pre;
  } else {
    CK(zcopy_batched(s, N, N, erc, ZMat{w.erci, N, N2}, Bl));
    CK(gj_inverse_batched(s, N, ZMat{w.erci, N, N2}, Bl, w.gj, info_sing));
  }
  if (urci_pre.p) {
    urci = urci_pre;
  } else {
    CK(zcopy_batched(s, N, N, urc, ZMat{w.urci, N, N2}, Bl));
    CK(gj_inverse_batched(s, N, ZMat{w.urci, N, N2}, Bl, w.gj, info_sing));
  }
  // P, Q (rcwa.cpp:264-285); Omega^2 = P Q (rcwa.cpp:288)
  CK(build_pq(s, N, erc, urc, erci, urci, Kx, Ky, M(b0), M(b1), Bl));
  // Omega^2 feeds the eigensolver, which amplifies its perturbations (P Q
  // cancels heavily), so it is formed in FP64 in both precision forms: a
  // split-precision P Q alone moves T by ~3e-6 on the 7x7-harmonic single
  // layer, the other split products together by ~3e-7 (~1e-6 at 21x21).
  CK(zgemm_batched(s, n, n, n, one, CM(b0), CM(b1), zero, none, M(b2), Bl));
  // eigendecomposition (rcwa.cpp:298-306)
  // b3 <- W^-1 of the unit-norm eigenvectors (formed inside the eigensolver
  // from the Schur vectors and the triangular eigenvectors of the Schur form)
  CK(eig_batched(s, n, M(b2), lam2, M(b3), Bl, w.eig, info_eig, c.eig_maxit,
                 marks, c.qr_log, true,phase));
  if (phase == 1) return RCWA_OK;
  return layer_stage_finish(c, N, Bl, v0, k0t, lam2, lam, ilam, X, S11, S12, w, info_sing, cond,
                            parts, on_part, marks);
}

// The layer stage after its eigensolver: from the eigenvectors' inverse W^-1
// (b3) and the eigenvalues lambda^2 to S11, S12.
static rcwa_status layer_stage_finish(Ctx& c, int N, int Bl, const zc* v0, const double* k0t,
                                      zc* lam2, zc* lam, zc* ilam, zc* X, ZMat S11, ZMat S12,
                                      LayerWork& w, int* info_sing, double* cond, int parts,
                                      const std::function<rcwa_status(int)>& on_part,
                                      hipEvent_t* marks) {
  const int n = 2 * N;
  const long long n2 = (long long)n * n;
  hipStream_t s = c.stream;
  auto CM = [&](const zc* p) { return ZCMat{p, n, n2}; };
  const zc one = zmake(1, 0), mone = zmake(-1, 0), zero = zmake(0, 0);
  const ZCMat none{nullptr, 0, 0};
  zc *b0 = w.b[0], *b1 = w.b[1], *b2 = w.b[2], *b3 = w.b[3], *b4 = w.b[4], *b5 = w.b[5];
  if (c.time_eig) {
    float th = 0, ts = 0, tv = 0;
    CKH(hipEventSynchronize(marks[3]));
    CKH(hipEventElapsedTime(&th, marks[0], marks[1]));
    CKH(hipEventElapsedTime(&ts, marks[1], marks[2]));
    CKH(hipEventElapsedTime(&tv, marks[2], marks[3]));
    c.stats.ms_eig += th + ts + tv;
    c.stats.ms_eig_hess += th;
    c.stats.ms_eig_schur += ts;
    c.stats.ms_eig_vec += tv;
    c.stats.n_eig_launches += 1;
    c.stats.n_eig_matrices += Bl;
  }
  // conditioning: W = Z X D^-1 has unit columns, so ||W||_F = sqrt(n) and
  // cond_2(W) <= sqrt(n) ||W^-1||_F
  if (cond) CK(zfrob_norms(s, n, n, CM(b3), std::sqrt((double)n), cond, Bl));
  // lambda, X (rcwa.cpp:317, 339)
  CK(layer_modes(s, n, lam2, lam, ilam, X, k0t, Bl));
  // A, B = W^-1 W0 +- V^-1 V0 (rcwa.cpp:326-338, W0 = I).  The reference
  // forms V = Q W diag(1/lambda) and inverts it; since P Q = W diag(lambda^2)
  // W^-1, V^-1 = diag(1/lambda) W^-1 P, so V^-1 V0 = diag(1/lambda) W^-1
  // (P V0): one product instead of a product and an inverse.  P V0 (V0 is
  // a 2x2 block of diagonals) -> b2; W^-1 is in b3; Y = W^-1 P V0 -> b0 (P
  // is not needed after this), FP64 in both precision forms (it carries the
  // eigenvectors); A -> b0, B -> b1.
  // From here on every matrix is independent of the others: with parts > 1
  // the batch is finished in that many sub-batches, and on_part(done) runs
  // after each (done = layers of the batch with their S-matrices formed), so
  // that a caller can queue work on the first layers while the last are
  // formed.
  const int np = std::max(1, std::min(parts, Bl));
  for (int q = 0; q < np; ++q) {
    const int j0 = (int)((long long)Bl * q / np), j1 = (int)((long long)Bl * (q + 1) / np);
    const int m = j1 - j0;
    if (m <= 0) continue;
    const size_t o = (size_t)j0 * n2;
    auto Mq = [&](zc* p) { return ZMat{p + o, n, n2}; };
    auto CMq = [&](const zc* p) { return ZCMat{p + o, n, n2}; };
    const zc* ilq = ilam + (size_t)j0 * n;
    const zc* Xq = X + (size_t)j0 * n;
    int* isq = info_sing + j0;
    const ZMat S11q{S11.p + (size_t)j0 * S11.stride, S11.ld, S11.stride};
    const ZMat S12q{S12.p + (size_t)j0 * S12.stride, S12.ld, S12.stride};
    CK(mul_v0(s, N, CMq(b0), v0 + (size_t)j0 * 4 * N, Mq(b2), m));
    CK(zgemm_batched(s, n, n, n, one, CMq(b3), CMq(b2), zero, none, Mq(b0), m));
    CK(build_ab(s, n, CMq(b3), Mq(b0), ilq, Mq(b1), m));
    // A^-1
    CK(zcopy_batched(s, n, n, CMq(b0), Mq(b2), m));
    CK(gj_inverse_batched(s, n, Mq(b2), m, w.gj, isq));
    // B A^-1 -> b3; C2 = A - (B A^-1) B -> b2 (rcwa.cpp:348, the reference's
    // B (A^-1 B) reassociated so that B A^-1 serves twice; A^-1 is not needed
    // after this)
    CK(c.gemm(n, n, n, one, CMq(b1), CMq(b2), zero, none, Mq(b3), m));
    CK(c.gemm(n, n, n, mone, CMq(b3), CMq(b1), one, CMq(b0), Mq(b2), m));
    // M1 = X B A^-1 X (in place)
    CK(zscale_rc(s, n, n, Mq(b3), Xq, n, Xq, n, m));
    // D = A - M1 B ; R11 = M1 A - B (rcwa.cpp:346-347)
    CK(c.gemm(n, n, n, mone, CMq(b3), CMq(b1), one, CMq(b0), Mq(b4), m));
    CK(c.gemm(n, n, n, one, CMq(b3), CMq(b0), mone, CMq(b1), Mq(b5), m));
    // S11 = D^-1 R11 ; S12 = D^-1 X C2
    CK(gj_inverse_batched(s, n, Mq(b4), m, w.gj, isq));
    CK(c.gemm(n, n, n, one, CMq(b4), CMq(b5), zero, none, S11q, m));
    CK(zscale_rc(s, n, n, Mq(b4), nullptr, 0, Xq, n, m));
    CK(c.gemm(n, n, n, one, CMq(b4), CMq(b2), zero, none, S12q, m));
    if (on_part) {
      if (rcwa_status st = on_part(j1)) return st;
    }
  }
  return RCWA_OK;
}

void LayerWork::carve(Ctx& c, int N, int Bl) {
  const size_t n = 2 * (size_t)N, n2 = n * n;
  for (int i = 0; i < 6; ++i) b[i] = (zc*)c.take((size_t)Bl * n2 * sizeof(zc));
  erci = (zc*)c.take((size_t)Bl * N * N * sizeof(zc));
  urci = (zc*)c.take((size_t)Bl * N * N * sizeof(zc));
  gj = c.take(gj_workspace_bytes((int)n, Bl));
  eig = c.take(eig_workspace_bytes((int)n, Bl));
}

// ------------------------------------------------------------------ star product
// S = A * B (rcwa.cpp:588-615) for a batch.  A's four slots receive S.
// Unless keep_slots is set, the slot pointers in A and tmp may be permuted
// instead of copying results; keep_slots is needed when the batch covers only
// a prefix of the slot buffers (the other entries must stay where they are).
rcwa_status star_product(Ctx& c, int n, int batch, zc* A[4], const ZCMat B[4], zc* tmp[5],
                         void* gjws, int* info_sing, bool keep_slots) {
  hipStream_t s = c.stream;
  const long long n2 = (long long)n * n;
  auto M = [&](zc* p) { return ZMat{p, n, n2}; };
  auto CM = [&](const zc* p) { return ZCMat{p, n, n2}; };
  const zc one = zmake(1, 0), mone = zmake(-1, 0), zero = zmake(0, 0);
  const ZCMat none{nullptr, 0, 0};
  zc *&A11 = A[0], *&A12 = A[1], *&A21 = A[2], *&A22 = A[3];
  const ZCMat &B11 = B[0], &B12 = B[1], &B21 = B[2], &B22 = B[3];
  // F1 = I - B11 A22, F2 = I - A22 B11 (rcwa.cpp:602-606), inverted in place
  CK(zset_identity(s, n, M(tmp[0]), batch, one));
  CK(c.gemm(n, n, n, mone, B11, CM(A22), one, CM(tmp[0]), M(tmp[0]), batch));
  CK(zset_identity(s, n, M(tmp[1]), batch, one));
  CK(c.gemm(n, n, n, mone, CM(A22), B11, one, CM(tmp[1]), M(tmp[1]), batch));
  CK(gj_inverse_batched(s, n, M(tmp[0]), batch, gjws, info_sing));
  CK(gj_inverse_batched(s, n, M(tmp[1]), batch, gjws, info_sing));
  // S11 = A11 + A12 (F1^-1 B11) A21 (rcwa.cpp:608)
  CK(c.gemm(n, n, n, one, CM(tmp[0]), B11, zero, none, M(tmp[2]), batch));
  CK(c.gemm(n, n, n, one, CM(A12), CM(tmp[2]), zero, none, M(tmp[3]), batch));
  CK(c.gemm(n, n, n, one, CM(tmp[3]), CM(A21), one, CM(A11), M(A11), batch));
  // S12 = A12 (F1^-1 B12) (rcwa.cpp:609)
  CK(c.gemm(n, n, n, one, CM(tmp[0]), B12, zero, none, M(tmp[2]), batch));
  CK(c.gemm(n, n, n, one, CM(A12), CM(tmp[2]), zero, none, M(tmp[3]), batch));
  if (keep_slots) CK(zcopy_batched(s, n, n, CM(tmp[3]), M(A12), batch));
  else std::swap(A12, tmp[3]);
  // S21 = B21 (F2^-1 A21) (rcwa.cpp:610)
  CK(c.gemm(n, n, n, one, CM(tmp[1]), CM(A21), zero, none, M(tmp[2]), batch));
  CK(c.gemm(n, n, n, one, B21, CM(tmp[2]), zero, none, M(tmp[4]), batch));
  if (keep_slots) CK(zcopy_batched(s, n, n, CM(tmp[4]), M(A21), batch));
  else std::swap(A21, tmp[4]);
  // S22 = B22 + B21 (F2^-1 A22) B12 (rcwa.cpp:611)
  CK(c.gemm(n, n, n, one, CM(tmp[1]), CM(A22), zero, none, M(tmp[2]), batch));
  CK(c.gemm(n, n, n, one, B21, CM(tmp[2]), zero, none, M(tmp[4]), batch));
  CK(c.gemm(n, n, n, one, CM(tmp[4]), B12, one, B22, M(A22), batch));
  return RCWA_OK;
}

// ------------------------------------------------------------------ solve pipeline
namespace {

// One S-matrix to compute: a device at one incidence, with the sources
// (polarisations) that share it.
struct Job {
  const rcwa_device* dev;
  const rcwa_params* par;
  rcwa_host::Incidence inc;
  rcwa_host::Media media;
  std::vector<int> srcs;   // source indices within the problem
  const rcwa_source* src;  // problem's source array
  rcwa_results* out;       // problem's output array
  rcwa_host::WaveVectors wv;
  std::vector<cplx> cols;  // S11, S21 columns at the source entries
  // a point of a sweep plan: [[eps]], [[mu]] are mixed from the plan's mask
  // matrices with the point's material values; jobs of equal `row` have
  // equal eps and mu
  const SweepPlan* plan = nullptr;
  const rcwa_c128* eps = nullptr;
  const rcwa_c128* mu = nullptr;  // nullptr: mu = 1 for every material
  int row = 0;
};

rcwa_status check_device(const rcwa_device* d, const rcwa_params* p) {
  if (!d || !p) return rcwa_set_error(RCWA_E_ARG, "null device or params");
  if (d->Nx <= 0 || d->Ny <= 0 || d->num_layers < 0)
    return rcwa_set_error(RCWA_E_ARG, "grid sizes must be positive");
  if (d->num_layers > 0 && (!d->er || !d->ur || !d->t))
    return rcwa_set_error(RCWA_E_ARG, "null er/ur/t");
  if (p->Nx_harmonics < 0 || p->Ny_harmonics < 0)
    return rcwa_set_error(RCWA_E_ARG, "negative harmonic count");
  if (!(d->Lx > 0) || !(d->Ly > 0)) return rcwa_set_error(RCWA_E_ARG, "periods must be > 0");
  if (!std::isfinite(d->Lx) || !std::isfinite(d->Ly))
    return rcwa_set_error(RCWA_E_NONFINITE, "non-finite period");
  for (int l = 0; l < d->num_layers; ++l)
    if (!std::isfinite(d->t[l]) || d->t[l] < 0)
      return rcwa_set_error(std::isfinite(d->t[l]) ? RCWA_E_ARG : RCWA_E_NONFINITE,
                            "layer " + std::to_string(l) + ": thickness must be finite and >= 0");
  for (rcwa_c128 z : {p->er_ref, p->ur_ref, p->er_trn, p->ur_trn})
    if (!std::isfinite(z.re) || !std::isfinite(z.im))
      return rcwa_set_error(RCWA_E_NONFINITE, "non-finite half-space permittivity/permeability");
  return RCWA_OK;
}

// Sources must be finite; lambda0 > 0.
rcwa_status check_source(const rcwa_source& sr) {
  if (!std::isfinite(sr.lambda0) || !std::isfinite(sr.theta) || !std::isfinite(sr.phi) ||
      !std::isfinite(sr.pte) || !std::isfinite(sr.ptm))
    return rcwa_set_error(RCWA_E_NONFINITE, "non-finite source parameter");
  if (!(sr.lambda0 > 0)) return rcwa_set_error(RCWA_E_ARG, "lambda0 must be > 0");
  return RCWA_OK;
}

// k_z = 0 at some order in a half-space, or in the gap medium when the
// device has layers: the reference divides by it (rcwa.cpp:225 in V0,
// rcwa.cpp:329 through lambda = 0, rcwa.cpp:541 and 560 in the z field
// components) and returns inf/NaN; here it is an error.
rcwa_status check_kz(const rcwa_host::WaveVectors& wv, bool has_layers, int hx, int hy) {
  const int N = (int)wv.Kx.size(), Q = 2 * hy + 1;
  for (int m = 0; m < N; ++m) {
    const cplx kz0 = 1.0 - wv.Kx[m] * wv.Kx[m] - wv.Ky[m] * wv.Ky[m];
    const char* where = wv.Kz_ref[m] == 0.0   ? "reflection half-space"
                        : wv.Kz_trn[m] == 0.0 ? "transmission half-space"
                        : (has_layers && kz0 == 0.0) ? "gap medium"
                                                     : nullptr;
    if (where)
      return rcwa_set_error(RCWA_E_SINGULAR,
                            std::string("k_z = 0 in the ") + where + " at order (p, q) = (" +
                                std::to_string(m / Q - hx) + ", " + std::to_string(m % Q - hy) +
                                ") (Wood anomaly)");
  }
  return RCWA_OK;
}

inline cplx C(rcwa_c128 z) { return cplx(z.re, z.im); }

}  // namespace

rcwa_status validate_problem(const rcwa_problem& pr) {
  if (rcwa_status st = check_device(pr.device, pr.params)) return st;
  if (pr.nsrc <= 0 || !pr.src) return rcwa_set_error(RCWA_E_ARG, "problem without sources");
  for (int i = 0; i < pr.nsrc; ++i)
    if (rcwa_status st = check_source(pr.src[i])) return st;
  return RCWA_OK;
}

rcwa_status check_wave_vectors(const rcwa_host::WaveVectors& wv, bool has_layers, int hx,
                               int hy) {
  return check_kz(wv, has_layers, hx, hy);
}

// HBM bytes of one chunk: G S-matrices with TL layers in total, layer
// sub-batch Bl.
size_t chunk_bytes(int N, int TL, int Lmax, int Nx, int Ny, int hx, int hy, bool host_grids,
                   int G, int Bl) {
  const size_t n = 2 * (size_t)N, n2 = n * n, T1 = std::max(TL, std::max(G, 1));
  const size_t nf = (size_t)(4 * hx + 1) * (4 * hy + 1);
  size_t z = 0;
  if (host_grids) z += 2 * (size_t)Lmax * Nx * Ny;  // one job's grids at a time
  z += (size_t)Lmax * (Ny * (4 * hx + 1) + nf);    // DFT rows and coefficients
  z += 2 * T1 * N * N;                              // erc, urc
  z += T1 * (6 * (size_t)N) + 4 * T1 * n;           // Kx, Ky, v0, lam2/lam/ilam/X
  z += 2 * T1 * n2;                                 // layer S11, S12
  z += 13 * (size_t)G * n2;                         // fold slots
  z += (size_t)G * 32 * N + (size_t)G * 4 * n;      // end diagonals, gathered columns
  size_t bytes = z * sizeof(zc) + T1 * (2 * sizeof(double) + 3 * sizeof(int)) + 64;
  bytes += (size_t)std::max(G, 1) * sizeof(int) + 256;  // the fold's refinement flags
  bytes += layer_work_bytes(N, Bl) + block_inverse_ws_bytes((int)n, std::max(G, 1)) +
           eig_workspace_bytes((int)n, Bl) + 64 * 256;
  return bytes;
}

// HBM bytes a chunk of G sweep points adds (M materials): the material
// values, and with `share` >= 0 distinct rows the inverses formed per row
// (a buffer for all layers and one for the rows, for eps and unless mu = 1
// for mu).  Zero for a chunk of plain problems (M = 0).
static size_t sweep_chunk_extra_bytes(int N, int L, int G, int M, int share, bool mu_unit) {
  if (M <= 0) return 0;
  const size_t N2 = (size_t)N * N, T1 = std::max((size_t)G * L, (size_t)std::max(G, 1));
  const size_t rows = (size_t)std::max(share, 0);
  size_t z = 2 * (size_t)G * M;
  if (share >= 0) z += (mu_unit ? 1 : 2) * T1 * N2 + rows * L * N2;
  return z * sizeof(zc) + 2 * (rows * L + T1) * sizeof(int) + 6 * 256;
}

// ------------------------------------------------------------------ fold
// The fold for the two source columns:
// S_glob = S_ref * S_0 * ... * S_{L-1} * S_trn (Redheffer products,
// rcwa.cpp:588-615), associated left to right.  The efficiencies need only
// the columns of S_glob's S11 and S21 at the two source entries
// (rcwa.cpp:512-514, 535, 553), so the running product P keeps P12 and P22
// as matrices and only those columns a = P11 c, b = P21 c.  With
// F = (I - X11 P22)^-1, one step P * X is
//   a += P12 F X11 b,   b = X21 (b + P22 F X11 b),
//   P12 = P12 F X12,    P22 = X22 + X21 P22 F X12
// (the push-through identity (I - P22 X11)^-1 P22 = P22 F turns the
// reference's second inverse into products): one inverse and five n^3
// products per layer instead of two and ten, and the last step (S_trn)
// needs only the columns.
//
// One step of iterative refinement for the solves with the explicit inverse
// F of M = I - X11 P22 (Y = F B, then Y += F (B - M Y)): with layer
// S-matrices of large entries M is ill-conditioned, and a product with an
// explicit inverse is then much less accurate than a backward-stable solve
// (the reference's LU solves, rcwa.cpp:602-606); the refined solve is as
// accurate as one.  Level 1 refines the column solves (two n x 2 products),
// level 2 also F X12 (two n^3 products).  Level 2 refines only the
// products whose M has ||M||_1 ||F||_1 above kRefineKappa1: below it the
// product with F is within about kappa sqrt(n) eps of the solve's accuracy,
// 1e-12 relative at n = 882, and the two n^3 products are skipped.
constexpr double kRefineKappa1 = 1e3;
int Fold::gemm(int M_, int N_, int K_, zc al, ZCMat A_, ZCMat B_, zc be, ZCMat D_, ZMat C_,
               int bt) {
  return on2 ? c->gemm2(M_, N_, K_, al, A_, B_, be, D_, C_, bt)
             : c->gemm(M_, N_, K_, al, A_, B_, be, D_, C_, bt);
}

// Every work slot, the column indices and the flags must be set before a
// fold call: the kernels index through them without further checks.
static rcwa_status fold_check(const Fold& f) {
  const zc* p[] = {f.P12, f.P22, f.P12n, f.P22n, f.Fm, f.Gm, f.Hm, f.Id, f.Mf, f.vu, f.X11, f.X21, f.va, f.vb};
  for (const zc* q : p)
    if (!q) return rcwa_set_error(RCWA_E_ARG, "fold work slot not set");
  if (!f.colidx || !f.gj || !f.c || !f.fs || !f.rflag)
    return rcwa_set_error(RCWA_E_ARG, "fold not set up");
  return RCWA_OK;
}

rcwa_status Fold::start(const zc* ends, long long estride, int cnt) {
  if (rcwa_status st = fold_check(*this)) return st;
  const long long n2 = (long long)n * n;
  const zc one = zmake(1, 0);
  auto Mt = [&](zc* p) { return ZMat{p, n, n2}; };
  auto CMt = [&](const zc* p) { return ZCMat{p, n, n2}; };
  CK(zset_identity(fs, n, ZMat{Id, n, n2}, 1, one));
  // P = S_ref: full S12, S22; columns of S11, S21
  CK(expand_block_diag(fs, N, ends + 0 * N, estride, Mt(X11), cnt));
  CK(expand_block_diag(fs, N, ends + 4 * (size_t)N, estride, Mt(P12), cnt));
  CK(expand_block_diag(fs, N, ends + 8 * (size_t)N, estride, Mt(X21), cnt));
  CK(expand_block_diag(fs, N, ends + 12 * (size_t)N, estride, Mt(P22), cnt));
  CK(gather_cols(fs, n, CMt(X11), colidx, 2, va, cnt));
  CK(gather_cols(fs, n, CMt(X21), colidx, 2, vb, cnt));
  return RCWA_OK;
}

rcwa_status Fold::step(int na, ZCMat x11, ZCMat x12, ZCMat x21, ZCMat x22, bool columns_only) {
  if (rcwa_status st = fold_check(*this)) return st;
  if (!info) return rcwa_set_error(RCWA_E_ARG, "fold without inverse flags");
  if (na > G) return rcwa_set_error(RCWA_E_ARG, "fold step over more products than it holds");
  const long long n2 = (long long)n * n;
  const zc one = zmake(1, 0), mone = zmake(-1, 0), zero = zmake(0, 0);
  const ZCMat none{nullptr, 0, 0};
  const ZCMat ident{Id, n, 0};  // stride 0: the same identity for every product
  auto Mt = [&](zc* p) { return ZMat{p, n, n2}; };
  auto CMt = [&](const zc* p) { return ZCMat{p, n, n2}; };
  auto Mv = [&](zc* p) { return ZMat{p, n, 2LL * n}; };
  auto CMv = [&](const zc* p) { return ZCMat{p, n, 2LL * n}; };
  zc* vw = vu + (size_t)G * 2 * n;
  zc* vy = vw + (size_t)G * 2 * n;
  const int refine = c->refine;
  CK(gemm(n, n, n, mone, x11, CMt(P22), one, ident, Mt(Fm), na));
  if (refine > 0) CK(zcopy_batched(fs, n, n, CMt(Fm), Mt(Mf), na));
  // RCWA_FOLD_INVERSE=gj: Gauss-Jordan for every size and batch (tests)
  const char* fi = getenv("RCWA_FOLD_INVERSE");
  if (fi && fi[0] == 'g')
    CK(gj_inverse_batched(fs, n, Mt(Fm), na, gj, info));
  else
    CK(block_inverse_batched(fs, n, Mt(Fm), na, gj, info));
  // the columns, with the old P12 and P22
  CK(gemm(n, 2, n, one, x11, CMv(vb), zero, none, Mv(vu), na));
  CK(gemm(n, 2, n, one, CMt(Fm), CMv(vu), zero, none, Mv(vw), na));
  if (refine > 0) {
    CK(gemm(n, 2, n, mone, CMt(Mf), CMv(vw), one, CMv(vu), Mv(vy), na));
    CK(gemm(n, 2, n, one, CMt(Fm), CMv(vy), one, CMv(vw), Mv(vw), na));
  }
  CK(gemm(n, 2, n, one, CMt(P12), CMv(vw), one, CMv(va), Mv(va), na));
  CK(gemm(n, 2, n, one, CMt(P22), CMv(vw), one, CMv(vb), Mv(vy), na));
  CK(gemm(n, 2, n, one, x21, CMv(vy), zero, none, Mv(vb), na));
  if (columns_only) return RCWA_OK;
  CK(gemm(n, n, n, one, CMt(Fm), x12, zero, none, Mt(Gm), na));
  if (refine > 1) {
    // Hm is free until the first refinement product: it holds the
    // estimate's partial column sums (far fewer than its G n^2 entries).
    // The refinement products are FP64 in both precision forms.
    CK(cond1_flags(fs, n, CMt(Mf), CMt(Fm), kRefineKappa1, rflag, na, (double*)Hm));
    CK(zgemm_batched_sel(fs, n, n, n, mone, CMt(Mf), CMt(Gm), one, x12, Mt(Hm), na, rflag));
    CK(zgemm_batched_sel(fs, n, n, n, one, CMt(Fm), CMt(Hm), one, CMt(Gm), Mt(Gm), na, rflag));
  }
  CK(gemm(n, n, n, one, CMt(P12), CMt(Gm), zero, none, Mt(P12n), na));
  CK(gemm(n, n, n, one, CMt(P22), CMt(Gm), zero, none, Mt(Hm), na));
  CK(gemm(n, n, n, one, x21, CMt(Hm), one, x22, Mt(P22n), na));
  if (na < G) {  // the products already past their last layer keep their P
    CK(zcopy_batched(fs, n, n, CMt(P12 + (size_t)na * n2), Mt(P12n + (size_t)na * n2), G - na));
    CK(zcopy_batched(fs, n, n, CMt(P22 + (size_t)na * n2), Mt(P22n + (size_t)na * n2), G - na));
  }
  std::swap(P12, P12n);
  std::swap(P22, P22n);
  return RCWA_OK;
}

rcwa_status Fold::finish(const zc* ends, long long estride, int cnt) {
  const long long n2 = (long long)n * n;
  auto Mt = [&](zc* p) { return ZMat{p, n, n2}; };
  auto CMt = [&](const zc* p) { return ZCMat{p, n, n2}; };
  CK(expand_block_diag(fs, N, ends + 16 * (size_t)N, estride, Mt(X11), cnt));
  CK(expand_block_diag(fs, N, ends + 24 * (size_t)N, estride, Mt(X21), cnt));
  return step(cnt, CMt(X11), CMt(X11), CMt(X21), CMt(X11), true);
}

// Computes the S-matrices and efficiencies of G jobs that share (hx, hy,
// Nx, Ny).  Jobs may have different layer counts: they are sorted by layer
// count (descending) and their layers are laid out layer-major, so that at
// fold step l the jobs still folding are a prefix of the job list and their
// layer-l S-matrices are contiguous.
static rcwa_status run_chunk(Ctx& c, std::vector<Job*> jobs) {
  std::stable_sort(jobs.begin(), jobs.end(),
                   [](const Job* a, const Job* b) { return a->dev->num_layers > b->dev->num_layers; });
  const int G = (int)jobs.size();
  const rcwa_device* d0 = jobs[0]->dev;
  const rcwa_params* p0 = jobs[0]->par;
  const int hx = p0->Nx_harmonics, hy = p0->Ny_harmonics;
  const int P = 2 * hx + 1, Q = 2 * hy + 1, N = P * Q, n = 2 * N;
  const int Nx = d0->Nx, Ny = d0->Ny;
  const bool host_grids = !d0->grids_on_device;
  const int Lmax = d0->num_layers;
  std::vector<int> act(Lmax + 1, 0), off(Lmax + 1, 0);  // active jobs and offsets per layer
  for (int l = 0; l < Lmax; ++l) {
    for (int g = 0; g < G; ++g) act[l] += jobs[g]->dev->num_layers > l;
    off[l + 1] = off[l] + act[l];
  }
  const int TL = off[Lmax];
  const long long n2 = (long long)n * n, N2 = (long long)N * N;
  const int nfx = 4 * hx + 1, nfy = 4 * hy + 1;
  hipStream_t s = c.stream;

  // sweep points: the distinct (eps, mu) rows among the jobs; when some row
  // repeats, [[eps]]^-1 (and [[mu]]^-1) are formed once per row and layer
  const SweepPlan* plan = jobs[0]->plan;
  const int Mm = plan ? plan->M : 0;
  bool mu_unit = true;
  std::vector<int> urow(G, 0), urep;  // job -> its row's index; row -> first job
  if (plan) {
    std::map<int, int> seen;
    for (int g = 0; g < G; ++g) {
      mu_unit = mu_unit && jobs[g]->mu == nullptr;
      auto it = seen.find(jobs[g]->row);
      if (it == seen.end()) {
        it = seen.emplace(jobs[g]->row, (int)urep.size()).first;
        urep.push_back(g);
      }
      urow[g] = it->second;
    }
  }
  const int U = (int)urep.size();
  const bool share = plan && U < G && Lmax > 0;
  const size_t extra = sweep_chunk_extra_bytes(N, Lmax, G, Mm, share ? U : -1, mu_unit);

  int Bl = std::max(TL, 1);
  while (Bl > 1 && chunk_bytes(N, TL, Lmax, Nx, Ny, hx, hy, host_grids, G, Bl) + extra > c.budget)
    Bl = (Bl + 1) / 2;
  // The eigensolver's kernels hold at most two matrices per CU at once, so a
  // layer batch beyond 2 x CUs only queues a second round behind the first;
  // for matrices whose work outweighs a launch (n >= 256), the layers are
  // split into batches of equal size of at most that many.  More batches also
  // let the fold of the earlier ones run beside the later ones.
  if (2 * N >= 256 && Bl > 2 * c.cus) {
    const int nb = (TL + 2 * c.cus - 1) / (2 * c.cus);
    Bl = (TL + nb - 1) / nb;
  }
  size_t need = chunk_bytes(N, TL, Lmax, Nx, Ny, hx, hy, host_grids, G, Bl) + extra;
  // The QR iteration runs one workgroup per matrix: a layer batch of at most
  // CUs / 2 matrices (n >= 256) leaves most of the chip idle while it runs,
  // so such a batch, when it is the only one, is formed as staggered groups
  // on two or three streams (FP64 form; the split form's products share one
  // slice workspace).  RCWA_LAYER_HALVES=0 turns this off, =1 on for any
  // single batch of two or more matrices.
  bool halves = c.stream2 && c.precision == RCWA_FP64 && Bl >= TL && TL >= 2;
  if (const char* he = getenv("RCWA_LAYER_HALVES"))
    halves = halves && he[0] == '1';
  else
    halves = halves && 2 * N >= 256 && TL <= c.cus / 2;
  // the first half (a first half of 30 or 40 % measured the same as 50 % on
  // the 41 x 41-harmonic, 50-layer stack)
  // Two groups by default; RCWA_LAYER_GROUPS=3 takes three (a fourth
  // stream).  Three measured 14.97-14.99 s against 15.21-15.35 s for two on
  // the 41 x 41-harmonic, 50-layer stack, but slowed per-call solves of
  // small requests (up to 25 x 25 harmonics, 1-30 layers): 2.37 against
  // 2.52 solves/s alone in a process, 1.73 against 2.72 after a server had
  // run in it (four streams leave no hardware queue to spare).
  int ngroups = halves ? 2 : 1;
  if (const char* ge = getenv("RCWA_LAYER_GROUPS"))
    if (ge[0] == '3' && halves && TL >= 3) ngroups = 3;
  auto gsize = [&](int g) { return (int)((long long)TL * (g + 1) / ngroups - (long long)TL * g / ngroups); };
  if (halves) {
    size_t two = 0;
    for (int g = 0; g < ngroups; ++g) two += layer_work_bytes(N, gsize(g));
    const size_t one = layer_work_bytes(N, Bl);
    if (two > one) {
      if (need + (two - one) <= c.budget)
        need += two - one;
      else
        halves = false;
    }
  }
  if (halves && !c.stream3) CKH(hipStreamCreateWithFlags(&c.stream3, hipStreamNonBlocking));
  if (ngroups > 2 && !c.stream4) CKH(hipStreamCreateWithFlags(&c.stream4, hipStreamNonBlocking));
  rcwa_status st = c.reserve(need);
  if (st) return st;

  // ---- host precompute (O(N) per job)
  const size_t T1 = std::max(TL, std::max(G, 1));
  std::vector<cplx> hKx((size_t)T1 * N), hKy((size_t)T1 * N), hv0((size_t)T1 * 4 * N);
  std::vector<double> hk0t(T1, 0.0);
  std::vector<int> hmap(T1, 0);
  std::vector<cplx> hends((size_t)G * 32 * N);
  for (int g = 0; g < G; ++g) {
    Job& j = *jobs[g];
    j.wv = rcwa_host::wave_vectors(j.dev->Lx, j.dev->Ly, j.inc, hx, hy, j.media);
    if ((st = check_kz(j.wv, j.dev->num_layers > 0, hx, hy)) != RCWA_OK) return st;
    std::vector<cplx> v0 = rcwa_host::gap_v0_diagonals(j.wv.Kx, j.wv.Ky);
    const double k0 = 2.0 * M_PI / j.inc.lambda0;
    for (int l = 0; l < j.dev->num_layers; ++l) {
      const size_t li = (size_t)off[l] + g;
      std::copy(j.wv.Kx.begin(), j.wv.Kx.end(), hKx.begin() + li * N);
      std::copy(j.wv.Ky.begin(), j.wv.Ky.end(), hKy.begin() + li * N);
      std::copy(v0.begin(), v0.end(), hv0.begin() + li * 4 * N);
      hk0t[li] = k0 * j.dev->t[l];
    }
    std::vector<cplx> er = rcwa_host::end_smatrix_diagonals(0, j.wv.Kx, j.wv.Ky, j.wv.Kz_ref,
                                                           j.media.er_ref, j.media.ur_ref);
    std::vector<cplx> et = rcwa_host::end_smatrix_diagonals(1, j.wv.Kx, j.wv.Ky, j.wv.Kz_trn,
                                                           j.media.er_trn, j.media.ur_trn);
    std::copy(er.begin(), er.end(), hends.begin() + (size_t)g * 32 * N);
    std::copy(et.begin(), et.end(), hends.begin() + (size_t)g * 32 * N + 16 * N);
  }

  // ---- device buffers
  zc* dgr_e = nullptr;
  zc* dgr_u = nullptr;
  if (host_grids && Lmax > 0) {
    dgr_e = (zc*)c.take((size_t)Lmax * Nx * Ny * sizeof(zc));
    dgr_u = (zc*)c.take((size_t)Lmax * Nx * Ny * sizeof(zc));
  }
  zc* rowbuf = (zc*)c.take((size_t)std::max(Lmax, 1) * Ny * nfx * sizeof(zc));
  zc* F = (zc*)c.take((size_t)std::max(Lmax, 1) * nfx * nfy * sizeof(zc));
  zc* erc = (zc*)c.take(T1 * N2 * sizeof(zc));
  zc* urc = (zc*)c.take(T1 * N2 * sizeof(zc));
  zc* dKx = (zc*)c.take(T1 * N * sizeof(zc));
  zc* dKy = (zc*)c.take(T1 * N * sizeof(zc));
  zc* dv0 = (zc*)c.take(T1 * 4 * N * sizeof(zc));
  double* dk0t = (double*)c.take(T1 * sizeof(double));
  int* dmap = (int*)c.take((size_t)(Lmax + 1) * G * sizeof(int) + 64);
  zc* lam2 = (zc*)c.take(T1 * n * sizeof(zc));
  zc* lam = (zc*)c.take(T1 * n * sizeof(zc));
  zc* ilam = (zc*)c.take(T1 * n * sizeof(zc));
  zc* Xd = (zc*)c.take(T1 * n * sizeof(zc));
  zc* S11l = (zc*)c.take(T1 * n2 * sizeof(zc));
  zc* S12l = (zc*)c.take(T1 * n2 * sizeof(zc));
  zc* slot[13];
  for (int i = 0; i < 13; ++i) slot[i] = (zc*)c.take((size_t)G * n2 * sizeof(zc));
  zc* dends = (zc*)c.take((size_t)G * 32 * N * sizeof(zc));
  zc* dcols = (zc*)c.take((size_t)G * 4 * n * sizeof(zc));
  int* dinfo = (int*)c.take(2 * T1 * sizeof(int));
  double* dcond = (double*)c.take(T1 * sizeof(double));
  int* dnonfin = (int*)c.take(sizeof(int));
  int* dcolidx = (int*)c.take(4 * sizeof(int));
  int* drflag = (int*)c.take((size_t)std::max(G, 1) * sizeof(int));
  void* gjf = c.take(block_inverse_ws_bytes(n, std::max(G, 1)));  // the fold's inverses (and the row inverses)
  LayerWork w, wg[3];
  if (halves) {
    for (int g = 0; g < ngroups; ++g) wg[g].carve(c, N, gsize(g));
  } else {
    w.carve(c, N, Bl);
  }
  zc *dcoef_e = nullptr, *dcoef_u = nullptr, *erci_all = nullptr, *urci_all = nullptr;
  zc* uinv = nullptr;
  int* dgmap = nullptr;
  if (plan) {
    dcoef_e = (zc*)c.take((size_t)G * Mm * sizeof(zc));
    dcoef_u = (zc*)c.take((size_t)G * Mm * sizeof(zc));
    if (share) {
      erci_all = (zc*)c.take(T1 * N2 * sizeof(zc));
      if (!mu_unit) urci_all = (zc*)c.take(T1 * N2 * sizeof(zc));
      uinv = (zc*)c.take((size_t)U * Lmax * N2 * sizeof(zc));
      dgmap = (int*)c.take(2 * ((size_t)U * Lmax + TL) * sizeof(int));
    }
  }
  CK(c.arena_check());

  // layer-major destination of job g's layer l
  std::vector<int> jmap((size_t)std::max(Lmax, 1) * G, 0);
  for (int g = 0; g < G; ++g)
    for (int l = 0; l < jobs[g]->dev->num_layers; ++l) jmap[(size_t)g * Lmax + l] = off[l] + g;

  CKH(hipMemsetAsync(dinfo, 0, 2 * T1 * sizeof(int), s));
  CKH(hipMemsetAsync(dnonfin, 0, sizeof(int), s));
  int* info_eig = dinfo;
  int* info_sing = dinfo + T1;
  if (TL > 0) {
    CKH(hipMemcpyAsync(dKx, hKx.data(), (size_t)TL * N * sizeof(zc), hipMemcpyHostToDevice, s));
    CKH(hipMemcpyAsync(dKy, hKy.data(), (size_t)TL * N * sizeof(zc), hipMemcpyHostToDevice, s));
    CKH(hipMemcpyAsync(dv0, hv0.data(), (size_t)TL * 4 * N * sizeof(zc), hipMemcpyHostToDevice, s));
    CKH(hipMemcpyAsync(dk0t, hk0t.data(), (size_t)TL * sizeof(double), hipMemcpyHostToDevice, s));
    CKH(hipMemcpyAsync(dmap, jmap.data(), jmap.size() * sizeof(int), hipMemcpyHostToDevice, s));
  }
  CKH(hipMemcpyAsync(dends, hends.data(), (size_t)G * 32 * N * sizeof(zc), hipMemcpyHostToDevice, s));
  const int hcol[4] = {N / 2, N / 2 + N, 0, 0};
  CKH(hipMemcpyAsync(dcolidx, hcol, sizeof(hcol), hipMemcpyHostToDevice, s));

  c.ev_on = true;
  CKH(hipEventRecord(c.ev[0], s));
  // ---- ConvMat for every layer (main.cpp:126-138)
  const size_t G2 = (size_t)Nx * Ny;
  if (plan && TL > 0) {
    // sweep points: [[eps]] = sum_m eps_m [[mask_m]] (ConvMat is linear in
    // the grid), [[mu]] likewise or I
    std::vector<zc> ce((size_t)G * Mm), cu((size_t)G * Mm);
    for (int g = 0; g < G; ++g)
      for (int m = 0; m < Mm; ++m) {
        ce[(size_t)g * Mm + m] = zmake(jobs[g]->eps[m].re, jobs[g]->eps[m].im);
        cu[(size_t)g * Mm + m] =
            jobs[g]->mu ? zmake(jobs[g]->mu[m].re, jobs[g]->mu[m].im) : zmake(1, 0);
      }
    CKH(hipMemcpyAsync(dcoef_e, ce.data(), ce.size() * sizeof(zc), hipMemcpyHostToDevice, s));
    CKH(hipMemcpyAsync(dcoef_u, cu.data(), cu.size() * sizeof(zc), hipMemcpyHostToDevice, s));
    CK(mix_convmat(s, N, Lmax, Mm, plan->dfirst, plan->dmat, plan->C, dcoef_e, dmap, erc, G));
    if (mu_unit)
      CK(zset_identity(s, N, ZMat{urc, N, N2}, TL, zmake(1, 0)));
    else
      CK(mix_convmat(s, N, Lmax, Mm, plan->dfirst, plan->dmat, plan->C, dcoef_u, dmap, urc, G));
    if (share) {
      // rows' inverses: gather each row's first job's matrices, invert them
      // (sub-batches of at most G, the size of the Gauss-Jordan workspace
      // at 2N), scatter them to every job of the row
      const int K1 = U * Lmax;
      std::vector<int> maps(2 * ((size_t)K1 + TL));
      int* sm1 = maps.data();
      int* dm1 = sm1 + K1;
      int* sm2 = dm1 + K1;
      int* dm2 = sm2 + TL;
      for (int u = 0; u < U; ++u)
        for (int l = 0; l < Lmax; ++l) {
          sm1[u * Lmax + l] = off[l] + urep[u];
          dm1[u * Lmax + l] = u * Lmax + l;
        }
      int k = 0;
      for (int l = 0; l < Lmax; ++l)
        for (int g = 0; g < act[l]; ++g, ++k) {
          sm2[k] = urow[g] * Lmax + l;
          dm2[k] = off[l] + g;
        }
      CKH(hipMemcpyAsync(dgmap, maps.data(), maps.size() * sizeof(int), hipMemcpyHostToDevice,
                         s));
      const int* dsm1 = dgmap;
      const int* ddm1 = dsm1 + K1;
      const int* dsm2 = ddm1 + K1;
      const int* ddm2 = dsm2 + TL;
      auto row_inverses = [&](const zc* src, zc* dst) -> rcwa_status {
        CK(gather_mats(s, N2, src, dsm1, uinv, ddm1, K1));
        for (int b0 = 0; b0 < K1; b0 += G)
          CK(gj_inverse_batched(s, N, ZMat{uinv + (size_t)b0 * N2, N, N2}, std::min(G, K1 - b0),
                                gjf, info_sing));
        CK(gather_mats(s, N2, uinv, dsm2, dst, ddm2, TL));
        return RCWA_OK;
      };
      if ((st = row_inverses(erc, erci_all)) != RCWA_OK) return st;
      if (!mu_unit && (st = row_inverses(urc, urci_all)) != RCWA_OK) return st;
    }
  }
  for (int g = 0; g < G && !plan; ++g) {
    const rcwa_device* d = jobs[g]->dev;
    const int L = d->num_layers;
    if (L == 0) continue;
    const zc *ge, *gu;
    if (host_grids) {
      CKH(hipMemcpyAsync(dgr_e, d->er, (size_t)L * G2 * sizeof(zc), hipMemcpyHostToDevice, s));
      CKH(hipMemcpyAsync(dgr_u, d->ur, (size_t)L * G2 * sizeof(zc), hipMemcpyHostToDevice, s));
      ge = dgr_e;
      gu = dgr_u;
    } else {
      ge = (const zc*)d->er;
      gu = (const zc*)d->ur;
    }
    const int* map = dmap + (size_t)g * Lmax;
    CK(convmat_dft(s, ge, Nx, Ny, hx, hy, F, L, rowbuf));
    CK(convmat_gather(s, F, Nx, Ny, hx, hy, ZMat{erc, N, N2}, L, map));
    CK(convmat_dft(s, gu, Nx, Ny, hx, hy, F, L, rowbuf));
    CK(convmat_gather(s, F, Nx, Ny, hx, hy, ZMat{urc, N, N2}, L, map));
  }
  // Non-finite grid values (or sweep materials) make every later stage
  // non-finite, and the QR iteration then runs to its sweep limit: checked
  // once here, on the convolution matrices, before the layer stage.
  if (TL > 0) {
    CK(zcount_nonfinite(s, erc, (long long)TL * N2, dnonfin));
    CK(zcount_nonfinite(s, urc, (long long)TL * N2, dnonfin));
    int hnf = 0;
    CKH(hipMemcpyAsync(&hnf, dnonfin, sizeof(int), hipMemcpyDeviceToHost, s));
    CKH(hipStreamSynchronize(s));
    if (hnf) {
      c.ev_on = false;
      return rcwa_set_error(RCWA_E_NONFINITE, "non-finite permittivity or permeability in the "
                                              "layer grids");
    }
  }
  CKH(hipEventRecord(c.ev[1], s));
  // ---- fold (main.cpp:143-170), on its own stream: fold step l needs only
  // layer l's S-matrices, so it runs as soon as the layer batch holding them
  // is done, beside the next batch's layer stage (the layer-major order puts
  // layer l of every job in one batch or two).  In the split form the fold's
  // products use their own slice workspace (Ctx::gemm2).
  const bool overlap = c.stream2 != nullptr;
  hipStream_t fs = overlap ? c.stream2 : s;
  std::vector<hipEvent_t> evs;  // evs[0]: inputs uploaded; then one per layer batch
  struct EventsGuard {
    std::vector<hipEvent_t>& v;
    ~EventsGuard() {
      for (hipEvent_t e : v) (void)hipEventDestroy(e);
    }
  } evs_guard{evs};
  auto fence = [&](hipStream_t from, hipStream_t to) -> rcwa_status {
    if (from == to) return RCWA_OK;
    hipEvent_t e;
    CKH(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    evs.push_back(e);
    CKH(hipEventRecord(e, from));
    CKH(hipStreamWaitEvent(to, e, 0));
    return RCWA_OK;
  };
  if ((st = fence(s, fs)) != RCWA_OK) return st;
  // The fold for the two source columns (Fold: S_ref, then one step per
  // layer, then S_trn).
  Fold fold;
  fold.c = &c;
  fold.fs = fs;
  fold.on2 = overlap;
  fold.N = N;
  fold.n = n;
  fold.G = G;
  fold.P12 = slot[0];
  fold.P22 = slot[1];
  fold.P12n = slot[2];
  fold.P22n = slot[3];
  fold.Fm = slot[4];
  fold.Gm = slot[5];
  fold.Hm = slot[6];
  fold.Id = slot[7];
  fold.vu = slot[8];  // three n x 2 column pairs per job (a slot holds G n^2 >= G 6n)
  fold.X11 = slot[9];
  fold.X21 = slot[10];
  fold.Mf = slot[11];
  fold.va = dcols;  // S11 columns, then S21 columns: read back by the host
  fold.vb = dcols + (size_t)G * 2 * n;
  fold.gj = gjf;
  fold.info = info_sing;
  fold.rflag = drflag;
  fold.colidx = dcolidx;
  if ((st = fold.start(dends, 32LL * N, G)) != RCWA_OK) return st;
  auto CMt = [&](const zc* p) { return ZCMat{p, n, n2}; };
  // A fold step over a layer launches the same kernels on the same buffers
  // as any other step with the same number of products still folding (na)
  // and the same side of the P12 / P22 ping-pong, once the layer's S11 and
  // S12 are staged in the fold's X11 / X21 slots (which hold S_ref and S_trn
  // only before the first and for the last step).  With RCWA_FOLD_GRAPHS=1
  // (FP64 form) each such step is captured once as a HIP graph and replayed:
  // one launch instead of the hundreds (thousands for n > 1024) of a step's
  // products and inverse.  Measured no faster than direct launches, so off
  // by default.  A graph is destroyed only once the fold stream has drained.
  struct StepGraphs {
    hipStream_t fs = nullptr;
    std::vector<hipGraphExec_t> ex{nullptr, nullptr}, retired;
    int na[2] = {-1, -1};
    ~StepGraphs() {
      bool any = false;
      for (hipGraphExec_t e : ex) any = any || e;
      if (!any && retired.empty()) return;
      (void)hipStreamSynchronize(fs);
      for (hipGraphExec_t e : ex)
        if (e) (void)hipGraphExecDestroy(e);
      for (hipGraphExec_t e : retired) (void)hipGraphExecDestroy(e);
    }
  } graphs;
  graphs.fs = fs;
  const char* genv = getenv("RCWA_FOLD_GRAPHS");
  const bool use_graphs = c.precision == RCWA_FP64 && genv && genv[0] == '1';
  zc* const ping = fold.P12;  // P12 on side 0 of the ping-pong
  auto layer_step = [&](int na, const zc* s11, const zc* s12) -> rcwa_status {
    if (!use_graphs)  // layers: S21 = S12, S22 = S11
      return fold.step(na, CMt(s11), CMt(s12), CMt(s12), CMt(s11), false);
    CK(zcopy_batched(fs, n, n, CMt(s11), ZMat{fold.X11, n, n2}, na));
    CK(zcopy_batched(fs, n, n, CMt(s12), ZMat{fold.X21, n, n2}, na));
    const ZCMat x11 = CMt(fold.X11), x12 = CMt(fold.X21);
    const int side = fold.P12 == ping ? 0 : 1;
    if (graphs.na[side] != na) {
      if (graphs.ex[side]) {  // may still run: destroyed with the others at the end
        graphs.retired.push_back(graphs.ex[side]);
        graphs.ex[side] = nullptr;
        graphs.na[side] = -1;
      }
      // the capture runs the step's host side (its ping-pong swap) and
      // records its launches; the graph launch below runs them
      hipGraph_t g = nullptr;
      CKH(hipStreamBeginCapture(fs, hipStreamCaptureModeThreadLocal));
      const rcwa_status st = fold.step(na, x11, x12, x12, x11, false);
      const hipError_t ec = hipStreamEndCapture(fs, &g);
      if (st != RCWA_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      CKH(ec);
      hipGraphExec_t ex = nullptr;
      const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      CKH(ei);
      graphs.ex[side] = ex;
      graphs.na[side] = na;
    } else {
      // a replay does on the host what the captured step did
      std::swap(fold.P12, fold.P12n);
      std::swap(fold.P22, fold.P22n);
    }
    CKH(hipGraphLaunch(graphs.ex[side], fs));
    return RCWA_OK;
  };
  int lf = 0;  // next fold step
  // ---- SMatrixLayer for every layer (main.cpp:151)
  // RCWA_PROGRESS set: a line on stderr as each layer batch completes (for
  // long single solves; the wait it adds is one stream synchronisation per
  // batch)
  const bool progress = getenv("RCWA_PROGRESS") != nullptr;
  // fold steps whose layers are among the first `upto` layer matrices, queued
  // on the fold stream behind the work queued so far on `from`, the stream
  // that formed them
  auto queue_fold = [&](int upto, hipStream_t from) -> rcwa_status {
    if (lf >= Lmax || off[lf + 1] > upto) return RCWA_OK;
    if (rcwa_status e = fence(from, fs)) return e;
    for (; lf < Lmax && off[lf + 1] <= upto; ++lf)
      if (rcwa_status e = layer_step(act[lf], S11l + (size_t)off[lf] * n2, S12l + (size_t)off[lf] * n2))
        return e;
    return RCWA_OK;
  };
  // With the fold on its own stream, each layer batch is finished (after its
  // eigensolver) in two halves, and the fold steps of the first half's
  // layers start while the second half is formed.
  const int parts = overlap ? 2 : 1;
  // The layer stage over layer matrices [l, l + m) on c.stream with work
  // buffers wk (phase and ev_group: layer_stage's).
  auto stage = [&](int l, int m, LayerWork& wk, int phase, int ev_group) -> rcwa_status {
    // sweep points: inverses formed per row above; [[mu]] = I is its own
    const ZCMat none_m{nullptr, 0, 0};
    const ZCMat erci_pre = erci_all ? ZCMat{erci_all + (size_t)l * N2, N, N2} : none_m;
    const ZCMat urci_pre = (plan && mu_unit) ? ZCMat{urc + (size_t)l * N2, N, N2}
                           : urci_all        ? ZCMat{urci_all + (size_t)l * N2, N, N2}
                                             : none_m;
    hipStream_t from = c.stream;
    return layer_stage(c, N, m, ZCMat{erc + (size_t)l * N2, N, N2},
                       ZCMat{urc + (size_t)l * N2, N, N2}, dKx + (size_t)l * N,
                       dKy + (size_t)l * N, dv0 + (size_t)l * 4 * N, dk0t + l,
                       lam2 + (size_t)l * n, lam + (size_t)l * n, ilam + (size_t)l * n,
                       Xd + (size_t)l * n, ZMat{S11l + (size_t)l * n2, n, n2},
                       ZMat{S12l + (size_t)l * n2, n, n2}, wk, info_eig + l, info_sing + l,
                       erci_pre, urci_pre, dcond + l, parts,
                       [&, l, from](int done) { return queue_fold(l + done, from); }, phase,
                       ev_group);
  };
  for (int l0 = 0; l0 < TL; l0 += Bl) {
    const int b = std::min(Bl, TL - l0);
    if (halves) {
      // The batch as staggered groups: group g's stage runs on its own
      // stream (c.stream, stream3, stream4) behind group g-1's Hessenberg
      // reduction, so that its matrices and Hessenberg reduction are formed
      // beside the earlier groups' QR iterations; each group's S-matrices are
      // then formed beside the later groups' QR iterations.  The host queues
      // every group up to its QR iteration before it waits on any.  The fold
      // stream follows each group in turn, so the fold steps of a layer whose
      // matrices straddle two groups follow both.
      hipStream_t gs[3] = {s, c.stream3, c.stream4};
      int g0[4] = {0, 0, 0, 0};
      for (int g = 0; g < ngroups; ++g) g0[g + 1] = g0[g] + gsize(g);
      for (int g = 0; g < ngroups && !st; ++g) {
        if (g > 0) {
          // c.ev[9 + 4 (g - 1)]: recorded after the previous group's
          // Hessenberg reduction, before its QR iteration
          const hipError_t he = hipStreamWaitEvent(gs[g], c.ev[9 + 4 * (g - 1)], 0);
          if (he != hipSuccess) {
            st = (rcwa_status)rcwa_internal_hip_fail(he, "hipStreamWaitEvent", __FILE__, __LINE__);
            break;
          }
        }
        c.stream = gs[g];
        st = stage(l0 + g0[g], g0[g + 1] - g0[g], wg[g], 1, g);
        c.stream = s;
      }
      for (int g = 0; g < ngroups && !st; ++g) {
        c.stream = gs[g];
        st = stage(l0 + g0[g], g0[g + 1] - g0[g], wg[g], 2, g);
        c.stream = s;
        if (!st) st = fence(gs[g], fs);
      }
      for (int g = 1; g < ngroups && !st; ++g) st = fence(gs[g], s);
      if (st) {
        for (int g = 1; g < ngroups; ++g) (void)hipStreamSynchronize(gs[g]);
        return st;
      }
    } else {
      st = stage(l0, b, w, 0, 0);
    }
    if (st) return st;
    if (progress) {
      CKH(hipStreamSynchronize(s));
      fprintf(stderr, "[rcwa] layer matrices %d/%d done\n", l0 + b, TL);
      fflush(stderr);
    }
    if ((st = queue_fold(l0 + b, s)) != RCWA_OK) return st;
  }
  CKH(hipEventRecord(c.ev[4], s));
  // * S_trn: only the columns
  if ((st = fold.finish(dends, 32LL * N, G)) != RCWA_OK) return st;
  if ((st = fence(fs, s)) != RCWA_OK) return st;
  CKH(hipEventRecord(c.ev[5], s));  // ms_fold: the fold's time not hidden by the layers
  // ---- S11 / S21 columns at the source entries (rcwa.cpp:512-514, 535, 553)
  std::vector<cplx> hcols((size_t)G * 4 * n);
  CKH(hipMemcpyAsync(hcols.data(), dcols, hcols.size() * sizeof(zc), hipMemcpyDeviceToHost, s));
  std::vector<int> hinfo(2 * T1);
  CKH(hipMemcpyAsync(hinfo.data(), dinfo, 2 * T1 * sizeof(int), hipMemcpyDeviceToHost, s));
  std::vector<double> hcond(T1, 0.0);
  if (TL > 0)
    CKH(hipMemcpyAsync(hcond.data(), dcond, (size_t)TL * sizeof(double), hipMemcpyDeviceToHost,
                       s));
  CKH(hipEventRecord(c.ev[6], s));
  CKH(hipStreamSynchronize(s));
  c.ev_on = false;
  float t01 = 0, t14 = 0, t45 = 0, t56 = 0;
  CKH(hipEventElapsedTime(&t01, c.ev[0], c.ev[1]));
  CKH(hipEventElapsedTime(&t14, c.ev[1], c.ev[4]));
  CKH(hipEventElapsedTime(&t45, c.ev[4], c.ev[5]));
  CKH(hipEventElapsedTime(&t56, c.ev[5], c.ev[6]));
  c.stats.ms_convmat += t01;
  c.stats.ms_layers += t14;
  c.stats.ms_fold += t45;
  c.stats.ms_out += t56;
  c.stats.ms_total += t01 + t14 + t45 + t56;
  c.stats.n_smatrices += G;
  c.stats.n_layers += TL;
  const double N3 = (double)N * N * N;
  c.stats.flops_algorithmic += (2384.0 * TL + 1621.0 * G) * N3;

  for (int i = 0; i < TL; ++i)
    if (hinfo[i]) return rcwa_set_error(RCWA_E_EIG_NOCONV, "eigensolver QR did not converge");
  for (size_t i = 0; i < T1; ++i)
    if (hinfo[T1 + i]) return rcwa_set_error(RCWA_E_SINGULAR, "exactly singular matrix inverted");

  // ---- efficiencies per source (host, O(N))
  for (int g = 0; g < G; ++g) {
    Job& j = *jobs[g];
    const cplx* s11c = hcols.data() + (size_t)g * 2 * n;
    const cplx* s21c = hcols.data() + (size_t)G * 2 * n + (size_t)g * 2 * n;
    j.cols.assign(s11c, s11c + 2 * n);
    j.cols.insert(j.cols.end(), s21c, s21c + 2 * n);
    double cond = 0;
    for (int l = 0; l < j.dev->num_layers; ++l) cond = std::max(cond, hcond[(size_t)off[l] + g]);
    if (!std::isfinite(cond))
      return rcwa_set_error(RCWA_E_NONFINITE, "non-finite eigenvector inverse");
    for (int si : j.srcs) {
      const rcwa_source& src = j.src[si];
      rcwa_host::Efficiencies e =
          rcwa_host::efficiencies(src.theta, src.pte, src.ptm, j.wv, j.media, s11c, s21c);
      rcwa_results& o = j.out[si];
      if (o.R) std::memcpy(o.R, e.R.data(), N * sizeof(double));
      if (o.T) std::memcpy(o.T, e.T.data(), N * sizeof(double));
      if (o.r) std::memcpy(o.r, e.r.data(), 3 * (size_t)N * sizeof(rcwa_c128));
      if (o.t) std::memcpy(o.t, e.t.data(), 3 * (size_t)N * sizeof(rcwa_c128));
      o.R_tot = e.R_tot;
      o.T_tot = e.T_tot;
      o.cond = cond;
      if (!std::isfinite(e.R_tot) || !std::isfinite(e.T_tot))
        return rcwa_set_error(RCWA_E_NONFINITE, "non-finite efficiencies");
    }
  }
  return RCWA_OK;
}

rcwa_status solve_problems(Ctx& c, const ProblemRef* probs, int nprob, rcwa_precision prec) {
  ZgemmTuneScope tune(c.autotune);
  if (rcwa_status st = c.set_precision(prec)) return st;
  c.last_cols.clear();
  std::vector<Job> jobs;
  for (int p = 0; p < nprob; ++p) {
    const rcwa_problem& pr = probs[p].p;
    rcwa_status st = check_device(pr.device, pr.params);
    if (st) return st;
    if (pr.nsrc <= 0 || !pr.src) return rcwa_set_error(RCWA_E_ARG, "problem without sources");
    std::vector<size_t> mine;
    for (int si = 0; si < pr.nsrc; ++si) {
      const rcwa_source& sr = pr.src[si];
      if ((st = check_source(sr)) != RCWA_OK) return st;
      bool found = false;
      for (size_t k : mine) {
        Job& j = jobs[k];
        if (j.inc.lambda0 == sr.lambda0 && j.inc.theta == sr.theta && j.inc.phi == sr.phi) {
          j.srcs.push_back(si);
          found = true;
          break;
        }
      }
      if (!found) {
        Job j;
        j.dev = pr.device;
        j.par = pr.params;
        j.inc = {sr.lambda0, sr.theta, sr.phi};
        j.media = {C(pr.params->er_ref), C(pr.params->ur_ref), C(pr.params->er_trn),
                   C(pr.params->ur_trn)};
        j.srcs.push_back(si);
        j.src = pr.src;
        j.out = probs[p].out;
        mine.push_back(jobs.size());
        jobs.push_back(j);
      }
    }
  }
  // bucket by shape; layer counts may differ within a bucket
  std::map<std::tuple<int, int, int, int, int>, std::vector<Job*>> groups;
  for (auto& j : jobs)
    groups[{j.par->Nx_harmonics, j.par->Ny_harmonics, j.dev->Nx, j.dev->Ny,
            j.dev->grids_on_device}]
        .push_back(&j);
  for (auto& kv : groups) {
    std::vector<Job*> v = kv.second;
    std::stable_sort(v.begin(), v.end(), [](const Job* a, const Job* b) {
      return a->dev->num_layers > b->dev->num_layers;
    });
    const Job& j0 = *v[0];
    const int hx = j0.par->Nx_harmonics, hy = j0.par->Ny_harmonics;
    const int N = (2 * hx + 1) * (2 * hy + 1);
    size_t i = 0;
    while (i < v.size()) {
      // largest prefix (at most max_chunk jobs) whose chunk fits the budget
      // with a layer sub-batch of min(its layers, kMinLayerBatch)
      size_t G = v.size() - i;
      if (c.max_chunk > 0) G = std::min(G, (size_t)c.max_chunk);
      for (;;) {
        int TL = 0;
        for (size_t k = i; k < i + G; ++k) TL += v[k]->dev->num_layers;
        const size_t need =
            chunk_bytes(N, TL, v[i]->dev->num_layers, j0.dev->Nx, j0.dev->Ny, hx, hy,
                        !j0.dev->grids_on_device, (int)G, std::max(1, std::min(TL, kMinLayerBatch)));
        if (need <= c.budget || G == 1) break;
        G = (G + 1) / 2;
      }
      std::vector<Job*> part(v.begin() + i, v.begin() + i + G);
      rcwa_status st = run_chunk(c, part);
      if (st) return st;
      i += G;
    }
  }
  for (auto& j : jobs) c.last_cols.insert(c.last_cols.end(), j.cols.begin(), j.cols.end());
  return RCWA_OK;
}

rcwa_status solve_sweep(Ctx& c, const SweepPlan& plan, int npoints, const rcwa_source* src,
                        int nsrc, const rcwa_c128* eps, const rcwa_c128* mu,
                        rcwa_precision prec, rcwa_results* out) {
  ZgemmTuneScope tune(c.autotune);
  if (npoints < 0 || (npoints > 0 && (!src || !eps || !out)) || nsrc <= 0)
    return rcwa_set_error(RCWA_E_ARG, "bad sweep arguments");
  if (rcwa_status st = c.set_precision(prec)) return st;
  c.last_cols.clear();
  const int M = plan.M, L = plan.dev.num_layers;
  for (size_t k = 0; k < (size_t)npoints * M; ++k)
    if (!std::isfinite(eps[k].re) || !std::isfinite(eps[k].im) ||
        (mu && (!std::isfinite(mu[k].re) || !std::isfinite(mu[k].im))))
      return rcwa_set_error(RCWA_E_NONFINITE, "non-finite material permittivity/permeability");
  // rows: points with bitwise equal eps (and mu) share one, numbered in
  // order of first appearance
  std::vector<int> row(npoints);
  std::map<std::string, int> rows;
  for (int p = 0; p < npoints; ++p) {
    std::string key((const char*)(eps + (size_t)p * M), (size_t)M * sizeof(rcwa_c128));
    if (mu) key.append((const char*)(mu + (size_t)p * M), (size_t)M * sizeof(rcwa_c128));
    row[p] = rows.emplace(key, (int)rows.size()).first->second;
  }
  const bool any_share = (int)rows.size() < npoints;
  std::vector<Job> jobs;
  for (int p = 0; p < npoints; ++p) {
    const rcwa_source* ps = src + (size_t)p * nsrc;
    const size_t first = jobs.size();
    for (int si = 0; si < nsrc; ++si) {
      const rcwa_source& sr = ps[si];
      if (rcwa_status st = check_source(sr)) return st;
      bool found = false;
      for (size_t k = first; k < jobs.size(); ++k) {
        Job& j = jobs[k];
        if (j.inc.lambda0 == sr.lambda0 && j.inc.theta == sr.theta && j.inc.phi == sr.phi) {
          j.srcs.push_back(si);
          found = true;
          break;
        }
      }
      if (found) continue;
      Job j;
      j.dev = &plan.dev;
      j.par = &plan.par;
      j.inc = {sr.lambda0, sr.theta, sr.phi};
      j.media = {C(plan.par.er_ref), C(plan.par.ur_ref), C(plan.par.er_trn), C(plan.par.ur_trn)};
      j.srcs.push_back(si);
      j.src = ps;
      j.out = out + (size_t)p * nsrc;
      j.plan = &plan;
      j.eps = eps + (size_t)p * M;
      j.mu = mu ? mu + (size_t)p * M : nullptr;
      j.row = row[p];
      jobs.push_back(j);
    }
  }
  if (jobs.empty()) return RCWA_OK;
  // equal rows next to each other, so that they land in one chunk
  std::vector<Job*> v;
  for (auto& j : jobs) v.push_back(&j);
  std::stable_sort(v.begin(), v.end(), [](const Job* a, const Job* b) { return a->row < b->row; });
  // the largest chunk that fits (with a layer sub-batch of at least
  // min(its layers, kMinLayerBatch)), then chunks of equal size
  const int hx = plan.par.Nx_harmonics, hy = plan.par.Ny_harmonics, N = plan.N;
  size_t G = v.size();
  if (c.max_chunk > 0) G = std::min(G, (size_t)c.max_chunk);
  for (;;) {
    const int TL = (int)G * L;
    const size_t need =
        chunk_bytes(N, TL, L, plan.dev.Nx, plan.dev.Ny, hx, hy, false, (int)G,
                    std::max(1, std::min(TL, kMinLayerBatch))) +
        sweep_chunk_extra_bytes(N, L, (int)G, M, any_share ? (int)G : -1, mu == nullptr);
    if (need <= c.budget || G == 1) break;
    G = (G + 1) / 2;
  }
  // sizes differ by at most one, so that no chunk runs a near-empty batch
  const size_t nch = (v.size() + G - 1) / G, base = v.size() / nch, rem = v.size() % nch;
  for (size_t k = 0, i = 0; k < nch; ++k) {
    const size_t g = base + (k < rem ? 1 : 0);
    std::vector<Job*> part(v.begin() + i, v.begin() + i + g);
    rcwa_status st = run_chunk(c, part);
    if (st) return st;
    i += g;
  }
  for (auto& j : jobs) c.last_cols.insert(c.last_cols.end(), j.cols.begin(), j.cols.end());
  return RCWA_OK;
}

// ====================================================================== C ABI
extern "C" {

const char* rcwa_last_error(void) { return g_last_error.c_str(); }
const char* rcwa_version(void) { return "rcwa_amd 0.1 (gfx950)"; }

rcwa_status rcwa_ctx_create(int device, size_t hbm_budget, rcwa_ctx** out) {
  if (!out) return rcwa_set_error(RCWA_E_ARG, "null out");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return rcwa_set_error(RCWA_E_HIP, "no HIP device available");
  if (device < 0 || device >= count) return rcwa_set_error(RCWA_E_ARG, "bad device index");
  Ctx* c = new Ctx();
  rcwa_status st = c->init(device, hbm_budget);
  if (st) {
    c->release();
    delete c;
    return st;
  }
  *out = reinterpret_cast<rcwa_ctx*>(c);
  return RCWA_OK;
}

rcwa_status rcwa_ctx_destroy(rcwa_ctx* ctx) {
  if (!ctx) return RCWA_OK;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->release();
  delete c;
  return RCWA_OK;
}

rcwa_status rcwa_ctx_set_eig_maxit(rcwa_ctx* ctx, int maxit) {
  if (!ctx || maxit < 0) return rcwa_set_error(RCWA_E_ARG, "bad argument");
  reinterpret_cast<Ctx*>(ctx)->eig_maxit = maxit;
  return RCWA_OK;
}

rcwa_status rcwa_ctx_set_qr_log(rcwa_ctx* ctx, long long rotations) {
  if (!ctx || rotations < 0) return rcwa_set_error(RCWA_E_ARG, "bad argument");
  reinterpret_cast<Ctx*>(ctx)->qr_log = rotations;
  return RCWA_OK;
}

rcwa_status rcwa_ctx_set_refine(rcwa_ctx* ctx, int level) {
  if (!ctx || level < 0 || level > 2) return rcwa_set_error(RCWA_E_ARG, "bad argument");
  reinterpret_cast<Ctx*>(ctx)->refine = level;
  return RCWA_OK;
}

rcwa_status rcwa_ctx_set_autotune(rcwa_ctx* ctx, int on) {
  if (!ctx) return rcwa_set_error(RCWA_E_ARG, "null ctx");
  reinterpret_cast<Ctx*>(ctx)->autotune = on != 0;
  return RCWA_OK;
}

int rcwa_autotune_entries(int* out, int max_entries) {
  return zgemm_tuned_entries(out, max_entries < 0 ? 0 : max_entries);
}

rcwa_status rcwa_zgemm_force_shape(int shape) {
  if (shape >= zgemm_shape_count()) return rcwa_set_error(RCWA_E_ARG, "no such tile shape");
  zgemm_force_shape(shape);
  return RCWA_OK;
}

rcwa_status rcwa_ctx_set_chunk(rcwa_ctx* ctx, int max_chunk) {
  if (!ctx) return rcwa_set_error(RCWA_E_ARG, "null ctx");
  reinterpret_cast<Ctx*>(ctx)->max_chunk = max_chunk;
  return RCWA_OK;
}

rcwa_status rcwa_ctx_set_timing(rcwa_ctx* ctx, int on) {
  if (!ctx) return rcwa_set_error(RCWA_E_ARG, "null ctx");
  reinterpret_cast<Ctx*>(ctx)->time_eig = on != 0;
  return RCWA_OK;
}

rcwa_status rcwa_ctx_get_stats(const rcwa_ctx* ctx, rcwa_stats* out) {
  if (!ctx || !out) return rcwa_set_error(RCWA_E_ARG, "null argument");
  *out = reinterpret_cast<const Ctx*>(ctx)->stats;
  return RCWA_OK;
}

rcwa_status rcwa_ctx_reset_stats(rcwa_ctx* ctx) {
  if (!ctx) return rcwa_set_error(RCWA_E_ARG, "null ctx");
  reinterpret_cast<Ctx*>(ctx)->stats = rcwa_stats{};
  return RCWA_OK;
}

rcwa_status rcwa_solve(rcwa_ctx* ctx, const rcwa_device* dev, const rcwa_params* params,
                       const rcwa_source* src, int nsrc, rcwa_precision precision,
                       rcwa_results* out) {
  rcwa_problem p{dev, params, src, nsrc};
  return rcwa_solve_batch(ctx, &p, 1, precision, out);
}

rcwa_status rcwa_solve_batch(rcwa_ctx* ctx, const rcwa_problem* probs, int nprob,
                             rcwa_precision precision, rcwa_results* out) {
  if (!ctx || !probs || nprob <= 0 || !out) return rcwa_set_error(RCWA_E_ARG, "bad argument");
  if (precision != RCWA_FP64 && precision != RCWA_SPLIT)
    return rcwa_set_error(RCWA_E_ARG, "unknown precision");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  CKH(hipSetDevice(c->device));
  std::vector<ProblemRef> ps(nprob);
  size_t off = 0;
  for (int i = 0; i < nprob; ++i) {
    ps[i].p = probs[i];
    ps[i].out = out + off;
    off += probs[i].nsrc;
  }
  return solve_problems(*c, ps.data(), nprob, precision);
}

rcwa_status rcwa_last_source_columns(rcwa_ctx* ctx, rcwa_c128* cols, size_t count) {
  if (!ctx || !cols) return rcwa_set_error(RCWA_E_ARG, "bad argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (count > c->last_cols.size()) return rcwa_set_error(RCWA_E_ARG, "count too large");
  std::memcpy((void*)cols, c->last_cols.data(), count * sizeof(rcwa_c128));
  return RCWA_OK;
}

}  // extern "C"
