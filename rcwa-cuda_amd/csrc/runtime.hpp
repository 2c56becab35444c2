// Internal host runtime types shared by the solve path and the server.
#pragma once
#include <string>
#include <vector>
#include <complex>
#include <functional>

#include "rcwa_amd.h"
#include "common.hpp"
#include "host_math.hpp"

rcwa_status rcwa_set_error(rcwa_status s, const std::string& msg);

// Split-precision complex GEMM (FP16 slices on the matrix cores, FP32
// accumulation); defined in split_gemm.hip.  ws holds zgemm_split_ws_bytes().
size_t zgemm_split_ws_bytes(int M, int N, int K, int batch);
int zgemm_split_batched_ws(hipStream_t s, int M, int N, int K, zc alpha, ZCMat A, ZCMat B,
                           zc beta, ZCMat D, ZMat C, int batch, void* ws);

// A solve context: one HIP stream, an HBM arena bounded by `budget`.
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // the fold of the batched solve, beside the layer stage
  // the second and third groups of a small layer batch, each staggered behind
  // the one before (the batched solve's layer stage).  Created on first use:
  // a process has four hardware queues, and streams beyond them share one
  // and run in order, so contexts that never take this path (the server's
  // two workers, two streams each) must not hold them.
  hipStream_t stream3 = nullptr;
  hipStream_t stream4 = nullptr;
  size_t budget = 0;
  char* arena = nullptr;
  void* split_ws = nullptr;  // operand slices of the split-precision GEMM
  size_t split_cap = 0;
  void* split_ws2 = nullptr;  // the same for the fold's products on stream2
  size_t split_cap2 = 0;
  size_t arena_cap = 0, arena_off = 0;
  int cus = 256;          // compute units of the device
  int eig_maxit = 30;
  long long qr_log = 0;  // QR rotation-log capacity per matrix (0 = default)
  int max_chunk = 0;
  int refine = 2;  // iterative refinement of the fold's solves (rcwa_ctx_set_refine)
  bool autotune = false;  // ZGEMM tile shapes tuned on first use (rcwa_ctx_set_autotune)
  rcwa_precision precision = RCWA_FP64;
  rcwa_stats stats{};
  std::vector<std::complex<double>> last_cols;
  hipEvent_t ev[20] = {};
  bool ev_on = false;
  bool time_eig = false;  // synchronise after each layer sub-batch to time the eigensolver

  rcwa_status init(int dev, size_t budget);
  void release();
  rcwa_status reserve(size_t bytes);  // arena of at least `bytes`; resets the bump pointer
  void* take(size_t bytes);
  // Sets the precision of the next solves; FP64 frees the split workspace.
  rcwa_status set_precision(rcwa_precision p);
  // Error if the buffers taken since reserve() overran the arena; called after
  // carving, before any kernel sees the buffers.
  rcwa_status arena_check() const;
  int gemm(int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta, ZCMat D, ZMat C, int batch);
  // The same on stream2 (the fold beside the layer stage), with its own split workspace.
  int gemm2(int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta, ZCMat D, ZMat C,
            int batch);
};

// Work buffers of the layer stage for a sub-batch of Bl layers.
struct LayerWork {
  zc* b[6];
  zc* erci;
  zc* urci;
  void* gj;
  void* eig;
  void carve(Ctx& c, int N, int Bl);
};

size_t layer_work_bytes(int N, int Bl);
// Layer sub-batch that chunk sizing reserves room for: the eigensolver runs
// one workgroup per matrix, so a launch needs about one matrix per CU (256 on
// MI355X) to fill the chip.  Admission keeps at least this much HBM for the
// layer stage before it adds structures to a chunk.
constexpr int kMinLayerBatch = 256;
size_t chunk_bytes(int N, int TL, int Lmax, int Nx, int Ny, int hx, int hy, bool host_grids,
                   int G, int Bl);

// erci_pre / urci_pre: [[eps]]^-1 / [[mu]]^-1 already formed (p != nullptr),
// else the stage inverts erc / urc itself.  cond (optional, Bl doubles)
// receives sqrt(n) ||W^-1||_F per layer, a bound on cond_2(W).  After the
// eigensolver the batch may be finished in `parts` sub-batches, on_part(done)
// running after each with the number of layers whose S-matrices are formed.
// The stage runs on c.stream.  phase (eig_batched's): 1 returns once the
// eigensolver's QR iteration is queued, 2 finishes the stage (same
// arguments), 0 does both.  The eigensolver's timing events (c.time_eig) are
// c.ev[8 + 4 * ev_group] to c.ev[11 + 4 * ev_group], ev_group 0 to 2.
rcwa_status layer_stage(Ctx& c, int N, int Bl, ZCMat erc, ZCMat urc, const zc* Kx, const zc* Ky,
                        const zc* v0, const double* k0t, zc* lam2, zc* lam, zc* ilam, zc* X,
                        ZMat S11, ZMat S12, LayerWork& w, int* info_eig, int* info_sing,
                        ZCMat erci_pre = ZCMat{nullptr, 0, 0},
                        ZCMat urci_pre = ZCMat{nullptr, 0, 0}, double* cond = nullptr,
                        int parts = 1, const std::function<rcwa_status(int)>& on_part = nullptr,
                        int phase = 0, int ev_group = 0);

rcwa_status star_product(Ctx& c, int n, int batch, zc* A[4], const ZCMat B[4], zc* tmp[5],
                         void* gjws, int* info_sing, bool keep_slots);

// S = A * B on device blocks (rcwa_star_product_batched and its
// allocation-free form); t[5] work slots, gj workspace, info cleared.
rcwa_status star_product_device(Ctx& c, int n, int batch, const rcwa_c128* const* A_dev,
                                const rcwa_c128* const* B_dev, rcwa_c128* const* S_dev, zc* t[5],
                                void* gj, int* dinfo);

// The Redheffer fold of G running products P = S_ref * S_0 * ... (contiguous
// batches of n x n blocks, n = 2N), keeping P12, P22 and the S11/S21 columns
// at the two source entries (va, vb: G x n x 2 each).  Work slots: P12n,
// P22n, Fm, Gm, Hm, Mf (G n^2 each), Id (n^2), X11, X21 (G n^2), vu (3 G n
// x 2).  Products run on fs (gemm2 when on2).
struct Fold {
  Ctx* c = nullptr;
  hipStream_t fs = nullptr;
  bool on2 = false;
  int N = 0, n = 0, G = 0;
  zc *P12 = nullptr, *P22 = nullptr, *P12n = nullptr, *P22n = nullptr, *Fm = nullptr,
     *Gm = nullptr, *Hm = nullptr, *Id = nullptr, *Mf = nullptr, *vu = nullptr, *X11 = nullptr,
     *X21 = nullptr, *va = nullptr, *vb = nullptr;
  void* gj = nullptr;        // block_inverse_ws_bytes(n, G) bytes
  int* info = nullptr;       // singular-pivot flags
  int* rflag = nullptr;      // G ints: which products refine F X12 (FP64 form)
  const int* colidx = nullptr;  // device {N/2, N/2 + N}
  int gemm(int M, int N, int K, zc alpha, ZCMat A, ZCMat B, zc beta, ZCMat D, ZMat C, int batch);
  // P = S_ref for the first cnt products, from their end diagonals (32 N
  // complex per product, estride apart: S_ref's four blocks then S_trn's)
  rcwa_status start(const zc* ends, long long estride, int cnt);
  // P <- P * X for the first na of the G products (the others keep P)
  rcwa_status step(int na, ZCMat x11, ZCMat x12, ZCMat x21, ZCMat x22, bool columns_only);
  // P <- P * S_trn for the first cnt products, the columns only
  rcwa_status finish(const zc* ends, long long estride, int cnt);
};

// Argument checks of one problem (devices, sources, media): the server
// applies them at submission, rcwa_solve* per call.
rcwa_status validate_problem(const rcwa_problem& p);

// A problem with the place its results go.
struct ProblemRef {
  rcwa_problem p;
  rcwa_results* out;
};

rcwa_status solve_problems(Ctx& c, const ProblemRef* probs, int nprob, rcwa_precision prec);

// A sweep plan (rcwa_sweep_create): a structure of num_layers layers made of
// M materials, with the convolution matrices C[k] (device, N x N each) of its
// (layer, material) masks; layer l's pairs are k = first[l] .. first[l+1]-1,
// of material mat[k].
struct SweepPlan {
  Ctx* ctx = nullptr;
  rcwa_device dev{};  // Nx, Ny, num_layers, Lx, Ly, t; no grids
  rcwa_params par{};
  std::vector<double> t;
  int M = 0, N = 0;
  std::vector<int> first, mat;
  zc* C = nullptr;  // device
  int* dfirst = nullptr;
  int* dmat = nullptr;
};

// Runs npoints points of a plan (rcwa_sweep_run's arguments).
rcwa_status solve_sweep(Ctx& c, const SweepPlan& plan, int npoints, const rcwa_source* src,
                        int nsrc, const rcwa_c128* eps, const rcwa_c128* mu,
                        rcwa_precision prec, rcwa_results* out);

// k_z = 0 in a half-space, or in the gap medium when has_layers: RCWA_E_SINGULAR.
rcwa_status check_wave_vectors(const rcwa_host::WaveVectors& wv, bool has_layers, int hx,
                               int hy);
