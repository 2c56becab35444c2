// Serving form of the RCWA solve path: a request loop with continuous
// batching on one GPU.  The reference has no counterpart (it runs one
// hard-coded solve, rcwa3d/src/main.cpp:35-229); each request's result is the
// one rcwa_solve gives it.
//
// Requests (any harmonic count, layer count, wavelength, angle) are checked
// and copied on submit and queued in arrival order.  Each of n_workers
// workers owns a solve context (its own two HIP streams and HBM arena,
// budget = hbm_budget / n_workers) and serves one size bucket (harmonics,
// grid) at a time as a set of flights -- one per request and incidence --
// each holding a slot for its running Redheffer product in HBM.  The worker
// loops over iterations:
//   1. admission: requests of its bucket that have arrived join as slots
//      free up (FIFO; a worker with no flights takes the oldest request's
//      bucket), so a request never waits for a batch to end;
//   2. the next layers of the flights in order of arrival, up to the
//      iteration's layer budget, go through ConvMat and the layer stage as
//      one batched launch sequence (the per-layer loop of main.cpp:146-164,
//      batched across requests);
//   3. on the second stream, each flight's running product is carried
//      through its layers of the iteration (the fold of main.cpp:153,
//      batched across flights), while the next iteration's layer stage runs
//      on the first; flights past their last layer take S_trn and leave;
//   4. results of the previous iteration's finished flights are published.
// While one worker serves a bucket the other serves the next, on its own
// streams, so their work overlaps on the device.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "rcwa_amd.h"
#include "host_math.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace {

struct Request {
  uint64_t ticket = 0;
  std::vector<rcwa_c128> er, ur;
  std::vector<double> t;
  rcwa_device dev{};
  rcwa_params par{};
  std::vector<rcwa_source> src;
  std::vector<double> R, T;
  std::vector<rcwa_c128> ra, ta;  // per-order amplitudes, 3N per source
  std::vector<rcwa_results> res;
  int N = 0;
  std::tuple<int, int, int, int> key;
  rcwa_status status = RCWA_E_NOTREADY;
  std::string err;
  double t_submit = 0, t_start = 0, t_finish = 0;
  int batch_size = 0;
  bool cancelled = false;
  bool taken = false;  // admitted into a batch (set under the server lock)
  // incidences: sources of equal (lambda, theta, phi), each solved as one
  // flight; the first `admitted` have joined a worker
  std::vector<std::vector<int>> groups;
  size_t admitted = 0;
  // the fields below change under the server lock
  int flights_left = 0;  // incidences not yet finished
  rcwa_status fail = RCWA_OK;
  std::string fail_msg;
};

// Ends a queued request with `st` (under the server lock).  A request whose
// first incidences already run keeps them: the others never start, and it
// ends with the last running one.
void drop_queued(Request* r, rcwa_status st, const std::string& msg, double now) {
  if (!r->taken) {
    r->taken = true;
    r->cancelled = st == RCWA_E_CANCELLED;
    r->status = st;
    r->err = msg;
    r->t_start = r->t_finish = now;
    return;
  }
  r->flights_left -= (int)(r->groups.size() - r->admitted);
  r->admitted = r->groups.size();
  if (r->fail == RCWA_OK) {
    r->fail = st;
    r->fail_msg = msg;
  }
  if (r->flights_left == 0 && r->status == RCWA_E_NOTREADY) {
    r->status = r->fail;
    r->err = r->fail_msg;
    r->t_finish = now;
  }
}

}  // namespace

struct rcwa_server {
  int device = 0;
  rcwa_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> pending;
  std::map<uint64_t, std::unique_ptr<Request>> all;
  std::vector<std::thread> workers;
  std::vector<std::unique_ptr<Ctx>> ctxs;
  bool stop = false;
  int waiters = 0;  // threads blocked in rcwa_server_wait
  uint64_t next_ticket = 1;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::string init_error;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  void worker(int wi);
};

// ====================================================================== flights
namespace {

using cplx = std::complex<double>;

#define ECK(expr)                                                                  \
  do {                                                                             \
    int _e = (int)(expr);                                                          \
    if (_e) {                                                                      \
      if (_e < 100) return (rcwa_status)_e;                                        \
      return rcwa_set_error(RCWA_E_HIP, std::string("HIP error ") +                \
                                            hipGetErrorString((hipError_t)_e) + " in " #expr); \
    }                                                                              \
  } while (0)

// One incidence of one request in flight: its layers go through the layer
// stage in order, some per iteration, and its running Redheffer product
// (P12, P22 and the two source columns of P11, P21) stays in its slot in HBM
// between iterations.
struct Flight {
  Request* r = nullptr;
  std::vector<int> srcs;  // sources of r with this incidence
  rcwa_host::Incidence inc{};
  rcwa_host::Media media{};
  rcwa_host::WaveVectors wv;
  std::vector<cplx> v0;  // gap-medium V0 diagonals, 4N
  int L = 0, next = 0, slot = -1;
  double cond = 0;
  bool started = false;  // its product holds S_ref
  bool leaving = false;  // takes S_trn in an iteration already queued
  bool failed = false;
  rcwa_status st = RCWA_OK;
  std::string err;
};

// The host side of one iteration, kept until its fold has been read back.
struct Iter {
  std::vector<Flight*> order;  // flights with layers this iteration, most layers first
  std::vector<int> cnt;        // their layer counts
  std::vector<int> off;        // layer-major row of step s: off[s] + order index
  std::vector<Flight*> fin;    // flights that take S_trn and leave with this iteration
  int rows = 0, nflights = 0;
  bool live = false;
  std::vector<cplx> hKx, hKy, hv0, hends;
  std::vector<double> hk0t;
  std::vector<int> hidx, hfidx, hmap;
};

// HBM layout of one bucket on one worker: S slots, layer batches of at most B
// matrices (layer stage launches of at most Bl), two parities of the layer
// outputs so that iteration i's fold reads its S-matrices while iteration
// i+1's layer stage writes the other set.
struct Session {
  int hx = 0, hy = 0, Nx = 0, Ny = 0, N = 0, n = 0, S = 0, B = 0, Bl = 0;
  zc *P12s = nullptr, *P22s = nullptr, *vas = nullptr, *vbs = nullptr, *endss = nullptr;  // per slot
  zc *erc = nullptr, *urc = nullptr, *dKx = nullptr, *dKy = nullptr, *dv0 = nullptr;
  zc *lam2 = nullptr, *lam = nullptr, *ilam = nullptr, *Xd = nullptr;
  zc *gre = nullptr, *gru = nullptr, *rowbuf = nullptr, *F = nullptr;
  double* dk0t = nullptr;
  int* dmap = nullptr;
  zc *S11l[2] = {}, *S12l[2] = {};
  int* dinfo[2] = {};     // eigensolver (B) and inverse (B) flags per row
  double* dcond[2] = {};  // per row
  int* didx[2] = {};      // gather/scatter index arrays
  zc* tends = nullptr;    // S x 32N: ends of the flights being started / finished
  LayerWork w{};
  Fold f;                 // work products of the fold, gathered from the slots
  int* finfo[2] = {};     // fold inverse flags per work product: the layer steps, then S_trn
  int* dcolidx = nullptr;  // {N/2, N/2 + N}: the source columns
  // pinned host copies of the read-backs
  zc* hcols[2] = {};      // S x 4n: S11 columns then S21 columns of finished flights
  int* hinfo[2] = {};     // 2B + 2S
  double* hcond[2] = {};  // B
  hipEvent_t ev_layers[2] = {}, ev_fold[2] = {};
  bool ev_fold_recorded[2] = {false, false};
};

size_t index_ints(int S) { return 8 * (size_t)S + 64; }

size_t session_bytes(int N, int Nx, int Ny, int hx, int hy, int S, int B, int Bl) {
  const size_t n = 2 * (size_t)N, n2 = n * n, N2 = (size_t)N * N, G2 = (size_t)Nx * Ny;
  const size_t nfx = 4 * (size_t)hx + 1, nfy = 4 * (size_t)hy + 1;
  size_t z = 0;
  z += (size_t)S * (2 * n2 + 4 * n + 32 * N);  // slots
  z += 2 * (size_t)B * N2;                      // erc, urc
  z += (size_t)B * (6 * N + 4 * n);             // Kx, Ky, v0, lam2/lam/ilam/X
  z += 2 * (size_t)B * G2;                      // grid staging
  z += (size_t)B * Ny * nfx + (size_t)B * nfx * nfy;  // DFT rows, coefficients
  z += 4 * (size_t)B * n2;                      // S11l/S12l, two parities
  z += (size_t)S * 32 * N;                      // ends of started / finished flights
  z += 10 * (size_t)S * n2 + n2 + 3 * (size_t)S * 2 * n + 2 * (size_t)S * 2 * n;  // fold work
  size_t bytes = z * sizeof(zc);
  bytes += (size_t)B * sizeof(double) + (size_t)B * sizeof(int);
  bytes += 2 * (2 * (size_t)B * sizeof(int) + (size_t)B * sizeof(double) +
                index_ints(S) * sizeof(int) + 2 * (size_t)S * sizeof(int));
  bytes += layer_work_bytes(N, Bl) + block_inverse_ws_bytes((int)n, S) + 64 * 256;
  bytes += (size_t)S * sizeof(int) + 256;  // the fold's refinement flags
  return bytes;
}

class Engine {
 public:
  Engine(rcwa_server* srv, Ctx& c) : srv_(srv), c_(c) {}
  ~Engine() { free_pinned(); }
  void run();

 private:
  rcwa_server* srv_;
  Ctx& c_;
  Session se_;
  std::tuple<int, int, int, int> key_{};
  bool open_ = false;
  std::vector<std::unique_ptr<Flight>> flights_;  // in arrival order
  std::vector<int> free_slots_;
  std::vector<Flight*> starting_;  // admitted, S_ref not yet formed
  Iter it_[2];
  int parity_ = 0;

  rcwa_status open_session(int hx, int hy, int Nx, int Ny);
  void free_pinned();
  int admit(bool block, bool can_switch);
  rcwa_status start_flights(std::vector<Flight*>& fl, Iter& iter);
  rcwa_status iterate(Iter& iter, int p);
  rcwa_status collect(Iter& iter, int p);
  void finish_flight(Flight* f, int batch);
  void fail_all(rcwa_status st, const std::string& msg);
};

void Engine::free_pinned() {
  for (int p = 0; p < 2; ++p) {
    if (se_.hcols[p]) (void)hipHostFree(se_.hcols[p]);
    if (se_.hinfo[p]) (void)hipHostFree(se_.hinfo[p]);
    if (se_.hcond[p]) (void)hipHostFree(se_.hcond[p]);
    se_.hcols[p] = nullptr;
    se_.hinfo[p] = nullptr;
    se_.hcond[p] = nullptr;
    if (se_.ev_layers[p]) (void)hipEventDestroy(se_.ev_layers[p]);
    if (se_.ev_fold[p]) (void)hipEventDestroy(se_.ev_fold[p]);
    se_.ev_layers[p] = se_.ev_fold[p] = nullptr;
  }
}

// Sizes the bucket's session to the worker's budget (largest layer batch of
// at most two matrices per CU for n >= 256, then slots) and carves it.
rcwa_status Engine::open_session(int hx, int hy, int Nx, int Ny) {
  free_pinned();
  Session& se = se_;
  se = Session{};
  se.hx = hx;
  se.hy = hy;
  se.Nx = Nx;
  se.Ny = Ny;
  se.N = (2 * hx + 1) * (2 * hy + 1);
  se.n = 2 * se.N;
  // slots (max_batch, at most 64), layer batches of two matrices per CU for
  // n >= 256 (the eigensolver holds at most two per CU) and four below;
  // beyond the budget the slots shrink to 16 first, then the layer batch,
  // then the slots
  const int maxb = srv_->cfg.max_batch > 0 ? srv_->cfg.max_batch : 64;
  int S = std::max(1, std::min(maxb, 64));
  int B = se.n >= 256 ? 2 * c_.cus : 4 * c_.cus;
  auto bytes = [&](int s_, int b_) {
    return session_bytes(se.N, se.Nx, se.Ny, se.hx, se.hy, s_, b_, b_);
  };
  while (bytes(S, B) > c_.budget && S > 16) S = (S + 1) / 2;
  while (bytes(S, B) > c_.budget && B > 1) B = (B + 1) / 2;
  while (bytes(S, B) > c_.budget && S > 1) S = (S + 1) / 2;
  se.S = S;
  se.B = B;
  se.Bl = B;
  if (rcwa_status st = c_.reserve(bytes(S, B))) return st;
  const size_t n = se.n, n2 = n * n, N = se.N, N2 = N * N, G2 = (size_t)se.Nx * se.Ny;
  const size_t nfx = 4 * (size_t)se.hx + 1, nfy = 4 * (size_t)se.hy + 1;
  auto Z = [&](size_t count) { return (zc*)c_.take(count * sizeof(zc)); };
  se.P12s = Z(S * n2);
  se.P22s = Z(S * n2);
  se.vas = Z(S * 2 * n);
  se.vbs = Z(S * 2 * n);
  se.endss = Z(S * 32 * N);
  se.erc = Z(B * N2);
  se.urc = Z(B * N2);
  se.dKx = Z(B * N);
  se.dKy = Z(B * N);
  se.dv0 = Z(B * 4 * N);
  se.lam2 = Z(B * n);
  se.lam = Z(B * n);
  se.ilam = Z(B * n);
  se.Xd = Z(B * n);
  se.gre = Z(B * G2);
  se.gru = Z(B * G2);
  se.rowbuf = Z(B * se.Ny * nfx);
  se.F = Z(B * nfx * nfy);
  se.dk0t = (double*)c_.take(B * sizeof(double));
  se.dmap = (int*)c_.take(B * sizeof(int));
  for (int p = 0; p < 2; ++p) {
    se.S11l[p] = Z(B * n2);
    se.S12l[p] = Z(B * n2);
    se.dinfo[p] = (int*)c_.take(2 * B * sizeof(int));
    se.dcond[p] = (double*)c_.take(B * sizeof(double));
    se.didx[p] = (int*)c_.take(index_ints(S) * sizeof(int));
    se.finfo[p] = (int*)c_.take(2 * S * sizeof(int));
  }
  se.tends = Z(S * 32 * N);
  Fold& f = se.f;
  f.c = &c_;
  f.on2 = c_.stream2 != nullptr;
  f.fs = f.on2 ? c_.stream2 : c_.stream;
  f.N = se.N;
  f.n = se.n;
  f.G = S;
  f.P12 = Z(S * n2);
  f.P22 = Z(S * n2);
  f.P12n = Z(S * n2);
  f.P22n = Z(S * n2);
  f.Fm = Z(S * n2);
  f.Gm = Z(S * n2);
  f.Hm = Z(S * n2);
  f.Mf = Z(S * n2);
  f.X11 = Z(S * n2);
  f.X21 = Z(S * n2);
  f.Id = Z(n2);
  f.vu = Z(3 * S * 2 * n);
  f.va = Z(S * 2 * n);
  f.vb = Z(S * 2 * n);
  f.gj = c_.take(block_inverse_ws_bytes(se.n, S));
  se.w.carve(c_, se.N, B);
  se.dcolidx = (int*)c_.take(4 * sizeof(int));
  f.rflag = (int*)c_.take((size_t)S * sizeof(int));
  if (rcwa_status st = c_.arena_check()) return st;
  const int hcol[4] = {se.N / 2, se.N / 2 + se.N, 0, 0};
  ECK(hipMemcpy(se.dcolidx, hcol, sizeof(hcol), hipMemcpyHostToDevice));
  f.colidx = se.dcolidx;
  for (int p = 0; p < 2; ++p) {
    ECK(hipHostMalloc((void**)&se.hcols[p], (size_t)S * 4 * n * sizeof(zc), 0));
    ECK(hipHostMalloc((void**)&se.hinfo[p], (2 * (size_t)B + 2 * S) * sizeof(int), 0));
    ECK(hipHostMalloc((void**)&se.hcond[p], (size_t)B * sizeof(double), 0));
    ECK(hipEventCreateWithFlags(&se.ev_layers[p], hipEventDisableTiming));
    ECK(hipEventCreateWithFlags(&se.ev_fold[p], hipEventDisableTiming));
  }
  free_slots_.clear();
  for (int q = S - 1; q >= 0; --q) free_slots_.push_back(q);
  ECK(zset_identity(f.fs, se.n, ZMat{f.Id, se.n, (long long)n2}, 1, zmake(1, 0)));
  open_ = true;
  return RCWA_OK;
}

// Admission (under the server lock): requests of the session's bucket, or,
// with no flights, of the oldest request's bucket, while slots are free.
// With block set, waits for work.  Returns the number of flights admitted,
// or -1 when the server stops with nothing left to do.
int Engine::admit(bool block, bool can_switch) {
  std::unique_lock<std::mutex> lk(srv_->mu);
  if (block)
    srv_->cv_work.wait(lk, [&] { return srv_->stop || !srv_->pending.empty(); });
  if (flights_.empty() && srv_->pending.empty()) return (srv_->stop && can_switch) ? -1 : 0;
  if (flights_.empty()) {
    const auto key = srv_->pending.front()->key;
    if (!open_ || key != key_) {
      // a new bucket only once no iteration of the old one is in flight
      if (!can_switch) return 0;
      key_ = key;
      lk.unlock();
      open_ = false;
      rcwa_status st = open_session(std::get<0>(key), std::get<1>(key), std::get<2>(key),
                                    std::get<3>(key));
      const std::string msg = st ? std::string(rcwa_last_error()) : std::string();
      lk.lock();
      if (st) {
        // the bucket does not fit this worker: its requests fail
        for (auto q = srv_->pending.begin(); q != srv_->pending.end();) {
          Request* r = *q;
          if (r->key == key_) {
            drop_queued(r, st, msg, srv_->now());
            q = srv_->pending.erase(q);
          } else {
            ++q;
          }
        }
        lk.unlock();
        srv_->cv_done.notify_all();
        return 0;
      }
    }
  }
  int added = 0;
  const double ts = srv_->now();
  // FIFO within the bucket: a request's incidences join as slots free up
  // (a request with more incidences than slots runs in parts), and the
  // requests behind it wait
  for (auto q = srv_->pending.begin(); q != srv_->pending.end() && !free_slots_.empty();) {
    Request* r = *q;
    if (r->key != key_) {
      ++q;
      continue;
    }
    if (!r->taken) {
      r->taken = true;
      r->t_start = ts;
      r->flights_left = (int)r->groups.size();
    }
    while (r->admitted < r->groups.size() && !free_slots_.empty()) {
      const std::vector<int>& g = r->groups[r->admitted++];
      auto f = std::make_unique<Flight>();
      f->r = r;
      f->srcs = g;
      const rcwa_source& s0 = r->src[g[0]];
      f->inc = {s0.lambda0, s0.theta, s0.phi};
      f->media = {cplx(r->par.er_ref.re, r->par.er_ref.im), cplx(r->par.ur_ref.re, r->par.ur_ref.im),
                  cplx(r->par.er_trn.re, r->par.er_trn.im), cplx(r->par.ur_trn.re, r->par.ur_trn.im)};
      f->L = r->dev.num_layers;
      f->slot = free_slots_.back();
      free_slots_.pop_back();
      starting_.push_back(f.get());
      flights_.push_back(std::move(f));
      ++added;
    }
    if (r->admitted < r->groups.size()) break;
    q = srv_->pending.erase(q);
  }
  return added;
}

// Forms S_ref in the slots of newly admitted flights (on the fold stream,
// ahead of the iteration's fold): wave vectors and end S-matrices on the
// host, expanded into the work products and scattered to the slots.
rcwa_status Engine::start_flights(std::vector<Flight*>& fl, Iter& iter) {
  Session& se = se_;
  const int N = se.N;
  std::vector<Flight*> ok;
  for (Flight* f : fl) {
    f->wv = rcwa_host::wave_vectors(f->r->dev.Lx, f->r->dev.Ly, f->inc, se.hx, se.hy, f->media);
    if (rcwa_status st = check_wave_vectors(f->wv, f->L > 0, se.hx, se.hy)) {
      f->failed = true;
      f->st = st;
      f->err = rcwa_last_error();
      continue;
    }
    f->v0 = rcwa_host::gap_v0_diagonals(f->wv.Kx, f->wv.Ky);
    f->started = true;
    ok.push_back(f);
  }
  fl.clear();
  if (ok.empty()) return RCWA_OK;
  const int cnt = (int)ok.size();
  iter.hends.assign((size_t)cnt * 32 * N, cplx(0, 0));
  iter.hidx.assign(2 * (size_t)cnt, 0);
  for (int k = 0; k < cnt; ++k) {
    Flight* f = ok[k];
    std::vector<cplx> er = rcwa_host::end_smatrix_diagonals(0, f->wv.Kx, f->wv.Ky, f->wv.Kz_ref,
                                                           f->media.er_ref, f->media.ur_ref);
    std::vector<cplx> et = rcwa_host::end_smatrix_diagonals(1, f->wv.Kx, f->wv.Ky, f->wv.Kz_trn,
                                                           f->media.er_trn, f->media.ur_trn);
    std::copy(er.begin(), er.end(), iter.hends.begin() + (size_t)k * 32 * N);
    std::copy(et.begin(), et.end(), iter.hends.begin() + (size_t)k * 32 * N + 16 * N);
    iter.hidx[k] = k;            // work product
    iter.hidx[cnt + k] = f->slot;  // slot
  }
  Fold& fo = se.f;
  hipStream_t fs = fo.fs;
  const long long n2 = (long long)se.n * se.n;
  // ends of the new flights to tends (contiguous) and to their slots
  ECK(hipMemcpyAsync(se.tends, iter.hends.data(), iter.hends.size() * sizeof(zc),
                     hipMemcpyHostToDevice, fs));
  int* dix = se.didx[parity_] + 4 * se.S;  // the upper half: admission indices
  ECK(hipMemcpyAsync(dix, iter.hidx.data(), iter.hidx.size() * sizeof(int),
                     hipMemcpyHostToDevice, fs));
  ECK(gather_mats(fs, 32LL * N, se.tends, dix, se.endss, dix + cnt, cnt));
  fo.G = cnt;
  if (rcwa_status st = fo.start(se.tends, 32LL * N, cnt)) return st;
  ECK(gather_mats(fs, n2, fo.P12, dix, se.P12s, dix + cnt, cnt));
  ECK(gather_mats(fs, n2, fo.P22, dix, se.P22s, dix + cnt, cnt));
  ECK(gather_mats(fs, 2LL * se.n, fo.va, dix, se.vas, dix + cnt, cnt));
  ECK(gather_mats(fs, 2LL * se.n, fo.vb, dix, se.vbs, dix + cnt, cnt));
  return RCWA_OK;
}

// Queues iteration `iter` (parity p): ConvMat and the layer stage of the
// flights' next layers on the first stream, their fold on the second.
rcwa_status Engine::iterate(Iter& iter, int p) {
  Session& se = se_;
  const int N = se.N, n = se.n, B = se.B, S = se.S;
  const long long n2 = (long long)n * n, N2 = (long long)N * N, G2 = (long long)se.Nx * se.Ny;
  hipStream_t s1 = c_.stream;
  Fold& fo = se.f;
  hipStream_t fs = fo.fs;
  iter.order.clear();
  iter.cnt.clear();
  iter.off.clear();
  iter.fin.clear();
  iter.rows = 0;
  iter.live = false;
  // the next layers of the flights, oldest first, up to B
  std::vector<std::pair<Flight*, int>> picks;
  int left = B;
  for (auto& up : flights_) {
    Flight* f = up.get();
    if (f->failed || !f->started || f->leaving) continue;
    const int c = std::min(f->L - f->next, left);
    if (c > 0) {
      picks.push_back({f, c});
      left -= c;
    }
    if (left == 0) break;
  }
  std::stable_sort(picks.begin(), picks.end(),
                   [](const std::pair<Flight*, int>& a, const std::pair<Flight*, int>& b) {
                     return a.second > b.second;
                   });
  for (auto& pc : picks) {
    iter.order.push_back(pc.first);
    iter.cnt.push_back(pc.second);
  }
  const int F = (int)iter.order.size();
  const int maxc = F ? iter.cnt[0] : 0;
  iter.off.assign(maxc + 1, 0);
  for (int st = 0; st < maxc; ++st) {
    int a = 0;
    for (int k = 0; k < F; ++k) a += iter.cnt[k] > st;
    iter.off[st + 1] = iter.off[st] + a;
  }
  iter.rows = iter.off[maxc];
  // flights past their last layer after this iteration take S_trn
  for (auto& up : flights_) {
    Flight* f = up.get();
    if (f->failed || !f->started || f->leaving) continue;
    int c = 0;
    for (int k = 0; k < F; ++k)
      if (iter.order[k] == f) c = iter.cnt[k];
    if (f->next + c == f->L) {
      f->leaving = true;
      iter.fin.push_back(f);
    }
  }
  const int Fin = (int)iter.fin.size();
  iter.nflights = std::max(F, Fin);
  if (F == 0 && Fin == 0) return RCWA_OK;
  iter.live = true;
  const int rows = iter.rows;
  if (rows > 0) {
    // per-row wave vectors, gap-medium diagonals and k0 t
    iter.hKx.assign((size_t)rows * N, cplx(0, 0));
    iter.hKy.assign((size_t)rows * N, cplx(0, 0));
    iter.hv0.assign((size_t)rows * 4 * N, cplx(0, 0));
    iter.hk0t.assign(rows, 0.0);
    iter.hmap.assign(rows, 0);
    int stage = 0;
    for (int k = 0; k < F; ++k) {
      Flight* f = iter.order[k];
      const double k0 = 2.0 * M_PI / f->inc.lambda0;
      for (int st = 0; st < iter.cnt[k]; ++st) {
        const int row = iter.off[st] + k;
        std::copy(f->wv.Kx.begin(), f->wv.Kx.end(), iter.hKx.begin() + (size_t)row * N);
        std::copy(f->wv.Ky.begin(), f->wv.Ky.end(), iter.hKy.begin() + (size_t)row * N);
        std::copy(f->v0.begin(), f->v0.end(), iter.hv0.begin() + (size_t)row * 4 * N);
        iter.hk0t[row] = k0 * f->r->t[f->next + st];
        iter.hmap[stage + st] = row;
      }
      stage += iter.cnt[k];
    }
    ECK(hipMemcpyAsync(se.dKx, iter.hKx.data(), iter.hKx.size() * sizeof(zc), hipMemcpyHostToDevice, s1));
    ECK(hipMemcpyAsync(se.dKy, iter.hKy.data(), iter.hKy.size() * sizeof(zc), hipMemcpyHostToDevice, s1));
    ECK(hipMemcpyAsync(se.dv0, iter.hv0.data(), iter.hv0.size() * sizeof(zc), hipMemcpyHostToDevice, s1));
    ECK(hipMemcpyAsync(se.dk0t, iter.hk0t.data(), rows * sizeof(double), hipMemcpyHostToDevice, s1));
    ECK(hipMemcpyAsync(se.dmap, iter.hmap.data(), rows * sizeof(int), hipMemcpyHostToDevice, s1));
    // ConvMat of every layer of the iteration (main.cpp:126-138), the grids
    // of each flight's layers staged in HBM
    stage = 0;
    for (int k = 0; k < F; ++k) {
      Flight* f = iter.order[k];
      const int c = iter.cnt[k];
      const size_t at = (size_t)f->next * G2;
      ECK(hipMemcpyAsync(se.gre + (size_t)stage * G2, f->r->er.data() + at, c * G2 * sizeof(zc),
                         hipMemcpyHostToDevice, s1));
      ECK(hipMemcpyAsync(se.gru + (size_t)stage * G2, f->r->ur.data() + at, c * G2 * sizeof(zc),
                         hipMemcpyHostToDevice, s1));
      ECK(convmat_dft(s1, se.gre + (size_t)stage * G2, se.Nx, se.Ny, se.hx, se.hy, se.F, c, se.rowbuf));
      ECK(convmat_gather(s1, se.F, se.Nx, se.Ny, se.hx, se.hy, ZMat{se.erc, N, N2}, c,
                         se.dmap + stage));
      ECK(convmat_dft(s1, se.gru + (size_t)stage * G2, se.Nx, se.Ny, se.hx, se.hy, se.F, c, se.rowbuf));
      ECK(convmat_gather(s1, se.F, se.Nx, se.Ny, se.hx, se.hy, ZMat{se.urc, N, N2}, c,
                         se.dmap + stage));
      stage += c;
    }
    // the layer outputs of parity p were last read by the fold two iterations back
    if (se.ev_fold_recorded[p]) ECK(hipStreamWaitEvent(s1, se.ev_fold[p], 0));
    ECK(hipMemsetAsync(se.dinfo[p], 0, 2 * (size_t)B * sizeof(int), s1));
    // SMatrixLayer for every row (main.cpp:151)
    const ZCMat none{nullptr, 0, 0};
    for (int r0 = 0; r0 < rows; r0 += se.Bl) {
      const int b = std::min(se.Bl, rows - r0);
      rcwa_status st = layer_stage(
          c_, N, b, ZCMat{se.erc + (size_t)r0 * N2, N, N2}, ZCMat{se.urc + (size_t)r0 * N2, N, N2},
          se.dKx + (size_t)r0 * N, se.dKy + (size_t)r0 * N, se.dv0 + (size_t)r0 * 4 * N,
          se.dk0t + r0, se.lam2 + (size_t)r0 * n, se.lam + (size_t)r0 * n, se.ilam + (size_t)r0 * n,
          se.Xd + (size_t)r0 * n, ZMat{se.S11l[p] + (size_t)r0 * n2, n, n2},
          ZMat{se.S12l[p] + (size_t)r0 * n2, n, n2}, se.w, se.dinfo[p] + r0,
          se.dinfo[p] + B + r0, none, none, se.dcond[p] + r0);
      if (st) return st;
    }
  }
  ECK(hipEventRecord(se.ev_layers[p], s1));
  ECK(hipStreamWaitEvent(fs, se.ev_layers[p], 0));
  // ---- the fold of the iteration (main.cpp:153), on the second stream
  iter.hfidx.assign(4 * (size_t)S, 0);
  int* hsm = iter.hfidx.data();
  int* hdm = hsm + S;
  int* hsf = hdm + S;
  int* hdf = hsf + S;
  for (int k = 0; k < F; ++k) {
    hsm[k] = iter.order[k]->slot;
    hdm[k] = k;
  }
  for (int k = 0; k < Fin; ++k) {
    hsf[k] = iter.fin[k]->slot;
    hdf[k] = k;
  }
  int* dx = se.didx[p];
  ECK(hipMemcpyAsync(dx, iter.hfidx.data(), 4 * (size_t)S * sizeof(int), hipMemcpyHostToDevice, fs));
  ECK(hipMemsetAsync(se.finfo[p], 0, 2 * (size_t)S * sizeof(int), fs));
  const int* dsm = dx;
  const int* ddm = dx + S;
  const int* dsf = dx + 2 * S;
  const int* ddf = dx + 3 * S;
  if (F > 0) {
    ECK(gather_mats(fs, n2, se.P12s, dsm, fo.P12, ddm, F));
    ECK(gather_mats(fs, n2, se.P22s, dsm, fo.P22, ddm, F));
    ECK(gather_mats(fs, 2LL * n, se.vas, dsm, fo.va, ddm, F));
    ECK(gather_mats(fs, 2LL * n, se.vbs, dsm, fo.vb, ddm, F));
    fo.G = F;
    fo.info = se.finfo[p];
    for (int st = 0; st < maxc; ++st) {
      const ZCMat x11{se.S11l[p] + (size_t)iter.off[st] * n2, n, n2};
      const ZCMat x12{se.S12l[p] + (size_t)iter.off[st] * n2, n, n2};
      // layers: S21 = S12, S22 = S11 (rcwa.cpp:351)
      if (rcwa_status e = fo.step(iter.off[st + 1] - iter.off[st], x11, x12, x12, x11, false))
        return e;
    }
    ECK(gather_mats(fs, n2, fo.P12, ddm, se.P12s, dsm, F));
    ECK(gather_mats(fs, n2, fo.P22, ddm, se.P22s, dsm, F));
    ECK(gather_mats(fs, 2LL * n, fo.va, ddm, se.vas, dsm, F));
    ECK(gather_mats(fs, 2LL * n, fo.vb, ddm, se.vbs, dsm, F));
  }
  if (Fin > 0) {
    // * S_trn (main.cpp:169-170), the columns only
    ECK(gather_mats(fs, n2, se.P12s, dsf, fo.P12, ddf, Fin));
    ECK(gather_mats(fs, n2, se.P22s, dsf, fo.P22, ddf, Fin));
    ECK(gather_mats(fs, 2LL * n, se.vas, dsf, fo.va, ddf, Fin));
    ECK(gather_mats(fs, 2LL * n, se.vbs, dsf, fo.vb, ddf, Fin));
    ECK(gather_mats(fs, 32LL * N, se.endss, dsf, se.tends, ddf, Fin));
    fo.G = Fin;
    fo.info = se.finfo[p] + S;
    if (rcwa_status e = fo.finish(se.tends, 32LL * N, Fin)) return e;
    ECK(hipMemcpyAsync(se.hcols[p], fo.va, (size_t)Fin * 2 * n * sizeof(zc), hipMemcpyDeviceToHost, fs));
    ECK(hipMemcpyAsync(se.hcols[p] + (size_t)S * 2 * n, fo.vb, (size_t)Fin * 2 * n * sizeof(zc),
                       hipMemcpyDeviceToHost, fs));
  }
  ECK(hipMemcpyAsync(se.hinfo[p], se.dinfo[p], 2 * (size_t)B * sizeof(int), hipMemcpyDeviceToHost, fs));
  ECK(hipMemcpyAsync(se.hinfo[p] + 2 * B, se.finfo[p], 2 * (size_t)S * sizeof(int),
                     hipMemcpyDeviceToHost, fs));
  if (rows > 0)
    ECK(hipMemcpyAsync(se.hcond[p], se.dcond[p], rows * sizeof(double), hipMemcpyDeviceToHost, fs));
  ECK(hipEventRecord(se.ev_fold[p], fs));
  se.ev_fold_recorded[p] = true;
  for (int k = 0; k < F; ++k) iter.order[k]->next += iter.cnt[k];
  return RCWA_OK;
}

// Waits for iteration `iter`'s fold and publishes what it finished.
rcwa_status Engine::collect(Iter& iter, int p) {
  Session& se = se_;
  const int B = se.B, S = se.S, n = se.n, N = se.N;
  if (!iter.live) return RCWA_OK;
  iter.live = false;
  ECK(hipEventSynchronize(se.ev_fold[p]));
  const int F = (int)iter.order.size();
  for (int k = 0; k < F; ++k) {
    Flight* f = iter.order[k];
    for (int st = 0; st < iter.cnt[k]; ++st) {
      const int row = iter.off[st] + k;
      const double cd = se.hcond[p][row];
      if (!std::isfinite(cd) && !f->failed) {
        f->failed = true;
        f->st = RCWA_E_NONFINITE;
        f->err = "non-finite permittivity or permeability, or eigenvector inverse";
      }
      f->cond = std::max(f->cond, cd);
      if (se.hinfo[p][row] && !f->failed) {
        f->failed = true;
        f->st = RCWA_E_EIG_NOCONV;
        f->err = "eigensolver QR did not converge";
      }
      if (se.hinfo[p][B + row] && !f->failed) {
        f->failed = true;
        f->st = RCWA_E_SINGULAR;
        f->err = "exactly singular matrix inverted";
      }
    }
    if (se.hinfo[p][2 * B + k] && !f->failed) {
      f->failed = true;
      f->st = RCWA_E_SINGULAR;
      f->err = "exactly singular matrix inverted";
    }
  }
  for (int k = 0; k < (int)iter.fin.size(); ++k) {
    Flight* f = iter.fin[k];
    if (se.hinfo[p][2 * B + S + k] && !f->failed) {
      f->failed = true;
      f->st = RCWA_E_SINGULAR;
      f->err = "exactly singular matrix inverted";
    }
    if (!f->failed) {
      // ComputeSourceModeCoeff .. ComputeDiffractionEfficiencies (rcwa.cpp:469-585)
      const cplx* s11c = reinterpret_cast<const cplx*>(se.hcols[p]) + (size_t)k * 2 * n;
      const cplx* s21c = reinterpret_cast<const cplx*>(se.hcols[p]) + (size_t)S * 2 * n +
                         (size_t)k * 2 * n;
      Request* r = f->r;
      for (int si : f->srcs) {
        const rcwa_source& src = r->src[si];
        rcwa_host::Efficiencies e =
            rcwa_host::efficiencies(src.theta, src.pte, src.ptm, f->wv, f->media, s11c, s21c);
        rcwa_results& o = r->res[si];
        std::memcpy(o.R, e.R.data(), N * sizeof(double));
        std::memcpy(o.T, e.T.data(), N * sizeof(double));
        std::memcpy(o.r, e.r.data(), 3 * (size_t)N * sizeof(rcwa_c128));
        std::memcpy(o.t, e.t.data(), 3 * (size_t)N * sizeof(rcwa_c128));
        o.R_tot = e.R_tot;
        o.T_tot = e.T_tot;
        o.cond = f->cond;
        if (!std::isfinite(e.R_tot) || !std::isfinite(e.T_tot)) {
          f->failed = true;
          f->st = RCWA_E_NONFINITE;
          f->err = "non-finite efficiencies";
        }
      }
    }
    finish_flight(f, iter.nflights);
  }
  // flights that failed before their last layer leave, unless the iteration
  // already queued names them (its copies may still read their grids): they
  // leave with that iteration's collection
  const Iter& queued = it_[p ^ 1];
  for (auto& up : flights_)
    if (up->failed && !up->leaving && up->slot >= 0) {
      bool named = false;
      if (queued.live)
        for (Flight* g : queued.order) named = named || g == up.get();
      if (named) continue;
      up->leaving = true;
      finish_flight(up.get(), iter.nflights);
    }
  return RCWA_OK;
}

// A flight's slot is free; its request completes with its last flight
// (flights of one request may run on both workers: the request's fields
// change under the server lock).  `batch`: flights of the flight's last
// iteration.
void Engine::finish_flight(Flight* f, int batch) {
  free_slots_.push_back(f->slot);
  f->slot = -1;
  Request* r = f->r;
  bool done = false;
  {
    std::lock_guard<std::mutex> lk(srv_->mu);
    r->batch_size = std::max(r->batch_size, batch);
    if (f->failed && r->fail == RCWA_OK) {
      r->fail = f->st;
      r->fail_msg = f->err;
    }
    if (--r->flights_left == 0 && r->status == RCWA_E_NOTREADY) {
      r->status = r->fail;
      r->err = r->fail_msg;
      r->t_finish = srv_->now();
      std::vector<rcwa_c128>().swap(r->er);
      std::vector<rcwa_c128>().swap(r->ur);
      done = true;
    }
  }
  if (done) srv_->cv_done.notify_all();
}

void Engine::fail_all(rcwa_status st, const std::string& msg) {
  for (auto& up : flights_) {
    if (up->slot < 0) continue;
    up->failed = true;
    up->st = st;
    up->err = msg;
    finish_flight(up.get(), 1);
  }
  flights_.clear();
  starting_.clear();
  it_[0].live = it_[1].live = false;
  open_ = false;
}

void Engine::run() {
  (void)hipSetDevice(srv_->device);
  ZgemmTuneScope tune(c_.autotune);
  (void)c_.set_precision(srv_->cfg.precision);  // checked at creation
  std::vector<std::unique_ptr<Flight>> retired;
  for (;;) {
    const bool idle = flights_.empty() && !it_[0].live && !it_[1].live;
    const int a = admit(idle, idle);
    if (a < 0) break;
    if (flights_.empty() && !it_[0].live && !it_[1].live) continue;
    const int p = parity_;
    rcwa_status st = RCWA_OK;
    if (!starting_.empty()) {
      st = start_flights(starting_, it_[p]);
      // flights refused at the start (k_z = 0) are in no iteration: they leave now
      for (auto& up : flights_)
        if (!st && up->failed && !up->started && up->slot >= 0) {
          up->leaving = true;
          finish_flight(up.get(), 1);
        }
    }
    if (!st) st = iterate(it_[p], p);
    if (!st) st = collect(it_[p ^ 1], p ^ 1);
    if (st) {
      const std::string msg = rcwa_last_error();
      (void)hipStreamSynchronize(c_.stream);
      if (c_.stream2) (void)hipStreamSynchronize(c_.stream2);
      fail_all(st, msg);
      continue;
    }
    parity_ ^= 1;
    // retire flights that left one collection ago (no queued iteration names them)
    retired.clear();
    for (auto q = flights_.begin(); q != flights_.end();) {
      if ((*q)->slot < 0) {
        retired.push_back(std::move(*q));
        q = flights_.erase(q);
      } else {
        ++q;
      }
    }
    if (!it_[0].live && !it_[1].live) retired.clear();
  }
  (void)hipStreamSynchronize(c_.stream);
  if (c_.stream2) (void)hipStreamSynchronize(c_.stream2);
}

}  // namespace

void rcwa_server::worker(int wi) {
  Engine e(this, *ctxs[wi]);
  e.run();
}

namespace {

rcwa_status fill(rcwa_server* s, Request* r, rcwa_results* out, rcwa_ticket_info* info) {
  if (info) *info = rcwa_ticket_info{r->t_submit, r->t_start, r->t_finish, r->batch_size};
  if (r->status != RCWA_OK) return rcwa_set_error(r->status, r->err);
  if (out) {
    for (size_t i = 0; i < r->src.size(); ++i) {
      if (out[i].R) std::memcpy(out[i].R, r->R.data() + i * r->N, r->N * sizeof(double));
      if (out[i].T) std::memcpy(out[i].T, r->T.data() + i * r->N, r->N * sizeof(double));
      if (out[i].r)
        std::memcpy(out[i].r, r->ra.data() + 3 * i * r->N, 3 * r->N * sizeof(rcwa_c128));
      if (out[i].t)
        std::memcpy(out[i].t, r->ta.data() + 3 * i * r->N, 3 * r->N * sizeof(rcwa_c128));
      out[i].R_tot = r->res[i].R_tot;
      out[i].T_tot = r->res[i].T_tot;
      out[i].cond = r->res[i].cond;
    }
  }
  (void)s;
  return RCWA_OK;
}

}  // namespace

extern "C" {

rcwa_status rcwa_server_create(int device, const rcwa_server_config* cfg, rcwa_server** out) {
  if (!out) return rcwa_set_error(RCWA_E_ARG, "null out");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return rcwa_set_error(RCWA_E_HIP, "no HIP device available");
  if (device < 0 || device >= count) return rcwa_set_error(RCWA_E_ARG, "bad device index");
  auto* s = new rcwa_server();
  s->device = device;
  if (cfg) s->cfg = *cfg;
  if (s->cfg.n_workers <= 0) s->cfg.n_workers = 2;
  if (s->cfg.precision != RCWA_FP64 && s->cfg.precision != RCWA_SPLIT) {
    delete s;
    return rcwa_set_error(RCWA_E_ARG, "unknown precision");
  }
  (void)hipSetDevice(device);
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    delete s;
    return rcwa_set_error(RCWA_E_HIP, "hipMemGetInfo failed");
  }
  const size_t budget = s->cfg.hbm_budget ? s->cfg.hbm_budget : (size_t)(0.6 * (double)fr);
  for (int i = 0; i < s->cfg.n_workers; ++i) {
    auto c = std::make_unique<Ctx>();
    rcwa_status st = c->init(device, budget / s->cfg.n_workers);
    c->autotune = s->cfg.autotune != 0;
    if (st) {
      for (auto& x : s->ctxs) x->release();
      delete s;
      return st;
    }
    s->ctxs.push_back(std::move(c));
  }
  for (int i = 0; i < s->cfg.n_workers; ++i) s->workers.emplace_back([s, i] { s->worker(i); });
  *out = s;
  return RCWA_OK;
}

rcwa_status rcwa_server_submit(rcwa_server* s, const rcwa_problem* p, uint64_t* ticket) {
  if (!s || !p || !ticket || !p->device || !p->params || !p->src || p->nsrc <= 0)
    return rcwa_set_error(RCWA_E_ARG, "bad argument");
  const rcwa_device& d = *p->device;
  if (d.grids_on_device) return rcwa_set_error(RCWA_E_ARG, "server takes host grids");
  if (rcwa_status st = validate_problem(*p)) return st;
  auto r = std::make_unique<Request>();
  const size_t G = (size_t)d.Nx * d.Ny * d.num_layers;
  r->er.assign(d.er, d.er + G);
  r->ur.assign(d.ur, d.ur + G);
  r->t.assign(d.t, d.t + d.num_layers);
  r->dev = d;
  r->dev.er = r->er.data();
  r->dev.ur = r->ur.data();
  r->dev.t = r->t.data();
  r->par = *p->params;
  r->src.assign(p->src, p->src + p->nsrc);
  r->N = (2 * r->par.Nx_harmonics + 1) * (2 * r->par.Ny_harmonics + 1);
  r->R.assign((size_t)p->nsrc * r->N, 0.0);
  r->T.assign((size_t)p->nsrc * r->N, 0.0);
  r->ra.assign(3 * (size_t)p->nsrc * r->N, rcwa_c128{0, 0});
  r->ta.assign(3 * (size_t)p->nsrc * r->N, rcwa_c128{0, 0});
  r->res.resize(p->nsrc);
  for (int i = 0; i < p->nsrc; ++i)
    r->res[i] = rcwa_results{r->R.data() + (size_t)i * r->N, r->T.data() + (size_t)i * r->N, 0, 0,
                             r->ra.data() + 3 * (size_t)i * r->N,
                             r->ta.data() + 3 * (size_t)i * r->N};
  r->key = {r->par.Nx_harmonics, r->par.Ny_harmonics, d.Nx, d.Ny};
  for (int si = 0; si < p->nsrc; ++si) {
    const rcwa_source& a = r->src[si];
    bool found = false;
    for (auto& g : r->groups) {
      const rcwa_source& b = r->src[g[0]];
      if (a.lambda0 == b.lambda0 && a.theta == b.theta && a.phi == b.phi) {
        g.push_back(si);
        found = true;
        break;
      }
    }
    if (!found) r->groups.push_back({si});
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->ticket = s->next_ticket++;
    r->t_submit = s->now();
    *ticket = r->ticket;
    s->pending.push_back(r.get());
    s->all[r->ticket] = std::move(r);
  }
  s->cv_work.notify_one();
  return RCWA_OK;
}

// A ticket's entry is erased once poll or wait has returned its final status
// (a result, an error or RCWA_E_CANCELLED): the server keeps no finished
// requests, and a later call with the same ticket reports an unknown ticket.
rcwa_status rcwa_server_poll(rcwa_server* s, uint64_t ticket, rcwa_results* out,
                             rcwa_ticket_info* info) {
  if (!s) return rcwa_set_error(RCWA_E_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return rcwa_set_error(RCWA_E_ARG, "unknown ticket");
  Request* r = it->second.get();
  if (r->status == RCWA_E_NOTREADY) return RCWA_E_NOTREADY;
  rcwa_status st = r->status == RCWA_E_CANCELLED
                       ? rcwa_set_error(RCWA_E_CANCELLED, "ticket cancelled")
                       : fill(s, r, out, info);
  s->all.erase(it);
  return st;
}

rcwa_status rcwa_server_wait(rcwa_server* s, uint64_t ticket, rcwa_results* out,
                             rcwa_ticket_info* info) {
  if (!s) return rcwa_set_error(RCWA_E_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return rcwa_set_error(RCWA_E_ARG, "unknown ticket");
  ++s->waiters;
  // cancel and destroy set the status too, so every wait ends; the entry is
  // looked up afresh because another thread collecting the same ticket may
  // have erased it meanwhile
  s->cv_done.wait(lk, [&] {
    it = s->all.find(ticket);
    return it == s->all.end() || it->second->status != RCWA_E_NOTREADY;
  });
  --s->waiters;
  rcwa_status st;
  if (it == s->all.end()) {
    st = rcwa_set_error(RCWA_E_ARG, "ticket collected by another call");
  } else {
    Request* r = it->second.get();
    st = r->status == RCWA_E_CANCELLED ? rcwa_set_error(RCWA_E_CANCELLED, "ticket cancelled")
                                       : fill(s, r, out, info);
    s->all.erase(it);
  }
  if (s->stop && s->waiters == 0) s->cv_done.notify_all();  // destroy waits for the waiters
  return st;
}

rcwa_status rcwa_server_cancel(rcwa_server* s, uint64_t ticket) {
  if (!s) return rcwa_set_error(RCWA_E_ARG, "null server");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return rcwa_set_error(RCWA_E_ARG, "unknown ticket");
    Request* r = it->second.get();
    if (r->status != RCWA_E_NOTREADY || r->taken)
      return rcwa_set_error(RCWA_E_ARG, "ticket already running or finished");
    for (auto q = s->pending.begin(); q != s->pending.end(); ++q)
      if (*q == r) {
        s->pending.erase(q);
        break;
      }
    r->cancelled = true;
    r->status = RCWA_E_CANCELLED;
  }
  s->cv_done.notify_all();
  return RCWA_OK;
}

double rcwa_server_clock(rcwa_server* s) { return s ? s->now() : 0.0; }

// Requests still queued end as RCWA_E_CANCELLED; the batches already on the
// GPU finish.  Threads blocked in rcwa_server_wait return before the server
// is freed; no call on the server may start once destroy has been called.
rcwa_status rcwa_server_destroy(rcwa_server* s) {
  if (!s) return RCWA_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (Request* r : s->pending)
      drop_queued(r, RCWA_E_CANCELLED, "server destroyed before the request started", s->now());
    s->pending.clear();
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.notify_all();
    s->cv_done.wait(lk, [&] { return s->waiters == 0; });
  }
  for (auto& c : s->ctxs) {
    (void)hipStreamSynchronize(c->stream);
    c->release();
  }
  delete s;
  return RCWA_OK;
}

}  // extern "C"
