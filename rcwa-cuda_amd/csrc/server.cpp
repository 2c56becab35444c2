// Serving form of the RCWA solve path: a request loop with continuous
// batching on one GPU.  The reference has no counterpart (it runs one
// hard-coded solve, rcwa3d/src/main.cpp:35-229); each request is solved by
// the same batched pipeline as rcwa_solve_batch.
//
// Requests (any harmonic count, layer count, wavelength, angle) are copied
// on submit and queued in arrival order.  Each of n_workers workers owns a
// solve context (its own HIP stream and HBM arena, budget = hbm_budget /
// n_workers) and loops:
//   1. take the oldest pending request; its size bucket is (harmonics, grid);
//   2. admit further pending requests of that bucket, oldest first, while the
//      chunk's HBM footprint (chunk_bytes: layers, S-matrix slots, fold
//      temporaries, workspaces for a layer sub-batch of kMinLayerBatch, so
//      that the layer launches still fill the chip) fits the worker's
//      budget, up to max_batch;
//   3. run them as one batched launch sequence (layers of all admitted
//      requests in one layer batch; the fold batched over requests, each
//      request leaving the fold after its own last layer);
//   4. publish results and timestamps.
// While one worker's batch is on the GPU, the other worker admits the
// requests that arrived meanwhile and enqueues them on its own stream, so
// their work overlaps on the device.
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
};

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

  void worker(int wi) {
    Ctx& c = *ctxs[wi];
    (void)hipSetDevice(device);
    for (;;) {
      std::vector<Request*> batch;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !pending.empty(); });
        if (stop && pending.empty()) return;
        // oldest request decides the bucket (cancelled requests have already
        // left the queue: rcwa_server_cancel removes them)
        if (pending.empty()) continue;
        const auto key = pending.front()->key;
        const int maxb = cfg.max_batch > 0 ? cfg.max_batch : 1 << 30;
        int TL = 0, Lmax = 0;
        for (auto it = pending.begin(); it != pending.end() && (int)batch.size() < maxb;) {
          Request* r = *it;
          if (r->key != key) {
            ++it;
            continue;
          }
          const int L = r->dev.num_layers;
          const size_t need = chunk_bytes(
              r->N, TL + L, std::max(Lmax, L), r->dev.Nx, r->dev.Ny, r->par.Nx_harmonics,
              r->par.Ny_harmonics, true, (int)batch.size() + 1,
              std::max(1, std::min(TL + L, kMinLayerBatch)));
          if (!batch.empty() && need > c.budget) break;  // HBM admission
          TL += L;
          Lmax = std::max(Lmax, L);
          batch.push_back(r);
          it = pending.erase(it);
        }
        // stamped inside the admission critical section, so that cancel
        // (which reads these under the same lock) never sees a request that
        // is already in a batch as still queued
        const double ts = now();
        for (Request* r : batch) {
          r->taken = true;
          r->t_start = ts;
        }
      }
      std::vector<ProblemRef> refs(batch.size());
      for (size_t i = 0; i < batch.size(); ++i) {
        Request* r = batch[i];
        refs[i].p = rcwa_problem{&r->dev, &r->par, r->src.data(), (int)r->src.size()};
        refs[i].out = r->res.data();
      }
      rcwa_status st = solve_problems(c, refs.data(), (int)refs.size(), cfg.precision);
      std::string msg = st ? std::string(rcwa_last_error()) : std::string();
      const double tf = now();
      {
        std::lock_guard<std::mutex> lk(mu);
        for (Request* r : batch) {
          r->status = st;
          r->err = msg;
          r->t_finish = tf;
          r->batch_size = (int)batch.size();
          // grids are no longer needed once solved
          std::vector<rcwa_c128>().swap(r->er);
          std::vector<rcwa_c128>().swap(r->ur);
        }
      }
      cv_done.notify_all();
    }
  }
};

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
  if (d.Nx <= 0 || d.Ny <= 0 || d.num_layers < 0 || (d.num_layers > 0 && (!d.er || !d.ur || !d.t)))
    return rcwa_set_error(RCWA_E_ARG, "bad device");
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
    for (Request* r : s->pending) {
      r->cancelled = true;
      r->status = RCWA_E_CANCELLED;
    }
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
