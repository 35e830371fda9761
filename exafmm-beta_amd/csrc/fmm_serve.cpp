// fmm_serve.cpp -- request loop for streams of independent Laplace systems
// (successive MD steps, many small problems), on top of the C ABI.
//
//   submit ──► plan builders (host threads: bounds, tree, lists, M2L groups)
//                  │ ready requests, in arrival order
//                  ▼
//   slots (one context + HIP stream each): take the oldest ready requests as
//   one batch (up to batch_bodies bodies / batch_requests requests), wait
//   until the batch's device footprint fits the budget next to the batches in
//   flight, fuse the batch's plans into one forest (fmm_plan_merge) so each
//   phase is one launch for the whole batch, upload, evaluate, copy back.
//
// Requests arriving while others run are planned at once and join the next
// batch a slot takes: batches form continuously, and batches of different
// slots overlap on the device (separate streams, default priority).  A
// system's arithmetic does not depend on its batch: its cells, lists, M2L
// tiles and chunks are the same as alone; only the order of FP64 atomic
// additions into L may differ.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "exafmm_amd.h"
#include "fmm_host.hpp"

namespace {

double clock_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
  int64_t id = 0;
  int n = 0;
  std::vector<double> x, q, trg;
  fmm_plan *plan = nullptr;
  int64_t footprint = 0;
  int status = 0;  // 0 queued, 1 planned, 2 running, 3 done
  int rc = FMM_OK;
  std::string err;
  fmm_request_info info{};
};

}  // namespace

struct fmm_server {
  fmm_params p{};
  fmm_server_config cfg{};
  double t0 = 0;
  std::mutex mu;
  std::condition_variable cv_plan, cv_ready, cv_done;
  std::deque<std::shared_ptr<Request>> to_plan;
  std::map<int64_t, std::shared_ptr<Request>> ready;  // planned, by arrival
  std::map<int64_t, std::shared_ptr<Request>> all;    // until fetched by fmm_server_wait
  int64_t next_id = 1;
  int64_t inflight_bytes = 0;
  int inflight_batches = 0;
  int waiters = 0;  // threads inside fmm_server_wait
  bool stop = false;
  std::vector<std::thread> builders, slots;
  fmm_server_stats stats{};
};

namespace {

// device bytes a plan needs in a context (bodies, cells with M and L, lists,
// groups); the contexts' buffers are grow-only, so this bounds a batch's share
int64_t footprint(const fmm_params &p, const fmm_stats &s) {
  const int64_t nt = (int64_t)p.P * (p.P + 1) / 2;
  int64_t b = (int64_t)s.n * (8 * 8 + 4);
  b += (int64_t)s.ncell * (2 * nt * 16 + (p.precision != FMM_FP64 ? nt * 8 + 16 : 0) + 80);
  b += s.m2l_pairs * (4 + 8) + s.p2p_pairs * 4 + s.m2l_pairs / 32 * 16;
  return b;
}

void builder_loop(fmm_server *s) {
  fmm::set_host_threads(1);  // requests are planned in parallel, each on one thread
  // plan_on_device: this builder's own context (stream) for the tree and lists
  fmm_ctx *bc = nullptr;
  if (s->cfg.plan_on_device && fmm_create(&bc, &s->p) != FMM_OK) {
    fmm_destroy(bc);
    bc = nullptr;  // plan on the host instead
  }
  struct Release {
    fmm_ctx *c;
    ~Release() { fmm_destroy(c); }
  } release{bc};
  for (;;) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_plan.wait(lk, [&] { return s->stop || !s->to_plan.empty(); });
      if (s->stop) return;
      r = s->to_plan.front();
      s->to_plan.pop_front();
    }
    fmm_plan *pl = nullptr;
    int rc = bc ? fmm_plan_build_device(bc, &pl, r->n, r->x.data(), s->p.ncrit, s->p.theta, s->p.nspawn)
                : fmm_plan_build(&pl, r->n, r->x.data(), s->p.ncrit, s->p.theta, s->p.nspawn);
    // M2L groups for the slots' contexts; free-space tiles are left to the
    // upload, which builds them on the device (for a batch of several
    // requests the merged forest is grouped on the host there)
    if (rc == FMM_OK && !(s->p.m2l_mode == FMM_M2L_TILES && s->p.images == 0)) rc = fmm_plan_prepare(pl, &s->p);
    fmm_stats st{};
    if (rc == FMM_OK) fmm_plan_stats(pl, &st);
    std::lock_guard<std::mutex> lk(s->mu);
    r->info.t_planned_ms = clock_ms() - s->t0;
    if (rc != FMM_OK) {
      fmm_plan_destroy(pl);
      r->rc = rc;
      r->err = "planning failed";
      r->status = 3;
      s->cv_done.notify_all();
      continue;
    }
    r->plan = pl;
    r->footprint = footprint(s->p, st);
    r->status = 1;
    s->ready[r->id] = r;
    s->cv_ready.notify_all();
  }
}

void fail_batch(fmm_server *s, std::vector<std::shared_ptr<Request>> &batch, int rc, const std::string &err) {
  for (auto &r : batch) {
    r->rc = rc;
    r->err = err;
  }
}

void slot_loop(fmm_server *s, int slot) {
  fmm_ctx *c = nullptr;
  int crc = fmm_create(&c, &s->p);
  for (;;) {
    std::vector<std::shared_ptr<Request>> batch;
    int64_t bytes = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      // admission: the oldest ready request must fit the budget beside the
      // batches in flight (always admitted when nothing else runs)
      s->cv_ready.wait(lk, [&] {
        if (s->stop) return true;
        if (s->ready.empty()) return false;
        const auto &r = s->ready.begin()->second;
        return s->inflight_batches == 0 || s->inflight_bytes + r->footprint <= s->cfg.device_budget;
      });
      if (s->stop) break;
      if (s->cfg.batch_delay_ms > 0) {
        // hold the batch open until it is full or its oldest request has
        // waited batch_delay_ms since planning
        const double t_open = s->ready.begin()->second->info.t_planned_ms + s->cfg.batch_delay_ms;
        auto full = [&] {
          int64_t b = 0;
          int k = 0;
          for (auto &kv : s->ready) {
            b += kv.second->n;
            if (++k >= s->cfg.batch_requests || b >= s->cfg.batch_bodies) return true;
          }
          return false;
        };
        while (!s->stop && !s->ready.empty() && !full()) {
          const double left = t_open - (clock_ms() - s->t0);
          if (left <= 0) break;
          s->cv_ready.wait_for(lk, std::chrono::duration<double, std::milli>(left));
        }
        if (s->stop) break;
        if (s->ready.empty()) continue;  // another slot took them
      }
      int64_t bodies = 0;
      for (auto it = s->ready.begin(); it != s->ready.end();) {
        const auto &r = it->second;
        const bool first = batch.empty();
        if (!first && ((int)batch.size() >= s->cfg.batch_requests || bodies + r->n > s->cfg.batch_bodies ||
                       s->inflight_bytes + bytes + r->footprint > s->cfg.device_budget))
          break;
        batch.push_back(r);
        bodies += r->n;
        bytes += r->footprint;
        it = s->ready.erase(it);
      }
      s->inflight_bytes += bytes;
      s->inflight_batches++;
      const double t = clock_ms() - s->t0;
      for (auto &r : batch) {
        r->status = 2;
        r->info.t_started_ms = t;
        r->info.batch_requests = (int)batch.size();
        r->info.batch_bodies = (int)bodies;
        r->info.slot = slot;
      }
    }
    // run the batch
    int rc = crc;
    std::string err = crc == FMM_OK ? "" : (c ? fmm_last_error(c) : "fmm_create failed");
    fmm_plan *merged = nullptr;
    const fmm_plan *pl = batch[0]->plan;
    std::vector<double> xq, qq;
    const double *xp = batch[0]->x.data(), *qp = batch[0]->q.data();
    int ntot = batch[0]->n;
    if (rc == FMM_OK && batch.size() > 1) {
      std::vector<const fmm_plan *> plans;
      ntot = 0;
      for (auto &r : batch) {
        plans.push_back(r->plan);
        xq.insert(xq.end(), r->x.begin(), r->x.end());
        qq.insert(qq.end(), r->q.begin(), r->q.end());
        ntot += r->n;
      }
      rc = fmm_plan_merge(&merged, (int)plans.size(), plans.data());
      if (rc != FMM_OK) err = "fmm_plan_merge failed";
      pl = merged;
      xp = xq.data();
      qp = qq.data();
    }
    std::vector<double> trg(4 * (size_t)ntot);
    if (rc == FMM_OK) rc = fmm_set_bodies_plan(c, pl, xp, qp);
    if (rc == FMM_OK) rc = fmm_evaluate(c);
    if (rc == FMM_OK) rc = fmm_get_targets(c, trg.data(), 1);
    if (rc != FMM_OK && err.empty()) err = fmm_last_error(c);
    fmm_stats st{};
    if (rc == FMM_OK) fmm_get_stats(c, &st);
    if (merged) fmm_plan_destroy(merged);
    if (rc != FMM_OK) fail_batch(s, batch, rc, err);
    // results per request (bodies of a batch are its requests' arrays in order)
    size_t off = 0;
    for (auto &r : batch) {
      if (rc == FMM_OK) r->trg.assign(trg.begin() + 4 * off, trg.begin() + 4 * (off + r->n));
      off += r->n;
      fmm_plan_destroy(r->plan);
      r->plan = nullptr;
      std::vector<double>().swap(r->x);
      std::vector<double>().swap(r->q);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    const double t = clock_ms() - s->t0;
    for (auto &r : batch) {
      r->info.t_done_ms = t;
      r->info.device_ms = st.dev_total_ms;
      r->status = 3;
    }
    s->inflight_bytes -= bytes;
    s->inflight_batches--;
    s->stats.batches++;
    s->stats.requests_done += (int64_t)batch.size();
    s->stats.bodies_done += ntot;
    s->stats.device_ms += st.dev_total_ms;
    s->stats.max_batch_requests = std::max(s->stats.max_batch_requests, (int)batch.size());
    s->cv_done.notify_all();
    s->cv_ready.notify_all();
  }
  if (c) fmm_destroy(c);
}

}  // namespace

extern "C" {

void fmm_server_default_config(fmm_server_config *cfg) {
  cfg->host_threads = 0;
  cfg->slots = 8;  // with one request per batch on tiles: 673 / 715 / 746 req/s at 6 / 8 / 10 slots on
                   // 16 host cores (builders take the rest; at 10 they fall behind)
  cfg->device_budget = (int64_t)32 << 30;
  cfg->batch_bodies = 1 << 20;
  cfg->batch_requests = 1;  // a batch of several is merged and grouped on the host: slower
  cfg->batch_delay_ms = 0.0;
  cfg->plan_on_device = 0;
}

int fmm_server_create(fmm_server **out, const fmm_params *p, const fmm_server_config *cfg) {
  if (!out || !p) return FMM_ERR_ARG;
  *out = nullptr;
  fmm_server_config c;
  fmm_server_default_config(&c);
  if (cfg) c = *cfg;
  if (c.slots < 1 || c.slots > 16 || c.batch_bodies < 1 || c.batch_requests < 1 || c.device_budget < 1 ||
      c.host_threads < 0 || !(c.batch_delay_ms >= 0))
    return FMM_ERR_ARG;
  if (p->images != 0) return FMM_ERR_UNSUPPORTED;  // batches are forests of free-space systems
  // validate the parameters once with a context (reports bad P, theta, ...)
  fmm_ctx *probe = nullptr;
  const int rc = fmm_create(&probe, p);
  fmm_destroy(probe);
  if (rc != FMM_OK) return rc;
  if (c.host_threads == 0) c.host_threads = std::max(1, std::min(16, (int)std::thread::hardware_concurrency() - c.slots));
  fmm_server *s = new (std::nothrow) fmm_server();
  if (!s) return FMM_ERR_MEMORY;
  s->p = *p;
  // AUTO, one request per batch: target tiles.  The builders then plan only
  // the tree and lists; a slot's upload builds the groups and tiles on the
  // device (grouped chunks are built on the host by the builders, about half
  // of a request's planning).  At 1000 req/s offered, 2^12-2^18-body
  // requests: tiles 715 req/s, chunks 580 req/s (8 and 5 slots).  Batches of
  // several requests: grouped chunks, as a merged forest's tiles are grouped
  // on the host, where mixed precision groups by level across systems of
  // different radii and runs the pairs left over on the per-target kernel
  // (results then depend on the batch, about 1e-8).  Split-FP16 runs on
  // tiles only.
  if (s->p.m2l_mode == FMM_M2L_AUTO && s->p.precision != FMM_MIXED16)
    s->p.m2l_mode = c.batch_requests == 1 ? FMM_M2L_TILES : FMM_M2L_GROUPED;
  s->cfg = c;
  s->t0 = clock_ms();
  for (int i = 0; i < c.host_threads; i++) s->builders.emplace_back(builder_loop, s);
  for (int i = 0; i < c.slots; i++) s->slots.emplace_back(slot_loop, s, i);
  *out = s;
  return FMM_OK;
}

int fmm_server_submit(fmm_server *s, int n, const double *x, const double *q, int64_t *id) {
  if (!s || !id || n < 0 || (n > 0 && (!x || !q))) return FMM_ERR_ARG;
  auto r = std::make_shared<Request>();
  r->n = n;
  r->x.assign(x, x + 3 * (size_t)n);
  r->q.assign(q, q + (size_t)n);
  r->info.n = n;
  std::lock_guard<std::mutex> lk(s->mu);
  r->id = s->next_id++;
  r->info.t_submit_ms = clock_ms() - s->t0;
  s->all[r->id] = r;
  s->to_plan.push_back(r);
  s->stats.requests_submitted++;
  *id = r->id;
  s->cv_plan.notify_one();
  return FMM_OK;
}

int fmm_server_wait(fmm_server *s, int64_t id, double *trg, fmm_request_info *info) {
  if (!s) return FMM_ERR_ARG;
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return FMM_ERR_ARG;
    r = it->second;
    s->waiters++;
    s->cv_done.wait(lk, [&] { return r->status == 3; });
    s->waiters--;
    s->all.erase(id);
    if (s->stop) s->cv_done.notify_all();  // fmm_server_destroy waits for the last waiter
  }
  if (info) *info = r->info;
  if (r->rc != FMM_OK) return r->rc;
  if (trg && r->n > 0) std::memcpy(trg, r->trg.data(), sizeof(double) * 4 * (size_t)r->n);
  return FMM_OK;
}

int fmm_server_get_stats(fmm_server *s, fmm_server_stats *out) {
  if (!s || !out) return FMM_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  return FMM_OK;
}

void fmm_server_destroy(fmm_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_plan.notify_all();
  s->cv_ready.notify_all();
  for (auto &t : s->builders) t.join();
  for (auto &t : s->slots) t.join();
  {
    // requests that never ran end with FMM_ERR_STATE; waits blocked on them
    // return before the server's mutex and condition variables go away
    std::unique_lock<std::mutex> lk(s->mu);
    for (auto &kv : s->all)
      if (kv.second->status != 3) {
        kv.second->rc = FMM_ERR_STATE;
        kv.second->err = "server destroyed before the request ran";
        kv.second->status = 3;
      }
    s->cv_done.notify_all();
    s->cv_done.wait(lk, [&] { return s->waiters == 0; });
  }
  for (auto &kv : s->ready) fmm_plan_destroy(kv.second->plan);
  delete s;
}

}  // extern "C"
