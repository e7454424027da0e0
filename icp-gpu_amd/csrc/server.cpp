// Serving loop (include/blsgpu_serve.h): continuous batching of verify requests over a
// fixed set of device slots.
//
// Life of a request: submit copies its inputs and queues it; the scheduler thread packs
// queued requests in arrival order into the next free slot (up to the slot's item
// capacity), copies them into the slot's pinned staging, and launches on the slot's
// stream: inputs H2D, the verification sequence (decode, -H(m) conversion, subgroup test
// on the slot's second stream, queue-fed Miller loop and final exponentiation), results
// D2H, and a completion event.  The scheduler polls the events of busy slots; when one
// has fired it stamps the completion time on each request of the step and frees the
// slot.  Steps of different slots run concurrently on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/blsgpu_serve.h"
#include "../../include/gpu_msm.h"
#include "context.hpp"
#include "kernels.hpp"
#include "verify_engine.hpp"

using namespace blsctx;

namespace {

using Clock = std::chrono::steady_clock;
double now_ms() { return std::chrono::duration<double, std::milli>(Clock::now().time_since_epoch()).count(); }

enum { Q_QUEUED = 0, Q_RUNNING = 1, Q_DONE = 2, Q_FAILED = 3 };

struct Request {
  int64_t id = 0;
  int n = 0;
  std::vector<uint8_t> sig;
  std::vector<uint64_t> nhm;
  std::vector<int32_t> key;
  std::vector<int32_t> res;
  std::vector<uint64_t> gt;
  double t_submit = 0, t_done = 0;
  int state = Q_QUEUED;
};

// bytes of device workspace one item needs in a slot
size_t item_bytes(bool want_gt) {
  return 48 + 96 + 4 + 2 * bls::POINT_WORDS * 4 + 4 + 4 + bls::MILLER_WORDS * 4 + 4 + (want_gt ? 576 : 0);
}

struct Slot {
  hipStream_t s = nullptr, s2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, done = nullptr;
  DevBuf comp, nhm, key, sig_pts, nhm_pts, status, sub, fbuf, res, gt, queue;
  PinnedBuf h_comp, h_nhm, h_key, h_res, h_gt;
  std::vector<Request*> reqs;
  int n = 0;
  bool busy = false;
};

}  // namespace

struct blsgpu_server {
  int cap = 0, max_reqs = 0;
  bool want_gt = false;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> pending;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
  std::vector<Slot> slots;
  std::thread th;
  bool stop = false;
  int64_t next_id = 0;
  double st_steps = 0, st_items = 0, st_completed = 0, st_max_step = 0, st_max_inflight = 0, st_failed = 0;

  int setup_slot(Slot& sl);
  static void release_slot(Slot& sl);
  int launch(Slot& sl);
  void finish(Slot& sl, bool ok);
  void loop();
};

int blsgpu_server::setup_slot(Slot& sl) {
  const size_t m = (size_t)cap;
  HIPCHK(hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&sl.s2, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&sl.fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&sl.join, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
  CHK(sl.comp.ensure(m * 48));
  CHK(sl.nhm.ensure(m * 96));
  CHK(sl.key.ensure(m * 4));
  CHK(sl.sig_pts.ensure(m * bls::POINT_WORDS * 4));
  CHK(sl.nhm_pts.ensure(m * bls::POINT_WORDS * 4));
  CHK(sl.status.ensure(m * 4));
  CHK(sl.sub.ensure(m * 4));
  CHK(sl.fbuf.ensure(m * bls::MILLER_WORDS * 4));
  CHK(sl.res.ensure(m * 4));
  if (want_gt) CHK(sl.gt.ensure(m * 576));
  CHK(sl.queue.ensure(64));
  CHK(sl.h_comp.ensure(m * 48));
  CHK(sl.h_nhm.ensure(m * 96));
  CHK(sl.h_key.ensure(m * 4));
  CHK(sl.h_res.ensure(m * 4));
  if (want_gt) CHK(sl.h_gt.ensure(m * 576));
  return 0;
}

// Frees what setup_slot made of the slot, once its streams are idle (also for a slot
// whose setup stopped part way).
void blsgpu_server::release_slot(Slot& sl) {
  if (sl.s) (void)hipStreamSynchronize(sl.s);
  if (sl.s2) (void)hipStreamSynchronize(sl.s2);
  for (DevBuf* b : {&sl.comp, &sl.nhm, &sl.key, &sl.sig_pts, &sl.nhm_pts, &sl.status, &sl.sub, &sl.fbuf, &sl.res, &sl.gt,
                    &sl.queue})
    if (b->p) (void)hipFree(b->p);
  for (PinnedBuf* b : {&sl.h_comp, &sl.h_nhm, &sl.h_key, &sl.h_res, &sl.h_gt})
    if (b->p) (void)hipHostFree(b->p);
  if (sl.s) (void)hipStreamDestroy(sl.s);
  if (sl.s2) (void)hipStreamDestroy(sl.s2);
  for (hipEvent_t e : {sl.fork, sl.join, sl.done})
    if (e) (void)hipEventDestroy(e);
  sl = Slot();
}

// Packs queued requests into the slot and launches its step (caller holds mu).
int blsgpu_server::launch(Slot& sl) {
  sl.reqs.clear();
  sl.n = 0;
  while (!pending.empty() && sl.n + pending.front()->n <= cap && (max_reqs <= 0 || (int)sl.reqs.size() < max_reqs)) {
    Request* r = pending.front();
    pending.pop_front();
    memcpy(sl.h_comp.as<uint8_t>() + (size_t)sl.n * 48, r->sig.data(), (size_t)r->n * 48);
    memcpy(sl.h_nhm.as<uint8_t>() + (size_t)sl.n * 96, r->nhm.data(), (size_t)r->n * 96);
    memcpy(sl.h_key.as<int32_t>() + sl.n, r->key.data(), (size_t)r->n * 4);
    r->state = Q_RUNNING;
    sl.reqs.push_back(r);
    sl.n += r->n;
  }
  if (sl.n == 0) return 0;
  const size_t m = (size_t)sl.n;
  std::lock_guard<std::mutex> lk(g.mu);  // key tables, configuration, launch-path caches
  HIPCHK(hipMemcpyAsync(sl.comp.p, sl.h_comp.p, m * 48, hipMemcpyHostToDevice, sl.s));
  HIPCHK(hipMemcpyAsync(sl.nhm.p, sl.h_nhm.p, m * 96, hipMemcpyHostToDevice, sl.s));
  HIPCHK(hipMemcpyAsync(sl.key.p, sl.h_key.p, m * 4, hipMemcpyHostToDevice, sl.s));
  const BatchedCfg c = cfg_for_locked(sl.n);
  if (bls::occupancy_grp(0, c.g_ml) <= 0 || bls::occupancy_grp(1, c.g_fe) <= 0) {
    g_err = "occupancy query failed";
    return -1;
  }
  bls::EngineArgs a{};
  a.M = sl.n;
  a.mode = 1;
  a.g_ml = c.g_ml;
  a.g_fe = c.g_fe;
  a.wpc_ml = c.wpc_ml;
  a.wpc_fe = c.wpc_fe;
  a.depth_ml = c.depth_ml;
  a.queue = sl.queue.as<int>();
  a.comp = sl.comp.as<uint8_t>();
  a.nhm_ref = sl.nhm.as<uint64_t>();
  a.key_idx = sl.key.as<int32_t>();
  a.keys = g.keys.as<int32_t>();
  a.nkeys = g.nkeys;
  a.gen_lines = g.gen_lines.as<int32_t>();
  a.sig_pts = sl.sig_pts.as<int32_t>();
  a.nhm_pts = sl.nhm_pts.as<int32_t>();
  a.status = sl.status.as<int32_t>();
  a.sub_status = sl.sub.as<int32_t>();
  a.s2 = sl.s2;
  a.fork = sl.fork;
  a.join = sl.join;
  a.fbuf = sl.fbuf.as<int32_t>();
  a.results = sl.res.as<int32_t>();
  a.gt = want_gt ? sl.gt.as<uint64_t>() : nullptr;
  CHK(bls::engine_compressed(a, sl.s));
  HIPCHK(hipMemcpyAsync(sl.h_res.p, sl.res.p, m * 4, hipMemcpyDeviceToHost, sl.s));
  if (want_gt) HIPCHK(hipMemcpyAsync(sl.h_gt.p, sl.gt.p, m * 576, hipMemcpyDeviceToHost, sl.s));
  HIPCHK(hipEventRecord(sl.done, sl.s));
  sl.busy = true;
  st_steps += 1;
  st_items += sl.n;
  st_max_step = std::max(st_max_step, (double)sl.n);
  return 0;
}

// Hands the step's results to its requests (caller holds mu).
void blsgpu_server::finish(Slot& sl, bool ok) {
  const double t = now_ms();
  int o = 0;
  for (Request* r : sl.reqs) {
    if (ok) {
      r->res.assign(sl.h_res.as<int32_t>() + o, sl.h_res.as<int32_t>() + o + r->n);
      if (want_gt) r->gt.assign(sl.h_gt.as<uint64_t>() + (size_t)o * 72, sl.h_gt.as<uint64_t>() + (size_t)(o + r->n) * 72);
    }
    r->t_done = t;
    r->state = ok ? Q_DONE : Q_FAILED;
    o += r->n;
    if (ok) st_completed += 1;
  }
  if (!ok) st_failed += 1;
  sl.reqs.clear();
  sl.n = 0;
  sl.busy = false;
  cv_done.notify_all();
}

void blsgpu_server::loop() {
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    int inflight = 0;
    for (Slot& sl : slots) {
      if (!sl.busy) continue;
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipSuccess) finish(sl, true);
      else if (q != hipErrorNotReady) finish(sl, false);
      else inflight++;
    }
    if (stop && inflight == 0) break;
    if (!stop) {
      for (Slot& sl : slots) {
        if (sl.busy || pending.empty()) continue;
        if (launch(sl) != 0) {
          // the step never reached the GPU in full: wait for what was queued, then fail it
          (void)hipStreamSynchronize(sl.s);
          (void)hipStreamSynchronize(sl.s2);
          finish(sl, false);
        } else if (sl.busy) {
          inflight++;
        }
      }
      st_max_inflight = std::max(st_max_inflight, (double)inflight);
    }
    if (inflight == 0 && pending.empty()) {
      if (stop) break;
      cv_work.wait(lk, [&] { return stop || !pending.empty(); });
    } else {
      cv_work.wait_for(lk, std::chrono::microseconds(25));
    }
  }
}

extern "C" {

blsgpu_server* blsgpu_server_create(int slots, int max_step_items, int64_t hbm_budget_bytes, int max_requests_per_step,
                                    int want_gt) {
  g_err.clear();
  if (slots < 1 || slots > 16 || max_step_items < 1 || hbm_budget_bytes <= 0 || max_requests_per_step < 0) {
    g_err = "invalid arguments";
    return nullptr;
  }
  {
    std::lock_guard<std::mutex> lk(g.mu);
    if (init_locked() != 0) return nullptr;
  }
  std::unique_ptr<blsgpu_server> srv(new blsgpu_server());
  srv->want_gt = want_gt != 0;
  srv->max_reqs = max_requests_per_step;
  const int64_t by_budget = hbm_budget_bytes / ((int64_t)slots * (int64_t)item_bytes(srv->want_gt));
  srv->cap = (int)std::min<int64_t>(max_step_items, by_budget);
  if (srv->cap < 1) {
    g_err = "HBM budget below one item per slot";
    return nullptr;
  }
  srv->slots.resize(slots);
  for (Slot& sl : srv->slots) {
    if (srv->setup_slot(sl) != 0) {
      for (Slot& b : srv->slots) blsgpu_server::release_slot(b);
      return nullptr;
    }
  }
  srv->th = std::thread([s = srv.get()] { s->loop(); });
  return srv.release();
}

int blsgpu_server_step_capacity(blsgpu_server* srv) { return srv ? srv->cap : -1; }

// t_arrival < 0: now.  The load generator passes the planned arrival time, so that its
// own lag behind the schedule counts as latency.
static int64_t submit_impl(blsgpu_server* srv, const uint8_t* sigs48, const uint64_t* neg_hms, const int32_t* key_idx,
                           int n, double t_arrival) {
  if (!srv || !sigs48 || !neg_hms || !key_idx || n < 1 || n > srv->cap) return -1;
  int nkeys;
  {
    std::lock_guard<std::mutex> lk(g.mu);
    nkeys = g.nkeys;
  }
  for (int i = 0; i < n; i++)
    if (key_idx[i] < 0 || key_idx[i] >= nkeys) return -1;
  std::unique_ptr<Request> r(new Request());
  r->n = n;
  r->sig.assign(sigs48, sigs48 + (size_t)n * 48);
  r->nhm.assign(neg_hms, neg_hms + (size_t)n * 12);
  r->key.assign(key_idx, key_idx + n);
  std::lock_guard<std::mutex> lk(srv->mu);
  if (srv->stop) return -1;
  r->id = srv->next_id++;
  r->t_submit = t_arrival < 0 ? now_ms() : t_arrival;
  const int64_t id = r->id;
  srv->pending.push_back(r.get());
  srv->reqs.emplace(id, std::move(r));
  srv->cv_work.notify_one();
  return id;
}

int64_t blsgpu_server_submit(blsgpu_server* srv, const uint8_t* sigs48, const uint64_t* neg_hms,
                             const int32_t* key_idx, int n) {
  return submit_impl(srv, sigs48, neg_hms, key_idx, n, -1.0);
}

static int take_locked(blsgpu_server* srv, int64_t id, int32_t* results, uint64_t* gt, double* latency_ms) {
  auto it = srv->reqs.find(id);
  if (it == srv->reqs.end()) return -1;
  Request* r = it->second.get();
  if (r->state == Q_QUEUED || r->state == Q_RUNNING) return 0;
  const bool ok = r->state == Q_DONE;
  if (ok) {
    if (results) memcpy(results, r->res.data(), (size_t)r->n * 4);
    if (gt && srv->want_gt) memcpy(gt, r->gt.data(), (size_t)r->n * 576);
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
  }
  srv->reqs.erase(it);
  return ok ? 1 : -1;
}

int blsgpu_server_poll(blsgpu_server* srv, int64_t id, int32_t* results, uint64_t* gt, double* latency_ms) {
  if (!srv) return -1;
  std::lock_guard<std::mutex> lk(srv->mu);
  return take_locked(srv, id, results, gt, latency_ms);
}

int blsgpu_server_wait(blsgpu_server* srv, int64_t id, int32_t* results, uint64_t* gt, double* latency_ms,
                       double timeout_ms) {
  if (!srv) return -1;
  std::unique_lock<std::mutex> lk(srv->mu);
  const auto until = Clock::now() + std::chrono::microseconds((int64_t)(std::max(timeout_ms, 0.0) * 1000));
  for (;;) {
    const int rc = take_locked(srv, id, results, gt, latency_ms);
    if (rc != 0) return rc;
    if (timeout_ms < 0) srv->cv_done.wait(lk);
    else if (srv->cv_done.wait_until(lk, until) == std::cv_status::timeout) return take_locked(srv, id, results, gt, latency_ms);
  }
}

int blsgpu_server_stats(blsgpu_server* srv, double out[8]) {
  if (!srv || !out) return -1;
  std::lock_guard<std::mutex> lk(srv->mu);
  out[0] = srv->st_steps;
  out[1] = srv->st_items;
  out[2] = srv->st_completed;
  out[3] = srv->st_max_step;
  out[4] = srv->st_max_inflight;
  out[5] = srv->cap;
  out[6] = (double)srv->slots.size();
  out[7] = srv->st_failed;
  return 0;
}

void blsgpu_server_destroy(blsgpu_server* srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
    for (Request* r : srv->pending) r->state = Q_FAILED;
    srv->pending.clear();
    srv->cv_work.notify_all();
  }
  if (srv->th.joinable()) srv->th.join();
  for (Slot& sl : srv->slots) blsgpu_server::release_slot(sl);
  delete srv;
}

int blsgpu_server_loadtest(blsgpu_server* srv, const uint8_t* pool_sigs48, const uint64_t* pool_neg_hms,
                           const int32_t* pool_key_idx, int pool_n, double items_per_s, double seconds,
                           double drain_s, double size_median, double size_sigma, int size_min, int size_max,
                           uint64_t seed, double out[11]) {
  g_err.clear();
  if (!srv || !pool_sigs48 || !pool_neg_hms || !pool_key_idx || pool_n < 1 || items_per_s <= 0 || seconds <= 0 ||
      size_median <= 0 || size_min < 1 || size_max < size_min || size_max > srv->cap || !out) {
    g_err = "invalid arguments";
    return -1;
  }
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  auto draw = [&]() {
    const double v = std::round(size_median * std::exp(size_sigma * nd(rng)));
    return (int)std::min<double>(size_max, std::max<double>(size_min, v));
  };
  // mean request size of this distribution, for the request arrival rate
  double mean_n = 0;
  {
    std::mt19937_64 r2(seed ^ 0x9e3779b97f4a7c15ULL);
    std::normal_distribution<double> n2(0.0, 1.0);
    const int K = 200000;
    for (int k = 0; k < K; k++) {
      const double v = std::round(size_median * std::exp(size_sigma * n2(r2)));
      mean_n += std::min<double>(size_max, std::max<double>(size_min, v));
    }
    mean_n /= K;
  }
  std::exponential_distribution<double> gap(items_per_s / mean_n / 1000.0);  // per ms
  // plan: arrival times and sizes
  std::vector<double> at;
  std::vector<int> sz;
  for (double t = gap(rng); t < seconds * 1000.0; t += gap(rng)) {
    at.push_back(t);
    sz.push_back(draw());
  }
  std::vector<uint8_t> sig;
  std::vector<uint64_t> nhm;
  std::vector<int32_t> key;
  std::vector<int64_t> ids(at.size(), -1);
  double items_sub = 0;
  int64_t off = 0;
  const double t0 = now_ms();
  for (size_t k = 0; k < at.size(); k++) {
    const int n = sz[k];
    sig.resize((size_t)n * 48);
    nhm.resize((size_t)n * 12);
    key.resize(n);
    for (int i = 0; i < n; i++) {
      const int64_t p = (off + i) % pool_n;
      memcpy(&sig[(size_t)i * 48], pool_sigs48 + p * 48, 48);
      memcpy(&nhm[(size_t)i * 12], pool_neg_hms + p * 12, 96);
      key[i] = pool_key_idx[p];
    }
    off = (off + n) % pool_n;
    // wait for this request's arrival time (sleep for long gaps, spin for short ones)
    for (;;) {
      const double dt = at[k] - (now_ms() - t0);
      if (dt <= 0) break;
      if (dt > 0.3) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)((dt - 0.2) * 1000)));
    }
    ids[k] = submit_impl(srv, sig.data(), nhm.data(), key.data(), n, t0 + at[k]);
    if (ids[k] < 0) {
      g_err = "submit refused";
      return -1;
    }
    items_sub += n;
  }
  const double t_sub_end = now_ms();
  const double deadline = t0 + (seconds + drain_s) * 1000.0;
  std::vector<double> lat(at.size(), std::numeric_limits<double>::infinity());
  double done_req = 0, done_items = 0, last_done = t0;
  for (size_t k = 0; k < at.size(); k++) {
    const double left = deadline - now_ms();
    double l = 0;
    const int rc = blsgpu_server_wait(srv, ids[k], nullptr, nullptr, &l, left > 0 ? left : 0);
    if (rc == 1) {
      lat[k] = l;
      done_req += 1;
      done_items += sz[k];
      last_done = std::max(last_done, now_ms());
    }
  }
  std::vector<double> sorted = lat;
  std::sort(sorted.begin(), sorted.end());
  auto pct = [&](double q) {
    if (sorted.empty()) return 0.0;
    size_t i = (size_t)std::ceil(q * sorted.size()) - 1;
    if (i >= sorted.size()) i = sorted.size() - 1;
    return sorted[i];
  };
  const double span = (last_done - t0) / 1000.0;
  out[0] = (double)at.size();
  out[1] = items_sub;
  out[2] = done_req;
  out[3] = done_items;
  out[4] = span;
  out[5] = span > 0 ? done_items / span : 0;
  out[6] = pct(0.50);
  out[7] = pct(0.99);
  out[8] = pct(0.999);
  out[9] = sorted.empty() ? 0 : sorted.back();
  out[10] = items_sub / std::max(1e-9, (t_sub_end - t0) / 1000.0);
  return 0;
}

}  // extern "C"
