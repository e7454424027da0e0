// Batched form of the verifier: many verify batches of uneven sizes in one submission.
//
// The items of all batches are concatenated and handed to one kernel sequence whose
// Miller-loop and final-exponentiation kernels are persistent: a grid sized to the
// device (workgroups per CU from the configuration, capped by occupancy) pops tiles of
// items from a device counter until none are left, so a batch of 3 items costs no
// launch of its own and the tail of one batch overlaps the head of the next.  The
// sequence (counter reset, decode, -H(m) conversion, subgroup test on a second stream,
// Miller loop, final exponentiation) is captured once as a hipGraph per input shape and
// replayed.  The lanes per item and workgroups per CU of each kernel come from
// blsgpu_set_batched_engine or from the autotuner (blsgpu_autotune_device), which times
// the candidates on a real job and keeps the winner per batch-size class.
//
// The reference has no batched form: its only entry point launches one kernel per call
// (gpu-crypto/cuda/api.cu:128-160), which the plain form here keeps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/gpu_msm.h"
#include "context.hpp"
#include "kernels.hpp"
#include "verify_engine.hpp"

namespace blsctx {

bool GraphKey::operator==(const GraphKey& o) const {
  for (int i = 0; i < 3; i++)
    if (in[i] != o.in[i]) return false;
  for (int i = 0; i < 2; i++)
    if (out[i] != o.out[i]) return false;
  for (int i = 0; i < 7; i++)
    if (ws[i] != o.ws[i]) return false;
  return keys == o.keys && nkeys == o.nkeys && M == o.M && cfg.g_ml == o.cfg.g_ml && cfg.g_fe == o.cfg.g_fe &&
         cfg.wpc_ml == o.cfg.wpc_ml && cfg.wpc_fe == o.cfg.wpc_fe && cfg.depth_ml == o.cfg.depth_ml;
}

namespace {

PinnedBuf h_comp, h_nhm, h_keyidx, h_res, h_gt;
hipEvent_t ev_in = nullptr, ev_out = nullptr;

int size_class(int M) {
  int c = 0;
  while ((1 << c) < M) c++;
  return c;
}

}  // namespace

BatchedCfg cfg_for_locked(int M) {
  BatchedCfg r = g.bcfg;
  if (!g.tuned.empty()) {
    const int c = size_class(M);
    auto best = g.tuned.begin();
    for (auto it = g.tuned.begin(); it != g.tuned.end(); ++it)
      if (std::abs(it->first - c) < std::abs(best->first - c)) best = it;
    r = best->second;
  }
  if (g.force_depth >= 0) r.depth_ml = g.force_depth;
  return r;
}

namespace {

bool valid_g_ml(int G) { return bls::grp_lanes_supported(0, G); }
bool valid_g_fe(int G) { return bls::grp_lanes_supported(1, G); }

int ensure_batched_ws(int M) {
  const size_t m = (size_t)M;
  CHK(g.b_sig_pts.ensure(m * bls::POINT_WORDS * 4));
  CHK(g.b_nhm_pts.ensure(m * bls::POINT_WORDS * 4));
  CHK(g.b_status.ensure(m * 4));
  CHK(g.b_sub.ensure(m * 4));
  CHK(g.b_fbuf.ensure(m * bls::MILLER_WORDS * 4));
  CHK(g.b_queue.ensure(64));
  CHK(g.b_park.ensure((size_t)bls::final_exp_park_words() * 4));
  return 0;
}

bls::EngineArgs batched_args(const uint8_t* comp, const uint64_t* nhm, const int32_t* keyidx, int32_t* res,
                             uint64_t* gt, int M, const BatchedCfg& c) {
  bls::EngineArgs a{};
  a.M = M;
  a.mode = 1;
  a.g_ml = c.g_ml;
  a.g_fe = c.g_fe;
  a.wpc_ml = c.wpc_ml;
  a.wpc_fe = c.wpc_fe;
  a.depth_ml = c.depth_ml;
  a.queue = g.b_queue.as<int>();
  a.comp = comp;
  a.nhm_ref = nhm;
  a.key_idx = keyidx;
  a.keys = g.keys.as<int32_t>();
  a.nkeys = g.nkeys;
  a.gen_lines = g.gen_lines.as<int32_t>();
  a.sig_pts = g.b_sig_pts.as<int32_t>();
  a.nhm_pts = g.b_nhm_pts.as<int32_t>();
  a.status = g.b_status.as<int32_t>();
  a.sub_status = g.b_sub.as<int32_t>();
  a.s2 = g.s2;
  a.fork = g.ev_fork;
  a.join = g.ev_join;
  a.fbuf = g.b_fbuf.as<int32_t>();
  a.park = g.b_park.p ? g.b_park.as<int32_t>() : nullptr;
  a.park_items = bls::final_exp_park_words() / (6 * 168);
  a.results = res;
  a.gt = gt;
  return a;
}

// Orders the library stream after the caller's stream (and back, in stream_return).
int stream_enter(hipStream_t caller) {
  if (!ev_in) HIPCHK(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
  if (!ev_out) HIPCHK(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
  if (caller == g.s) return 0;
  HIPCHK(hipEventRecord(ev_in, caller));
  HIPCHK(hipStreamWaitEvent(g.s, ev_in, 0));
  return 0;
}
int stream_return(hipStream_t caller) {
  if (caller == g.s) return 0;
  HIPCHK(hipEventRecord(ev_out, g.s));
  HIPCHK(hipStreamWaitEvent(caller, ev_out, 0));
  return 0;
}

// Runs the batched sequence on g.s for device-resident inputs, through the cached graph
// when graphs are enabled.
int run_batched_locked(const uint8_t* comp, const uint64_t* nhm, const int32_t* keyidx, int32_t* res, uint64_t* gt,
                       int M) {
  CHK(ensure_batched_ws(M));
  const BatchedCfg c = cfg_for_locked(M);
  const bls::EngineArgs a = batched_args(comp, nhm, keyidx, res, gt, M, c);
  // occupancy queries happen here, outside any capture
  if (bls::occupancy_grp(0, c.g_ml) <= 0 || bls::occupancy_grp(1, c.g_fe) <= 0) {
    g_err = "occupancy query failed";
    return -1;
  }
  if (!g.use_graph) return bls::engine_compressed(a, g.s);
  GraphKey k{};
  k.in[0] = comp;
  k.in[1] = nhm;
  k.in[2] = keyidx;
  k.out[0] = res;
  k.out[1] = gt;
  k.ws[0] = a.sig_pts;
  k.ws[1] = a.nhm_pts;
  k.ws[2] = a.status;
  k.ws[3] = a.sub_status;
  k.ws[4] = a.fbuf;
  k.ws[5] = a.queue;
  k.ws[6] = a.gen_lines;
  k.keys = a.keys;
  k.nkeys = a.nkeys;
  k.M = M;
  k.cfg = c;
  if (!g.have_graph || !(k == g.gkey)) {
    if (g.gexec) {
      HIPCHK(hipGraphExecDestroy(g.gexec));
      g.gexec = nullptr;
    }
    g.have_graph = false;
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(g.s, hipStreamCaptureModeThreadLocal));
    const int rc = bls::engine_compressed(a, g.s);
    const hipError_t ec = hipStreamEndCapture(g.s, &graph);
    if (rc != 0 || ec != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      g_err = "graph capture of the batched sequence failed";
      return -1;
    }
    const hipError_t ei = hipGraphInstantiate(&g.gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(ei);
    g.gkey = k;
    g.have_graph = true;
    g.graph_builds++;
  }
  HIPCHK(hipGraphLaunch(g.gexec, g.s));
  g.graph_launches++;
  return 0;
}

int check_offsets(const int32_t* off, int nb, int* M) {
  if (!off || nb <= 0 || off[0] != 0) return -1;
  for (int b = 0; b < nb; b++)
    if (off[b + 1] < off[b]) return -1;
  *M = off[nb];
  return *M > 0 ? 0 : -1;
}

}  // namespace
}  // namespace blsctx

using namespace blsctx;

extern "C" {

int blsgpu_verify_batches_device(const uint8_t* d_sigs48, const uint64_t* d_neg_hms, const int32_t* d_key_idx,
                                 const int32_t* batch_off, int nbatches, int32_t* d_results, uint64_t* d_gt,
                                 void* stream) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  int M = 0;
  if (!d_sigs48 || !d_neg_hms || !d_key_idx || !d_results || check_offsets(batch_off, nbatches, &M)) {
    g_err = "invalid arguments";
    return -1;
  }
  CHK(init_locked());
  const hipStream_t caller = stream ? (hipStream_t)stream : g.s;
  CHK(claim_workspace_locked(g.s));
  CHK(stream_enter(caller));
  CHK(run_batched_locked(d_sigs48, d_neg_hms, d_key_idx, d_results, d_gt, M));
  return stream_return(caller);
}

int blsgpu_verify_batches(const uint8_t* const* sigs48, const uint64_t* const* neg_hms, const int32_t* const* key_idx,
                          const int32_t* sizes, int nbatches, int32_t* const* results, uint64_t* const* gt_out) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (!sigs48 || !neg_hms || !key_idx || !sizes || !results || nbatches <= 0) {
    g_err = "invalid arguments";
    return -1;
  }
  size_t M = 0;
  for (int b = 0; b < nbatches; b++) {
    if (sizes[b] < 0 || (sizes[b] > 0 && (!sigs48[b] || !neg_hms[b] || !key_idx[b] || !results[b]))) {
      g_err = "invalid arguments";
      return -1;
    }
    M += (size_t)sizes[b];
  }
  if (M == 0 || M > (size_t)INT32_MAX / 72) {
    g_err = "invalid arguments";
    return -1;
  }
  CHK(init_locked());
  CHK(claim_workspace_locked(g.s));
  const bool want_gt = gt_out != nullptr;
  CHK(h_comp.ensure(M * 48));
  CHK(h_nhm.ensure(M * 96));
  CHK(h_keyidx.ensure(M * 4));
  CHK(h_res.ensure(M * 4));
  if (want_gt) CHK(h_gt.ensure(M * 576));
  size_t o = 0;
  for (int b = 0; b < nbatches; b++) {
    const size_t n = (size_t)sizes[b];
    if (!n) continue;
    for (size_t i = 0; i < n; i++)
      if (key_idx[b][i] < 0 || key_idx[b][i] >= g.nkeys) {
        g_err = "key index out of range";
        return -1;
      }
    memcpy(h_comp.as<uint8_t>() + o * 48, sigs48[b], n * 48);
    memcpy(h_nhm.as<uint8_t>() + o * 96, neg_hms[b], n * 96);
    memcpy(h_keyidx.as<int32_t>() + o, key_idx[b], n * 4);
    o += n;
  }
  CHK(g.b_comp.ensure(M * 48));
  CHK(g.b_nhm.ensure(M * 96));
  CHK(g.b_keyidx.ensure(M * 4));
  CHK(g.b_res.ensure(M * 4));
  if (want_gt) CHK(g.b_gt.ensure(M * 576));
  HIPCHK(hipMemcpyAsync(g.b_comp.p, h_comp.p, M * 48, hipMemcpyHostToDevice, g.s));
  HIPCHK(hipMemcpyAsync(g.b_nhm.p, h_nhm.p, M * 96, hipMemcpyHostToDevice, g.s));
  HIPCHK(hipMemcpyAsync(g.b_keyidx.p, h_keyidx.p, M * 4, hipMemcpyHostToDevice, g.s));
  CHK(run_batched_locked(g.b_comp.as<uint8_t>(), g.b_nhm.as<uint64_t>(), g.b_keyidx.as<int32_t>(),
                         g.b_res.as<int32_t>(), want_gt ? g.b_gt.as<uint64_t>() : nullptr, (int)M));
  HIPCHK(hipMemcpyAsync(h_res.p, g.b_res.p, M * 4, hipMemcpyDeviceToHost, g.s));
  if (want_gt) HIPCHK(hipMemcpyAsync(h_gt.p, g.b_gt.p, M * 576, hipMemcpyDeviceToHost, g.s));
  HIPCHK(hipStreamSynchronize(g.s));
  o = 0;
  for (int b = 0; b < nbatches; b++) {
    const size_t n = (size_t)sizes[b];
    if (!n) continue;
    memcpy(results[b], h_res.as<int32_t>() + o, n * 4);
    if (want_gt && gt_out[b]) memcpy(gt_out[b], h_gt.as<uint8_t>() + o * 576, n * 576);
    o += n;
  }
  return 0;
}

int blsgpu_set_batched_engine(int g_miller, int g_final_exp, int wpc_miller, int wpc_final_exp, int use_graph) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (!valid_g_ml(g_miller) || !valid_g_fe(g_final_exp) || wpc_miller < -1 || wpc_final_exp < -1) {
    g_err = "unsupported engine configuration";
    return -1;
  }
  g.bcfg = BatchedCfg{g_miller, g_final_exp, wpc_miller, wpc_final_exp};
  g.tuned.clear();
  g.use_graph = use_graph != 0;
  return 0;
}

int blsgpu_get_batched_engine(int M, int cfg[5]) {
  std::lock_guard<std::mutex> lk(g.mu);
  if (!cfg) return -1;
  const BatchedCfg c = cfg_for_locked(M > 0 ? M : 1);
  cfg[0] = c.g_ml;
  cfg[1] = c.g_fe;
  cfg[2] = c.wpc_ml;
  cfg[3] = c.wpc_fe;
  cfg[4] = g.use_graph ? 1 : 0;
  return 0;
}

int blsgpu_set_line_depth(int depth) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (depth < -1 || depth > 2) {
    g_err = "key-line depth must be -1, 0, 1 or 2";
    return -1;
  }
  g.force_depth = depth;
  return 0;
}

int blsgpu_get_line_depth(int form, int M) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if ((form != 0 && form != 1) || M <= 0) {
    g_err = "invalid arguments";
    return -1;
  }
  if (form == 1 && g.mode != 1) return 0;
  bls::EngineArgs a{};
  a.M = M;
  a.mode = 1;
  int dummy_queue = 0;
  if (form == 1) {
    a.g_ml = g.g_ml;
    a.depth_ml = plain_depth_locked();
  } else {
    const BatchedCfg c = cfg_for_locked(M);
    a.g_ml = c.g_ml;
    a.wpc_ml = c.wpc_ml;
    a.depth_ml = c.depth_ml;
    a.queue = &dummy_queue;
  }
  const bool tiles = !(a.queue && a.wpc_ml >= 0);
  return bls::engine_miller_depth(a, tiles && bls::grp_rider_items(0, a.g_ml, M) > 0);
}

int blsgpu_batched_stats(int out[2]) {
  std::lock_guard<std::mutex> lk(g.mu);
  if (!out) return -1;
  out[0] = g.graph_builds;
  out[1] = g.graph_launches;
  return 0;
}

int blsgpu_autotune_device(const uint8_t* d_sigs48, const uint64_t* d_neg_hms, const int32_t* d_key_idx, int M,
                           int form, void* stream, int cfg_out[4], double ms_out[2]) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (!d_sigs48 || !d_neg_hms || !d_key_idx || M <= 0 || (form != 0 && form != 1)) {
    g_err = "invalid arguments";
    return -1;
  }
  CHK(init_locked());
  const hipStream_t caller = stream ? (hipStream_t)stream : g.s;
  CHK(claim_workspace_locked(g.s));
  CHK(stream_enter(caller));
  CHK(ensure_batched_ws(M));
  CHK(g.b_res.ensure((size_t)M * 4));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  struct Ev {
    hipEvent_t a, b;
    ~Ev() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } guard{e0, e1};
  BatchedCfg c = g.bcfg;
  c.depth_ml = g.force_depth;
  bls::EngineArgs a =
      batched_args(d_sigs48, d_neg_hms, d_key_idx, g.b_res.as<int32_t>(), nullptr, M, c);
  CHK(bls::engine_prepare(a, g.s));
  // times one phase (1 final exponentiation, 2 the whole sequence) of configuration a: the
  // median of three runs after one untimed run (the whole sequence overlaps two streams,
  // and the best of a few runs can be a run where the concurrent kernel happened to
  // interfere less than it does in steady state).  A candidate whose first timed run is
  // 25 % slower than `bound` (the best so far) is not run again.
  // For the batched form with graphs on, the whole sequence runs from a graph of it, as
  // the batched entry points run it: the two concurrent kernels (the subgroup test and
  // the Miller loop) reach the dispatcher in another order from a graph than from two
  // streams, and a persistent grid whose SIMDs the test's wavefronts took first starts
  // part of its workgroups late (a 12-lane Miller loop at five workgroups per CU timed
  // 20.3 ms from the streams and ran 22.0 ms from the graph; the 32-lane one 20.3 ms both
  // ways).
  auto time_phase = [&](int which, double* ms, double bound) -> int {
    struct Exec {
      hipGraphExec_t e = nullptr;
      ~Exec() {
        if (e) (void)hipGraphExecDestroy(e);
      }
    } ex;
    if (which == 2 && form == 0 && g.use_graph) {
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(g.s, hipStreamCaptureModeThreadLocal));
      const int rc = bls::engine_compressed(a, g.s);
      const hipError_t ec = hipStreamEndCapture(g.s, &graph);
      if (rc != 0 || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        g_err = "graph capture of a tuning candidate failed";
        return -1;
      }
      const hipError_t ei = hipGraphInstantiate(&ex.e, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIPCHK(ei);
    }
    float t[3];
    int n = 0;
    for (int r = 0; r < 4; r++) {
      HIPCHK(hipEventRecord(e0, g.s));
      if (ex.e) HIPCHK(hipGraphLaunch(ex.e, g.s));
      else CHK(which == 1 ? bls::engine_final_exp(a, g.s) : bls::engine_compressed(a, g.s));
      HIPCHK(hipEventRecord(e1, g.s));
      HIPCHK(hipEventSynchronize(e1));
      if (r == 0) continue;
      HIPCHK(hipEventElapsedTime(&t[n], e0, e1));
      n++;
      if (r == 1 && t[0] > 1.25 * bound) break;
    }
    std::sort(t, t + n);
    *ms = t[n / 2];
    return 0;
  };
  const int gml[] = {8, 12, 13, 16, 20, 24, 32}, gfe[] = {9, 12, 16, 18, 24, 27, 32};
  // 1. the final exponentiation alone, on the Miller output of the current configuration;
  // the fastest candidates (three for the plain form, two for the batched form, which
  // has more workgroups-per-CU variants) go on to step 2
  CHK(bls::engine_miller(a, g.s));
  struct Cand {
    double ms;
    int G, w;
  };
  std::vector<Cand> fe;
  for (int G : gfe) {
    const int occ = form == 0 ? bls::occupancy_grp(1, G) : -1;
    for (int w = -1; w <= occ; w++) {
      if (w == 0) continue;  // the same grid as w == occ
      a.g_fe = G;
      a.wpc_fe = w;
      double t = 0;
      CHK(time_phase(1, &t, fe.empty() ? 1e30 : std::min_element(fe.begin(), fe.end(), [](const Cand& x, const Cand& y) {
                                                   return x.ms < y.ms;
                                                 })->ms));
      fe.push_back({t, G, w});
    }
  }
  std::sort(fe.begin(), fe.end(), [](const Cand& x, const Cand& y) { return x.ms < y.ms; });
  const size_t keep = form == 0 ? 2 : 3;
  if (fe.size() > keep) fe.resize(keep);
  const double best_fe = fe[0].ms;
  // 2. every Miller candidate with each of those, inside the whole sequence (decode, the
  // Miller loop and the final exponentiation, with the subgroup test running alongside
  // one of them): the concurrent test competes for the SIMDs, which favours pairs that
  // leave some free, so the two kernels are not timed alone
  double best_ml = 1e30;
  for (const Cand& f : fe) {
    a.g_fe = f.G;
    a.wpc_fe = f.w;
    for (int G : gml) {
      const int occ = form == 0 ? bls::occupancy_grp(0, G) : -1;
      for (int w = -1; w <= occ; w++) {
        if (w == 0) continue;
        a.g_ml = G;
        a.wpc_ml = w;
        double t = 0;
        CHK(time_phase(2, &t, best_ml));
        if (t < best_ml) {
          best_ml = t;
          c.g_ml = G;
          c.wpc_ml = w;
          c.g_fe = f.G;
          c.wpc_fe = f.w;
        }
      }
    }
  }
  // 3. the key-line depth of the chosen Miller configuration (unless one is forced):
  // every depth its instantiation holds, in the whole sequence
  if (g.force_depth < 0) {
    a.g_ml = c.g_ml;
    a.wpc_ml = c.wpc_ml;
    a.g_fe = c.g_fe;
    a.wpc_fe = c.wpc_fe;
    const bool tiles = !(a.queue && a.wpc_ml >= 0);
    const bool riders = tiles && bls::grp_rider_items(0, a.g_ml, M) > 0;
    for (int d = 0; d <= 2; d++) {
      a.depth_ml = d;
      if (bls::engine_miller_depth(a, riders) != d) continue;  // runs as a smaller depth
      double t = 0;
      CHK(time_phase(2, &t, best_ml));
      if (t < best_ml) {
        best_ml = t;
        c.depth_ml = d;
      }
    }
    if (c.depth_ml < 0) {
      a.depth_ml = -1;
      c.depth_ml = bls::engine_miller_depth(a, riders);
    }
  }
  if (form == 0) {
    g.tuned[size_class(M)] = c;
  } else {
    g.mode = 1;
    g.g_ml = c.g_ml;
    g.g_fe = c.g_fe;
    g.depth_ml = c.depth_ml;
  }
  if (cfg_out) {
    cfg_out[0] = c.g_ml;
    cfg_out[1] = c.g_fe;
    cfg_out[2] = c.wpc_ml;
    cfg_out[3] = c.wpc_fe;
  }
  if (ms_out) {
    ms_out[0] = best_ml;
    ms_out[1] = best_fe;
  }
  return stream_return(caller);
}

}  // extern "C"
