// Process-wide state of libgpu_msm.so shared by its entry points (api.cpp, batched.cpp,
// server.cpp): the device, its streams, grow-only workspaces, the resident key tables
// and the engine configuration.  Every field is guarded by Ctx::mu.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace blsctx {

// message of the last failure on the calling thread (blsgpu_last_error)
extern thread_local std::string g_err;

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      ::blsctx::g_err = std::string(#x) + ": " + hipGetErrorString(e_);             \
      return -1;                                                                    \
    }                                                                               \
  } while (0)
#define CHK(x)                                                                      \
  do {                                                                              \
    if ((x) != 0) {                                                                 \
      if (::blsctx::g_err.empty()) ::blsctx::g_err = std::string(#x) + " failed";   \
      return -1;                                                                    \
    }                                                                               \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t n) {
    if (n <= cap) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = n + n / 4 + 4096;
    if (hipMalloc(&p, want) != hipSuccess) {
      g_err = "hipMalloc failed";
      p = nullptr;
      return -1;
    }
    cap = want;
    return 0;
  }
  template <class T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

// pinned host memory (hipHostMalloc), grow-only
struct PinnedBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t n) {
    if (n <= cap) return 0;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = n + n / 4 + 4096;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
      g_err = "hipHostMalloc failed";
      p = nullptr;
      return -1;
    }
    cap = want;
    return 0;
  }
  template <class T>
  T* as() const { return reinterpret_cast<T*>(p); }
};
// Configuration of the queue-fed (persistent) form: lanes per item of each kernel and
// workgroups per CU (-1 = one workgroup per tile without a queue, 0 = occupancy limit).
struct BatchedCfg {
  int g_ml = 12, g_fe = 16, wpc_ml = 0, wpc_fe = 0;
  int depth_ml = -1;  // key-line depth of the Miller loop (-1 = the default of g_ml)
};

// One instantiated graph of the batched form and everything its nodes captured.
struct GraphKey {
  const void* in[3];
  void* out[2];
  const void* ws[7];
  const void* keys;
  int nkeys, M;
  BatchedCfg cfg;
  bool operator==(const GraphKey& o) const;
};

struct Ctx {
  std::mutex mu;
  int state = 0;  // 0 = not initialized, 1 = ready, -1 = no device
  int device = 0;
  hipStream_t s = nullptr;
  // per-call staging
  DevBuf sig_ref, nhm_ref, g2_ref, pk_ref, comp, keyidx;
  DevBuf msg_bytes, msg_offs, dst_bytes;  // messages and DST of the hashing entry points
  PinnedBuf stage;                         // host staging of the blocking entry points
  // the legacy entry point's key tables pass through a ring of pinned slots (kRingTables
  // tables each): ev_ring[k] is recorded after the copy that reads slot k
  static constexpr int kRingSlots = 3, kRingTables = 428;
  PinnedBuf ring;
  hipEvent_t ev_ring[kRingSlots] = {};
  bool ring_used[kRingSlots] = {};
  int ring_next = 0;
  hipEvent_t ev_part[2] = {};  // a part of the legacy batch is in device memory
  DevBuf sig_pts, nhm_pts, g2_lines, pk_lines, status, sub_status, fbuf, res, gt;
  hipStream_t s2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_h2c = nullptr;
  // fixed workspace of blsgpu_verify_compressed_device_noalloc: the device address of the
  // file-scope array in api.cpp on g.device (from hipGetSymbolAddress at init), the event
  // recorded after the last call that used it and the stream that call ran on
  void* scratch = nullptr;
  hipEvent_t ev_scratch = nullptr;
  hipStream_t scratch_last = nullptr;
  // resident
  DevBuf gen_lines;  // G2 generator lines (internal), computed at init
  DevBuf keys;       // registered key tables (internal, key-major)
  int nkeys = 0;
  // superseded key tables: kept until the process ends, since launches of other
  // streams (the serving loop's) may still read them
  std::vector<void*> retired_keys;
  // device-form calls: the stream the workspace was last used on, and timing events
  hipStream_t last = nullptr;
  bool profiling = false;
  std::vector<hipEvent_t> ev_free;
  std::vector<std::vector<hipEvent_t>> ev_pending;
  double phase_ms[4] = {0, 0, 0, 0};
  int phase_calls = 0;
  // the legacy entry point's host-to-device copies, kernels and device-to-host copies
  double legacy_ms[3] = {0, 0, 0};
  int legacy_calls = 0;
  // engine configuration of the plain form (blsgpu_set_engine), its Miller key-line depth
  // (autotuned, -1 = the default) and a depth forced on every form (blsgpu_set_line_depth)
  int mode = 1, g_ml = 12, g_fe = 16;
  int depth_ml = -1, force_depth = -1;
  // batched (queue-fed) form: its own workspace, the configuration by batch-size class
  // (ceil(log2 M) -> autotuned configuration) and the graph of the last call
  DevBuf b_comp, b_nhm, b_keyidx, b_res, b_gt;
  DevBuf b_sig_pts, b_nhm_pts, b_status, b_sub, b_fbuf, b_queue;
  DevBuf b_park;  // homes of the compact final exponentiation (bls::final_exp_park_words())
  BatchedCfg bcfg;
  std::map<int, BatchedCfg> tuned;
  bool use_graph = true;
  bool have_graph = false;
  GraphKey gkey{};
  hipGraphExec_t gexec = nullptr;
  int graph_builds = 0, graph_launches = 0;
};
extern Ctx g;

int init_locked();
// wait for the stream that last used the shared workspace if it differs from s
int claim_workspace_locked(hipStream_t s);
// configuration of the batched form for a job of M items (autotuned class or default),
// with the forced key-line depth applied
BatchedCfg cfg_for_locked(int M);
// key-line depth of the plain form's Miller loop (forced, else autotuned, else -1)
inline int plain_depth_locked() { return g.force_depth >= 0 ? g.force_depth : g.depth_ml; }

}  // namespace blsctx
