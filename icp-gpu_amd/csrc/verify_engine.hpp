// Kernel sequences for one verification batch.  Buffers are supplied by the caller
// (the library context or a device-resident caller); the engine allocates nothing.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bls {

struct EngineArgs {
  int M = 0;
  // inputs (device pointers)
  const uint64_t* sig_ref = nullptr;  // legacy: M x 12 u64 uncompressed signatures
  const uint8_t* comp = nullptr;      // fused: M x 48 compressed signatures
  const uint64_t* nhm_ref = nullptr;  // M x 12 u64, -H(m)
  const uint64_t* g2_ref = nullptr;   // legacy: generator lines, 2448 u64
  const uint64_t* pk_ref = nullptr;   // legacy: M x 2448 u64 per-item lines
  const int32_t* key_idx = nullptr;   // fused: M key handles
  const int32_t* keys = nullptr;      // fused: resident key tables (internal)
  int nkeys = 0;                      // fused: number of resident keys
  const int32_t* gen_lines = nullptr; // fused: generator table (internal)
  // workspace
  int32_t* sig_pts = nullptr;   // M x 28
  int32_t* nhm_pts = nullptr;   // M x 28
  int32_t* g2_lines = nullptr;  // 5712 (legacy)
  int32_t* pk_lines = nullptr;  // legacy: the M items' internal line tables, 5712 words each
  int32_t* item_idx = nullptr;  // legacy: M words, item i's table index (i)
  int32_t* status = nullptr;    // M (fused): decode status
  int32_t* sub_status = nullptr; // M (fused): subgroup status
  int32_t* fbuf = nullptr;      // M x 168 Miller values: word-major (SoA) in mode 0, item-major in mode 1
  // outputs
  int32_t* results = nullptr;  // M
  uint64_t* gt = nullptr;      // M x 72 (nullable)
  // tuning: mode 0 = one thread per item, 1 = G lanes per item (g_ml Miller, g_fe final exp)
  int block = 64;
  int mode = 1;
  int g_ml = 16;
  int g_fe = 16;
  // persistent form: two device counters (Miller, final exp) that the sequence zeroes
  // itself, and the workgroups per CU of each kernel (0 = occupancy limit, -1 = that
  // kernel is launched with one workgroup per tile and no queue)
  int* queue = nullptr;
  int wpc_ml = 0;
  int wpc_fe = 0;
  // homes of the compact final exponentiation (final_exp_park_words() words) for park_items
  // resident items; null: queue-fed launches keep the full-LDS kernel
  int32_t* park = nullptr;
  long park_items = 0;
  // key-line depth of the Miller loop (launch_miller_grp; -1 = the default)
  int depth_ml = -1;
  // optional timing: events recorded before each phase and at the end (5 events)
  hipEvent_t* ev = nullptr;
  // second stream for the subgroup test, and the fork/join events (fused path)
  hipStream_t s2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  // set when -H(m) is produced on another stream (hashed on the device): the engine
  // waits for it before reading nhm_ref
  hipEvent_t nhm_ready = nullptr;
};

int engine_legacy(const EngineArgs& a, hipStream_t s);
// The legacy sequence in parts, so that the Miller loop of one part runs while the next
// part's tables are still being copied in: engine_legacy_setup (on the copy stream cs:
// the generator's table and the item-index array), then per part of items
// [first, first + n) engine_legacy_part (its points and tables converted on cs, `ready`
// recorded there, then its Miller loop on s after that event), then
// engine_legacy_finish (the final exponentiation of all M on s).  Mode 0 runs one part.
// engine_legacy is the three on one stream.
int engine_legacy_setup(const EngineArgs& a, hipStream_t cs);
int engine_legacy_part(const EngineArgs& a, int first, int n, hipStream_t cs, hipEvent_t ready, hipStream_t s);
int engine_legacy_finish(const EngineArgs& a, hipStream_t s);
int engine_compressed(const EngineArgs& a, hipStream_t s);
// decode (and optionally subgroup-check) a.comp into a.sig_pts / a.status
int engine_decompress(const EngineArgs& a, int check_subgroup, hipStream_t s);
// the two halves of engine_compressed, for the autotuner: decode, -H(m) conversion and
// the subgroup test (all on s); then the Miller loop alone and the final exponentiation alone
int engine_prepare(const EngineArgs& a, hipStream_t s);
int engine_miller(const EngineArgs& a, hipStream_t s);
int engine_final_exp(const EngineArgs& a, hipStream_t s);
// the key-line depth the Miller launch of a (with or without subgroup riders) runs
int engine_miller_depth(const EngineArgs& a, bool riders);

}  // namespace bls
