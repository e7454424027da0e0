// C ABI of libgpu_msm.so (declared in include/gpu_msm.h): host orchestration over HIP.
//
// Differences from the reference's host side (gpu-crypto/cuda/api.cu:115-160,
// ic-integration/gpu_crypto/cuda/decompress.cu:191-234), by design:
//   * one process-wide context per library with grow-only device buffers instead of
//     five hipMalloc/hipFree pairs per call;
//   * a private non-blocking stream and a mutex, so concurrent callers are safe;
//   * every HIP return code is checked; any failure returns -1 (the caller's CPU
//     fallback) and records a message for blsgpu_last_error();
//   * per-key line tables can be registered once and stay resident in HBM.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpu_msm.h"
#include "context.hpp"
#include "host_copy.hpp"
#include "kernels.hpp"
#include "verify_engine.hpp"

namespace blsctx {

thread_local std::string g_err;
Ctx g;

namespace {
// Standard G2 generator in ic_bls12_381's Montgomery form (R = 2^384).
const uint64_t G2_GEN_REF[24] = {
    0xf5f28fa202940a10ULL, 0xb3f5fb2687b4961aULL, 0xa1a893b53e2ae580ULL, 0x9894999d1a3caee9ULL,
    0x6f67b7631863366bULL, 0x058191924350bcd7ULL, 0xa5a9c0759e23f606ULL, 0xaaa0c59dbccd60c3ULL,
    0x3bb17e18e2867806ULL, 0x1b1ab6cc8541b367ULL, 0xc2b6ed0ef2158547ULL, 0x11922a097360edf3ULL,
    0x4c730af860494c4aULL, 0x597cfa1f5e369c5aULL, 0xe7e6856caa0a635aULL, 0xbbefb5e96e0d495fULL,
    0x07d3a975f0ef25a2ULL, 0x0083fd8e7e80dae5ULL, 0xadc0fc92df64b05dULL, 0x18aa270a2b1461dcULL,
    0x86adac6a3be4eba0ULL, 0x79495c4ec93da33aULL, 0xe7175850a43ccaedULL, 0x0b2bc2a163de1bf2ULL};


}  // namespace

// Fixed workspace of blsgpu_verify_compressed_device_noalloc, for callers that cannot
// allocate: per item the decoded signature and -H(m) (28 words each), the decode and
// subgroup statuses and the Miller value (168 words), each part laid out for
// BLSGPU_NOALLOC_MAX_ITEMS items at a 256-byte aligned offset.  Every word a call reads
// is written earlier in the same call (decode writes every status, the subgroup test
// every subgroup status, the Miller loop the value of every item the final
// exponentiation reads), so nothing is cleared between calls.  In host code `scratch`
// names the host-side shadow of the array, not device memory: host code uses only the
// address hipGetSymbolAddress gives at init (g.scratch).
namespace {
constexpr size_t NA_MAX = BLSGPU_NOALLOC_MAX_ITEMS;
constexpr size_t NA_SIG = 0;
constexpr size_t NA_NHM = NA_SIG + NA_MAX * bls::POINT_WORDS * 4;
constexpr size_t NA_STATUS = NA_NHM + NA_MAX * bls::POINT_WORDS * 4;
constexpr size_t NA_SUB = NA_STATUS + NA_MAX * 4;
constexpr size_t NA_FBUF = NA_SUB + NA_MAX * 4;
constexpr size_t NA_BYTES = NA_FBUF + NA_MAX * bls::MILLER_WORDS * 4;
static_assert(NA_NHM % 256 == 0 && NA_STATUS % 256 == 0 && NA_SUB % 256 == 0 && NA_FBUF % 256 == 0,
              "aligned parts");
static_assert(NA_BYTES == BLSGPU_NOALLOC_SCRATCH_BYTES, "the size the header states");
}  // namespace
static __device__ __align__(256) uint8_t scratch[NA_BYTES];

// Small host arrays of a blocking call travel through one pinned staging area: one CPU
// copy, then DMA from pinned memory (the runtime stages pageable copies itself, at a
// fixed cost per copy that dominated these calls).  Every such call synchronizes before
// it returns, so the area is free again at the next one.  Large arrays (the per-item line
// tables of the legacy entry point) are copied directly.
struct Staging {
  static constexpr size_t kMax = size_t(8) << 20;
  size_t off = 0;
  struct Out {
    void* host;
    size_t at, n;
  };
  std::vector<Out> outs;
  static size_t span(size_t n) { return n <= kMax ? (n + 255) & ~size_t(255) : 0; }
  int begin(size_t total) {
    off = 0;
    outs.clear();
    return g.stage.ensure(total + 4096);
  }
  int h2d(void* dev, const void* host, size_t n, hipStream_t s) {
    if (n == 0) return 0;
    const void* src = host;
    if (n <= kMax && off + span(n) <= g.stage.cap) {
      uint8_t* p = g.stage.as<uint8_t>() + off;
      std::memcpy(p, host, n);
      off += span(n);
      src = p;
    }
    if (hipMemcpyAsync(dev, src, n, hipMemcpyHostToDevice, s) != hipSuccess) {
      g_err = "host to device copy failed";
      return -1;
    }
    return 0;
  }
  int d2h(void* host, const void* dev, size_t n, hipStream_t s) {
    if (n == 0) return 0;
    void* dstp = host;
    if (n <= kMax && off + span(n) <= g.stage.cap) {
      outs.push_back({host, off, n});
      dstp = g.stage.as<uint8_t>() + off;
      off += span(n);
    }
    if (hipMemcpyAsync(dstp, dev, n, hipMemcpyDeviceToHost, s) != hipSuccess) {
      g_err = "device to host copy failed";
      return -1;
    }
    return 0;
  }
  // after the stream has synchronized
  void finish() {
    for (const Out& o : outs) std::memcpy(o.host, g.stage.as<uint8_t>() + o.at, o.n);
  }
};

// Workspace buffers are shared by all calls; before they are reused (or grown) from a
// different stream, wait for the stream that last used them.
int claim_workspace_locked(hipStream_t s) {
  if (g.last && g.last != s) HIPCHK(hipStreamSynchronize(g.last));
  g.last = s;
  return 0;
}

int init_locked() {
  if (g.state == 1) return 0;
  if (g.state == -1) {
    g_err = "no HIP device";
    return -1;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    g.state = -1;
    g_err = "no HIP device";
    return -1;
  }
  HIPCHK(hipSetDevice(g.device));
  // each object is made once: a set-up that failed part way resumes where it stopped
  // on the next call instead of making (and leaking) the earlier objects again
  auto stream = [](hipStream_t& st) { return st ? hipSuccess : hipStreamCreateWithFlags(&st, hipStreamNonBlocking); };
  auto event = [](hipEvent_t& e) { return e ? hipSuccess : hipEventCreateWithFlags(&e, hipEventDisableTiming); };
  HIPCHK(stream(g.s));
  HIPCHK(stream(g.s2));
  HIPCHK(event(g.ev_fork));
  HIPCHK(event(g.ev_join));
  HIPCHK(event(g.ev_h2c));
  for (hipEvent_t& e : g.ev_ring) HIPCHK(event(e));
  for (hipEvent_t& e : g.ev_part) HIPCHK(event(e));
  // the fixed workspace's device address on this device (g.device is current)
  void* sp = nullptr;
  HIPCHK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  if (!sp) {
    g_err = "no device address for the fixed workspace";
    return -1;
  }
  g.scratch = sp;
  HIPCHK(event(g.ev_scratch));
  // generator lines, once
  CHK(g.g2_ref.ensure(24 * 8));
  CHK(g.gen_lines.ensure(bls::KEY_TABLE_WORDS * 4));
  HIPCHK(hipMemcpyAsync(g.g2_ref.p, G2_GEN_REF, 24 * 8, hipMemcpyHostToDevice, g.s));
  CHK(bls::launch_g2_prepare(g.g2_ref.as<uint64_t>(), g.gen_lines.as<int32_t>(), 1, g.s));
  HIPCHK(hipStreamSynchronize(g.s));
  g.state = 1;
  return 0;
}

// Tables [first, first + n) of the caller's key-line array (2448 u64 each) to g.pk_ref on
// the copy stream cs, through the pinned ring: host_copy fills a slot once the copy that
// last read it has completed, then an asynchronous copy takes the slot to the device, so
// that filling slot k overlaps the transfer of slot k - 1.
static int stage_tables_locked(const uint64_t* pkc, int first, int n, hipStream_t cs) {
  constexpr size_t tb = 2448 * 8;
  constexpr int per = Ctx::kRingTables;
  for (int c = first; c < first + n; c += per) {
    const int k = first + n - c < per ? first + n - c : per;
    const int slot = g.ring_next;
    g.ring_next = (g.ring_next + 1) % Ctx::kRingSlots;
    if (g.ring_used[slot]) HIPCHK(hipEventSynchronize(g.ev_ring[slot]));
    uint8_t* h = g.ring.as<uint8_t>() + (size_t)slot * per * tb;
    host_copy(h, pkc + (size_t)c * 2448, (size_t)k * tb);
    HIPCHK(hipMemcpyAsync(g.pk_ref.as<uint8_t>() + (size_t)c * tb, h, (size_t)k * tb, hipMemcpyHostToDevice, cs));
    HIPCHK(hipEventRecord(g.ev_ring[slot], cs));
    g.ring_used[slot] = true;
  }
  return 0;
}

// Batches of at least this many items run the legacy sequence in two parts (the Miller
// loop of the first half overlaps the copy of the second half's tables).
constexpr int kLegacySplit = 1024;

int legacy_verify_locked(const uint64_t* sigs, const uint64_t* nhms, const uint64_t* g2c,
                         const uint64_t* pkc, int32_t* results, uint64_t* gt_out, int M) {
  CHK(init_locked());
  CHK(claim_workspace_locked(g.s));
  const size_t m = (size_t)M;
  CHK(g.sig_ref.ensure(m * 96));
  CHK(g.nhm_ref.ensure(m * 96));
  CHK(g.g2_ref.ensure(2448 * 8));
  CHK(g.pk_ref.ensure(m * 2448 * 8));
  CHK(g.sig_pts.ensure(m * bls::POINT_WORDS * 4));
  CHK(g.nhm_pts.ensure(m * bls::POINT_WORDS * 4));
  CHK(g.g2_lines.ensure(bls::KEY_TABLE_WORDS * 4));
  CHK(g.pk_lines.ensure(m * bls::KEY_TABLE_WORDS * 4));
  CHK(g.keyidx.ensure(m * 4));
  CHK(g.res.ensure(m * 4));
  CHK(g.fbuf.ensure(m * bls::MILLER_WORDS * 4));
  if (gt_out) CHK(g.gt.ensure(m * 72 * 8));
  CHK(g.ring.ensure((size_t)Ctx::kRingSlots * Ctx::kRingTables * 2448 * 8));
  Staging st;
  CHK(st.begin(2 * Staging::span(m * 96) + Staging::span(2448 * 8) + Staging::span(m * 4) +
               (gt_out ? Staging::span(m * 72 * 8) : 0)));
  // copies in and conversions on cs, pairing kernels and copies out on g.s
  hipStream_t cs = g.s2;
  // with profiling on: events at the first copy in, when the last key table has landed,
  // after the kernels and after the copies out
  hipEvent_t ev[4] = {};
  if (g.profiling) {
    for (hipEvent_t& e : ev) {
      if (!g.ev_free.empty()) {
        e = g.ev_free.back();
        g.ev_free.pop_back();
      } else {
        HIPCHK(hipEventCreate(&e));
      }
    }
    HIPCHK(hipEventRecord(ev[0], cs));
  }
  CHK(st.h2d(g.sig_ref.p, sigs, m * 96, cs));
  CHK(st.h2d(g.nhm_ref.p, nhms, m * 96, cs));
  CHK(st.h2d(g.g2_ref.p, g2c, 2448 * 8, cs));
  bls::EngineArgs a{};
  a.M = M;
  a.mode = g.mode;
  a.g_ml = g.g_ml;
  a.g_fe = g.g_fe;
  a.depth_ml = plain_depth_locked();
  a.sig_ref = g.sig_ref.as<uint64_t>();
  a.nhm_ref = g.nhm_ref.as<uint64_t>();
  a.g2_ref = g.g2_ref.as<uint64_t>();
  a.pk_ref = g.pk_ref.as<uint64_t>();
  a.sig_pts = g.sig_pts.as<int32_t>();
  a.nhm_pts = g.nhm_pts.as<int32_t>();
  a.g2_lines = g.g2_lines.as<int32_t>();
  a.pk_lines = g.pk_lines.as<int32_t>();
  a.item_idx = g.keyidx.as<int32_t>();
  a.fbuf = g.fbuf.as<int32_t>();
  a.results = g.res.as<int32_t>();
  a.gt = gt_out ? g.gt.as<uint64_t>() : nullptr;
  CHK(bls::engine_legacy_setup(a, cs));
  const int parts = g.mode == 1 && M >= kLegacySplit ? 2 : 1;
  for (int p = 0, first = 0; p < parts; p++) {
    const int n = p + 1 < parts ? M / parts : M - first;
    CHK(stage_tables_locked(pkc, first, n, cs));
    if (p + 1 == parts && ev[1]) HIPCHK(hipEventRecord(ev[1], cs));
    CHK(bls::engine_legacy_part(a, first, n, cs, g.ev_part[p], g.s));
    first += n;
  }
  CHK(bls::engine_legacy_finish(a, g.s));
  if (ev[2]) HIPCHK(hipEventRecord(ev[2], g.s));
  CHK(st.d2h(results, g.res.p, m * 4, g.s));
  if (gt_out) CHK(st.d2h(gt_out, g.gt.p, m * 72 * 8, g.s));
  if (ev[3]) HIPCHK(hipEventRecord(ev[3], g.s));
  // every operation on cs precedes an event that g.s waited for
  HIPCHK(hipStreamSynchronize(g.s));
  st.finish();
  if (ev[0]) {
    for (int k = 0; k < 3; k++) {
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
      g.legacy_ms[k] += t;
    }
    g.legacy_calls++;
    for (hipEvent_t e : ev) g.ev_free.push_back(e);
  }
  return 0;
}

int grow_keys_locked(int add) {
  size_t need = (size_t)(g.nkeys + add) * bls::KEY_TABLE_WORDS * 4;
  if (need <= g.keys.cap) return 0;
  CHK(claim_workspace_locked(g.s));
  size_t want = need * 2;
  void* np = nullptr;
  HIPCHK(hipMalloc(&np, want));
  if (g.nkeys) HIPCHK(hipMemcpyAsync(np, g.keys.p, (size_t)g.nkeys * bls::KEY_TABLE_WORDS * 4,
                                      hipMemcpyDeviceToDevice, g.s));
  HIPCHK(hipStreamSynchronize(g.s));
  if (g.keys.p) g.retired_keys.push_back(g.keys.p);
  g.keys.p = np;
  g.keys.cap = want;
  return 0;
}

}  // namespace blsctx

using namespace blsctx;

extern "C" {

int gpu_crypto_init(void) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  return init_locked();
}

int gpu_batch_bls_verify(const uint64_t* sigs, const uint64_t* neg_hms, const uint64_t* g2_coeffs,
                         const uint64_t* pk_coeffs, int32_t* results, int M) {
  return blsgpu_batch_verify_gt(sigs, neg_hms, g2_coeffs, pk_coeffs, results, nullptr, M);
}

int blsgpu_batch_verify_gt(const uint64_t* sigs, const uint64_t* neg_hms, const uint64_t* g2_coeffs,
                           const uint64_t* pk_coeffs, int32_t* results, uint64_t* gt_out, int M) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (M <= 0 || !sigs || !neg_hms || !g2_coeffs || !pk_coeffs || !results) {
    g_err = "invalid arguments";
    return -1;
  }
  return legacy_verify_locked(sigs, neg_hms, g2_coeffs, pk_coeffs, results, gt_out, M);
}

int blsgpu_g1_decompress_status(const uint8_t* in48, uint8_t* out96, int32_t* status, int n, int check_subgroup) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (n <= 0 || !in48) {
    g_err = "invalid arguments";
    return -1;
  }
  CHK(init_locked());
  CHK(claim_workspace_locked(g.s));
  const size_t m = (size_t)n;
  CHK(g.comp.ensure(m * 48));
  CHK(g.sig_pts.ensure(m * bls::POINT_WORDS * 4));
  CHK(g.status.ensure(m * 4));
  CHK(g.sub_status.ensure(m * 4));
  CHK(g.sig_ref.ensure(m * 96));
  Staging st;
  CHK(st.begin(Staging::span(m * 48) + Staging::span(m * 96) + Staging::span(m * 4)));
  CHK(st.h2d(g.comp.p, in48, m * 48, g.s));
  {
    bls::EngineArgs a{};
    a.M = n;
    a.comp = g.comp.as<uint8_t>();
    a.sig_pts = g.sig_pts.as<int32_t>();
    a.status = g.status.as<int32_t>();
    a.sub_status = g.sub_status.as<int32_t>();
    CHK(bls::engine_decompress(a, check_subgroup, g.s));
  }
  CHK(bls::launch_points_to_ref(g.sig_pts.as<int32_t>(), g.status.as<int32_t>(), g.sig_ref.as<uint64_t>(), n, g.s));
  if (out96) CHK(st.d2h(out96, g.sig_ref.p, m * 96, g.s));
  if (status) CHK(st.d2h(status, g.status.p, m * 4, g.s));
  HIPCHK(hipStreamSynchronize(g.s));
  st.finish();
  return 0;
}

int gpu_g1_decompress_batch(const unsigned char* compressed, unsigned char* out_points, int n) {
  if (n <= 0 || !compressed || !out_points) return -1;
  std::vector<int32_t> st((size_t)n);
  if (blsgpu_g1_decompress_status(compressed, out_points, st.data(), n, 1) != 0) return -1;
  for (int i = 0; i < n; i++)
    if (st[i] != 0) {
      g_err = "batch holds a point that is not a valid non-identity G1 element";
      return -1;
    }
  return 0;
}

int gpu_g1_decompress(const unsigned char* compressed, unsigned char* out_point) {
  return gpu_g1_decompress_batch(compressed, out_point, 1);
}

int gpu_msm_g1(const unsigned char*, const unsigned char*, size_t, unsigned char*) {
  g_err = "G1 MSM is not provided by this library";
  return -1;
}
int gpu_jacobian_to_affine(const unsigned char*, unsigned char*) {
  g_err = "G1 MSM is not provided by this library";
  return -1;
}
int gpu_batch_sha256(const uint8_t*, uint8_t*, int, int) {
  g_err = "SHA-256 is not provided by this library";
  return -1;
}

int blsgpu_register_key_affine(const uint64_t pk[24]) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (!pk) return -1;
  CHK(init_locked());
  CHK(grow_keys_locked(1));
  CHK(g.g2_ref.ensure(24 * 8));
  HIPCHK(hipMemcpyAsync(g.g2_ref.p, pk, 24 * 8, hipMemcpyHostToDevice, g.s));
  CHK(bls::launch_g2_prepare(g.g2_ref.as<uint64_t>(), g.keys.as<int32_t>() + (size_t)g.nkeys * bls::KEY_TABLE_WORDS,
                             1, g.s));
  HIPCHK(hipStreamSynchronize(g.s));
  return g.nkeys++;
}

int blsgpu_register_keys_affine(const uint64_t* pks, int n, int32_t* handles) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (!pks || n <= 0) {
    g_err = "invalid arguments";
    return -1;
  }
  CHK(init_locked());
  CHK(grow_keys_locked(n));
  CHK(claim_workspace_locked(g.s));
  CHK(g.pk_ref.ensure((size_t)n * 24 * 8));
  HIPCHK(hipMemcpyAsync(g.pk_ref.p, pks, (size_t)n * 24 * 8, hipMemcpyHostToDevice, g.s));
  CHK(bls::launch_g2_prepare(g.pk_ref.as<uint64_t>(), g.keys.as<int32_t>() + (size_t)g.nkeys * bls::KEY_TABLE_WORDS,
                             n, g.s));
  HIPCHK(hipStreamSynchronize(g.s));
  const int first = g.nkeys;
  g.nkeys += n;
  if (handles)
    for (int i = 0; i < n; i++) handles[i] = first + i;
  return first;
}

int blsgpu_register_keys_compressed(const uint8_t* pks96, int n, int32_t* handles, int32_t* status,
                                    int check_subgroup) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (!pks96 || n <= 0) {
    g_err = "invalid arguments";
    return -1;
  }
  CHK(init_locked());
  CHK(claim_workspace_locked(g.s));
  const size_t row = 24 * 8;
  CHK(g.comp.ensure((size_t)n * 96));
  CHK(g.pk_ref.ensure((size_t)n * row * 2));  // decoded points, then the accepted ones packed
  CHK(g.status.ensure((size_t)n * 4));
  HIPCHK(hipMemcpyAsync(g.comp.p, pks96, (size_t)n * 96, hipMemcpyHostToDevice, g.s));
  CHK(bls::launch_g2_decompress(g.comp.as<uint8_t>(), g.pk_ref.as<uint64_t>(), g.status.as<int32_t>(), n,
                                check_subgroup, g.s));
  std::vector<int32_t> st(n);
  HIPCHK(hipMemcpyAsync(st.data(), g.status.p, (size_t)n * 4, hipMemcpyDeviceToHost, g.s));
  HIPCHK(hipStreamSynchronize(g.s));
  // accepted keys (status 0: decoded, not the identity) packed in order, one copy per run
  uint8_t* dec = g.pk_ref.as<uint8_t>();
  uint8_t* packed = dec + (size_t)n * row;
  int k = 0;
  for (int i = 0; i < n;) {
    if (st[i] != 0) {
      i++;
      continue;
    }
    int j = i;
    while (j < n && st[j] == 0) j++;
    HIPCHK(hipMemcpyAsync(packed + (size_t)k * row, dec + (size_t)i * row, (size_t)(j - i) * row,
                          hipMemcpyDeviceToDevice, g.s));
    k += j - i;
    i = j;
  }
  const int first = g.nkeys;
  if (k > 0) {
    CHK(grow_keys_locked(k));
    CHK(bls::launch_g2_prepare(reinterpret_cast<const uint64_t*>(packed),
                               g.keys.as<int32_t>() + (size_t)g.nkeys * bls::KEY_TABLE_WORDS, k, g.s));
    HIPCHK(hipStreamSynchronize(g.s));
    g.nkeys += k;
  }
  for (int i = 0, h = first; i < n; i++) {
    if (status) status[i] = st[i];
    if (handles) handles[i] = st[i] == 0 ? h++ : -1;
  }
  return k;
}

int blsgpu_register_key_coeffs(const uint64_t coeffs[2448]) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (!coeffs) return -1;
  CHK(init_locked());
  CHK(grow_keys_locked(1));
  CHK(g.pk_ref.ensure(2448 * 8));
  HIPCHK(hipMemcpyAsync(g.pk_ref.p, coeffs, 2448 * 8, hipMemcpyHostToDevice, g.s));
  CHK(bls::launch_lines_from_ref(g.pk_ref.as<uint64_t>(),
                                 g.keys.as<int32_t>() + (size_t)g.nkeys * bls::KEY_TABLE_WORDS, 1, g.s));
  HIPCHK(hipStreamSynchronize(g.s));
  return g.nkeys++;
}

int blsgpu_key_coeffs(int handle, uint64_t coeffs[2448]) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (!coeffs || handle < 0 || handle >= g.nkeys) {
    g_err = "invalid key handle";
    return -1;
  }
  CHK(g.pk_ref.ensure(2448 * 8));
  CHK(bls::launch_table_to_ref(g.keys.as<int32_t>() + (size_t)handle * bls::KEY_TABLE_WORDS,
                               g.pk_ref.as<uint64_t>(), g.s));
  HIPCHK(hipMemcpyAsync(coeffs, g.pk_ref.p, 2448 * 8, hipMemcpyDeviceToHost, g.s));
  HIPCHK(hipStreamSynchronize(g.s));
  return 0;
}

int blsgpu_num_keys(void) {
  std::lock_guard<std::mutex> lk(g.mu);
  return g.nkeys;
}
// The tables are retired, not reused: verifications already enqueued (device-form calls
// on callers' streams, serving steps) may still read them, and the next registration
// must not overwrite rows they read.  A retired buffer lives until the process ends.
void blsgpu_reset_keys(void) {
  std::lock_guard<std::mutex> lk(g.mu);
  if (g.keys.p) g.retired_keys.push_back(g.keys.p);
  g.keys.p = nullptr;
  g.keys.cap = 0;
  g.nkeys = 0;
}

// messages (bytes msgs[offsets[i] .. offsets[i+1]) of n) and the DST: the offsets and
// the DST length (1..255, RFC 9380 sec. 5.3.1) checked; returns the bytes to copy, or -1
static int64_t check_messages(const uint8_t* msgs, const int64_t* offsets, int n, const uint8_t* dst,
                              int dst_len) {
  if (!offsets || !dst || dst_len < 1 || dst_len > 255 || offsets[0] < 0) {
    g_err = "invalid message arguments";
    return -1;
  }
  for (int i = 0; i < n; i++)
    if (offsets[i + 1] < offsets[i] || offsets[i + 1] - offsets[i] > (1 << 24)) {
      g_err = "message offsets must be non-decreasing, messages at most 16 MiB";
      return -1;
    }
  if (offsets[n] > 0 && !msgs) {
    g_err = "invalid message arguments";
    return -1;
  }
  return offsets[n];
}
// the checked messages and DST, copied to the device on g.s
static int upload_messages_locked(Staging& st, const uint8_t* msgs, const int64_t* offsets, int n,
                                  const uint8_t* dst, int dst_len) {
  const size_t total = (size_t)offsets[n];
  CHK(g.msg_bytes.ensure(total > 0 ? total : 1));
  CHK(g.msg_offs.ensure((size_t)(n + 1) * 8));
  CHK(g.dst_bytes.ensure(256));
  CHK(st.h2d(g.msg_bytes.p, msgs, total, g.s));
  CHK(st.h2d(g.msg_offs.p, offsets, (size_t)(n + 1) * 8, g.s));
  CHK(st.h2d(g.dst_bytes.p, dst, (size_t)dst_len, g.s));
  return 0;
}

int blsgpu_hash_to_g1(const uint8_t* msgs, const int64_t* offsets, int n, const uint8_t* dst, int dst_len,
                      uint64_t* out, int32_t* is_identity, int negate) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (n <= 0 || !out) {
    g_err = "invalid arguments";
    return -1;
  }
  CHK(init_locked());
  CHK(claim_workspace_locked(g.s));
  const int64_t msg_total = check_messages(msgs, offsets, n, dst, dst_len);
  if (msg_total < 0) return -1;
  Staging st;
  CHK(st.begin(Staging::span((size_t)msg_total) + Staging::span((size_t)(n + 1) * 8) + 256 +
               Staging::span((size_t)n * 96) + Staging::span((size_t)n * 4)));
  CHK(upload_messages_locked(st, msgs, offsets, n, dst, dst_len));
  CHK(g.nhm_ref.ensure((size_t)n * 96));
  CHK(g.status.ensure((size_t)n * 4));
  CHK(bls::launch_hash_to_g1(g.msg_bytes.as<uint8_t>(), g.msg_offs.as<int64_t>(), n, g.dst_bytes.as<uint8_t>(),
                             dst_len, g.nhm_ref.as<uint64_t>(), g.status.as<int32_t>(), negate, g.s));
  CHK(st.d2h(out, g.nhm_ref.p, (size_t)n * 96, g.s));
  if (is_identity) CHK(st.d2h(is_identity, g.status.p, (size_t)n * 4, g.s));
  HIPCHK(hipStreamSynchronize(g.s));
  st.finish();
  return 0;
}

// shared by blsgpu_verify_compressed (neg_hms given) and blsgpu_verify_messages (the
// messages hashed on the device, on the second stream while the signatures decode)
static int verify_compressed_locked(const uint8_t* sigs48, const uint64_t* neg_hms, const uint8_t* msgs,
                                    const int64_t* offsets, const uint8_t* dst, int dst_len,
                                    const int32_t* key_idx, int32_t* results, uint64_t* gt_out, int M) {
  CHK(init_locked());
  CHK(claim_workspace_locked(g.s));
  for (int i = 0; i < M; i++)
    if (key_idx[i] < 0 || key_idx[i] >= g.nkeys) {
      g_err = "key index out of range";
      return -1;
    }
  const size_t m = (size_t)M;
  CHK(g.comp.ensure(m * 48));
  CHK(g.nhm_ref.ensure(m * 96));
  CHK(g.keyidx.ensure(m * 4));
  CHK(g.sig_pts.ensure(m * bls::POINT_WORDS * 4));
  CHK(g.nhm_pts.ensure(m * bls::POINT_WORDS * 4));
  CHK(g.status.ensure(m * 4));
  CHK(g.sub_status.ensure(m * 4));
  CHK(g.res.ensure(m * 4));
  CHK(g.fbuf.ensure(m * bls::MILLER_WORDS * 4));
  if (gt_out) CHK(g.gt.ensure(m * 72 * 8));
  const int64_t msg_total = neg_hms ? 0 : check_messages(msgs, offsets, M, dst, dst_len);
  if (msg_total < 0) return -1;
  Staging st;
  CHK(st.begin(Staging::span(m * 48) + Staging::span(m * 96) + Staging::span((size_t)msg_total) +
               Staging::span((m + 1) * 8) + 256 + 2 * Staging::span(m * 4) +
               (gt_out ? Staging::span(m * 72 * 8) : 0)));
  CHK(st.h2d(g.comp.p, sigs48, m * 48, g.s));
  if (neg_hms) CHK(st.h2d(g.nhm_ref.p, neg_hms, m * 96, g.s));
  else CHK(upload_messages_locked(st, msgs, offsets, M, dst, dst_len));
  CHK(st.h2d(g.keyidx.p, key_idx, m * 4, g.s));
  bls::EngineArgs a{};
  if (!neg_hms) {  // -H(m) on the second stream, beside the signature decoding
    HIPCHK(hipEventRecord(g.ev_fork, g.s));
    HIPCHK(hipStreamWaitEvent(g.s2, g.ev_fork, 0));
    CHK(bls::launch_hash_to_g1(g.msg_bytes.as<uint8_t>(), g.msg_offs.as<int64_t>(), M, g.dst_bytes.as<uint8_t>(),
                               dst_len, g.nhm_ref.as<uint64_t>(), nullptr, 1, g.s2));
    HIPCHK(hipEventRecord(g.ev_h2c, g.s2));
    a.nhm_ready = g.ev_h2c;
  }
  a.M = M;
  a.mode = g.mode;
  a.g_ml = g.g_ml;
  a.g_fe = g.g_fe;
  a.depth_ml = plain_depth_locked();
  a.comp = g.comp.as<uint8_t>();
  a.nhm_ref = g.nhm_ref.as<uint64_t>();
  a.key_idx = g.keyidx.as<int32_t>();
  a.keys = g.keys.as<int32_t>();
  a.nkeys = g.nkeys;
  a.gen_lines = g.gen_lines.as<int32_t>();
  a.sig_pts = g.sig_pts.as<int32_t>();
  a.nhm_pts = g.nhm_pts.as<int32_t>();
  a.status = g.status.as<int32_t>();
  a.sub_status = g.sub_status.as<int32_t>();
  a.s2 = g.s2;
  a.fork = g.ev_fork;
  a.join = g.ev_join;
  a.fbuf = g.fbuf.as<int32_t>();
  a.results = g.res.as<int32_t>();
  a.gt = gt_out ? g.gt.as<uint64_t>() : nullptr;
  CHK(bls::engine_compressed(a, g.s));
  CHK(st.d2h(results, g.res.p, m * 4, g.s));
  if (gt_out) CHK(st.d2h(gt_out, g.gt.p, m * 72 * 8, g.s));
  HIPCHK(hipStreamSynchronize(g.s));
  st.finish();
  return 0;
}

int blsgpu_verify_compressed(const uint8_t* sigs48, const uint64_t* neg_hms, const int32_t* key_idx,
                             int32_t* results, uint64_t* gt_out, int M) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (M <= 0 || !sigs48 || !neg_hms || !key_idx || !results) {
    g_err = "invalid arguments";
    return -1;
  }
  return verify_compressed_locked(sigs48, neg_hms, nullptr, nullptr, nullptr, 0, key_idx, results, gt_out, M);
}

int blsgpu_verify_messages(const uint8_t* sigs48, const uint8_t* msgs, const int64_t* offsets, const uint8_t* dst,
                           int dst_len, const int32_t* key_idx, int32_t* results, uint64_t* gt_out, int M) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (M <= 0 || !sigs48 || !key_idx || !results) {
    g_err = "invalid arguments";
    return -1;
  }
  return verify_compressed_locked(sigs48, nullptr, msgs, offsets, dst, dst_len, key_idx, results, gt_out, M);
}

int blsgpu_verify_compressed_device(const uint8_t* d_sigs48, const uint64_t* d_neg_hms, const int32_t* d_key_idx,
                                    int32_t* d_results, uint64_t* d_gt, int M, void* stream) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (M <= 0 || !d_sigs48 || !d_neg_hms || !d_key_idx || !d_results) {
    g_err = "invalid arguments";
    return -1;
  }
  CHK(init_locked());
  hipStream_t s = stream ? (hipStream_t)stream : g.s;
  CHK(claim_workspace_locked(s));
  const size_t m = (size_t)M;
  CHK(g.sig_pts.ensure(m * bls::POINT_WORDS * 4));
  CHK(g.nhm_pts.ensure(m * bls::POINT_WORDS * 4));
  CHK(g.status.ensure(m * 4));
  CHK(g.sub_status.ensure(m * 4));
  CHK(g.fbuf.ensure(m * bls::MILLER_WORDS * 4));
  bls::EngineArgs a{};
  a.M = M;
  a.mode = g.mode;
  a.g_ml = g.g_ml;
  a.g_fe = g.g_fe;
  a.depth_ml = plain_depth_locked();
  a.comp = d_sigs48;
  a.nhm_ref = d_neg_hms;
  a.key_idx = d_key_idx;
  a.keys = g.keys.as<int32_t>();
  a.nkeys = g.nkeys;
  a.gen_lines = g.gen_lines.as<int32_t>();
  a.sig_pts = g.sig_pts.as<int32_t>();
  a.nhm_pts = g.nhm_pts.as<int32_t>();
  a.status = g.status.as<int32_t>();
  a.sub_status = g.sub_status.as<int32_t>();
  a.s2 = g.s2;
  a.fork = g.ev_fork;
  a.join = g.ev_join;
  a.fbuf = g.fbuf.as<int32_t>();
  a.results = d_results;
  a.gt = d_gt;
  std::vector<hipEvent_t> ev;
  if (g.profiling) {
    for (int k = 0; k < 5; k++) {
      hipEvent_t e;
      if (!g.ev_free.empty()) {
        e = g.ev_free.back();
        g.ev_free.pop_back();
      } else {
        HIPCHK(hipEventCreate(&e));
      }
      ev.push_back(e);
    }
    a.ev = ev.data();
  }
  CHK(bls::engine_compressed(a, s));
  if (g.profiling) g.ev_pending.push_back(ev);
  return 0;
}

int blsgpu_verify_compressed_device_noalloc(const uint8_t* d_sigs48, const uint64_t* d_neg_hms,
                                            const int32_t* d_key_idx, int32_t* d_results, uint64_t* d_gt, int M,
                                            void* stream) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  if (M <= 0 || !d_sigs48 || !d_neg_hms || !d_key_idx || !d_results) {
    g_err = "invalid arguments";
    return -1;
  }
  if (M > BLSGPU_NOALLOC_MAX_ITEMS) {
    g_err = "M exceeds BLSGPU_NOALLOC_MAX_ITEMS (8192 items), the capacity of the fixed workspace";
    return -1;
  }
  // set-up (which allocates) belongs to gpu_crypto_init, never to this entry point
  if (g.state != 1 || !g.scratch) {
    g_err = "library not initialized: call gpu_crypto_init() first";
    return -1;
  }
  hipStream_t s = stream ? (hipStream_t)stream : g.s;
  // the previous call's kernels on the array finish first: stream order on the same
  // stream, an event wait across streams (no host wait)
  if (g.scratch_last && g.scratch_last != s) HIPCHK(hipStreamWaitEvent(s, g.ev_scratch, 0));
  uint8_t* base = static_cast<uint8_t*>(g.scratch);
  bls::EngineArgs a{};
  a.M = M;
  a.mode = g.mode;
  a.g_ml = g.g_ml;
  a.g_fe = g.g_fe;
  a.depth_ml = plain_depth_locked();
  a.comp = d_sigs48;
  a.nhm_ref = d_neg_hms;
  a.key_idx = d_key_idx;
  a.keys = g.keys.as<int32_t>();
  a.nkeys = g.nkeys;
  a.gen_lines = g.gen_lines.as<int32_t>();
  a.sig_pts = reinterpret_cast<int32_t*>(base + NA_SIG);
  a.nhm_pts = reinterpret_cast<int32_t*>(base + NA_NHM);
  a.status = reinterpret_cast<int32_t*>(base + NA_STATUS);
  a.sub_status = reinterpret_cast<int32_t*>(base + NA_SUB);
  a.fbuf = reinterpret_cast<int32_t*>(base + NA_FBUF);
  a.s2 = g.s2;
  a.fork = g.ev_fork;
  a.join = g.ev_join;
  a.results = d_results;
  a.gt = d_gt;
  CHK(bls::engine_compressed(a, s));
  HIPCHK(hipEventRecord(g.ev_scratch, s));
  g.scratch_last = s;
  return 0;
}

int blsgpu_set_engine(int mode, int g_miller, int g_final_exp) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  const bool ok_ml = bls::grp_lanes_supported(0, g_miller), ok_fe = bls::grp_lanes_supported(1, g_final_exp);
  if ((mode != 0 && mode != 1) || (mode == 1 && (!ok_ml || !ok_fe))) {
    g_err = "unsupported engine configuration";
    return -1;
  }
  g.mode = mode;
  if (mode == 1) {
    g.g_ml = g_miller;
    g.g_fe = g_final_exp;
    g.depth_ml = -1;  // a tuned depth belongs to the engine it was tuned with
  }
  return 0;
}

int blsgpu_get_engine(int* mode, int* g_miller, int* g_final_exp) {
  std::lock_guard<std::mutex> lk(g.mu);
  if (mode) *mode = g.mode;
  if (g_miller) *g_miller = g.g_ml;
  if (g_final_exp) *g_final_exp = g.g_fe;
  return 0;
}

int blsgpu_set_profiling(int on) {
  std::lock_guard<std::mutex> lk(g.mu);
  g.profiling = on != 0;
  for (int k = 0; k < 4; k++) g.phase_ms[k] = 0;
  g.phase_calls = 0;
  for (int k = 0; k < 3; k++) g.legacy_ms[k] = 0;
  g.legacy_calls = 0;
  return 0;
}

int blsgpu_legacy_phase_times(double ms[3], int* calls) {
  std::lock_guard<std::mutex> lk(g.mu);
  for (int k = 0; k < 3; k++) ms[k] = g.legacy_ms[k];
  if (calls) *calls = g.legacy_calls;
  return 0;
}

int blsgpu_phase_times(double ms[4], int* calls) {
  std::lock_guard<std::mutex> lk(g.mu);
  g_err.clear();
  for (auto& ev : g.ev_pending) {
    HIPCHK(hipEventSynchronize(ev[4]));
    for (int k = 0; k < 4; k++) {
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
      g.phase_ms[k] += t;
    }
    g.phase_calls++;
    for (hipEvent_t e : ev) g.ev_free.push_back(e);
  }
  g.ev_pending.clear();
  for (int k = 0; k < 4; k++) ms[k] = g.phase_ms[k];
  if (calls) *calls = g.phase_calls;
  return 0;
}

const char* blsgpu_last_error(void) { return g_err.c_str(); }

}  // extern "C"
