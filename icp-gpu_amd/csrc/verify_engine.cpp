// Kernel sequences for one verification batch (see verify_engine.hpp).
#include "verify_engine.hpp"

#include "kernels.hpp"

namespace bls {

static int miller(const EngineArgs& a, const int32_t* gen, const int32_t* keys, const int32_t* key_idx, int nkeys,
                  int32_t* status, hipStream_t s, const SubgroupRiders& riders = {}) {
  if (a.mode == 0)
    return launch_miller(a.sig_pts, a.nhm_pts, gen, keys, key_idx, nkeys, status, a.fbuf, a.M, a.block, s);
  return launch_miller_grp(a.g_ml, a.sig_pts, a.nhm_pts, gen, keys, key_idx, nkeys, status, a.fbuf, a.M,
                           a.queue && a.wpc_ml >= 0 ? a.queue : nullptr, a.wpc_ml, a.depth_ml, riders, s);
}
static int final_exp(const EngineArgs& a, const int32_t* status, const int32_t* sub_status, hipStream_t s,
                     const SubgroupRiders& riders = {}) {
  if (a.mode == 0) {
    if (sub_status && launch_merge_status(const_cast<int32_t*>(status), sub_status, a.M, s)) return -1;
    return launch_final_exp(a.fbuf, status, a.results, a.gt, a.M, a.block, s);
  }
  return launch_final_exp_grp(a.g_fe, a.fbuf, status, sub_status, a.results, a.gt, a.M,
                              a.queue && a.wpc_fe >= 0 ? a.queue + 1 : nullptr, a.wpc_fe, riders, a.park,
                              a.park_items, s);
}

int engine_legacy_setup(const EngineArgs& a, hipStream_t cs) {
  if (launch_lines_from_ref(a.g2_ref, a.g2_lines, 1, cs)) return -1;
  // the M per-item tables are M keys, item i using key i
  return launch_iota(a.item_idx, a.M, cs);
}
int engine_legacy_part(const EngineArgs& a, int first, int n, hipStream_t cs, hipEvent_t ready, hipStream_t s) {
  // thread-per-item Miller values are word-major with stride M: that mode runs one part
  if (first < 0 || n <= 0 || first + n > a.M || (a.mode == 0 && n != a.M)) return -1;
  // the part's points form SoA blocks of their own (stride n) at its items' offset
  int32_t* sp = a.sig_pts + (size_t)first * POINT_WORDS;
  int32_t* np = a.nhm_pts + (size_t)first * POINT_WORDS;
  if (launch_points_from_ref(a.sig_ref + (size_t)first * 12, sp, n, cs)) return -1;
  if (launch_points_from_ref(a.nhm_ref + (size_t)first * 12, np, n, cs)) return -1;
  if (launch_lines_from_ref(a.pk_ref + (size_t)first * 2448, a.pk_lines + (size_t)first * KEY_TABLE_WORDS, n, cs))
    return -1;
  if (cs != s && (!ready || hipEventRecord(ready, cs) != hipSuccess || hipStreamWaitEvent(s, ready, 0) != hipSuccess))
    return -1;
  EngineArgs b = a;
  b.M = n;
  b.sig_pts = sp;
  b.nhm_pts = np;
  b.fbuf = a.fbuf + (size_t)first * MILLER_WORDS;
  return miller(b, a.g2_lines, a.pk_lines, a.item_idx + first, a.M, nullptr, s);
}
int engine_legacy_finish(const EngineArgs& a, hipStream_t s) { return final_exp(a, nullptr, nullptr, s); }
int engine_legacy(const EngineArgs& a, hipStream_t s) {
  if (engine_legacy_setup(a, s) || engine_legacy_part(a, 0, a.M, s, nullptr, s)) return -1;
  return engine_legacy_finish(a, s);
}

static inline int mark(const EngineArgs& a, int k, hipStream_t s) {
  return a.ev ? (hipEventRecord(a.ev[k], s) == hipSuccess ? 0 : -1) : 0;
}

// Otherwise: decode -> fork: [subgroup test on s2] || [Miller loop -> final exp on s] -> join ->
// subgroup verdicts applied.  The subgroup test does not feed the pairing, so in the
// lanes-per-item mode it overlaps the pairing kernels (which run at a raised wave
// priority, so the test takes the issue slots they leave).  It starts with the Miller
// loop, unless the Miller loop's one-workgroup-per-tile grid puts a wavefront on every
// SIMD and the final exponentiation's does not: then it starts with the final
// exponentiation, on the SIMDs that one leaves free (a wavefront sharing a SIMD with a
// lone pairing wavefront slows it by a third).  The thread-per-item mode joins before
// its final exponentiation, which reads the merged status.
static int launch_subgroup_fork(const EngineArgs& a, bool fork, hipStream_t s) {
  if (!fork) return launch_g1_subgroup(a.sig_pts, a.status, a.sub_status, a.M, s);
  if (hipEventRecord(a.fork, s) != hipSuccess || hipStreamWaitEvent(a.s2, a.fork, 0) != hipSuccess) return -1;
  if (launch_g1_subgroup(a.sig_pts, a.status, a.sub_status, a.M, a.s2)) return -1;
  return hipEventRecord(a.join, a.s2) == hipSuccess ? 0 : -1;
}
// decode -> Miller loop -> final exp, the subgroup test riding on the first pairing
// launch whose one-workgroup-per-tile grid leaves SIMDs enough for it: the Miller loop
// (the final exponentiation then reads the subgroup status), else the final
// exponentiation (its verdicts are corrected afterwards).
static int engine_riders(const EngineArgs& a, const SubgroupRiders& rml, const SubgroupRiders& rfe, hipStream_t s) {
  if (mark(a, 2, s) || miller(a, a.gen_lines, a.keys, a.key_idx, a.nkeys, a.status, s, rml)) return -1;
  if (mark(a, 3, s) || final_exp(a, a.status, rml.per ? a.sub_status : nullptr, s, rfe)) return -1;
  if (rfe.per && launch_apply_subgroup(a.results, a.gt, a.status, a.sub_status, a.M, s)) return -1;
  return mark(a, 4, s);
}

int engine_compressed(const EngineArgs& a, hipStream_t s) {
  if (mark(a, 0, s)) return -1;
  if (a.queue && a.mode == 1 && launch_zero_words(a.queue, 2, s)) return -1;
  if (launch_g1_decode(a.comp, a.sig_pts, a.status, a.M, s)) return -1;
  if (a.nhm_ready && hipStreamWaitEvent(s, a.nhm_ready, 0) != hipSuccess) return -1;
  if (mark(a, 1, s) || launch_points_from_ref(a.nhm_ref, a.nhm_pts, a.M, s)) return -1;
  const bool tiles_ml = !(a.queue && a.wpc_ml >= 0), tiles_fe = !(a.queue && a.wpc_fe >= 0);
  if (a.mode == 1) {
    SubgroupRiders rml, rfe;
    const int per_ml = tiles_ml ? grp_rider_items(0, a.g_ml, a.M) : 0;
    const int per_fe = !per_ml && tiles_fe ? grp_rider_items(1, a.g_fe, a.M) : 0;
    if (per_ml) rml = {a.sig_pts, a.status, a.sub_status, per_ml};
    if (per_fe) rfe = {a.sig_pts, a.status, a.sub_status, per_fe};
    if (per_ml || per_fe) return engine_riders(a, rml, rfe, s);
  }
  const bool fork = a.s2 && a.fork && a.join;
  const bool late = a.mode == 1;
  const bool with_fe = late && fork && tiles_ml && tiles_fe && grp_fills_simds(a.g_ml, a.M) &&
                       !grp_fills_simds(a.g_fe, a.M);
  if (!with_fe && launch_subgroup_fork(a, fork, s)) return -1;
  if (mark(a, 2, s) || miller(a, a.gen_lines, a.keys, a.key_idx, a.nkeys, a.status, s)) return -1;
  if (with_fe && launch_subgroup_fork(a, fork, s)) return -1;
  if (fork && !late && hipStreamWaitEvent(s, a.join, 0) != hipSuccess) return -1;
  if (mark(a, 3, s) || final_exp(a, a.status, late ? nullptr : a.sub_status, s)) return -1;
  if (late) {
    if (fork && hipStreamWaitEvent(s, a.join, 0) != hipSuccess) return -1;
    if (launch_apply_subgroup(a.results, a.gt, a.status, a.sub_status, a.M, s)) return -1;
  }
  return mark(a, 4, s);
}

int engine_prepare(const EngineArgs& a, hipStream_t s) {
  if (launch_g1_decode(a.comp, a.sig_pts, a.status, a.M, s)) return -1;
  if (launch_points_from_ref(a.nhm_ref, a.nhm_pts, a.M, s)) return -1;
  return launch_g1_subgroup(a.sig_pts, a.status, a.sub_status, a.M, s);
}
int engine_miller(const EngineArgs& a, hipStream_t s) {
  if (a.queue && a.mode == 1 && launch_zero_words(a.queue, 1, s)) return -1;
  return miller(a, a.gen_lines, a.keys, a.key_idx, a.nkeys, a.status, s);
}
int engine_final_exp(const EngineArgs& a, hipStream_t s) {
  if (a.queue && a.mode == 1 && launch_zero_words(a.queue + 1, 1, s)) return -1;
  return final_exp(a, a.status, a.sub_status, s);
}

int engine_miller_depth(const EngineArgs& a, bool riders) {
  if (a.mode != 1) return 0;
  return miller_grp_depth(a.g_ml, a.queue && a.wpc_ml >= 0, a.wpc_ml, riders, a.depth_ml, a.M);
}

int engine_decompress(const EngineArgs& a, int check_subgroup, hipStream_t s) {
  if (launch_g1_decode(a.comp, a.sig_pts, a.status, a.M, s)) return -1;
  if (!check_subgroup) return 0;
  if (launch_g1_subgroup(a.sig_pts, a.status, a.sub_status, a.M, s)) return -1;
  return launch_merge_status(a.status, a.sub_status, a.M, s);
}

}  // namespace bls
