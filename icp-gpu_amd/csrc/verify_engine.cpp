// Kernel sequences for one verification batch (see verify_engine.hpp).
#include "verify_engine.hpp"

#include "kernels.hpp"

namespace bls {

int engine_legacy(const EngineArgs& a, hipStream_t s) {
  if (launch_points_from_ref(a.sig_ref, a.sig_pts, a.M, s)) return -1;
  if (launch_points_from_ref(a.nhm_ref, a.nhm_pts, a.M, s)) return -1;
  if (launch_lines_from_ref(a.g2_ref, a.g2_lines, 1, 0, s)) return -1;
  if (launch_lines_from_ref(a.pk_ref, a.pk_lines, a.M, 1, s)) return -1;
  if (launch_miller(a.sig_pts, a.nhm_pts, a.g2_lines, a.pk_lines, nullptr, 0, nullptr, a.fbuf, a.M, a.block, s))
    return -1;
  return launch_final_exp(a.fbuf, nullptr, a.results, a.gt, a.M, a.block, s);
}

static inline int mark(const EngineArgs& a, int k, hipStream_t s) {
  return a.ev ? (hipEventRecord(a.ev[k], s) == hipSuccess ? 0 : -1) : 0;
}

int engine_compressed(const EngineArgs& a, hipStream_t s) {
  if (mark(a, 0, s) || launch_g1_decompress(a.comp, a.sig_pts, a.status, a.M, 1, s)) return -1;
  if (mark(a, 1, s) || launch_points_from_ref(a.nhm_ref, a.nhm_pts, a.M, s)) return -1;
  if (mark(a, 2, s) ||
      launch_miller(a.sig_pts, a.nhm_pts, a.gen_lines, a.keys, a.key_idx, a.nkeys, a.status, a.fbuf, a.M,
                    a.block, s))
    return -1;
  if (mark(a, 3, s) || launch_final_exp(a.fbuf, a.status, a.results, a.gt, a.M, a.block, s)) return -1;
  return mark(a, 4, s);
}

}  // namespace bls
