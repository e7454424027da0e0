// Host-side dispatch over the lane counts of the lanes-per-item kernels: each G's
// launchers live in the translation unit kernels_grp.hip built with -DGRP_LANES=G.
#include "grp_config.hpp"

namespace bls {

int grp_num_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 1;
  }
  return cus;
}

int launch_miller_grp(int G, const int32_t* sig, const int32_t* nhm, const int32_t* gen, const int32_t* keys,
                      const int32_t* key_idx, int nkeys, int32_t* status, int32_t* fout, int M,
                      int* queue, int wg_per_cu, int depth, const SubgroupRiders& riders, hipStream_t s) {
  switch (G) {
#define ML_CASE(g)                                                                                                  \
  case g:                                                                                                           \
    return grp_launch_ml<g>(sig, nhm, gen, keys, key_idx, nkeys, status, fout, M, queue, wg_per_cu, depth,        \
                            riders, s);
    BLS_GRP_ML_LANES(ML_CASE)
#undef ML_CASE
    default: return -1;
  }
}
int miller_grp_depth(int G, bool queue, int wg_per_cu, bool riders, int depth, int M) {
  switch (G) {
#define ML_CASE(g) case g: return grp_ml_depth<g>(queue, wg_per_cu, riders, depth, M);
    BLS_GRP_ML_LANES(ML_CASE)
#undef ML_CASE
    default: return -1;
  }
}
int launch_final_exp_grp(int G, const int32_t* fin, const int32_t* status, const int32_t* sub_status, int32_t* result,
                         uint64_t* gt, int M, int* queue, int wg_per_cu, const SubgroupRiders& riders, int32_t* park,
                         long park_items, hipStream_t s) {
  switch (G) {
#define FE_CASE(g)                                                                                        \
  case g:                                                                                                 \
    return grp_launch_fe<g>(fin, status, sub_status, result, gt, M, queue, wg_per_cu, riders, park, park_items, s);
    BLS_GRP_FE_LANES(FE_CASE)
#undef FE_CASE
    default: return -1;
  }
}

bool grp_lanes_supported(int which, int G) {
  switch (G) {
#define ML_CASE(g) case g:
    BLS_GRP_ML_LANES(ML_CASE)
#undef ML_CASE
    if (which == 0) return true;
    break;
    default: break;
  }
  switch (G) {
#define FE_CASE(g) case g:
    BLS_GRP_FE_LANES(FE_CASE)
#undef FE_CASE
    return which == 1;
    default: return false;
  }
}

// wavefronts of a one-workgroup-per-tile launch of M items, against the GPU's SIMDs
static int grp_waves_t(int G, int M) {
  switch (G) {
#define WV_CASE(g) case g: return (M + Cfg<g>::IPW - 1) / Cfg<g>::IPW * Cfg<g>::W;
    BLS_GRP_ALL_LANES(WV_CASE)
#undef WV_CASE
    default: return 0;
  }
}
bool grp_fills_simds(int G, int M) { return grp_waves_t(G, M) >= 4 * grp_num_cus(); }

static int grp_threads(int G) {
  switch (G) {
#define TH_CASE(g) case g: return Cfg<g>::THREADS;
    BLS_GRP_ALL_LANES(TH_CASE)
#undef TH_CASE
    default: return 0;
  }
}
static bool grp_rides(int which, int G) {
  switch (G) {
#define RD_CASE(g) case g: return rides<g>(which);
    BLS_GRP_ALL_LANES(RD_CASE)
#undef RD_CASE
    default: return false;
  }
}
// Full workgroups of riders (one item per thread), all on SIMDs that the tiles leave
// without a wavefront.  A rider wavefront's run time (two 64-step double-and-add chains,
// about half of a pairing wavefront's) does not lengthen the launch.
int grp_rider_items(int which, int G, int M) {
  if (M <= 0 || !grp_lanes_supported(which, G) || !grp_rides(which, G)) return 0;
  const int threads = grp_threads(G);
  const long free_waves = 4L * grp_num_cus() - grp_waves_t(G, M);
  const long rider_waves = (long)(M + threads - 1) / threads * (threads / 64);
  return free_waves > 0 && rider_waves <= free_waves ? threads : 0;
}

int occupancy_grp(int which, int G) {
  if (!grp_lanes_supported(which, G)) return -1;
  if (which == 0) {
    switch (G) {
#define ML_CASE(g) case g: return grp_occ_ml<g>();
      BLS_GRP_ML_LANES(ML_CASE)
#undef ML_CASE
      default: return -1;
    }
  }
  // the final exponentiation's compact form (12 lanes) fits more workgroups per CU
  switch (G) {
#define FE_CASE(g) case g: return grp_occ_fe_compact<g>() > grp_occ_fe<g>() ? grp_occ_fe_compact<g>() : grp_occ_fe<g>();
    BLS_GRP_FE_LANES(FE_CASE)
#undef FE_CASE
    default: return -1;
  }
}

// homes of 6 x 168 words per resident item of the compact 12-lane final exponentiation
long final_exp_park_words() {
  return (long)grp_num_cus() * grp_occ_fe_compact<12>() * Cfg<12>::IPW * 6 * 168;
}

}  // namespace bls
