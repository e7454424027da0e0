// Configurations of the lanes-per-item kernels (kernels_grp.hip), shared by the kernel
// translation units (one per lane count G, built with -DGRP_LANES=G) and the host-side
// dispatch (grp_dispatch.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace bls {

// Lane counts with a Miller-loop kernel and with a final-exponentiation kernel.
#define BLS_GRP_ML_LANES(X) X(8) X(12) X(13) X(16) X(20) X(24) X(32)
#define BLS_GRP_FE_LANES(X) X(9) X(12) X(16) X(18) X(24) X(27) X(32)
#define BLS_GRP_ALL_LANES(X) X(8) X(9) X(12) X(13) X(16) X(18) X(20) X(24) X(27) X(32)

// Layout of a lanes-per-item configuration G: W wavefronts per workgroup, L = G / W lanes
// of each wavefront per item, IPW = 64 / L items per workgroup.  Lane s (0 <= s < G) of
// an item sits in wavefront s / L at position item * L + s % L.
template <int G>
struct Cfg {
  static constexpr int W = (G == 24 || G == 32) ? 2 : 1;
  static constexpr int L = G / W;
  static constexpr int IPW = 64 / L;
  static constexpr int THREADS = 64 * W;
  static_assert(L * W == G, "lanes per item");
};

// Lane counts whose one-workgroup-per-tile grid leaves SIMDs free at batch sizes near a
// full GPU (8 / 9 / 12 lanes: 8, 7, 5 items per wavefront) have a second instantiation
// of each kernel that carries the subgroup test (RIDE); its register budget is that of
// one wavefront per SIMD, which the subgroup test's double-and-add uses.
template <int G>
constexpr bool rides(int which) { return which == 0 ? (G == 8 || G == 12) : (G == 9 || G == 12); }
// Lane counts with an instantiation compiled for one wavefront per SIMD: those that carry
// the subgroup test, and the 32-lane Miller loop, whose tile launches of up to 2048 items
// (one wavefront per SIMD) would otherwise share SIMDs two waves to one while others
// stay empty (the two-wave budget lets the dispatcher place them so).
template <int G>
constexpr bool one_wave_inst(int which) { return rides<G>(which) || (which == 0 && G == 32); }

// Key-line depth of the Miller loop: how many steps ahead of its use each step's line of
// the item's key is read from memory into registers (0: read at the step that stores it
// to LDS; 1: one step ahead; 2: two steps ahead).  A depth is available where the held
// words (LW = ceil(84 / G) per lane and step) fit the kernel's register budget: LW <= 4
// per step held with the two-wave budget, LW <= 7 with the one-wave budget.
template <int G>
constexpr bool ml_depth_ok(int D, bool one_wave) {
  constexpr int LW = (84 + G - 1) / G;
  return D == 0 || (D >= 1 && D <= 2 && D * LW <= (one_wave ? 14 : 8) && LW <= (one_wave ? 7 : 4));
}
// the depth a launch runs: the requested one if available, else the largest available
// below it; requested < 0 = the default (1 where available, else 0)
template <int G>
constexpr int ml_depth_for(bool one_wave, int requested) {
  int d = requested < 0 ? 1 : requested > 2 ? 2 : requested;
  while (d > 0 && !ml_depth_ok<G>(d, one_wave)) d--;
  return d;
}

// Per-G launchers and occupancy queries, instantiated in the translation unit of G.
template <int G>
int grp_launch_ml(const int32_t* sig, const int32_t* nhm, const int32_t* gen, const int32_t* keys,
                  const int32_t* key_idx, int nkeys, int32_t* status, int32_t* fout, int M,
                  int* queue, int wg_per_cu, int depth, const SubgroupRiders& riders, hipStream_t s);
// the depth grp_launch_ml runs for these arguments (M items)
template <int G>
int grp_ml_depth(bool queue, int wg_per_cu, bool riders, int depth, int M);
// park / park_items: homes of the compact final exponentiation (12 lanes, queue-fed
// launches of more workgroups per CU than the full-LDS kernel holds), for park_items
// resident items; without them such a launch runs the full-LDS kernel at its occupancy
template <int G>
int grp_launch_fe(const int32_t* fin, const int32_t* status, const int32_t* sub_status, int32_t* result, uint64_t* gt,
                  int M, int* queue, int wg_per_cu, const SubgroupRiders& riders, int32_t* park, long park_items,
                  hipStream_t s);
template <int G>
int grp_occ_ml();
template <int G>
int grp_occ_fe();
// resident workgroups per CU of the compact final exponentiation (0 where G has none)
template <int G>
int grp_occ_fe_compact();
// the compact 12-lane final exponentiation, in a translation unit of its own
// (kernels_grp.hip built with GRP_KIND 2): its occupancy, and a queue-fed launch of grid
// workgroups whose items have homes in park (6 x 168 words per item)
int fe_compact_occ();
int fe_compact_launch(int grid, const int32_t* fin, const int32_t* status, const int32_t* sub_status, int32_t* result,
                      uint64_t* gt, int M, int* queue, int32_t* park, hipStream_t s);

// number of compute units of the current device (queried once)
int grp_num_cus();

}  // namespace bls
