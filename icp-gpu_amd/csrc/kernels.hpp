// Host-side launch wrappers of the verification kernels (kernels.hip).
// All pointers are device pointers; every wrapper returns 0 or -1 (launch error).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bls {

constexpr int KEY_TABLE_WORDS = 68 * 84;  // internal G2Prepared table of one key
constexpr int POINT_WORDS = 28;           // internal affine G1 point
constexpr int MILLER_WORDS = 168;         // internal Fp12

int launch_points_from_ref(const uint64_t* in, int32_t* out, int M, hipStream_t s);
// M tables of 2448 u64 (reference form) -> M internal tables of KEY_TABLE_WORDS, one after the other
int launch_lines_from_ref(const uint64_t* in, int32_t* out, int M, hipStream_t s);
int launch_table_to_ref(const int32_t* t, uint64_t* out, hipStream_t s);
int launch_g2_prepare(const uint64_t* q, int32_t* out, int K, hipStream_t s);
// n messages (bytes msgs[offs[i] .. offs[i+1])) -> H(m) (RFC 9380 hash to G1 with the
// given DST), or -H(m) when negate != 0, as n x 12 u64 affine points in reference form;
// status[i] (nullable) = 1 for the identity
int launch_hash_to_g1(const uint8_t* msgs, const int64_t* offs, int n, const uint8_t* dst, int dlen,
                      uint64_t* out, int32_t* status, int negate, hipStream_t s);
// K compressed G2 points (96 bytes each) -> reference-form affine points + decode status
int launch_g2_decompress(const uint8_t* in, uint64_t* q_ref, int32_t* status, int K, int check_subgroup,
                         hipStream_t s);
int launch_points_to_ref(const int32_t* pts, const int32_t* status, uint64_t* out, int M, hipStream_t s);
int launch_miller(const int32_t* sig, const int32_t* nhm, const int32_t* g2lines, const int32_t* pklines,
                  const int32_t* key_idx, int nkeys, int32_t* sig_status, int32_t* fout, int M, int block,
                  hipStream_t s);
int launch_final_exp(const int32_t* fin, const int32_t* sig_status, int32_t* result, uint64_t* gt_out, int M,
                     int block, hipStream_t s);

// The G1 subgroup test of a batch carried by a pairing launch: `per` items (one thread
// each) in every workgroup past the launch's tiles, out[i] = subgroup status of item i of
// the decoded points pts / status (per = 0: none).
struct SubgroupRiders {
  const int32_t* pts = nullptr;
  const int32_t* status = nullptr;
  int32_t* out = nullptr;
  int per = 0;
};

// lanes-per-pairing kernels (kernels_grp.hip); G = lanes per item (Miller: 8, 12, 13, 16, 20,
// 24, 32; final exponentiation: 9, 12, 16, 18, 24, 27, 32; 24 and 32 run two wavefronts per
// workgroup).  Return -1 for an unsupported G or a launch error.
// queue == nullptr: one workgroup per tile of its items, then ceil(M / riders.per)
// workgroups of subgroup tests.  Otherwise a persistent grid (at most wg_per_cu
// workgroups per CU, 0 = the occupancy limit) pops tiles from the device counter *queue,
// which must be zero at launch; riders must then be empty.
// depth: the Miller loop's key-line depth (steps ahead its key lines are read into
// registers: 0, 1, 2; -1 = the default of G); a depth the instantiation's register budget
// does not hold runs as the largest one below it (miller_grp_depth gives the one it runs).
int launch_miller_grp(int G, const int32_t* sig, const int32_t* nhm, const int32_t* gen, const int32_t* keys,
                      const int32_t* key_idx, int nkeys, int32_t* status, int32_t* fout, int M,
                      int* queue, int wg_per_cu, int depth, const SubgroupRiders& riders, hipStream_t s);
int miller_grp_depth(int G, bool queue, int wg_per_cu, bool riders, int depth, int M);
int launch_final_exp_grp(int G, const int32_t* fin, const int32_t* status, const int32_t* sub_status, int32_t* result,
                         uint64_t* gt, int M, int* queue, int wg_per_cu, const SubgroupRiders& riders, int32_t* park,
                         long park_items, hipStream_t s);
// int32 words of the homes a queue-fed final exponentiation may need for the values it
// keeps out of LDS (its compact form), at any workgroups per CU; 0 if none
long final_exp_park_words();
// Items per subgroup-test workgroup that a one-workgroup-per-tile launch of M items at G
// lanes per item (which: 0 Miller, 1 final exp) can carry on the SIMDs its tiles leave
// without a wavefront; 0 if those SIMDs cannot hold the test of M items.
int grp_rider_items(int which, int G, int M);
// whether a lanes-per-item kernel (which: 0 Miller, 1 final exp) exists for G lanes
bool grp_lanes_supported(int which, int G);
// whether one workgroup per tile of M items at G lanes per item gives every SIMD a wavefront
bool grp_fills_simds(int G, int M);
// resident workgroups per CU of a grouped kernel (which: 0 Miller, 1 final exp), -1 if G is unsupported
int occupancy_grp(int which, int G);
// p[0..n) = 0 by a kernel (safe inside captured graphs)
int launch_zero_words(int* p, int n, hipStream_t s);
// p[i] = i for i < n
int launch_iota(int32_t* p, int n, hipStream_t s);

// G1 decoding in two parts (kernels_g1.hip): decode (flags, range, square root, sign)
// and the subgroup test, which may run concurrently with the Miller loop.
int launch_g1_decode(const uint8_t* in, int32_t* pts, int32_t* status, int M, hipStream_t s);
int launch_g1_subgroup(const int32_t* pts, const int32_t* status, int32_t* sub_status, int M, hipStream_t s);
int launch_merge_status(int32_t* status, const int32_t* sub_status, int M, hipStream_t s);
// result := sub_status (and GT := 0) for items that decoded but failed the subgroup test
int launch_apply_subgroup(int32_t* result, uint64_t* gt, const int32_t* status, const int32_t* sub_status, int M,
                          hipStream_t s);

}  // namespace bls
