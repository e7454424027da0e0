// G1 signature decoding, split in two kernels so that the subgroup test can run
// concurrently with the Miller loop (which needs only the decoded coordinates):
//   k_g1_decode   flags, x < p, y = sqrt(x^3 + 4) with the sign choice -> point + status
//   k_g1_subgroup [x^2] P == (beta x, -y) on decoded points -> subgroup status (four
//                 lanes per item; k_g1_subgroup_lane: one, for large launches)
// Semantics as ic_bls12_381 G1Affine::from_compressed; the reference GPU decoder
// (ic-integration/gpu_crypto/cuda/decompress.cu:140-188) performs the decoding steps
// without any of the validity checks.  Compiled with the field multiplier inlined.
#include <hip/hip_runtime.h>

#include "curve.hpp"
#include "kernels.hpp"

namespace bls {

__device__ __forceinline__ void st_soa(i32* base, int w0, int M, int i, const Fp& v) {
#pragma unroll
  for (int j = 0; j < NL; j++) base[(size_t)(w0 + j) * M + i] = v.l[j];
}

// Lanes past the last item decode it again and store nothing (here and in the other
// thread-per-item kernels): a wavefront with 8 or fewer active lanes issues about three
// times slower per instruction on this GPU (a dependent chain of Fp products measured
// 6,020 cycles per product against 2,018 with 16 lanes or more), so a small batch's one
// partial wavefront would take three times as long.
// Above 256 wavefronts, workgroups of four wavefronts: the four go to the four SIMDs of
// one CU, so a launch of up to 1,024 wavefronts keeps one per SIMD.  With one-wavefront
// workgroups (212 registers, so two fit a SIMD) the dispatcher doubled them up on some
// SIMDs while others stayed empty: 34,368 items decoded in 0.58-0.70 ms instead of 0.38
// (batched form 21.27 -> 21.11 ms, 65,536 items 39.48 -> 38.91 ms).  Up to 256
// wavefronts, one-wavefront workgroups spread over that many CUs: four to a CU made the
// 4096-item decode 1 % slower.
constexpr int DECODE_BLOCK = 256;
__global__ __launch_bounds__(DECODE_BLOCK, 1) void k_g1_decode(const uint8_t* __restrict__ in, i32* __restrict__ pts,
                                                  i32* __restrict__ status, int M) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = t < M;
  const int i = live ? t : M - 1;
  u32 be[12];
  const uint8_t* b = in + (size_t)i * 48;
#pragma unroll
  for (int j = 0; j < 12; j++)
    be[j] = ((u32)b[4 * j] << 24) | ((u32)b[4 * j + 1] << 16) | ((u32)b[4 * j + 2] << 8) | b[4 * j + 3];
  Aff<Fp> P;
  const int32_t st = g1_decompress(be, P);
  if (!live) return;
  st_soa(pts, 0, M, i, P.x);
  st_soa(pts, NL, M, i, P.y);
  status[i] = st;
}

// The test of g1_subgroup_status on a lane quad: lanes 4k .. 4k + 3 of the grid hold
// item k and split the products of [|x|]([|x|]P) (quad_mul_absx: complete projective
// formulas, two product latencies per doubling where dbl-2009-l on one lane takes
// seven); [x^2]P == (beta x, -y) is then tested projectively, X == beta x Z and
// Y == -y Z with Z != 0, on every lane, and lane 0 of the quad writes the status.
// Quads past the last item repeat it, and items that did not decode test the generator
// instead, with no result taken from either: every lane runs the whole test (see
// k_g1_decode), and the four lanes of a quad always run together.
__global__ __launch_bounds__(64, 1) void k_g1_subgroup(const i32* __restrict__ pts, const i32* __restrict__ status,
                                                    i32* __restrict__ sub_status, int M) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = (t >> 2) < M;
  const int i = live ? t >> 2 : M - 1;
  const int q = t & 3;
  const bool decoded = status[i] == DEC_OK;
  Fp x = fp_const(G1_GEN_X_M), y = fp_const(G1_GEN_Y_M);
  if (decoded) {
    x = ld_soa(pts, 0, M, i);
    y = ld_soa(pts, NL, M, i);
  }
  Prj b = {x, y, fp_one()};
#pragma unroll 1
  for (int k = 0; k < 2; k++) b = quad_mul_absx(b, q);
  bool in = false;
  if (!is_zero(b.z)) in = eq(b.x, mul(mul(x, fp_const(BETA_M)), b.z)) && eq(b.y, neg(mul(y, b.z)));
  if (live && q == 0) sub_status[i] = in || !decoded ? DEC_OK : DEC_NOT_IN_SUBGROUP;
}

// The same test on one lane per point, for launches past one wavefront per SIMD in
// quads (more than 16,384 points: the batched forms, where the test runs beside the
// Miller loop and takes issue cycles from it).  No exchange and no repeated products:
// two thirds of the quad's instructions per point (9.9 k against 14.8 k VALU), which is
// what counts once every SIMD is busy.
__global__ __launch_bounds__(64, 1) void k_g1_subgroup_lane(const i32* __restrict__ pts, const i32* __restrict__ status,
                                                         i32* __restrict__ sub_status, int M) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = t < M;
  const int32_t st = g1_subgroup_status<FieldInline>(pts, status, M, live ? t : M - 1);
  if (live) sub_status[t] = st;
}

// status := sub_status where status is OK (for entry points that report decoding alone)
__global__ void k_merge_status(i32* __restrict__ status, const i32* __restrict__ sub_status, int M) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  if (status[i] == DEC_OK && sub_status[i] != DEC_OK) status[i] = sub_status[i];
}

// After a final exponentiation that ran before the subgroup test finished: items that
// decoded but are not in G1 get the subgroup status as their result, and a zero GT
// (the value every rejected item exports).
__global__ void k_apply_subgroup(int32_t* __restrict__ result, uint64_t* __restrict__ gt, const i32* __restrict__ status,
                                 const i32* __restrict__ sub_status, int M) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  if (status[i] != DEC_OK || sub_status[i] == DEC_OK) return;
  result[i] = sub_status[i];
  if (gt) {
    for (int j = 0; j < 72; j++) gt[(size_t)i * 72 + j] = 0;
  }
}

int launch_apply_subgroup(int32_t* result, uint64_t* gt, const int32_t* status, const int32_t* sub_status, int M,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_apply_subgroup, dim3((M + 255) / 256), dim3(256), 0, s, result, gt, status, sub_status, M);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
__global__ void k_zero_words(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
int launch_zero_words(int* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_words, dim3((n + 63) / 64), dim3(64), 0, s, p, n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
__global__ void k_iota(int32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}
int launch_iota(int32_t* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_g1_decode(const uint8_t* in, int32_t* pts, int32_t* status, int M, hipStream_t s) {
  const int block = M <= 256 * 64 ? 64 : DECODE_BLOCK;
  hipLaunchKernelGGL(k_g1_decode, dim3((M + block - 1) / block), dim3(block), 0, s, in, pts, status, M);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
// quads up to one wavefront per SIMD (16,384 points), one lane per point beyond
int launch_g1_subgroup(const int32_t* pts, const int32_t* status, int32_t* sub_status, int M, hipStream_t s) {
  if (M > 16384)
    hipLaunchKernelGGL(k_g1_subgroup_lane, dim3((M + 63) / 64), dim3(64), 0, s, pts, status, sub_status, M);
  else
    hipLaunchKernelGGL(k_g1_subgroup, dim3((4 * (size_t)M + 63) / 64), dim3(64), 0, s, pts, status, sub_status, M);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_merge_status(int32_t* status, const int32_t* sub_status, int M, hipStream_t s) {
  hipLaunchKernelGGL(k_merge_status, dim3((M + 255) / 256), dim3(256), 0, s, status, sub_status, M);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace bls
