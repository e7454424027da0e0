// Public-key registration kernels: G2 decompression with the subgroup test, and the line
// tables of prepared keys (G2Prepared), four lanes per key.  A translation unit of their
// own so that the field product is inlined: through the out-of-line product of
// kernels.hip every call saved and restored the live registers in private memory
// (2.4 KB per lane in the decompression).
#include <hip/hip_runtime.h>

#include "curve.hpp"
#include "kernels.hpp"

namespace bls {

constexpr int LINE_WORDS = 84;
constexpr int KEY_WORDS = NUM_COEFFS * LINE_WORDS;  // 5712

__device__ __forceinline__ void st_fp_aos(i32* p, const Fp& v) {
#pragma unroll
  for (int j = 0; j < NL; j++) p[j] = v.l[j];
}

// G2Prepared of K affine points given in reference form (24 u64: x.c0 x.c1 y.c0 y.c1)
// -> key-major internal tables (K x 5712 words).  Four lanes per key (lanes 4k ..
// 4k + 3): the steps' products split over the quad (g2_prepare_quad), and lane q stores
// Fp components q and q + 4 of every line.
// Quads past the last key (here and in k_g2_decompress) repeat it and store nothing: a
// wavefront with 8 or fewer active lanes issues about three times slower per
// instruction on this GPU, which would triple a single key's registration, and the four
// lanes of a quad always run together.
__global__ __launch_bounds__(64, 1) void k_g2_prepare(const u64* __restrict__ q, i32* __restrict__ out, int K) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = (t >> 2) < K;
  const int i = live ? t >> 2 : K - 1;
  const int qd = t & 3;
  const u64* s = q + (size_t)i * 24;
  Aff<Fp2> Q = {{ld_ref(s), ld_ref(s + 6)}, {ld_ref(s + 12), ld_ref(s + 18)}};
  i32* dst = out + (size_t)i * KEY_WORDS;
  g2_prepare_quad(Q, qd, [&](int k, const Line& L) {
    if (live) {
      i32* d = dst + k * LINE_WORDS;
      st_fp_aos(d + qd * NL, pick4(qd, L.c0.c0, L.c0.c1, L.c1.c0, L.c1.c1));
      if (qd < 2) st_fp_aos(d + (qd + 4) * NL, pick(qd == 1, L.c2.c1, L.c2.c0));
    }
  });
}

// compressed G2 public keys (K x 96 bytes) -> affine points in reference form (K x 24
// u64: x.c0 x.c1 y.c0 y.c1; zero where the encoding is rejected) + status per key,
// with the subgroup test (check_subgroup != 0).  Four lanes per key: each lane decodes
// the key (the square roots are one chain per key), then the quad runs the subgroup
// test with its products split four ways (g2_subgroup_quad); keys that did not decode
// test the zero point, with the result ignored, so that every lane runs the same code.
// Lane 0 of the quad stores.
__global__ __launch_bounds__(64, 1) void k_g2_decompress(const uint8_t* __restrict__ in, u64* __restrict__ q_ref, i32* __restrict__ status,
                                int K, int check_subgroup) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = (t >> 2) < K;
  const int i = live ? t >> 2 : K - 1;
  const int qd = t & 3;
  u32 be[24];
  const uint8_t* b = in + (size_t)i * 96;
#pragma unroll
  for (int j = 0; j < 24; j++)
    be[j] = ((u32)b[4 * j] << 24) | ((u32)b[4 * j + 1] << 16) | ((u32)b[4 * j + 2] << 8) | b[4 * j + 3];
  Aff<Fp2> Q;
  int32_t st = g2_decompress(be, Q);
  if (check_subgroup && !g2_subgroup_quad(Q, qd) && st == DEC_OK) st = DEC_NOT_IN_SUBGROUP;
  if (!live || qd != 0) return;
  u64* o = q_ref + (size_t)i * 24;
  if (st == DEC_OK) {
    st_ref(o, Q.x.c0);
    st_ref(o + 6, Q.x.c1);
    st_ref(o + 12, Q.y.c0);
    st_ref(o + 18, Q.y.c1);
  } else {
    for (int j = 0; j < 24; j++) o[j] = 0;
  }
  status[i] = st;
}

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

int launch_g2_decompress(const uint8_t* in, uint64_t* q_ref, int32_t* status, int K, int check_subgroup,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_g2_decompress, dim3(nblk(4LL * K, 64)), dim3(64), 0, s, in, q_ref, status, K, check_subgroup);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_g2_prepare(const uint64_t* q, int32_t* out, int K, hipStream_t s) {
  hipLaunchKernelGGL(k_g2_prepare, dim3(nblk(4LL * K, 64)), dim3(64), 0, s, q, out, K);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
}  // namespace bls
