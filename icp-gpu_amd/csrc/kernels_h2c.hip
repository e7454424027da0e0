// Hash to G1 (RFC 9380 BLS12381G1_XMD:SHA-256_SSWU_RO_), one thread per message.
#include <hip/hip_runtime.h>

#include "h2c.hpp"
#include "kernels.hpp"

namespace bls {

__device__ __forceinline__ void st_ref_h2c(u64* p, const Fp& a) {
  u32 w[12];
  to_ref(a, w);
#pragma unroll
  for (int j = 0; j < 6; j++) p[j] = (u64)w[2 * j] | ((u64)w[2 * j + 1] << 32);
}

// message i = msgs[offs[i] .. offs[i+1]) -> H(m) (or -H(m) when negate != 0) as an
// affine point in reference form (12 u64: x then y); status[i] = 0, or 1 and zeros for
// the identity (nullable)
__global__ void k_hash_to_g1(const uint8_t* __restrict__ msgs, const int64_t* __restrict__ offs, int n,
                             const uint8_t* __restrict__ dst, int dlen, u64* __restrict__ out,
                             i32* __restrict__ status, int negate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t o = offs[i];
  Fp x, y;
  const bool ok = hash_to_g1(msgs + o, (int)(offs[i + 1] - o), dst, dlen, x, y);
  u64* q = out + (size_t)i * 12;
  if (ok) {
    st_ref_h2c(q, x);
    st_ref_h2c(q + 6, negate ? neg(y) : y);
  } else {
    for (int j = 0; j < 12; j++) q[j] = 0;
  }
  if (status) status[i] = ok ? 0 : 1;
}

int launch_hash_to_g1(const uint8_t* msgs, const int64_t* offs, int n, const uint8_t* dst, int dlen,
                      uint64_t* out, int32_t* status, int negate, hipStream_t s) {
  hipLaunchKernelGGL(k_hash_to_g1, dim3((n + 63) / 64), dim3(64), 0, s, msgs, offs, n, dst, dlen, out, status,
                     negate);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace bls
