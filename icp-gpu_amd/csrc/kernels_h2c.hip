// Hash to G1 (RFC 9380 BLS12381G1_XMD:SHA-256_SSWU_RO_), one thread per message.
#include <hip/hip_runtime.h>

#include "h2c.hpp"
#include "kernels.hpp"

namespace bls {

// message i = msgs[offs[i] .. offs[i+1]) -> H(m) (or -H(m) when negate != 0) as an
// affine point in reference form (12 u64: x then y); status[i] = 0, or 1 and zeros for
// the identity (nullable).  Two lanes per message (adjacent lanes of a wavefront): both
// expand the message, lane 0 maps u0 and lane 1 u1 onto the curve at the same time, the
// pair swaps its points across and both add them and clear the cofactor together (the
// doublings split over the pair); lane 0 stores.
constexpr int H2C_BLOCK = 64;
__global__ __launch_bounds__(H2C_BLOCK, 1) void k_hash_to_g1(const uint8_t* __restrict__ msgs,
                                                         const int64_t* __restrict__ offs, int n,
                                                         const uint8_t* __restrict__ dst, int dlen,
                                                         u64* __restrict__ out, i32* __restrict__ status,
                                                         int negate) {
  const int t = blockIdx.x * H2C_BLOCK + threadIdx.x;
  // pairs past the last message hash it again and store nothing: every lane of the
  // wavefront stays busy (one with 8 or fewer active lanes issues about three times
  // slower per instruction on this GPU)
  const bool live = (t >> 1) < n;
  const int i = live ? t >> 1 : n - 1;
  const bool odd = t & 1;
  const int64_t o = offs[i];
  Fp u0, u1;
  hash_to_field2(msgs + o, (int)(offs[i + 1] - o), dst, dlen, u0, u1);
  const Jac<Fp> q = map_to_g1_curve<FieldInline>(odd ? u1 : u0);
  Jac<Fp> other;
  other.x = pair_swap(q.x);
  other.y = pair_swap(q.y);
  other.z = pair_swap(q.z);
  // Q0 and Q1 by selects, and one call: the pair must run the doublings together (a
  // branch per lane parity would mask the partner off under every DPP read)
  const Jac<Fp> q0 = {pick(odd, other.x, q.x), pick(odd, other.y, q.y), pick(odd, other.z, q.z)};
  const Jac<Fp> q1 = {pick(odd, q.x, other.x), pick(odd, q.y, other.y), pick(odd, q.z, other.z)};
  Fp x, y;
  const bool ok = clear_cofactor_affine(q0, q1, odd, x, y);
  if (odd || !live) return;
  u64* r = out + (size_t)i * 12;
  if (ok) {
    st_ref(r, x);
    st_ref(r + 6, negate ? neg(y) : y);
  } else {
    for (int j = 0; j < 12; j++) r[j] = 0;
  }
  if (status) status[i] = ok ? 0 : 1;
}

int launch_hash_to_g1(const uint8_t* msgs, const int64_t* offs, int n, const uint8_t* dst, int dlen,
                      uint64_t* out, int32_t* status, int negate, hipStream_t s) {
  const long long threads = 2LL * n;
  hipLaunchKernelGGL(k_hash_to_g1, dim3((unsigned)((threads + H2C_BLOCK - 1) / H2C_BLOCK)), dim3(H2C_BLOCK), 0, s, msgs,
                     offs, n, dst, dlen, out, status, negate);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace bls
