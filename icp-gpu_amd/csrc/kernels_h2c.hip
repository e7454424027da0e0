// Hash to G1 (RFC 9380 BLS12381G1_XMD:SHA-256_SSWU_RO_), four lanes per message.
#include <hip/hip_runtime.h>

#include "h2c.hpp"
#include "kernels.hpp"

namespace bls {

// message i = msgs[offs[i] .. offs[i+1]) -> H(m) (or -H(m) when negate != 0) as an
// affine point in reference form (12 u64: x then y); status[i] = 0, or 1 and zeros for
// the identity (nullable).  Four lanes per message (a quad of a wavefront): all expand
// the message; lanes 0 | 1 map u0 | u1 onto the curve and evaluate the isogeny's x map,
// lanes 2 | 3 the same field elements and its y map; the quad then adds the two points
// and clears the cofactor together (the products of each step split over its lanes);
// lane 0 stores.
constexpr int H2C_BLOCK = 64;
__global__ __launch_bounds__(H2C_BLOCK, 1) void k_hash_to_g1(const uint8_t* __restrict__ msgs,
                                                         const int64_t* __restrict__ offs, int n,
                                                         const uint8_t* __restrict__ dst, int dlen,
                                                         u64* __restrict__ out, i32* __restrict__ status,
                                                         int negate) {
  const int t = blockIdx.x * H2C_BLOCK + threadIdx.x;
  // quads past the last message hash it again and store nothing: every lane of the
  // wavefront stays busy (one with 8 or fewer active lanes issues about three times
  // slower per instruction on this GPU), and the four lanes of a quad run together
  const bool live = (t >> 2) < n;
  const int i = live ? t >> 2 : n - 1;
  const int q = t & 3;
  const int64_t o = offs[i];
  Fp u0, u1;
  hash_to_field2(msgs + o, (int)(offs[i + 1] - o), dst, dlen, u0, u1);
  const Prj mine = map_to_g1_curve<FieldInline>((q & 1) ? u1 : u0, q);
  const Prj other = {pair_swap(mine.x), pair_swap(mine.y), pair_swap(mine.z)};
  // Q0 and Q1 by selects, and one call: the quad must run the steps together (a branch
  // per lane would mask lanes off under a DPP read)
  const bool odd = q & 1;
  const Prj q0 = {pick(odd, other.x, mine.x), pick(odd, other.y, mine.y), pick(odd, other.z, mine.z)};
  const Prj q1 = {pick(odd, mine.x, other.x), pick(odd, mine.y, other.y), pick(odd, mine.z, other.z)};
  Fp x, y;
  const bool ok = clear_cofactor_affine(q0, q1, q, x, y);
  if (q != 0 || !live) return;
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
  const long long threads = 4LL * n;
  hipLaunchKernelGGL(k_hash_to_g1, dim3((unsigned)((threads + H2C_BLOCK - 1) / H2C_BLOCK)), dim3(H2C_BLOCK), 0, s, msgs,
                     offs, n, dst, dlen, out, status, negate);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace bls
