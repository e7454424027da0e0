// The device field arithmetic (icp-gpu_amd/csrc/field.hpp) compiled for the host with a
// stand-in HIP header, exported through a C ABI for the CPU tests.  Values cross the
// boundary in the reference form (6 x u64 little-endian, Montgomery R = 2^384).
#include <stdint.h>

#include "field.hpp"

using namespace bls;

static Fp in(const uint64_t a[6]) {
  u32 w[12];
  for (int j = 0; j < 6; j++) {
    w[2 * j] = (u32)a[j];
    w[2 * j + 1] = (u32)(a[j] >> 32);
  }
  return from_ref(w);
}
static void out(const Fp& x, uint64_t r[6]) {
  u32 w[12];
  to_ref(x, w);
  for (int j = 0; j < 6; j++) r[j] = (uint64_t)w[2 * j] | ((uint64_t)w[2 * j + 1] << 32);
}

extern "C" {
void fh_mul(const uint64_t a[6], const uint64_t b[6], uint64_t r[6]) { out(mul(in(a), in(b)), r); }
void fh_sqr(const uint64_t a[6], uint64_t r[6]) { out(sqr(in(a)), r); }
void fh_inv(const uint64_t a[6], uint64_t r[6]) { out(inv_fp(in(a)), r); }
void fh_inv_uniform(const uint64_t a[6], uint64_t r[6]) { out(inv_fp_uniform(in(a)), r); }
void fh_inv_fermat(const uint64_t a[6], uint64_t r[6]) { out(inv(in(a)), r); }
// inverse of a value given by raw internal limbs (any N2 operand), for bound tests
void fh_inv_limbs(const int32_t l[14], uint64_t r[6]) {
  Fp x;
  for (int j = 0; j < NL; j++) x.l[j] = l[j];
  out(inv_fp(x), r);
}
}
