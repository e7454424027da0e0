// Multi-Miller loop over precomputed G2 line coefficients and the final exponentiation.
//
// Loop structure, coefficient order and line-evaluation argument order follow
// gpu-crypto/cuda/api.cu:22-52 (ic_bls12_381's multi_miller_loop with G2Prepared):
// 62 iterations over the bits of BLS_X >> 1 below the leading one; each iteration
// evaluates the doubling line (and the addition line on set bits) for every pair, then
// squares; a final doubling line follows; the result is conjugated because x < 0.
// The line for coefficient (c0, c1, c2) at P is mul_by_014(c2, c1 * P.x, c0 * P.y).
// The final exponentiation is the reference's chain (api.cu:54-78).
#pragma once
#include "tower.hpp"

namespace bls {

constexpr int NUM_COEFFS = 68;  // 63 doubling + 5 addition lines

struct G1Aff {
  Fp x, y;
};
struct Line {
  Fp2 c0, c1, c2;
};

__device__ __forceinline__ Fp12 ell(const Fp12& f, const Line& L, const G1Aff& p) {
  return mul_by_014(f, L.c2, mul_fp(L.c1, p.x), mul_fp(L.c0, p.y));
}

// bits of BLS_X >> 1 below its leading one, MSB first: bit i (61..0) of 0x6900800000008000
__device__ __forceinline__ bool miller_bit(int b) { return ((BLS_X >> 1) >> b) & 1; }

// Two-pair Miller loop; LineSrc(k) returns line k of each pair: L1(k), L2(k).
// skip1/skip2 drop a pair whose G1 point is the identity (ic_bls12_381 multi_miller_loop
// leaves f unchanged for such a pair).
template <class Src1, class Src2>
__device__ __forceinline__ Fp12 miller2(const G1Aff& p1, Src1 L1, const G1Aff& p2, Src2 L2,
                                        bool skip1 = false, bool skip2 = false) {
  Fp12 f = fp12_one();
  int ci = 0;
  for (int b = 61; b >= 0; b--) {
    if (!skip1) f = ell(f, L1(ci), p1);
    if (!skip2) f = ell(f, L2(ci), p2);
    ci++;
    if (miller_bit(b)) {
      if (!skip1) f = ell(f, L1(ci), p1);
      if (!skip2) f = ell(f, L2(ci), p2);
      ci++;
    }
    f = sqr(f);
  }
  if (!skip1) f = ell(f, L1(ci), p1);
  if (!skip2) f = ell(f, L2(ci), p2);
  return conj(f);
}

// f^x for x = -BLS_X, elements of the cyclotomic subgroup (api.cu:54-63)
__device__ __forceinline__ Fp12 cyc_exp(const Fp12& f) {
  Fp12 t = f;  // leading bit of BLS_X
  for (int i = 62; i >= 0; i--) {
    t = cyc_sqr(t);
    if ((BLS_X >> i) & 1) t = mul(t, f);
  }
  return conj(t);
}

// easy part f^((p^6-1)(p^2+1)) (api.cu:66-67)
__device__ __forceinline__ Fp12 final_exp_easy(const Fp12& f) {
  Fp12 t2 = mul(conj(f), inv(f));
  return mul(frob(frob(t2)), t2);
}
// hard part (api.cu:68-77) on the output of the easy part
__device__ __forceinline__ Fp12 final_exp_hard(const Fp12& t2in) {
  Fp12 t2 = t2in;
  Fp12 f = t2;
  Fp12 t1 = conj(cyc_sqr(t2));
  Fp12 t3 = cyc_exp(t2);
  Fp12 t4 = mul(t1, t3);
  t1 = cyc_exp(t4);
  t4 = conj(t4);
  f = mul(f, t4);
  t4 = cyc_sqr(t3);
  Fp12 t0 = cyc_exp(t1);
  t3 = mul(t3, t0);
  t3 = frob(frob(t3));
  f = mul(f, t3);
  t4 = mul(t4, cyc_exp(t0));
  f = mul(f, cyc_exp(t4));
  t4 = mul(t4, conj(t2));
  t2 = mul(t2, t1);
  t2 = frob(frob(frob(t2)));
  f = mul(f, t2);
  t4 = frob(t4);
  f = mul(f, t4);
  return f;
}

}  // namespace bls
