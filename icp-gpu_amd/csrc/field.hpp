// BLS12-381 base field Fp on gfx950: 14 signed 28-bit limbs, Montgomery radix R' = 2^392.
//
// The reference computes Fp with sppark's mont_t on 12 x 32-bit limbs, R = 2^384,
// canonical after every operation (gpu-crypto/cuda/vendor/sppark/mont_t.cuh:402-422,
// 632-668).  This file computes the same field with a layout chosen for CDNA:
//   * radix 2^28 leaves 8 bits of headroom per 64-bit column, so a Montgomery
//     product is pure product scanning: one v_mad_i64_i32 per limb product, no
//     per-product carry handling (the 32-bit CIOS form needs ~3.4 extra VALU
//     instructions per product on this ISA);
//   * additions and subtractions are limb-wise and lazy (no carries);
//   * norm() restores the storage invariant with one carry pass and a fold of the
//     top limb by an estimated multiple of p.
// Values are only congruent mod p, never canonical, inside the pipeline.  canon()
// produces the unique representative in [0, p) wherever a result leaves the
// device (GT export, verdict, decompressed coordinates) or is compared, so every
// exported value is bit-identical to the reference's canonical encoding.
//
// Bounds (the proofs are in the comments of each routine):
//   N1 ("normalized"): limbs 0..12 in [0, 2^28), |limb 13| < 2^18.4, |value| < 2^382.4.
//     Every mul/sqr/norm output is N1.
//   N2: the sum or difference of two N1 values (|limbs| < 2^29, |value| < 2^383.4).
//   mul/sqr accept N2 operands.  Any expression with more terms must go through norm().
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "field_constants.hpp"

#ifndef BLS_MUL_INLINE
#define BLS_MUL_INLINE __forceinline__
#endif

namespace bls {

typedef uint32_t u32;
typedef uint64_t u64;
typedef int32_t i32;
typedef int64_t i64;

constexpr int NL = 14;
constexpr i32 MASK28 = (1 << 28) - 1;

struct Fp {
  i32 l[NL];
};

__device__ __forceinline__ Fp fp_const(const i32 (&c)[NL]) {
  Fp r;
#pragma unroll
  for (int i = 0; i < NL; i++) r.l[i] = c[i];
  return r;
}
__device__ __forceinline__ Fp fp_zero() {
  Fp r;
#pragma unroll
  for (int i = 0; i < NL; i++) r.l[i] = 0;
  return r;
}
__device__ __forceinline__ Fp fp_one() { return fp_const(ONE_M); }

// ---- lazy limb-wise linear operations (no carries) ----
__device__ __forceinline__ Fp add(const Fp& a, const Fp& b) {
  Fp r;
#pragma unroll
  for (int i = 0; i < NL; i++) r.l[i] = a.l[i] + b.l[i];
  return r;
}
__device__ __forceinline__ Fp sub(const Fp& a, const Fp& b) {
  Fp r;
#pragma unroll
  for (int i = 0; i < NL; i++) r.l[i] = a.l[i] - b.l[i];
  return r;
}
__device__ __forceinline__ Fp neg(const Fp& a) {
  Fp r;
#pragma unroll
  for (int i = 0; i < NL; i++) r.l[i] = -a.l[i];
  return r;
}
__device__ __forceinline__ Fp dbl(const Fp& a) { return add(a, a); }

// ---- normalization: carry pass + fold of the top limb ----
// Input: any limbs with |limb| < 2^30.  Output: N1.
// Carry pass: limbs 0..12 -> [0, 2^28) (arithmetic shifts give floor division, so
// negative limbs borrow correctly); the top limb t absorbs the carry, |t| < 2^31.
// Fold: q = floor(t * 40323 / 2^32) ~ t / 106513 where 106513 = P_L[13]; subtract
// q*p.  The top becomes t - q*P13 - borrow with |q| < 2^14.3, so |top| < 2^18.4.
// fold_top() is the fold alone, for a value already carried (limbs 0..12 in
// [0, 2^28), |top| < 2^31); norm() is carry pass + fold.
__device__ __forceinline__ Fp fold_top(const Fp& a) {
  Fp r = a;
  const i32 t = a.l[NL - 1];
  const i32 q = (i32)(((i64)t * 40323) >> 32);
  i64 acc = 0;
#pragma unroll
  for (int i = 0; i < NL - 1; i++) {
    acc += (i64)r.l[i] - (i64)q * P_L[i];
    r.l[i] = (i32)acc & MASK28;
    acc >>= 28;
  }
  r.l[NL - 1] = t - q * P_L[NL - 1] + (i32)acc;
  return r;
}
__device__ __forceinline__ Fp norm(const Fp& a) {
  Fp r;
  i32 c = 0;
#pragma unroll
  for (int i = 0; i < NL - 1; i++) {
    i32 v = a.l[i] + c;
    r.l[i] = v & MASK28;
    c = v >> 28;
  }
  r.l[NL - 1] = a.l[NL - 1] + c;
  return fold_top(r);
}

// Carry pass only: limbs 0..12 -> [0, 2^28), value unchanged (top limb absorbs the
// carry).  Enough to make a short lazy sum a valid multiplier operand.
__device__ __forceinline__ Fp carry(const Fp& a) {
  Fp r;
  i32 c = 0;
#pragma unroll
  for (int i = 0; i < NL - 1; i++) {
    i32 v = a.l[i] + c;
    r.l[i] = v & MASK28;
    c = v >> 28;
  }
  r.l[NL - 1] = a.l[NL - 1] + c;
  return r;
}

// ---- Montgomery product  a*b*2^-392 (mod p) ----
// Product scanning over 27 columns with the reduction interleaved (the m_k*p terms of
// column k are added in the same pass).  Column k holds <= 14 products |a_i b_j| < 2^58
// (N2 operands) plus <= 14 products m_i p_j < 2^56 plus a carry < 2^35, so
// |column| < 2^62.2: no overflow of the signed 64-bit accumulator.
// Each column is one chain of v_mad_i64_i32 / v_mad_u64_u32 into a single 64-bit
// accumulator that starts from the incoming carry: no separate partial sums to add.  A
// wavefront issues a dependent 64-bit multiply-add no slower than an independent one
// on this GPU (about 6.4 cycles apart with one wave per SIMD, measured), so the chain
// costs no latency and saves the partial-sum additions (471 VALU instructions per
// product instead of 543).
// Output: limbs 0..12 in [0, 2^28); value = (ab + Mp)/2^392 with 0 <= M < 2^392, so
// |value| < |ab|/2^392 + p < 2^380.8 for N2 inputs: N1.
__device__ BLS_MUL_INLINE Fp mul(const Fp& a, const Fp& b) {
  u32 m[NL];
  Fp r;
  i64 acc = 0;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    i64 t = acc;
#pragma unroll
    for (int i = 0; i <= k; i++) t += (i64)a.l[i] * b.l[k - i];
#pragma unroll
    for (int i = 0; i < k; i++) t += (i64)(i32)m[i] * P_L[k - i];
    m[k] = ((u32)t * NPRIME) & (u32)MASK28;
    t += (i64)(i32)m[k] * P_L[0];
    acc = t >> 28;  // exact: t == 0 mod 2^28
  }
#pragma unroll
  for (int k = NL; k < 2 * NL - 1; k++) {
    i64 t = acc;
#pragma unroll
    for (int i = k - (NL - 1); i < NL; i++) t += (i64)a.l[i] * b.l[k - i];
#pragma unroll
    for (int i = k - (NL - 1); i < NL; i++) t += (i64)(i32)m[i] * P_L[k - i];
    r.l[k - NL] = (i32)t & MASK28;
    acc = t >> 28;
  }
  r.l[NL - 1] = (i32)acc;
  return r;
}

// Squaring: cross products a_i a_j (i != j) once, with one operand doubled.
// |2 a_i| < 2^30, so a column holds <= 7 terms < 2^59 plus a square < 2^58 plus the
// reduction terms: |column| < 2^62.3.  Same single-chain columns as mul().
__device__ BLS_MUL_INLINE Fp sqr(const Fp& a) {
  i32 d[NL];
#pragma unroll
  for (int i = 0; i < NL; i++) d[i] = a.l[i] * 2;
  u32 m[NL];
  Fp r;
  i64 acc = 0;
#pragma unroll
  for (int k = 0; k < 2 * NL - 1; k++) {
    i64 t = acc;
    const int lo = k < NL ? 0 : k - (NL - 1);
#pragma unroll
    for (int i = lo; 2 * i < k; i++) t += (i64)d[i] * a.l[k - i];
    if ((k & 1) == 0) t += (i64)a.l[k / 2] * a.l[k / 2];
    if (k < NL) {
#pragma unroll
      for (int i = 0; i < k; i++) t += (i64)(i32)m[i] * P_L[k - i];
      m[k] = ((u32)t * NPRIME) & (u32)MASK28;
      t += (i64)(i32)m[k] * P_L[0];
      acc = t >> 28;
    } else {
#pragma unroll
      for (int i = k - (NL - 1); i < NL; i++) t += (i64)(i32)m[i] * P_L[k - i];
      r.l[k - NL] = (i32)t & MASK28;
      acc = t >> 28;
    }
  }
  r.l[NL - 1] = (i32)acc;
  return r;
}

// ---- canonical representative in [0, p) ----
// norm() gives |v| < 2^382.4 < 3.2p; add 4p, then conditionally subtract 4p, 2p, p.
__device__ __forceinline__ Fp sub_if_nonneg(const Fp& v, int k) {
  // v normalized nonnegative with limbs < 2^28; compute v - 2^k p exactly.
  Fp t;
  i64 acc = 0;
#pragma unroll
  for (int i = 0; i < NL; i++) {
    // 2^k * p in 28-bit limbs: shift P_L by k bits with carries between limbs
    i64 pk = ((i64)P_L[i] << k) & MASK28;
    if (i > 0) pk |= ((i64)P_L[i - 1] << k) >> 28;
    if (i == NL - 1) pk = ((i64)P_L[i] << k) | (((i64)P_L[i - 1] << k) >> 28);
    acc += (i64)v.l[i] - pk;
    if (i < NL - 1) {
      t.l[i] = (i32)acc & MASK28;
      acc >>= 28;
    } else {
      t.l[i] = (i32)acc;
    }
  }
  return t.l[NL - 1] >= 0 ? t : v;
}
__device__ __forceinline__ Fp canon(const Fp& a) {
  Fp v = norm(a);
  // + 4p, then carry-propagate so limbs are in [0, 2^28) and the value is positive
  Fp w;
  i64 acc = 0;
#pragma unroll
  for (int i = 0; i < NL; i++) {
    i64 pk = ((i64)P_L[i] << 2) & MASK28;
    if (i > 0) pk |= ((i64)P_L[i - 1] << 2) >> 28;
    if (i == NL - 1) pk = ((i64)P_L[i] << 2) | (((i64)P_L[i - 1] << 2) >> 28);
    acc += (i64)v.l[i] + pk;
    if (i < NL - 1) {
      w.l[i] = (i32)acc & MASK28;
      acc >>= 28;
    } else {
      w.l[i] = (i32)acc;
    }
  }
  w = sub_if_nonneg(w, 2);
  w = sub_if_nonneg(w, 1);
  w = sub_if_nonneg(w, 0);
  return w;
}
__device__ __forceinline__ bool is_zero(const Fp& a) {
  Fp c = canon(a);
  i32 acc = 0;
#pragma unroll
  for (int i = 0; i < NL; i++) acc |= c.l[i];
  return acc == 0;
}
__device__ __forceinline__ bool eq(const Fp& a, const Fp& b) { return is_zero(sub(a, b)); }

// ---- conversions at the device boundary ----
// 384-bit little-endian word array -> 14 x 28-bit limbs (no reduction).
__device__ __forceinline__ Fp unpack384(const u32 w[12]) {
  Fp r;
#pragma unroll
  for (int i = 0; i < NL; i++) {
    const int bit = 28 * i, wi = bit >> 5, sh = bit & 31;
    u64 lo = w[wi];
    u64 hi = (wi + 1 < 12) ? (u64)w[wi + 1] : 0;
    r.l[i] = (i32)(((lo | (hi << 32)) >> sh) & MASK28);
  }
  return r;
}
// canonical limbs -> 12 little-endian 32-bit words
__device__ __forceinline__ void pack384(const Fp& c, u32 w[12]) {
#pragma unroll
  for (int j = 0; j < 12; j++) {
    const int bit = 32 * j, li = bit / 28, sh = bit % 28;
    u64 v = (u64)(u32)c.l[li] >> sh;
    if (li + 1 < NL) v |= (u64)(u32)c.l[li + 1] << (28 - sh);
    if (li + 2 < NL && 56 - sh < 32) v |= (u64)(u32)c.l[li + 2] << (56 - sh);
    w[j] = (u32)v;
  }
}
// reference Montgomery form (R = 2^384, as stored by ic_bls12_381 and the C ABI)
// -> internal form: v*2^400*2^-392 = x*2^392.
__device__ __forceinline__ Fp from_ref(const u32 w[12]) { return mul(unpack384(w), fp_const(C_TO_INT)); }
__device__ __forceinline__ void to_ref(const Fp& a, u32 w[12]) {
  pack384(canon(mul(a, fp_const(C_TO_REF))), w);
}
// one Fp in the boundary form: 6 little-endian u64 words, Montgomery R = 2^384
__device__ __forceinline__ Fp ld_ref(const u64* p) {
  u32 w[12];
#pragma unroll
  for (int j = 0; j < 6; j++) {
    const u64 v = p[j];
    w[2 * j] = (u32)v;
    w[2 * j + 1] = (u32)(v >> 32);
  }
  return from_ref(w);
}
__device__ __forceinline__ void st_ref(u64* p, const Fp& a) {
  u32 w[12];
  to_ref(a, w);
#pragma unroll
  for (int j = 0; j < 6; j++) p[j] = (u64)w[2 * j] | ((u64)w[2 * j + 1] << 32);
}
// plain integer (e.g. decoded from big-endian bytes) -> internal Montgomery form
__device__ __forceinline__ Fp from_plain(const Fp& v) { return mul(v, fp_const(C_R2_INT)); }
// internal -> canonical plain integer
__device__ __forceinline__ Fp to_plain(const Fp& a) {
  Fp one = fp_zero();
  one.l[0] = 1;
  return canon(mul(a, one));
}

// ---- exponentiation by a fixed public exponent (12 x 32-bit words, LE) ----
__device__ __forceinline__ Fp pow_fixed(const Fp& a, const u32 (&e)[12]) {
  Fp r = fp_one(), b = a;
  for (int w = 0; w < 12; w++) {
    u32 bits = e[w];
    for (int i = 0; i < 32; i++) {
      if (bits & 1) r = mul(r, b);
      b = sqr(b);
      bits >>= 1;
    }
  }
  return r;
}
// Fermat inverse a^(p-2) (the reference's fp_inv, gpu-crypto/cuda/field_sppark.cuh:40-50)
__device__ __forceinline__ Fp inv(const Fp& a) { return pow_fixed(a, INV_EXP); }
// a^e for an exponent given as a 4-bit sliding window (a table of gen_constants.py);
// sqrt_cand: a^((p+1)/4) (ic-integration/gpu_crypto/cuda/decompress.cu:16-40), the
// same exponent by the 4-bit sliding window of SQRT_WIN (379 squarings and 79
// multiplications instead of 384 and 229): odd powers a, a^3, .., a^15 first, then one
// multiplication per window.  The window walk is the same on every lane (the exponent is
// a constant), so the digit selection is a scalar branch around register copies and
// the loop holds a single multiplier body.  The odd powers are eight named values, not
// an array: a table indexed by the (uniform, but not constant) digit went to private
// memory, one scratch load per limb at every window.
template <int LEN>
__device__ __forceinline__ Fp pow_win(const Fp& a, const uint8_t (&win)[LEN][2]) {
  const Fp a2 = sqr(a);
  const Fp p1 = a, p3 = mul(p1, a2), p5 = mul(p3, a2), p7 = mul(p5, a2);
  const Fp p9 = mul(p7, a2), p11 = mul(p9, a2), p13 = mul(p11, a2), p15 = mul(p13, a2);
  auto pick = [&](int d) -> Fp {
    switch (__builtin_amdgcn_readfirstlane(d) >> 1) {
      case 0: return p1;
      case 1: return p3;
      case 2: return p5;
      case 3: return p7;
      case 4: return p9;
      case 5: return p11;
      case 6: return p13;
      default: return p15;
    }
  };
  Fp r = pick(win[0][1]);
#pragma unroll 1
  for (int w = 1; w < LEN; w++) {
    const int nsq = win[w][0], d = win[w][1];
#pragma unroll 1
    for (int q = 0; q < nsq; q++) r = sqr(r);
    if (d) r = mul(r, pick(d));
  }
  return r;
}
__device__ __forceinline__ Fp sqrt_cand(const Fp& a) { return pow_win(a, SQRT_WIN); }

// ---- inverse by batched divsteps (Bernstein-Yang "safegcd", eprint 2019/266) ----
// Variable time (every input here is public).  Works on the canonical value
// c = a R' mod p, held in the same 14-limb radix-2^28 form as Fp: f = p, g = c, and
// d, e with d c = f, e c = g (mod p).  Each batch runs 28 divsteps on the low 28 bits
// of f and g alone (the decisions depend on nothing else), which yields a 2x2 matrix
// T with entries of magnitude <= 2^28; then (f, g) = T (f, g) / 2^28 exactly, and
// (d, e) = T (d, e) / 2^28 mod p, the division made exact by adding m p with
// m = -(T (d, e))_0 p^-1 mod 2^28.  g reaches 0 after about 29 batches for 381-bit
// inputs; then f = +-1 and d f = c^-1.  The Montgomery form of a^-1 is
// c^-1 R'^2 = mont(c^-1, R'^3).  All limb products are < 2^56 and every column sum
// fits an i64.  The reference inverts by Fermat exponentiation
// (gpu-crypto/cuda/field_sppark.cuh:40-50); the result is the same field element.
constexpr u32 P_INV28 = 0x30003;  // p^-1 mod 2^28
static_assert(((u64)(u32)P_L[0] * P_INV28) % (1u << 28) == 1, "P_INV28 is not p^-1 mod 2^28");

// 28 divsteps on the low bits; eta = -delta.  Zeros of g are skipped in bulk; an odd
// g takes one step that either swaps (f, g) <- (g, -f) or not and then adds f to g.
__device__ __forceinline__ void sg_divsteps(int& eta, u32 f, u32 g, i32& u, i32& v, i32& q, i32& r) {
  u = 1;
  v = 0;
  q = 0;
  r = 1;
  int i = 28;
  for (int guard = 0; guard < 64; guard++) {
    const int z = __builtin_ctz(g | (0xffffffffu << i));  // <= i
    g >>= z;
    u <<= z;
    v <<= z;
    eta -= z;
    i -= z;
    if (i == 0) break;
    if (eta < 0) {
      eta = -eta;
      const u32 t = f;
      f = g;
      g = 0u - t;
      const i32 tu = u, tv = v;
      u = q;
      v = r;
      q = -tu;
      r = -tv;
    }
    g += f;
    q += u;
    r += v;
  }
}

// (x, y) <- (a x + b y, c x + d y) / 2^28 for normalized limb vectors (limbs 0..12 in
// [0, 2^28), signed top); with_p adds m p to each row first so the division is exact.
__device__ __forceinline__ void sg_update(i32 (&x)[NL], i32 (&y)[NL], i32 a, i32 b, i32 c, i32 d, bool with_p) {
  i64 cx = (i64)a * x[0] + (i64)b * y[0];
  i64 cy = (i64)c * x[0] + (i64)d * y[0];
  i32 mx = 0, my = 0;
  if (with_p) {
    mx = (i32)((0u - (u32)cx) * P_INV28 & (u32)MASK28);
    my = (i32)((0u - (u32)cy) * P_INV28 & (u32)MASK28);
    if (mx >= (1 << 27)) mx -= 1 << 28;
    if (my >= (1 << 27)) my -= 1 << 28;
    cx += (i64)mx * P_L[0];
    cy += (i64)my * P_L[0];
  }
  cx >>= 28;
  cy >>= 28;
#pragma unroll
  for (int j = 1; j < NL; j++) {
    cx += (i64)a * x[j] + (i64)b * y[j];
    cy += (i64)c * x[j] + (i64)d * y[j];
    if (with_p) {
      cx += (i64)mx * P_L[j];
      cy += (i64)my * P_L[j];
    }
    x[j - 1] = (i32)cx & MASK28;
    y[j - 1] = (i32)cy & MASK28;
    cx >>= 28;
    cy >>= 28;
  }
  x[NL - 1] = (i32)cx;
  y[NL - 1] = (i32)cy;
}

// 0 maps to 0.  UNIFORM: the caller guarantees that every active lane holds the same
// value (in practice, one active lane); the divstep decisions then run on the scalar
// unit, with scalar branches, from values read out of the first active lane.
template <bool UNIFORM>
__device__ __noinline__ Fp inv_fp_t(const Fp& a) {
  const Fp c = canon(a);
  i32 f[NL], g[NL], d[NL], e[NL];
#pragma unroll
  for (int j = 0; j < NL; j++) {
    f[j] = P_L[j];
    g[j] = c.l[j];
    d[j] = 0;
    e[j] = j == 0;
  }
  int eta = -1;
  for (int batch = 0; batch < 64; batch++) {  // about 29 are needed; the cap only bounds the loop
    i32 any = 0;
#pragma unroll
    for (int j = 0; j < NL; j++) any |= g[j];
    u32 f0 = (u32)f[0], g0 = (u32)g[0];
    if (UNIFORM) {
      any = __builtin_amdgcn_readfirstlane(any);
      f0 = __builtin_amdgcn_readfirstlane(f0);
      g0 = __builtin_amdgcn_readfirstlane(g0);
    }
    if (any == 0) break;
    i32 u, v, q, r;
    sg_divsteps(eta, f0, g0, u, v, q, r);
    sg_update(f, g, u, v, q, r, false);
    sg_update(d, e, u, v, q, r, true);
  }
  Fp x;
  const bool negf = f[NL - 1] < 0;  // f = -1
#pragma unroll
  for (int j = 0; j < NL; j++) x.l[j] = negf ? -d[j] : d[j];
  return mul(x, fp_const(C_R3_INT));
}
__device__ __forceinline__ Fp inv_fp(const Fp& a) { return inv_fp_t<false>(a); }
__device__ __forceinline__ Fp inv_fp_uniform(const Fp& a) { return inv_fp_t<true>(a); }

// canonical plain value > (p-1)/2 (decompress.cu:107-135)
__device__ __forceinline__ bool lex_largest(const Fp& a) {
  u32 w[12];
  pack384(to_plain(a), w);
  bool gt = false, decided = false;
#pragma unroll
  for (int j = 11; j >= 0; j--) {
    if (!decided && w[j] != HALF_P32[j]) {
      gt = w[j] > HALF_P32[j];
      decided = true;
    }
  }
  return gt;
}

}  // namespace bls
