// Extension tower Fp2 = Fp[u]/(u^2+1), Fp6 = Fp2[v]/(v^3-(1+u)), Fp12 = Fp6[w]/(w^2-v).
//
// Same tower, basis and element order as the reference (gpu-crypto/cuda/field_sppark.cuh:52-197;
// Fp12 stored c0.c0.c0, c0.c0.c1, ..., c1.c2.c1).  Because every value is only determined
// mod p, any formula for a product gives the same canonical result; the formulas here
// follow the reference's Karatsuba shapes so that operation counts match the roofline
// accounting, with the normalization discipline of field.hpp:
//   * every function returns components in N1 (norm() applied to each output);
//   * fp2_mul/fp2_sqr accept N2 components (sums of two N1 values);
//   * fp6/fp12 functions take N1 inputs; sums passed to them are carried first.
#pragma once
#include "field.hpp"

namespace bls {

struct Fp2 {
  Fp c0, c1;
};
struct Fp6 {
  Fp2 c0, c1, c2;
};
struct Fp12 {
  Fp6 c0, c1;
};

// ---------------- Fp2 ----------------
__device__ __forceinline__ Fp2 fp2_zero() { return {fp_zero(), fp_zero()}; }
__device__ __forceinline__ Fp2 fp2_one() { return {fp_one(), fp_zero()}; }
__device__ __forceinline__ Fp2 add(const Fp2& a, const Fp2& b) { return {add(a.c0, b.c0), add(a.c1, b.c1)}; }
__device__ __forceinline__ Fp2 sub(const Fp2& a, const Fp2& b) { return {sub(a.c0, b.c0), sub(a.c1, b.c1)}; }
__device__ __forceinline__ Fp2 neg(const Fp2& a) { return {neg(a.c0), neg(a.c1)}; }
__device__ __forceinline__ Fp2 dbl(const Fp2& a) { return add(a, a); }
__device__ __forceinline__ Fp2 conj(const Fp2& a) { return {a.c0, neg(a.c1)}; }
__device__ __forceinline__ Fp2 norm(const Fp2& a) { return {norm(a.c0), norm(a.c1)}; }
__device__ __forceinline__ Fp2 carry(const Fp2& a) { return {carry(a.c0), carry(a.c1)}; }
// times the non-residue 1+u: (c0 - c1) + (c0 + c1)u   (field_sppark.cuh:63)
__device__ __forceinline__ Fp2 mul_nr(const Fp2& a) { return {sub(a.c0, a.c1), add(a.c0, a.c1)}; }
__device__ __forceinline__ Fp2 mul_fp(const Fp2& a, const Fp& s) { return {mul(a.c0, s), mul(a.c1, s)}; }

// Karatsuba (field_sppark.cuh:66-69); accepts N2 components.
__device__ __forceinline__ Fp2 mul(const Fp2& a, const Fp2& b) {
  Fp t0 = mul(a.c0, b.c0);
  Fp t1 = mul(a.c1, b.c1);
  Fp s = mul(carry(add(a.c0, a.c1)), carry(add(b.c0, b.c1)));
  return {norm(sub(t0, t1)), norm(sub(sub(s, t0), t1))};
}
// (a0+a1)(a0-a1) + 2 a0 a1 u (field_sppark.cuh:70-79); accepts N2 components.
__device__ __forceinline__ Fp2 sqr(const Fp2& a) {
  Fp t = mul(a.c0, a.c1);
  Fp c0 = mul(carry(add(a.c0, a.c1)), carry(sub(a.c0, a.c1)));
  return {c0, norm(add(t, t))};
}
__device__ __forceinline__ Fp2 inv(const Fp2& a) {
  Fp n = norm(add(sqr(a.c0), sqr(a.c1)));
  Fp i = inv(n);
  return {mul(a.c0, i), neg(mul(a.c1, i))};
}

// ---------------- Fp6 ----------------
__device__ __forceinline__ Fp6 fp6_zero() { return {fp2_zero(), fp2_zero(), fp2_zero()}; }
__device__ __forceinline__ Fp6 fp6_one() { return {fp2_one(), fp2_zero(), fp2_zero()}; }
__device__ __forceinline__ Fp6 add(const Fp6& a, const Fp6& b) { return {add(a.c0, b.c0), add(a.c1, b.c1), add(a.c2, b.c2)}; }
__device__ __forceinline__ Fp6 sub(const Fp6& a, const Fp6& b) { return {sub(a.c0, b.c0), sub(a.c1, b.c1), sub(a.c2, b.c2)}; }
__device__ __forceinline__ Fp6 neg(const Fp6& a) { return {neg(a.c0), neg(a.c1), neg(a.c2)}; }
__device__ __forceinline__ Fp6 norm(const Fp6& a) { return {norm(a.c0), norm(a.c1), norm(a.c2)}; }
__device__ __forceinline__ Fp6 carry(const Fp6& a) { return {carry(a.c0), carry(a.c1), carry(a.c2)}; }
// times v: (nr * c2, c0, c1)
__device__ __forceinline__ Fp6 mul_v(const Fp6& a) { return {mul_nr(a.c2), a.c0, a.c1}; }

// field_sppark.cuh:93-100
__device__ __forceinline__ Fp6 mul(const Fp6& a, const Fp6& b) {
  Fp2 aa = mul(a.c0, b.c0), bb = mul(a.c1, b.c1), cc = mul(a.c2, b.c2);
  Fp2 t0 = mul(add(a.c1, a.c2), add(b.c1, b.c2));
  Fp2 t1 = mul(add(a.c0, a.c1), add(b.c0, b.c1));
  Fp2 t2 = mul(add(a.c0, a.c2), add(b.c0, b.c2));
  Fp2 c0 = add(aa, mul_nr(sub(sub(t0, bb), cc)));
  Fp2 c1 = add(sub(sub(t1, aa), bb), mul_nr(cc));
  Fp2 c2 = add(sub(sub(t2, aa), cc), bb);
  return {norm(c0), norm(c1), norm(c2)};
}
// field_sppark.cuh:102-110
__device__ __forceinline__ Fp6 inv(const Fp6& f) {
  Fp2 c0s = sqr(f.c0), c1s = sqr(f.c1), c2s = sqr(f.c2);
  Fp2 c01 = mul(f.c0, f.c1), c02 = mul(f.c0, f.c2), c12 = mul(f.c1, f.c2);
  Fp2 t0 = norm(sub(c0s, mul_nr(c12)));
  Fp2 t1 = norm(sub(mul_nr(c2s), c01));
  Fp2 t2 = norm(sub(c1s, c02));
  Fp2 x = norm(mul_nr(norm(add(mul(f.c2, t1), mul(f.c1, t2)))));
  Fp2 sc = norm(add(mul(f.c0, t0), x));
  Fp2 si = inv(sc);
  return {mul(t0, si), mul(t1, si), mul(t2, si)};
}
// Frobenius (field_sppark.cuh:112-121); constants from closed forms (1+u)^((p-1)/3) etc.
__device__ __forceinline__ Fp6 frob(const Fp6& f) {
  Fp2 c0 = conj(f.c0);
  Fp2 c1 = conj(f.c1);
  Fp2 c2 = conj(f.c2);
  // c1 * (0 + g u) = (-g c1.c1) + (g c1.c0) u ; c2 * (h + 0 u) = h c2
  Fp g = fp_const(FROB6_C1_1), h = fp_const(FROB6_C2_0);
  Fp2 n1 = {neg(mul(c1.c1, g)), mul(c1.c0, g)};
  Fp2 n2 = {mul(c2.c0, h), mul(c2.c1, h)};
  return {c0, n1, n2};
}
// sparse products (field_sppark.cuh:123-134)
__device__ __forceinline__ Fp6 mul_by_01(const Fp6& s, const Fp2& c0, const Fp2& c1) {
  Fp2 aa = mul(s.c0, c0), bb = mul(s.c1, c1);
  Fp2 o0 = add(mul_nr(mul(s.c2, c1)), aa);
  Fp2 o1 = sub(sub(mul(add(c0, c1), add(s.c0, s.c1)), aa), bb);
  Fp2 o2 = add(mul(s.c2, c0), bb);
  return {norm(o0), norm(o1), norm(o2)};
}
__device__ __forceinline__ Fp6 mul_by_1(const Fp6& s, const Fp2& c1) {
  return {norm(mul_nr(mul(s.c2, c1))), mul(s.c0, c1), mul(s.c1, c1)};
}

// ---------------- Fp12 ----------------
__device__ __forceinline__ Fp12 fp12_one() { return {fp6_one(), fp6_zero()}; }
__device__ __forceinline__ Fp12 conj(const Fp12& a) { return {a.c0, neg(a.c1)}; }

// field_sppark.cuh:140-145
__device__ __forceinline__ Fp12 mul(const Fp12& a, const Fp12& b) {
  Fp6 aa = mul(a.c0, b.c0), bb = mul(a.c1, b.c1);
  Fp6 s = mul(carry(add(a.c0, a.c1)), carry(add(b.c0, b.c1)));
  return {norm(add(aa, mul_v(bb))), norm(sub(sub(s, aa), bb))};
}
// field_sppark.cuh:147-150
__device__ __forceinline__ Fp12 sqr(const Fp12& a) {
  Fp6 ab = mul(a.c0, a.c1);
  Fp6 s1 = carry(add(a.c0, a.c1));
  Fp6 s2 = carry(add(a.c0, mul_v(a.c1)));
  Fp6 t = mul(s2, s1);
  return {norm(sub(sub(t, ab), mul_v(ab))), norm(add(ab, ab))};
}
// field_sppark.cuh:152-156
__device__ __forceinline__ Fp12 inv(const Fp12& f) {
  Fp6 t = norm(sub(mul(f.c0, f.c0), mul_v(mul(f.c1, f.c1))));
  Fp6 ti = inv(t);
  return {mul(f.c0, ti), neg(mul(f.c1, ti))};
}
// field_sppark.cuh:158-166
__device__ __forceinline__ Fp12 frob(const Fp12& f) {
  Fp6 c0 = frob(f.c0), c1 = frob(f.c1);
  Fp2 co = {fp_const(FROB12_0), fp_const(FROB12_1)};
  return {c0, {mul(c1.c0, co), mul(c1.c1, co), mul(c1.c2, co)}};
}
// sparse product with a line: nonzero at w-basis positions 0, 1 and 4
// (field_sppark.cuh:168-172)
__device__ __forceinline__ Fp12 mul_by_014(const Fp12& f, const Fp2& c0, const Fp2& c1, const Fp2& c4) {
  Fp6 aa = mul_by_01(f.c0, c0, c1);
  Fp6 bb = mul_by_1(f.c1, c4);
  Fp6 s = carry(add(f.c1, f.c0));
  Fp6 t = mul_by_01(s, c0, carry(add(c1, c4)));
  return {norm(add(mul_v(bb), aa)), norm(sub(sub(t, aa), bb))};
}
// Granger-Scott squaring for elements of the cyclotomic subgroup (field_sppark.cuh:174-197).
__device__ __forceinline__ void fp4_sqr(const Fp2& a, const Fp2& b, Fp2& o0, Fp2& o1) {
  Fp2 t0 = sqr(a), t1 = sqr(b);
  o0 = norm(add(mul_nr(t1), t0));
  o1 = norm(sub(sub(sqr(carry(add(a, b))), t0), t1));
}
__device__ __forceinline__ Fp12 cyc_sqr(const Fp12& f) {
  Fp2 z0 = f.c0.c0, z4 = f.c0.c1, z3 = f.c0.c2, z2 = f.c1.c0, z1 = f.c1.c1, z5 = f.c1.c2;
  Fp2 A0, A1, B0, B1, C0, C1;
  fp4_sqr(z0, z1, A0, A1);
  fp4_sqr(z2, z3, B0, B1);
  fp4_sqr(z4, z5, C0, C1);
  // z' = 3A - 2z or 3A + 2z; folded by norm() so values cannot compound
  z0 = norm(add(dbl(sub(A0, z0)), A0));
  z1 = norm(add(dbl(add(A1, z1)), A1));
  z4 = norm(add(dbl(sub(B0, z4)), B0));
  z5 = norm(add(dbl(add(B1, z5)), B1));
  Fp2 nC1 = norm(mul_nr(C1));
  z2 = norm(add(dbl(add(nC1, z2)), nC1));
  z3 = norm(add(dbl(sub(C0, z3)), C0));
  return {{z0, z4, z3}, {z2, z1, z5}};
}

__device__ __forceinline__ bool is_one(const Fp12& f) {
  const Fp* v = &f.c0.c0.c0;
  bool ok = eq(v[0], fp_one());
#pragma unroll
  for (int i = 1; i < 12; i++) ok = ok && is_zero(v[i]);
  return ok;
}

}  // namespace bls
