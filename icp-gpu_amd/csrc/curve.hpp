// Curve arithmetic for G1 (y^2 = x^3 + 4 over Fp) and G2 (y^2 = x^3 + 4(1+u) over Fp2):
// Jacobian group law, G1 decompression with validity checks, the G1 subgroup test and
// the G2Prepared line coefficients consumed by the Miller loop.
#pragma once
#include "pairing.hpp"

namespace bls {

__device__ __forceinline__ bool is_zero(const Fp2& a) { return is_zero(a.c0) && is_zero(a.c1); }
__device__ __forceinline__ bool eq(const Fp2& a, const Fp2& b) { return is_zero(sub(a, b)); }
__device__ __forceinline__ Fp2 canon(const Fp2& a) { return {canon(a.c0), canon(a.c1)}; }
__device__ __forceinline__ Fp sqr_any(const Fp& a) { return sqr(a); }
__device__ __forceinline__ Fp2 sqr_any(const Fp2& a) { return sqr(a); }

template <class F>
struct Jac {
  F x, y, z;
};
template <class F>
struct Aff {
  F x, y;
};

// The field products of the group law: inline (FieldInline), or through one out-of-line
// multiplier and squarer (FpCalls, Fp only), which keeps the code of a long
// double-and-add loop small (the subgroup test runs beside the pairing kernels, on the
// CUs whose instruction cache their loops use).
struct FieldInline {
  template <class F>
  static __device__ __forceinline__ F mul(const F& a, const F& b) { return bls::mul(a, b); }
  template <class F>
  static __device__ __forceinline__ F sqr(const F& a) { return sqr_any(a); }
};
// (operands by value: they travel in registers, not through the stack)
static __device__ __noinline__ Fp mul_call(Fp a, Fp b) { return mul(a, b); }
static __device__ __noinline__ Fp sqr_call(Fp a) { return sqr(a); }
struct FpCalls {
  static __device__ __forceinline__ Fp mul(const Fp& a, const Fp& b) { return mul_call(a, b); }
  static __device__ __forceinline__ Fp sqr(const Fp& a) { return sqr_call(a); }
};

// dbl-2009-l (a = 0).  Z = 0 (identity) maps to Z3 = 0.
template <class F, class M = FieldInline>
__device__ __forceinline__ Jac<F> jdbl(const Jac<F>& p) {
  F A = M::sqr(p.x), B = M::sqr(p.y), C = M::sqr(B);
  F D = norm(dbl(sub(sub(M::sqr(carry(add(p.x, B))), A), C)));
  F E = norm(add(dbl(A), A));
  F Fv = M::sqr(E);
  Jac<F> r;
  r.x = norm(sub(Fv, dbl(D)));
  F C8 = norm(dbl(dbl(dbl(C))));  // 8C: normalized before the subtraction (limb range)
  r.y = norm(sub(M::mul(E, sub(D, r.x)), C8));
  r.z = norm(dbl(M::mul(p.y, p.z)));
  return r;
}
// madd-2007-bl: p Jacobian, q affine (not the identity).
template <class F, class M = FieldInline>
__device__ __forceinline__ Jac<F> jadd_aff(const Jac<F>& p, const Aff<F>& q, const F& one) {
  if (is_zero(p.z)) return {q.x, q.y, one};
  F Z1Z1 = M::sqr(p.z);
  F U2 = M::mul(q.x, Z1Z1);
  F S2 = M::mul(q.y, M::mul(p.z, Z1Z1));
  F H = norm(sub(U2, p.x));
  F R = norm(dbl(sub(S2, p.y)));
  if (is_zero(H)) {
    if (is_zero(R)) return jdbl<F, M>(p);
    Jac<F> o;
    o.x = one; o.y = one; o.z = norm(sub(one, one));
    return o;
  }
  F HH = M::sqr(H);
  F I = norm(dbl(dbl(HH)));
  F J = M::mul(H, I);
  F V = M::mul(p.x, I);
  Jac<F> r;
  r.x = norm(sub(sub(sub(M::sqr(R), J), V), V));
  r.y = norm(sub(M::mul(R, sub(V, r.x)), dbl(M::mul(p.y, J))));
  r.z = norm(sub(sub(M::sqr(carry(add(p.z, H))), Z1Z1), HH));
  return r;
}
template <class F>
__device__ __forceinline__ Aff<F> to_affine(const Jac<F>& p) {
  F zi = inv(p.z);
  F zi2 = sqr_any(zi);
  return {mul(p.x, zi2), mul(p.y, mul(zi2, zi))};
}
// scalar (8 x 32-bit words, little-endian) times affine point, MSB first
template <class F>
__device__ __forceinline__ Jac<F> smul(const Aff<F>& q, const u32 k[8], const F& one) {
  Jac<F> acc;
  acc.x = one; acc.y = one; acc.z = norm(sub(one, one));
  for (int i = 255; i >= 0; i--) {
    acc = jdbl(acc);
    if ((k[i >> 5] >> (i & 31)) & 1) acc = jadd_aff(acc, q, one);
  }
  return acc;
}


// ---- G2: membership and decompression (public keys) ----
// Square root in Fp2 by the norm (p = 3 mod 4; eprint 2012/685, Algorithm 8, the
// "complex method"), with two exponentiations in Fp where two exponentiations in Fp2
// took 2.6 times the products.  For a = a0 + a1 u:
//   lambda = sqrt(a0^2 + a1^2) (a is a square in Fp2 exactly when its norm is one in Fp);
//   delta = (a0 + lambda) / 2 and t = delta^((p+1)/4).  t^2 = delta gives the root
//   (t, a1 / (2t)); t^2 = -delta gives (a1 / (2t), t): the product of the two candidate
//   deltas (a0 +- lambda) / 2 is -(a1/2)^2, so exactly one is a square (-1 is not one) and
//   a1 / (2t) is the root of the other.  a1 = 0 takes delta = a0 and a zero partner
//   coordinate: (t, 0) when a0 is a square, else (0, t).
// The second exponentiation is s = delta^((p-3)/4), the same length: t = delta s, and
// t s = delta^((p-1)/2) = +-1, so 1/t = s (t s) needs no inversion (a zero delta gives
// zero, as the inversion gave; it occurs with a1 != 0 only for a non-square a).
// ok = the result squares back to a.  Which of the two roots comes out does not matter:
// the caller picks the sign by the sort flag.
__device__ __forceinline__ Fp2 sqrt2(const Fp2& a, bool& ok) {
  const bool a1z = is_zero(a.c1);
  const Fp n = norm(add(sqr(a.c0), sqr(a.c1)));
  const Fp lambda = sqrt_cand(n);
  const Fp delta = a1z ? a.c0 : mul(norm(add(a.c0, lambda)), fp_const(HALF_M));
  const Fp s = pow_win(delta, P34_WIN);
  const Fp t = mul(delta, s);
  const bool sq = eq(sqr(t), delta);
  Fp w = fp_zero();
  if (!a1z) w = mul(mul(a.c1, mul(s, mul(t, s))), fp_const(HALF_M));  // a1 / (2t)
  Fp2 r = sq ? Fp2{t, w} : Fp2{w, t};
  r = norm(r);
  ok = eq(sqr(r), a);
  return r;
}
// y lexicographically largest in Fp2 (the sort flag): y.c1 largest, or y.c1 == 0 and
// y.c0 largest
__device__ __forceinline__ bool lex_largest2(const Fp2& a) {
  return lex_largest(a.c1) || (is_zero(a.c1) && lex_largest(a.c0));
}
// psi(x, y) = (conj(x) cx, conj(y) cy), the untwist-Frobenius-twist endomorphism
__device__ __forceinline__ Aff<Fp2> g2_psi(const Aff<Fp2>& q) {
  const Fp2 cx = {fp_const(PSI_X0), fp_const(PSI_X1)}, cy = {fp_const(PSI_Y0), fp_const(PSI_Y1)};
  return {mul(conj(q.x), cx), mul(conj(q.y), cy)};
}

// ---- G1 decompression (ZCash encoding: bit7 compressed, bit6 infinity, bit5 sort) ----
enum DecodeStatus : int32_t {
  DEC_OK = 0,
  DEC_IDENTITY = 1,
  DEC_BAD_FLAGS = -1,
  DEC_X_NOT_CANONICAL = -2,
  DEC_NOT_ON_CURVE = -3,
  DEC_NOT_IN_SUBGROUP = -4,
  DEC_BAD_KEY = -5,  // set by the Miller kernel: key index outside the registry
};

// Jacobian + Jacobian (add-2007-bl), a = 0, with the doubling and identity cases.
template <class M = FieldInline>
__device__ __forceinline__ Jac<Fp> jadd_jac(const Jac<Fp>& p, const Jac<Fp>& q) {
  if (is_zero(p.z)) return q;
  if (is_zero(q.z)) return p;
  Fp z1z1 = M::sqr(p.z), z2z2 = M::sqr(q.z);
  Fp u1 = M::mul(p.x, z2z2), u2 = M::mul(q.x, z1z1);
  Fp s1 = M::mul(p.y, M::mul(q.z, z2z2)), s2 = M::mul(q.y, M::mul(p.z, z1z1));
  Fp h = norm(sub(u2, u1));
  Fp r = norm(dbl(sub(s2, s1)));
  if (is_zero(h)) {
    if (is_zero(r)) return jdbl<Fp, M>(p);
    return {fp_one(), fp_one(), fp_zero()};
  }
  Fp i = M::sqr(norm(dbl(h)));
  Fp j = M::mul(h, i);
  Fp v = M::mul(u1, i);
  Jac<Fp> o;
  o.x = norm(sub(sub(sub(M::sqr(r), j), v), v));
  o.y = norm(sub(M::mul(r, sub(v, o.x)), dbl(M::mul(s1, j))));
  o.z = M::mul(norm(sub(sub(M::sqr(carry(add(p.z, q.z))), z1z1), z2z2)), h);
  return o;
}
template <class M = FieldInline>
__device__ __forceinline__ Jac<Fp> mul_absx_jac(const Jac<Fp>& q) {
  Jac<Fp> acc = q;
#pragma unroll 1
  for (int i = 62; i >= 0; i--) {
    acc = jdbl<Fp, M>(acc);
    if ((BLS_X >> i) & 1) acc = jadd_jac<M>(acc, q);
  }
  return acc;
}
// ---- [|x|] P on a lane quad: complete projective formulas ----
// (the cofactor chain of the hash to G1 and the standalone subgroup test)
// Homogeneous coordinates (X : Y : Z) with x = X / Z, y = Y / Z; the identity is
// (0 : 1 : 0).  The complete formulas of Renes, Costello and Batina (eprint 2015/1060,
// algorithm 9 for a doubling and algorithm 7 for an addition, a = 0, b3 = 3b = 12) hold
// for every pair of points (P + P, P + (-P), the identity), and their products come in
// two steps of independent products.  The four lanes of a quad (lanes 4k .. 4k + 3, all
// holding the same points) take one product each per step and read the others' by DPP:
// a doubling is two product latencies (Y^2 | YZ | Z^2 | XY, then four products of
// those) where dbl-2009-l on one lane is seven, an addition four (six products per
// step, the last two of each step computed twice).  All four lanes of a quad take every
// branch together (their values are the same), so a DPP read always finds its source
// lane active.
__device__ __forceinline__ Fp pair_swap(const Fp& a) {  // lane i reads lane i ^ 1
  Fp r;
#pragma unroll
  for (int j = 0; j < NL; j++) r.l[j] = __builtin_amdgcn_update_dpp(a.l[j], a.l[j], 0xB1, 0xF, 0xF, false);
  return r;
}
template <int K>
__device__ __forceinline__ Fp quad_get(const Fp& a) {  // every lane reads lane K of its quad
  constexpr int ctrl = K | (K << 2) | (K << 4) | (K << 6);
  Fp r;
#pragma unroll
  for (int j = 0; j < NL; j++) r.l[j] = __builtin_amdgcn_update_dpp(a.l[j], a.l[j], ctrl, 0xF, 0xF, false);
  return r;
}
__device__ __forceinline__ Fp pick(bool c, const Fp& a, const Fp& b) {
  Fp r;
#pragma unroll
  for (int j = 0; j < NL; j++) r.l[j] = c ? a.l[j] : b.l[j];
  return r;
}
__device__ __forceinline__ Fp pick4(int q, const Fp& a0, const Fp& a1, const Fp& a2, const Fp& a3) {
  return pick(q & 2, pick(q & 1, a3, a2), pick(q & 1, a1, a0));
}
// c0 a0 + c1 a1 (+ c2 a2) for N1 values and small integer coefficients (|c| <= 128,
// uniform or per lane): one carry pass in 64 bits, then the fold of the top limb.  Limbs
// 0..12 sum to below 3 * 2^35; the top limb to below 3 * 2^25.4 < 2^31: N1 out.
__device__ __forceinline__ Fp lin3(const Fp& a0, int c0, const Fp& a1, int c1, const Fp& a2, int c2) {
  Fp r;
  i64 acc = 0;
#pragma unroll
  for (int i = 0; i < NL - 1; i++) {
    acc += (i64)c0 * a0.l[i] + (i64)c1 * a1.l[i] + (i64)c2 * a2.l[i];
    r.l[i] = (i32)acc & MASK28;
    acc >>= 28;
  }
  r.l[NL - 1] = (i32)(acc + (i64)c0 * a0.l[NL - 1] + (i64)c1 * a1.l[NL - 1] + (i64)c2 * a2.l[NL - 1]);
  return fold_top(r);
}
__device__ __forceinline__ Fp lin2(const Fp& a0, int c0, const Fp& a1, int c1) {
  Fp r;
  i64 acc = 0;
#pragma unroll
  for (int i = 0; i < NL - 1; i++) {
    acc += (i64)c0 * a0.l[i] + (i64)c1 * a1.l[i];
    r.l[i] = (i32)acc & MASK28;
    acc >>= 28;
  }
  r.l[NL - 1] = (i32)(acc + (i64)c0 * a0.l[NL - 1] + (i64)c1 * a1.l[NL - 1]);
  return fold_top(r);
}

struct Prj {
  Fp x, y, z;
};
// 2P on the quad (q = the lane's index in it).  X, Z in N1 or N2, Y in N2; out: X, Z
// N1, Y N2 (a sum of two products).
//   step 1: Y^2 | YZ | Z^2 | XY
//   step 2: 8 b3 Y^2 Z^2 | 8 Y^3 Z | (Y^2 - 3 b3 Z^2)(Y^2 + b3 Z^2) | 2 XY (Y^2 - 3 b3 Z^2)
//   X3 = the fourth, Y3 = the first + the third, Z3 = the second
__device__ __forceinline__ Prj quad_dbl(const Prj& p, int q) {
  const Fp s1 = mul(pick(q == 3, p.x, pick(q == 2, p.z, p.y)), pick(q == 1 || q == 2, p.z, p.y));
  const Fp y2 = quad_get<0>(s1), yz = quad_get<1>(s1), z2 = quad_get<2>(s1), xy = quad_get<3>(s1);
  // the lane's operands: c0 Y^2 + c1 YZ + c2 Z^2 and d0 Y^2 + d1 Z^2 + d2 XY
  const int c0 = q >= 2 ? 1 : 0, c1 = q == 1 ? 8 : 0, c2 = q == 0 ? 96 : (q >= 2 ? -36 : 0);
  const int d0 = q == 3 ? 0 : 1, d1 = q == 2 ? 12 : 0, d2 = q == 3 ? 2 : 0;
  const Fp s2 = mul(lin3(y2, c0, yz, c1, z2, c2), lin3(y2, d0, z2, d1, xy, d2));
  return {quad_get<3>(s2), add(quad_get<0>(s2), quad_get<2>(s2)), quad_get<1>(s2)};
}
// P + R on the quad, N1 coordinates in and out.
//   step 1: t0 = X1 X2 | t1 = Y1 Y2 | t2 = Z1 Z2 | t3 = (X1 + Y1)(X2 + Y2), then
//           t4 = (Y1 + Z1)(Y2 + Z2) | t5 = (X1 + Z1)(X2 + Z2) on lanes 0 | 1 (2 | 3 repeat)
//   with xy = t3 - t0 - t1, yz = t4 - t1 - t2, xz = t5 - t0 - t2:
//   step 2: A = yz b3 xz | B = xy (t1 - b3 t2) | C = b3 xz 3 t0 | D = (t1 - b3 t2)(t1 + b3 t2),
//           then E = 3 t0 xy | F = (t1 + b3 t2) yz
//   X3 = B - A, Y3 = D + C, Z3 = F + E
__device__ __forceinline__ Prj quad_add(const Prj& p, const Prj& r, int q) {
  const bool h = q & 1;
  const Fp s1 = mul(pick4(q, p.x, p.y, p.z, add(p.x, p.y)), pick4(q, r.x, r.y, r.z, add(r.x, r.y)));
  const Fp s1b = mul(pick(h, add(p.x, p.z), add(p.y, p.z)), pick(h, add(r.x, r.z), add(r.y, r.z)));
  const Fp t0 = quad_get<0>(s1), t1 = quad_get<1>(s1), t2 = quad_get<2>(s1), t3 = quad_get<3>(s1);
  const Fp t4 = quad_get<0>(s1b), t5 = quad_get<1>(s1b);
  const Fp xy = lin3(t3, 1, t0, -1, t1, -1), yz = lin3(t4, 1, t1, -1, t2, -1);
  const Fp xz3 = lin3(t5, 12, t0, -12, t2, -12);  // b3 xz
  const Fp ym = lin2(t1, 1, t2, -12), yp = lin2(t1, 1, t2, 12), x3 = lin2(t0, 3, t0, 0);
  const Fp s2 = mul(pick4(q, yz, xy, xz3, ym), pick4(q, xz3, ym, x3, yp));
  const Fp s2b = mul(pick(h, yp, x3), pick(h, yz, xy));
  const Fp A = quad_get<0>(s2), B = quad_get<1>(s2), C = quad_get<2>(s2), D = quad_get<3>(s2);
  const Fp E = quad_get<0>(s2b), F = quad_get<1>(s2b);
  return {norm(sub(B, A)), norm(add(D, C)), norm(add(F, E))};
}
// [|x|] P on the quad (P in N1), N1 out
__device__ __forceinline__ Prj quad_mul_absx(const Prj& p, int q) {
  Prj acc = p;
#pragma unroll 1
  for (int i = 62; i >= 0; i--) {
    acc = quad_dbl(acc, q);
    if ((BLS_X >> i) & 1) {
      acc.y = norm(acc.y);
      acc = quad_add(acc, p, q);
    }
  }
  acc.y = norm(acc.y);
  return acc;
}

// word w0 + j of item i in a word-major (SoA) array of M items
__device__ __forceinline__ Fp ld_soa(const i32* base, int w0, int M, int i) {
  Fp r;
#pragma unroll
  for (int j = 0; j < NL; j++) r.l[j] = base[(size_t)(w0 + j) * M + i];
  return r;
}
// Subgroup status of decoded item i of M (points word-major: x in words 0..13, y in
// 14..27): DEC_OK for items that did not decode (their status stands) and for points of
// G1, DEC_NOT_IN_SUBGROUP otherwise.  [x^2]P is computed as [|x|]([|x|]P) without
// leaving Jacobian coordinates, and compared with (beta x, -y) projectively:
// X == beta x Z^2 and Y == -y Z^3.  G1 membership of a point on the curve is
// phi(P) == [-x^2]P with phi(x, y) = (beta x, y) (eprint 2021/1130 sec. 6; ic_bls12_381's
// G1Affine::is_torsion_free).
// The field products go through FpCalls by default: beside the pairing kernels (the
// riders of a pairing launch) one copy of the multiplier keeps the loop small in the
// instruction cache they share; the standalone one-lane kernel inlines them.
// An item that did not decode tests the generator instead, its result ignored: every
// lane of the wavefront runs the whole test (a wavefront with 8 or fewer active lanes
// issues about three times slower per instruction on this GPU).
template <class Mul = FpCalls>
__device__ __forceinline__ int32_t g1_subgroup_status(const i32* pts, const i32* status, int M, int i) {
  const bool decoded = status[i] == DEC_OK;
  Fp x = fp_const(G1_GEN_X_M), y = fp_const(G1_GEN_Y_M);
  if (decoded) {
    x = ld_soa(pts, 0, M, i);
    y = ld_soa(pts, NL, M, i);
  }
  Jac<Fp> p1 = {x, y, fp_one()};
  Jac<Fp> a = mul_absx_jac<Mul>(p1);
  Jac<Fp> b = mul_absx_jac<Mul>(a);
  bool in = false;
  if (!is_zero(b.z)) {
    Fp z2 = Mul::sqr(b.z);
    in = eq(b.x, Mul::mul(Mul::mul(x, fp_const(BETA_M)), z2)) && eq(b.y, neg(Mul::mul(y, Mul::mul(z2, b.z))));
  }
  return in || !decoded ? DEC_OK : DEC_NOT_IN_SUBGROUP;
}
// Decode 48 big-endian bytes given as 12 big-endian-packed 32-bit words (w[0] holds bytes
// 0..3).  Semantics of ic_bls12_381 G1Affine::from_compressed: the reference GPU kernel
// (ic-integration/gpu_crypto/cuda/decompress.cu:140-188) performs the flag handling, the
// x^3+4 square root by (p+1)/4 and the sign choice, but none of the validity checks,
// which this routine adds except the subgroup test: this is from_compressed_unchecked,
// and the test runs apart on the decoded points (g1_subgroup_status, k_g1_subgroup).
__device__ __forceinline__ int32_t g1_decompress(const u32 be[12], Aff<Fp>& out) {
  const u32 top = be[0] >> 24;
  const bool cflag = top & 0x80, iflag = top & 0x40, sflag = top & 0x20;
  u32 w[12];  // little-endian 32-bit words of x
#pragma unroll
  for (int j = 0; j < 12; j++) w[j] = be[11 - j];
  w[11] &= 0x1fffffffu;
  out.x = fp_zero();
  out.y = fp_zero();
  if (!cflag) return DEC_BAD_FLAGS;
  if (iflag) {
    u32 acc = 0;
#pragma unroll
    for (int j = 0; j < 12; j++) acc |= w[j];
    return (!sflag && acc == 0) ? DEC_IDENTITY : DEC_BAD_FLAGS;
  }
  // x < p
  bool lt = false, decided = false;
#pragma unroll
  for (int j = 11; j >= 0; j--) {
    if (!decided && w[j] != P32[j]) {
      lt = w[j] < P32[j];
      decided = true;
    }
  }
  if (!lt) return DEC_X_NOT_CANONICAL;
  Fp x = from_plain(unpack384(w));
  Fp y2 = norm(add(mul(sqr(x), x), fp_const(B_G1_M)));
  Fp y = sqrt_cand(y2);
  if (!eq(sqr(y), y2)) return DEC_NOT_ON_CURVE;
  if (lex_largest(y) != sflag) y = neg(y);
  out.x = x;
  out.y = y;
  return DEC_OK;
}
// 96 bytes given as 24 big-endian-packed words: x.c1 (words 0..11, flags in the top
// three bits) then x.c0.  Semantics of ic_bls12_381 G2Affine::from_compressed_unchecked
// (the reference has no G2 decompression; its callers pass prepared lines); the
// subgroup test (g2_subgroup_quad) runs apart.
__device__ __forceinline__ int32_t g2_decompress(const u32 be[24], Aff<Fp2>& out) {
  const u32 top = be[0] >> 24;
  const bool cflag = top & 0x80, iflag = top & 0x40, sflag = top & 0x20;
  u32 w1[12], w0[12];  // little-endian 32-bit words of x.c1 and x.c0
#pragma unroll
  for (int j = 0; j < 12; j++) {
    w1[j] = be[11 - j];
    w0[j] = be[23 - j];
  }
  w1[11] &= 0x1fffffffu;
  out.x = {fp_zero(), fp_zero()};
  out.y = {fp_zero(), fp_zero()};
  if (!cflag) return DEC_BAD_FLAGS;
  if (iflag) {
    u32 acc = 0;
#pragma unroll
    for (int j = 0; j < 12; j++) acc |= w1[j] | w0[j];
    return (!sflag && acc == 0) ? DEC_IDENTITY : DEC_BAD_FLAGS;
  }
  auto below_p = [](const u32 (&w)[12]) {
    bool lt = false, decided = false;
#pragma unroll
    for (int j = 11; j >= 0; j--) {
      if (!decided && w[j] != P32[j]) {
        lt = w[j] < P32[j];
        decided = true;
      }
    }
    return lt;
  };
  if (!below_p(w1) || !below_p(w0)) return DEC_X_NOT_CANONICAL;
  const Fp2 x = {from_plain(unpack384(w0)), from_plain(unpack384(w1))};
  const Fp b4 = fp_const(B_G2_M);
  const Fp2 y2 = norm(add(mul(sqr(x), x), Fp2{b4, b4}));
  bool ok;
  Fp2 y = sqrt2(y2, ok);
  if (!ok) return DEC_NOT_ON_CURVE;
  if (lex_largest2(y) != sflag) y = neg(y);
  out.x = x;
  out.y = norm(y);
  return DEC_OK;
}

// 48-byte compressed encoding of an affine point (big-endian words as above)
__device__ __forceinline__ void g1_compress(const Aff<Fp>& p, bool identity, u32 be[12]) {
  if (identity) {
#pragma unroll
    for (int j = 0; j < 12; j++) be[j] = 0;
    be[0] = 0xc0000000u;
    return;
  }
  u32 w[12];
  pack384(to_plain(p.x), w);
#pragma unroll
  for (int j = 0; j < 12; j++) be[j] = w[11 - j];
  be[0] |= 0x80000000u;
  if (lex_largest(p.y)) be[0] |= 0x20000000u;
}

// ---- G2Prepared line coefficients (cuda/test_miller_v2.cu:152-228, eprint 2010/354
// Alg. 26 and 27).  The projective representation of R is not canonical, so the
// coefficient values depend on this exact formula sequence; it is the one that
// reproduces cuda/g2_coeffs.h and cuda/pk42_coeffs.h.  On a lane quad (four lanes
// holding the same key): each step's independent Fp2 products are split four ways, the
// steps' values read across the quad by DPP, so a doubling step is three product
// latencies (eleven products) and an addition step five (sixteen); every lane then holds
// the whole line and the next R. ----
struct G2Proj {
  Fp2 x, y, z;
};
template <int K>
__device__ __forceinline__ Fp2 quad_get2(const Fp2& a) {
  return {quad_get<K>(a.c0), quad_get<K>(a.c1)};
}
__device__ __forceinline__ Fp2 pick2(bool c, const Fp2& a, const Fp2& b) { return {pick(c, a.c0, b.c0), pick(c, a.c1, b.c1)}; }
__device__ __forceinline__ Fp2 pick4_2(int q, const Fp2& a0, const Fp2& a1, const Fp2& a2, const Fp2& a3) {
  return {pick4(q, a0.c0, a1.c0, a2.c0, a3.c0), pick4(q, a0.c1, a1.c1, a2.c1, a3.c1)};
}
__device__ __forceinline__ Line doubling_step_quad(G2Proj& r, int q) {
  // X^2 | Y^2 | Z^2 | (Z + Y)^2
  const Fp2 s1 = sqr(pick4_2(q, r.x, r.y, r.z, carry(add(r.z, r.y))));
  const Fp2 tmp0 = quad_get2<0>(s1), tmp1 = quad_get2<1>(s1), zsq = quad_get2<2>(s1), zy2 = quad_get2<3>(s1);
  const Fp2 tmp4 = norm(add(dbl(tmp0), tmp0));
  const Fp2 tmp6 = carry(add(r.x, tmp4));
  // tmp1^2 | (tmp1 + X)^2 | tmp4^2 | tmp6^2
  const Fp2 s2 = sqr(pick4_2(q, tmp1, carry(add(tmp1, r.x)), tmp4, tmp6));
  const Fp2 tmp2 = quad_get2<0>(s2), t13 = quad_get2<1>(s2), tmp5 = quad_get2<2>(s2), t66 = quad_get2<3>(s2);
  const Fp2 tmp3 = norm(dbl(sub(t13, add(tmp0, tmp2))));
  const Fp2 nx = norm(sub(sub(tmp5, tmp3), tmp3));
  const Fp2 nz = norm(sub(zy2, add(tmp1, zsq)));
  const Fp2 c2 = norm(sub(sub(t66, add(tmp0, tmp5)), dbl(dbl(tmp1))));
  // (tmp3 - nx) tmp4 | tmp4 zsq | nz zsq (lane 3 repeats lane 0)
  const Fp2 d = sub(tmp3, nx);
  const Fp2 s3 = mul(pick4_2(q, d, tmp4, nz, d), pick4_2(q, tmp4, zsq, zsq, tmp4));
  r.x = nx;
  r.y = norm(sub(quad_get2<0>(s3), norm(dbl(dbl(dbl(tmp2))))));
  r.z = nz;
  const Fp2 c1 = norm(neg(dbl(quad_get2<1>(s3))));
  const Fp2 c0 = norm(dbl(quad_get2<2>(s3)));
  return {c0, c1, c2};
}
__device__ __forceinline__ Line addition_step_quad(G2Proj& r, const Aff<Fp2>& qa, int q) {
  const bool h = q & 1;
  // Z^2 | qy^2 | (qy + Z)^2 (lane 3 repeats lane 0)
  const Fp2 s1 = sqr(pick4_2(q, r.z, qa.y, carry(add(qa.y, r.z)), r.z));
  const Fp2 zsq = quad_get2<0>(s1), ysq = quad_get2<1>(s1), yz2 = quad_get2<2>(s1);
  // t0 = zsq qx | t1 = ((qy + Z)^2 - qy^2 - zsq) zsq (lanes 2, 3 repeat 0, 1)
  const Fp2 s2 = mul(pick2(h, norm(sub(yz2, add(ysq, zsq))), zsq), pick2(h, zsq, qa.x));
  const Fp2 t2 = norm(sub(quad_get2<0>(s2), r.x));
  const Fp2 t6 = norm(sub(sub(quad_get2<1>(s2), r.y), r.y));
  // t3 = t2^2 | t9 = t6 qx | t6^2 | (Z + t2)^2
  const Fp2 zt = carry(add(r.z, t2));
  const Fp2 s3 = mul(pick4_2(q, t2, t6, t6, zt), pick4_2(q, t2, qa.x, t6, zt));
  const Fp2 t3 = quad_get2<0>(s3), t9 = quad_get2<1>(s3), t66 = quad_get2<2>(s3), zt2 = quad_get2<3>(s3);
  const Fp2 t4 = norm(dbl(dbl(t3)));
  const Fp2 nz = norm(sub(zt2, add(zsq, t3)));
  const Fp2 t10 = carry(add(qa.y, nz));
  // t5 = t4 t2 | t7 = t4 X | t10^2 | nz^2
  const Fp2 s4 = mul(pick4_2(q, t4, t4, t10, nz), pick4_2(q, t2, r.x, t10, nz));
  const Fp2 t5 = quad_get2<0>(s4), t7 = quad_get2<1>(s4), t1010 = quad_get2<2>(s4), nz2 = quad_get2<3>(s4);
  const Fp2 nx = norm(sub(sub(sub(t66, t5), t7), t7));
  // t8 = (t7 - nx) t6 | Y t5 (lanes 2, 3 repeat 0, 1)
  const Fp2 s5 = mul(pick2(h, r.y, sub(t7, nx)), pick2(h, t5, t6));
  r.x = nx;
  r.y = norm(sub(quad_get2<0>(s5), dbl(quad_get2<1>(s5))));
  r.z = nz;
  const Fp2 c2 = norm(sub(dbl(t9), norm(sub(sub(t1010, ysq), nz2))));
  const Fp2 c0 = norm(dbl(nz));
  const Fp2 c1 = norm(dbl(neg(t6)));
  return {c0, c1, c2};
}
// Emit the 68 lines of q in Miller-loop order through sink(k, line), on the quad.
template <class Sink>
__device__ __forceinline__ void g2_prepare_quad(const Aff<Fp2>& qa, int q, Sink sink) {
  G2Proj r = {qa.x, qa.y, fp2_one()};
  int n = 0;
#pragma unroll 1
  for (int b = 61; b >= 0; b--) {
    sink(n++, doubling_step_quad(r, q));
    if (miller_bit(b)) sink(n++, addition_step_quad(r, qa, q));
  }
  sink(n++, doubling_step_quad(r, q));
}

// ---- G2 membership on the quad: psi(Q) == [x]Q = -[|x|]Q (eprint 2021/1130 sec. 4,
// eprint 2022/352; ic_bls12_381's G2Affine::is_torsion_free).  [|x|]Q by the complete
// projective formulas of quad_dbl / quad_add over Fp2 for the twist y^2 = x^3 + 4(1 + u)
// (b3 = 12(1 + u), so the b3 terms become integer combinations of an Fp2 value's two
// components), then X == psi_x Z and -Y == psi_y Z with Z != 0. ----
struct Prj2 {
  Fp2 x, y, z;
};
// c0 a0 + c1 a1 + c2 a2 + c3 a3 for N1 values, |c| <= 128 (as lin3): N1 out
__device__ __forceinline__ Fp lin4(const Fp& a0, int c0, const Fp& a1, int c1, const Fp& a2, int c2, const Fp& a3,
                                   int c3) {
  Fp r;
  i64 acc = 0;
#pragma unroll
  for (int i = 0; i < NL - 1; i++) {
    acc += (i64)c0 * a0.l[i] + (i64)c1 * a1.l[i] + (i64)c2 * a2.l[i] + (i64)c3 * a3.l[i];
    r.l[i] = (i32)acc & MASK28;
    acc >>= 28;
  }
  r.l[NL - 1] = (i32)(acc + (i64)c0 * a0.l[NL - 1] + (i64)c1 * a1.l[NL - 1] + (i64)c2 * a2.l[NL - 1] +
                      (i64)c3 * a3.l[NL - 1]);
  return fold_top(r);
}
// k0 a + k1 (1 + u) b + k2 c for Fp2 values with N1 components ((1 + u) b =
// (b0 - b1) + (b0 + b1) u): N1 components out
__device__ __forceinline__ Fp2 lin_xi(const Fp2& a, int k0, const Fp2& b, int k1, const Fp2& c, int k2) {
  return {lin4(a.c0, k0, b.c0, k1, b.c1, -k1, c.c0, k2), lin4(a.c1, k0, b.c0, k1, b.c1, k1, c.c1, k2)};
}
__device__ __forceinline__ Fp2 lin3_2(const Fp2& a0, int c0, const Fp2& a1, int c1, const Fp2& a2, int c2) {
  return {lin3(a0.c0, c0, a1.c0, c1, a2.c0, c2), lin3(a0.c1, c0, a1.c1, c1, a2.c1, c2)};
}
// 2Q (as quad_dbl; Y out N2)
__device__ __forceinline__ Prj2 quad_dbl2(const Prj2& p, int q) {
  const Fp2 s1 = mul(pick2(q == 3, p.x, pick2(q == 2, p.z, p.y)), pick2(q == 1 || q == 2, p.z, p.y));
  const Fp2 y2 = quad_get2<0>(s1), yz = quad_get2<1>(s1), z2 = quad_get2<2>(s1), xy = quad_get2<3>(s1);
  // c0 Y^2 + c2 (1 + u) Z^2 + c1 YZ and d0 Y^2 + d1 (1 + u) Z^2 + d2 XY
  const int c0 = q >= 2 ? 1 : 0, c1 = q == 1 ? 8 : 0, c2 = q == 0 ? 96 : (q >= 2 ? -36 : 0);
  const int d0 = q == 3 ? 0 : 1, d1 = q == 2 ? 12 : 0, d2 = q == 3 ? 2 : 0;
  const Fp2 s2 = mul(lin_xi(y2, c0, z2, c2, yz, c1), lin_xi(y2, d0, z2, d1, xy, d2));
  return {quad_get2<3>(s2), add(quad_get2<0>(s2), quad_get2<2>(s2)), quad_get2<1>(s2)};
}
// P + R (as quad_add; N1 components in and out)
__device__ __forceinline__ Prj2 quad_add2(const Prj2& p, const Prj2& r, int q) {
  const bool h = q & 1;
  const Fp2 s1 = mul(pick4_2(q, p.x, p.y, p.z, add(p.x, p.y)), pick4_2(q, r.x, r.y, r.z, add(r.x, r.y)));
  const Fp2 s1b = mul(pick2(h, add(p.x, p.z), add(p.y, p.z)), pick2(h, add(r.x, r.z), add(r.y, r.z)));
  const Fp2 t0 = quad_get2<0>(s1), t1 = quad_get2<1>(s1), t2 = quad_get2<2>(s1), t3 = quad_get2<3>(s1);
  const Fp2 t4 = quad_get2<0>(s1b), t5 = quad_get2<1>(s1b);
  const Fp2 xy = lin3_2(t3, 1, t0, -1, t1, -1), yz = lin3_2(t4, 1, t1, -1, t2, -1);
  const Fp2 xz3 = lin_xi(t5, 0, lin3_2(t5, 1, t0, -1, t2, -1), 12, t5, 0);  // b3 xz
  const Fp2 ym = lin_xi(t1, 1, t2, -12, t2, 0), yp = lin_xi(t1, 1, t2, 12, t2, 0);
  const Fp2 x3 = lin3_2(t0, 3, t0, 0, t0, 0);
  const Fp2 s2 = mul(pick4_2(q, yz, xy, xz3, ym), pick4_2(q, xz3, ym, x3, yp));
  const Fp2 s2b = mul(pick2(h, yp, x3), pick2(h, yz, xy));
  const Fp2 A = quad_get2<0>(s2), B = quad_get2<1>(s2), C = quad_get2<2>(s2), D = quad_get2<3>(s2);
  const Fp2 E = quad_get2<0>(s2b), F = quad_get2<1>(s2b);
  return {norm(sub(B, A)), norm(add(D, C)), norm(add(F, E))};
}
// whether a point on the twist lies in G2 (every lane of the quad, same answer)
__device__ __forceinline__ bool g2_subgroup_quad(const Aff<Fp2>& qa, int q) {
  const Prj2 p = {qa.x, qa.y, fp2_one()};
  Prj2 acc = p;
#pragma unroll 1
  for (int i = 62; i >= 0; i--) {
    acc = quad_dbl2(acc, q);
    if ((BLS_X >> i) & 1) {
      acc.y = norm(acc.y);
      acc = quad_add2(acc, p, q);
    }
  }
  const Aff<Fp2> ps = g2_psi(qa);
  return !is_zero(acc.z) && eq(acc.x, mul(ps.x, acc.z)) && eq(neg(acc.y), mul(ps.y, acc.z));
}

}  // namespace bls
