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
static __device__ __noinline__ Fp2 mul2_call(Fp2 a, Fp2 b) { return mul(a, b); }
static __device__ __noinline__ Fp2 sqr2_call(Fp2 a) { return sqr(a); }
struct Fp2Calls {
  static __device__ __forceinline__ Fp2 mul(const Fp2& a, const Fp2& b) { return mul2_call(a, b); }
  static __device__ __forceinline__ Fp2 sqr(const Fp2& a) { return sqr2_call(a); }
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
template <class F, class M = FieldInline>
__device__ __forceinline__ bool jeq_aff(const Jac<F>& p, const Aff<F>& q) {
  if (is_zero(p.z)) return false;
  F z2 = M::sqr(p.z);
  return eq(p.x, M::mul(q.x, z2)) && eq(p.y, M::mul(q.y, M::mul(z2, p.z)));
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
// "complex method"), with two square roots in Fp and one inversion where two
// exponentiations in Fp2 took 2.6 times the products.  For a = a0 + a1 u:
//   lambda = sqrt(a0^2 + a1^2) (a is a square in Fp2 exactly when its norm is one in Fp);
//   delta = (a0 + lambda) / 2 and t = delta^((p+1)/4).  t^2 = delta gives the root
//   (t, a1 / (2t)); t^2 = -delta gives (a1 / (2t), t): the product of the two candidate
//   deltas (a0 +- lambda) / 2 is -(a1/2)^2, so exactly one is a square (-1 is not one) and
//   a1 / (2t) is the root of the other.  a1 = 0 takes delta = a0 and a zero partner
//   coordinate: (t, 0) when a0 is a square, else (0, t).
// ok = the result squares back to a.  Which of the two roots comes out does not matter:
// the caller picks the sign by the sort flag.
__device__ __forceinline__ Fp2 sqrt2(const Fp2& a, bool& ok) {
  const bool a1z = is_zero(a.c1);
  const Fp n = norm(add(sqr(a.c0), sqr(a.c1)));
  const Fp lambda = sqrt_cand(n);
  const Fp delta = a1z ? a.c0 : mul(norm(add(a.c0, lambda)), fp_const(HALF_M));
  const Fp t = sqrt_cand(delta);
  const bool sq = eq(sqr(t), delta);
  Fp w = fp_zero();
  if (!a1z) w = mul(a.c1, inv_fp(dbl(t)));  // t != 0: delta = 0 only when a1 = 0
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
// G2 membership of a point on the twist: psi(Q) == [x]Q = -[|x|]Q (eprint 2021/1130
// sec. 4, eprint 2022/352; ic_bls12_381's G2Affine::is_torsion_free)
__device__ __forceinline__ bool g2_in_subgroup(const Aff<Fp2>& q) {
  const Fp2 one = {fp_one(), fp_zero()};
  Jac<Fp2> acc = {q.x, q.y, one};
  for (int i = 62; i >= 0; i--) {
    acc = jdbl<Fp2, Fp2Calls>(acc);
    if ((BLS_X >> i) & 1) acc = jadd_aff<Fp2, Fp2Calls>(acc, q, one);
  }
  acc.y = neg(acc.y);
  return jeq_aff<Fp2, Fp2Calls>(acc, g2_psi(q));
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
// 48-byte compressed encoding of an affine point (big-endian words as above)
// 96 bytes given as 24 big-endian-packed words: x.c1 (words 0..11, flags in the top
// three bits) then x.c0.  Semantics of ic_bls12_381 G2Affine::from_compressed (the
// reference has no G2 decompression; its callers pass prepared lines):
// check_subgroup=false gives from_compressed_unchecked.
__device__ __forceinline__ int32_t g2_decompress(const u32 be[24], Aff<Fp2>& out, bool check_subgroup) {
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
  if (check_subgroup && !g2_in_subgroup(out)) return DEC_NOT_IN_SUBGROUP;
  return DEC_OK;
}

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
// reproduces cuda/g2_coeffs.h and cuda/pk42_coeffs.h. ----
struct G2Proj {
  Fp2 x, y, z;
};
__device__ __forceinline__ Line doubling_step(G2Proj& r) {
  Fp2 tmp0 = sqr(r.x);
  Fp2 tmp1 = sqr(r.y);
  Fp2 tmp2 = sqr(tmp1);
  Fp2 tmp3 = norm(dbl(sub(sqr(carry(add(tmp1, r.x))), add(tmp0, tmp2))));
  Fp2 tmp4 = norm(add(dbl(tmp0), tmp0));
  Fp2 tmp6 = carry(add(r.x, tmp4));
  Fp2 tmp5 = sqr(tmp4);
  Fp2 zsq = sqr(r.z);
  Fp2 nx = norm(sub(sub(tmp5, tmp3), tmp3));
  Fp2 nz = norm(sub(sqr(carry(add(r.z, r.y))), add(tmp1, zsq)));
  Fp2 ny = norm(sub(mul(sub(tmp3, nx), tmp4), norm(dbl(dbl(dbl(tmp2))))));
  r.x = nx; r.y = ny; r.z = nz;
  Fp2 c1 = norm(neg(dbl(mul(tmp4, zsq))));
  Fp2 c2 = norm(sub(sub(sqr(tmp6), add(tmp0, tmp5)), dbl(dbl(tmp1))));
  Fp2 c0 = norm(dbl(mul(r.z, zsq)));
  return {c0, c1, c2};
}
__device__ __forceinline__ Line addition_step(G2Proj& r, const Aff<Fp2>& q) {
  Fp2 zsq = sqr(r.z);
  Fp2 ysq = sqr(q.y);
  Fp2 t0 = mul(zsq, q.x);
  Fp2 t1 = mul(norm(sub(sqr(carry(add(q.y, r.z))), add(ysq, zsq))), zsq);
  Fp2 t2 = norm(sub(t0, r.x));
  Fp2 t3 = sqr(t2);
  Fp2 t4 = norm(dbl(dbl(t3)));
  Fp2 t5 = mul(t4, t2);
  Fp2 t6 = norm(sub(sub(t1, r.y), r.y));
  Fp2 t9 = mul(t6, q.x);
  Fp2 t7 = mul(t4, r.x);
  Fp2 nx = norm(sub(sub(sub(sqr(t6), t5), t7), t7));
  Fp2 nz = norm(sub(sqr(carry(add(r.z, t2))), add(zsq, t3)));
  Fp2 t10 = carry(add(q.y, nz));
  Fp2 t8 = mul(sub(t7, nx), t6);
  Fp2 ny = norm(sub(t8, dbl(mul(r.y, t5))));
  r.x = nx; r.y = ny; r.z = nz;
  t10 = norm(sub(sub(sqr(t10), ysq), sqr(nz)));
  Fp2 c2 = norm(sub(dbl(t9), t10));
  Fp2 c0 = norm(dbl(nz));
  Fp2 c1 = norm(dbl(neg(t6)));
  return {c0, c1, c2};
}
// Emit the 68 lines of q in Miller-loop order through sink(k, line).
template <class Sink>
__device__ __forceinline__ void g2_prepare(const Aff<Fp2>& q, Sink sink) {
  G2Proj r = {q.x, q.y, fp2_one()};
  int n = 0;
  for (int b = 61; b >= 0; b--) {
    sink(n++, doubling_step(r));
    if (miller_bit(b)) sink(n++, addition_step(r, q));
  }
  sink(n++, doubling_step(r));
}

}  // namespace bls
