// Hash to G1 on the device: RFC 9380, suite BLS12381G1_XMD:SHA-256_SSWU_RO_ (what
// ic_bls12_381's HashToCurve<ExpandMsgXmd<Sha256>> computes; the reference hashes on
// the CPU before its GPU step, ic-integration/patches/ic-gpu-consensus.patch:454-457,
// with the DST of ic-bench/src/main.rs:68).  Four lanes per message (kernels_h2c.hip):
//   expand_message_xmd (SHA-256) -> two field elements (64 bytes each, mod p)
//   -> simplified SWU onto E': y^2 = x^3 + A'x + B' (x kept as a fraction, y affine
//      from one exponentiation) -> the 11-isogeny to E (rational maps evaluated
//      homogeneously, result in projective coordinates)
//   -> Q0 + Q1 -> [1 - x](Q0 + Q1) (cofactor clearing, complete formulas) -> affine.
#pragma once
#include "curve.hpp"

namespace bls {

// ---- SHA-256 (FIPS 180-4) ----
constexpr u32 SHA256_K[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
constexpr u32 SHA256_IV[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                              0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};

// one SHA-256 compression with a rolling 16-word schedule; host and device (the
// synthetic workload generator hashes on the host with the same function)
__host__ __device__ __forceinline__ u32 rotr32(u32 x, int n) { return (x >> n) | (x << (32 - n)); }

__host__ __device__ __forceinline__ void sha256_compress(u32 (&h)[8], u32 (&w)[16]) {
  u32 a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], k = h[7];
#pragma unroll
  for (int i = 0; i < 64; i++) {
    if (i >= 16) {
      const u32 w15 = w[(i - 15) & 15], w2 = w[(i - 2) & 15];
      w[i & 15] += (rotr32(w15, 7) ^ rotr32(w15, 18) ^ (w15 >> 3)) + w[(i - 7) & 15] +
                   (rotr32(w2, 17) ^ rotr32(w2, 19) ^ (w2 >> 10));
    }
    const u32 t1 = k + (rotr32(e, 6) ^ rotr32(e, 11) ^ rotr32(e, 25)) + ((e & f) ^ (~e & g)) + SHA256_K[i] + w[i & 15];
    const u32 t2 = (rotr32(a, 2) ^ rotr32(a, 13) ^ rotr32(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
    k = g;
    g = f;
    f = e;
    e = d + t1;
    d = c;
    c = b;
    b = a;
    a = t1 + t2;
  }
  h[0] += a;
  h[1] += b;
  h[2] += c;
  h[3] += d;
  h[4] += e;
  h[5] += f;
  h[6] += g;
  h[7] += k;
}

// SHA-256 of an n-byte message whose first PRE words (block 0, big-endian) are pre[]
// and whose other bytes come from gen(pos); the padding is appended here.
template <int PRE, class Gen>
__device__ __forceinline__ void sha256_gen(int n, const u32 (&pre)[8], Gen gen, u32 (&out)[8]) {
#pragma unroll
  for (int j = 0; j < 8; j++) out[j] = SHA256_IV[j];
  const int total = ((n + 9 + 63) >> 6) << 6;
  const u64 bits = (u64)n * 8;
#pragma unroll 1
  for (int blk = 0; blk < total; blk += 64) {
    u32 w[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
      if (j < PRE && blk == 0) {
        w[j] = pre[j < 8 ? j : 0];
        continue;
      }
      u32 v = 0;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int pos = blk + 4 * j + q;
        u32 by;
        if (pos < n) by = gen(pos);
        else if (pos == n) by = 0x80;
        else if (pos >= total - 8) by = (u32)(bits >> (8 * (total - 1 - pos))) & 0xff;
        else by = 0;
        v = (v << 8) | by;
      }
      w[j] = v;
    }
    sha256_compress(out, w);
  }
}

// ---- expand_message_xmd (RFC 9380 sec. 5.3.1) for 128 bytes: b_1 .. b_4 ----
// b_0 = H(Z_pad || msg || I2OSP(128, 2) || 0 || DST'), b_1 = H(b_0 || 1 || DST'),
// b_i = H((b_0 xor b_(i-1)) || i || DST') with DST' = DST || I2OSP(len(DST), 1).
__device__ __forceinline__ void expand_xmd128(const uint8_t* msg, int len, const uint8_t* dst, int dlen,
                                              u32 (&ub)[4][8]) {
  const u32 none[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  u32 b0[8];
  const int n0 = 64 + len + 3 + dlen + 1;
  sha256_gen<0>(n0, none, [&](int pos) -> u32 {
    if (pos < 64) return 0;
    int k = pos - 64;
    if (k < len) return msg[k];
    k -= len;
    if (k < 3) return k == 1 ? 0x80u : 0u;  // I2OSP(128, 2) || I2OSP(0, 1)
    k -= 3;
    return k < dlen ? (u32)dst[k] : (u32)dlen;
  }, b0);
  const int ni = 32 + 1 + dlen + 1;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    u32 pre[8];
#pragma unroll
    for (int j = 0; j < 8; j++) pre[j] = i == 0 ? b0[j] : b0[j] ^ ub[i > 0 ? i - 1 : 0][j];
    sha256_gen<8>(ni, pre, [&](int pos) -> u32 {
      const int k = pos - 33;
      if (k < 0) return (u32)(i + 1);
      return k < dlen ? (u32)dst[k] : (u32)dlen;
    }, ub[i]);
  }
}

// 64 big-endian bytes (16 words, w[0] most significant) mod p, internal form:
// lo (the last 48 bytes) R' + hi (the first 16) 2^384 R'
__device__ __forceinline__ Fp field_from_512(const u32 (&hiw)[8], const u32 (&low)[8]) {
  u32 lo[12], hi[12];
#pragma unroll
  for (int j = 0; j < 12; j++) {
    const int k = 15 - j;  // big-endian word index of little-endian word j
    lo[j] = k < 8 ? hiw[k] : low[k - 8];
    hi[j] = j < 4 ? hiw[3 - j] : 0;
  }
  return norm(add(mul(unpack384(lo), fp_const(C_R2_INT)), mul(unpack384(hi), fp_const(C_H2F_HI))));
}

// parity of the canonical value (sgn0 of RFC 9380 sec. 4.1 for Fp)
__device__ __forceinline__ u32 sgn0(const Fp& a) { return (u32)to_plain(a).l[0] & 1; }

// ---- simplified SWU (RFC 9380 sec. 6.6.2) onto E', then the 11-isogeny to E ----
// x1 = xn / xd with xn = -B'(t + 1), xd = A' t for t = Z^2 u^4 + Z u^2 (xn = B',
// xd = Z A' when t = 0); g(x1) = U / V with U = xn^3 + A' xn xd^2 + B' xd^3, V = xd^3.
// y1 = (U V^3)^((p-3)/4) U V squares to U / V exactly when U / V is a square, and to
// -U / V otherwise; then x2 = Z u^2 x1 and y2 = y1 Z u^3 sqrt(-Z) (g(x2) = Z^3 u^6 g(x1)).
// The isogeny's maps, evaluated at x = xn / xd as forms of degree 15 (the x map's
// numerator and denominator, of degrees 11 and 10, padded with leading zero
// coefficients: XN xd^4 and XD xd^5), give x = XN xd^4 / (XD xd^5) = XN / (xd XD) and
// y = y YN / YD, so the projective result is X = XN' YD, Y = y YN XD', Z = XD' YD (XN',
// XD' the padded forms); Z = 0 (a point of the isogeny's kernel) gives the identity
// (0 : 1 : 0).  On a lane quad holding one message: lanes 0 | 1 map u0 | u1 and evaluate
// the x map's two forms, lanes 2 | 3 the same field elements and the y map's, and the
// pairs (0, 2), (1, 3) exchange their forms and split the final products.
// (out of line, by value; field products inline, measured 12 % faster than through
// FpCalls here)
__device__ __forceinline__ Fp quad_swap2(const Fp& a) {  // lane i reads lane i ^ 2
  Fp r;
#pragma unroll
  for (int j = 0; j < NL; j++) r.l[j] = __builtin_amdgcn_update_dpp(a.l[j], a.l[j], 0x4E, 0xF, 0xF, false);
  return r;
}
template <class M>
__device__ __noinline__ Prj map_to_g1_curve(Fp u, int q) {
  const Fp A = fp_const(SSWU_A_M), B = fp_const(SSWU_B_M), Zc = fp_const(SSWU_Z_M);
  const Fp zu2 = M::mul(Zc, M::sqr(u));
  const Fp t = norm(add(M::sqr(zu2), zu2));
  Fp xn, xd;
  if (is_zero(t)) {
    xn = B;
    xd = M::mul(Zc, A);
  } else {
    xn = M::mul(neg(B), norm(add(t, fp_one())));
    xd = M::mul(A, t);
  }
  const Fp xd2 = M::sqr(xd), xd3 = M::mul(xd2, xd);
  const Fp U = norm(add(add(M::mul(M::sqr(xn), xn), M::mul(M::mul(A, xn), xd2)), M::mul(B, xd3)));
  const Fp UV = M::mul(U, xd3);
  const Fp y1 = M::mul(pow_win(M::mul(UV, M::sqr(xd3)), P34_WIN), UV);
  const bool square = eq(M::mul(M::sqr(y1), xd3), U);
  Fp x = xn, y = y1;
  if (!square) {
    x = M::mul(zu2, xn);
    y = M::mul(M::mul(y1, M::mul(zu2, u)), fp_const(SSWU_SQRT_MZ_M));
  }
  if (sgn0(y) != sgn0(u)) y = neg(y);
  // the isogeny: numerator N and denominator D of the lane's map, one pass over the
  // powers xd^j, j = 1..15 (coefficient 15 - j at step j)
  const bool ymap = q >= 2;
  Fp pw = fp_one();
  Fp N = pick(ymap, fp_const(ISO_YNUM[15]), fp_zero()), D = pick(ymap, fp_one(), fp_zero());
#pragma unroll 1
  for (int j = 1; j <= 15; j++) {
    const int k = 15 - j;
    Fp cxn = fp_zero(), cxd = fp_zero();
    if (k <= 11) cxn = fp_const(ISO_XNUM[k]);
    if (k <= 9) cxd = fp_const(ISO_XDEN[k]);
    if (k == 10) cxd = fp_one();
    pw = M::mul(pw, xd);
    N = norm(add(M::mul(N, x), M::mul(pick(ymap, fp_const(ISO_YNUM[k]), cxn), pw)));
    D = norm(add(M::mul(D, x), M::mul(pick(ymap, fp_const(ISO_YDEN[k]), cxd), pw)));
  }
  const Fp No = quad_swap2(N), Do = quad_swap2(D);
  const Fp XN = pick(ymap, No, N), XD = pick(ymap, Do, D), YN = pick(ymap, N, No), YD = pick(ymap, D, Do);
  // lanes 0, 1: XN' YD, then XD' YD; lanes 2, 3: y YN, then y YN XD'
  const Fp s1 = M::mul(pick(ymap, y, XN), pick(ymap, YN, YD));
  const Fp s2 = M::mul(pick(ymap, s1, XD), pick(ymap, XD, YD));
  const Fp s1o = quad_swap2(s1), s2o = quad_swap2(s2);
  Prj r = {pick(ymap, s1o, s1), pick(ymap, s2, s2o), pick(ymap, s2o, s2)};
  if (is_zero(r.z)) r = {fp_zero(), fp_one(), fp_zero()};
  return r;
}

// the two field elements u0, u1 of hash_to_field (RFC 9380 sec. 5.2, L = 64)
__device__ __forceinline__ void hash_to_field2(const uint8_t* msg, int len, const uint8_t* dst, int dlen, Fp& u0,
                                               Fp& u1) {
  u32 ub[4][8];
  expand_xmd128(msg, len, dst, dlen, ub);
  u0 = field_from_512(ub[0], ub[1]);
  u1 = field_from_512(ub[2], ub[3]);
}
// [1 - x](Q0 + Q1) (cofactor clearing, sec. 7 with h_eff of sec. 8.8.1) as affine
// (x, y) in internal form, on every lane of the quad; false for the identity
__device__ __forceinline__ bool clear_cofactor_affine(const Prj& q0, const Prj& q1, int q, Fp& x, Fp& y) {
  const Prj r = quad_add(q0, q1, q);
  const Prj h = quad_add(quad_mul_absx(r, q), r, q);  // [|x| + 1] R = [1 - x] R
  if (is_zero(h.z)) return false;
  const Fp zi = inv_fp(h.z);
  x = mul(h.x, zi);
  y = mul(h.y, zi);
  return true;
}

}  // namespace bls
