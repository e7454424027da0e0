/*
 * CPU oracle for batch BLS12-381 signature verification.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * CPU verification path (ic_bls12_381 0.10.1 as driven by
 * oracle/src/main.rs:403-420 and gpu-crypto/src/lib.rs:142-146) and of the
 * reference's GPU formulas (gpu-crypto/cuda/api.cu:22-91,
 * gpu-crypto/cuda/field_sppark.cuh:29-197, cuda/test_miller_v2.cu:152-271,
 * ic-integration/gpu_crypto/cuda/decompress.cu:16-188).  Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark may load it, and
 * only as the checker; the GPU library never links it.
 *
 * Representation: Fp is 6 x u64 little-endian limbs in Montgomery form with
 * R = 2^384, always fully reduced to [0, p).  This is ic_bls12_381's internal
 * form and the byte layout of the reference C ABI.  Fp12 is 72 u64 in the
 * order c0.c0.c0, c0.c0.c1, c0.c1.c0, ..., c1.c2.c1.
 *
 * Parity pins (see tests/test_oracle_*.py): Fp against Python big-integer
 * closed forms; tower against schoolbook polynomial arithmetic and the
 * reference's toy KATs; Frobenius constants against their closed forms;
 * G2Prepared against cuda/g2_coeffs.h and cuda/pk42_coeffs.h; Miller loop
 * against cuda/g2_coeffs.h:551-565 and cuda/test_bls_verify.cu:147; final
 * exponentiation against lib/test_cyc_sq.cu:141-153 plus bilinearity and
 * GT^r = 1; decompression by round trips, the Euler criterion and r*P = O by
 * brute force; an identity signature in the fused path by e(O, G2) = 1 (its GT
 * equals e(-H(m), pk) and the conjugate of e(sig, G2) for the valid signature).
 * Parity unpinned: G1 from_compressed error classes beyond accept/reject,
 * hash_to_curve against ic_bls12_381 itself (pinned by RFC 9380's vectors), and
 * the reference's ChaCha20 key stream (the oracle uses its own seeded
 * generator, see orc_rand_*).
 */
#ifndef BLS_ORACLE_H
#define BLS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- field level (all operands Montgomery u64[6]) ---- */
void orc_fp_mul(const uint64_t a[6], const uint64_t b[6], uint64_t out[6]);
void orc_fp_add(const uint64_t a[6], const uint64_t b[6], uint64_t out[6]);
void orc_fp_sub(const uint64_t a[6], const uint64_t b[6], uint64_t out[6]);
void orc_fp_neg(const uint64_t a[6], uint64_t out[6]);
void orc_fp_inv(const uint64_t a[6], uint64_t out[6]);
int  orc_fp_sqrt(const uint64_t a[6], uint64_t out[6]); /* 1 if a square */
void orc_fp_to_mont(const uint64_t canon[6], uint64_t out[6]);
void orc_fp_from_mont(const uint64_t mont[6], uint64_t out[6]);

void orc_fp2_mul(const uint64_t a[12], const uint64_t b[12], uint64_t out[12]);
void orc_fp2_sqr(const uint64_t a[12], uint64_t out[12]);
void orc_fp2_inv(const uint64_t a[12], uint64_t out[12]);
void orc_fp6_mul(const uint64_t a[36], const uint64_t b[36], uint64_t out[36]);
void orc_fp12_mul(const uint64_t a[72], const uint64_t b[72], uint64_t out[72]);
void orc_fp12_sqr(const uint64_t a[72], uint64_t out[72]);
void orc_fp12_cyc_sqr(const uint64_t a[72], uint64_t out[72]);
void orc_fp12_inv(const uint64_t a[72], uint64_t out[72]);
void orc_fp12_frob(const uint64_t a[72], uint64_t out[72]);
void orc_fp12_conj(const uint64_t a[72], uint64_t out[72]);
void orc_fp12_mul_by_014(const uint64_t f[72], const uint64_t c0[12],
                         const uint64_t c1[12], const uint64_t c4[12], uint64_t out[72]);

/* ---- pairing ---- */
/* G2Prepared line coefficients of an affine G2 point (u64[24]: x.c0 x.c1 y.c0 y.c1),
 * 68 x (c0,c1,c2) in Fp2 = 2448 u64, Miller-loop order. */
void orc_g2_prepare(const uint64_t q[24], uint64_t coeffs[2448]);
/* two-pair Miller loop over prepared coefficients (api.cu:30-52). */
void orc_miller2(const uint64_t p1[12], const uint64_t q1c[2448],
                 const uint64_t p2[12], const uint64_t q2c[2448], uint64_t out[72]);
void orc_final_exp(const uint64_t f[72], uint64_t out[72]);
void orc_miller1(const uint64_t p1[12], const uint64_t q1c[2448], uint64_t out[72]);
void orc_pairing(const uint64_t p[12], const uint64_t q[24], uint64_t out[72]);
/* e(sig,G2)*e(neg_hm,pk) == 1 ; gt_out (nullable) receives the GT element. */
int  orc_verify(const uint64_t sig[12], const uint64_t neg_hm[12], const uint64_t g2c[2448],
                const uint64_t pkc[2448], uint64_t gt_out[72]);
/* legacy-ABI batch: pk coefficients AoS M x 2448; threads <= 0 means all cores. */
int  orc_verify_batch(const uint64_t* sigs, const uint64_t* neg_hms, const uint64_t* g2c,
                      const uint64_t* pkc, int32_t* results, uint64_t* gts, int m, int threads);
/* batch keyed by index: pk coefficients come from a table of nkeys keys. */
int  orc_verify_batch_keyed(const uint64_t* sigs, const uint64_t* neg_hms, const uint64_t* g2c,
                            const uint64_t* key_coeffs, const int32_t* key_idx,
                            int32_t* results, uint64_t* gts, int m, int threads);

/* fused CPU path: decompress each 48-byte signature with all checks, then verify;
 * results: 1 valid, 0 invalid, negative ORC_DEC_* code for a rejected encoding;
 * gts (nullable, m x 72): GT of each item, zeros for a rejected encoding.  An identity
 * signature is paired as the identity: GT = e(neg_hm, pk), verdict GT == 1. */
int  orc_verify_compressed_batch(const uint8_t* sig48, const uint64_t* neg_hms, const uint64_t* g2c,
                                 const uint64_t* key_coeffs, const int32_t* key_idx,
                                 int32_t* results, uint64_t* gts, int m, int threads);

/* ---- groups (affine Montgomery; identity encoded by *inf / return code) ---- */
#define ORC_DEC_OK 0
#define ORC_DEC_IDENTITY 1
#define ORC_DEC_BAD_FLAGS (-1)
#define ORC_DEC_X_NOT_CANONICAL (-2)
#define ORC_DEC_NOT_ON_CURVE (-3)
#define ORC_DEC_NOT_IN_SUBGROUP (-4)
int  orc_g1_decompress(const uint8_t in[48], uint64_t out[12]);
/* decompress without the subgroup check (on-curve only) */
int  orc_g1_decompress_unchecked(const uint8_t in[48], uint64_t out[12]);
void orc_g1_compress(const uint64_t p[12], int inf, uint8_t out[48]);
int  orc_g1_is_on_curve(const uint64_t p[12]);
int  orc_g1_in_subgroup(const uint64_t p[12]);
/* scalar (32 bytes little-endian, any 256-bit value) times affine point;
 * returns 1 if the result is the identity (then out is zeroed). */
int  orc_g1_mul(const uint64_t p[12], const uint8_t k[32], uint64_t out[12]);
int  orc_g2_mul(const uint64_t q[24], const uint8_t k[32], uint64_t out[24]);
int  orc_g2_in_subgroup(const uint64_t q[24]);
/* G2 compression as ic_bls12_381 (96 bytes: x.c1 then x.c0 big-endian, flags in byte
 * 0); decompress returns an ORC_DEC_* code, the checked form also [r]Q == O */
void orc_g2_compress(const uint64_t q[24], int inf, uint8_t out[96]);
int  orc_g2_decompress(const uint8_t in[96], uint64_t out[24]);
int  orc_g2_decompress_unchecked(const uint8_t in[96], uint64_t out[24]);
int  orc_g2_is_on_curve(const uint64_t q[24]);
void orc_g1_neg(const uint64_t p[12], uint64_t out[12]);
void orc_g1_generator(uint64_t out[12]);
void orc_g2_generator(uint64_t out[24]);

/* ---- hashing and workload ---- */
void orc_sha256(const uint8_t* msg, size_t len, uint8_t out[32]);
/* H(m) = s*G1, s = LE(SHA-256("ic-bls-test-hash" || msg)) with byte 31 &= 0x0f
 * (oracle/src/main.rs:383-397). */
void orc_hash_to_g1(const uint8_t* msg, size_t len, uint64_t out[12]);
/* deterministic secret key in [1, r) from (seed, index) via splitmix64 + rejection */
void orc_rand_scalar(uint64_t seed, uint64_t index, uint8_t out[32]);
void orc_rand_bytes(uint64_t seed, uint64_t index, uint8_t* out, size_t len);
/* x in [0,p) canonical bytes (big-endian, 48) such that x^3+4 is a non-square
 * (want_square=0) or a square (want_square=1), searched from (seed,index). */
void orc_rand_x(uint64_t seed, uint64_t index, int want_square, uint64_t x_mont[6]);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
