/*
 * libblsgpu_workload.so: the synthetic workload generator and the multiply-add issue
 * probe used by bench.py and the tests.  A library of its own, built beside
 * libgpu_msm.so and not linked into it: the verifier a replica loads holds no key
 * generator or signer.  No counterpart in the reference.
 * The seeded streams match the test oracle's generator, so the same (seed, index)
 * gives the same keys, messages and encodings.
 */
#ifndef BLSGPU_WORKLOAD_H
#define BLSGPU_WORKLOAD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BLSGPU_WL_VALID 0
#define BLSGPU_WL_WRONG_MSG 1     /* valid point, signature over another message */
#define BLSGPU_WL_NO_SQRT 2       /* x^3 + 4 is not a square: decompression rejects */
#define BLSGPU_WL_NON_SUBGROUP 3  /* on the curve, outside the r-order subgroup */

/* n secret keys in [1, r), 32 bytes little-endian each (host only) */
void blsgpu_wl_secret_keys(uint64_t seed, int n, uint8_t* sk32);
/* the 32-byte message of item `index` (host only) */
void blsgpu_wl_message(uint64_t seed, int64_t index, uint8_t msg32[32]);
/* pk = sk * G2 on the device, reference form (n x 24 u64) */
int blsgpu_wl_public_keys(const uint8_t* sk32, int n, uint64_t* pk_ref);
/* n items: compressed signature (48 B) and -H(m) (12 u64, reference form) per item;
 * item i signs message i with key key_idx[i] unless its category says otherwise */
int blsgpu_wl_items(uint64_t seed, int n, const uint8_t* sk32, int nkeys, const int32_t* key_idx,
                    const int8_t* category, uint8_t* sig48, uint64_t* nhm_ref);
/* as blsgpu_wl_items, but item i signs message msg_idx[i] (messages shared by several
 * signers); msg_idx NULL = item index */
int blsgpu_wl_items_msg(uint64_t seed, int n, const uint8_t* sk32, int nkeys, const int32_t* key_idx,
                        const int64_t* msg_idx, const int8_t* category, uint8_t* sig48, uint64_t* nhm_ref);
/* compressed signatures sk32[key_idx[i]] * H_i of given G1 points (n x 12 u64,
 * reference form), e.g. hashes from blsgpu_hash_to_g1 (device) */
int blsgpu_wl_sign_points(const uint64_t* h_ref, int n, const uint8_t* sk32, int nkeys, const int32_t* key_idx,
                          uint8_t* sig48);
/* measured v_mad_u64_u32 issue rate over all CUs (products per second) */
int blsgpu_probe_mad_rate(double* mads_per_s, double* ms);
const char* blsgpu_wl_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
