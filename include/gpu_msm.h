/*
 * libgpu_msm.so -- MI355X (gfx950) batch BLS12-381 signature verification.
 *
 * Drop-in for the reference's GPU crypto ABI.  The first block of entry points
 * keeps the exact names, argument meaning, data layout and return convention of
 * the reference (all pointers are caller-owned host memory, calls are
 * synchronous, 0 = success, -1 = failure, on which callers fall back to the CPU).
 * Unlike the reference, every entry point is thread-safe.
 *
 * Data layout (ic_bls12_381 internal form, as marshalled by the reference's Rust
 * crate, ic-integration/gpu_crypto/lib.rs:45-69):
 *   Fp        6 x u64 little-endian, Montgomery form with R = 2^384, reduced to [0, p)
 *   G1 point  x | y = 12 x u64 (no infinity flag)
 *   G2 point  x.c0 | x.c1 | y.c0 | y.c1 = 24 x u64
 *   line      (c0, c1, c2) in Fp2^3 = 36 x u64; one key = 68 lines = 2448 x u64
 *   GT        12 Fp = 72 x u64 in the order c0.c0.c0, c0.c0.c1, c0.c1.c0, ..., c1.c2.c1
 */
#ifndef GPU_MSM_H
#define GPU_MSM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ===================== reference ABI ===================== */

/* Replaces gpu-crypto/cuda/api.cu:118-124.  0 when a GPU is present, -1 otherwise. */
int gpu_crypto_init(void);

/* Replaces gpu-crypto/cuda/api.cu:128-160 (Rust binding ic-integration/gpu_crypto/lib.rs:20-27).
 * Verifies e(sigs[i], G2) * e(neg_hms[i], pk_i) == 1 for i < M, with pk_i given by its
 * 68 precomputed lines pk_coeffs + i*2448 and the generator's lines in g2_coeffs.
 * results[i] = 1 (valid) or 0 (invalid).  Returns 0, or -1 on any device error or
 * M <= 0 (the reference's M = 0 launch also fails).  The key tables (19,584 B per item)
 * are copied through pinned staging by the calling thread and three library threads;
 * the arrays are only read, and only during the call. */
int gpu_batch_bls_verify(const uint64_t* sigs, const uint64_t* neg_hms, const uint64_t* g2_coeffs,
                         const uint64_t* pk_coeffs, int32_t* results, int M);

/* Replaces ic-integration/gpu_crypto/cuda/decompress.cu:198-226.  n x 48-byte ZCash-
 * compressed G1 points -> n x 96 bytes (x | y, ic_bls12_381 internal form).  Returns -1
 * when n <= 0 or when ANY point is not a valid non-identity G1 element (bad flags,
 * x >= p, x^3+4 not a square, not in the r-order subgroup, or the identity encoding):
 * the caller then decodes the batch on the CPU and gets ic_bls12_381's exact error. */
int gpu_g1_decompress_batch(const unsigned char* compressed, unsigned char* out_points, int n);

/* Replaces decompress.cu:229-231: gpu_g1_decompress_batch with n = 1. */
int gpu_g1_decompress(const unsigned char* compressed, unsigned char* out_point);

/* Replace ic-integration/gpu_crypto/cuda/msm.cu:32-56.  G1 MSM is not part of this
 * library; both symbols exist because the caller's dlopen requires all three of
 * gpu_msm_g1, gpu_jacobian_to_affine and gpu_g1_decompress_batch
 * (ic-integration/gpu_crypto/lib.rs:104-116).  Always -1 (caller uses the CPU). */
int gpu_msm_g1(const unsigned char* points, const unsigned char* scalars, size_t n, unsigned char* out);
int gpu_jacobian_to_affine(const unsigned char* jac, unsigned char* aff);

/* Replaces gpu-crypto/cuda/api.cu:165-186 (linked by lib.rs:14-19).  Out of scope:
 * always -1, the precedent being the reference's own stub build (BUILD.bazel:38). */
int gpu_batch_sha256(const uint8_t* chunks, uint8_t* hashes, int n, int chunk_size);

/* ===================== extensions (no reference counterpart) ===================== */

/* Per-item codes of the extended verify entry points. */
#define BLSGPU_VALID 1
#define BLSGPU_INVALID 0
#define BLSGPU_ERR_BAD_FLAGS (-1)       /* signature encoding flags */
#define BLSGPU_ERR_X_NOT_CANONICAL (-2) /* signature x >= p */
#define BLSGPU_ERR_NOT_ON_CURVE (-3)    /* x^3 + 4 has no square root */
#define BLSGPU_ERR_NOT_IN_SUBGROUP (-4) /* point outside the r-order subgroup */

/* gpu_batch_bls_verify plus the GT element of every item (M x 72 u64, nullable). */
int blsgpu_batch_verify_gt(const uint64_t* sigs, const uint64_t* neg_hms, const uint64_t* g2_coeffs,
                           const uint64_t* pk_coeffs, int32_t* results, uint64_t* gt_out, int M);

/* Key registry: per-key line tables stay resident in HBM.  Returns a handle >= 0 or -1.
 * from_affine computes G2Prepared on the GPU from an affine G2 point (24 u64);
 * from_coeffs takes the 68 precomputed lines (2448 u64). */
int blsgpu_register_key_affine(const uint64_t pk[24]);
int blsgpu_register_key_coeffs(const uint64_t coeffs[2448]);
/* n keys from affine points (n x 24 u64) in one launch; handles[i] (nullable) receives
 * the handle of key i.  Returns the first handle, or -1. */
int blsgpu_register_keys_affine(const uint64_t* pks, int n, int32_t* handles);
/* n keys from compressed G2 points (n x 96 bytes: x.c1 then x.c0, big-endian, flags in
 * the top three bits of byte 0), decompressed, validated and prepared on the GPU with
 * ic_bls12_381's G2Affine::from_compressed semantics (check_subgroup = 0 gives
 * from_compressed_unchecked).  status[i] (nullable): 0 accepted, 1 the identity, or a
 * negative BLSGPU_ERR_* code; handles[i] (nullable): the key's handle, -1 unless status
 * is 0 (the identity is not registered).  Accepted keys get consecutive handles in input
 * order.  Returns the number of keys registered, or -1. */
int blsgpu_register_keys_compressed(const uint8_t* pks96, int n, int32_t* handles, int32_t* status,
                                    int check_subgroup);
/* copies a registered key's lines back out (2448 u64, reference form) */
int blsgpu_key_coeffs(int handle, uint64_t coeffs[2448]);
int blsgpu_num_keys(void);
void blsgpu_reset_keys(void);

/* Fused path: M compressed signatures (M x 48 bytes) are decompressed and validated on
 * the device and verified against e(.,G2) * e(neg_hms[i], key[key_idx[i]]).
 * results[i]: BLSGPU_VALID, BLSGPU_INVALID or a negative BLSGPU_ERR_* decode code.
 * An identity signature is decoded and paired as the identity (ic_bls12_381 semantics).
 * gt_out nullable (zeros for items with a decode error). */
int blsgpu_verify_compressed(const uint8_t* sigs48, const uint64_t* neg_hms, const int32_t* key_idx,
                             int32_t* results, uint64_t* gt_out, int M);

/* Hash to G1 on the GPU: RFC 9380 suite BLS12381G1_XMD:SHA-256_SSWU_RO_ with the
 * caller's DST (1..255 bytes; the reference's benchmark uses
 * "BLS_SIG_BLS12381G2_XMD:SHA-256_SSWU_RO_NUL_", ic-bench/src/main.rs:68), the
 * computation ic_bls12_381's hash_to_curve performs on the CPU before the reference's
 * GPU step (patch:454-457).  Message i is msgs[offsets[i] .. offsets[i+1]) (offsets: n+1
 * entries).  out: n x 12 u64, H(m) affine in the G1 boundary form, or -H(m) when
 * negate != 0; is_identity (nullable) flags the (never observed) identity, output zeros. */
int blsgpu_hash_to_g1(const uint8_t* msgs, const int64_t* offsets, int n, const uint8_t* dst, int dst_len,
                      uint64_t* out, int32_t* is_identity, int negate);

/* Fused wire-format path: compressed signatures, messages and key handles in; H(m) is
 * computed on the device (as blsgpu_hash_to_g1, beside the signature decoding) and the
 * rest is blsgpu_verify_compressed (same results and gt_out). */
int blsgpu_verify_messages(const uint8_t* sigs48, const uint8_t* msgs, const int64_t* offsets, const uint8_t* dst,
                           int dst_len, const int32_t* key_idx, int32_t* results, uint64_t* gt_out, int M);

/* Decompress with per-point status (n x 96 bytes out; status[i] as BLSGPU_ERR_*, 0 ok,
 * 1 identity).  check_subgroup = 0 skips the subgroup test. */
int blsgpu_g1_decompress_status(const uint8_t* in48, uint8_t* out96, int32_t* status, int n,
                                int check_subgroup);

/* Device-resident form of blsgpu_verify_compressed: every array is device memory and
 * the work is enqueued on `stream` (a hipStream_t; NULL = the library's stream) without
 * waiting.  Items whose key index is out of range get BLSGPU_ERR_BAD_KEY.  Workspace
 * comes from the library's pool; calls are serialized by the library's lock. */
#define BLSGPU_ERR_BAD_KEY (-5)
int blsgpu_verify_compressed_device(const uint8_t* d_sigs48, const uint64_t* d_neg_hms,
                                    const int32_t* d_key_idx, int32_t* d_results, uint64_t* d_gt,
                                    int M, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * blsgpu_verify_compressed_device (same arguments, results and GT) with the library's
 * grow-only workspace replaced by a fixed device array of BLSGPU_NOALLOC_SCRATCH_BYTES
 * (7,405,568 bytes: 904 per item for BLSGPU_NOALLOC_MAX_ITEMS items) that each GPU the
 * library runs on holds until the process ends.  It calls no hipMalloc, hipFree or other
 * allocating routine; set-up is gpu_crypto_init's (called first, else -1).  M above
 * BLSGPU_NOALLOC_MAX_ITEMS returns -1 with a message.  Every call on a GPU shares that
 * GPU's array: the library orders its own calls (a call on another stream waits, on the
 * device, for the previous call's kernels), but work that overlaps on one GPU outside
 * that order races on the array, e.g. a captured graph of this call replayed beside
 * other calls of it. */
#define BLSGPU_NOALLOC_MAX_ITEMS 8192
#define BLSGPU_NOALLOC_SCRATCH_BYTES 7405568
int blsgpu_verify_compressed_device_noalloc(const uint8_t* d_sigs48, const uint64_t* d_neg_hms,
                                            const int32_t* d_key_idx, int32_t* d_results, uint64_t* d_gt,
                                            int M, void* stream);

/* Batched form: nbatches verify batches in one submission, with the same per-item
 * semantics as blsgpu_verify_compressed.  The Miller-loop and final-exponentiation
 * kernels are persistent (a device-sized grid pops tiles of items from a device work
 * queue), and the whole kernel sequence is captured once as a hipGraph per input shape
 * and replayed.  No counterpart in the reference (its entry point launches one kernel
 * per batch, gpu-crypto/cuda/api.cu:128-160).
 *   _device: inputs and outputs are device pointers; batch b is items
 *     [batch_off[b], batch_off[b+1]) of the concatenated arrays (batch_off: host array
 *     of nbatches + 1 non-decreasing entries, batch_off[0] = 0); the work is ordered on
 *     `stream` (NULL = the library's stream) and the call returns without waiting.
 *   host form: batch b has sizes[b] items at sigs48[b], neg_hms[b], key_idx[b] and its
 *     results go to results[b] (gt_out nullable, else gt_out[b] nullable per batch);
 *     synchronous. */
int blsgpu_verify_batches_device(const uint8_t* d_sigs48, const uint64_t* d_neg_hms, const int32_t* d_key_idx,
                                 const int32_t* batch_off, int nbatches, int32_t* d_results, uint64_t* d_gt,
                                 void* stream);
int blsgpu_verify_batches(const uint8_t* const* sigs48, const uint64_t* const* neg_hms,
                          const int32_t* const* key_idx, const int32_t* sizes, int nbatches,
                          int32_t* const* results, uint64_t* const* gt_out);
/* Batched-form configuration: lanes per item of the Miller loop (8, 12, 13, 16, 20, 24, 32) and of
 * the final exponentiation (9, 12, 16, 18, 24, 27, 32), workgroups per CU of each (0 = occupancy
 * limit, -1 = one workgroup per tile, no queue), and graph capture on/off.  Clears the
 * autotuned table. */
int blsgpu_set_batched_engine(int g_miller, int g_final_exp, int wpc_miller, int wpc_final_exp, int use_graph);
/* configuration used for a job of M items: {g_miller, g_final_exp, wpc_miller, wpc_final_exp, use_graph} */
int blsgpu_get_batched_engine(int M, int cfg[5]);
/* out[0] graphs instantiated, out[1] graph launches */
int blsgpu_batched_stats(int out[2]);
/* Key-line depth of the lanes-per-item Miller loop (coefficient-tile depth; no
 * counterpart in the reference): how many steps ahead each step's line of the item's key
 * is read from memory into registers, 0 (read at the step that stores it to LDS), 1 or 2.
 * The autotuner picks it for the Miller configuration it chooses.  set: force a depth on
 * every form (-1 = back to the autotuned or default depth); a depth that the chosen
 * kernel's register budget cannot hold runs as the largest one below it.  get: the depth
 * a call of the plain form (form 1) or the batched form (form 0) of M items runs, or -1. */
int blsgpu_set_line_depth(int depth);
int blsgpu_get_line_depth(int form, int M);
/* Autotuner on a device-resident job of M items, timed with HIP events (the median of three runs;
 * a candidate 25 % slower than the best so far on its first run is not run again).
 * It times every final-exponentiation candidate alone, then every Miller-loop candidate
 * with each of the fastest final-exponentiation candidates inside the whole kernel
 * sequence (the subgroup test runs concurrently with the pairing kernels, so their best
 * configuration depends on it).
 * form 0 (batched): candidates are lanes per item x workgroups per CU; the fastest pair
 *   is recorded for M's size class (ceil(log2 M)); later batched calls and serving steps
 *   use the entry of the nearest class.
 * form 1 (plain): lanes-per-item candidates with one workgroup per tile; the fastest
 *   pair becomes the plain form's engine (as blsgpu_set_engine(1, ...)).
 * Then, unless a depth is forced (blsgpu_set_line_depth), every key-line depth the chosen
 * Miller kernel holds is timed the same way and the fastest is kept with the pair
 * (blsgpu_get_line_depth reports it).
 * cfg_out (nullable) = {g_miller, g_final_exp, wpc_miller, wpc_final_exp}; ms_out
 * (nullable) = {best whole-sequence time, best final-exponentiation time} in ms.
 * Synchronous. */
int blsgpu_autotune_device(const uint8_t* d_sigs48, const uint64_t* d_neg_hms, const int32_t* d_key_idx, int M,
                           int form, void* stream, int cfg_out[4], double ms_out[2]);

/* Engine selection: mode 0 = one thread per verification (thread-per-item kernels),
 * mode 1 = a group of G lanes of one wavefront per verification (g_miller in
 * {8, 12, 13, 16, 20, 24, 32} for the Miller loop, g_final_exp in {9, 12, 16, 18, 24, 27, 32} for the final
 * exponentiation).  The
 * default is mode 1 with 12/16. */
int blsgpu_set_engine(int mode, int g_miller, int g_final_exp);
int blsgpu_get_engine(int* mode, int* g_miller, int* g_final_exp);

/* Per-kernel timing with HIP events recorded between the kernels of every device-form
 * call while enabled.  phases: 0 decompress, 1 -H(m) conversion, 2 Miller loop,
 * 3 final exponentiation.  ms[k] = summed milliseconds; *calls = calls accumulated. */
int blsgpu_set_profiling(int on);
int blsgpu_phase_times(double ms[4], int* calls);
/* The same for gpu_batch_bls_verify / blsgpu_batch_verify_gt calls made while profiling
 * is on: ms[0] from the first host-to-device copy of the caller's arrays until the last
 * key table has landed, ms[1] from then until the kernels end (batches of 1024 items or
 * more run the Miller loop of their first half during the copies of the second half,
 * so part of the kernel time falls in ms[0]), ms[2] the device-to-host copies of the
 * results, summed over *calls calls. */
int blsgpu_legacy_phase_times(double ms[3], int* calls);

/* Last error message of the calling thread ("" if none). */
const char* blsgpu_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
