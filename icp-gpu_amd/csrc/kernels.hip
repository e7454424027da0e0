// Batch BLS verification kernels for gfx950 (one thread per item).
//
// Device data formats (all little-endian):
//   internal Fp  : 14 x i32 limbs (field.hpp), value congruent to x * 2^392 mod p
//   G1 points    : SoA over items, word w (0..27 = x limbs, y limbs) of item i at
//                  pts[w * M + i]  (coalesced across a wavefront)
//   line tables  : one key = 68 lines x (c0, c1, c2 in Fp2) = 68 x 84 words; per-key
//                  tables and the per-item tables of the legacy ABI are stored key-major
//                  (one table after the other, read with 16-byte loads)
//   Miller value : SoA, 168 words (12 Fp) per item
#include <hip/hip_runtime.h>

#include "curve.hpp"
#include "kernels.hpp"

namespace bls {

constexpr int LINE_WORDS = 84;
constexpr int KEY_WORDS = NUM_COEFFS * LINE_WORDS;  // 5712
constexpr int PT_WORDS = 2 * NL;                    // 28
constexpr int F12_WORDS = 12 * NL;                  // 168

__device__ __forceinline__ Fp ld_fp_soa(const i32* base, int w0, int M, int i) {
  Fp r;
#pragma unroll
  for (int j = 0; j < NL; j++) r.l[j] = base[(size_t)(w0 + j) * M + i];
  return r;
}
__device__ __forceinline__ void st_fp_soa(i32* base, int w0, int M, int i, const Fp& v) {
#pragma unroll
  for (int j = 0; j < NL; j++) base[(size_t)(w0 + j) * M + i] = v.l[j];
}
__device__ __forceinline__ Fp ld_fp_aos(const i32* p) {
  Fp r;
#pragma unroll
  for (int j = 0; j < NL; j++) r.l[j] = p[j];
  return r;
}
__device__ __forceinline__ void st_fp_aos(i32* p, const Fp& v) {
#pragma unroll
  for (int j = 0; j < NL; j++) p[j] = v.l[j];
}

// ---- line sources for the Miller loop ----
struct LinesAoS {  // one key table (same for every lane or chosen per lane)
  const i32* t;
  __device__ __forceinline__ Line operator()(int k) const {
    const int4* q = reinterpret_cast<const int4*>(t + k * LINE_WORDS);
    i32 w[LINE_WORDS];
#pragma unroll
    for (int j = 0; j < LINE_WORDS / 4; j++) {
      int4 v = q[j];
      w[4 * j] = v.x; w[4 * j + 1] = v.y; w[4 * j + 2] = v.z; w[4 * j + 3] = v.w;
    }
    Line L;
    Fp* f = &L.c0.c0;
#pragma unroll
    for (int e = 0; e < 6; e++)
#pragma unroll
      for (int j = 0; j < NL; j++) f[e].l[j] = w[e * NL + j];
    return L;
  }
};
// ---- kernels ----

// internal SoA points -> reference form x | y (M x 12 u64); identity status -> zeros
__global__ void k_points_to_ref(const i32* __restrict__ pts, const i32* __restrict__ status, u64* __restrict__ out,
                                int M) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  u64* o = out + (size_t)i * 12;
  if (status && status[i] != DEC_OK) {
#pragma unroll
    for (int j = 0; j < 12; j++) o[j] = 0;
    return;
  }
  st_ref(o, ld_fp_soa(pts, 0, M, i));
  st_ref(o + 6, ld_fp_soa(pts, NL, M, i));
}

// two-pair Miller loop.  Pair 1: (sig, G2 generator lines); pair 2: (-H(m), pk lines).
// pk lines come from the key table selected by key_idx[i] (the legacy ABI's per-item
// tables come as M keys with key_idx[i] = i).  sig_status (nullable): DEC_IDENTITY drops pair 1, any other
// non-OK status skips the item (its verdict is set by the final kernel).
__global__ void k_miller(const i32* __restrict__ sig, const i32* __restrict__ nhm, const i32* __restrict__ g2lines,
                         const i32* __restrict__ pklines, const int32_t* __restrict__ key_idx, int nkeys,
                         i32* __restrict__ sig_status, i32* __restrict__ fout, int M) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  bool skip1 = false;
  if (sig_status) {
    int32_t st = sig_status[i];
    if (st != DEC_OK && st != DEC_IDENTITY) return;
    skip1 = st == DEC_IDENTITY;
  }
  if (key_idx[i] < 0 || key_idx[i] >= nkeys) {
    if (sig_status) sig_status[i] = DEC_BAD_KEY;
    return;
  }
  G1Aff p1 = {ld_fp_soa(sig, 0, M, i), ld_fp_soa(sig, NL, M, i)};
  G1Aff p2 = {ld_fp_soa(nhm, 0, M, i), ld_fp_soa(nhm, NL, M, i)};
  Fp12 f;
  LinesAoS gen{g2lines};
  const LinesAoS pk{pklines + (size_t)key_idx[i] * KEY_WORDS};
  f = miller2(p1, gen, p2, pk, skip1, false);
  const Fp* v = &f.c0.c0.c0;
#pragma unroll
  for (int e = 0; e < 12; e++) st_fp_soa(fout, e * NL, M, i, v[e]);
}

// final exponentiation + identity test.  result[i] = 1 valid, 0 invalid, or the
// (negative) decode status of the signature when sig_status is given.  gt_out
// (nullable): M x 72 u64 in reference form (ic_bls12_381 Gt layout).
__global__ void k_final_exp(const i32* __restrict__ fin, const i32* __restrict__ sig_status,
                            int32_t* __restrict__ result, u64* __restrict__ gt_out, int M) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  if (sig_status) {
    int32_t st = sig_status[i];
    if (st != DEC_OK && st != DEC_IDENTITY) {
      result[i] = st;
      if (gt_out) {
        for (int j = 0; j < 72; j++) gt_out[(size_t)i * 72 + j] = 0;
      }
      return;
    }
  }
  Fp12 f;
  Fp* v = &f.c0.c0.c0;
#pragma unroll
  for (int e = 0; e < 12; e++) v[e] = ld_fp_soa(fin, e * NL, M, i);
  Fp12 r = final_exp_hard(final_exp_easy(f));
  result[i] = is_one(r) ? 1 : 0;
  if (gt_out) {
    const Fp* rv = &r.c0.c0.c0;
#pragma unroll
    for (int e = 0; e < 12; e++) st_ref(gt_out + (size_t)i * 72 + e * 6, rv[e]);
  }
}

// internal key table -> reference form (2448 u64), one thread per Fp
__global__ void k_table_to_ref(const i32* __restrict__ t, u64* __restrict__ out) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= NUM_COEFFS * 6) return;
  st_ref(out + (size_t)f * 6, ld_fp_aos(t + f * NL));
}

// ---- launch wrappers (host) ----
static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

int launch_table_to_ref(const int32_t* t, uint64_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_table_to_ref, dim3(nblk(NUM_COEFFS * 6, 64)), dim3(64), 0, s, t, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_points_to_ref(const int32_t* pts, const int32_t* status, uint64_t* out, int M, hipStream_t s) {
  hipLaunchKernelGGL(k_points_to_ref, dim3(nblk(M, 64)), dim3(64), 0, s, pts, status, out, M);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_miller(const int32_t* sig, const int32_t* nhm, const int32_t* g2lines, const int32_t* pklines,
                  const int32_t* key_idx, int nkeys, int32_t* sig_status, int32_t* fout, int M, int block,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_miller, dim3(nblk(M, block)), dim3(block), 0, s, sig, nhm, g2lines, pklines, key_idx,
                     nkeys, sig_status, fout, M);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_final_exp(const int32_t* fin, const int32_t* sig_status, int32_t* result, uint64_t* gt_out, int M,
                     int block, hipStream_t s) {
  hipLaunchKernelGGL(k_final_exp, dim3(nblk(M, block)), dim3(block), 0, s, fin, sig_status, result, gt_out, M);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace bls
