// Synthetic workload generator and device probes of libblsgpu_workload.so (declared in
// blsgpu_workload.h beside this file).
//
// Generates the benchmark configurations' inputs on the device: secret keys from a
// seeded splitmix64 stream, pk = sk*G2, H(m) = s*G1 with s from SHA-256 as in the
// reference's oracle (oracle/src/main.rs:383-397), signatures sk*H(m), and the invalid
// categories (signature over another message, x^3+4 without a square root, on-curve
// points outside the r-order subgroup).  The streams are defined so that the same
// (seed, index) gives the same bytes as the CPU oracle's generator in the tests.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "blsgpu_workload.h"
#include "h2c.hpp"

namespace bls {

// ---------------- splitmix64 streams (keyed by seed and index) ----------------
__host__ __device__ inline u64 smix(u64& s) {
  u64 z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__host__ __device__ inline u64 stream_start(u64 seed, u64 index, u64 salt) {
  return seed * 0xd1342543de82ef95ULL + index * 0x9e3779b97f4a7c15ULL + salt;
}
static void rand_bytes(u64 seed, u64 index, uint8_t* out, size_t len) {
  u64 s = stream_start(seed, index, 1);
  for (size_t i = 0; i < len; i += 8) {
    u64 v = smix(s);
    for (size_t b = 0; b < 8 && i + b < len; b++) out[i + b] = (uint8_t)(v >> (8 * b));
  }
}
static void rand_scalar(u64 seed, u64 index, uint8_t out[32]) {
  static const u64 R[4] = {0xffffffff00000001ULL, 0x53bda402fffe5bfeULL, 0x3339d80809a1d805ULL,
                           0x73eda753299d7d48ULL};
  u64 s = stream_start(seed, index, 0x5ca1ab1eULL);
  for (;;) {
    u64 w[4];
    for (int i = 0; i < 4; i++) w[i] = smix(s);
    w[3] &= 0x7fffffffffffffffULL;
    bool lt = false, nz = (w[0] | w[1] | w[2] | w[3]) != 0;
    for (int i = 3; i >= 0; i--) {
      if (w[i] < R[i]) { lt = true; break; }
      if (w[i] > R[i]) break;
    }
    if (lt && nz) {
      for (int i = 0; i < 32; i++) out[i] = (uint8_t)(w[i / 8] >> (8 * (i % 8)));
      return;
    }
  }
}

// SHA-256 on the host, through the compression function of the device hash (h2c.hpp)
static void sha256(const uint8_t* msg, size_t len, uint8_t out[32]) {
  u32 h[8];
  for (int k = 0; k < 8; k++) h[k] = SHA256_IV[k];
  std::vector<uint8_t> buf(msg, msg + len);
  buf.push_back(0x80);
  while (buf.size() % 64 != 56) buf.push_back(0);
  const uint64_t bits = (uint64_t)len * 8;
  for (int k = 7; k >= 0; k--) buf.push_back((uint8_t)(bits >> (8 * k)));
  for (size_t off = 0; off < buf.size(); off += 64) {
    u32 w[16];
    for (int i = 0; i < 16; i++)
      w[i] = (u32)buf[off + 4 * i] << 24 | (u32)buf[off + 4 * i + 1] << 16 | (u32)buf[off + 4 * i + 2] << 8 |
             buf[off + 4 * i + 3];
    sha256_compress(h, w);
  }
  for (int k = 0; k < 8; k++) {
    out[4 * k] = (uint8_t)(h[k] >> 24); out[4 * k + 1] = (uint8_t)(h[k] >> 16);
    out[4 * k + 2] = (uint8_t)(h[k] >> 8); out[4 * k + 3] = (uint8_t)h[k];
  }
}
// s = LE(SHA-256("ic-bls-test-hash" || msg)) with byte 31 &= 0x0f
static void hash_scalar(const uint8_t* msg, size_t len, uint8_t s[32]) {
  std::vector<uint8_t> b(16 + len);
  memcpy(b.data(), "ic-bls-test-hash", 16);
  memcpy(b.data() + 16, msg, len);
  sha256(b.data(), b.size(), s);
  s[31] &= 0x0f;
}

// ---------------- device side ----------------
__device__ __forceinline__ Aff<Fp> g1_generator() {
  // G1 generator, canonical x, y (big-endian hex 17f1d3a7...c6bb, 08b3f481...e7e1) as
  // 12 little-endian 32-bit words, converted to the internal form
  const u32 X[12] = {0xdb22c6bbu, 0xfb3af00au, 0xf97a1aefu, 0x6c55e83fu, 0x171bac58u, 0xa14e3a3fu,
                     0x9774b905u, 0xc3688c4fu, 0x4fa9ac0fu, 0x2695638cu, 0x3197d794u, 0x17f1d3a7u};
  const u32 Y[12] = {0x46c5e7e1u, 0x0caa2329u, 0xa2888ae4u, 0xd03cc744u, 0x2c04b3edu, 0x00db18cbu,
                     0xd5d00af6u, 0xfcf5e095u, 0x741d8ae4u, 0xa09e30edu, 0xe3aaa0f1u, 0x08b3f481u};
  return {from_plain(unpack384(X)), from_plain(unpack384(Y))};
}
__device__ __forceinline__ Aff<Fp2> g2_generator() {
  const u32 XC0[12] = {0xc121bdb8u, 0xd48056c8u, 0xa805bbefu, 0x0bac0326u, 0x7ae3d177u, 0xb4510b64u,
                       0xfa403b02u, 0xc6e47ad4u, 0x2dc51051u, 0x26080527u, 0xf08f0a91u, 0x024aa2b2u};
  const u32 XC1[12] = {0x5d042b7eu, 0xe5ac7d05u, 0x13945d57u, 0x334cf112u, 0xdc7f5049u, 0xb5da61bbu,
                       0x9920b61au, 0x596bd0d0u, 0x88274f65u, 0x7dacd3a0u, 0x52719f60u, 0x13e02b60u};
  const u32 YC0[12] = {0x08b82801u, 0xe1935486u, 0x3baca289u, 0x923ac9ccu, 0x5160d12cu, 0x6d429a69u,
                       0x8cbdd3a7u, 0xadfd9baau, 0xda2e351au, 0x8cc9cdc6u, 0x727d6e11u, 0x0ce5d527u};
  const u32 YC1[12] = {0xf05f79beu, 0xaaa9075fu, 0x5cec1da1u, 0x3f370d27u, 0x572e99abu, 0x267492abu,
                       0x85a763afu, 0xcb3e287eu, 0x2bc28b99u, 0x32acd2b0u, 0x2ea734ccu, 0x0606c4a0u};
  return {{from_plain(unpack384(XC0)), from_plain(unpack384(XC1))},
          {from_plain(unpack384(YC0)), from_plain(unpack384(YC1))}};
}
__device__ __forceinline__ void ld_scalar(const uint8_t* b, u32 k[8]) {
#pragma unroll
  for (int j = 0; j < 8; j++)
    k[j] = (u32)b[4 * j] | ((u32)b[4 * j + 1] << 8) | ((u32)b[4 * j + 2] << 16) | ((u32)b[4 * j + 3] << 24);
}

__global__ __launch_bounds__(64, 1) void k_wl_keys(const uint8_t* __restrict__ sk, u64* __restrict__ pk, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u32 k[8];
  ld_scalar(sk + (size_t)i * 32, k);
  Aff<Fp2> q = to_affine(smul(g2_generator(), k, fp2_one()));
  u64* o = pk + (size_t)i * 24;
  st_ref(o, q.x.c0); st_ref(o + 6, q.x.c1); st_ref(o + 12, q.y.c0); st_ref(o + 18, q.y.c1);
}

// one thread per item: H = hs*G1, nhm = -H, sig by category
__global__ __launch_bounds__(64, 1) void k_wl_items(const uint8_t* __restrict__ hs, const uint8_t* __restrict__ hs_other,
                           const uint8_t* __restrict__ sk, const int32_t* __restrict__ key_idx,
                           const int8_t* __restrict__ cat, u64 seed, uint8_t* __restrict__ sig48,
                           u64* __restrict__ nhm, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u32 k[8], kk[8];
  ld_scalar(hs + (size_t)i * 32, k);
  Aff<Fp> H = to_affine(smul(g1_generator(), k, fp_one()));
  st_ref(nhm + (size_t)i * 12, H.x);
  st_ref(nhm + (size_t)i * 12 + 6, neg(H.y));
  const int c = cat[i];
  u32 be[12];
  if (c == BLSGPU_WL_VALID || c == BLSGPU_WL_WRONG_MSG) {
    Aff<Fp> base = H;
    if (c == BLSGPU_WL_WRONG_MSG) {
      ld_scalar(hs_other + (size_t)i * 32, k);
      base = to_affine(smul(g1_generator(), k, fp_one()));
    }
    ld_scalar(sk + (size_t)key_idx[i] * 32, kk);
    Aff<Fp> S = to_affine(smul(base, kk, fp_one()));
    g1_compress(S, false, be);
  } else {
    // x from the (seed, i) stream with x^3+4 a non-square (NO_SQRT) or a square
    // (NON_SUBGROUP, decoded with the smaller root, sort flag 0)
    const bool want_sq = c == BLSGPU_WL_NON_SUBGROUP;
    u64 s = stream_start(seed, (u64)i, 0xbadc0ffeeULL);
    u32 w[12];
    for (;;) {
      u64 raw[6];
      for (int j = 0; j < 6; j++) raw[j] = smix(s);
      raw[5] &= 0x1fffffffffffffffULL;
      for (int j = 0; j < 6; j++) { w[2 * j] = (u32)raw[j]; w[2 * j + 1] = (u32)(raw[j] >> 32); }
      bool lt = false, decided = false;
      for (int j = 11; j >= 0; j--)
        if (!decided && w[j] != P32[j]) { lt = w[j] < P32[j]; decided = true; }
      if (!lt) continue;
      Fp x = from_plain(unpack384(w));
      Fp y2 = norm(add(mul(sqr(x), x), fp_const(B_G1_M)));
      Fp y = sqrt_cand(y2);
      if (eq(sqr(y), y2) == want_sq) break;
    }
    for (int j = 0; j < 12; j++) be[j] = w[11 - j];
    be[0] |= 0x80000000u;
  }
  uint8_t* o = sig48 + (size_t)i * 48;
  for (int j = 0; j < 12; j++) {
    o[4 * j] = (uint8_t)(be[j] >> 24); o[4 * j + 1] = (uint8_t)(be[j] >> 16);
    o[4 * j + 2] = (uint8_t)(be[j] >> 8); o[4 * j + 3] = (uint8_t)be[j];
  }
}

// signatures sk[key_idx[i]] * H_i of given points H_i (reference form, 12 u64 each),
// compressed (48 bytes)
__global__ __launch_bounds__(64, 1) void k_wl_sign_points(const u64* __restrict__ h_ref, const uint8_t* __restrict__ sk,
                                 const int32_t* __restrict__ key_idx, uint8_t* __restrict__ sig48, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u32 kk[8];
  const u64* h = h_ref + (size_t)i * 12;
  const Aff<Fp> H = {ld_ref(h), ld_ref(h + 6)};
  ld_scalar(sk + (size_t)key_idx[i] * 32, kk);
  u32 be[12];
  g1_compress(to_affine(smul(H, kk, fp_one())), false, be);
  uint8_t* o = sig48 + (size_t)i * 48;
  for (int j = 0; j < 12; j++) {
    o[4 * j] = (uint8_t)(be[j] >> 24); o[4 * j + 1] = (uint8_t)(be[j] >> 16);
    o[4 * j + 2] = (uint8_t)(be[j] >> 8); o[4 * j + 3] = (uint8_t)be[j];
  }
}

// v_mad_u64_u32 issue-rate probe: 8 independent chains per lane, 64 per loop trip
__global__ void k_probe_mad(u64* out, u32 seed, int iters) {
  u32 x = seed + threadIdx.x, y = seed * 3 + blockIdx.x;
  u64 a0 = x, a1 = x + 1, a2 = x + 2, a3 = x + 3, a4 = x + 4, a5 = x + 5, a6 = x + 6, a7 = x + 7;
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int r = 0; r < 8; r++) {
      a0 = (u64)(u32)a0 * y + a0; a1 = (u64)(u32)a1 * y + a1; a2 = (u64)(u32)a2 * y + a2;
      a3 = (u64)(u32)a3 * y + a3; a4 = (u64)(u32)a4 * y + a4; a5 = (u64)(u32)a5 * y + a5;
      a6 = (u64)(u32)a6 * y + a6; a7 = (u64)(u32)a7 * y + a7;
    }
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
}

}  // namespace bls

using namespace bls;

namespace {
thread_local std::string wl_err;
std::mutex wl_mu;
#define WLCHK(x)                                                          \
  do {                                                                    \
    hipError_t e_ = (x);                                                  \
    if (e_ != hipSuccess) {                                               \
      wl_err = std::string(#x) + ": " + hipGetErrorString(e_);            \
      return -1;                                                          \
    }                                                                     \
  } while (0)
struct Scoped {
  std::vector<void*> ptrs;
  ~Scoped() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t alloc(T** p, size_t bytes) {
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
};
}  // namespace

extern "C" {

void blsgpu_wl_secret_keys(uint64_t seed, int n, uint8_t* sk32) {
  for (int i = 0; i < n; i++) rand_scalar(seed, (u64)i, sk32 + (size_t)i * 32);
}

void blsgpu_wl_message(uint64_t seed, int64_t index, uint8_t msg32[32]) { rand_bytes(seed, (u64)index, msg32, 32); }

int blsgpu_wl_public_keys(const uint8_t* sk32, int n, uint64_t* pk_ref) {
  std::lock_guard<std::mutex> lk(wl_mu);
  if (n <= 0) return -1;
  Scoped m;
  uint8_t* dsk;
  u64* dpk;
  WLCHK(m.alloc(&dsk, (size_t)n * 32));
  WLCHK(m.alloc(&dpk, (size_t)n * 24 * 8));
  WLCHK(hipMemcpy(dsk, sk32, (size_t)n * 32, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_wl_keys, dim3((n + 63) / 64), dim3(64), 0, 0, dsk, dpk, n);
  WLCHK(hipGetLastError());
  WLCHK(hipMemcpy(pk_ref, dpk, (size_t)n * 24 * 8, hipMemcpyDeviceToHost));
  return 0;
}

int blsgpu_wl_sign_points(const uint64_t* h_ref, int n, const uint8_t* sk32, int nkeys, const int32_t* key_idx,
                          uint8_t* sig48) {
  std::lock_guard<std::mutex> lk(wl_mu);
  if (n <= 0 || nkeys <= 0) return -1;
  for (int i = 0; i < n; i++)
    if (key_idx[i] < 0 || key_idx[i] >= nkeys) {
      wl_err = "bad key index";
      return -1;
    }
  Scoped m;
  u64* dh;
  uint8_t *dsk, *dsig;
  int32_t* dki;
  WLCHK(m.alloc(&dh, (size_t)n * 96));
  WLCHK(m.alloc(&dsk, (size_t)nkeys * 32));
  WLCHK(m.alloc(&dki, (size_t)n * 4));
  WLCHK(m.alloc(&dsig, (size_t)n * 48));
  WLCHK(hipMemcpy(dh, h_ref, (size_t)n * 96, hipMemcpyHostToDevice));
  WLCHK(hipMemcpy(dsk, sk32, (size_t)nkeys * 32, hipMemcpyHostToDevice));
  WLCHK(hipMemcpy(dki, key_idx, (size_t)n * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_wl_sign_points, dim3((n + 63) / 64), dim3(64), 0, 0, dh, dsk, dki, dsig, n);
  WLCHK(hipGetLastError());
  WLCHK(hipMemcpy(sig48, dsig, (size_t)n * 48, hipMemcpyDeviceToHost));
  return 0;
}

int blsgpu_wl_items(uint64_t seed, int n, const uint8_t* sk32, int nkeys, const int32_t* key_idx,
                    const int8_t* category, uint8_t* sig48, uint64_t* nhm_ref) {
  return blsgpu_wl_items_msg(seed, n, sk32, nkeys, key_idx, nullptr, category, sig48, nhm_ref);
}

int blsgpu_wl_items_msg(uint64_t seed, int n, const uint8_t* sk32, int nkeys, const int32_t* key_idx,
                        const int64_t* msg_idx, const int8_t* category, uint8_t* sig48, uint64_t* nhm_ref) {
  std::lock_guard<std::mutex> lk(wl_mu);
  if (n <= 0 || nkeys <= 0) return -1;
  for (int i = 0; i < n; i++)
    if (key_idx[i] < 0 || key_idx[i] >= nkeys || category[i] < 0 || category[i] > 3) {
      wl_err = "bad key index or category";
      return -1;
    }
  std::vector<uint8_t> hs((size_t)n * 32), ho((size_t)n * 32, 0);
  uint8_t msg[32];
  for (int i = 0; i < n; i++) {
    if (msg_idx && msg_idx[i] < 0) {
      wl_err = "bad message index";
      return -1;
    }
    rand_bytes(seed, 1000000ULL + (u64)(msg_idx ? msg_idx[i] : i), msg, 32);
    hash_scalar(msg, 32, &hs[(size_t)i * 32]);
    if (category[i] == BLSGPU_WL_WRONG_MSG) {
      rand_bytes(seed, 2000000ULL + (u64)i, msg, 32);
      hash_scalar(msg, 32, &ho[(size_t)i * 32]);
    }
  }
  Scoped m;
  uint8_t *dhs, *dho, *dsk, *dsig;
  int32_t* dki;
  int8_t* dcat;
  u64* dnhm;
  WLCHK(m.alloc(&dhs, (size_t)n * 32));
  WLCHK(m.alloc(&dho, (size_t)n * 32));
  WLCHK(m.alloc(&dsk, (size_t)nkeys * 32));
  WLCHK(m.alloc(&dsig, (size_t)n * 48));
  WLCHK(m.alloc(&dki, (size_t)n * 4));
  WLCHK(m.alloc(&dcat, (size_t)n));
  WLCHK(m.alloc(&dnhm, (size_t)n * 96));
  WLCHK(hipMemcpy(dhs, hs.data(), hs.size(), hipMemcpyHostToDevice));
  WLCHK(hipMemcpy(dho, ho.data(), ho.size(), hipMemcpyHostToDevice));
  WLCHK(hipMemcpy(dsk, sk32, (size_t)nkeys * 32, hipMemcpyHostToDevice));
  WLCHK(hipMemcpy(dki, key_idx, (size_t)n * 4, hipMemcpyHostToDevice));
  WLCHK(hipMemcpy(dcat, category, (size_t)n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_wl_items, dim3((n + 63) / 64), dim3(64), 0, 0, dhs, dho, dsk, dki, dcat, (u64)seed, dsig,
                     dnhm, n);
  WLCHK(hipGetLastError());
  WLCHK(hipMemcpy(sig48, dsig, (size_t)n * 48, hipMemcpyDeviceToHost));
  WLCHK(hipMemcpy(nhm_ref, dnhm, (size_t)n * 96, hipMemcpyDeviceToHost));
  return 0;
}

int blsgpu_probe_mad_rate(double* mads_per_s, double* ms) {
  std::lock_guard<std::mutex> lk(wl_mu);
  int dev = 0, cus = 0;
  WLCHK(hipGetDevice(&dev));
  WLCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int B = 256, G = cus * 8, iters = 2000;
  Scoped m;
  u64* out;
  WLCHK(m.alloc(&out, (size_t)G * B * 8));  // one u64 per thread of the grid
  hipEvent_t e0, e1;
  WLCHK(hipEventCreate(&e0));
  WLCHK(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_probe_mad, dim3(G), dim3(B), 0, 0, out, 7u, iters);
  WLCHK(hipGetLastError());
  WLCHK(hipDeviceSynchronize());
  const int reps = 5;
  WLCHK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_probe_mad, dim3(G), dim3(B), 0, 0, out, 7u + r, iters);
  WLCHK(hipEventRecord(e1, 0));
  WLCHK(hipEventSynchronize(e1));
  float t = 0;
  WLCHK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  double per = t / reps;
  if (ms) *ms = per;
  *mads_per_s = (double)G * B * iters * 64 / (per * 1e-3);
  return 0;
}

const char* blsgpu_wl_last_error(void) { return wl_err.c_str(); }

}  // extern "C"
