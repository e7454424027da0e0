// Caller-side harness for the certifier: the share loop of one certification height
// (ic-integration/patches/ic-gpu-consensus.patch:98-120, where each pre-validated share
// goes through crypto.verify one at a time) restated in C++ against libgpu_msm.so.
//
//   * the library is used as the replica loads it (lib.rs:92-122, 38-41): dlopen, the
//     three-symbol check, gpu_crypto_init() == 0;
//   * the committee's 96-byte public keys are registered once
//     (blsgpu_register_keys_compressed; a key that does not decode gets no handle);
//   * all shares of the height are verified in one blsgpu_verify_messages call over the
//     certified content (hashed to G1 on the device with the caller's DST);
//   * actions as in the loop: 1 MoveToValidated, 0 HandleInvalid (bad signature or a
//     signature that does not decode), -1 no action (the signer's key is unusable: the
//     loop's ValidationFailed branch continues);
//   * any library failure sends the whole height to the CPU path, here the test
//     oracle: orc_g2_decompress / orc_g2_prepare for the key, orc_g1_decompress and
//     orc_verify for the share, with -H(content) supplied by the caller (the CPU path
//     hashes with ic_bls12_381).
//
// Usage: certifier_harness <libgpu_msm.so> <liboracle.so> <input.bin> <output.bin>
//   input:  int32 N shares, int32 K keys, int32 content bytes, int32 DST bytes, then the
//           content, the DST, K x 96 B keys, N x (int32 signer, 48 B signature),
//           12 u64 -H(content) (reference form, for the CPU path)
//   output: int32 path (1 = GPU, 0 = CPU fallback), then N int32 actions
#include <dlfcn.h>
#include <stdint.h>

#include <cstdio>
#include <vector>

typedef int (*init_fn)(void);
typedef int (*reg_fn)(const uint8_t*, int, int32_t*, int32_t*, int);
typedef int (*verify_msgs_fn)(const uint8_t*, const uint8_t*, const int64_t*, const uint8_t*, int, const int32_t*,
                              int32_t*, uint64_t*, int);
typedef void (*reset_fn)(void);
typedef int (*dec1_fn)(const uint8_t*, uint64_t*);
typedef int (*dec2_fn)(const uint8_t*, uint64_t*);
typedef void (*prep_fn)(const uint64_t*, uint64_t*);
typedef void (*gen2_fn)(uint64_t*);
typedef int (*orc_verify_fn)(const uint64_t*, const uint64_t*, const uint64_t*, const uint64_t*, uint64_t*);

static bool gpu_height(void* h, int N, int K, const std::vector<uint8_t>& content, const std::vector<uint8_t>& dst,
                       const std::vector<uint8_t>& keys, const std::vector<int32_t>& signer,
                       const std::vector<uint8_t>& sigs, std::vector<int32_t>& act) {
  if (!h || !dlsym(h, "gpu_msm_g1") || !dlsym(h, "gpu_jacobian_to_affine") || !dlsym(h, "gpu_g1_decompress_batch"))
    return false;
  init_fn init = (init_fn)dlsym(h, "gpu_crypto_init");
  reg_fn reg = (reg_fn)dlsym(h, "blsgpu_register_keys_compressed");
  verify_msgs_fn verify = (verify_msgs_fn)dlsym(h, "blsgpu_verify_messages");
  reset_fn reset = (reset_fn)dlsym(h, "blsgpu_reset_keys");
  if (!init || !reg || !verify || !reset || init() != 0) return false;
  reset();
  std::vector<int32_t> handle(K, -1), status(K, 0);
  if (reg(keys.data(), K, handle.data(), status.data(), 1) < 0) return false;
  // shares whose signer has a usable key, in order
  std::vector<int> idx;
  for (int i = 0; i < N; i++) act[i] = -1;
  for (int i = 0; i < N; i++)
    if (signer[i] >= 0 && signer[i] < K && handle[signer[i]] >= 0) idx.push_back(i);
  const int M = (int)idx.size();
  if (M == 0) return true;
  std::vector<uint8_t> s48((size_t)M * 48), msgs((size_t)M * content.size());
  std::vector<int64_t> offs(M + 1);
  std::vector<int32_t> kidx(M), res(M);
  for (int j = 0; j < M; j++) {
    const int i = idx[j];
    for (int b = 0; b < 48; b++) s48[(size_t)j * 48 + b] = sigs[(size_t)i * 48 + b];
    for (size_t b = 0; b < content.size(); b++) msgs[(size_t)j * content.size() + b] = content[b];
    offs[j] = (int64_t)j * (int64_t)content.size();
    kidx[j] = handle[signer[i]];
  }
  offs[M] = (int64_t)M * (int64_t)content.size();
  if (verify(s48.data(), msgs.data(), offs.data(), dst.data(), (int)dst.size(), kidx.data(), res.data(), nullptr, M))
    return false;
  for (int j = 0; j < M; j++) act[idx[j]] = res[j] == 1 ? 1 : 0;
  return true;
}

static bool cpu_height(void* o, int N, int K, const std::vector<uint8_t>& keys, const std::vector<int32_t>& signer,
                       const std::vector<uint8_t>& sigs, const std::vector<uint64_t>& nhm, std::vector<int32_t>& act) {
  dec1_fn dec1 = o ? (dec1_fn)dlsym(o, "orc_g1_decompress") : nullptr;
  dec2_fn dec2 = o ? (dec2_fn)dlsym(o, "orc_g2_decompress") : nullptr;
  prep_fn prep = o ? (prep_fn)dlsym(o, "orc_g2_prepare") : nullptr;
  gen2_fn gen2 = o ? (gen2_fn)dlsym(o, "orc_g2_generator") : nullptr;
  orc_verify_fn vfy = o ? (orc_verify_fn)dlsym(o, "orc_verify") : nullptr;
  if (!dec1 || !dec2 || !prep || !gen2 || !vfy) return false;
  std::vector<uint64_t> g2(24), g2c(2448), lines((size_t)K * 2448);
  gen2(g2.data());
  prep(g2.data(), g2c.data());
  std::vector<int> usable(K, 0);
  for (int k = 0; k < K; k++) {
    uint64_t q[24];
    if (dec2(&keys[(size_t)k * 96], q) != 0) continue;  // identity and rejected encodings
    prep(q, &lines[(size_t)k * 2448]);
    usable[k] = 1;
  }
  for (int i = 0; i < N; i++) {
    if (signer[i] < 0 || signer[i] >= K || !usable[signer[i]]) {
      act[i] = -1;
      continue;
    }
    uint64_t s[12];
    if (dec1(&sigs[(size_t)i * 48], s) != 0) {
      act[i] = 0;
      continue;
    }
    act[i] = vfy(s, nhm.data(), g2c.data(), &lines[(size_t)signer[i] * 2448], nullptr) ? 1 : 0;
  }
  return true;
}

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s lib oracle in out\n", argv[0]);
    return 2;
  }
  FILE* f = fopen(argv[3], "rb");
  if (!f) return 2;
  int32_t hdr[4];
  if (fread(hdr, 4, 4, f) != 4 || hdr[0] <= 0 || hdr[1] <= 0 || hdr[2] < 0 || hdr[3] < 1 || hdr[3] > 255) return 2;
  const int N = hdr[0], K = hdr[1];
  std::vector<uint8_t> content(hdr[2]), dst(hdr[3]), keys((size_t)K * 96), sigs((size_t)N * 48);
  std::vector<int32_t> signer(N);
  std::vector<uint64_t> nhm(12);
  bool ok = fread(content.data(), 1, content.size(), f) == content.size() &&
            fread(dst.data(), 1, dst.size(), f) == dst.size() && fread(keys.data(), 1, keys.size(), f) == keys.size();
  for (int i = 0; ok && i < N; i++)
    ok = fread(&signer[i], 4, 1, f) == 1 && fread(&sigs[(size_t)i * 48], 1, 48, f) == 48;
  ok = ok && fread(nhm.data(), 8, 12, f) == 12;
  fclose(f);
  if (!ok) return 2;

  std::vector<int32_t> act(N, -1);
  int path = gpu_height(dlopen(argv[1], RTLD_NOW | RTLD_GLOBAL), N, K, content, dst, keys, signer, sigs, act) ? 1 : 0;
  if (!path && !cpu_height(dlopen(argv[2], RTLD_NOW), N, K, keys, signer, sigs, nhm, act)) return 3;
  f = fopen(argv[4], "wb");
  if (!f) return 2;
  fwrite(&path, 4, 1, f);
  fwrite(act.data(), 4, act.size(), f);
  fclose(f);
  return 0;
}
