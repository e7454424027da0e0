// Caller-side harness: the sequence a replica runs against libgpu_msm.so, restated in
// C++ so that the drop-in can be exercised the way the reference's Rust crate uses it.
//
//   * load_msm_lib (ic-integration/gpu_crypto/lib.rs:92-122): dlopen the library and use
//     it only if gpu_msm_g1, gpu_jacobian_to_affine and gpu_g1_decompress_batch all resolve;
//   * init (lib.rs:38-41): gpu_crypto_init() == 0 means a device is present;
//   * batch_bls_verify (lib.rs:238-281): one gpu_batch_bls_verify call, result == 1 -> true;
//   * bls_verify / verify_bls_signature_batch_gpu fallback (lib.rs:285-304): any failure
//     sends the whole batch to the CPU path, here the test oracle's orc_verify.
//
// Usage: replica_harness <libgpu_msm.so> <liboracle.so> <input.bin> <output.bin>
//   input:  int32 M, then M x 12 u64 sigs, M x 12 u64 -H(m), 2448 u64 generator lines,
//           M x 2448 u64 key lines
//   output: int32 path (1 = GPU, 0 = CPU fallback), then M int32 verdicts (1/0)
#include <dlfcn.h>
#include <stdint.h>

#include <cstdio>
#include <vector>

typedef int (*init_fn)(void);
typedef int (*verify_fn)(const uint64_t*, const uint64_t*, const uint64_t*, const uint64_t*, int32_t*, int);
typedef int (*orc_verify_fn)(const uint64_t*, const uint64_t*, const uint64_t*, const uint64_t*, uint64_t*);

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s lib oracle in out\n", argv[0]);
    return 2;
  }
  FILE* f = fopen(argv[3], "rb");
  if (!f) return 2;
  int32_t M = 0;
  if (fread(&M, 4, 1, f) != 1 || M <= 0) return 2;
  std::vector<uint64_t> sig((size_t)M * 12), nhm((size_t)M * 12), g2c(2448), pkc((size_t)M * 2448);
  bool ok = fread(sig.data(), 8, sig.size(), f) == sig.size() && fread(nhm.data(), 8, nhm.size(), f) == nhm.size() &&
            fread(g2c.data(), 8, g2c.size(), f) == g2c.size() && fread(pkc.data(), 8, pkc.size(), f) == pkc.size();
  fclose(f);
  if (!ok) return 2;

  std::vector<int32_t> res(M, 0);
  int path = 0;
  void* h = dlopen(argv[1], RTLD_NOW | RTLD_GLOBAL);
  if (h && dlsym(h, "gpu_msm_g1") && dlsym(h, "gpu_jacobian_to_affine") && dlsym(h, "gpu_g1_decompress_batch")) {
    init_fn init = (init_fn)dlsym(h, "gpu_crypto_init");
    verify_fn verify = (verify_fn)dlsym(h, "gpu_batch_bls_verify");
    if (init && verify && init() == 0 &&
        verify(sig.data(), nhm.data(), g2c.data(), pkc.data(), res.data(), M) == 0)
      path = 1;
  }
  if (!path) {
    void* o = dlopen(argv[2], RTLD_NOW);
    orc_verify_fn cpu = o ? (orc_verify_fn)dlsym(o, "orc_verify") : nullptr;
    if (!cpu) return 3;
    for (int i = 0; i < M; i++)
      res[i] = cpu(&sig[(size_t)i * 12], &nhm[(size_t)i * 12], g2c.data(), &pkc[(size_t)i * 2448], nullptr) ? 1 : 0;
  }
  f = fopen(argv[4], "wb");
  if (!f) return 2;
  fwrite(&path, 4, 1, f);
  for (int i = 0; i < M; i++) {
    const int32_t v = res[i] == 1 ? 1 : 0;  // the caller maps r == 1 to true
    fwrite(&v, 4, 1, f);
  }
  fclose(f);
  return 0;
}
