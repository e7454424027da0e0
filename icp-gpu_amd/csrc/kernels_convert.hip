// Conversions of the legacy ABI's inputs from the reference's boundary form (6 u64 per
// field element, Montgomery R = 2^384) to the internal one, run on every legacy call.
// A translation unit of their own so that the Montgomery product is inlined: a call to
// it saves registers to scratch memory in every thread, which tripled these kernels'
// memory traffic.
#include <hip/hip_runtime.h>

#include "curve.hpp"
#include "kernels.hpp"

namespace bls {

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

__device__ __forceinline__ void st_fp_soa(i32* base, int w0, int M, int i, const Fp& v) {
#pragma unroll
  for (int j = 0; j < NL; j++) base[(size_t)(w0 + j) * M + i] = v.l[j];
}

// legacy ABI points: M x 12 u64 (x | y, reference Montgomery form) -> internal SoA
__global__ void k_points_from_ref(const u64* __restrict__ in, i32* __restrict__ out, int M) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  st_fp_soa(out, 0, M, i, ld_ref(in + (size_t)i * 12));
  st_fp_soa(out, NL, M, i, ld_ref(in + (size_t)i * 12 + 6));
}

// legacy ABI line tables: M x 2448 u64 (68 x 6 Fp, reference form) -> internal tables,
// one after the other (the key-major layout of the resident keys).  Both sides are flat
// arrays of field elements (element g at in + 6 g and out + 14 g), so a workgroup takes
// LINES_PER_WG consecutive elements: their input words are read into LDS with one
// coalesced load per lane and word, each thread converts one element, and the output
// words leave LDS as contiguous 16-byte stores (a thread's own 14 words are 56 bytes
// apart from its neighbours', which wrote partial cache lines from every store).
constexpr int LINES_PER_WG = 256;
__global__ __launch_bounds__(LINES_PER_WG) void k_lines_from_ref(const u64* __restrict__ in, i32* __restrict__ out,
                                                                 int M) {
  __shared__ u64 win[LINES_PER_WG * 6];
  __shared__ __align__(16) i32 wout[LINES_PER_WG * NL];
  const size_t n = (size_t)M * (NUM_COEFFS * 6);
  const size_t g0 = (size_t)blockIdx.x * LINES_PER_WG;
  const int cnt = n - g0 < (size_t)LINES_PER_WG ? (int)(n - g0) : LINES_PER_WG;
  const u64* src = in + g0 * 6;
  for (int k = threadIdx.x; k < cnt * 6; k += LINES_PER_WG) win[k] = src[k];
  __syncthreads();
  if ((int)threadIdx.x < cnt) {
    u32 w[12];
#pragma unroll
    for (int j = 0; j < 6; j++) {
      const u64 v = win[threadIdx.x * 6 + j];
      w[2 * j] = (u32)v;
      w[2 * j + 1] = (u32)(v >> 32);
    }
    const Fp v = from_ref(w);
#pragma unroll
    for (int l = 0; l < NL; l++) wout[threadIdx.x * NL + l] = v.l[l];
  }
  __syncthreads();
  // out + 14 g0 is 16-byte aligned (g0 is a multiple of 256; the launcher checks out)
  const int words = cnt * NL;
  int4* d4 = reinterpret_cast<int4*>(out + g0 * NL);
  const int4* s4 = reinterpret_cast<const int4*>(wout);
  for (int k = threadIdx.x; k < words / 4; k += LINES_PER_WG) d4[k] = s4[k];
  for (int k = (words & ~3) + threadIdx.x; k < words; k += LINES_PER_WG) out[g0 * NL + k] = wout[k];
}

int launch_points_from_ref(const uint64_t* in, int32_t* out, int M, hipStream_t s) {
  hipLaunchKernelGGL(k_points_from_ref, dim3(nblk(M, 64)), dim3(64), 0, s, in, out, M);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_lines_from_ref(const uint64_t* in, int32_t* out, int M, hipStream_t s) {
  if (M <= 0 || (reinterpret_cast<uintptr_t>(out) & 15)) return -1;
  long long n = (long long)M * NUM_COEFFS * 6;
  hipLaunchKernelGGL(k_lines_from_ref, dim3(nblk(n, LINES_PER_WG)), dim3(LINES_PER_WG), 0, s, in, out, M);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
}  // namespace bls
