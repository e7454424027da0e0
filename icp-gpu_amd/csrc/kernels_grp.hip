// Lanes-per-pairing kernels: a group of G lanes of one wavefront verifies one item.
//
// Each item owns a block of Fp "slots" in LDS (SLOT = 14 words, one per limb).  Every
// extension-field operation is a bilinear schedule (schedules_gen.hpp): the group's
// lanes split its independent Fp products, then its output combinations.  Most ops run
// through one round interpreter (run_op, a single multiplier body); the hot ones (the
// Miller loop's line ops and square, the cyclotomic square and product of the final
// exponentiation) run as straight-line code from the same tables taken as compile-time
// constants (run_op_c), each emitted once.
//
// A workgroup is W wavefronts holding 64 / L items, L = G / W lanes of each wavefront
// per item (Cfg in grp_config.hpp): one wavefront for G <= 27; two for G = 24 and 32, so
// that two wavefronts share each SIMD when the items exactly fill the GPU at one
// workgroup per tile (a lone wavefront issues at most one instruction per ~6 cycles,
// two together one per ~4.9).  Lanes past the last full group idle but take part in the
// barriers.  The Miller value passes between the two kernels item-major (168
// consecutive words per item), so the G lanes of an item read and write whole cache
// lines whichever XCD runs the workgroup.
//
// This file is compiled once per lane count and kernel (-DGRP_LANES=G -DGRP_KIND=0 for
// the Miller loop, 1 for the final exponentiation, 2 with GRP_LANES=12 for the compact
// final exponentiation); the launchers are instantiated at the end, the dispatch over G
// is in grp_dispatch.cpp.
#include <hip/hip_runtime.h>

#include "curve.hpp"
#include "grp_config.hpp"
#include "kernels.hpp"
#include "schedules_gen.hpp"

#if !defined(GRP_LANES) || !defined(GRP_KIND)
#error "kernels_grp.hip is compiled once per lane count and kernel: -DGRP_LANES=<G> -DGRP_KIND=<0|1|2>"
#endif

namespace bls {

constexpr int SLOT = 14;

// All LDS of these kernels is one dynamic array; device functions address it by word
// offset so that they stay in the LDS address space across calls.
extern __shared__ __align__(16) i32 g_lds[];

// Slots are 56 bytes apart, so the slots that the lanes of one LDS access read start at
// 8-byte offsets spread over the banks instead of all at 64-byte boundaries (with
// 64-byte slots, lanes reading different slots collided on the same banks); a slot moves
// as three ds_read2_b64 / ds_write2_b64 and one 8-byte access.
__device__ __forceinline__ Fp lds_ld(int off) {
  const int2* q = reinterpret_cast<const int2*>(g_lds + off);
  Fp r;
#pragma unroll
  for (int j = 0; j < 7; j++) {
    const int2 v = q[j];
    r.l[2 * j] = v.x;
    r.l[2 * j + 1] = v.y;
  }
  return r;
}
__device__ __forceinline__ void lds_st(int off, const Fp& v) {
  int2* q = reinterpret_cast<int2*>(g_lds + off);
#pragma unroll
  for (int j = 0; j < 7; j++) q[j] = make_int2(v.l[2 * j], v.l[2 * j + 1]);
}

// word offsets of the four operand regions of a schedule: A, B, T (products), K (constants)
struct Regions {
  int a, b, t, k;
  __device__ __forceinline__ int base(int reg) const { return reg == 0 ? a : reg == 1 ? b : reg == 2 ? t : k; }
  static_assert(SLOT == 14, "region-relative terms hold slot * 14");
  // the same, branch-free (reg differs between lanes)
  __device__ __forceinline__ int select(u32 reg) const {
    const int m1 = -(int)(reg & 1), m2 = -(int)((reg >> 1) & 1);
    const int lo = a ^ ((a ^ b) & m1), hi = t ^ ((t ^ k) & m1);
    return lo ^ ((lo ^ hi) & m2);
  }
};

// ---- the round interpreter ----
// A lane's record of one round is held as 32-bit words (two halfword terms each).
constexpr int REC_WORDS = 4 * sched::MAX_OUT_Q;
static_assert(sched::PROD_Q <= sched::MAX_OUT_Q, "record buffer");

template <int RW>
__device__ __forceinline__ u32 rec_hw(const u32 (&rec)[RW], int k) {
  return (k & 1) ? (rec[k >> 1] >> 16) : (rec[k >> 1] & 0xffffu);
}
// the record tables, addressed as global memory (a generic pointer would make the record
// loads flat loads, which the LDS wait counter also counts)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef const u32x4 __attribute__((address_space(1))) GRec;
typedef const uint32_t __attribute__((address_space(1))) GWord;

// this lane's record of a round: q 16-byte units at desc + sub * q (q is uniform)
__device__ __forceinline__ void load_record(u32 (&rec)[REC_WORDS], GRec* __restrict__ desc, u32 at, int q,
                                            int sub) {
  GRec* p = desc + at + (u32)sub * q;
#pragma unroll
  for (int u = 0; u < sched::MAX_OUT_Q; u++) {
    if (u < q) {
      const u32x4 v = p[u];
      rec[4 * u] = v.x;
      rec[4 * u + 1] = v.y;
      rec[4 * u + 2] = v.z;
      rec[4 * u + 3] = v.w;
    }
  }
}

// acc += sum of the n terms at halfwords FIRST.. of the record (n uniform, n <= MAXN).
// Terms are processed four at a time, their LDS loads issued before the arithmetic;
// each term is a slot times a signed coefficient, accumulated in 64 bits per limb
// (one v_mad_i64_i32 per limb).  Zero-coefficient terms are padding.
template <int FIRST, int MAXN>
__device__ __forceinline__ void accumulate(i64 (&acc)[NL], const u32 (&rec)[REC_WORDS], int n, const Regions& R) {
#pragma unroll
  for (int c = 0; c < MAXN; c += 4) {
    if (c < n) {
      Fp x[4];
      i32 k[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (c + j < MAXN && c + j < n) {
          const u32 t = rec_hw(rec, FIRST + c + j);
          x[j] = lds_ld(R.select(t >> 6) + (int)(t & 63) * SLOT);
          k[j] = (i32)(int8_t)(t >> 8);
        }
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (c + j < MAXN && c + j < n) {
#pragma unroll
          for (int i = 0; i < NL; i++) acc[i] += (i64)x[j].l[i] * k[j];
        }
      }
    }
  }
}

// 64-bit limb sums -> limbs 0..12 in [0, 2^28), the top limb signed (value unchanged)
__device__ __forceinline__ Fp carry64(const i64 (&acc)[NL]) {
  Fp r;
  i64 c = 0;
#pragma unroll
  for (int i = 0; i < NL - 1; i++) {
    const i64 v = acc[i] + c;
    r.l[i] = (i32)v & MASK28;
    c = v >> 28;
  }
  r.l[NL - 1] = (i32)(acc[NL - 1] + c);
  return r;
}

template <class P>
__device__ __forceinline__ P* uniform_ptr(P* p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (P*)(((uint64_t)hi << 32) | lo);
}

// One operation for the item of this lane's group, as rounds of the per-G tables:
// product rounds (each lane multiplies at most one pair of operand sums into T), a
// barrier, output rounds (each lane forms at most one output in D), a barrier.  Every
// lane of the wavefront runs the same instruction stream whatever its record holds.
// D may alias A (every op used in place reads no other input component in its
// outputs).  Not inlined: one copy of the multiplier serves every call site.
// One instantiation per register budget of the calling kernels (WPE waves per SIMD), so
// that each copy is compiled within its callers' budget.
template <int WPE>
__device__ __noinline__ void run_op_impl(sched::COpRounds* ops, sched::CRound* rounds, GRec* __restrict__ desc,
                                         int op, int sub, bool active, int A, int B, int D, int T, int K) {
  // the table pointers, the op and the round shapes are the same on every lane: make
  // that visible to the compiler so that they live in scalar registers, the tables are
  // read with scalar loads and every branch on them is a scalar branch
  ops = uniform_ptr(ops);
  rounds = uniform_ptr(rounds);
  desc = uniform_ptr(desc);
  op = __builtin_amdgcn_readfirstlane(op);
  K = __builtin_amdgcn_readfirstlane(K);
  const Regions R{A, B, T, K};
  const int first = ops[op].first, np = ops[op].np, nr = np + ops[op].no;
  u32 cur[REC_WORDS], nxt[REC_WORDS];
  if (nr > 0) load_record(cur, desc, rounds[first].desc, rounds[first].q, sub);
  for (int r = 0; r < nr; r++) {
    const int na = rounds[first + r].na, nb = rounds[first + r].nb, fl = rounds[first + r].flags;
    if (r + 1 < nr) load_record(nxt, desc, rounds[first + r + 1].desc, rounds[first + r + 1].q, sub);
    const u32 dst = rec_hw(cur, 0);
    // a wavefront none of whose lanes has a product or output in this round skips it
    // (the records of a round list the busy lanes first, so with two wavefronts per
    // item the second one is often idle)
    const bool work = __ballot(dst != sched::IDLE) != 0;
    i64 acc[NL];
#pragma unroll
    for (int i = 0; i < NL; i++) acc[i] = 0;
    if (r < np) {
      if (work) {
        Fp a, b;
        if (fl & 1) {
          const u32 t = rec_hw(cur, 1);
          a = lds_ld(R.select(t >> 6) + (int)(t & 63) * SLOT);
        } else {
          accumulate<1, 8>(acc, cur, na, R);
          a = carry64(acc);
#pragma unroll
          for (int i = 0; i < NL; i++) acc[i] = 0;
        }
        if (fl & 2) {
          const u32 t = rec_hw(cur, sched::B_AT);
          b = lds_ld(R.select(t >> 6) + (int)(t & 63) * SLOT);
        } else {
          accumulate<sched::B_AT, 8>(acc, cur, nb, R);
          b = carry64(acc);
        }
        const Fp p = mul(a, b);
        if (active && dst != sched::IDLE) lds_st(T + (int)dst * SLOT, p);
      }
      if (r + 1 == np) __syncthreads();
    } else if (work) {
      accumulate<1, 2 * REC_WORDS - 1>(acc, cur, na, R);
      const Fp o = fold_top(carry64(acc));
      if (active && dst != sched::IDLE) lds_st(D + (int)dst * SLOT, o);
    }
    if (r + 1 < nr) {
#pragma unroll
      for (int w = 0; w < REC_WORDS; w++) cur[w] = nxt[w];
    }
  }
  __syncthreads();
}
// WPE: the waves per SIMD that the calling kernel is compiled for (its register budget)
template <int G, int WPE>
__device__ __forceinline__ void run_op(int op, int sub, bool active, int A, int B, int D, int T, int K) {
  run_op_impl<WPE>(sched::Tab<G>::ops(), sched::Tab<G>::rounds(), (GRec*)sched::Tab<G>::desc(), op, sub, active, A, B, D,
                         T, K);
}

// ---- ops with compile-time round shapes ----
// The hot ops (the cyclotomic square and the Fp12 product of the final exponentiation,
// the line ops and the square of the Miller loop) run as straight-line code built from
// the same tables taken as compile-time constants (Tab<G>::OPS_C / ROUNDS_C): term
// counts, copy flags and record sizes are immediates, so a round has no scalar
// bookkeeping, no per-term branches, no accumulator clearing (the first term sets it)
// and no record copies.  The records themselves (which slots, which coefficients) are
// still read per lane from DESC.

// A uniform 1 the compiler cannot see through: code guarded by it forms basic blocks
// of its own, which the scheduler does not interleave (one block per group of terms
// keeps their LDS loads and arithmetic together, as in the interpreter).
__device__ __forceinline__ int opaque_one() {
  int v;
  asm volatile("s_mov_b32 %0, 1" : "=s"(v));
  return v;
}


// Term k of the record: its slot's LDS word offset and its coefficient.  REG is the
// region of every term of the list (the slot offset relative to it is in the record)
// or sched::MIXED (region, slot and coefficient fields as in the interpreter's tables).
template <int REG, int RW>
__device__ __forceinline__ int term_at(const u32 (&rec)[RW], int k, const Regions& R, i32& coef) {
  const u32 t = rec_hw(rec, k);
  if constexpr (REG == sched::MIXED) {
    coef = (i32)(int8_t)(t >> 8);
    return R.select(t >> 6) + (int)(t & 63) * SLOT;
  } else if constexpr (REG >= sched::PAIR) {  // two regions: bit 10 picks the second
    coef = ((i32)(t << 16)) >> 27;  // bits 11..15, signed
    const int b0 = R.base((REG - sched::PAIR) / 4), b1 = R.base((REG - sched::PAIR) % 4);
    return ((t & 1024u) ? b1 : b0) + (int)(t & 1023u);
  } else {
    coef = ((i32)(t << 16)) >> 26;  // bits 10..15, signed
    return R.base(REG) + (int)(t & 1023u);
  }
}

// the sum of the N >= 1 terms at halfwords FIRST.. of the record, carried (CARRY), or,
// for a sum of weight <= 2 of normalized slots (limbs below 2^29, a valid multiplier
// operand as it is), the low words of the limb sums
template <int FIRST, int N, int REG, int RW, bool CARRY = true>
__device__ __forceinline__ Fp sum_terms(const u32 (&rec)[RW], const Regions& R) {
  static_assert(N >= 1 && FIRST + N <= 2 * RW, "terms of a busy round");
  i64 acc[NL];
#pragma unroll
  for (int c = 0; c < N; c += 4) {
    if (c == 0 || opaque_one()) {
      Fp x[4];
      i32 k[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (c + j < N) x[j] = lds_ld(term_at<REG>(rec, FIRST + c + j, R, k[j]));
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (c + j < N) {
#pragma unroll
          for (int i = 0; i < NL; i++) {
            const i64 v = (i64)x[j].l[i] * k[j];
            acc[i] = c + j == 0 ? v : acc[i] + v;
          }
        }
      }
    }
  }
  if constexpr (CARRY) {
    return carry64(acc);
  } else {
    Fp r;
#pragma unroll
    for (int i = 0; i < NL; i++) r.l[i] = (i32)acc[i];
    return r;
  }
}
// a one-term operand with coefficient 1 on every busy lane
template <int AT, int REG, int RW>
__device__ __forceinline__ Fp load_term(const u32 (&rec)[RW], const Regions& R) {
  i32 coef;
  return lds_ld(term_at<REG>(rec, AT, R, coef));
}

// round RI of op OP at G lanes: its shape, and this lane's record
template <int G, int OP, int RI>
struct RoundC {
  static constexpr sched::OpRounds o = sched::Tab<G>::OPS_C[OP];
  static constexpr sched::Round rd = sched::Tab<G>::ROUNDS_C[o.first + RI];
  // records in DESC_S (region-relative terms, lane-major) or in the interpreter's DESC
  static constexpr sched::SRound s0 = sched::Tab<G>::SROUNDS_C[o.first + RI];
  static constexpr bool S = s0.desc != 0xffffffffu;
  static constexpr sched::SRound sr = S ? s0 : sched::SRound{rd.desc, rd.q, sched::MIXED, sched::MIXED, 255, 255};
  static constexpr int NR = o.np + o.no, Q = rd.q, RW = 4 * Q;
  // whether the round reads halfword k of the record: dst, then its A and B terms
  // (product rounds) or its terms (output rounds)
  static constexpr bool reads(int k) {
    if (k == 0) return true;
    if (RI < o.np) return (k >= 1 && k < 1 + rd.na) || (k >= sched::B_AT && k < sched::B_AT + rd.nb);
    return k < 1 + rd.na;
  }
  // Only the words the round reads are loaded (adjacent ones merge into wide loads): a
  // loaded word the round never reads would free its register at once, and reusing it
  // would make the wavefront wait for the load.
  static __device__ __forceinline__ void load(GRec* __restrict__ desc, int sub, u32 (&rec)[RW]) {
    const GWord* p = (const GWord*)(desc + sr.desc + (u32)sub * sr.stride);
#pragma unroll
    for (int w = 0; w < RW; w++) rec[w] = reads(2 * w) || reads(2 * w + 1) ? p[w] : 0u;
  }
};

// Round RI with its record in rec; the record of round RI + 1 is read from memory before
// this round's arithmetic, so that its latency is hidden behind it.
template <int G, int OP, int RI>
__device__ __forceinline__ void static_round(GRec* __restrict__ desc, int sub, bool active, const Regions& R, int D,
                                             const u32 (&rec)[RoundC<G, OP, RI>::RW]) {
  using C = RoundC<G, OP, RI>;
  constexpr sched::OpRounds o = C::o;
  constexpr sched::Round rd = C::rd;
  [[maybe_unused]] u32 nxt[RI + 1 < C::NR ? RoundC<G, OP, (RI + 1 < C::NR ? RI + 1 : RI)>::RW : 1];
  if constexpr (RI + 1 < C::NR) RoundC<G, OP, RI + 1>::load(desc, sub, nxt);
  const u32 dst = rec[0] & 0xffffu;
  // a wavefront none of whose lanes has a product or output in this round skips it
  const bool work = __ballot(dst != sched::IDLE) != 0;
  // In the Miller loop a lane with no product or output in the round (or of an inactive
  // item) sits it out: measured 1.6 % faster (1.276 -> 1.250 ms per 4,096 items at 32
  // lanes).  In the final exponentiation it runs the round on its padding record, which
  // measured 4 % faster than masking it.
  constexpr bool MASK = GRP_KIND == 0;
  const bool busy = work && (!MASK || (active && dst != sched::IDLE));
  if constexpr (RI < o.np) {
    if (busy) {
      Fp a, b;
      if constexpr (rd.flags & 1) a = load_term<1, C::sr.ra>(rec, R);
      else a = sum_terms<1, rd.na, C::sr.ra, C::RW, (C::sr.wa > 2)>(rec, R);
      if constexpr (rd.flags & 2) b = load_term<sched::B_AT, C::sr.rb>(rec, R);
      else b = sum_terms<sched::B_AT, rd.nb, C::sr.rb, C::RW, (C::sr.wb > 2)>(rec, R);
      const Fp pr = mul(a, b);
      if (MASK || (active && dst != sched::IDLE)) lds_st(R.t + (int)dst * SLOT, pr);
    }
    if constexpr (RI + 1 == o.np) __syncthreads();
  } else {
    if (busy) {
      const Fp v = fold_top(sum_terms<1, rd.na, C::sr.ra>(rec, R));
      if (MASK || (active && dst != sched::IDLE)) lds_st(D + (int)dst * SLOT, v);
    }
  }
  if constexpr (RI + 1 < C::NR) static_round<G, OP, RI + 1>(desc, sub, active, R, D, nxt);
}
// run_op for an op known at compile time (same contract as run_op).  Inlined: its
// callers hold it at one call site each inside a function entered rarely (the Miller
// loop's kernel, the final exponentiation's cyclotomic power), as a call of its own
// would save and restore the callee-saved registers it uses on every op.
template <int G, int OP>
__device__ __forceinline__ GRec* op_desc() {
  if constexpr (RoundC<G, OP, 0>::S) return uniform_ptr((GRec*)sched::Tab<G>::sdesc());
  else return uniform_ptr((GRec*)sched::Tab<G>::desc());
}
// the record of an op's first round, for callers that read it once ahead of a loop
template <int G, int OP>
struct FirstRec {
  u32 w[RoundC<G, OP, 0>::RW];
  __device__ __forceinline__ void load(int sub) { RoundC<G, OP, 0>::load(op_desc<G, OP>(), sub, w); }
};
template <int G, int OP>
__device__ __forceinline__ void run_op_c(const FirstRec<G, OP>& first, int sub, bool active, int A, int B, int D, int T,
                                         int K) {
  constexpr sched::OpRounds o = sched::Tab<G>::OPS_C[OP];
  static_assert(o.np + o.no > 0, "an op has rounds");
  K = __builtin_amdgcn_readfirstlane(K);
  const Regions R{A, B, T, K};
  static_round<G, OP, 0>(op_desc<G, OP>(), sub, active, R, D, first.w);
  __syncthreads();
}
template <int G, int OP>
__device__ __forceinline__ void run_op_c(int sub, bool active, int A, int B, int D, int T, int K) {
  FirstRec<G, OP> first;
  first.load(sub);
  run_op_c<G, OP>(first, sub, active, A, B, D, T, K);
}

__device__ __forceinline__ void load_consts(int K) {
  for (int w = threadIdx.x; w < sched::NUM_CONSTS * NL; w += blockDim.x) {
    const int c = w / NL, l = w % NL;
    g_lds[K + c * SLOT + l] = sched::KCONST[c][l];
  }
}

// ---------------- Miller loop ----------------
// slots per item: F 0..11 | L1 12..17 | L2 18..23 | P 24..27 (P1.x P1.y P2.x P2.y) |
// S1 28..33 | S2 34..39 | T 40..; the constants follow the last item.  The merged line
// ops address L1 .. S2 as one region (B 0..27) and write S1 | S2 at f + 28 .. f + 39.
constexpr int ML_SLOTS = 40 + sched::MAX_PRODUCTS;

template <int G>
constexpr size_t miller_lds_bytes() {
  return (size_t)(Cfg<G>::IPW * ML_SLOTS + sched::NUM_CONSTS) * SLOT * 4;
}

template <int G, int WPE, int D>
__device__ __forceinline__ void miller_tile(int tile, const i32* __restrict__ sig, const i32* __restrict__ nhm,
                                            const i32* __restrict__ gen_lines, const i32* __restrict__ keys,
                                            const int32_t* __restrict__ key_idx, int nkeys,
                                            i32* __restrict__ status,
                                            i32* __restrict__ fout, int M) {
  constexpr int IPW = Cfg<G>::IPW, L = Cfg<G>::L;
  const int ln = threadIdx.x & 63, lid = ln / L, sub = (int)(threadIdx.x >> 6) * L + ln % L;
  const bool in_wg = lid < IPW;
  const int item = tile * IPW + lid;
  const bool valid = in_wg && item < M;
  const int it = valid ? item : 0;
  const int base = (in_wg ? lid : 0) * ML_SLOTS * SLOT;
  const int F = base, L1 = base + 12 * SLOT, L2 = base + 18 * SLOT, P = base + 24 * SLOT;
  const int S1 = base + 28 * SLOT, S2 = base + 34 * SLOT, T = base + 40 * SLOT;
  const int K = IPW * ML_SLOTS * SLOT;
  bool skip_all = !valid, skip1 = false;
  int kidx = 0;
  if (valid) {
    if (status) {
      const int32_t st = status[it];
      skip_all = st != DEC_OK && st != DEC_IDENTITY;
      skip1 = st == DEC_IDENTITY;
    }
    if (key_idx) {
      const int k = key_idx[it];
      if (k < 0 || k >= nkeys) {
        if (!skip_all && sub == 0 && status) status[it] = DEC_BAD_KEY;
        skip_all = true;
      } else {
        kidx = k;
      }
    }
  }
  const bool act = !skip_all, act1 = act && !skip1;
  if (in_wg) {
    for (int w = sub; w < 4 * NL; w += G) {
      const int e = w / NL, l = w % NL;
      const i32* src = e < 2 ? sig : nhm;
      g_lds[P + e * SLOT + l] = src[(size_t)((e & 1) * NL + l) * M + it];
    }
    for (int w = sub; w < 12 * NL; w += G) {
      const int e = w / NL, l = w % NL;
      g_lds[F + e * SLOT + l] = 0;
    }
  }
  __syncthreads();
  if (in_wg && sub == 0) lds_st(F, fp_one());
  // this item's key table (the legacy ABI's per-item tables come as M keys with
  // key_idx[i] = i), addressed as a 32-bit word offset from the uniform base: one
  // register per lane instead of a 64-bit pointer (the tables stay below 2^31 words)
  const u32 koff = (u32)kidx * (u32)KEY_TABLE_WORDS;
  __syncthreads();

  // lines of step ci into L1 | L2: generator (same for every item) and this item's key;
  // only items that use them read them (a skipped item may have no valid key table)
  auto load_lines = [&](int ci) {
    if (act) {
      for (int w = sub; w < 2 * 84; w += G) {
        const int which = w / 84, ww = w % 84, e = ww / NL, l = ww % NL;
        i32 v;
        if (which == 0) v = gen_lines[ci * 84 + ww];
        else v = keys[koff + (u32)(ci * 84 + ww)];
        g_lds[(which == 0 ? L1 : L2) + e * SLOT + l] = v;
      }
    }
  };
  // the same in two halves for the key's line (read from HBM: the generator's lines are
  // shared by every item and stay in L2), so that it is read from memory D steps before
  // it goes to LDS: lane sub holds words sub + k G (k < LW) of the 84 of each step held
  // (lv the next one to store, lv2 the one after at depth 2).  At depth 0 store_lines
  // loads the step itself.
  constexpr int LW = (84 + G - 1) / G;
  static_assert(ml_depth_ok<G>(D, WPE == 1), "key-line depth beyond the register budget");
  constexpr bool PREFETCH = D >= 1;
  auto fetch_lines = [&](int ci, i32(&v)[LW]) {
    if (PREFETCH && act) {
#pragma unroll
      for (int k = 0; k < LW; k++) {
        const int w = sub + k * G;
        if (w < 84) v[k] = keys[koff + (u32)(ci * 84 + w)];
      }
    }
  };
  auto store_lines = [&](int ci, const i32(&v)[LW]) {
    if (!PREFETCH) {
      load_lines(ci);
    } else if (act) {
      for (int w = sub; w < 84; w += G) g_lds[L1 + (w / NL) * SLOT + w % NL] = gen_lines[ci * 84 + w];
#pragma unroll
      for (int k = 0; k < LW; k++) {
        const int w = sub + k * G;
        if (w < 84) g_lds[L2 + (w / NL) * SLOT + w % NL] = v[k];
      }
    }
  };
  int ci = 0;
  if (!__syncthreads_or(act && !act1)) {
    // Every item of the tile pairs both pairs: each step multiplies f by the product of
    // its two lines (OP_LL, then OP_FL_SCALE), and the scaling of the next step's lines
    // rides along with f * LL.  Region B of both ops is the line block L1 | L2 | P | S.
    constexpr int STEPS = KEY_TABLE_WORDS / 84;
    i32 lv[LW], lv2[D == 2 ? LW : 1];
    // step `next` to LDS, and the key lines held in registers advanced by one step
    auto advance = [&](int next) {
      store_lines(next, lv);
      if constexpr (D == 1) {
        if (next + 1 < STEPS) fetch_lines(next + 1, lv);
      } else if constexpr (D == 2) {
#pragma unroll
        for (int k = 0; k < LW; k++) lv[k] = lv2[k];
        if (next + 2 < STEPS) fetch_lines(next + 2, lv2);
      }
    };
    fetch_lines(0, lv);
    if constexpr (D == 2) fetch_lines(1, lv2);
    advance(0);
    __syncthreads();
    run_op<G, WPE>(sched::OP_SCALE2, sub, act, L1, P, S1, T, K);
    // L1 | L2 are free once S holds their scaled form: the raw lines of the next step
    // go to LDS before an op that ends in a barrier, ahead of the op that scales them,
    // and the lines of the steps after are read from memory meanwhile
    advance(1);
    run_op_c<G, sched::OP_LL>(sub, act, 0, L1, S1, T, K);
    // For b = 61 .. 0: a line step, a second one if bit b of |x| is set, the square of f;
    // then the last line step.  A line step multiplies f by LL, the product of its two
    // lines, and scales the next step's lines (OP_FL_SCALE); the next LL is formed with
    // the square of f when one follows (OP_SQR_LL), else on its own.  One loop, so that
    // each op's straight-line code is emitted once.
    bool second = false;
    FirstRec<G, sched::OP_FL_SCALE> fl0;  // first-round records, the same every step
    FirstRec<G, sched::OP_SQR_LL> sq0;
    fl0.load(sub);
    sq0.load(sub);
    for (int b = 61;;) {
      if (b < 0) {
        run_op<G, WPE>(sched::OP_FL, sub, act, F, L1, F, T, K);
        break;
      }
      run_op_c<G, sched::OP_FL_SCALE>(fl0, sub, act, F, L1, F, T, K);
      ci++;
      if (ci + 1 < STEPS) advance(ci + 1);
      if (!second && miller_bit(b)) {
        second = true;
        run_op_c<G, sched::OP_LL>(sub, act, 0, L1, S1, T, K);
        continue;
      }
      second = false;
      run_op_c<G, sched::OP_SQR_LL>(sq0, sub, act, F, L1, F, T, K);
      b--;
    }
  } else {
    // a tile with an identity signature: its pair-1 lines are skipped item by item
    auto step = [&]() {
      load_lines(ci++);
      __syncthreads();
      run_op<G, WPE>(sched::OP_SCALE2, sub, act, L1, P, S1, T, K);  // S1 | S2 (both pairs)
      run_op<G, WPE>(sched::OP_ELL, sub, act1, F, S1, F, T, K);
      run_op<G, WPE>(sched::OP_ELL, sub, act, F, S2, F, T, K);
    };
    for (int b = 61; b >= 0; b--) {
      step();
      if (miller_bit(b)) step();
      run_op<G, WPE>(sched::OP_SQR, sub, act, F, 0, F, T, K);
    }
    step();
  }
  run_op<G, WPE>(sched::OP_CONJ, sub, act, F, 0, F, T, K);
  if (valid && act) {
    for (int w = sub; w < 12 * NL; w += G) {
      const int e = w / NL, l = w % NL;
      fout[(size_t)item * (12 * NL) + w] = g_lds[F + e * SLOT + l];
    }
  }
  __syncthreads();  // the next tile reuses this tile's LDS
}

// Next tile of a workgroup: blockIdx.x alone (one tile per workgroup), or, with a work
// queue, tiles popped from a device counter until every tile is taken.  Every wave of
// the workgroup leaves at the same tile index, so the barriers inside stay uniform.
__device__ __forceinline__ int next_tile(int* queue, int round) {
  __shared__ int s_tile;
  if (!queue) return round == 0 ? (int)blockIdx.x : INT_MAX;
  if (threadIdx.x == 0) s_tile = atomicAdd(queue, 1);
  __syncthreads();
  const int t = s_tile;
  __syncthreads();
  return t;
}

// A workgroup past a launch's tiles: thread t tests item k * per + t of the batch for
// membership of G1 (the subgroup test needs only the decoded signature, and no pairing
// value needs it).  These workgroups run at the default wave priority on the SIMDs the
// tiles leave free.
// Lanes with no item of their own repeat the workgroup's last one and store nothing, so
// that every lane runs the test (a wavefront with few active lanes issues slower).
__device__ __forceinline__ void subgroup_rider(const i32* pts, const i32* status, i32* out, int M, int per, int k) {
  const int t = (int)threadIdx.x;
  const int i = k * per + t;
  const bool live = t < per && i < M;
  const int last = (k * per + per < M ? k * per + per : M) - 1;
  const int32_t s = g1_subgroup_status(pts, status, M, live ? i : last);
  if (live) out[i] = s;
}

template <int G, bool RIDE, int D>
__global__ __launch_bounds__(Cfg<G>::THREADS, RIDE ? 1 : 2) void k_miller_grp(const i32* __restrict__ sig, const i32* __restrict__ nhm,
                                                   const i32* __restrict__ gen_lines, const i32* __restrict__ keys,
                                                   const int32_t* __restrict__ key_idx, int nkeys,
                                                   i32* __restrict__ status,
                                                   i32* __restrict__ fout, int M, int* queue, SubgroupRiders riders) {
  constexpr int IPW = Cfg<G>::IPW;
  const int ntiles = (M + IPW - 1) / IPW;
  const int K = IPW * ML_SLOTS * SLOT;
  if constexpr (RIDE) {
    if (!queue && (int)blockIdx.x >= ntiles) {
      subgroup_rider(riders.pts, riders.status, riders.out, M, riders.per, (int)blockIdx.x - ntiles);
      return;
    }
  }
  // raised wave priority: a concurrent subgroup-test wave on the same SIMD issues in the
  // cycles this wave leaves idle instead of halving its issue rate
  __builtin_amdgcn_s_setprio(2);
  load_consts(K);
  for (int round = 0;; round++) {
    const int tile = next_tile(queue, round);
    if (tile >= ntiles) break;
    miller_tile<G, RIDE ? 1 : 2, D>(tile, sig, nhm, gen_lines, keys, key_idx, nkeys, status, fout, M);
  }
}

// ---------------- final exponentiation ----------------
// slots per item: R0..R5 (12 each) 0..71 | T 72..; then the constants, then the per-item
// identity flags (12 words per item).
constexpr int FE_SLOTS = 72 + sched::MAX_PRODUCTS;

template <int G>
constexpr size_t final_exp_lds_bytes() {
  return ((size_t)(Cfg<G>::IPW * FE_SLOTS + sched::NUM_CONSTS) * SLOT + Cfg<G>::IPW * 12) * 4;
}

// dst = src^|x| conjugated (= src^x on the cyclotomic subgroup).  Not inlined: one copy
// of the straight-line cyclotomic square and product serves its five call sites.
template <int G, int WPE>
__device__ __noinline__ void cyc_exp_grp(int sub, bool act, int src, int dst, int T, int K) {
  run_op<G, WPE>(sched::OP_COPY, sub, act, src, 0, dst, T, K);
  FirstRec<G, sched::OP_CYC_SQR> sq0;  // the same every iteration
  sq0.load(sub);
  // (the product's first record is read at each product: held across the loop it made
  // the 12-lane kernel 7 % slower, 1.48 -> 1.58 ms per 4,096 items)
  for (int i = 62; i >= 0; i--) {
    run_op_c<G, sched::OP_CYC_SQR>(sq0, sub, act, dst, 0, dst, T, K);
    if ((BLS_X >> i) & 1) run_op_c<G, sched::OP_MUL>(sub, act, dst, src, dst, T, K);
  }
  run_op<G, WPE>(sched::OP_CONJ, sub, act, dst, 0, dst, T, K);
}

template <int G, int WPE>
__device__ __forceinline__ void final_exp_tile(int tile, const i32* __restrict__ fin, const i32* __restrict__ status,
                                               const i32* __restrict__ sub_status, int32_t* __restrict__ result,
                                               u64* __restrict__ gt_out, int M) {
  constexpr int IPW = Cfg<G>::IPW, L = Cfg<G>::L;
  const int ln = threadIdx.x & 63, lid = ln / L, sub = (int)(threadIdx.x >> 6) * L + ln % L;
  const bool in_wg = lid < IPW;
  const int item = tile * IPW + lid;
  const bool valid = in_wg && item < M;
  const int it = valid ? item : 0;
  const int base = (in_wg ? lid : 0) * FE_SLOTS * SLOT;
  const int R0 = base, R1 = base + 12 * SLOT, R2 = base + 24 * SLOT, R3 = base + 36 * SLOT;
  const int R4 = base + 48 * SLOT, R5 = base + 60 * SLOT, T = base + 72 * SLOT;
  const int K = IPW * FE_SLOTS * SLOT;
  const int FLAGS = K + sched::NUM_CONSTS * SLOT + (in_wg ? lid : 0) * 12;
  int32_t st = DEC_OK;
  if (valid && status) st = status[it];
  if (valid && sub_status && st == DEC_OK) st = sub_status[it];
  const bool act = valid && (st == DEC_OK || st == DEC_IDENTITY);
  if (in_wg) {
    for (int w = sub; w < 12 * NL; w += G) {
      const int e = w / NL, l = w % NL;
      g_lds[R0 + e * SLOT + l] = act ? fin[(size_t)it * (12 * NL) + w] : 0;
    }
    // The shared inversion below reads the norm slot R3[2] of every item of the tile,
    // and the ops store nothing for an inactive item: its slot would hold whatever an
    // earlier workgroup or kernel left in this LDS (words outside the multiplier's input
    // range corrupt the product, and so every item's inverse).  Zero counts as 1 there.
    if (!act) {
      for (int l = sub; l < NL; l += G) g_lds[R3 + 2 * SLOT + l] = 0;
    }
  }
  __syncthreads();

  // easy part: f^((p^6-1)(p^2+1)) with f^-1 from the norm chain Fp12 -> Fp6 -> Fp2 -> Fp
  run_op<G, WPE>(sched::OP_N12, sub, act, R0, 0, R1, T, K);            // R1[0..5] = c0^2 - v c1^2
  run_op<G, WPE>(sched::OP_N6A, sub, act, R1, 0, R2, T, K);            // R2[0..5] = (t0, t1, t2)
  run_op<G, WPE>(sched::OP_N6B, sub, act, R1, R2, R3, T, K);           // R3[0..1] = sc
  run_op<G, WPE>(sched::OP_N2, sub, act, R3, 0, R3 + 2 * SLOT, T, K);  // R3[2] = n
  // One inversion per workgroup (Montgomery's trick over its IPW items), on every lane:
  // code that a handful of lanes of a wavefront run alone issues about three times slower
  // per instruction on this GPU (a dependent chain of products measured 6,020 cycles per
  // product with 8 active lanes or fewer, 2,018 with 16 or more), so no step below runs
  // on fewer than half a wavefront.
  //  (a) the lanes of each item test its n for zero (inactive items; a zero counts as 1
  //      in the products and gets a zero inverse);
  //  (b) lanes 0..31 of each wave form the prefix products P_q = n_0 .. n_(q-1) into
  //      R3[6] of item q, lanes 32..63 the suffix products S_q = n_(q+1) .. n_(IPW-1)
  //      into R3[7], one product per step on both halves;
  //  (c) every lane inverts the product of all n by batched divsteps (inv_fp_uniform:
  //      the values are the same on every lane, both halves' products canonicalised);
  //  (d) item q's inverse inv * P_q * S_q goes to R3[3].
  {
    const Fp n_own = lds_ld(R3 + 2 * SLOT);
    const bool z = is_zero(n_own);
    __syncthreads();
    if (in_wg && sub == 0 && z) lds_st(R3 + 2 * SLOT, fp_one());
    __syncthreads();
    const bool pre = (threadIdx.x & 63) < 32;
    Fp acc = fp_one();
    for (int k = 0; k < IPW; k++) {
      const int bq = (pre ? k : IPW - 1 - k) * FE_SLOTS * SLOT + 36 * SLOT;
      if ((threadIdx.x & 31) == 0 && threadIdx.x < 64) lds_st(bq + (pre ? 6 : 7) * SLOT, acc);
      acc = mul(acc, lds_ld(bq + 2 * SLOT));
    }
    const Fp inv_all = inv_fp_uniform(acc);
    __syncthreads();
    const Fp inv_own = mul(mul(inv_all, lds_ld(R3 + 6 * SLOT)), lds_ld(R3 + 7 * SLOT));
    if (in_wg && sub == 0) lds_st(R3 + 3 * SLOT, z ? fp_zero() : inv_own);
  }
  __syncthreads();
  run_op<G, WPE>(sched::OP_INV2, sub, act, R3, R3 + 3 * SLOT, R3 + 4 * SLOT, T, K);  // R3[4..5] = sc^-1
  run_op<G, WPE>(sched::OP_INV6, sub, act, R2, R3 + 4 * SLOT, R4, T, K);             // R4[0..5] = t^-1
  run_op<G, WPE>(sched::OP_INV12, sub, act, R0, R4, R5, T, K);                       // R5 = f^-1
  run_op<G, WPE>(sched::OP_MUL_CONJA, sub, act, R0, R5, R1, T, K);                   // R1 = f^(p^6-1)
  run_op<G, WPE>(sched::OP_FROB2, sub, act, R1, 0, R2, T, K);
  run_op<G, WPE>(sched::OP_MUL, sub, act, R2, R1, R2, T, K);                         // R2 = t2

  // hard part (api.cu:68-77).  Registers: F=R0 T1=R1 T2=R2 T3=R3 T4=R4 T0=R5
  run_op<G, WPE>(sched::OP_COPY, sub, act, R2, 0, R0, T, K);       // f = t2
  run_op<G, WPE>(sched::OP_CYC_SQR, sub, act, R2, 0, R1, T, K);
  run_op<G, WPE>(sched::OP_CONJ, sub, act, R1, 0, R1, T, K);       // t1 = conj(cyc_sqr(t2))
  cyc_exp_grp<G, WPE>(sub, act, R2, R3, T, K);                     // t3 = t2^x
  run_op<G, WPE>(sched::OP_MUL, sub, act, R1, R3, R4, T, K);       // t4 = t1 t3
  cyc_exp_grp<G, WPE>(sub, act, R4, R1, T, K);                     // t1 = t4^x
  run_op<G, WPE>(sched::OP_CONJ, sub, act, R4, 0, R4, T, K);       // t4 = conj(t4)
  run_op<G, WPE>(sched::OP_MUL, sub, act, R0, R4, R0, T, K);       // f *= t4
  run_op<G, WPE>(sched::OP_CYC_SQR, sub, act, R3, 0, R4, T, K);    // t4 = cyc_sqr(t3)
  cyc_exp_grp<G, WPE>(sub, act, R1, R5, T, K);                     // t0 = t1^x
  run_op<G, WPE>(sched::OP_MUL, sub, act, R3, R5, R3, T, K);       // t3 *= t0
  run_op<G, WPE>(sched::OP_FROB2, sub, act, R3, 0, R3, T, K);      // t3 = frob^2(t3)
  run_op<G, WPE>(sched::OP_MUL, sub, act, R0, R3, R0, T, K);       // f *= t3
  cyc_exp_grp<G, WPE>(sub, act, R5, R3, T, K);                     // tmp = t0^x
  run_op<G, WPE>(sched::OP_MUL, sub, act, R4, R3, R4, T, K);       // t4 *= tmp
  cyc_exp_grp<G, WPE>(sub, act, R4, R3, T, K);                     // tmp = t4^x
  run_op<G, WPE>(sched::OP_MUL, sub, act, R0, R3, R0, T, K);       // f *= tmp
  run_op<G, WPE>(sched::OP_CONJ, sub, act, R2, 0, R5, T, K);       // conj(t2)
  run_op<G, WPE>(sched::OP_MUL, sub, act, R4, R5, R4, T, K);       // t4 *= conj(t2)
  run_op<G, WPE>(sched::OP_MUL, sub, act, R2, R1, R2, T, K);       // t2 *= t1
  run_op<G, WPE>(sched::OP_FROB3, sub, act, R2, 0, R2, T, K);      // t2 = frob^3(t2)
  run_op<G, WPE>(sched::OP_MUL, sub, act, R0, R2, R0, T, K);       // f *= t2
  run_op<G, WPE>(sched::OP_FROB1, sub, act, R4, 0, R4, T, K);      // t4 = frob(t4)
  run_op<G, WPE>(sched::OP_MUL, sub, act, R0, R4, R0, T, K);       // f *= t4

  // identity test and export: component c handled by lane c mod G
  if (in_wg) {
    for (int c = sub; c < 12; c += G) {
      int ok = 1;
      if (act) {
        const Fp v = canon(lds_ld(R0 + c * SLOT));
        const Fp want = c == 0 ? canon(fp_one()) : fp_zero();
        for (int l = 0; l < NL; l++) ok &= v.l[l] == want.l[l];
        if (gt_out) st_ref(gt_out + (size_t)it * 72 + c * 6, lds_ld(R0 + c * SLOT));
      } else if (valid && gt_out) {
        for (int j = 0; j < 6; j++) gt_out[(size_t)it * 72 + c * 6 + j] = 0;
      }
      g_lds[FLAGS + c] = ok;
    }
  }
  __syncthreads();
  if (valid && sub == 0) {
    int all = 1;
    for (int k = 0; k < 12; k++) all &= g_lds[FLAGS + k];
    result[it] = act ? all : st;
  }
  __syncthreads();  // the next tile reuses this tile's LDS
}

template <int G, bool RIDE>
__global__ __launch_bounds__(Cfg<G>::THREADS, RIDE ? 1 : 2) void k_final_exp_grp(const i32* __restrict__ fin, const i32* __restrict__ status,
                                                      const i32* __restrict__ sub_status, int32_t* __restrict__ result,
                                                      u64* __restrict__ gt_out, int M, int* queue, SubgroupRiders riders) {
  constexpr int IPW = Cfg<G>::IPW;
  const int ntiles = (M + IPW - 1) / IPW;
  const int K = IPW * FE_SLOTS * SLOT;
  if constexpr (RIDE) {
    if (!queue && (int)blockIdx.x >= ntiles) {
      subgroup_rider(riders.pts, riders.status, riders.out, M, riders.per, (int)blockIdx.x - ntiles);
      return;
    }
  }
  __builtin_amdgcn_s_setprio(2);
  load_consts(K);
  for (int round = 0;; round++) {
    const int tile = next_tile(queue, round);
    if (tile >= ntiles) break;
    final_exp_tile<G, RIDE ? 1 : 2>(tile, fin, status, sub_status, result, gt_out, M);
  }
}

// ---------------- final exponentiation in compact LDS ----------------
// (a translation unit of its own, GRP_KIND 2: next to the full-LDS kernel its code moved
// that kernel's code and made it 8 % slower, 1.45 -> 1.57 ms per 4,096 items)
#if GRP_KIND == 2
// The same exponentiation with 66 slots per item instead of 134, so that twice as many
// workgroups fit a CU (two wavefronts per SIMD at 12 lanes: 8 workgroups of 19.5 KB).
// Per item: two Fp12 areas X | Y (0..11, 12..23) and a scratch region S (24..65) that
// holds the products of an op, or, in an Fp12 product, the first half s = (a0 + a1)
// (b0 + b1) in its first 6 slots (OP_MUL_S) and the products of OP_MUL_PQ after them.
// The values of the hard part that the current op does not touch wait in device memory:
// each item has six homes of 168 words in PARK (its workgroup's slot of the persistent
// grid, blockIdx.x * IPW + item), and lane s of the item moves component s, so every
// word goes back to the lane that stored it.
constexpr int FEC_SLOTS = 66;
constexpr int FEC_X = 0, FEC_Y = 12 * SLOT, FEC_S = 24 * SLOT;
constexpr int FEC_HOMES = 6, FEC_HOME_WORDS = 12 * NL;  // homes per item, words per home
enum { H_F = 0, H_T0, H_T1, H_T2, H_T3, H_T4 };

template <int G>
constexpr size_t final_exp_compact_lds_bytes() {
  return ((size_t)(Cfg<G>::IPW * FEC_SLOTS + sched::NUM_CONSTS) * SLOT + Cfg<G>::IPW * 12) * 4;
}

// a *= b, the product split in two ops through S: straight-line (in the cyclotomic
// powers) or through the interpreter (the glue between them, run once per tile)
template <int G>
__device__ __forceinline__ void mul_compact(int sub, bool act, int a, int b, int S, int K) {
  run_op_c<G, sched::OP_MUL_S>(sub, act, a, b, S, S + 6 * SLOT, K);
  run_op_c<G, sched::OP_MUL_PQ>(sub, act, a, b, a, S, K);
}
template <int G, int WPE>
__device__ __forceinline__ void mul_compact_i(int sub, bool act, int a, int b, int S, int K) {
  run_op<G, WPE>(sched::OP_MUL_S, sub, act, a, b, S, S + 6 * SLOT, K);
  run_op<G, WPE>(sched::OP_MUL_PQ, sub, act, a, b, a, S, K);
}

// dst = src^|x| conjugated, as cyc_exp_grp, with the compact product
template <int G, int WPE>
__device__ __noinline__ void cyc_exp_compact(int sub, bool act, int src, int dst, int S, int K) {
  run_op<G, WPE>(sched::OP_COPY, sub, act, src, 0, dst, S, K);
  FirstRec<G, sched::OP_CYC_SQR> sq0;
  sq0.load(sub);
  for (int i = 62; i >= 0; i--) {
    run_op_c<G, sched::OP_CYC_SQR>(sq0, sub, act, dst, 0, dst, S, K);
    if ((BLS_X >> i) & 1) mul_compact<G>(sub, act, dst, src, S, K);
  }
  run_op<G, WPE>(sched::OP_CONJ, sub, act, dst, 0, dst, S, K);
}

template <int G, int WPE>
__device__ __forceinline__ void final_exp_tile_compact(int tile, const i32* __restrict__ fin, const i32* __restrict__ status,
                                                       const i32* __restrict__ sub_status, int32_t* __restrict__ result,
                                                       u64* __restrict__ gt_out, i32* __restrict__ park, int M) {
  constexpr int IPW = Cfg<G>::IPW, L = Cfg<G>::L;
  static_assert(G == L && G == 12, "one wavefront per item group, a lane per Fp12 component");
  const int ln = threadIdx.x & 63, lid = ln / L, sub = ln % L;
  const bool in_wg = lid < IPW;
  const int item = tile * IPW + lid;
  const bool valid = in_wg && item < M;
  const int it = valid ? item : 0;
  const int base = (in_wg ? lid : 0) * FEC_SLOTS * SLOT;
  const int X = base + FEC_X, Y = base + FEC_Y, S = base + FEC_S;
  const int K = IPW * FEC_SLOTS * SLOT;
  const int FLAGS = K + sched::NUM_CONSTS * SLOT + (in_wg ? lid : 0) * 12;
  // this lane's word of home h: the item's slot in the grid, then component sub
  i32* const home = park + ((size_t)blockIdx.x * IPW + (in_wg ? lid : 0)) * FEC_HOMES * FEC_HOME_WORDS + sub * NL;
  auto store_home = [&](int area, int h) {
    if (in_wg) {
#pragma unroll
      for (int l = 0; l < NL; l++) home[h * FEC_HOME_WORDS + l] = g_lds[area + sub * SLOT + l];
    }
  };
  auto load_home = [&](int h, int area) {
    if (in_wg) {
#pragma unroll
      for (int l = 0; l < NL; l++) g_lds[area + sub * SLOT + l] = home[h * FEC_HOME_WORDS + l];
    }
    __syncthreads();
  };
  int32_t st = DEC_OK;
  if (valid && status) st = status[it];
  if (valid && sub_status && st == DEC_OK) st = sub_status[it];
  const bool act = valid && (st == DEC_OK || st == DEC_IDENTITY);
  if (in_wg) {
    for (int w = sub; w < 12 * NL; w += G) {
      const int e = w / NL, l = w % NL;
      g_lds[X + e * SLOT + l] = act ? fin[(size_t)it * (12 * NL) + w] : 0;
    }
  }
  __syncthreads();

  // easy part (as final_exp_tile): f in X; the norm chain in Y and X; f waits at home
  run_op<G, WPE>(sched::OP_N12, sub, act, X, 0, Y, S, K);                 // Y[0..5] = c0^2 - v c1^2
  store_home(X, H_F);
  run_op<G, WPE>(sched::OP_N6A, sub, act, Y, 0, Y + 6 * SLOT, S, K);     // Y[6..11] = (t0, t1, t2)
  run_op<G, WPE>(sched::OP_N6B, sub, act, Y, Y + 6 * SLOT, X, S, K);     // X[0..1] = sc
  run_op<G, WPE>(sched::OP_N2, sub, act, X, 0, X + 2 * SLOT, S, K);      // X[2] = n
  // one inversion per workgroup, as in final_exp_tile, on X[2] (an inactive item's X
  // holds the zeros it was loaded with: its n counts as 1 and gets a zero inverse)
  {
    const Fp n_own = lds_ld(X + 2 * SLOT);
    const bool z = is_zero(n_own);
    __syncthreads();
    if (in_wg && sub == 0 && z) lds_st(X + 2 * SLOT, fp_one());
    __syncthreads();
    const bool pre = (threadIdx.x & 63) < 32;
    Fp acc = fp_one();
    for (int k = 0; k < IPW; k++) {
      const int bq = (pre ? k : IPW - 1 - k) * FEC_SLOTS * SLOT + FEC_X;
      if ((threadIdx.x & 31) == 0 && threadIdx.x < 64) lds_st(bq + (pre ? 6 : 7) * SLOT, acc);
      acc = mul(acc, lds_ld(bq + 2 * SLOT));
    }
    const Fp inv_all = inv_fp_uniform(acc);
    __syncthreads();
    const Fp inv_own = mul(mul(inv_all, lds_ld(X + 6 * SLOT)), lds_ld(X + 7 * SLOT));
    if (in_wg && sub == 0) lds_st(X + 3 * SLOT, z ? fp_zero() : inv_own);
  }
  __syncthreads();
  run_op<G, WPE>(sched::OP_INV2, sub, act, X, X + 3 * SLOT, X + 4 * SLOT, S, K);  // X[4..5] = sc^-1
  run_op<G, WPE>(sched::OP_INV6, sub, act, Y + 6 * SLOT, X + 4 * SLOT, Y, S, K);  // Y[0..5] = t^-1
  load_home(H_F, X);
  run_op<G, WPE>(sched::OP_INV12, sub, act, X, Y, X, S, K);                      // X = f^-1
  load_home(H_F, Y);
  run_op<G, WPE>(sched::OP_CONJ, sub, act, Y, 0, Y, S, K);
  mul_compact_i<G, WPE>(sub, act, Y, X, S, K);                                          // Y = f^(p^6-1)
  run_op<G, WPE>(sched::OP_FROB2, sub, act, Y, 0, X, S, K);
  mul_compact_i<G, WPE>(sub, act, X, Y, S, K);                                          // X = t2

  // hard part (api.cu:68-77), the same sequence as final_exp_tile with the values
  // f t0 t1 t2 t3 t4 moved between X, Y and their homes
  store_home(X, H_F);                                                            // f = t2
  run_op<G, WPE>(sched::OP_CYC_SQR, sub, act, X, 0, Y, S, K);
  run_op<G, WPE>(sched::OP_CONJ, sub, act, Y, 0, Y, S, K);                       // Y = t1
  store_home(Y, H_T1);
  cyc_exp_compact<G, WPE>(sub, act, X, Y, S, K);                                 // Y = t3 = t2^x
  store_home(X, H_T2);
  load_home(H_T1, X);
  mul_compact_i<G, WPE>(sub, act, X, Y, S, K);                                          // X = t4 = t1 t3
  store_home(Y, H_T3);
  cyc_exp_compact<G, WPE>(sub, act, X, Y, S, K);                                 // Y = t1 = t4^x
  run_op<G, WPE>(sched::OP_CONJ, sub, act, X, 0, X, S, K);                       // X = conj(t4)
  store_home(Y, H_T1);
  load_home(H_F, Y);
  mul_compact_i<G, WPE>(sub, act, Y, X, S, K);                                          // Y = f *= t4
  store_home(Y, H_F);
  load_home(H_T3, X);
  run_op<G, WPE>(sched::OP_CYC_SQR, sub, act, X, 0, Y, S, K);                    // Y = t4 = cyc_sqr(t3)
  store_home(Y, H_T4);
  load_home(H_T1, X);
  cyc_exp_compact<G, WPE>(sub, act, X, Y, S, K);                                 // Y = t0 = t1^x
  load_home(H_T3, X);
  mul_compact_i<G, WPE>(sub, act, X, Y, S, K);                                          // X = t3 *= t0
  run_op<G, WPE>(sched::OP_FROB2, sub, act, X, 0, X, S, K);                      // X = t3 = frob^2(t3)
  store_home(Y, H_T0);
  load_home(H_F, Y);
  mul_compact_i<G, WPE>(sub, act, Y, X, S, K);                                          // Y = f *= t3
  store_home(Y, H_F);
  load_home(H_T0, X);
  cyc_exp_compact<G, WPE>(sub, act, X, Y, S, K);                                 // Y = tmp = t0^x
  load_home(H_T4, X);
  mul_compact_i<G, WPE>(sub, act, X, Y, S, K);                                          // X = t4 *= tmp
  cyc_exp_compact<G, WPE>(sub, act, X, Y, S, K);                                 // Y = tmp = t4^x
  store_home(X, H_T4);
  load_home(H_F, X);
  mul_compact_i<G, WPE>(sub, act, X, Y, S, K);                                          // X = f *= tmp
  store_home(X, H_F);
  load_home(H_T4, X);
  load_home(H_T2, Y);
  run_op<G, WPE>(sched::OP_CONJ, sub, act, Y, 0, Y, S, K);                       // Y = conj(t2)
  mul_compact_i<G, WPE>(sub, act, X, Y, S, K);                                          // X = t4 *= conj(t2)
  run_op<G, WPE>(sched::OP_CONJ, sub, act, Y, 0, Y, S, K);                       // Y = t2
  store_home(X, H_T4);
  load_home(H_T1, X);
  mul_compact_i<G, WPE>(sub, act, Y, X, S, K);                                          // Y = t2 *= t1
  run_op<G, WPE>(sched::OP_FROB3, sub, act, Y, 0, Y, S, K);                      // Y = t2 = frob^3(t2)
  load_home(H_F, X);
  mul_compact_i<G, WPE>(sub, act, X, Y, S, K);                                          // X = f *= t2
  load_home(H_T4, Y);
  run_op<G, WPE>(sched::OP_FROB1, sub, act, Y, 0, Y, S, K);                      // Y = t4 = frob(t4)
  mul_compact_i<G, WPE>(sub, act, X, Y, S, K);                                          // X = f *= t4

  // identity test and export, as final_exp_tile
  if (in_wg) {
    for (int c = sub; c < 12; c += G) {
      int ok = 1;
      if (act) {
        const Fp v = canon(lds_ld(X + c * SLOT));
        const Fp want = c == 0 ? canon(fp_one()) : fp_zero();
        for (int l = 0; l < NL; l++) ok &= v.l[l] == want.l[l];
        if (gt_out) st_ref(gt_out + (size_t)it * 72 + c * 6, lds_ld(X + c * SLOT));
      } else if (valid && gt_out) {
        for (int j = 0; j < 6; j++) gt_out[(size_t)it * 72 + c * 6 + j] = 0;
      }
      g_lds[FLAGS + c] = ok;
    }
  }
  __syncthreads();
  if (valid && sub == 0) {
    int all = 1;
    for (int k = 0; k < 12; k++) all &= g_lds[FLAGS + k];
    result[it] = act ? all : st;
  }
  __syncthreads();  // the next tile reuses this tile's LDS
}

// queue-fed persistent grid only (its homes are indexed by workgroup), two waves per SIMD
template <int G>
__global__ __launch_bounds__(Cfg<G>::THREADS, 2) void k_final_exp_compact(const i32* __restrict__ fin, const i32* __restrict__ status,
                                                           const i32* __restrict__ sub_status, int32_t* __restrict__ result,
                                                           u64* __restrict__ gt_out, int M, int* queue, i32* __restrict__ park) {
  constexpr int IPW = Cfg<G>::IPW;
  const int ntiles = (M + IPW - 1) / IPW;
  const int K = IPW * FEC_SLOTS * SLOT;
  __builtin_amdgcn_s_setprio(2);
  load_consts(K);
  for (int round = 0;; round++) {
    const int tile = next_tile(queue, round);
    if (tile >= ntiles) break;
    final_exp_tile_compact<G, 2>(tile, fin, status, sub_status, result, gt_out, park, M);
  }
}
#endif  // GRP_KIND == 2

// ---------------- launch wrappers ----------------
// Resident workgroups per CU of each kernel instantiation, queried once (callers hold
// the library lock) so that no occupancy query happens during graph capture.
template <int G>
int grp_occ_ml() {
  static int occ = 0;
  if (!occ && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_miller_grp<G, false, ml_depth_for<G>(false, -1)>, Cfg<G>::THREADS, miller_lds_bytes<G>()) != hipSuccess || occ <= 0)) occ = 1;
  return occ;
}
template <int G>
int grp_occ_fe() {
  static int occ = 0;
  if (!occ && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_final_exp_grp<G, false>, Cfg<G>::THREADS, final_exp_lds_bytes<G>()) != hipSuccess || occ <= 0)) occ = 1;
  return occ;
}
template <int G>
int grp_occ_fe_compact() {
  if constexpr (G == 12) return fe_compact_occ();
  else return 0;
}
// Grid of a queue-fed launch: at most wg_per_cu workgroups per CU (0 = as many as
// fit), never more than the occupancy limit, never more than there are tiles.
static int queue_grid(int occ, int ntiles, int wg_per_cu) {
  const int per = wg_per_cu > 0 && wg_per_cu < occ ? wg_per_cu : occ;
  const long g = (long)grp_num_cus() * per;
  return (int)(g < ntiles ? g : ntiles);
}

// Whether a launch puts at most one wave on each SIMD when its waves are spread evenly:
// one workgroup per tile (M up to a few thousand items), or a persistent grid of at most
// four single-wave workgroups per CU.
static bool at_most_one_wave_per_simd(int* queue, int occ, int wg_per_cu, int waves_per_wg) {
  if (!queue) return true;
  const int per = wg_per_cu > 0 && wg_per_cu < occ ? wg_per_cu : occ;
  return per * waves_per_wg <= 4;
}

// grid of a launch: the persistent grid with a queue, else one workgroup per tile plus
// the subgroup-test workgroups (riders with per <= threads of a workgroup, no queue)
static int launch_grid(int occ, int ntiles, int threads, int* queue, int wg_per_cu, const SubgroupRiders& r, int M,
                       int* grid) {
  if (queue) {
    if (r.per) return -1;
    *grid = queue_grid(occ, ntiles, wg_per_cu);
    return 0;
  }
  if (r.per < 0 || r.per > threads || (r.per && (!r.pts || !r.status || !r.out))) return -1;
  *grid = ntiles + (r.per ? (M + r.per - 1) / r.per : 0);
  return 0;
}

// one instantiation of the Miller kernel, at the depth the launch resolved
template <int G, bool RIDE, int D>
static int launch_ml_kernel(int grid, const int32_t* sig, const int32_t* nhm, const int32_t* gen, const int32_t* keys,
                            const int32_t* key_idx, int nkeys, int32_t* status, int32_t* fout,
                            int M, int* queue, const SubgroupRiders& riders, hipStream_t s) {
  hipLaunchKernelGGL((k_miller_grp<G, RIDE, D>), dim3(grid), dim3(Cfg<G>::THREADS), miller_lds_bytes<G>(), s, sig, nhm, gen,
                     keys, key_idx, nkeys, status, fout, M, queue, riders);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
template <int G, bool RIDE>
static int launch_ml_depth(int depth, int grid, const int32_t* sig, const int32_t* nhm, const int32_t* gen,
                           const int32_t* keys, const int32_t* key_idx, int nkeys,
                           int32_t* status, int32_t* fout, int M, int* queue, const SubgroupRiders& riders,
                           hipStream_t s) {
  constexpr bool one = RIDE;
  switch (depth) {
    case 2:
      if constexpr (ml_depth_ok<G>(2, one))
        return launch_ml_kernel<G, RIDE, 2>(grid, sig, nhm, gen, keys, key_idx, nkeys, status, fout, M, queue,
                                            riders, s);
      return -1;
    case 1:
      if constexpr (ml_depth_ok<G>(1, one))
        return launch_ml_kernel<G, RIDE, 1>(grid, sig, nhm, gen, keys, key_idx, nkeys, status, fout, M, queue,
                                            riders, s);
      return -1;
    default:
      return launch_ml_kernel<G, RIDE, 0>(grid, sig, nhm, gen, keys, key_idx, nkeys, status, fout, M, queue,
                                          riders, s);
  }
}
// whether a launch of M items runs the one-wave instantiation: with riders, or when its
// waves (one workgroup per tile, or the persistent grid) fit one on each SIMD
template <int G>
static bool ml_one_wave(int* queue, int occ, int wg_per_cu, bool riders, int M) {
  if constexpr (one_wave_inst<G>(0)) {
    if (riders) return true;
    if (!queue) return (long)((M + Cfg<G>::IPW - 1) / Cfg<G>::IPW) * Cfg<G>::W <= 4L * grp_num_cus();
    return at_most_one_wave_per_simd(queue, occ, wg_per_cu, Cfg<G>::W);
  }
  return false;
}
template <int G>
int grp_ml_depth(bool queue, int wg_per_cu, bool riders, int depth, int M) {
  int q = 0;
  return ml_depth_for<G>(ml_one_wave<G>(queue ? &q : nullptr, queue ? grp_occ_ml<G>() : 0, wg_per_cu, riders, M),
                         depth);
}

template <int G>
int grp_launch_ml(const int32_t* sig, const int32_t* nhm, const int32_t* gen, const int32_t* keys,
                  const int32_t* key_idx, int nkeys, int32_t* status, int32_t* fout, int M,
                  int* queue, int wg_per_cu, int depth, const SubgroupRiders& riders, hipStream_t s) {
  constexpr int IPW = Cfg<G>::IPW;
  const int ntiles = (M + IPW - 1) / IPW;
  int grid = 0;
  if (riders.per && !rides<G>(0)) return -1;
  const int occ = queue ? grp_occ_ml<G>() : 0;
  if (launch_grid(occ, ntiles, Cfg<G>::THREADS, queue, wg_per_cu, riders, M, &grid)) return -1;
  // At most one wave per SIMD wanted: the instantiation compiled for one wave per SIMD
  // where there is one (its registers exceed half a SIMD's, so no SIMD takes two of its
  // waves while another SIMD of the CU is empty; with the two-wave instantiation the
  // dispatcher doubles waves up, which measured 2.47 instead of 1.47 ms for the 12-lane
  // final exponentiation of 4096 items).
  const bool one = ml_one_wave<G>(queue, occ, wg_per_cu, riders.per != 0, M);
  const int d = ml_depth_for<G>(one, depth);
  if constexpr (one_wave_inst<G>(0)) {
    if (one)
      return launch_ml_depth<G, true>(d, grid, sig, nhm, gen, keys, key_idx, nkeys, status, fout, M, queue,
                                      riders, s);
  }
  return launch_ml_depth<G, false>(d, grid, sig, nhm, gen, keys, key_idx, nkeys, status, fout, M, queue,
                                   riders, s);
}
template <int G>
int grp_launch_fe(const int32_t* fin, const int32_t* status, const int32_t* sub_status, int32_t* result, uint64_t* gt,
                  int M, int* queue, int wg_per_cu, const SubgroupRiders& riders, int32_t* park, long park_items,
                  hipStream_t s) {
  constexpr int IPW = Cfg<G>::IPW;
  const int ntiles = (M + IPW - 1) / IPW;
  int grid = 0;
  if (riders.per && !rides<G>(1)) return -1;
  const int occ = queue ? grp_occ_fe<G>() : 0;
  // more workgroups per CU than the full-LDS kernel holds: the compact one, whose homes
  // are indexed by workgroup, so only on a queue-fed grid whose items all have homes
  if constexpr (G == 12) {
    const int occ_c = queue ? fe_compact_occ() : 0;
    if (queue && park && wg_per_cu > occ && occ_c > occ) {
      grid = queue_grid(occ_c, ntiles, wg_per_cu);
      if ((long)grid * IPW <= park_items)
        return fe_compact_launch(grid, fin, status, sub_status, result, gt, M, queue, park, s);
    }
  }
  if (launch_grid(occ, ntiles, Cfg<G>::THREADS, queue, wg_per_cu, riders, M, &grid)) return -1;
  if constexpr (rides<G>(1)) {
    if (riders.per || at_most_one_wave_per_simd(queue, occ, wg_per_cu, Cfg<G>::W)) {
      hipLaunchKernelGGL((k_final_exp_grp<G, true>), dim3(grid), dim3(Cfg<G>::THREADS), final_exp_lds_bytes<G>(), s, fin,
                         status, sub_status, result, gt, M, queue, riders);
      return hipGetLastError() == hipSuccess ? 0 : -1;
    }
  }
  hipLaunchKernelGGL((k_final_exp_grp<G, false>), dim3(grid), dim3(Cfg<G>::THREADS), final_exp_lds_bytes<G>(), s, fin, status,
                     sub_status, result, gt, M, queue, riders);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// the instantiations of this translation unit: lane count GRP_LANES, kernel GRP_KIND
// (0 Miller, 1 final exponentiation; one kernel per unit, so that neither kernel's
// register allocation depends on the other's code)
#define BLS_IS(g) || GRP_LANES == g
#if GRP_KIND == 0 && (0 BLS_GRP_ML_LANES(BLS_IS))
template int grp_launch_ml<GRP_LANES>(const int32_t*, const int32_t*, const int32_t*, const int32_t*, const int32_t*, int,
                                      int32_t*, int32_t*, int, int*, int, int, const SubgroupRiders&, hipStream_t);
template int grp_ml_depth<GRP_LANES>(bool, int, bool, int, int);
template int grp_occ_ml<GRP_LANES>();
#endif
#if GRP_KIND == 1 && (0 BLS_GRP_FE_LANES(BLS_IS))
template int grp_launch_fe<GRP_LANES>(const int32_t*, const int32_t*, const int32_t*, int32_t*, uint64_t*, int, int*, int,
                                      const SubgroupRiders&, int32_t*, long, hipStream_t);
template int grp_occ_fe<GRP_LANES>();
template int grp_occ_fe_compact<GRP_LANES>();
#endif
#if GRP_KIND == 2
int fe_compact_occ() {
  static int occ = -1;
  if (occ < 0 && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_final_exp_compact<12>, Cfg<12>::THREADS,
                                                               final_exp_compact_lds_bytes<12>()) != hipSuccess ||
                  occ < 0))
    occ = 0;
  return occ;
}
int fe_compact_launch(int grid, const int32_t* fin, const int32_t* status, const int32_t* sub_status, int32_t* result,
                      uint64_t* gt, int M, int* queue, int32_t* park, hipStream_t s) {
  hipLaunchKernelGGL((k_final_exp_compact<12>), dim3(grid), dim3(Cfg<12>::THREADS), final_exp_compact_lds_bytes<12>(),
                     s, fin, status, sub_status, result, gt, M, queue, park);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
#endif
#undef BLS_IS

}  // namespace bls
