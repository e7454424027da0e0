#!/usr/bin/env python3
"""Benchmark: batch BLS12-381 signature verification on one MI355X.

A step is one pass of the verification path over one batch: the plain GPU form on a
uniform batch of 4096 verifications with 4096 distinct keys and 64 invalid items
(32 signatures over another message, 16 x-coordinates without a square root, 16
on-curve points outside the r-order subgroup).  Signatures arrive compressed (48 B) and
are decompressed and validated on the device; -H(m) arrives in ic_bls12_381's
Montgomery form; the 4096 per-key line tables are resident in HBM (registered once,
outside the timed region), as are all inputs when the timed region starts.

Beside that headline step the same run measures the other forms on their workloads:
  batched     256 verify batches of geometric sizes (mean 128, [1, 2048]) over a 13-key
              pool, 1024 messages each signed by a random 1..13 signers, 1/128 invalid:
              the plain form (one kernel sequence per batch) against the batched form
              (one graph-replayed sequence with queue-fed persistent kernels, autotuned);
  at_scale    65536 verifications over 40 keys and 4096 messages, lognormal batch sizes
              (median 64, sigma 1.2, [1, 8192]), batched form (and one plain pass);
  wire        the fused wire-format path on a uniform 4096 batch: compressed signatures,
              32-byte messages and key handles in, H(m) hashed on the device (RFC 9380 with
              the reference benchmark's DST), against the compressed path fed -H(m) from
              the host; host entry points, copies included;
  legacy      the reference's unmodified entry point gpu_batch_bls_verify on the uniform
              4096 batch as host arrays (uncompressed signatures, -H(m), 2448 u64 of key
              lines per item): the call's time and its split into copies in, kernels and
              copies out;
  decompress  gpu_g1_decompress_batch (decode + subgroup test) on 448 and 4096 points;
  keys        blsgpu_register_keys_compressed: 4096 public keys from their 96-byte
              encodings (G2 decompression, subgroup test, line tables) per call;
  serving     Poisson arrivals at 2k, 20k and 100k verifications/s, lognormal request
              sizes (median 16, sigma 1.0, [1, 512]) over 40 keys, 1/256 invalid, under a
              4 GB HBM budget: p50/p99/p999 latency and sustained throughput of the
              continuous-batching server (60 s per rate) and of the plain
              one-request-per-step form (5 s per rate: it saturates).

Prints one JSON line (see the keys below).  The CPU baseline is the test oracle (a C
restatement of the reference's CPU path) timed on the host's cores on bounded samples of
the 64-item CPU reference workload and of the 4096 batch; its verdicts are compared with
the GPU's on the 4096 sample.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "icp-gpu_amd"))

# algorithmic work per verification as the reference implements it (Fp multiplications)
FP_MULS_MILLER = 8080
FP_MULS_FINAL_EXP = 8861
FP_MULS_DECOMPRESS = 616
PRODUCTS_PER_FP_MUL = 288  # 12x12 limb products + 12x12 reduction products (32-bit limbs)
# MI355X: 256 CUs at 2.4 GHz, one 32x32->64-bit multiply-add per lane per cycle, the rate
# of the published 78.6 TFLOP/s fp64 vector FMA (39.3 T FMA/s)
MI355X_CUS, MI355X_GHZ = 256, 2.4
PEAK_PRODUCTS = MI355X_CUS * 64 * MI355X_GHZ * 1e9
CPU_PUBLISHED_US = 348.0  # ic_bls12_381 verify on the reference's host (cuda/bench_bls_verify.cu:167,228)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample of the batch")
    ap.add_argument("--cpu-c1-seconds", type=float, default=5.0, help="budget of the 64-item CPU workload's sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--engine", default="auto",
                    help="mode:lanes_miller:lanes_final_exp, or auto (the library's autotuner on this batch)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 256 MiB device copy (calibrates HBM byte counters under rocprofv3)")
    ap.add_argument("--forms", default="plain,wire,legacy,decompress,keys,batched,at_scale,serving",
                    help="comma list of forms to measure beside the headline plain form")
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="length of each continuous serving run")
    ap.add_argument("--plain-serve-seconds", type=float, default=5.0,
                    help="length of each run of the plain one-request-per-step form")
    ap.add_argument("--serve-rates", default="2000,20000,100000")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner for the batched forms")
    ap.add_argument("--only", default="",
                    help="run only this form (legacy, decompress or keys) on the uniform batch and print its JSON "
                         "(for profiling runs; no headline step)")
    ap.add_argument("--batched-engine", default="",
                    help="lanes_miller:lanes_final_exp:wg_per_cu_miller:wg_per_cu_final_exp for the batched forms "
                         "instead of the autotuner (e.g. 12:12:5:4)")
    return ap.parse_args()


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def host_cpu():
    """what the CPU baseline runs on: logical CPUs, the CPUs this process may use, the
    cgroup's CPU quota (cores) and the model name"""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    affinity = len(os.sched_getaffinity(0))
    usable = affinity if quota is None else max(1, min(affinity, math.ceil(quota)))
    return {"nproc": os.cpu_count(), "affinity": affinity, "cgroup_quota_cores": quota, "usable": usable,
            "model": model}


def cpu_baseline(w, gpu_res, c1, budget_s, c1_budget_s):
    """The oracle (C restatement of the reference's CPU path: decompress with every
    check, then e(sig, G2) e(-H(m), pk) == 1) on the host's cores.  Key line tables are
    precomputed outside the timing, as the GPU's are resident.  Measures single-thread
    latency on the 64-item CPU reference workload, then all-core throughput on samples
    of that workload and of the 4096 batch, each repeated to about its budget."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    cpu = host_cpu()
    threads = cpu["usable"]
    g2c = O.g2_prepare(O.g2_gen())

    def prepared(wk):
        keys = np.unique(wk["key_idx"])
        remap = np.zeros(int(keys.max()) + 1, np.int32)
        remap[keys] = np.arange(len(keys), dtype=np.int32)
        kc = np.ascontiguousarray(np.array([O.g2_prepare(wk["pk"][k]) for k in keys]))
        return remap, kc

    def run(wk, remap, kc, idx, nthreads):
        ki = np.ascontiguousarray(remap[wk["key_idx"][idx]])
        s = np.ascontiguousarray(wk["sig"][idx])
        h = np.ascontiguousarray(wk["nhm"][idx])
        t0 = time.perf_counter()
        res, _ = O.verify_compressed_batch(s, h, g2c, kc, ki, threads=nthreads, want_gt=False)
        return res, time.perf_counter() - t0

    def sample(n, per, budget):
        want = max(threads, int(budget * threads / per))
        reps = max(1, -(-want // n))
        idx = np.arange(want) if want < n else np.tile(np.arange(n), reps)
        return idx, reps, want >= n

    # the 64-item CPU reference workload: all 64 on one thread, then all cores
    r1, k1 = prepared(c1)
    res1, t_single = run(c1, r1, k1, np.arange(len(c1["sig"])), 1)
    per1 = t_single / len(c1["sig"])
    assert (res1 == 1).all(), "the 64-item CPU workload is all valid"
    idx1, reps1, _ = sample(len(c1["sig"]), per1, c1_budget_s)
    _, t1 = run(c1, r1, k1, idx1, threads)
    # the 4096 batch on all cores
    r2, k2 = prepared(w)
    n = len(w["sig"])
    idx2, reps2, whole = sample(n, per1, budget_s)
    res2, t2 = run(w, r2, k2, idx2, threads)
    agree = bool((res2 == gpu_res[idx2]).all())
    return {
        "value": len(idx2) / t2,
        "unit": "verifications/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d verifications: the %d-item batch %s (decompress with every check, then verify; key lines "
                  "precomputed), %.1f s on %d threads; verdicts agree with the GPU's: %s"
                  % (len(idx2), n, "repeated %d times" % reps2 if whole else "(first %d items)" % len(idx2), t2,
                     threads, agree),
        "host": cpu,
        "single_thread_ms_per_verify": per1 * 1e3,
        "single_thread_note": "the 64 valid items of the CPU reference workload on one thread; the reference "
                              "publishes %.0f us per verification for ic_bls12_381 on its own host "
                              "(cuda/bench_bls_verify.cu:167,228), so this C restatement runs %.1fx that per core"
                              % (CPU_PUBLISHED_US, per1 * 1e6 / CPU_PUBLISHED_US),
        "c1_64_valid": {"verif_per_s": len(idx1) / t1, "verifications": len(idx1), "seconds": t1,
                        "threads": threads, "sample": "the 64 items repeated %d times" % reps1},
        "verdicts_agree": agree,
    }


def measure_batched(torch, lib, blsgpu, w, handles, steps, plain_steps, tune, fixed=""):
    """plain form (one kernel sequence per batch) vs batched form (one graph-replayed
    sequence, queue-fed persistent kernels) on the same device-resident batches"""
    from blsgpu import workload
    dev = torch.device("cuda:0")
    sizes = np.asarray(w["sizes"], np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    n, nb = int(off[-1]), len(sizes)
    d_sig = torch.from_numpy(w["sig"].reshape(-1)).to(dev)
    d_nhm = torch.from_numpy(w["nhm"].view(np.int64).reshape(-1)).to(dev)
    d_key = torch.from_numpy(handles[w["key_idx"]].astype(np.int32)).to(dev)
    d_res_b = torch.zeros(n, dtype=torch.int32, device=dev)
    d_res_p = torch.zeros(n, dtype=torch.int32, device=dev)
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    offp = _p(off)
    tuned = None
    if fixed:
        g = [int(v) for v in fixed.split(":")]
        blsgpu.set_batched_engine(g[0], g[1], g[2], g[3])
        tuned = {"lanes_miller": g[0], "lanes_final_exp": g[1], "wg_per_cu_miller": g[2], "wg_per_cu_final_exp": g[3],
                 "fixed": True}
    elif tune:
        cfg = np.zeros(4, np.int32)
        ms = np.zeros(2, np.float64)
        t_t = time.perf_counter()
        if lib.blsgpu_autotune_device(d_sig.data_ptr(), d_nhm.data_ptr(), d_key.data_ptr(), n, 0, sp, _p(cfg), _p(ms)):
            raise RuntimeError("autotune: " + blsgpu.last_error())
        tuned = {"lanes_miller": int(cfg[0]), "lanes_final_exp": int(cfg[1]), "wg_per_cu_miller": int(cfg[2]),
                 "wg_per_cu_final_exp": int(cfg[3]), "sequence_ms": float(ms[0]), "final_exp_ms": float(ms[1]),
                 "tuning_s": round(time.perf_counter() - t_t, 2)}
        log("autotuned for %d items: %s" % (n, tuned))

    def batched():
        if lib.blsgpu_verify_batches_device(d_sig.data_ptr(), d_nhm.data_ptr(), d_key.data_ptr(), offp, nb,
                                            d_res_b.data_ptr(), None, sp):
            raise RuntimeError(blsgpu.last_error())

    def plain():
        for b in range(nb):
            if not sizes[b]:
                continue
            o = int(off[b])
            if lib.blsgpu_verify_compressed_device(d_sig.data_ptr() + 48 * o, d_nhm.data_ptr() + 96 * o,
                                                   d_key.data_ptr() + 4 * o, d_res_p.data_ptr() + 4 * o, None,
                                                   int(sizes[b]), sp):
                raise RuntimeError(blsgpu.last_error())

    def timed(fn, k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / k

    batched()
    batched()
    plain()
    ms_b = timed(batched, steps)
    ms_p = timed(plain, plain_steps)
    want = workload.expected_codes(w["cat"])
    rb, rp = d_res_b.cpu().numpy(), d_res_p.cpu().numpy()
    return {
        "verifications": n, "batches": nb, "batch_size_min": int(sizes.min()), "batch_size_max": int(sizes.max()),
        "batched_verif_per_s": n / (ms_b * 1e-3), "plain_verif_per_s": n / (ms_p * 1e-3),
        "batched_ms": ms_b, "plain_ms": ms_p, "speedup": ms_p / ms_b,
        "steps_batched": steps, "steps_plain": plain_steps,
        "engine": tuned or dict(zip(("lanes_miller", "lanes_final_exp", "wg_per_cu_miller", "wg_per_cu_final_exp",
                                     "graph"), blsgpu.get_batched_engine(n))),
        "graphs": dict(zip(("instantiated", "launches"), blsgpu.batched_stats())),
        "verdicts_as_expected": bool((rb == want).all() and (rp == want).all()),
        "batched_equals_plain": bool((rb == rp).all()),
    }


IC_DST = b"BLS_SIG_BLS12381G2_XMD:SHA-256_SSWU_RO_NUL_"  # ic-bench/src/main.rs:68


def measure_wire(lib, blsgpu, sk, handles, key_idx, seed, steps):
    from blsgpu import workload
    n = len(key_idx)
    msgs = [workload.message(seed, i) for i in range(n)]
    h, _ = blsgpu.hash_to_g1(msgs, IC_DST)
    sig = workload.sign_points(h, sk, key_idx)
    want = np.ones(n, np.int32)
    for i in range(0, n, 64):  # 1/64 of the messages changed after signing
        msgs[i] = bytes([msgs[i][0] ^ 1]) + msgs[i][1:]
        want[i] = 0
    offs = np.zeros(n + 1, np.int64)
    offs[1:] = np.cumsum([len(m) for m in msgs])
    mb = np.frombuffer(b"".join(msgs), np.uint8).copy()
    dst = np.frombuffer(IC_DST, np.uint8).copy()
    hk = np.ascontiguousarray(handles[np.asarray(key_idx)], np.int32)
    res = np.zeros(n, np.int32)
    nh = np.zeros((n, 12), np.uint64)
    ident = np.zeros(n, np.int32)

    def wire():
        if lib.blsgpu_verify_messages(_p(sig), _p(mb), _p(offs), _p(dst), len(IC_DST), _p(hk), _p(res), None, n):
            raise RuntimeError(blsgpu.last_error())

    def hash_only():
        if lib.blsgpu_hash_to_g1(_p(mb), _p(offs), n, _p(dst), len(IC_DST), _p(nh), _p(ident), 1):
            raise RuntimeError(blsgpu.last_error())

    def compressed():
        if lib.blsgpu_verify_compressed(_p(sig), _p(nh), _p(hk), _p(res), None, n):
            raise RuntimeError(blsgpu.last_error())

    def timed(fn):
        fn()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        return (time.perf_counter() - t) * 1e3 / steps

    ms_wire = timed(wire)
    ok = bool((res == want).all())
    ms_hash = timed(hash_only)
    ms_comp = timed(compressed)
    ok = ok and bool((res == want).all())
    return {"verif_per_s": n / (ms_wire * 1e-3), "ms_per_step": ms_wire, "hash_ms": ms_hash,
            "compressed_ms": ms_comp, "steps": steps, "verdicts_as_expected": ok,
            "workload": "%d verifications, %d keys, distinct 32-byte messages, 1/64 signed over another message; "
                        "host entry points (copies included): blsgpu_verify_messages, blsgpu_hash_to_g1 alone, "
                        "blsgpu_verify_compressed fed the same -H(m)" % (n, int(sk.shape[0]))}


def measure_legacy(lib, blsgpu, w, handles, steps):
    """gpu_batch_bls_verify as the replica calls it (gpu-crypto/cuda/api.cu:128-160): host
    arrays of uncompressed signatures (ic_bls12_381's affine form without the flag),
    -H(m), the generator's lines and 2448 u64 of lines per item"""
    from blsgpu import workload
    n = len(w["sig"])
    pts, st = blsgpu.decompress_status(w["sig"], check_subgroup=False)  # undecodable encodings: zeros
    pkc = np.empty((n, 2448), np.uint64)
    for i in range(n):
        pkc[i] = blsgpu.key_coeffs(int(handles[w["key_idx"][i]]))
    hg = blsgpu.register_key_affine(workload.g2_generator())
    g2c = blsgpu.key_coeffs(hg)
    nhm = np.ascontiguousarray(w["nhm"])
    res = np.zeros(n, np.int32)
    args = (_p(pts), _p(nhm), _p(g2c), _p(pkc), _p(res), n)
    if lib.gpu_batch_bls_verify(*args):
        raise RuntimeError(blsgpu.last_error())
    lib.blsgpu_set_profiling(1)
    t0 = time.perf_counter()
    for _ in range(steps):
        if lib.gpu_batch_bls_verify(*args):
            raise RuntimeError(blsgpu.last_error())
    ms = (time.perf_counter() - t0) * 1e3 / steps
    split = np.zeros(3, np.float64)
    calls = ctypes.c_int(0)
    lib.blsgpu_legacy_phase_times(_p(split), ctypes.byref(calls))
    lib.blsgpu_set_profiling(0)
    split /= max(calls.value, 1)
    # without any flag the legacy layout decodes nothing itself: only category 0 verifies
    want = (w["cat"] == 0).astype(np.int32)
    in_bytes = pts.nbytes + nhm.nbytes + g2c.nbytes + pkc.nbytes
    return {"verif_per_s": n / (ms * 1e-3), "ms_per_call": ms, "steps": steps,
            "h2d_ms": float(split[0]), "kernels_after_h2d_ms": float(split[1]), "d2h_ms": float(split[2]),
            "h2d_share": float(split[0] / ms), "h2d_bytes": int(in_bytes),
            "h2d_GB_per_s": in_bytes / (split[0] * 1e-3) / 1e9 if split[0] > 0 else None,
            "verdicts_as_expected": bool((res == want).all()),
            "workload": "the uniform %d batch as host arrays: %d uncompressed signatures and -H(m) (96 B each), "
                        "the generator's lines, %d x 2448 u64 of key lines (19,584 B per item, in pageable "
                        "memory); blocking call, copies included; h2d_ms runs from the first copy in until the "
                        "last key table has landed (the first half's Miller loop runs during the second half's "
                        "copies), kernels_after_h2d_ms from then until the kernels end" % (n, n, n)}


def measure_decompress(lib, blsgpu, points48, steps):
    """gpu_g1_decompress_batch (ic-integration/gpu_crypto/cuda/decompress.cu:198-226) on
    valid points: decode (flags, x < p, square root, sign) and the subgroup test"""
    out = []
    for n in (448, 4096):
        src = np.ascontiguousarray(points48[:n])
        dst = np.zeros((n, 96), np.uint8)
        if lib.gpu_g1_decompress_batch(_p(src), _p(dst), n):
            raise RuntimeError(blsgpu.last_error())
        t0 = time.perf_counter()
        for _ in range(steps):
            if lib.gpu_g1_decompress_batch(_p(src), _p(dst), n):
                raise RuntimeError(blsgpu.last_error())
        ms = (time.perf_counter() - t0) * 1e3 / steps
        # the same points without the subgroup test (the reference's decoder checks
        # nothing): blsgpu_g1_decompress_status with check_subgroup = 0
        st = np.zeros(n, np.int32)
        t0 = time.perf_counter()
        for _ in range(steps):
            if lib.blsgpu_g1_decompress_status(_p(src), _p(dst), _p(st), n, 0):
                raise RuntimeError(blsgpu.last_error())
        ms_nosub = (time.perf_counter() - t0) * 1e3 / steps
        if not (st == 0).all():
            raise RuntimeError("a valid point was rejected")
        out.append({"points": n, "ms_per_call": ms, "points_per_s": n / (ms * 1e-3), "steps": steps,
                    "ms_per_call_without_subgroup_test": ms_nosub})
    return {"runs": out,
            "workload": "valid compressed signatures of the uniform batch; blocking host entry points, copies "
                        "included: gpu_g1_decompress_batch (every check, the subgroup test included), and "
                        "blsgpu_g1_decompress_status without the subgroup test; the reference publishes "
                        "448 points in 0.98 ms on an RTX PRO 6000 (README.md:243), with none of the checks"}


def measure_keys(lib, blsgpu, pk, steps):
    """blsgpu_register_keys_compressed: keys from their 96-byte encodings, decompressed,
    subgroup-checked and prepared (68 lines each) on the GPU"""
    from blsgpu import workload
    enc = workload.g2_compress(pk)
    n = len(enc)
    times = []
    for r in range(steps + 1):
        lib.blsgpu_reset_keys()
        t0 = time.perf_counter()
        h, st = blsgpu.register_keys_compressed(enc)
        times.append((time.perf_counter() - t0) * 1e3)
        if not (st == 0).all() or h.tolist() != list(range(n)):
            raise RuntimeError("key registration did not accept every key")
    ms = float(np.mean(times[1:]))
    lib.blsgpu_reset_keys()
    return {"keys": n, "ms_per_call": ms, "keys_per_s": n / (ms * 1e-3), "steps": steps,
            "workload": "%d distinct public keys as 96-byte compressed G2 points; blocking host entry point "
                        "(copy in, decompression with the subgroup test, line tables, copy of the statuses)" % n}


def measure_serving(blsgpu, pool, handles, rates, seconds, plain_seconds, seed):
    """continuous-batching server vs the plain one-request-per-step form, per arrival rate"""
    keys = handles[pool["key_idx"]].astype(np.int32)
    out = []
    for rate in rates:
        row = {"offered_verif_per_s": rate}
        for name, kw, secs, drain in (
                ("continuous", dict(slots=3, max_step_items=8192), seconds, 2.0),
                ("plain_per_request", dict(slots=1, max_step_items=512, max_requests_per_step=1), plain_seconds, 2.0)):
            log("serving %s at %d verif/s for %.0f s" % (name, rate, secs))
            with blsgpu.Server(hbm_budget_bytes=4 << 30, **kw) as srv:
                r = srv.loadtest(pool["sig"], pool["nhm"], keys, items_per_s=rate, seconds=secs, drain_s=drain,
                                 size_median=16, size_sigma=1.0, size_min=1, size_max=512, seed=seed + rate)
                st = srv.stats()
            r = {k: (None if isinstance(v, float) and not np.isfinite(v) else v) for k, v in r.items()}
            r["seconds"] = secs
            r["steps"] = int(st["steps"])
            r["mean_step_items"] = st["items"] / max(st["steps"], 1)
            r["max_steps_in_flight"] = int(st["max_inflight"])
            row[name] = r
        out.append(row)
    return out


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summaries
    (a line 'traffic_bytes_per_launch <kernel> <bytes>' in profiles/*.md), or None."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for name in sorted(os.listdir(d)):
        if not name.endswith(".md"):
            continue
        for line in open(os.path.join(d, name)):
            f = line.strip().strip("`").split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
                return {"bytes": float(f[2]), "source": "profiles/" + name}
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X (no multi-GPU path); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    log("importing torch")
    import torch
    import blsgpu
    from blsgpu import workload

    lib = blsgpu.load()
    if blsgpu.init() != 0:
        raise RuntimeError("no GPU: " + blsgpu.last_error())

    n = args.batch
    log("generating workload n=%d seed=%d" % (n, args.seed))
    t_gen = time.perf_counter()
    w = workload.uniform(n, args.seed)
    sk, pk, key_idx, cat, sig, nhm = w["sk"], w["pk"], w["key_idx"], w["cat"], w["sig"], w["nhm"]
    lib.blsgpu_reset_keys()
    handles = blsgpu.register_keys_affine(pk)
    t_gen = time.perf_counter() - t_gen

    if args.only:
        mode, gml, gfe = (int(x) for x in (args.engine if args.engine != "auto" else "1:32:12").split(":"))
        blsgpu.set_engine(mode, gml, gfe)
        if args.copy_probe:
            src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda:0")
            torch.empty_like(src).copy_(src)
            torch.cuda.synchronize()
        if args.only == "legacy":
            r = measure_legacy(lib, blsgpu, w, handles, args.steps)
        elif args.only == "decompress":
            r = measure_decompress(lib, blsgpu, np.concatenate([sig[cat <= 1], workload.valid_only(64, args.seed + 1)
                                                                ["sig"]]), args.steps)
        elif args.only == "keys":
            r = measure_keys(lib, blsgpu, pk, args.steps)
        else:
            raise SystemExit("--only takes legacy, decompress or keys")
        print(json.dumps({"only": args.only, "engine": args.engine, args.only: r}))
        return

    dev = torch.device("cuda:0")
    d_sig = torch.from_numpy(sig.reshape(-1)).to(dev)
    d_nhm = torch.from_numpy(nhm.view(np.int64).reshape(-1)).to(dev)
    d_key = torch.from_numpy(handles[key_idx]).to(dev)
    d_res = torch.zeros(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)

    def step():
        rc = lib.blsgpu_verify_compressed_device(d_sig.data_ptr(), d_nhm.data_ptr(), d_key.data_ptr(),
                                                 d_res.data_ptr(), None, n, sp)
        if rc != 0:
            raise RuntimeError(blsgpu.last_error())

    tuning = None
    if args.engine == "auto":
        # the library's autotuner, plain form: lanes per item of each kernel, timed on this batch
        cfg = np.zeros(4, np.int32)
        ms = np.zeros(2, np.float64)
        t_t = time.perf_counter()
        if lib.blsgpu_autotune_device(d_sig.data_ptr(), d_nhm.data_ptr(), d_key.data_ptr(), n, 1, sp, _p(cfg), _p(ms)):
            raise RuntimeError("autotune: " + blsgpu.last_error())
        tuning = {"sequence_ms": float(ms[0]), "final_exp_ms": float(ms[1]),
                  "tuning_s": round(time.perf_counter() - t_t, 2)}
        args.engine = "1:%d:%d" % (cfg[0], cfg[1])
    mode, gml, gfe = (int(x) for x in args.engine.split(":"))
    blsgpu.set_engine(mode, gml, gfe)
    log("engine %s" % args.engine)
    log("warmup %d steps" % args.warmup)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib.blsgpu_set_profiling(1)
    torch.cuda.synchronize()
    log("timing %d steps" % args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phase = np.zeros(4, np.float64)
    calls = ctypes.c_int(0)
    lib.blsgpu_phase_times(_p(phase), ctypes.byref(calls))
    lib.blsgpu_set_profiling(0)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n / (ms_per_step * 1e-3)

    if args.copy_probe:
        src = torch.empty(256 << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    res = d_res.cpu().numpy()
    want = workload.expected_codes(cat)
    ok = bool((res == want).all())

    # roofline of the dominant kernel: algorithmic 32-bit limb products / its duration
    # (HIP events recorded between the kernels on the launch stream)
    phase_avg = phase / max(calls.value, 1)
    names = ["g1_decompress", "neg_hm_convert", "miller_loop", "final_exp"]
    decoded = int((want >= 0).sum())
    work = [n * FP_MULS_DECOMPRESS, 0, decoded * FP_MULS_MILLER, decoded * FP_MULS_FINAL_EXP]
    dom = int(np.argmax(phase_avg))
    mads, _ = workload.probe_mad_rate()
    achieved = work[dom] * PRODUCTS_PER_FP_MUL / (phase_avg[dom] * 1e-3)
    roofline = {
        "bound": "valu",
        "kernel": names[dom],
        "achieved": achieved / 1e12,
        "peak": PEAK_PRODUCTS / 1e12,
        "unit": "T(32x32-bit products)/s",
        "frac": achieved / PEAK_PRODUCTS,
        "traffic": None,
        "traffic_source": None,
        "work_per_launch": "%d decoded items x %d Fp multiplications x %d limb products"
                           % (decoded, FP_MULS_FINAL_EXP if dom == 3 else FP_MULS_MILLER, PRODUCTS_PER_FP_MUL),
        "peak_note": "256 CUs x 64 lanes x 2.4 GHz, one 32x32-bit multiply-add per lane per cycle (the published "
                     "78.6 TFLOP/s fp64 FMA rate); the in-run issue probe measured %.2f T/s" % (mads / 1e12),
        "peak_measured": mads / 1e12,
        "kernel_ms": {k: round(float(v), 4) for k, v in zip(names, phase_avg)},
        "whole_job_frac": value * (FP_MULS_MILLER + FP_MULS_FINAL_EXP) * PRODUCTS_PER_FP_MUL / PEAK_PRODUCTS,
    }
    tr = read_traffic("%s_g%d" % (names[dom], gfe if names[dom] == "final_exp" else gml))
    if tr:
        roofline["traffic"] = tr["bytes"]
        roofline["traffic_source"] = tr["source"]
    out = {
        "metric": "BLS sig verifications/sec on 1 GPU, batched vs plain; p50/p99 latency at load",
        "value": value,
        "unit": "verifications/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 limbs (14 x 28-bit signed, exact 381-bit integer arithmetic)",
        "data": "synthetic: seeded keys/messages generated on the device (seed %d)" % args.seed,
        "config": {"workload": "plain GPU form, uniform batch of %d verifications, %d distinct keys, "
                               "64 invalid (32 wrong-message, 16 no-sqrt, 16 off-subgroup), "
                               "compressed signatures, inputs resident in HBM" % (n, n),
                   "global_batch": n, "parallelism": "single GPU"},
        "verdicts_as_expected": ok,
        "roofline": roofline,
        "setup_s": round(t_gen, 2),
        "engine": {"mode": mode, "lanes_miller": gml, "lanes_final_exp": gfe, "autotuned": tuning},
    }
    forms = set(f for f in args.forms.split(",") if f)
    t_forms = time.perf_counter()
    if "wire" in forms:
        log("wire-format path")
        out["wire"] = measure_wire(lib, blsgpu, sk, handles, key_idx, args.seed + 6, 10)
        log("wire: %s" % {k: out["wire"][k] for k in ("verif_per_s", "hash_ms", "compressed_ms")})
    if "legacy" in forms:
        log("legacy entry point on host arrays")
        out["legacy"] = measure_legacy(lib, blsgpu, w, handles, 5)
        log("legacy: %s" % {k: out["legacy"][k] for k in ("verif_per_s", "ms_per_call", "h2d_ms", "kernels_after_h2d_ms")})
    c1 = None
    if "decompress" in forms or not args.no_cpu:
        c1 = workload.valid_only(64, args.seed + 1)
    if "decompress" in forms:
        log("G1 decompression drop-in")
        pts = np.concatenate([sig[cat <= 1], c1["sig"]])
        out["decompress"] = measure_decompress(lib, blsgpu, pts, 20)
        log("decompress: %s" % out["decompress"]["runs"])
    if "keys" in forms:
        log("key registration from compressed encodings")
        out["keys"] = measure_keys(lib, blsgpu, pk, 3)
        log("keys: %s" % {k: out["keys"][k] for k in ("ms_per_call", "keys_per_s")})
        lib.blsgpu_reset_keys()
        handles = blsgpu.register_keys_affine(pk)
    if "batched" in forms or "at_scale" in forms or "serving" in forms:
        blsgpu.set_batched_engine(12, 16, 0, 0, True)
    if "batched" in forms:
        log("batched form: generating the 256-batch workload")
        w3 = workload.batched(args.seed + 3)
        lib.blsgpu_reset_keys()
        h3 = blsgpu.register_keys_affine(w3["pk"])
        out["batched"] = measure_batched(torch, lib, blsgpu, w3, h3, args.batched_steps, 2, not args.no_tune,
                                         args.batched_engine)
        out["batched"]["workload"] = ("256 batches, geometric sizes mean 128 on [1,2048], 13 signers, 1024 messages "
                                      "each signed by 1..13 signers (%d signatures), 1/128 invalid"
                                      % w3["signatures"])
        log("batched: %s" % {k: out["batched"][k] for k in ("batched_verif_per_s", "plain_verif_per_s", "speedup")})
    w4 = None
    if "at_scale" in forms or "serving" in forms:
        log("at-scale workload: 65536 items over 40 keys")
        w4 = workload.at_scale(args.seed + 4)
        lib.blsgpu_reset_keys()
        h4 = blsgpu.register_keys_affine(w4["pk"])
    if "at_scale" in forms:
        out["at_scale"] = measure_batched(torch, lib, blsgpu, w4, h4, max(2, args.batched_steps // 2), 1,
                                          not args.no_tune, args.batched_engine)
        out["at_scale"]["workload"] = ("65536 verifications, 40 signer keys, 4096 messages, lognormal batch sizes "
                                       "(median 64, sigma 1.2, [1,8192]), 1/128 invalid")
        log("at_scale: %s" % {k: out["at_scale"][k] for k in ("batched_verif_per_s", "plain_verif_per_s", "speedup")})
    if "serving" in forms:
        pool = workload.serving_pool(args.seed + 5, w4["sk"])
        rates = [int(r) for r in args.serve_rates.split(",") if r]
        out["serving"] = {
            "workload": "Poisson arrivals, request sizes lognormal (median 16, sigma 1.0, [1,512]), 40 keys, fresh "
                        "messages (pool of %d items, 1/256 invalid), 4 GB HBM budget; continuous server %.0f s per "
                        "rate, plain per-request form %.0f s per rate"
                        % (len(pool["key_idx"]), args.serve_seconds, args.plain_serve_seconds),
            "latency": "from each request's scheduled arrival to the host seeing its results; null = not finished "
                       "before the drain deadline",
            "runs": measure_serving(blsgpu, pool, h4, rates, args.serve_seconds, args.plain_serve_seconds, args.seed),
        }
    if forms - {"plain"}:
        out["forms_s"] = round(time.perf_counter() - t_forms, 1)
    if not args.no_cpu:
        log("cpu baseline (budgets %.0f s + %.0f s)" % (args.cpu_seconds, args.cpu_c1_seconds))
        out["cpu_baseline"] = cpu_baseline(w, res, c1, args.cpu_seconds, args.cpu_c1_seconds)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
