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
  batched   256 verify batches of geometric sizes (mean 128, [1, 2048]) over a 13-key
            pool, 1024 messages each signed by a random 1..13 signers, 1/128 invalid:
            the plain form (one kernel sequence per batch) against the batched form
            (one graph-replayed sequence with queue-fed persistent kernels, autotuned);
  at_scale  65536 verifications over 40 keys and 4096 messages, lognormal batch sizes
            (median 64, sigma 1.2, [1, 8192]), batched form (and one plain pass);
  wire      the fused wire-format path on a uniform 4096 batch: compressed signatures,
            32-byte messages and key handles in, H(m) hashed on the device (RFC 9380 with
            the reference benchmark's DST), against the compressed path fed -H(m) from
            the host; host entry points, copies included;
  serving   Poisson arrivals at 2k, 20k and 100k verifications/s, lognormal request
            sizes (median 16, sigma 1.0, [1, 512]) over 40 keys, 1/256 invalid, under a
            4 GB HBM budget: p50/p99/p999 latency and sustained throughput of the
            continuous-batching server and of the plain one-request-per-step form.

Prints one JSON line (see the keys below).  The CPU baseline is the test oracle (a C
restatement of the reference's CPU path) timed on the host's cores on a bounded sample
of the same batch; its verdicts are also compared with the GPU's on that sample.
"""
import argparse
import ctypes
import json
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
MI355X_CUS, MI355X_GHZ = 256, 2.4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--engine", default="auto",
                    help="mode:lanes_miller:lanes_final_exp, or auto (the library's autotuner on this batch)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 256 MiB device copy (calibrates HBM byte counters under rocprofv3)")
    ap.add_argument("--forms", default="plain,wire,batched,at_scale,serving",
                    help="comma list of forms to measure beside the headline plain form")
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--serve-seconds", type=float, default=5.0, help="length of each serving run (60 for the full runs)")
    ap.add_argument("--serve-rates", default="2000,20000,100000")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner for the batched forms")
    ap.add_argument("--batched-engine", default="",
                    help="lanes_miller:lanes_final_exp:wg_per_cu_miller:wg_per_cu_final_exp for the batched forms "
                         "instead of the autotuner (e.g. 12:12:5:4)")
    return ap.parse_args()


def make_workload(lib, n, seed):
    """uniform batch: n distinct keys, item i signed by key i, 64 invalid (if n >= 128)"""
    rng = np.random.default_rng(seed)
    cat = np.zeros(n, np.int8)
    n_bad = 64 if n >= 128 else max(n // 16, 0)
    if n_bad:
        pos = rng.choice(n, size=n_bad, replace=False)
        half, q = n_bad // 2, n_bad // 4
        cat[pos[:half]] = 1            # wrong message
        cat[pos[half:half + q]] = 2    # x^3+4 not a square
        cat[pos[half + q:]] = 3        # off the subgroup
    sk = np.zeros((n, 32), np.uint8)
    lib.blsgpu_wl_secret_keys(ctypes.c_uint64(seed), n, sk.ctypes.data_as(ctypes.c_void_p))
    pk = np.zeros((n, 24), np.uint64)
    if lib.blsgpu_wl_public_keys(sk.ctypes.data_as(ctypes.c_void_p), n, pk.ctypes.data_as(ctypes.c_void_p)):
        raise RuntimeError("public keys: " + lib.blsgpu_wl_last_error().decode())
    key_idx = np.arange(n, dtype=np.int32)
    sig = np.zeros((n, 48), np.uint8)
    nhm = np.zeros((n, 12), np.uint64)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    if lib.blsgpu_wl_items(ctypes.c_uint64(seed), n, p(sk), n, p(key_idx), p(cat), p(sig), p(nhm)):
        raise RuntimeError("items: " + lib.blsgpu_wl_last_error().decode())
    return sk, pk, key_idx, cat, sig, nhm


def expected_codes(cat):
    return np.select([cat == 0, cat == 1, cat == 2, cat == 3], [1, 0, -3, -4]).astype(np.int32)


def cpu_baseline(sig, nhm, pk, key_idx, gpu_res, budget_s):
    """Oracle (CPU restatement of the reference's path) on a bounded sample: the batch,
    repeated until the sample holds about budget_s seconds of work on the host's cores
    (key line tables precomputed outside the timing, as the GPU's are)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    cores = min(16, os.cpu_count() or 1)
    g2c = O.g2_prepare(O.g2_gen())
    keys = np.unique(key_idx)
    remap = np.zeros(int(keys.max()) + 1, np.int32)
    remap[keys] = np.arange(len(keys), dtype=np.int32)
    kc = np.ascontiguousarray(np.array([O.g2_prepare(pk[k]) for k in keys]))

    def run(idx, threads):
        ki = np.ascontiguousarray(remap[key_idx[idx]])
        s = np.ascontiguousarray(sig[idx])
        h = np.ascontiguousarray(nhm[idx])
        res = np.zeros(len(idx), np.int32)
        t0 = time.perf_counter()
        O.lib().orc_verify_compressed_batch(O.ptr(s), O.ptr(h), O.ptr(g2c), O.ptr(kc), O.ptr(ki), O.ptr(res), None,
                                           len(idx), threads)
        return res, time.perf_counter() - t0
    _, t1 = run(np.arange(4), 1)
    per = t1 / 4
    n = len(sig)
    want = max(cores, int(budget_s * cores / per))
    reps = max(1, -(-want // n))
    idx = np.arange(want) if want < n else np.tile(np.arange(n), reps)
    res, t = run(idx, cores)
    agree = bool((res == gpu_res[idx]).all())
    return {
        "value": len(idx) / t,
        "unit": "verifications/s",
        "cores": cores,
        "kind": "port",
        "sample": "%d verifications: the %d-item batch %s (decompress+checks+verify, key lines precomputed), %.1f s "
                  "on %d threads; single-thread %.2f ms/verify; verdicts agree with GPU: %s"
                  % (len(idx), n, "repeated %d times" % reps if want >= n else "(first %d items)" % len(idx), t,
                     cores, per * 1e3, agree),
        "single_thread_ms": per * 1e3,
        "verdicts_agree": agree,
    }


def gen_items(wl, seed, sk, key_idx, msg_idx, cat):
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    n = len(key_idx)
    sig = np.zeros((n, 48), np.uint8)
    nhm = np.zeros((n, 12), np.uint64)
    key_idx = np.ascontiguousarray(key_idx, np.int32)
    msg_idx = np.ascontiguousarray(msg_idx, np.int64)
    cat = np.ascontiguousarray(cat, np.int8)
    if wl.blsgpu_wl_items_msg(ctypes.c_uint64(seed), n, p(sk), sk.shape[0], p(key_idx), p(msg_idx), p(cat), p(sig),
                              p(nhm)):
        raise RuntimeError("items: " + wl.blsgpu_wl_last_error().decode())
    return sig, nhm


def gen_keys(wl, seed, nkeys):
    sk = np.zeros((nkeys, 32), np.uint8)
    wl.blsgpu_wl_secret_keys(ctypes.c_uint64(seed), nkeys, sk.ctypes.data_as(ctypes.c_void_p))
    pk = np.zeros((nkeys, 24), np.uint64)
    if wl.blsgpu_wl_public_keys(sk.ctypes.data_as(ctypes.c_void_p), nkeys, pk.ctypes.data_as(ctypes.c_void_p)):
        raise RuntimeError("public keys: " + wl.blsgpu_wl_last_error().decode())
    return sk, pk


def mixed_invalid(rng, n, every):
    """categories: 1/every of the items invalid, cycling wrong-message, no-sqrt, off-subgroup"""
    cat = np.zeros(n, np.int8)
    bad = rng.choice(n, size=n // every, replace=False)
    cat[bad] = 1 + (np.arange(len(bad)) % 3)
    return cat


def batched_workload(wl, seed):
    """256 batches, geometric sizes (mean 128, [1, 2048]); 13 signers; 1024 messages each
    signed by a random subset of 1..13 signers; items drawn from those signatures; 1/128 invalid"""
    rng = np.random.default_rng(seed)
    sk, pk = gen_keys(wl, seed, 13)
    pool = []
    for m in range(1024):
        k = int(rng.integers(1, 14))
        for s in rng.choice(13, size=k, replace=False):
            pool.append((m, int(s)))
    pool = np.array(pool)
    sizes = np.clip(rng.geometric(1.0 / 128, size=256), 1, 2048)
    n = int(sizes.sum())
    pick = rng.integers(0, len(pool), n)
    msg_idx, key_idx = pool[pick, 0], pool[pick, 1]
    cat = mixed_invalid(rng, n, 128)
    sig, nhm = gen_items(wl, seed, sk, key_idx, msg_idx, cat)
    return dict(sk=sk, pk=pk, sizes=sizes, key_idx=key_idx, cat=cat, sig=sig, nhm=nhm, signatures=len(pool))


def at_scale_workload(wl, seed, total=65536):
    """40 keys, 4096 messages, lognormal batch sizes (median 64, sigma 1.2, [1, 8192])"""
    rng = np.random.default_rng(seed)
    sk, pk = gen_keys(wl, seed, 40)
    sizes = []
    while sum(sizes) < total:
        sizes.append(int(np.clip(round(np.exp(np.log(64) + 1.2 * rng.normal())), 1, 8192)))
    sizes[-1] -= sum(sizes) - total
    sizes = np.array([x for x in sizes if x > 0])
    key_idx = rng.integers(0, 40, total)
    msg_idx = rng.integers(0, 4096, total)
    cat = mixed_invalid(rng, total, 128)
    sig, nhm = gen_items(wl, seed, sk, key_idx, msg_idx, cat)
    return dict(sk=sk, pk=pk, sizes=sizes, key_idx=key_idx, cat=cat, sig=sig, nhm=nhm)


def serving_pool(wl, seed, sk, n=65536):
    """fresh random messages over the given keys, 1/256 invalid"""
    rng = np.random.default_rng(seed)
    key_idx = rng.integers(0, sk.shape[0], n)
    msg_idx = 5_000_000 + np.arange(n)
    cat = mixed_invalid(rng, n, 256)
    sig, nhm = gen_items(wl, seed, sk, key_idx, msg_idx, cat)
    return dict(key_idx=key_idx, cat=cat, sig=sig, nhm=nhm)


def measure_batched(torch, lib, blsgpu, w, handles, steps, plain_steps, tune, fixed=""):
    """plain form (one kernel sequence per batch) vs batched form (one graph-replayed
    sequence, queue-fed persistent kernels) on the same device-resident batches"""
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
    offp = off.ctypes.data_as(ctypes.c_void_p)
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
        if lib.blsgpu_autotune_device(d_sig.data_ptr(), d_nhm.data_ptr(), d_key.data_ptr(), n, 0, sp,
                                      cfg.ctypes.data_as(ctypes.c_void_p), ms.ctypes.data_as(ctypes.c_void_p)):
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
    want = expected_codes(w["cat"])
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


def measure_wire(lib, wl, blsgpu, sk, handles, key_idx, seed, steps):
    vp = ctypes.c_void_p
    p = lambda a: a.ctypes.data_as(vp)
    n = len(key_idx)
    msgs = []
    buf = np.zeros(32, np.uint8)
    for i in range(n):
        wl.blsgpu_wl_message(ctypes.c_uint64(seed), ctypes.c_int64(i), p(buf))
        msgs.append(buf.tobytes())
    h, _ = blsgpu.hash_to_g1(msgs, IC_DST)
    sig = np.zeros((n, 48), np.uint8)
    kidx = np.ascontiguousarray(key_idx, np.int32)
    wl.blsgpu_wl_sign_points.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, vp, vp]
    if wl.blsgpu_wl_sign_points(p(h), n, p(sk), sk.shape[0], p(kidx), p(sig)):
        raise RuntimeError("sign: " + wl.blsgpu_wl_last_error().decode())
    want = np.ones(n, np.int32)
    for i in range(0, n, 64):  # 1/64 of the messages changed after signing
        msgs[i] = bytes([msgs[i][0] ^ 1]) + msgs[i][1:]
        want[i] = 0
    offs = np.zeros(n + 1, np.int64)
    offs[1:] = np.cumsum([len(m) for m in msgs])
    mb = np.frombuffer(b"".join(msgs), np.uint8).copy()
    dst = np.frombuffer(IC_DST, np.uint8).copy()
    hk = np.ascontiguousarray(handles[kidx], np.int32)
    res = np.zeros(n, np.int32)
    nh = np.zeros((n, 12), np.uint64)
    ident = np.zeros(n, np.int32)

    def wire():
        if lib.blsgpu_verify_messages(p(sig), p(mb), p(offs), p(dst), len(IC_DST), p(hk), p(res), None, n):
            raise RuntimeError(blsgpu.last_error())

    def hash_only():
        if lib.blsgpu_hash_to_g1(p(mb), p(offs), n, p(dst), len(IC_DST), p(nh), p(ident), 1):
            raise RuntimeError(blsgpu.last_error())

    def compressed():
        if lib.blsgpu_verify_compressed(p(sig), p(nh), p(hk), p(res), None, n):
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


def measure_serving(blsgpu, pool, handles, rates, seconds, seed):
    """continuous-batching server vs the plain one-request-per-step form, per arrival rate"""
    keys = handles[pool["key_idx"]].astype(np.int32)
    out = []
    for rate in rates:
        row = {"offered_verif_per_s": rate}
        for name, kw, drain in (("continuous", dict(slots=3, max_step_items=8192), max(2.0, seconds)),
                                ("plain_per_request", dict(slots=1, max_step_items=512, max_requests_per_step=1),
                                 2.0)):
            log("serving %s at %d verif/s for %.0f s" % (name, rate, seconds))
            with blsgpu.Server(hbm_budget_bytes=4 << 30, **kw) as srv:
                r = srv.loadtest(pool["sig"], pool["nhm"], keys, items_per_s=rate, seconds=seconds, drain_s=drain,
                                 size_median=16, size_sigma=1.0, size_min=1, size_max=512, seed=seed + rate)
                st = srv.stats()
            r = {k: (None if isinstance(v, float) and not np.isfinite(v) else v) for k, v in r.items()}
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


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X (no multi-GPU path); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    log("importing torch")
    import torch
    import blsgpu

    lib = blsgpu.load()
    wl = ctypes.CDLL(blsgpu.LIB_PATH)
    wl.blsgpu_wl_last_error.restype = ctypes.c_char_p
    vp = ctypes.c_void_p
    wl.blsgpu_probe_mad_rate.argtypes = [vp, vp]
    lib.blsgpu_verify_compressed_device.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int, vp]
    lib.blsgpu_phase_times.argtypes = [vp, vp]
    if blsgpu.init() != 0:
        raise RuntimeError("no GPU: " + blsgpu.last_error())

    n = args.batch
    log("generating workload n=%d seed=%d" % (n, args.seed))
    t_gen = time.perf_counter()
    sk, pk, key_idx, cat, sig, nhm = make_workload(wl, n, args.seed)
    lib.blsgpu_reset_keys()
    handles = blsgpu.register_keys_affine(pk)
    t_gen = time.perf_counter() - t_gen

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
        if lib.blsgpu_autotune_device(d_sig.data_ptr(), d_nhm.data_ptr(), d_key.data_ptr(), n, 1, sp,
                                      cfg.ctypes.data_as(vp), ms.ctypes.data_as(vp)):
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
    lib.blsgpu_phase_times(phase.ctypes.data_as(vp), ctypes.byref(calls))
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
    want = expected_codes(cat)
    ok = bool((res == want).all())

    # roofline of the dominant kernel: algorithmic 32-bit limb products / duration
    phase_avg = phase / max(calls.value, 1)
    names = ["g1_decompress", "neg_hm_convert", "miller_loop", "final_exp"]
    decoded = int((want >= 0).sum())
    work = [n * FP_MULS_DECOMPRESS, 0, decoded * FP_MULS_MILLER, decoded * FP_MULS_FINAL_EXP]
    dom = int(np.argmax(phase_avg))
    mads = ctypes.c_double(0)
    pms = ctypes.c_double(0)
    wl.blsgpu_probe_mad_rate(ctypes.byref(mads), ctypes.byref(pms))
    achieved = work[dom] * PRODUCTS_PER_FP_MUL / (phase_avg[dom] * 1e-3)
    roofline = {
        "bound": "valu",
        "kernel": names[dom],
        "achieved": achieved / 1e9,
        "peak": mads.value / 1e9,
        "unit": "G(32x32-bit products)/s",
        "frac": achieved / mads.value,
        "traffic": None,
        "traffic_source": None,
        "peak_theoretical": MI355X_CUS * 64 * MI355X_GHZ,
        "kernel_ms": {k: round(float(v), 4) for k, v in zip(names, phase_avg)},
        "whole_job_frac": value * (FP_MULS_MILLER + FP_MULS_FINAL_EXP) * PRODUCTS_PER_FP_MUL / mads.value,
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
        "dtype": "u32 limbs (exact 381-bit integer)",
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
        out["wire"] = measure_wire(lib, wl, blsgpu, sk, handles, key_idx, args.seed + 6, 10)
        log("wire: %s" % {k: out["wire"][k] for k in ("verif_per_s", "hash_ms", "compressed_ms")})
    if "batched" in forms or "at_scale" in forms or "serving" in forms:
        blsgpu.set_batched_engine(12, 16, 0, 0, True)
    if "batched" in forms:
        log("batched form: generating the 256-batch workload")
        w3 = batched_workload(wl, args.seed + 3)
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
        w4 = at_scale_workload(wl, args.seed + 4)
        lib.blsgpu_reset_keys()
        h4 = blsgpu.register_keys_affine(w4["pk"])
    if "at_scale" in forms:
        out["at_scale"] = measure_batched(torch, lib, blsgpu, w4, h4, max(2, args.batched_steps // 2), 1,
                                          not args.no_tune, args.batched_engine)
        out["at_scale"]["workload"] = ("65536 verifications, 40 signer keys, 4096 messages, lognormal batch sizes "
                                       "(median 64, sigma 1.2, [1,8192]), 1/128 invalid")
        log("at_scale: %s" % {k: out["at_scale"][k] for k in ("batched_verif_per_s", "plain_verif_per_s", "speedup")})
    if "serving" in forms:
        pool = serving_pool(wl, args.seed + 5, w4["sk"])
        rates = [int(r) for r in args.serve_rates.split(",") if r]
        out["serving"] = {
            "workload": "Poisson arrivals, request sizes lognormal (median 16, sigma 1.0, [1,512]), 40 keys, fresh "
                        "messages (pool of %d items, 1/256 invalid), 4 GB HBM budget, %.0f s per run"
                        % (len(pool["key_idx"]), args.serve_seconds),
            "latency": "from each request's scheduled arrival to the host seeing its results; null = not finished "
                       "before the drain deadline",
            "runs": measure_serving(blsgpu, pool, h4, rates, args.serve_seconds, args.seed),
        }
    if forms - {"plain"}:
        out["forms_s"] = round(time.perf_counter() - t_forms, 1)
    if not args.no_cpu:
        log("cpu baseline (budget %.0f s)" % args.cpu_seconds)
        out["cpu_baseline"] = cpu_baseline(sig, nhm, pk, key_idx, res, args.cpu_seconds)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
