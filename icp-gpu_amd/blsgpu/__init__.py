"""ctypes binding of libgpu_msm.so (the C ABI declared in include/gpu_msm.h).

This is the Python counterpart of the reference's Rust FFI crate
(ic-integration/gpu_crypto/lib.rs:10-28, 92-122): it loads the shared library the
same way a replica would, checks that the symbols the replica's dlopen requires are
present, and exposes numpy-friendly wrappers.  Loading fails loudly when the library
is missing; there is no CPU fallback here (the caller owns that decision, as in the
reference's verify_bls_signature_batch_gpu).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libgpu_msm.so")

# Symbols the reference's callers bind (lib.rs:10-28 statically, lib.rs:104-116 by dlsym)
REFERENCE_SYMBOLS = (
    "gpu_crypto_init",
    "gpu_batch_bls_verify",
    "gpu_batch_sha256",
    "gpu_msm_g1",
    "gpu_jacobian_to_affine",
    "gpu_g1_decompress_batch",
    "gpu_g1_decompress",
)
EXTENSION_SYMBOLS = (
    "blsgpu_batch_verify_gt",
    "blsgpu_register_key_affine",
    "blsgpu_register_key_coeffs",
    "blsgpu_key_coeffs",
    "blsgpu_num_keys",
    "blsgpu_reset_keys",
    "blsgpu_verify_compressed",
    "blsgpu_g1_decompress_status",
    "blsgpu_last_error",
    "blsgpu_set_engine",
    "blsgpu_get_engine",
    "blsgpu_verify_compressed_device",
    "blsgpu_verify_compressed_device_noalloc",
    "blsgpu_set_profiling",
    "blsgpu_phase_times",
    "blsgpu_legacy_phase_times",
    "blsgpu_register_keys_affine",
    "blsgpu_register_keys_compressed",
    "blsgpu_hash_to_g1",
    "blsgpu_verify_messages",
    "blsgpu_verify_batches",
    "blsgpu_verify_batches_device",
    "blsgpu_set_batched_engine",
    "blsgpu_get_batched_engine",
    "blsgpu_batched_stats",
    "blsgpu_set_line_depth",
    "blsgpu_get_line_depth",
    "blsgpu_autotune_device",
    "blsgpu_server_create",
    "blsgpu_server_step_capacity",
    "blsgpu_server_submit",
    "blsgpu_server_poll",
    "blsgpu_server_wait",
    "blsgpu_server_stats",
    "blsgpu_server_destroy",
    "blsgpu_server_loadtest",
)

VALID, INVALID = 1, 0
ERR_BAD_FLAGS, ERR_X_NOT_CANONICAL, ERR_NOT_ON_CURVE, ERR_NOT_IN_SUBGROUP = -1, -2, -3, -4

_lib = None


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def load(path=None):
    """Load libgpu_msm.so once; raise if it or any required symbol is missing."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("BLSGPU_LIB", LIB_PATH)
    if not os.path.exists(path):
        raise RuntimeError("libgpu_msm.so not built: %s (run __graft_entry__.build())" % path)
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    for name in REFERENCE_SYMBOLS + EXTENSION_SYMBOLS:
        if not hasattr(lib, name):
            raise RuntimeError("libgpu_msm.so lacks symbol %s" % name)
    vp, i32, u64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64
    i64, dbl = ctypes.c_int64, ctypes.c_double
    sig = {
        "gpu_crypto_init": ([], i32),
        "gpu_batch_bls_verify": ([vp, vp, vp, vp, vp, i32], i32),
        "gpu_batch_sha256": ([vp, vp, i32, i32], i32),
        "gpu_msm_g1": ([vp, vp, ctypes.c_size_t, vp], i32),
        "gpu_jacobian_to_affine": ([vp, vp], i32),
        "gpu_g1_decompress_batch": ([vp, vp, i32], i32),
        "gpu_g1_decompress": ([vp, vp], i32),
        "blsgpu_batch_verify_gt": ([vp, vp, vp, vp, vp, vp, i32], i32),
        "blsgpu_register_key_affine": ([vp], i32),
        "blsgpu_register_key_coeffs": ([vp], i32),
        "blsgpu_key_coeffs": ([i32, vp], i32),
        "blsgpu_num_keys": ([], i32),
        "blsgpu_reset_keys": ([], None),
        "blsgpu_verify_compressed": ([vp, vp, vp, vp, vp, i32], i32),
        "blsgpu_g1_decompress_status": ([vp, vp, vp, i32, i32], i32),
        "blsgpu_last_error": ([], ctypes.c_char_p),
        "blsgpu_set_engine": ([i32, i32, i32], i32),
        "blsgpu_get_engine": ([vp, vp, vp], i32),
        "blsgpu_verify_compressed_device": ([vp, vp, vp, vp, vp, i32, vp], i32),
        "blsgpu_verify_compressed_device_noalloc": ([vp, vp, vp, vp, vp, i32, vp], i32),
        "blsgpu_set_profiling": ([i32], i32),
        "blsgpu_phase_times": ([vp, vp], i32),
        "blsgpu_legacy_phase_times": ([vp, vp], i32),
        "blsgpu_register_keys_affine": ([vp, i32, vp], i32),
        "blsgpu_register_keys_compressed": ([vp, i32, vp, vp, i32], i32),
        "blsgpu_hash_to_g1": ([vp, vp, i32, vp, i32, vp, vp, i32], i32),
        "blsgpu_verify_messages": ([vp, vp, vp, vp, i32, vp, vp, vp, i32], i32),
        "blsgpu_verify_batches": ([vp, vp, vp, vp, i32, vp, vp], i32),
        "blsgpu_verify_batches_device": ([vp, vp, vp, vp, i32, vp, vp, vp], i32),
        "blsgpu_set_batched_engine": ([i32, i32, i32, i32, i32], i32),
        "blsgpu_get_batched_engine": ([i32, vp], i32),
        "blsgpu_batched_stats": ([vp], i32),
        "blsgpu_set_line_depth": ([i32], i32),
        "blsgpu_get_line_depth": ([i32, i32], i32),
        "blsgpu_autotune_device": ([vp, vp, vp, i32, i32, vp, vp, vp], i32),
        "blsgpu_server_create": ([i32, i32, i64, i32, i32], vp),
        "blsgpu_server_step_capacity": ([vp], i32),
        "blsgpu_server_submit": ([vp, vp, vp, vp, i32], i64),
        "blsgpu_server_poll": ([vp, i64, vp, vp, vp], i32),
        "blsgpu_server_wait": ([vp, i64, vp, vp, vp, dbl], i32),
        "blsgpu_server_stats": ([vp, vp], i32),
        "blsgpu_server_destroy": ([vp], None),
        "blsgpu_server_loadtest": ([vp, vp, vp, vp, i32, dbl, dbl, dbl, dbl, dbl, i32, i32, u64, vp], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def last_error():
    return load().blsgpu_last_error().decode()


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: %s" % (what, last_error()))


def init():
    return load().gpu_crypto_init()


def batch_verify(sigs, neg_hms, g2_coeffs, pk_coeffs, want_gt=False):
    """Legacy ABI: sigs, neg_hms (M,12) u64; g2_coeffs (2448,) u64; pk_coeffs (M,2448) u64."""
    sigs = np.ascontiguousarray(sigs, dtype=np.uint64)
    neg_hms = np.ascontiguousarray(neg_hms, dtype=np.uint64)
    g2_coeffs = np.ascontiguousarray(g2_coeffs, dtype=np.uint64)
    pk_coeffs = np.ascontiguousarray(pk_coeffs, dtype=np.uint64)
    m = sigs.shape[0]
    res = np.zeros(m, dtype=np.int32)
    gt = np.zeros((m, 72), dtype=np.uint64) if want_gt else None
    rc = load().blsgpu_batch_verify_gt(_p(sigs), _p(neg_hms), _p(g2_coeffs), _p(pk_coeffs), _p(res), _p(gt), m)
    _check(rc, "blsgpu_batch_verify_gt")
    return (res, gt) if want_gt else res


def decompress_status(comp, check_subgroup=True):
    comp = np.ascontiguousarray(comp, dtype=np.uint8).reshape(-1, 48)
    n = comp.shape[0]
    out = np.zeros((n, 12), dtype=np.uint64)
    st = np.zeros(n, dtype=np.int32)
    _check(load().blsgpu_g1_decompress_status(_p(comp), _p(out), _p(st), n, int(check_subgroup)),
           "blsgpu_g1_decompress_status")
    return out, st


def register_key_affine(pk):
    pk = np.ascontiguousarray(pk, dtype=np.uint64)
    h = load().blsgpu_register_key_affine(_p(pk))
    if h < 0:
        raise RuntimeError("register_key_affine: " + last_error())
    return h


def register_key_coeffs(coeffs):
    coeffs = np.ascontiguousarray(coeffs, dtype=np.uint64)
    h = load().blsgpu_register_key_coeffs(_p(coeffs))
    if h < 0:
        raise RuntimeError("register_key_coeffs: " + last_error())
    return h


def key_coeffs(handle):
    out = np.zeros(2448, dtype=np.uint64)
    _check(load().blsgpu_key_coeffs(handle, _p(out)), "blsgpu_key_coeffs")
    return out


def verify_compressed(sigs48, neg_hms, key_idx, want_gt=False):
    sigs48 = np.ascontiguousarray(sigs48, dtype=np.uint8).reshape(-1, 48)
    neg_hms = np.ascontiguousarray(neg_hms, dtype=np.uint64)
    key_idx = np.ascontiguousarray(key_idx, dtype=np.int32)
    m = sigs48.shape[0]
    res = np.zeros(m, dtype=np.int32)
    gt = np.zeros((m, 72), dtype=np.uint64) if want_gt else None
    _check(load().blsgpu_verify_compressed(_p(sigs48), _p(neg_hms), _p(key_idx), _p(res), _p(gt), m),
           "blsgpu_verify_compressed")
    return (res, gt) if want_gt else res


def set_engine(mode, g_miller=16, g_final_exp=16):
    _check(load().blsgpu_set_engine(mode, g_miller, g_final_exp), "blsgpu_set_engine")


def get_engine():
    m, a, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    load().blsgpu_get_engine(ctypes.byref(m), ctypes.byref(a), ctypes.byref(b))
    return m.value, a.value, b.value


def register_keys_affine(pks):
    """Registers n keys (n, 24) u64 in one launch; returns their handles."""
    pks = np.ascontiguousarray(pks, dtype=np.uint64).reshape(-1, 24)
    h = np.zeros(pks.shape[0], dtype=np.int32)
    if load().blsgpu_register_keys_affine(_p(pks), pks.shape[0], _p(h)) < 0:
        raise RuntimeError("register_keys_affine: " + last_error())
    return h


def register_keys_compressed(pks96, check_subgroup=True):
    """Decompresses, validates and registers n compressed G2 keys ((n, 96) u8) on the
    GPU.  Returns (handles, status): handle -1 and status 1 (identity) or a negative
    decode code for keys that were not registered."""
    pks96 = np.ascontiguousarray(pks96, dtype=np.uint8).reshape(-1, 96)
    n = pks96.shape[0]
    h = np.zeros(n, dtype=np.int32)
    st = np.zeros(n, dtype=np.int32)
    if load().blsgpu_register_keys_compressed(_p(pks96), n, _p(h), _p(st), int(bool(check_subgroup))) < 0:
        raise RuntimeError("register_keys_compressed: " + last_error())
    return h, st


def _pack_messages(msgs):
    offs = np.zeros(len(msgs) + 1, dtype=np.int64)
    offs[1:] = np.cumsum([len(m) for m in msgs])
    buf = np.frombuffer(b"".join(bytes(m) for m in msgs) or b"\0", dtype=np.uint8).copy()
    return buf, offs


def hash_to_g1(msgs, dst, negate=False):
    """H(m) for each message (RFC 9380 BLS12381G1_XMD:SHA-256_SSWU_RO_ with `dst`) on the
    GPU: (n, 12) u64 affine points in the boundary form, and the identity flags."""
    buf, offs = _pack_messages(msgs)
    d = np.frombuffer(bytes(dst), dtype=np.uint8).copy()
    n = len(msgs)
    out = np.zeros((n, 12), dtype=np.uint64)
    ident = np.zeros(n, dtype=np.int32)
    _check(load().blsgpu_hash_to_g1(_p(buf), _p(offs), n, _p(d), len(dst), _p(out), _p(ident), int(bool(negate))),
           "blsgpu_hash_to_g1")
    return out, ident


def verify_messages(sigs48, msgs, dst, key_idx, want_gt=False):
    """Fused wire-format path: compressed signatures, messages (hashed on the GPU) and key
    handles; results as verify_compressed."""
    sigs48 = np.ascontiguousarray(sigs48, dtype=np.uint8).reshape(-1, 48)
    key_idx = np.ascontiguousarray(key_idx, dtype=np.int32)
    buf, offs = _pack_messages(msgs)
    d = np.frombuffer(bytes(dst), dtype=np.uint8).copy()
    m = sigs48.shape[0]
    assert len(msgs) == m and key_idx.shape[0] == m
    res = np.zeros(m, dtype=np.int32)
    gt = np.zeros((m, 72), dtype=np.uint64) if want_gt else None
    _check(load().blsgpu_verify_messages(_p(sigs48), _p(buf), _p(offs), _p(d), len(dst), _p(key_idx), _p(res),
                                         _p(gt), m), "blsgpu_verify_messages")
    return (res, gt) if want_gt else res


def verify_batches(batches, want_gt=False):
    """Batched form, host arrays.  batches: list of (sigs48 (n,48) u8, neg_hms (n,12) u64,
    key_idx (n,) i32).  Returns a list of per-batch results (and GT arrays)."""
    nb = len(batches)
    keep = []
    sp, hp, kp, rp, gp = [(ctypes.c_void_p * nb)() for _ in range(5)]
    sizes = np.zeros(nb, dtype=np.int32)
    outs, gts = [], []
    for b, (sig, nhm, kidx) in enumerate(batches):
        sig = np.ascontiguousarray(sig, dtype=np.uint8).reshape(-1, 48)
        nhm = np.ascontiguousarray(nhm, dtype=np.uint64).reshape(-1, 12)
        kidx = np.ascontiguousarray(kidx, dtype=np.int32)
        n = sig.shape[0]
        sizes[b] = n
        r = np.zeros(n, dtype=np.int32)
        g = np.zeros((n, 72), dtype=np.uint64) if want_gt else None
        keep += [sig, nhm, kidx]
        sp[b], hp[b], kp[b], rp[b] = sig.ctypes.data, nhm.ctypes.data, kidx.ctypes.data, r.ctypes.data
        if want_gt:
            gp[b] = g.ctypes.data
        outs.append(r)
        gts.append(g)
    _check(load().blsgpu_verify_batches(ctypes.cast(sp, ctypes.c_void_p), ctypes.cast(hp, ctypes.c_void_p),
                                        ctypes.cast(kp, ctypes.c_void_p), _p(sizes), nb,
                                        ctypes.cast(rp, ctypes.c_void_p),
                                        ctypes.cast(gp, ctypes.c_void_p) if want_gt else None),
           "blsgpu_verify_batches")
    return (outs, gts) if want_gt else outs


def set_batched_engine(g_miller=12, g_final_exp=16, wpc_miller=0, wpc_final_exp=0, use_graph=True):
    _check(load().blsgpu_set_batched_engine(g_miller, g_final_exp, wpc_miller, wpc_final_exp, int(use_graph)),
           "blsgpu_set_batched_engine")


def get_batched_engine(m=1):
    cfg = np.zeros(5, dtype=np.int32)
    load().blsgpu_get_batched_engine(m, _p(cfg))
    return tuple(int(v) for v in cfg)


def set_line_depth(depth):
    """force the Miller loop's key-line depth (0, 1, 2) on every form; -1 = autotuned/default"""
    _check(load().blsgpu_set_line_depth(int(depth)), "blsgpu_set_line_depth")


def get_line_depth(form, m):
    """key-line depth a call of the plain form (form=1) or batched form (form=0) of m items runs"""
    return load().blsgpu_get_line_depth(int(form), int(m))


def batched_stats():
    out = np.zeros(2, dtype=np.int32)
    load().blsgpu_batched_stats(_p(out))
    return int(out[0]), int(out[1])


class Server:
    """Serving loop (include/blsgpu_serve.h)."""

    def __init__(self, slots=3, max_step_items=8192, hbm_budget_bytes=4 << 30, max_requests_per_step=0,
                 want_gt=False):
        lib = load()
        self.want_gt = want_gt
        self.h = lib.blsgpu_server_create(slots, max_step_items, hbm_budget_bytes, max_requests_per_step,
                                          int(want_gt))
        if not self.h:
            raise RuntimeError("blsgpu_server_create: " + last_error())
        self.capacity = lib.blsgpu_server_step_capacity(self.h)

    def submit(self, sigs48, neg_hms, key_idx):
        sigs48 = np.ascontiguousarray(sigs48, dtype=np.uint8).reshape(-1, 48)
        neg_hms = np.ascontiguousarray(neg_hms, dtype=np.uint64).reshape(-1, 12)
        key_idx = np.ascontiguousarray(key_idx, dtype=np.int32)
        rid = load().blsgpu_server_submit(self.h, _p(sigs48), _p(neg_hms), _p(key_idx), sigs48.shape[0])
        if rid < 0:
            raise RuntimeError("blsgpu_server_submit refused the request")
        return rid, sigs48.shape[0]

    def wait(self, req, timeout_ms=-1.0):
        rid, n = req
        res = np.zeros(n, dtype=np.int32)
        gt = np.zeros((n, 72), dtype=np.uint64) if self.want_gt else None
        lat = ctypes.c_double()
        rc = load().blsgpu_server_wait(self.h, rid, _p(res), _p(gt), ctypes.byref(lat), timeout_ms)
        if rc != 1:
            raise RuntimeError("request %d: %s" % (rid, "time-out" if rc == 0 else "failed"))
        return res, gt, lat.value

    def stats(self):
        out = np.zeros(8, dtype=np.float64)
        load().blsgpu_server_stats(self.h, _p(out))
        keys = ("steps", "items", "completed", "max_step", "max_inflight", "capacity", "slots", "failed")
        return dict(zip(keys, out.tolist()))

    def loadtest(self, pool_sigs48, pool_neg_hms, pool_key_idx, items_per_s, seconds, drain_s=2.0,
                 size_median=16.0, size_sigma=1.0, size_min=1, size_max=512, seed=7):
        pool_sigs48 = np.ascontiguousarray(pool_sigs48, dtype=np.uint8).reshape(-1, 48)
        pool_neg_hms = np.ascontiguousarray(pool_neg_hms, dtype=np.uint64).reshape(-1, 12)
        pool_key_idx = np.ascontiguousarray(pool_key_idx, dtype=np.int32)
        out = np.zeros(11, dtype=np.float64)
        _check(load().blsgpu_server_loadtest(self.h, _p(pool_sigs48), _p(pool_neg_hms), _p(pool_key_idx),
                                             pool_sigs48.shape[0], items_per_s, seconds, drain_s, size_median,
                                             size_sigma, size_min, size_max, ctypes.c_uint64(seed), _p(out)),
               "blsgpu_server_loadtest")
        keys = ("requests", "items", "requests_done", "items_done", "span_s", "sustained_items_per_s",
                "p50_ms", "p99_ms", "p999_ms", "max_ms", "offered_items_per_s")
        return dict(zip(keys, out.tolist()))

    def close(self):
        if self.h:
            load().blsgpu_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
