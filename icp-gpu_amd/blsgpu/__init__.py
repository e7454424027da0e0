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
    "blsgpu_set_profiling",
    "blsgpu_phase_times",
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
        "blsgpu_set_profiling": ([i32], i32),
        "blsgpu_phase_times": ([vp, vp], i32),
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
