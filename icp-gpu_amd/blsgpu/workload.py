"""Seeded synthetic workloads of the benchmark configurations, generated on the device
by libblsgpu_workload.so (declared in csrc/blsgpu_workload.h).

That library is bench and test infrastructure: it is built beside libgpu_msm.so but is
not part of it, so a replica that loads the verifier loads no key generator or signer.
The streams match the test oracle's generator (same seed and index, same bytes).

Configurations (keys pk = sk G2, H(m) = s G1 with s from SHA-256 of the message as in
the reference's oracle/src/main.rs:383-397, signatures sk H(m)):
  uniform     n items, n distinct keys and messages; 64 invalid (n >= 128): 32 over
              another message, 16 x without a square root, 16 off the r-order subgroup
  batched     256 batches, geometric sizes (mean 128, [1, 2048]), 13 signers, 1024
              messages each signed by a random 1..13 signers, 1/128 invalid
  at_scale    65536 items over 40 keys and 4096 messages, lognormal batch sizes
              (median 64, sigma 1.2, [1, 8192]), 1/128 invalid
  serving     a pool of items over given keys with fresh messages, 1/256 invalid
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WORKLOAD_LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libblsgpu_workload.so")

WRONG_MSG, NO_SQRT, NON_SUBGROUP = 1, 2, 3

_wl = None


def lib():
    """libblsgpu_workload.so, loaded once; raises if it is not built."""
    global _wl
    if _wl is None:
        if not os.path.exists(WORKLOAD_LIB_PATH):
            raise RuntimeError("libblsgpu_workload.so not built: %s" % WORKLOAD_LIB_PATH)
        w = ctypes.CDLL(WORKLOAD_LIB_PATH)
        vp, i32, u64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64
        w.blsgpu_wl_secret_keys.argtypes = [u64, i32, vp]
        w.blsgpu_wl_secret_keys.restype = None
        w.blsgpu_wl_message.argtypes = [u64, ctypes.c_int64, vp]
        w.blsgpu_wl_message.restype = None
        w.blsgpu_wl_public_keys.argtypes = [vp, i32, vp]
        w.blsgpu_wl_items.argtypes = [u64, i32, vp, i32, vp, vp, vp, vp]
        w.blsgpu_wl_items_msg.argtypes = [u64, i32, vp, i32, vp, vp, vp, vp, vp]
        w.blsgpu_wl_sign_points.argtypes = [vp, i32, vp, i32, vp, vp]
        w.blsgpu_probe_mad_rate.argtypes = [vp, vp]
        w.blsgpu_wl_last_error.restype = ctypes.c_char_p
        _wl = w
    return _wl


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _err(what):
    return RuntimeError("%s: %s" % (what, lib().blsgpu_wl_last_error().decode()))


def expected_codes(cat):
    """the fused path's result for each category: 1 valid, 0 wrong message, -3 no
    square root, -4 off the subgroup"""
    return np.select([cat == 0, cat == 1, cat == 2, cat == 3], [1, 0, -3, -4]).astype(np.int32)


def gen_keys(seed, nkeys):
    """nkeys secret keys (32 bytes LE) and public keys (24 u64, reference form)"""
    sk = np.zeros((nkeys, 32), np.uint8)
    lib().blsgpu_wl_secret_keys(seed, nkeys, _p(sk))
    pk = np.zeros((nkeys, 24), np.uint64)
    if lib().blsgpu_wl_public_keys(_p(sk), nkeys, _p(pk)):
        raise _err("public keys")
    return sk, pk


def gen_items(seed, sk, key_idx, msg_idx, cat):
    """compressed signatures (n x 48) and -H(m) (n x 12 u64): item i signs message
    msg_idx[i] with key key_idx[i] unless its category says otherwise"""
    n = len(key_idx)
    sig = np.zeros((n, 48), np.uint8)
    nhm = np.zeros((n, 12), np.uint64)
    key_idx = np.ascontiguousarray(key_idx, np.int32)
    msg_idx = np.ascontiguousarray(msg_idx, np.int64)
    cat = np.ascontiguousarray(cat, np.int8)
    if lib().blsgpu_wl_items_msg(seed, n, _p(sk), sk.shape[0], _p(key_idx), _p(msg_idx), _p(cat), _p(sig), _p(nhm)):
        raise _err("items")
    return sig, nhm


def mixed_invalid(rng, n, every):
    """categories: n // every items invalid, cycling wrong-message, no-sqrt, off-subgroup"""
    cat = np.zeros(n, np.int8)
    bad = rng.choice(n, size=n // every, replace=False)
    cat[bad] = 1 + (np.arange(len(bad)) % 3)
    return cat


def uniform(n, seed):
    """n distinct keys, item i signed by key i over message i; 64 invalid when n >= 128
    (32 wrong message, 16 no square root, 16 off the subgroup), else n // 16"""
    rng = np.random.default_rng(seed)
    cat = np.zeros(n, np.int8)
    n_bad = 64 if n >= 128 else max(n // 16, 0)
    if n_bad:
        pos = rng.choice(n, size=n_bad, replace=False)
        half, q = n_bad // 2, n_bad // 4
        cat[pos[:half]] = WRONG_MSG
        cat[pos[half:half + q]] = NO_SQRT
        cat[pos[half + q:]] = NON_SUBGROUP
    sk, pk = gen_keys(seed, n)
    key_idx = np.arange(n, dtype=np.int32)
    sig = np.zeros((n, 48), np.uint8)
    nhm = np.zeros((n, 12), np.uint64)
    if lib().blsgpu_wl_items(seed, n, _p(sk), n, _p(key_idx), _p(cat), _p(sig), _p(nhm)):
        raise _err("items")
    return dict(sk=sk, pk=pk, key_idx=key_idx, cat=cat, sig=sig, nhm=nhm, sizes=np.array([n]))


def valid_only(n, seed):
    """n valid items over n distinct keys and messages (the CPU reference path's shape)"""
    sk, pk = gen_keys(seed, n)
    key_idx = np.arange(n, dtype=np.int32)
    cat = np.zeros(n, np.int8)
    sig, nhm = gen_items(seed, sk, key_idx, key_idx, cat)
    return dict(sk=sk, pk=pk, key_idx=key_idx, cat=cat, sig=sig, nhm=nhm, sizes=np.array([n]))


def batched(seed):
    """256 batches, geometric sizes (mean 128, [1, 2048]); 13 signers; 1024 messages each
    signed by a random subset of 1..13 signers; items drawn from those signatures; 1/128
    invalid"""
    rng = np.random.default_rng(seed)
    sk, pk = gen_keys(seed, 13)
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
    sig, nhm = gen_items(seed, sk, key_idx, msg_idx, cat)
    return dict(sk=sk, pk=pk, sizes=sizes, key_idx=key_idx.astype(np.int32), cat=cat, sig=sig, nhm=nhm,
                signatures=len(pool))


def at_scale(seed, total=65536):
    """40 keys, 4096 messages, lognormal batch sizes (median 64, sigma 1.2, [1, 8192])"""
    rng = np.random.default_rng(seed)
    sk, pk = gen_keys(seed, 40)
    sizes = []
    while sum(sizes) < total:
        sizes.append(int(np.clip(round(np.exp(np.log(64) + 1.2 * rng.normal())), 1, 8192)))
    sizes[-1] -= sum(sizes) - total
    sizes = np.array([x for x in sizes if x > 0])
    key_idx = rng.integers(0, 40, total).astype(np.int32)
    msg_idx = rng.integers(0, 4096, total)
    cat = mixed_invalid(rng, total, 128)
    sig, nhm = gen_items(seed, sk, key_idx, msg_idx, cat)
    return dict(sk=sk, pk=pk, sizes=sizes, key_idx=key_idx, cat=cat, sig=sig, nhm=nhm)


def serving_pool(seed, sk, n=65536):
    """n items over the given keys with fresh random messages, 1/256 invalid"""
    rng = np.random.default_rng(seed)
    key_idx = rng.integers(0, sk.shape[0], n).astype(np.int32)
    msg_idx = 5_000_000 + np.arange(n)
    cat = mixed_invalid(rng, n, 256)
    sig, nhm = gen_items(seed, sk, key_idx, msg_idx, cat)
    return dict(key_idx=key_idx, cat=cat, sig=sig, nhm=nhm)


def message(seed, index):
    """the 32-byte message of item `index`"""
    buf = np.zeros(32, np.uint8)
    lib().blsgpu_wl_message(seed, index, _p(buf))
    return buf.tobytes()


def sign_points(h_ref, sk, key_idx):
    """compressed signatures sk[key_idx[i]] * H_i of given G1 points (n x 12 u64)"""
    n = len(key_idx)
    sig = np.zeros((n, 48), np.uint8)
    kidx = np.ascontiguousarray(key_idx, np.int32)
    if lib().blsgpu_wl_sign_points(_p(np.ascontiguousarray(h_ref)), n, _p(sk), sk.shape[0], _p(kidx), _p(sig)):
        raise _err("sign")
    return sig


def probe_mad_rate():
    """measured v_mad_u64_u32 issue rate over all CUs (products/s) and the probe's ms"""
    mads, ms = ctypes.c_double(0), ctypes.c_double(0)
    if lib().blsgpu_probe_mad_rate(ctypes.byref(mads), ctypes.byref(ms)):
        raise _err("probe")
    return mads.value, ms.value


_P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
_R_INV = pow(1 << 384, -1, _P)


def _canon(limbs):
    return sum(int(v) << (64 * i) for i, v in enumerate(limbs)) * _R_INV % _P


def g2_compress(pks):
    """96-byte compressed encodings (ZCash format, as ic_bls12_381's G2Affine::
    to_compressed: x.c1 then x.c0 big-endian, compression flag, sort flag when y is the
    lexicographically largest root) of affine G2 points in reference form (n x 24 u64:
    x.c0 x.c1 y.c0 y.c1, Montgomery, R = 2^384)"""
    pks = np.asarray(pks, np.uint64).reshape(-1, 24)
    out = np.zeros((pks.shape[0], 96), np.uint8)
    half = (_P - 1) // 2
    for i, q in enumerate(pks):
        xc0, xc1, yc0, yc1 = (_canon(q[6 * k:6 * k + 6]) for k in range(4))
        b = bytearray(xc1.to_bytes(48, "big") + xc0.to_bytes(48, "big"))
        b[0] |= 0x80
        if yc1 > half or (yc1 == 0 and yc0 > half):
            b[0] |= 0x20
        out[i] = np.frombuffer(bytes(b), np.uint8)
    return out


def g2_generator():
    """the G2 generator in reference form (24 u64), as the public key of secret key 1"""
    sk = np.zeros((1, 32), np.uint8)
    sk[0, 0] = 1
    pk = np.zeros((1, 24), np.uint64)
    if lib().blsgpu_wl_public_keys(_p(sk), 1, _p(pk)):
        raise _err("public keys")
    return pk[0]
