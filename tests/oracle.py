"""Test-side loader for the CPU oracle (oracle/_build/liboracle.so) plus workload
helpers built on it.  Only tests, the smoke check and the benchmark's CPU-baseline leg
use this module, as the checker."""
import ctypes
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
FIXTURES = os.path.join(ROOT, "tests", "golden", "fixtures.json")

P = 0x1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f6241eabfffeb153ffffb9feffffffffaaab
R_ORDER = 0x73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001
R384 = 1 << 384

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = ctypes.CDLL(ORACLE_PATH)
        for name in dir(_lib):
            pass
        _lib.orc_num_threads.restype = ctypes.c_int
    return _lib


def fixtures():
    with open(FIXTURES) as f:
        return json.load(f)


def u64(hexlist):
    return np.array([int(h, 16) for h in hexlist], dtype=np.uint64)


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# ---- integer <-> limb helpers ----
def to_int(limbs):
    return sum(int(v) << (64 * i) for i, v in enumerate(limbs))


def from_int(x, n=6):
    return np.array([(x >> (64 * i)) & ((1 << 64) - 1) for i in range(n)], dtype=np.uint64)


def mont(x):
    """canonical integer -> reference Montgomery limbs"""
    return from_int(x * R384 % P)


def unmont(limbs):
    return to_int(limbs) * pow(R384, -1, P) % P


def scalar_bytes(k):
    return np.frombuffer(int(k).to_bytes(32, "little"), dtype=np.uint8).copy()


# ---- oracle wrappers ----
def g1_gen():
    o = np.zeros(12, np.uint64)
    lib().orc_g1_generator(ptr(o))
    return o


def g2_gen():
    o = np.zeros(24, np.uint64)
    lib().orc_g2_generator(ptr(o))
    return o


def g1_mul(p, k):
    o = np.zeros(12, np.uint64)
    inf = lib().orc_g1_mul(ptr(np.ascontiguousarray(p)), ptr(scalar_bytes(k)), ptr(o))
    return o, bool(inf)


def g2_mul(q, k):
    o = np.zeros(24, np.uint64)
    inf = lib().orc_g2_mul(ptr(np.ascontiguousarray(q)), ptr(scalar_bytes(k)), ptr(o))
    return o, bool(inf)


def g1_neg(p):
    o = np.zeros(12, np.uint64)
    lib().orc_g1_neg(ptr(np.ascontiguousarray(p)), ptr(o))
    return o


def g2_prepare(q):
    o = np.zeros(2448, np.uint64)
    lib().orc_g2_prepare(ptr(np.ascontiguousarray(q)), ptr(o))
    return o


def g1_compress(p, inf=False):
    o = np.zeros(48, np.uint8)
    lib().orc_g1_compress(ptr(np.ascontiguousarray(p)), int(inf), ptr(o))
    return o


def g1_decompress(b, check=True):
    o = np.zeros(12, np.uint64)
    f = lib().orc_g1_decompress if check else lib().orc_g1_decompress_unchecked
    rc = f(ptr(np.ascontiguousarray(b, dtype=np.uint8)), ptr(o))
    return rc, o


def g2_compress(q, inf=False):
    o = np.zeros(96, np.uint8)
    lib().orc_g2_compress(ptr(np.ascontiguousarray(q)), int(inf), ptr(o))
    return o


def g2_decompress(b, check=True):
    o = np.zeros(24, np.uint64)
    f = lib().orc_g2_decompress if check else lib().orc_g2_decompress_unchecked
    rc = f(ptr(np.ascontiguousarray(b, dtype=np.uint8)), ptr(o))
    return rc, o


def g2_is_on_curve(q):
    return bool(lib().orc_g2_is_on_curve(ptr(np.ascontiguousarray(q))))


def g2_in_subgroup(q):
    return bool(lib().orc_g2_in_subgroup(ptr(np.ascontiguousarray(q))))


def g2_off_subgroup_point(seed, idx):
    """an on-curve G2 point outside the r-subgroup: x from seeded bytes until
    x^3 + 4(1 + u) is a square (the cofactor is ~2^381, so such a point misses G2)"""
    k = 0
    while True:
        xb = bytearray(rand_bytes(seed, 3_000_000 + 1000 * idx + k, 96))
        xb[0] &= 0x0F
        xb[48] &= 0x0F
        xb[0] |= 0x80
        rc, q = g2_decompress(np.frombuffer(bytes(xb), np.uint8), check=False)
        if rc == 0 and not g2_in_subgroup(q):
            return bytes(xb), q
        k += 1


def hash_to_g1(msg):
    m = np.frombuffer(bytes(msg), dtype=np.uint8).copy()
    o = np.zeros(12, np.uint64)
    lib().orc_hash_to_g1(ptr(m), ctypes.c_size_t(len(msg)), ptr(o))
    return o


def rand_scalar(seed, idx):
    o = np.zeros(32, np.uint8)
    lib().orc_rand_scalar(ctypes.c_uint64(seed), ctypes.c_uint64(idx), ptr(o))
    return int.from_bytes(o.tobytes(), "little")


def rand_bytes(seed, idx, n):
    o = np.zeros(n, np.uint8)
    lib().orc_rand_bytes(ctypes.c_uint64(seed), ctypes.c_uint64(idx), ptr(o), ctypes.c_size_t(n))
    return o.tobytes()


def rand_x(seed, idx, square):
    o = np.zeros(6, np.uint64)
    lib().orc_rand_x(ctypes.c_uint64(seed), ctypes.c_uint64(idx), int(square), ptr(o))
    return o


def verify(sig, nhm, g2c, pkc):
    gt = np.zeros(72, np.uint64)
    ok = lib().orc_verify(ptr(np.ascontiguousarray(sig)), ptr(np.ascontiguousarray(nhm)),
                          ptr(np.ascontiguousarray(g2c)), ptr(np.ascontiguousarray(pkc)), ptr(gt))
    return bool(ok), gt


def verify_batch(sigs, nhms, g2c, pkc, threads=0, want_gt=True):
    m = sigs.shape[0]
    res = np.zeros(m, np.int32)
    gts = np.zeros((m, 72), np.uint64) if want_gt else None
    lib().orc_verify_batch(ptr(sigs), ptr(nhms), ptr(g2c), ptr(pkc), ptr(res), ptr(gts), m, threads)
    return res, gts


def verify_batch_keyed(sigs, nhms, g2c, key_coeffs, key_idx, threads=0, want_gt=True):
    m = sigs.shape[0]
    res = np.zeros(m, np.int32)
    gts = np.zeros((m, 72), np.uint64) if want_gt else None
    key_idx = np.ascontiguousarray(key_idx, dtype=np.int32)
    lib().orc_verify_batch_keyed(ptr(sigs), ptr(nhms), ptr(g2c), ptr(key_coeffs), ptr(key_idx), ptr(res),
                                 ptr(gts), m, threads)
    return res, gts


def verify_compressed_batch(sig48, nhms, g2c, key_coeffs, key_idx, threads=0, want_gt=True):
    """The oracle's fused path: decode each 48-byte signature with every check, then
    verify against the keyed lines.  Results 1/0 or the negative decode code; GT of
    each item (zeros for a rejected encoding; e(-H(m), pk) for an identity signature)."""
    m = sig48.shape[0]
    res = np.zeros(m, np.int32)
    gts = np.zeros((m, 72), np.uint64) if want_gt else None
    sig48 = np.ascontiguousarray(sig48, dtype=np.uint8)
    nhms = np.ascontiguousarray(nhms, dtype=np.uint64)
    key_coeffs = np.ascontiguousarray(key_coeffs, dtype=np.uint64)
    key_idx = np.ascontiguousarray(key_idx, dtype=np.int32)
    lib().orc_verify_compressed_batch(ptr(sig48), ptr(nhms), ptr(np.ascontiguousarray(g2c)), ptr(key_coeffs),
                                      ptr(key_idx), ptr(res), ptr(gts), m, threads)
    return res, gts


IDENTITY_G1 = np.array([0xC0] + [0] * 47, np.uint8)  # compressed encoding of the G1 identity


# ---- seeded workload generation (the configs' shapes, small sizes for tests) ----
def make_keys(seed, n):
    """n key pairs: sk, pk (affine G2, 24 u64) and pk lines (2448 u64)."""
    g2 = g2_gen()
    sks, pks, coeffs = [], [], []
    for i in range(n):
        sk = rand_scalar(seed, i)
        pk, _ = g2_mul(g2, sk)
        sks.append(sk)
        pks.append(pk)
        coeffs.append(g2_prepare(pk))
    return sks, np.array(pks), np.array(coeffs)


def sign(sk, msg):
    h = hash_to_g1(msg)
    s, _ = g1_mul(h, sk)
    return s, h


def make_batch(seed, n_items, sks, key_idx, invalid=None, msg_idx=None):
    """Signatures over fresh 32-byte messages.  invalid: dict index -> category in
    {'wrong_msg', 'no_sqrt', 'non_subgroup'}.  Returns compressed sigs (n,48), uncompressed
    sigs (n,12) (zeros where the encoding is undecodable), neg_hms (n,12), messages.
    msg_idx: message index of each item (default: the item index)."""
    invalid = invalid or {}
    comp = np.zeros((n_items, 48), np.uint8)
    unc = np.zeros((n_items, 12), np.uint64)
    nhm = np.zeros((n_items, 12), np.uint64)
    msgs = []
    for i in range(n_items):
        msg = rand_bytes(seed, 1_000_000 + (i if msg_idx is None else int(msg_idx[i])), 32)
        msgs.append(msg)
        sk = sks[key_idx[i]]
        cat = invalid.get(i)
        h = hash_to_g1(msg)
        nhm[i] = g1_neg(h)
        if cat == "wrong_msg":
            s, _ = sign(sk, rand_bytes(seed, 2_000_000 + i, 32))
            unc[i] = s
            comp[i] = g1_compress(s)
        elif cat == "no_sqrt":
            x = rand_x(seed, i, 0)
            xb = int(unmont(x)).to_bytes(48, "big")
            comp[i] = np.frombuffer(xb, np.uint8)
            comp[i, 0] |= 0x80
        elif cat == "non_subgroup":
            x = rand_x(seed, i, 1)
            xb = bytearray(int(unmont(x)).to_bytes(48, "big"))
            xb[0] |= 0x80
            rc, pt = g1_decompress(np.frombuffer(bytes(xb), np.uint8), check=False)
            assert rc == 0
            unc[i] = pt
            comp[i] = np.frombuffer(bytes(xb), np.uint8)
        else:
            s, _ = sign(sk, msg)
            unc[i] = s
            comp[i] = g1_compress(s)
    return comp, unc, nhm, msgs
