"""Test-side helpers around the device workload generator and the oracle: seeded
workloads of the batching and serving shapes, and the oracle's expected verdicts and GT
elements for them."""
import ctypes

import numpy as np

import blsgpu
import oracle as O
from blsgpu import workload


def wl_lib():
    return workload.lib()


def make_workload(seed, nkeys, n, n_msgs, invalid_every=0):
    """n items over nkeys keys and n_msgs shared messages, generated on the device.
    Every invalid_every-th item (if > 0) is invalid, cycling through the three
    categories.  Returns dict(sk, pk, key_idx, msg_idx, cat, sig, nhm)."""
    lib = wl_lib()
    rng = np.random.default_rng(seed)
    sk = np.zeros((nkeys, 32), np.uint8)
    lib.blsgpu_wl_secret_keys(ctypes.c_uint64(seed), nkeys, O.ptr(sk))
    pk = np.zeros((nkeys, 24), np.uint64)
    assert lib.blsgpu_wl_public_keys(O.ptr(sk), nkeys, O.ptr(pk)) == 0, lib.blsgpu_wl_last_error()
    key_idx = rng.integers(0, nkeys, n).astype(np.int32)
    msg_idx = rng.integers(0, n_msgs, n).astype(np.int64)
    cat = np.zeros(n, np.int8)
    if invalid_every:
        bad = np.arange(invalid_every // 2, n, invalid_every)
        cat[bad] = 1 + (np.arange(len(bad)) % 3)
    sig = np.zeros((n, 48), np.uint8)
    nhm = np.zeros((n, 12), np.uint64)
    assert lib.blsgpu_wl_items_msg(ctypes.c_uint64(seed), n, O.ptr(sk), nkeys, O.ptr(key_idx), O.ptr(msg_idx),
                                   O.ptr(cat), O.ptr(sig), O.ptr(nhm)) == 0, lib.blsgpu_wl_last_error()
    return dict(sk=sk, pk=pk, key_idx=key_idx, msg_idx=msg_idx, cat=cat, sig=sig, nhm=nhm)


def oracle_expected(sig48, nhm, pks, key_idx, threads=0):
    """Verdicts (1/0 or the negative decode code) and GT elements of the oracle's fused
    path (zeros for rejected encodings; an identity signature is paired as the identity,
    so its GT is e(-H(m), pk) and its verdict 0)."""
    coeffs = np.array([O.g2_prepare(p) for p in pks])
    g2c = O.g2_prepare(O.g2_gen())
    return O.verify_compressed_batch(sig48, nhm, g2c, coeffs, key_idx, threads=threads)
