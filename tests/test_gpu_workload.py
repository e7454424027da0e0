"""The device workload generator (include/blsgpu_workload.h) against the oracle's
generator: the same seeds give bit-identical keys, signatures and -H(m)."""
import ctypes

import numpy as np
import pytest

import blsgpu
import oracle as O
from blsgpu import workload

pytestmark = pytest.mark.gpu


def _wl():
    return workload.lib()


def test_generator_matches_oracle_streams():
    print("case: workload generator seed 31 n=12", flush=True)
    assert blsgpu.init() == 0
    lib = _wl()
    seed, nkeys, n = 31, 3, 12
    sk = np.zeros((nkeys, 32), np.uint8)
    lib.blsgpu_wl_secret_keys(ctypes.c_uint64(seed), nkeys, O.ptr(sk))
    osks, opks, _ = O.make_keys(seed, nkeys)
    assert [int.from_bytes(s.tobytes(), "little") for s in sk] == osks
    pk = np.zeros((nkeys, 24), np.uint64)
    assert lib.blsgpu_wl_public_keys(O.ptr(sk), nkeys, O.ptr(pk)) == 0, lib.blsgpu_wl_last_error()
    assert (pk == opks).all()
    key_idx = (np.arange(n) % nkeys).astype(np.int32)
    cat = np.array([0, 1, 0, 2, 3, 0, 1, 0, 0, 3, 2, 0], np.int8)
    sig = np.zeros((n, 48), np.uint8)
    nhm = np.zeros((n, 12), np.uint64)
    assert lib.blsgpu_wl_items(ctypes.c_uint64(seed), n, O.ptr(sk), nkeys, O.ptr(key_idx), O.ptr(cat), O.ptr(sig),
                               O.ptr(nhm)) == 0, lib.blsgpu_wl_last_error()
    names = {1: "wrong_msg", 2: "no_sqrt", 3: "non_subgroup"}
    inv = {i: names[int(c)] for i, c in enumerate(cat) if c}
    comp, _, onhm, _ = O.make_batch(seed, n, osks, key_idx, inv)
    assert (sig == comp).all()
    assert (nhm == onhm).all()
    # and the fused path classifies them as the categories say
    blsgpu.load().blsgpu_reset_keys()
    hs = [blsgpu.register_key_affine(p) for p in pk]
    res = blsgpu.verify_compressed(sig, nhm, np.array(hs, np.int32)[key_idx])
    want = np.select([cat == 0, cat == 1, cat == 2, cat == 3], [1, 0, -3, -4])
    assert (res == want).all()
