"""GPU parity: libgpu_msm.so on the device against the CPU oracle and the reference's
fixtures.  Bit-exact for every verdict, decoded coordinate, line table and GT element."""
import numpy as np
import pytest

import blsgpu
import oracle as O
from gpu_workload import make_workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    assert blsgpu.init() == 0, blsgpu.last_error()


def _fx():
    return O.fixtures()


def test_reference_verify_vector():
    # sk = 42: sig = 294 G1, -H(m) = -7 G1 (cuda/test_bls_verify.cu:112-172)
    fx = _fx()
    sig, nhm = O.u64(fx["sig_294_g1"])[None], O.u64(fx["neg_7_g1"])[None]
    g2c, pkc = O.u64(fx["g2_gen_coeffs"]), O.u64(fx["pk42_coeffs"])[None]
    print("case: reference verify vector", flush=True)
    res, gt = blsgpu.batch_verify(sig, nhm, g2c, pkc, want_gt=True)
    assert res.tolist() == [1]
    one = np.zeros(72, np.uint64)
    one[:6] = O.u64(fx["mont_one"])
    assert (gt[0] == one).all()
    # invalid: the generator as signature (gpu-crypto/src/lib.rs:226-242)
    res, gt = blsgpu.batch_verify(O.g1_gen()[None], nhm, g2c, pkc, want_gt=True)
    ok, ogt = O.verify(O.g1_gen(), nhm[0], g2c, pkc[0])
    assert res.tolist() == [0] and not ok
    assert (gt[0] == ogt).all()


def test_legacy_batch_matches_oracle():
    print("case: legacy batch M=40", flush=True)
    sks, pks, coeffs = O.make_keys(11, 5)
    key_idx = np.arange(40) % 5
    inv = {3: "wrong_msg", 17: "wrong_msg", 30: "non_subgroup"}
    comp, unc, nhm, _ = O.make_batch(11, 40, sks, key_idx, inv)
    g2c = O.g2_prepare(O.g2_gen())
    pkc = coeffs[key_idx]
    res, gt = blsgpu.batch_verify(unc, nhm, g2c, pkc, want_gt=True)
    ores, ogt = O.verify_batch(unc, nhm, g2c, np.ascontiguousarray(pkc))
    assert (res == ores).all()
    assert (gt == ogt).all()
    assert res.sum() == 37


def test_key_registration_lines_match_reference_fixtures():
    print("case: key registration", flush=True)
    fx = _fx()
    h = blsgpu.register_key_affine(O.g2_gen())
    assert (blsgpu.key_coeffs(h) == O.u64(fx["g2_gen_coeffs"])).all()
    q42, _ = O.g2_mul(O.g2_gen(), 42)
    h = blsgpu.register_key_affine(q42)
    assert (blsgpu.key_coeffs(h) == O.u64(fx["pk42_coeffs"])).all()
    h2 = blsgpu.register_key_coeffs(O.u64(fx["pk42_coeffs"]))
    assert (blsgpu.key_coeffs(h2) == O.u64(fx["pk42_coeffs"])).all()


def test_decompression_matches_oracle():
    print("case: decompression", flush=True)
    rng = np.random.default_rng(5)
    pts = []
    for i in range(24):
        p, _ = O.g1_mul(O.g1_gen(), int(rng.integers(1, 1 << 62)) * 7919 + i)
        pts.append(O.g1_compress(p))
    # rejects of every class
    x = O.unmont(O.rand_x(3, 0, 0)); b = bytearray(x.to_bytes(48, "big")); b[0] |= 0x80
    pts.append(np.frombuffer(bytes(b), np.uint8))                      # no sqrt
    x = O.unmont(O.rand_x(3, 1, 1)); b = bytearray(x.to_bytes(48, "big")); b[0] |= 0xa0
    pts.append(np.frombuffer(bytes(b), np.uint8))                      # off subgroup
    b = bytearray(O.P.to_bytes(48, "big")); b[0] |= 0x80
    pts.append(np.frombuffer(bytes(b), np.uint8))                      # x >= p
    pts.append(np.zeros(48, np.uint8))                                 # no compression flag
    ident = np.zeros(48, np.uint8); ident[0] = 0xc0; pts.append(ident)  # identity
    bad_ident = ident.copy(); bad_ident[0] = 0xe0; pts.append(bad_ident)
    comp = np.array(pts)
    out, st = blsgpu.decompress_status(comp)
    for i in range(len(pts)):
        rc, q = O.g1_decompress(comp[i])
        assert st[i] == rc, i
        if rc == 0:
            assert (out[i] == q).all(), i
    # drop-in entry: all-valid batch succeeds, any invalid point makes it return -1
    lib = blsgpu.load()
    good = np.ascontiguousarray(comp[:24])
    o = np.zeros((24, 12), np.uint64)
    assert lib.gpu_g1_decompress_batch(O.ptr(good), O.ptr(o), 24) == 0
    assert (o == out[:24]).all()
    o2 = np.zeros((len(pts), 12), np.uint64)
    assert lib.gpu_g1_decompress_batch(O.ptr(np.ascontiguousarray(comp)), O.ptr(o2), len(pts)) == -1


@pytest.mark.parametrize("n", [16384, 16385])
def test_subgroup_test_at_its_layout_switch(n):
    """The decompression entry point at the batch size where the subgroup test changes
    layout (four lanes per point up to 16,384 points, one lane per point beyond): every
    status against the categories the device generator planted (no square root, off the
    subgroup; wrong-message signatures are valid points), with off-subgroup encodings also
    first and last in the batch, the coordinates equal to the unchecked decoding's, and a
    sample against the oracle."""
    print("case: subgroup layouts, n=%d" % n, flush=True)
    w = make_workload(17, 3, n, 32, invalid_every=61)
    comp, cat = w["sig"].copy(), w["cat"].copy()
    off = int(np.nonzero(cat == 3)[0][0])
    for j in (0, n - 1):
        comp[j], cat[j] = comp[off], 3
    want = np.select([cat == 2, cat == 3], [-3, -4], 0)
    out, st = blsgpu.decompress_status(comp)
    assert (st == want).all(), np.nonzero(st != want)
    out_u, st_u = blsgpu.decompress_status(comp, check_subgroup=False)
    assert (st_u == np.where(cat == 2, -3, 0)).all()
    ok = st == 0
    assert (out[ok] == out_u[ok]).all()
    for i in sorted(set(list(range(0, n, 1021)) + [off, n - 2, n - 1])):
        rc, q = O.g1_decompress(comp[i])
        assert st[i] == rc, i
        if rc == 0:
            assert (out[i] == q).all(), i


def test_fused_compressed_path_matches_oracle():
    print("case: fused compressed M=64", flush=True)
    blsgpu.load().blsgpu_reset_keys()
    sks, pks, coeffs = O.make_keys(21, 4)
    handles = [blsgpu.register_key_affine(pk) for pk in pks]
    key_idx = np.array([i % 4 for i in range(64)], np.int32)
    inv = {5: "wrong_msg", 9: "no_sqrt", 22: "non_subgroup", 40: "wrong_msg"}
    comp, unc, nhm, _ = O.make_batch(21, 64, sks, key_idx, inv)
    res, gt = blsgpu.verify_compressed(comp, nhm, np.array(handles, np.int32)[key_idx], want_gt=True)
    g2c = O.g2_prepare(O.g2_gen())
    for i in range(64):
        rc, _ = O.g1_decompress(comp[i])
        if rc != 0:
            assert res[i] == rc and (gt[i] == 0).all(), i
            continue
        ok, ogt = O.verify(unc[i], nhm[i], g2c, coeffs[key_idx[i]])
        assert res[i] == int(ok), i
        assert (gt[i] == ogt).all(), i
    assert (res == 1).sum() == 60
