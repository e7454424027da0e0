"""Every engine configuration (one thread per item; G lanes per item for the Miller loop
and the final exponentiation) against the oracle: verdicts and GT bit-exact, on the
legacy ABI and on the fused compressed path, at ragged batch sizes."""
import numpy as np
import pytest

import blsgpu
import oracle as O

pytestmark = pytest.mark.gpu

CONFIGS = [(1, 16, 16), (1, 8, 16), (1, 32, 32), (1, 12, 12), (1, 13, 18), (1, 20, 27), (1, 16, 9), (1, 24, 24), (0, 16, 16)]


@pytest.fixture(scope="module")
def data():
    assert blsgpu.init() == 0, blsgpu.last_error()
    sks, pks, coeffs = O.make_keys(41, 6)
    n = 37  # not a multiple of any group count
    key_idx = np.array([(i * 5) % 6 for i in range(n)], np.int32)
    inv = {2: "wrong_msg", 11: "no_sqrt", 19: "non_subgroup", 30: "wrong_msg"}
    comp, unc, nhm, _ = O.make_batch(41, n, sks, key_idx, inv)
    g2c = O.g2_prepare(O.g2_gen())
    want = []
    wgt = []
    for i in range(n):
        rc, _ = O.g1_decompress(comp[i])
        if rc != 0:
            want.append(rc)
            wgt.append(np.zeros(72, np.uint64))
        else:
            ok, gt = O.verify(unc[i], nhm[i], g2c, coeffs[key_idx[i]])
            want.append(int(ok))
            wgt.append(gt)
    return dict(pks=pks, coeffs=coeffs, key_idx=key_idx, comp=comp, unc=unc, nhm=nhm, g2c=g2c,
                want=np.array(want), wgt=np.array(wgt))


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "mode%d_ml%d_fe%d" % c)
def test_fused_path(cfg, data):
    print("case: fused path engine", cfg, flush=True)
    blsgpu.set_engine(*cfg)
    blsgpu.load().blsgpu_reset_keys()
    hs = np.array([blsgpu.register_key_affine(pk) for pk in data["pks"]], np.int32)
    res, gt = blsgpu.verify_compressed(data["comp"], data["nhm"], hs[data["key_idx"]], want_gt=True)
    assert (res == data["want"]).all(), (res, data["want"])
    assert (gt == data["wgt"]).all()


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "mode%d_ml%d_fe%d" % c)
def test_legacy_path(cfg, data):
    print("case: legacy path engine", cfg, flush=True)
    blsgpu.set_engine(*cfg)
    ok = data["want"] >= 0
    unc = np.ascontiguousarray(data["unc"][ok])
    nhm = np.ascontiguousarray(data["nhm"][ok])
    pkc = np.ascontiguousarray(data["coeffs"][data["key_idx"][ok]])
    res, gt = blsgpu.batch_verify(unc, nhm, data["g2c"], pkc, want_gt=True)
    # non-subgroup points are accepted by the legacy layout (no checks) and simply fail
    ores, ogt = O.verify_batch(unc, nhm, data["g2c"], pkc)
    assert (res == ores).all()
    assert (gt == ogt).all()


@pytest.mark.parametrize("cfg", [(1, 32, 12), (1, 16, 16)], ids=lambda c: "mode%d_ml%d_fe%d" % c)
def test_legacy_two_parts_odd_size(cfg):
    """1031 items (over the 1024 at which the lanes-per-item legacy sequence splits the
    batch): parts of 515 and 516 items, each one full and one partial chunk of the pinned
    ring; every verdict and GT element against the oracle."""
    print("case: legacy two parts, 1031 items, engine", cfg, flush=True)
    n = 1031
    sks, pks, coeffs = O.make_keys(47, 7)
    key_idx = np.array([(i * 3) % 7 for i in range(n)], np.int32)
    inv = {0: "wrong_msg", 514: "wrong_msg", 515: "non_subgroup", 700: "no_sqrt", 1030: "wrong_msg"}
    comp, unc, nhm, _ = O.make_batch(47, n, sks, key_idx, inv)
    g2c = O.g2_prepare(O.g2_gen())
    pkc = np.ascontiguousarray(coeffs[key_idx])
    ores, ogt = O.verify_batch(unc, nhm, g2c, pkc, threads=16)
    assert ores.sum() == n - len(inv)
    blsgpu.set_engine(*cfg)
    res, gt = blsgpu.batch_verify(unc, nhm, g2c, pkc, want_gt=True)
    assert (res == ores).all(), np.nonzero(res != ores)
    assert (gt == ogt).all(), np.nonzero((gt != ogt).any(axis=1))


def test_bad_key_index_is_reported():
    print("case: bad key index", flush=True)
    blsgpu.set_engine(1, 16, 16)
    blsgpu.load().blsgpu_reset_keys()
    sks, pks, coeffs = O.make_keys(43, 1)
    h = blsgpu.register_key_affine(pks[0])
    comp, unc, nhm, _ = O.make_batch(43, 3, sks, np.zeros(3, np.int32))
    # host entry point validates indices before any launch
    with pytest.raises(RuntimeError):
        blsgpu.verify_compressed(comp, nhm, np.array([h, h + 5, h], np.int32))


def test_restore_default_engine():
    blsgpu.set_engine(1, 12, 16)
    assert blsgpu.get_engine() == (1, 12, 16)


@pytest.mark.parametrize("cfg", [(1, 12, 16), (1, 16, 12), (1, 16, 16)], ids=lambda c: "mode%d_ml%d_fe%d" % c)
def test_subgroup_test_placements(cfg):
    """The subgroup test carried by the Miller launch (12 lanes), by the final
    exponentiation launch (12 lanes) and on the second stream (16/16): off-subgroup
    signatures at the edges of the 64-item rider workgroups, among undecodable and
    wrong-message items, over a batch of five rider workgroups."""
    print("case: subgroup test placement", cfg, flush=True)
    n = 300
    sks, pks, _ = O.make_keys(43, 5)
    key_idx = np.array([(i * 3) % 5 for i in range(n)], np.int32)
    inv = {0: "non_subgroup", 63: "non_subgroup", 64: "non_subgroup", 100: "wrong_msg", 130: "non_subgroup",
           200: "no_sqrt", 255: "non_subgroup", 256: "wrong_msg", 299: "non_subgroup"}
    comp, _, nhm, _ = O.make_batch(43, n, sks, key_idx, inv)
    want = []
    for i in range(n):
        rc, _ = O.g1_decompress(comp[i])
        want.append(rc if rc != 0 else (0 if inv.get(i) == "wrong_msg" else 1))
    want = np.array(want)
    assert (want == -4).sum() == 6 and (want == 1).sum() == n - 9
    blsgpu.set_engine(*cfg)
    blsgpu.load().blsgpu_reset_keys()
    hs = np.array([blsgpu.register_key_affine(pk) for pk in pks], np.int32)
    res, gt = blsgpu.verify_compressed(comp, nhm, hs[key_idx], want_gt=True)
    assert (res == want).all(), np.nonzero(res != want)
    assert not gt[want < 0].any()
