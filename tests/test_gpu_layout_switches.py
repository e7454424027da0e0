"""The plain form at the batch sizes where its launches change layout, against the oracle
bit for bit (every verdict and GT element), at Miller 32 lanes / final exponentiation 12
lanes (the engine the autotuner picks for the 4096 batch):
  * the Miller loop's tiles (two wavefronts of four items each at 32 lanes) run the
    one-wave kernel while they make at most one wavefront per SIMD (M <= 2 * 4 * CUs) and
    the two-wave kernel beyond;
  * the subgroup test rides on the final exponentiation's launch while its workgroups
    (one per 64 items) fit the SIMDs that the tiles (one wavefront per 5 items) leave
    free, ceil(M / 5) + ceil(M / 64) <= 4 * CUs, and runs on the second stream beside
    the pairing kernels beyond.
Each size is checked on both sides of its switch.  The batches are prefixes of one
uniform workload (distinct keys and messages; wrong-message, no-square-root and
off-subgroup items among them)."""
import numpy as np
import pytest

import blsgpu
import oracle as O
from blsgpu import workload

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _cus():
    return torch.cuda.get_device_properties(0).multi_processor_count


def _rider_limit(cus):
    # largest M whose final-exponentiation tiles and rider workgroups fit 4 * CUs SIMDs
    m = 1
    while -(-(m + 1) // 5) + -(-(m + 1) // 64) <= 4 * cus:
        m += 1
    return m


@pytest.fixture(scope="module")
def data():
    assert blsgpu.init() == 0, blsgpu.last_error()
    cus = _cus()
    one_wave = 2 * 4 * cus
    riders = _rider_limit(cus)
    sizes = sorted({one_wave, one_wave + 1, riders, riders + 1})
    n = sizes[-1]
    print("case: layout switches, %d CUs, sizes %s" % (cus, sizes), flush=True)
    w = workload.uniform(n, 4242)
    coeffs = np.array([O.g2_prepare(p) for p in w["pk"]])
    g2c = O.g2_prepare(O.g2_gen())
    w["want"], w["wgt"] = O.verify_compressed_batch(w["sig"], w["nhm"], g2c, coeffs, w["key_idx"], threads=16)
    assert (w["want"] == workload.expected_codes(w["cat"])).all()
    blsgpu.load().blsgpu_reset_keys()
    w["h"] = blsgpu.register_keys_affine(w["pk"])
    w["sizes"] = sizes
    return w


@pytest.mark.parametrize("k", range(4), ids=["one_wave", "one_wave+1", "riders", "riders+1"])
def test_plain_form_across_layout_switch(data, k):
    m = data["sizes"][k]
    cat = data["cat"][:m]
    # every category in the prefix, so that a misplaced subgroup test shows in a verdict
    assert {1, 2, 3} <= set(int(c) for c in cat), m
    print("case: plain form, M = %d" % m, flush=True)
    saved = blsgpu.get_engine()
    try:
        blsgpu.set_engine(1, 32, 12)
        res, gt = blsgpu.verify_compressed(data["sig"][:m], data["nhm"][:m], data["h"][data["key_idx"][:m]],
                                           want_gt=True)
    finally:
        blsgpu.set_engine(*saved)
    assert (res == data["want"][:m]).all(), np.nonzero(res != data["want"][:m])
    assert (gt == data["wgt"][:m]).all(), np.nonzero((gt != data["wgt"][:m]).any(axis=1))
