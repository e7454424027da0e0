"""Identity signatures (the 0xc0 encoding) on the GPU: a tile holding one takes the
Miller loop's one-line-at-a-time branch that drops the item's first pair.  Identity
signatures sit first, in the middle and last in tiles of 5 items (12 or 24 lanes per
item), 4 (13, 16 or 32 lanes), 3 (20 lanes) and 8 (8 lanes), and one tile is identity
signatures only; the plain form at every Miller lane count and in the thread-per-item
mode, the batched form and the serving loop all give the oracle's verdict (0) and GT
(e(-H(m), pk), all 72 words) for them, and the oracle's results for the items around
them."""
import numpy as np
import pytest

import blsgpu
import oracle as O
from gpu_workload import make_workload, oracle_expected

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

N = 40
# positions in tiles of 5: 0 first, 9 last, 12 middle, 19 last, 24 last, 30..34 a whole
# tile, 39 the batch's last item; in tiles of 4: 0 first, 9 second, 12 first, 19 last,
# 24 first, 32..34, 39 last
IDENT = [0, 9, 12, 19, 24, 30, 31, 32, 33, 34, 39]


@pytest.fixture(scope="module")
def data():
    assert blsgpu.init() == 0, blsgpu.last_error()
    w = make_workload(91, 6, N, 25, invalid_every=7)
    for i in IDENT:
        w["sig"][i] = O.IDENTITY_G1
    blsgpu.load().blsgpu_reset_keys()
    w["h"] = blsgpu.register_keys_affine(w["pk"])
    w["want"], w["wgt"] = oracle_expected(w["sig"], w["nhm"], w["pk"], w["key_idx"])
    assert (w["want"][IDENT] == 0).all()
    assert (w["want"] == 1).sum() > 10  # valid items around them
    return w


@pytest.mark.parametrize("g_ml", [8, 12, 13, 16, 20, 24, 32])
def test_plain_form_every_miller_lane_count(data, g_ml):
    print("case: identity signatures, plain form, Miller G=%d, final exp G=12" % g_ml, flush=True)
    saved = blsgpu.get_engine()
    try:
        blsgpu.set_engine(1, g_ml, 12)
        res, gt = blsgpu.verify_compressed(data["sig"], data["nhm"], data["h"][data["key_idx"]], want_gt=True)
    finally:
        blsgpu.set_engine(*saved)
    assert (res == data["want"]).all(), np.nonzero(res != data["want"])
    assert (gt == data["wgt"]).all(), np.nonzero((gt != data["wgt"]).any(axis=1))


def test_thread_per_item_mode(data):
    print("case: identity signatures, thread-per-item mode", flush=True)
    saved = blsgpu.get_engine()
    try:
        blsgpu.set_engine(0)
        res, gt = blsgpu.verify_compressed(data["sig"], data["nhm"], data["h"][data["key_idx"]], want_gt=True)
    finally:
        blsgpu.set_engine(*saved)
    assert (res == data["want"]).all()
    assert (gt == data["wgt"]).all()


@pytest.mark.parametrize("cfg", [(12, 12, 0, 0), (32, 12, 1, 1), (12, 12, -1, -1), (32, 24, -1, 0)],
                         ids=lambda c: "ml%d_fe%d_w%d_%d" % c)
def test_batched_form(data, cfg):
    print("case: identity signatures, batched form", cfg, flush=True)
    blsgpu.set_batched_engine(*cfg, True)
    sizes = [3, 10, 1, 14, 12]
    off = np.concatenate([[0], np.cumsum(sizes)])
    batches = [(data["sig"][off[b]:off[b + 1]], data["nhm"][off[b]:off[b + 1]],
                data["h"][data["key_idx"][off[b]:off[b + 1]]]) for b in range(len(sizes))]
    res, gts = blsgpu.verify_batches(batches, want_gt=True)
    blsgpu.set_batched_engine(12, 16, 0, 0, True)
    for b in range(len(sizes)):
        assert (res[b] == data["want"][off[b]:off[b + 1]]).all(), b
        assert (gts[b] == data["wgt"][off[b]:off[b + 1]]).all(), b


def test_serving_loop(data):
    print("case: identity signatures, serving loop with 2 slots", flush=True)
    blsgpu.set_batched_engine(12, 12, 0, 0, True)
    reqs = [(0, 5), (5, 8), (13, 1), (14, 17), (31, 9)]
    with blsgpu.Server(slots=2, max_step_items=64, hbm_budget_bytes=1 << 28, want_gt=True) as srv:
        ids = [srv.submit(data["sig"][o:o + n], data["nhm"][o:o + n], data["h"][data["key_idx"][o:o + n]])
               for o, n in reqs]
        for (o, n), rid in zip(reqs, ids):
            res, gt, _ = srv.wait(rid, timeout_ms=60000)
            assert (res == data["want"][o:o + n]).all(), o
            assert (gt == data["wgt"][o:o + n]).all(), o
    blsgpu.set_batched_engine(12, 16, 0, 0, True)


def test_legacy_abi_identity_without_flag(data):
    """The legacy ABI takes uncompressed points without an infinity flag (the caller's
    marshalling drops it, ic-integration/gpu_crypto/lib.rs:45-53): ic_bls12_381's
    identity, (0, 1) in its affine form, then reaches the pairing as that point.  The GPU
    computes the oracle's two-pair result for exactly those words (verdict 0)."""
    print("case: legacy ABI with the flagless identity point", flush=True)
    fx = O.fixtures()
    g2c = O.u64(fx["g2_gen_coeffs"])
    pkc = np.stack([O.u64(fx["pk42_coeffs"])] * 2)
    ident = np.zeros(12, np.uint64)
    ident[6:] = O.mont(1)
    sig = np.stack([ident, O.u64(fx["sig_294_g1"])])
    nhm = np.stack([O.u64(fx["neg_7_g1"])] * 2)
    res, gt = blsgpu.batch_verify(sig, nhm, g2c, pkc, want_gt=True)
    ores, ogt = O.verify_batch(sig, nhm, g2c, pkc)
    assert res.tolist() == [0, 1]
    assert (res == ores).all() and (gt == ogt).all()
