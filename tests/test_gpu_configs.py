"""The benchmark configurations at their full sizes against the oracle, every verdict
and every GT element bit for bit:
  uniform 4096   4096 distinct keys and messages, 64 invalid (32 wrong message, 16 no
                 square root, 16 off the subgroup), plain form through the blocking
                 compressed entry point and through the device-resident form that the
                 benchmark times, at the library's default engine and at Miller 32 /
                 final exponentiation 12 lanes;
  at scale       65536 verifications over 40 keys and 4096 messages, lognormal batch
                 sizes (median 64, sigma 1.2, [1, 8192]; seed 1006 draws 487 batches,
                 the largest 6478 items), 1/128 invalid, the batched form (queue-fed
                 persistent kernels, graph replay).
The oracle runs on 16 host threads (about 0.5 s of wall time for the 4096 batch and 7 s
for the 65536 items on the GPU host's cores)."""
import numpy as np
import pytest

import blsgpu
import oracle as O
from blsgpu import workload

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ORACLE_THREADS = 16


def _oracle(w):
    w["coeffs"] = np.array([O.g2_prepare(p) for p in w["pk"]])
    w["g2c"] = O.g2_prepare(O.g2_gen())
    return O.verify_compressed_batch(w["sig"], w["nhm"], w["g2c"], w["coeffs"], w["key_idx"], threads=ORACLE_THREADS)


@pytest.fixture(scope="module")
def uniform():
    assert blsgpu.init() == 0, blsgpu.last_error()
    print("case: generating the uniform 4096 batch (seed 20261203)", flush=True)
    w = workload.uniform(4096, 20261203)
    w["want"], w["wgt"] = _oracle(w)
    assert (w["want"] == workload.expected_codes(w["cat"])).all()
    return w


@pytest.fixture(scope="module")
def at_scale():
    assert blsgpu.init() == 0, blsgpu.last_error()
    print("case: generating the 65536-item at-scale workload (seed 1006)", flush=True)
    w = workload.at_scale(1006)
    assert len(w["sizes"]) == 487 and int(w["sizes"].max()) == 6478
    w["want"], w["wgt"] = _oracle(w)
    assert (w["want"] == workload.expected_codes(w["cat"])).all()
    return w


@pytest.mark.parametrize("engine", [None, (1, 32, 12)], ids=["default", "ml32_fe12"])
def test_uniform_4096_plain_form(uniform, engine):
    print("case: uniform 4096, plain form, engine", engine, flush=True)
    w = uniform
    saved = blsgpu.get_engine()
    try:
        if engine:
            blsgpu.set_engine(*engine)
        blsgpu.load().blsgpu_reset_keys()
        h = blsgpu.register_keys_affine(w["pk"])
        res, gt = blsgpu.verify_compressed(w["sig"], w["nhm"], h[w["key_idx"]], want_gt=True)
        assert (res == w["want"]).all(), np.nonzero(res != w["want"])
        assert (gt == w["wgt"]).all(), np.nonzero((gt != w["wgt"]).any(axis=1))
        # the device-resident form the benchmark times, on a torch stream
        dev = torch.device("cuda:0")
        d_sig = torch.from_numpy(w["sig"].reshape(-1)).to(dev)
        d_nhm = torch.from_numpy(w["nhm"].view(np.int64).reshape(-1)).to(dev)
        d_key = torch.from_numpy(h[w["key_idx"]].astype(np.int32)).to(dev)
        d_res = torch.full((4096,), 7, dtype=torch.int32, device=dev)
        d_gt = torch.zeros((4096, 72), dtype=torch.int64, device=dev)
        rc = blsgpu.load().blsgpu_verify_compressed_device(d_sig.data_ptr(), d_nhm.data_ptr(), d_key.data_ptr(),
                                                           d_res.data_ptr(), d_gt.data_ptr(), 4096,
                                                           torch.cuda.current_stream().cuda_stream)
        assert rc == 0, blsgpu.last_error()
        torch.cuda.synchronize()
        assert (d_res.cpu().numpy() == w["want"]).all()
        assert (d_gt.cpu().numpy().view(np.uint64) == w["wgt"]).all()
    finally:
        blsgpu.set_engine(*saved)


@pytest.mark.parametrize("engine", [(1, 32, 12), (1, 12, 12), (0, 16, 16)],
                         ids=["ml32_fe12_two_parts", "ml12_fe12_two_parts", "thread_per_item_one_part"])
def test_uniform_4096_legacy_abi(uniform, engine):
    """gpu_batch_bls_verify as a replica calls it: host arrays of uncompressed signatures
    (zeros where the encoding does not decode: the flagless identity), -H(m), the
    generator's lines and 2448 u64 of lines per item.  In the lanes-per-item mode the
    4096 items run as two parts of 2048, the Miller loop of the first during the copies
    of the second, whose key tables cross the three-slot pinned ring in five chunks each
    (slots reused); the thread-per-item mode runs one part.  Compared with the oracle's
    verification of the same arrays, every verdict and GT element."""
    print("case: uniform 4096, legacy ABI, engine", engine, flush=True)
    w = uniform
    if "legacy" not in w:
        unc, _ = blsgpu.decompress_status(w["sig"], check_subgroup=False)
        pkc = np.ascontiguousarray(w["coeffs"][w["key_idx"]])
        ores, ogt = O.verify_batch(unc, w["nhm"], w["g2c"], pkc, threads=ORACLE_THREADS)
        w["legacy"] = (unc, pkc, ores, ogt)
    unc, pkc, ores, ogt = w["legacy"]
    # every valid item verifies; the others (wrong message, zeros, off-subgroup points) fail
    assert (ores == (w["cat"] == 0)).all()
    saved = blsgpu.get_engine()
    try:
        blsgpu.set_engine(*engine)
        res, gt = blsgpu.batch_verify(unc, w["nhm"], w["g2c"], pkc, want_gt=True)
        assert (res == ores).all(), np.nonzero(res != ores)
        assert (gt == ogt).all(), np.nonzero((gt != ogt).any(axis=1))
        # the same arrays again (the ring's slots now hold the previous call's tables)
        res2 = blsgpu.batch_verify(unc, w["nhm"], w["g2c"], pkc)
        assert (res2 == ores).all()
    finally:
        blsgpu.set_engine(*saved)


@pytest.mark.parametrize("cfg", [(12, 12, 0, 0), (32, 24, 0, 0), (32, 12, 4, 8)], ids=lambda c: "ml%d_fe%d_w%d_%d" % c)
def test_at_scale_65536_batched_form(at_scale, cfg):
    print("case: at scale 65536 items / 487 batches / 40 keys, batched form", cfg, flush=True)
    w = at_scale
    blsgpu.load().blsgpu_reset_keys()
    h = blsgpu.register_keys_affine(w["pk"])
    blsgpu.set_batched_engine(*cfg, True)
    try:
        dev = torch.device("cuda:0")
        n = len(w["key_idx"])
        off = np.concatenate([[0], np.cumsum(w["sizes"])]).astype(np.int32)
        assert off[-1] == n
        d_sig = torch.from_numpy(w["sig"].reshape(-1)).to(dev)
        d_nhm = torch.from_numpy(w["nhm"].view(np.int64).reshape(-1)).to(dev)
        d_key = torch.from_numpy(h[w["key_idx"]].astype(np.int32)).to(dev)
        d_res = torch.full((n,), 7, dtype=torch.int32, device=dev)
        d_gt = torch.zeros((n, 72), dtype=torch.int64, device=dev)
        rc = blsgpu.load().blsgpu_verify_batches_device(d_sig.data_ptr(), d_nhm.data_ptr(), d_key.data_ptr(),
                                                        O.ptr(off), len(off) - 1, d_res.data_ptr(), d_gt.data_ptr(),
                                                        torch.cuda.current_stream().cuda_stream)
        assert rc == 0, blsgpu.last_error()
        torch.cuda.synchronize()
        res = d_res.cpu().numpy()
        gt = d_gt.cpu().numpy().view(np.uint64)
        assert (res == w["want"]).all(), np.nonzero(res != w["want"])
        assert (gt == w["wgt"]).all(), np.nonzero((gt != w["wgt"]).any(axis=1))
    finally:
        blsgpu.set_batched_engine(12, 16, 0, 0, True)
