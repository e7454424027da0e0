"""The batched form (persistent queue-fed kernels, hipGraph replay, autotuner) against
the oracle: verdicts and GT elements bit-exact for uneven batches, for every
lanes-per-item / workgroups-per-CU configuration, and across graph replays."""
import numpy as np
import pytest

import blsgpu
import oracle as O
from gpu_workload import make_workload, oracle_expected

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def small():
    assert blsgpu.init() == 0, blsgpu.last_error()
    w = make_workload(71, 5, 97, 20, invalid_every=9)
    w["want"], w["wgt"] = oracle_expected(w["sig"], w["nhm"], w["pk"], w["key_idx"])
    return w


@pytest.fixture(scope="module")
def large():
    # more tiles than one workgroup per CU can take at once, so every queue-fed
    # workgroup runs several tiles when wpc = 1
    assert blsgpu.init() == 0, blsgpu.last_error()
    w = make_workload(72, 13, 1600, 300, invalid_every=128)
    w["want"], w["wgt"] = oracle_expected(w["sig"], w["nhm"], w["pk"], w["key_idx"])
    return w


def _register(w):
    blsgpu.load().blsgpu_reset_keys()
    return blsgpu.register_keys_affine(w["pk"])


def _split(n, sizes):
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    assert off[-1] == n
    return off


def test_register_keys_in_one_launch_matches_oracle_lines(small):
    print("case: register 5 keys in one launch", flush=True)
    h = _register(small)
    assert h.tolist() == list(range(5))
    for i in range(5):
        assert (blsgpu.key_coeffs(int(h[i])) == O.g2_prepare(small["pk"][i])).all()


def test_host_batches_uneven_sizes(small):
    print("case: host batched form, 8 uneven batches of 97 items", flush=True)
    h = _register(small)
    blsgpu.set_batched_engine(12, 16, 0, 0, True)
    sizes = [1, 30, 0, 7, 2, 40, 1, 16]
    off = _split(97, sizes)
    batches = [(small["sig"][off[b]:off[b + 1]], small["nhm"][off[b]:off[b + 1]],
                h[small["key_idx"][off[b]:off[b + 1]]]) for b in range(len(sizes))]
    res, gts = blsgpu.verify_batches(batches, want_gt=True)
    for b in range(len(sizes)):
        assert (res[b] == small["want"][off[b]:off[b + 1]]).all(), b
        assert (gts[b] == small["wgt"][off[b]:off[b + 1]]).all(), b


def _device_run(w, h, off, stream=None):
    dev = "cuda"
    sig = torch.from_numpy(w["sig"]).to(dev)
    nhm = torch.from_numpy(w["nhm"].view(np.int64)).to(dev)
    ki = torch.from_numpy(h[w["key_idx"]].astype(np.int32)).to(dev)
    n = w["sig"].shape[0]
    res = torch.zeros(n, dtype=torch.int32, device=dev)
    gt = torch.zeros((n, 72), dtype=torch.int64, device=dev)
    return sig, nhm, ki, res, gt


def _call_device(bufs, off):
    sig, nhm, ki, res, gt = bufs
    offa = np.ascontiguousarray(off, np.int32)
    stream = torch.cuda.current_stream().cuda_stream
    rc = blsgpu.load().blsgpu_verify_batches_device(sig.data_ptr(), nhm.data_ptr(), ki.data_ptr(), O.ptr(offa),
                                                    len(off) - 1, res.data_ptr(), gt.data_ptr(), stream)
    assert rc == 0, blsgpu.last_error()
    torch.cuda.synchronize()
    return res.cpu().numpy().copy(), gt.cpu().numpy().view(np.uint64).copy()


@pytest.mark.parametrize("cfg", [(12, 16, 0, 0), (12, 16, 1, 1), (13, 18, 2, 1), (16, 16, -1, -1), (8, 32, 1, 2),
                                 (32, 12, 0, 1), (32, 12, 4, 8), (12, 12, 5, 6)],
                         ids=lambda c: "ml%d_fe%d_w%d_%d" % c)
def test_device_batches_queue_configs(cfg, large):
    # 12-lane final exponentiations at more than four workgroups per CU run its compact
    # form (two Fp12 values in LDS, the others in device memory)
    print("case: device batched form 1600 items, config", cfg, flush=True)
    h = _register(large)
    blsgpu.set_batched_engine(*cfg, True)
    rng = np.random.default_rng(5)
    sizes = []
    left = 1600
    while left:
        s = int(min(left, rng.geometric(1 / 60)))
        sizes.append(s)
        left -= s
    off = _split(1600, sizes)
    bufs = _device_run(large, h, off)
    res, gt = _call_device(bufs, off)
    assert (res == large["want"]).all()
    assert (gt == large["wgt"]).all()


def test_graph_replays_are_exact_and_cached(small):
    print("case: graph replay x3 with eager work between", flush=True)
    h = _register(small)
    blsgpu.set_batched_engine(12, 16, 1, 1, True)
    off = _split(97, [50, 47])
    bufs = _device_run(small, h, off)
    b0 = blsgpu.batched_stats()
    for rep in range(3):
        bufs[3].fill_(7)
        bufs[4].fill_(7)
        res, gt = _call_device(bufs, off)
        assert (res == small["want"]).all(), rep
        assert (gt == small["wgt"]).all(), rep
        # eager library work between replays
        r2 = blsgpu.verify_compressed(small["sig"][:5], small["nhm"][:5], h[small["key_idx"][:5]])
        assert (r2 == small["want"][:5]).all()
    b1 = blsgpu.batched_stats()
    assert b1[0] - b0[0] == 1  # one instantiation
    assert b1[1] - b0[1] == 3  # three launches


def test_autotuner_picks_a_valid_config_and_stays_exact(small):
    print("case: autotune on 97 items", flush=True)
    h = _register(small)
    blsgpu.set_batched_engine(12, 16, 0, 0, True)
    sig, nhm, ki, res, gt = _device_run(small, h, None)
    cfg = np.zeros(4, np.int32)
    ms = np.zeros(2, np.float64)
    rc = blsgpu.load().blsgpu_autotune_device(sig.data_ptr(), nhm.data_ptr(), ki.data_ptr(), 97, 0,
                                              torch.cuda.current_stream().cuda_stream, O.ptr(cfg), O.ptr(ms))
    assert rc == 0, blsgpu.last_error()
    assert cfg[0] in (8, 12, 13, 16, 20, 24, 32) and cfg[1] in (9, 12, 16, 18, 24, 27, 32)
    assert (ms > 0).all()
    assert blsgpu.get_batched_engine(97)[:4] == tuple(int(c) for c in cfg)
    res2, gt2 = _call_device((sig, nhm, ki, res, gt), _split(97, [97]))
    assert (res2 == small["want"]).all()
    assert (gt2 == small["wgt"]).all()
    blsgpu.set_batched_engine(12, 16, 0, 0, True)
    # plain form: sets the plain engine's lanes per item
    saved = blsgpu.get_engine()
    rc = blsgpu.load().blsgpu_autotune_device(sig.data_ptr(), nhm.data_ptr(), ki.data_ptr(), 97, 1,
                                              torch.cuda.current_stream().cuda_stream, O.ptr(cfg), O.ptr(ms))
    assert rc == 0, blsgpu.last_error()
    assert cfg[2] == -1 and cfg[3] == -1
    assert blsgpu.get_engine() == (1, int(cfg[0]), int(cfg[1]))
    assert blsgpu.get_line_depth(1, 97) in (0, 1, 2)  # the tuned key-line depth
    r3, g3 = blsgpu.verify_compressed(small["sig"], small["nhm"], h[small["key_idx"]], want_gt=True)
    assert (r3 == small["want"]).all() and (g3 == small["wgt"]).all()
    blsgpu.set_engine(*saved)


@pytest.mark.parametrize("depth", [0, 1, 2])
def test_key_line_depths_match_oracle(large, depth):
    """Every key-line depth (how many steps ahead the Miller loop reads its key lines into
    registers) gives the oracle's verdicts and GT elements, in the plain form at 32, 12
    (the one-wave kernel that carries the subgroup test) and 24 lanes and in the batched
    form; the depth each launch runs is the requested one where its kernel holds it."""
    print("case: key-line depth", depth, flush=True)
    h = _register(large)
    saved = blsgpu.get_engine()
    n = 400
    blsgpu.set_line_depth(depth)
    try:
        for eng, avail in (((1, 32, 12), 2), ((1, 12, 12), 2), ((1, 24, 12), 2), ((1, 16, 16), 0)):
            blsgpu.set_engine(*eng)
            assert blsgpu.get_line_depth(1, n) == min(depth, avail), eng
            res, gt = blsgpu.verify_compressed(large["sig"][:n], large["nhm"][:n], h[large["key_idx"][:n]],
                                               want_gt=True)
            assert (res == large["want"][:n]).all(), eng
            assert (gt == large["wgt"][:n]).all(), eng
        blsgpu.set_batched_engine(32, 12, 0, 0, True)
        assert blsgpu.get_line_depth(0, 1600) == depth
        off = _split(1600, [700, 1, 899])
        res, gt = _call_device(_device_run(large, h, off), off)
        assert (res == large["want"]).all() and (gt == large["wgt"]).all()
    finally:
        blsgpu.set_line_depth(-1)
        blsgpu.set_engine(*saved)
        blsgpu.set_batched_engine(12, 16, 0, 0, True)
