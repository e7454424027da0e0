"""Thread safety: the reference's callers verify from concurrent threads (rayon, and the
thread-local batch queue of ic-integration/patches/ic-gpu-consensus.patch:324).  Six
host threads call the legacy entry point, the compressed entry point, decompression and
key registration at the same time while a serving loop runs its own requests; every
result equals the oracle's.  Also: a key reset and a new registration while
device-form work that reads the old key tables is still queued."""
import threading

import numpy as np
import pytest

import blsgpu
import oracle as O
from gpu_workload import make_workload, oracle_expected

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def data():
    assert blsgpu.init() == 0, blsgpu.last_error()
    w = make_workload(101, 6, 60, 30, invalid_every=11)
    blsgpu.load().blsgpu_reset_keys()
    w["h"] = blsgpu.register_keys_affine(w["pk"])
    w["want"], w["wgt"] = oracle_expected(w["sig"], w["nhm"], w["pk"], w["key_idx"])
    # the legacy ABI's inputs: uncompressed signatures of the items that decode, and each
    # item's key lines (2448 u64 per item)
    ok = np.array([O.g1_decompress(s)[0] == 0 for s in w["sig"]])
    idx = np.nonzero(ok)[0]
    w["leg_idx"] = idx
    w["leg_sig"] = np.array([O.g1_decompress(w["sig"][i])[1] for i in idx])
    w["leg_nhm"] = np.ascontiguousarray(w["nhm"][idx])
    w["coeffs"] = np.array([O.g2_prepare(p) for p in w["pk"]])
    w["leg_pkc"] = np.ascontiguousarray(w["coeffs"][w["key_idx"][idx]])
    w["g2c"] = O.g2_prepare(O.g2_gen())
    return w


def test_threads_and_server_at_once(data):
    print("case: 6 host threads + a 2-slot server, 3 rounds each", flush=True)
    w = data
    errors = []
    h = w["h"]

    def legacy():
        for _ in range(3):
            res, gt = blsgpu.batch_verify(w["leg_sig"], w["leg_nhm"], w["g2c"], w["leg_pkc"], want_gt=True)
            assert (res == w["want"][w["leg_idx"]]).all()
            assert (gt == w["wgt"][w["leg_idx"]]).all()

    def compressed(lo, hi):
        def run():
            for _ in range(3):
                res, gt = blsgpu.verify_compressed(w["sig"][lo:hi], w["nhm"][lo:hi], h[w["key_idx"][lo:hi]],
                                                   want_gt=True)
                assert (res == w["want"][lo:hi]).all()
                assert (gt == w["wgt"][lo:hi]).all()
        return run

    def decompress():
        for _ in range(3):
            out, st = blsgpu.decompress_status(w["sig"])
            for i in range(len(w["sig"])):
                rc, pt = O.g1_decompress(w["sig"][i])
                assert st[i] == rc
                if rc == 0:
                    assert (out[i] == pt).all()

    def register():
        for r in range(3):
            hs = blsgpu.register_keys_affine(w["pk"][:3])
            for k in range(3):
                assert (blsgpu.key_coeffs(int(hs[k])) == w["coeffs"][k]).all()
            res, gt = blsgpu.verify_compressed(w["sig"][:20], w["nhm"][:20],
                                               np.where(w["key_idx"][:20] < 3, hs[np.minimum(w["key_idx"][:20], 2)],
                                                        h[w["key_idx"][:20]]).astype(np.int32), want_gt=True)
            assert (res == w["want"][:20]).all() and (gt == w["wgt"][:20]).all()

    def guarded(fn):
        def run():
            try:
                fn()
            except Exception as e:  # collected and re-raised in the main thread
                errors.append(e)
        return run

    blsgpu.set_batched_engine(12, 12, 0, 0, True)
    with blsgpu.Server(slots=2, max_step_items=64, hbm_budget_bytes=1 << 28, want_gt=True) as srv:
        def serve():
            for _ in range(3):
                reqs = [(0, 7), (7, 20), (27, 3), (30, 30)]
                ids = [srv.submit(w["sig"][o:o + n], w["nhm"][o:o + n], h[w["key_idx"][o:o + n]]) for o, n in reqs]
                for (o, n), rid in zip(reqs, ids):
                    res, gt, _ = srv.wait(rid, timeout_ms=60000)
                    assert (res == w["want"][o:o + n]).all()
                    assert (gt == w["wgt"][o:o + n]).all()

        threads = [threading.Thread(target=guarded(f)) for f in
                   (legacy, compressed(0, 30), compressed(30, 60), decompress, register, serve)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=100)
        assert not any(t.is_alive() for t in threads), "a thread did not finish"
    blsgpu.set_batched_engine(12, 16, 0, 0, True)
    if errors:
        raise errors[0]


def test_reset_and_register_while_device_work_is_queued(data):
    """Device-form work reads the key tables when its kernels run, after the call
    returns.  A reset followed by a registration must not overwrite the rows that
    queued work still reads: the queued results stay those of the old keys."""
    print("case: reset + register while device-form work is queued", flush=True)
    w = data
    lib = blsgpu.load()
    lib.blsgpu_reset_keys()
    h = blsgpu.register_keys_affine(w["pk"])
    dev = torch.device("cuda:0")
    reps = 8  # the same 60 items several times over, so the queued work lasts a while
    sig = torch.from_numpy(np.tile(w["sig"], (reps, 1)).reshape(-1)).to(dev)
    nhm = torch.from_numpy(np.tile(w["nhm"], (reps, 1)).view(np.int64).reshape(-1)).to(dev)
    key = torch.from_numpy(np.tile(h[w["key_idx"]], reps).astype(np.int32)).to(dev)
    n = 60 * reps
    res = torch.zeros(n, dtype=torch.int32, device=dev)
    gt = torch.zeros((n, 72), dtype=torch.int64, device=dev)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        # the stream first waits on a long torch job, so the verification is still queued
        # when the keys are reset and re-registered
        big = torch.randn(4096, 4096, device=dev)
        for _ in range(8):
            big = big @ big / 4096
        rc = lib.blsgpu_verify_compressed_device(sig.data_ptr(), nhm.data_ptr(), key.data_ptr(), res.data_ptr(),
                                                 gt.data_ptr(), n, s.cuda_stream)
        assert rc == 0, blsgpu.last_error()
    lib.blsgpu_reset_keys()
    # other keys take the handles 0.. of the old ones
    others = np.ascontiguousarray(w["pk"][::-1])
    for k in range(len(others)):
        assert blsgpu.register_key_affine(others[k]) == k
    s.synchronize()
    torch.cuda.synchronize()
    r = res.cpu().numpy()
    g = gt.cpu().numpy().view(np.uint64)
    assert (r == np.tile(w["want"], reps)).all()
    assert (g == np.tile(w["wgt"], (reps, 1))).all()
    # and the new registrations hold the other keys
    for k in range(len(others)):
        assert (blsgpu.key_coeffs(k) == O.g2_prepare(others[k])).all()
    lib.blsgpu_reset_keys()
