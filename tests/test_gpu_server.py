"""The serving loop: requests that share steps with others (continuous batching over
several slots) get exactly the verdicts and GT elements they get alone, and the plain
one-request-per-step form agrees; the load generator completes a short run."""
import numpy as np
import pytest

import blsgpu
from gpu_workload import make_workload, oracle_expected

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool():
    assert blsgpu.init() == 0, blsgpu.last_error()
    w = make_workload(81, 40, 400, 120, invalid_every=17)
    blsgpu.load().blsgpu_reset_keys()
    w["h"] = blsgpu.register_keys_affine(w["pk"])
    w["want"], w["wgt"] = oracle_expected(w["sig"], w["nhm"], w["pk"], w["key_idx"])
    return w


def _requests(n_total, seed):
    rng = np.random.default_rng(seed)
    out, o = [], 0
    while o < n_total:
        s = int(min(n_total - o, 60, max(1, round(np.exp(np.log(12) + rng.normal())))))
        out.append((o, s))
        o += s
    return out


@pytest.mark.parametrize("slots,max_reqs", [(3, 0), (1, 1), (2, 0)])
def test_mixed_requests_equal_alone(pool, slots, max_reqs):
    print("case: server slots=%d max_requests_per_step=%d" % (slots, max_reqs), flush=True)
    blsgpu.set_batched_engine(12, 16, 0, 0, True)
    reqs = _requests(400, 3)
    with blsgpu.Server(slots=slots, max_step_items=96, hbm_budget_bytes=1 << 30, max_requests_per_step=max_reqs,
                       want_gt=True) as srv:
        ids = [srv.submit(pool["sig"][o:o + n], pool["nhm"][o:o + n], pool["h"][pool["key_idx"][o:o + n]])
               for o, n in reqs]
        for (o, n), rid in zip(reqs, ids):
            res, gt, lat = srv.wait(rid, timeout_ms=60000)
            assert (res == pool["want"][o:o + n]).all(), o
            assert (gt == pool["wgt"][o:o + n]).all(), o
            assert lat >= 0
            # the same request alone through the plain fused entry point
            alone, agt = blsgpu.verify_compressed(pool["sig"][o:o + n], pool["nhm"][o:o + n],
                                                  pool["h"][pool["key_idx"][o:o + n]], want_gt=True)
            assert (alone == res).all() and (agt == gt).all()
        st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["failed"] == 0
    if max_reqs == 1:
        assert st["steps"] == len(reqs)
    else:
        assert st["steps"] < len(reqs)  # requests shared steps
        assert st["max_step"] > max(n for _, n in reqs)


def test_submit_rejects_bad_requests(pool):
    with blsgpu.Server(slots=1, max_step_items=8, hbm_budget_bytes=1 << 30) as srv:
        with pytest.raises(RuntimeError):
            srv.submit(pool["sig"][:9], pool["nhm"][:9], pool["h"][pool["key_idx"][:9]])  # above capacity
        with pytest.raises(RuntimeError):
            srv.submit(pool["sig"][:2], pool["nhm"][:2], np.array([0, 999], np.int32))  # unknown key


def test_budget_caps_step_capacity(pool):
    with blsgpu.Server(slots=2, max_step_items=100000, hbm_budget_bytes=2 * 1056 * 50) as srv:
        assert srv.capacity == 50


def test_loadtest_short_run(pool):
    print("case: loadtest 0.5 s at 2000 items/s", flush=True)
    with blsgpu.Server(slots=3, max_step_items=2048, hbm_budget_bytes=1 << 30) as srv:
        r = srv.loadtest(pool["sig"], pool["nhm"], pool["h"][pool["key_idx"]], items_per_s=2000, seconds=0.5,
                         drain_s=10.0, size_median=16, size_sigma=1.0, size_min=1, size_max=256, seed=1)
    assert r["requests"] > 0
    assert r["requests_done"] == r["requests"]
    assert 0 < r["p50_ms"] <= r["p99_ms"] <= r["p999_ms"] <= r["max_ms"] < 10000
