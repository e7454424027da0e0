"""blsgpu_verify_compressed_device_noalloc (the entry point for callers that cannot
allocate: its workspace is a fixed device array) against
blsgpu_verify_compressed_device (the library's grow-only workspace): verdicts and all 72
GT words bit for bit on two calls in a row, the second smaller, then a third on another
stream; sizes above the array's capacity are refused with a message; the oracle checks
the smaller call's results."""
import numpy as np
import pytest

import blsgpu
import oracle as O
from blsgpu import workload
from gpu_workload import oracle_expected

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MAX_ITEMS = 8192


def _dev(w, h, lo, hi):
    dev = torch.device("cuda:0")
    n = hi - lo
    return dict(
        sig=torch.from_numpy(np.ascontiguousarray(w["sig"][lo:hi]).reshape(-1)).to(dev),
        nhm=torch.from_numpy(np.ascontiguousarray(w["nhm"][lo:hi]).view(np.int64).reshape(-1)).to(dev),
        key=torch.from_numpy(np.ascontiguousarray(h[w["key_idx"][lo:hi]]).astype(np.int32)).to(dev),
        res=torch.full((n,), 7, dtype=torch.int32, device=dev),
        gt=torch.full((n, 72), 7, dtype=torch.int64, device=dev),
        n=n)


def _call(fn, b, stream):
    rc = fn(b["sig"].data_ptr(), b["nhm"].data_ptr(), b["key"].data_ptr(), b["res"].data_ptr(), b["gt"].data_ptr(),
            b["n"], stream)
    assert rc == 0, blsgpu.last_error()


def _out(b):
    return b["res"].cpu().numpy().copy(), b["gt"].cpu().numpy().view(np.uint64).copy()


def test_noalloc_equals_device_form_on_two_calls_then_another_stream():
    assert blsgpu.init() == 0, blsgpu.last_error()
    lib = blsgpu.load()
    print("case: no-alloc entry point, 4096 then 1000 items, then 333 on a second stream", flush=True)
    w = workload.uniform(4096, 77)
    for i in (5, 600, 601):  # identity signatures ride along
        w["sig"][i] = O.IDENTITY_G1
    lib.blsgpu_reset_keys()
    h = blsgpu.register_keys_affine(w["pk"])
    s = torch.cuda.current_stream().cuda_stream
    for lo, hi in ((0, 4096), (100, 1100)):
        a, b = _dev(w, h, lo, hi), _dev(w, h, lo, hi)
        _call(lib.blsgpu_verify_compressed_device, a, s)
        _call(lib.blsgpu_verify_compressed_device_noalloc, b, s)
        torch.cuda.synchronize()
        ra, ga = _out(a)
        rb, gb = _out(b)
        assert (ra != 7).all()
        assert (rb == ra).all(), np.nonzero(rb != ra)
        assert (gb == ga).all(), np.nonzero((gb != ga).any(axis=1))
    # the smaller call against the oracle
    want, wgt = oracle_expected(w["sig"][100:1100], w["nhm"][100:1100], w["pk"], w["key_idx"][100:1100])
    assert (rb == want).all() and (gb == wgt).all()
    # a third call on another stream, right after one on the first: ordered by the library
    other = torch.cuda.Stream()
    c1, c2 = _dev(w, h, 2000, 4096), _dev(w, h, 3000, 3333)
    _call(lib.blsgpu_verify_compressed_device_noalloc, c1, s)
    _call(lib.blsgpu_verify_compressed_device_noalloc, c2, other.cuda_stream)
    torch.cuda.synchronize()
    r1, g1 = _out(c1)
    r2, g2 = _out(c2)
    assert (r2 == r1[1000:1333]).all() and (g2 == g1[1000:1333]).all()
    assert (r1 == workload.expected_codes(w["cat"][2000:4096])).all()


def test_noalloc_refuses_sizes_above_capacity():
    assert blsgpu.init() == 0, blsgpu.last_error()
    lib = blsgpu.load()
    rc = lib.blsgpu_verify_compressed_device_noalloc(8, 8, 8, 8, None, MAX_ITEMS + 1, None)
    assert rc == -1
    assert "8192" in blsgpu.last_error()
    assert lib.blsgpu_verify_compressed_device_noalloc(8, 8, 8, 8, None, 0, None) == -1
