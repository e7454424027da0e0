"""GPU hash to G1 (blsgpu_hash_to_g1) against RFC 9380's vectors and the oracle
(oracle/h2c.py) over message and DST lengths that cross SHA-256 block boundaries, and
the fused wire-format path (blsgpu_verify_messages) against the compressed path fed
with the oracle's -H(m)."""
import os
import sys

import numpy as np
import pytest

import blsgpu
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import h2c  # noqa: E402
from test_oracle_h2c import G1_DST, G1_VECTORS  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    assert blsgpu.init() == 0, blsgpu.last_error()


def _affine(row):
    return (O.unmont(row[:6]), O.unmont(row[6:12]))


def test_rfc_vectors_on_gpu():
    print("case: hash_to_g1 RFC vectors", flush=True)
    out, ident = blsgpu.hash_to_g1([m for m, _, _ in G1_VECTORS], G1_DST)
    assert ident.tolist() == [0] * len(G1_VECTORS)
    for row, (_, x, y) in zip(out, G1_VECTORS):
        assert _affine(row) == (x, y)


def _messages(seed, lengths):
    return [bytes(O.rand_bytes(seed, 7_000_000 + i, n)) if n else b"" for i, n in enumerate(lengths)]


# b_0 hashes 64 + len + 3 + len(DST) + 1 bytes: these lengths put its padding on both
# sides of every block edge for the DSTs below
LENGTHS = [0, 1, 2, 3, 8, 9, 10, 11, 12, 20, 31, 32, 33, 55, 56, 57, 63, 64, 65, 72, 73, 74, 75, 76, 120, 127,
           128, 129, 200, 300]


@pytest.mark.parametrize("dst", [h2c.IC_DST, b"X", bytes(range(1, 256)), G1_DST],
                         ids=["ic-dst", "1-byte-dst", "255-byte-dst", "rfc-dst"])
def test_matches_oracle_over_lengths(dst):
    print("case: hash_to_g1 vs oracle, DST length %d" % len(dst), flush=True)
    msgs = _messages(len(dst), LENGTHS)
    out, ident = blsgpu.hash_to_g1(msgs, dst)
    neg, _ = blsgpu.hash_to_g1(msgs, dst, negate=True)
    assert not ident.any()
    for i, m in enumerate(msgs):
        x, y = h2c.hash_to_g1(m, dst)
        assert _affine(out[i]) == (x, y), (i, len(m))
        assert _affine(neg[i]) == (x, (-y) % O.P), i


@pytest.mark.parametrize("n", [1, 3, 17])
def test_small_batches_against_oracle(n):
    # one quad per message: a single message leaves 60 lanes of its wavefront repeating
    # it, 3 and 17 end a wavefront part way through; every output against the oracle
    print("case: hash_to_g1 n=%d" % n, flush=True)
    msgs = _messages(90 + n, [(7 * i) % 70 for i in range(n)])
    out, ident = blsgpu.hash_to_g1(msgs, h2c.IC_DST)
    assert not ident.any()
    for i, m in enumerate(msgs):
        assert _affine(out[i]) == h2c.hash_to_g1(m, h2c.IC_DST), (n, i)


def test_fused_wire_path():
    print("case: verify_messages M=40", flush=True)
    blsgpu.load().blsgpu_reset_keys()
    sks, pks, _ = O.make_keys(61, 4)
    handles = blsgpu.register_keys_affine(pks)
    n = 40
    key_idx = (np.arange(n) % 4).astype(np.int32)
    msgs = _messages(61, [32 + (i % 7) * 9 for i in range(n)])
    sigs = np.zeros((n, 48), np.uint8)
    nhm = np.zeros((n, 12), np.uint64)
    wrong = {5, 17, 33}
    for i in range(n):
        x, y = h2c.hash_to_g1(msgs[i], h2c.IC_DST)
        hm = np.concatenate([O.mont(x), O.mont(y)])
        nhm[i] = O.g1_neg(hm)
        signed = hm
        if i in wrong:  # a signature over another message
            x2, y2 = h2c.hash_to_g1(msgs[i] + b"!", h2c.IC_DST)
            signed = np.concatenate([O.mont(x2), O.mont(y2)])
        s, _ = O.g1_mul(signed, sks[key_idx[i]])
        sigs[i] = O.g1_compress(s)
    sigs[9, 0] &= 0x7F  # a decode error passes through
    res, gt = blsgpu.verify_messages(sigs, msgs, h2c.IC_DST, handles[key_idx], want_gt=True)
    want = np.ones(n, np.int32)
    want[sorted(wrong)] = 0
    want[9] = -1
    assert (res == want).all(), res
    res2, gt2 = blsgpu.verify_compressed(sigs, nhm, handles[key_idx], want_gt=True)
    assert (res2 == res).all() and (gt2 == gt).all()


def test_rejects_bad_message_arguments():
    with pytest.raises(RuntimeError):
        blsgpu.hash_to_g1([b"abc"], b"")  # the DST must be 1..255 bytes
    with pytest.raises(RuntimeError):
        blsgpu.hash_to_g1([b"abc"], bytes(256))
    sig = np.zeros((1, 48), np.uint8)
    with pytest.raises(RuntimeError):
        blsgpu.verify_messages(sig, [b"abc"], b"", np.zeros(1, np.int32))


def test_large_batch_on_curve_and_sampled_against_oracle():
    # 3,000 messages (188 workgroups of 16 lane quads, the last one half full): every
    # output on E and in the r-torsion ([r]P checked on a sample), a sample against the
    # oracle
    print("case: hash_to_g1 n=3000", flush=True)
    n = 3000
    msgs = [bytes(O.rand_bytes(97, 8_000_000 + i, 32)) for i in range(n)]
    out, ident = blsgpu.hash_to_g1(msgs, h2c.IC_DST)
    assert not ident.any()
    pts = [_affine(row) for row in out]
    for x, y in pts:
        assert (y * y - x ** 3 - 4) % O.P == 0
    assert len(set(pts)) == n
    for i in list(range(0, n, 211)) + [n - 1]:
        assert pts[i] == h2c.hash_to_g1(msgs[i], h2c.IC_DST), i
    for i in (0, n // 2, n - 1):
        assert h2c.ec_mul(pts[i], O.R_ORDER, 0) is None
