"""GPU G2 key registration from 96-byte compressed public keys
(blsgpu_register_keys_compressed): per-key decode status equal to the oracle's
from_compressed / from_compressed_unchecked decision, line tables equal to the oracle's
G2Prepared of the decoded point, consecutive handles for the accepted keys, and
signatures verifying through those handles."""
import numpy as np
import pytest

import blsgpu
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    assert blsgpu.init() == 0, blsgpu.last_error()


def _neg(q):
    n = q.copy()
    for a, b in ((12, 18), (18, 24)):
        n[a:b] = O.mont((O.P - O.unmont(q[a:b])) % O.P)
    return n


def _special_encodings():
    """every rejected form, the identity, both sort flags and an off-subgroup point"""
    g = O.g2_gen()
    enc, names = [], []

    def add(name, b):
        names.append(name)
        enc.append(np.frombuffer(bytes(b), np.uint8))

    good = bytearray(O.g2_compress(g))
    add("generator", good)
    add("negated generator", O.g2_compress(_neg(g)))
    ident = bytearray(96)
    ident[0] = 0xC0
    add("identity", ident)
    b = bytearray(good)
    b[0] &= 0x7F
    add("compression flag clear", b)
    b = bytearray(good)
    b[0] |= 0x40
    add("identity flag with x", b)
    b = bytearray(ident)
    b[0] |= 0x20
    add("identity with sort flag", b)
    b = bytearray(O.P.to_bytes(48, "big") + bytes(good[48:]))
    b[0] |= 0x80
    add("x.c1 = p", b)
    b = bytearray(bytes(good[:48]) + O.P.to_bytes(48, "big"))
    add("x.c0 = p", b)
    b = bytearray((O.P - 1).to_bytes(48, "big") + (O.P - 1).to_bytes(48, "big"))
    b[0] |= 0x80
    add("x = (p-1)(1+u)", b)
    k = 0
    while True:  # an x with no point on the curve
        xb = bytearray(O.rand_bytes(41, 5_000_000 + k, 96))
        xb[0] = (xb[0] & 0x0F) | 0x80
        xb[48] &= 0x0F
        if O.g2_decompress(np.frombuffer(bytes(xb), np.uint8), check=False)[0] == -3:
            add("not on curve", xb)
            break
        k += 1
    for i in range(2):
        xb, _ = O.g2_off_subgroup_point(41, i)
        add("off subgroup %d" % i, xb)
        nb = bytearray(xb)
        nb[0] ^= 0x20
        add("off subgroup %d, other root" % i, nb)
    # x = x0 + 2u with Im(x^3) = x1 (3 x0^2 - x1^2) = -4, so that y^2 = x^3 + 4(1 + u) lies
    # in Fp: the square root's zero-imaginary-part case.  The two roots x0 give y^2 a
    # square in Fp (y real) and a non-square (y purely imaginary); both sort flags.
    r = (4 - 4 * pow(2, -1, O.P)) * pow(3, -1, O.P) % O.P
    assert pow(r, (O.P - 1) // 2, O.P) == 1
    x0 = pow(r, (O.P + 1) // 4, O.P)
    for x0s in (x0, O.P - x0):
        real = pow((x0s ** 3 - 12 * x0s + 4) % O.P, (O.P - 1) // 2, O.P) == 1
        for sflag in (0, 0x20):
            xb = bytearray((2).to_bytes(48, "big") + x0s.to_bytes(48, "big"))
            xb[0] |= 0x80 | sflag
            add("y^2 in Fp, y %s, sort flag %d" % ("real" if real else "imaginary", sflag >> 5), xb)
    return names, np.stack(enc)


def _random_encodings(seed, n):
    """random 96-byte strings with random flag bits: x >= p, off-curve, off-subgroup"""
    out = []
    for i in range(n):
        b = bytearray(O.rand_bytes(seed, 6_000_000 + i, 96))
        b[0] = (b[0] & 0x1F) | (0x80 if i % 8 else 0) | (0x20 if i % 3 == 0 else 0)
        if i % 5:
            b[48] &= 0x1F
        out.append(np.frombuffer(bytes(b), np.uint8))
    return np.stack(out)


def _check_against_oracle(enc, check_subgroup):
    blsgpu.load().blsgpu_reset_keys()
    h, st = blsgpu.register_keys_compressed(enc, check_subgroup=check_subgroup)
    nxt = 0
    for i in range(enc.shape[0]):
        rc, q = O.g2_decompress(enc[i], check=check_subgroup)
        assert st[i] == rc, (i, st[i], rc)
        if rc == 0:
            assert h[i] == nxt, i
            nxt += 1
            assert (blsgpu.key_coeffs(int(h[i])) == O.g2_prepare(q)).all(), i
        else:
            assert h[i] == -1, i
    assert blsgpu.load().blsgpu_num_keys() == nxt
    return h, st


@pytest.mark.parametrize("check_subgroup", [True, False])
def test_special_encodings_match_oracle(check_subgroup):
    names, enc = _special_encodings()
    print("case: G2 special encodings, check_subgroup=%s" % check_subgroup, flush=True)
    h, st = _check_against_oracle(enc, check_subgroup)
    got = dict(zip(names, st.tolist()))
    assert got["generator"] == 0 and got["negated generator"] == 0
    assert got["identity"] == 1
    for name in ("compression flag clear", "identity flag with x", "identity with sort flag"):
        assert got[name] == -1, name
    for name in ("x.c1 = p", "x.c0 = p"):
        assert got[name] == -2, name
    assert got["not on curve"] == -3
    for name in ("off subgroup 0", "off subgroup 1"):
        assert got[name] == (-4 if check_subgroup else 0), name


def test_random_encodings_match_oracle():
    enc = _random_encodings(43, 96)
    print("case: G2 random encodings n=96", flush=True)
    _, st = _check_against_oracle(enc, True)
    _, st_u = _check_against_oracle(enc, False)
    # the sample reaches every decision
    for code in (-1, -2, -3, -4):
        assert (st == code).any(), code
    assert ((st == -4) == (st_u == 0)).all()


def test_registered_keys_verify_signatures():
    print("case: G2 compressed keys, verify M=48", flush=True)
    sks, pks, coeffs = O.make_keys(45, 6)
    enc = np.stack([O.g2_compress(pk) for pk in pks])
    # a rejected key between accepted ones shifts nothing: handles stay in input order
    bad = bytearray(enc[2])
    bad[0] &= 0x7F
    enc = np.concatenate([enc[:2], np.frombuffer(bytes(bad), np.uint8)[None], enc[2:]])
    blsgpu.load().blsgpu_reset_keys()
    h, st = blsgpu.register_keys_compressed(enc)
    assert st.tolist() == [0, 0, -1, 0, 0, 0, 0]
    handles = h[[0, 1, 3, 4, 5, 6]]
    assert handles.tolist() == [0, 1, 2, 3, 4, 5]
    for k in range(6):
        assert (blsgpu.key_coeffs(int(handles[k])) == coeffs[k]).all(), k
    key_idx = np.arange(48, dtype=np.int32) % 6
    inv = {7: "wrong_msg", 30: "wrong_msg"}
    comp, unc, nhm, _ = O.make_batch(45, 48, sks, key_idx, inv)
    res = blsgpu.verify_compressed(comp, nhm, handles[key_idx].astype(np.int32))
    want = np.ones(48, np.int32)
    want[[7, 30]] = 0
    assert (res == want).all()
