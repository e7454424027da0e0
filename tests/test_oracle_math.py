"""The CPU oracle against what the mathematics fixes: big-integer closed forms,
schoolbook tower products (tests/pyfield.py), Frobenius as the p-th power,
bilinearity, the group order, and brute-force group checks."""
import hashlib
import random

import numpy as np
import pytest

import oracle as O
import pyfield as F

P = O.P
RI = pow(O.R384, -1, P)
EDGE = [0, 1, 2, P - 1, P - 2, (P - 1) // 2, O.R384 % P, (1 << 380) % P]


def _fp(x):
    return O.mont(x)


def _rand_tower(rng, n):
    return [(rng.randrange(P), rng.randrange(P)) for _ in range(n)]


def _tower_arr(t):
    return np.array(F.tower_to_limbs72(t) if len(t) == 6 else
                    [v for c in t for x in c for v in O.from_int(x * O.R384 % P)], dtype=np.uint64)


@pytest.mark.parametrize("op", ["mul", "add", "sub"])
def test_fp_ops_closed_form(op):
    rng = random.Random(1)
    vals = EDGE + [rng.randrange(P) for _ in range(40)]
    f = getattr(O.lib(), "orc_fp_" + op)
    for a in vals:
        b = vals[rng.randrange(len(vals))]
        out = np.zeros(6, np.uint64)
        f(O.ptr(_fp(a)), O.ptr(_fp(b)), O.ptr(out))
        want = {"mul": a * b, "add": a + b, "sub": a - b}[op] % P
        assert (out == _fp(want)).all(), (op, a, b)


def test_fp_neg_inv_sqrt():
    rng = random.Random(2)
    for a in EDGE + [rng.randrange(P) for _ in range(20)]:
        out = np.zeros(6, np.uint64)
        O.lib().orc_fp_neg(O.ptr(_fp(a)), O.ptr(out))
        assert (out == _fp(-a % P)).all()
        if a:
            O.lib().orc_fp_inv(O.ptr(_fp(a)), O.ptr(out))
            assert (out == _fp(pow(a, -1, P))).all()
        ok = O.lib().orc_fp_sqrt(O.ptr(_fp(a)), O.ptr(out))
        euler = a == 0 or pow(a, (P - 1) // 2, P) == 1
        assert bool(ok) == euler
        if ok:
            assert O.unmont(out) ** 2 % P == a


def test_toy_tower_kats():
    # cuda/test_fp2.cu:110-124, cuda/test_fp6.cu:118-158, cuda/test_fp12.cu:124-163
    k = O.fixtures()["toy_kats"]
    def arr(ints):
        return np.concatenate([_fp(x % P) for x in ints])
    out = np.zeros(12, np.uint64)
    O.lib().orc_fp2_mul(O.ptr(arr(k["fp2_mul"]["a"])), O.ptr(arr(k["fp2_mul"]["b"])), O.ptr(out))
    assert (out == arr(k["fp2_mul"]["out"])).all()
    out6 = np.zeros(36, np.uint64)
    O.lib().orc_fp6_mul(O.ptr(arr(k["fp6_mul"]["a"])), O.ptr(arr(k["fp6_mul"]["b"])), O.ptr(out6))
    assert (out6 == arr(k["fp6_mul"]["out"])).all()
    out12 = np.zeros(72, np.uint64)
    O.lib().orc_fp12_mul(O.ptr(arr(k["fp12_mul"]["a"])), O.ptr(arr(k["fp12_mul"]["b"])), O.ptr(out12))
    for idx, val in k["fp12_mul"]["out_fp2"].items():
        i = int(idx)
        assert (out12[12 * i:12 * i + 12] == arr(val)).all()


def test_fp12_mul_sqr_inv_against_schoolbook():
    rng = random.Random(3)
    for _ in range(6):
        a, b = _rand_tower(rng, 6), _rand_tower(rng, 6)
        A, B = np.array(F.tower_to_limbs72(a), np.uint64), np.array(F.tower_to_limbs72(b), np.uint64)
        want = F.f12_to_tower(F.f12mul_w(F.f12_from_tower(a), F.f12_from_tower(b)))
        out = np.zeros(72, np.uint64)
        O.lib().orc_fp12_mul(O.ptr(A), O.ptr(B), O.ptr(out))
        assert F.limbs72_to_tower(out) == want
        sq = F.f12_to_tower(F.f12mul_w(F.f12_from_tower(a), F.f12_from_tower(a)))
        O.lib().orc_fp12_sqr(O.ptr(A), O.ptr(out))
        assert F.limbs72_to_tower(out) == sq
        O.lib().orc_fp12_inv(O.ptr(A), O.ptr(out))
        prod = F.f12mul_w(F.f12_from_tower(a), F.f12_from_tower(F.limbs72_to_tower(out)))
        assert prod == [(1, 0)] + [(0, 0)] * 5


def test_frobenius_is_pth_power():
    rng = random.Random(4)
    a = _rand_tower(rng, 6)
    A = np.array(F.tower_to_limbs72(a), np.uint64)
    out = np.zeros(72, np.uint64)
    O.lib().orc_fp12_frob(O.ptr(A), O.ptr(out))
    assert F.limbs72_to_tower(out) == F.f12_to_tower(F.f12pow_w(F.f12_from_tower(a), P))
    # frob^6 = conj, frob^12 = id (cuda/test_frobenius.cu:112-125)
    x = A.copy()
    for i in range(12):
        O.lib().orc_fp12_frob(O.ptr(x.copy()), O.ptr(x))
        if i == 5:
            c = np.zeros(72, np.uint64)
            O.lib().orc_fp12_conj(O.ptr(A), O.ptr(c))
            assert (x == c).all()
    assert (x == A).all()


def test_mul_by_014_is_sparse_product():
    rng = random.Random(5)
    f = _rand_tower(rng, 6)
    c0, c1, c4 = _rand_tower(rng, 3)
    F_ = np.array(F.tower_to_limbs72(f), np.uint64)
    conv = lambda c: np.concatenate([_fp(c[0]), _fp(c[1])])
    out = np.zeros(72, np.uint64)
    O.lib().orc_fp12_mul_by_014(O.ptr(F_), O.ptr(conv(c0)), O.ptr(conv(c1)), O.ptr(conv(c4)), O.ptr(out))
    # line = c0 + c1 v + c4 v w   (tower positions c0.c0, c0.c1, c1.c1)
    line = [c0, c1, (0, 0), (0, 0), c4, (0, 0)]
    want = F.f12_to_tower(F.f12mul_w(F.f12_from_tower(f), F.f12_from_tower(line)))
    assert F.limbs72_to_tower(out) == want


def _pairing(p, q):
    out = np.zeros(72, np.uint64)
    O.lib().orc_pairing(O.ptr(p), O.ptr(q), O.ptr(out))
    return out


def test_single_pair_miller_matches_reference_value():
    # cuda/g2_coeffs.h:551-565 is the Miller loop of e(G1, G2)
    fx = O.fixtures()
    ml = np.zeros(72, np.uint64)
    O.lib().orc_miller1(O.ptr(O.g1_gen()), O.ptr(O.u64(fx["g2_gen_coeffs"])), O.ptr(ml))
    assert (ml == O.u64(fx["miller_g1_g2"])).all()
    assert (_pairing(O.g1_gen(), O.g2_gen()) == O.u64(fx["pairing_g1_g2"])).all()


def test_bilinearity_and_order():
    e = F.f12_from_tower(F.limbs72_to_tower(_pairing(O.g1_gen(), O.g2_gen())))
    p3, _ = O.g1_mul(O.g1_gen(), 3)
    q5, _ = O.g2_mul(O.g2_gen(), 5)
    e35 = F.f12_from_tower(F.limbs72_to_tower(_pairing(p3, q5)))
    assert e35 == F.f12pow_w(e, 15)
    assert F.f12pow_w(e, O.R_ORDER) == [(1, 0)] + [(0, 0)] * 5
    assert e != [(1, 0)] + [(0, 0)] * 5


def _py_g1_mul(k, pt):
    def add(a, b):
        if a is None:
            return b
        if b is None:
            return a
        (x1, y1), (x2, y2) = a, b
        if x1 == x2:
            if (y1 + y2) % P == 0:
                return None
            lam = 3 * x1 * x1 * pow(2 * y1, -1, P) % P
        else:
            lam = (y2 - y1) * pow(x2 - x1, -1, P) % P
        x3 = (lam * lam - x1 - x2) % P
        return (x3, (lam * (x1 - x3) - y1) % P)
    r = None
    while k:
        if k & 1:
            r = add(r, pt)
        pt = add(pt, pt)
        k >>= 1
    return r


def test_decompression_round_trip_and_rejects():
    rng = random.Random(6)
    for i in range(8):
        p, _ = O.g1_mul(O.g1_gen(), rng.randrange(1, O.R_ORDER))
        b = O.g1_compress(p)
        rc, q = O.g1_decompress(b)
        assert rc == 0 and (q == p).all()
        # encoding: x big-endian with flag bits, sort flag = y > (p-1)/2
        x, y = O.unmont(p[:6]), O.unmont(p[6:])
        assert int.from_bytes(b.tobytes(), "big") & ((1 << 381) - 1) == x
        assert bool(b[0] & 0x20) == (y > (P - 1) // 2) and b[0] & 0x80
    # x^3 + 4 with no square root (Euler criterion)
    x = O.unmont(O.rand_x(9, 0, 0))
    assert pow((x ** 3 + 4) % P, (P - 1) // 2, P) == P - 1
    b = bytearray(x.to_bytes(48, "big")); b[0] |= 0x80
    assert O.g1_decompress(np.frombuffer(bytes(b), np.uint8))[0] == -3
    # on the curve, outside the subgroup: r*P != O by brute force in Python
    x = O.unmont(O.rand_x(9, 1, 1))
    b = bytearray(x.to_bytes(48, "big")); b[0] |= 0x80
    rc, q = O.g1_decompress(np.frombuffer(bytes(b), np.uint8), check=False)
    assert rc == 0
    assert _py_g1_mul(O.R_ORDER, (O.unmont(q[:6]), O.unmont(q[6:]))) is not None
    assert O.g1_decompress(np.frombuffer(bytes(b), np.uint8))[0] == -4
    # a subgroup point passes r*P == O in the same model
    p, _ = O.g1_mul(O.g1_gen(), 12345)
    assert _py_g1_mul(O.R_ORDER, (O.unmont(p[:6]), O.unmont(p[6:]))) is None
    # x >= p, bad flags, identity
    b = bytearray(P.to_bytes(48, "big")); b[0] |= 0x80
    assert O.g1_decompress(np.frombuffer(bytes(b), np.uint8))[0] == -2
    assert O.g1_decompress(np.frombuffer(bytes(48), np.uint8))[0] == -1
    ident = np.zeros(48, np.uint8); ident[0] = 0xc0
    assert O.g1_decompress(ident)[0] == 1
    ident[0] = 0xe0
    assert O.g1_decompress(ident)[0] == -1


def test_hash_to_g1_definition():
    # oracle/src/main.rs:383-397: s = LE(SHA-256("ic-bls-test-hash" || m)) with byte 31 &= 0x0f
    out = np.zeros(32, np.uint8)
    O.lib().orc_sha256(O.ptr(np.frombuffer(b"abc", np.uint8).copy()), 3, O.ptr(out))
    assert out.tobytes().hex() == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"
    msg = b"test-message-0"
    h = bytearray(hashlib.sha256(b"ic-bls-test-hash" + msg).digest())
    h[31] &= 0x0F
    s = int.from_bytes(bytes(h), "little")
    want, _ = O.g1_mul(O.g1_gen(), s)
    assert (O.hash_to_g1(msg) == want).all()


def test_signature_scheme_end_to_end():
    sks, pks, coeffs = O.make_keys(7, 2)
    g2c = O.g2_prepare(O.g2_gen())
    msg = b"m" * 32
    s, h = O.sign(sks[0], msg)
    ok, _ = O.verify(s, O.g1_neg(h), g2c, coeffs[0])
    assert ok
    ok, _ = O.verify(s, O.g1_neg(h), g2c, coeffs[1])
    assert not ok


def test_identity_signature_is_paired_as_the_identity():
    """The fused path's identity branch (a 0xc0 signature).  ic_bls12_381's
    multi_miller_loop leaves f unchanged for an identity term, and e(O, G2) = 1, so the
    item's GT must be e(-H(m), pk) alone, and (bilinearity, with sig = sk H(m) the valid
    signature) also the inverse, i.e. the conjugate, of e(sig, G2).  A branch that paired
    the identity as the flagless point (0, 0) or (0, 1) (the legacy marshalling drops the
    flag) would multiply f by the lines' constant terms and fail both checks."""
    sks, pks, coeffs = O.make_keys(11, 3)
    g2c = O.g2_prepare(O.g2_gen())
    comp, unc, nhm, msgs = O.make_batch(11, 6, sks, [0, 1, 2, 0, 1, 2])
    ident = [0, 3, 5]  # first, middle and last items take the identity signature
    for i in ident:
        comp[i] = O.IDENTITY_G1
    key_idx = np.array([0, 1, 2, 0, 1, 2], np.int32)
    res, gts = O.verify_compressed_batch(comp, nhm, g2c, coeffs, key_idx, threads=2)
    one = np.zeros(72, np.uint64)
    one[:6] = O.mont(1)
    for i in range(6):
        k = int(key_idx[i])
        if i in ident:
            assert res[i] == 0, i
            assert (gts[i] == _pairing(nhm[i], pks[k])).all(), i
            assert not (gts[i] == one).all()
            conj = np.zeros(72, np.uint64)
            O.lib().orc_fp12_conj(O.ptr(_pairing(unc[i], O.g2_gen())), O.ptr(conj))
            assert (gts[i] == conj).all(), i
        else:
            # the other items of the batch keep the two-pair result: valid, GT = 1
            assert res[i] == 1 and (gts[i] == one).all(), i
    # an identity signature fails against every key, whatever the message
    res2, _ = O.verify_compressed_batch(np.stack([O.IDENTITY_G1] * 3), nhm[:3], g2c, coeffs,
                                        np.arange(3, dtype=np.int32), threads=1)
    assert res2.tolist() == [0, 0, 0]
    # rejected encodings export zeros
    bad = np.stack([O.IDENTITY_G1, O.IDENTITY_G1])
    bad[1, 0] = 0xE0  # identity with the sort flag: bad flags
    res3, gts3 = O.verify_compressed_batch(bad, nhm[:2], g2c, coeffs, np.zeros(2, np.int32), threads=1)
    assert res3.tolist() == [0, -1]
    assert (gts3[1] == 0).all()
