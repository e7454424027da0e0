"""G2 point compression in the oracle (the encoding of ic_bls12_381's G2Affine
to_compressed / from_compressed, which key registration from 96-byte public keys
decodes): pinned by the published compressed form of the standard generator, round
trips, the sort-flag rule, every rejected encoding, an independent solvability test
(Euler's criterion on the Fp norm) and the brute-force [r]Q == O subgroup decision."""
import random

import numpy as np
import pytest

import oracle as O

P = O.P
# the standard BLS12-381 G2 generator (x = x.c0 + x.c1 u), published with the curve
GEN_X_C0 = 0x024AA2B2F08F0A91260805272DC51051C6E47AD4FA403B02B4510B647AE3D1770BAC0326A805BBEFD48056C8C121BDB8
GEN_X_C1 = 0x13E02B6052719F607DACD3A088274F65596BD0D09920B61AB5DA61BBDC7F5049334CF11213945D57E5AC7D055D042B7E


def test_generator_encoding():
    g = O.g2_gen()
    want = bytearray(GEN_X_C1.to_bytes(48, "big") + GEN_X_C0.to_bytes(48, "big"))
    want[0] |= 0x80  # compressed; the generator's y is not lexicographically largest
    assert bytes(O.g2_compress(g)) == bytes(want)
    assert bytes(want[:4]).hex() == "93e02b60"
    rc, q = O.g2_decompress(np.frombuffer(bytes(want), np.uint8))
    assert rc == 0 and (q == g).all()
    # -G: the same x with the sort flag
    neg = g.copy()
    ny0, ny1 = (P - O.unmont(g[12:18])) % P, (P - O.unmont(g[18:24])) % P
    neg[12:18], neg[18:24] = O.mont(ny0), O.mont(ny1)
    enc = O.g2_compress(neg)
    assert enc[0] == want[0] | 0x20 and bytes(enc[1:]) == bytes(want[1:])
    rc, q = O.g2_decompress(enc)
    assert rc == 0 and (q == neg).all()


def test_round_trips_and_sort_flag():
    rnd = random.Random(5)
    g = O.g2_gen()
    for i in range(6):
        q, inf = O.g2_mul(g, rnd.randrange(1, O.R_ORDER))
        assert not inf
        enc = O.g2_compress(q)
        rc, back = O.g2_decompress(enc)
        assert rc == 0 and (back == q).all()
        y1, y0 = O.unmont(q[18:24]), O.unmont(q[12:18])
        largest = y1 > (P - 1) // 2 or (y1 == 0 and y0 > (P - 1) // 2)
        assert bool(enc[0] & 0x20) == largest


def _norm_is_square(x0, x1):
    # a = x^3 + 4(1 + u) in Fp2 = Fp[u]/(u^2 + 1); a is a square iff its norm is one in Fp
    x2 = ((x0 * x0 - x1 * x1) % P, 2 * x0 * x1 % P)
    x3 = ((x2[0] * x0 - x2[1] * x1) % P, (x2[0] * x1 + x2[1] * x0) % P)
    a0, a1 = (x3[0] + 4) % P, (x3[1] + 4) % P
    n = (a0 * a0 + a1 * a1) % P
    return n == 0 or pow(n, (P - 1) // 2, P) == 1


def test_solvability_and_off_subgroup_points():
    rnd = random.Random(9)
    seen = {0: 0, -3: 0}
    for i in range(24):
        x0, x1 = rnd.randrange(P), rnd.randrange(P)
        enc = bytearray(x1.to_bytes(48, "big") + x0.to_bytes(48, "big"))
        enc[0] |= 0x80
        rc, q = O.g2_decompress(np.frombuffer(bytes(enc), np.uint8), check=False)
        assert rc in (0, -3)
        assert (rc == 0) == _norm_is_square(x0, x1)
        seen[rc] += 1
        if rc == 0:
            assert O.g2_is_on_curve(q)
            # a random curve point misses the r-subgroup: the checked form rejects it
            assert not O.g2_in_subgroup(q)
            assert O.g2_decompress(np.frombuffer(bytes(enc), np.uint8))[0] == -4
    assert seen[0] > 4 and seen[-3] > 4


def test_y_squared_in_fp():
    """x = x0 + 2u with Im(x^3) = -4, so y^2 = x^3 + 4(1 + u) has no u part: the decoded
    y is real when y^2 is a square in Fp and purely imaginary otherwise; it squares back
    to y^2 and follows the sort flag (the case the device's square root by the norm
    handles apart)"""
    r = (4 - 4 * pow(2, -1, P)) * pow(3, -1, P) % P
    x0 = pow(r, (P + 1) // 4, P)
    assert x0 * x0 % P == r
    kinds = set()
    for x0s in (x0, P - x0):
        y2 = (x0s ** 3 - 12 * x0s + 4) % P  # Re(x^3) + 4 with x1 = 2
        assert (3 * x0s * x0s * 2 - 8 + 4) % P == 0  # Im(x^3) + 4 = 0
        for sflag in (0, 0x20):
            enc = bytearray((2).to_bytes(48, "big") + x0s.to_bytes(48, "big"))
            enc[0] |= 0x80 | sflag
            rc, q = O.g2_decompress(np.frombuffer(bytes(enc), np.uint8), check=False)
            assert rc == 0 and O.g2_is_on_curve(q)
            y0, y1 = O.unmont(q[12:18]), O.unmont(q[18:24])
            if pow(y2, (P - 1) // 2, P) == 1:
                assert y1 == 0 and y0 * y0 % P == y2
                kinds.add("real")
            else:
                assert y0 == 0 and (P - y1 * y1) % P == y2
                kinds.add("imaginary")
            largest = y1 > (P - 1) // 2 or (y1 == 0 and y0 > (P - 1) // 2)
            assert largest == bool(sflag)
    assert kinds == {"real", "imaginary"}


def test_rejected_encodings():
    g = O.g2_gen()
    enc = bytearray(O.g2_compress(g))
    no_c = bytearray(enc)
    no_c[0] &= 0x7F
    assert O.g2_decompress(np.frombuffer(bytes(no_c), np.uint8))[0] == -1
    inf = bytearray(96)
    inf[0] = 0xC0
    assert O.g2_decompress(np.frombuffer(bytes(inf), np.uint8))[0] == 1
    bad_inf = bytearray(inf)
    bad_inf[0] |= 0x20
    assert O.g2_decompress(np.frombuffer(bytes(bad_inf), np.uint8))[0] == -1
    bad_inf = bytearray(inf)
    bad_inf[95] = 1
    assert O.g2_decompress(np.frombuffer(bytes(bad_inf), np.uint8))[0] == -1
    for half in (0, 48):  # x.c1 or x.c0 >= p
        big = bytearray(enc)
        big[half:half + 48] = P.to_bytes(48, "big")
        big[0] |= 0x80
        assert O.g2_decompress(np.frombuffer(bytes(big), np.uint8))[0] == -2
