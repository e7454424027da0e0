"""The hash-to-G1 oracle (oracle/h2c.py) against RFC 9380's published test vectors, and
the device's isogeny constants (derived by gen_constants.py as rational maps) against
the oracle's point-formula isogeny."""
import os
import random
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "icp-gpu_amd", "csrc"))
import h2c  # noqa: E402

P = h2c.P
# RFC 9380 appendix K.1: expand_message_xmd(SHA-256), DST below, len_in_bytes = 0x20
XMD_DST = b"QUUX-V01-CS02-with-expander-SHA256-128"
XMD_VECTORS = [
    (b"", "68a985b87eb6b46952128911f2a4412bbc302a9d759667f87f7a21d803f07235"),
    (b"abc", "d8ccab23b5985ccea865c6c97b6e5b8350e794e603b4b97902f53a8a0d605615"),
]
# RFC 9380 appendix J.9.1: suite BLS12381G1_XMD:SHA-256_SSWU_RO_
G1_DST = b"QUUX-V01-CS02-with-BLS12381G1_XMD:SHA-256_SSWU_RO_"
G1_U_EMPTY = (
    0x0ba14bd907ad64a016293ee7c2d276b8eae71f25a4b941eece7b0d89f17f75cb3ae5438a614fb61d6835ad59f29c564f,
    0x019b9bd7979f12657976de2884c7cce192b82c177c80e0ec604436a7f538d231552f0d96d9f7babe5fa3b19b3ff25ac9,
)
G1_VECTORS = [
    (b"",
     0x052926add2207b76ca4fa57a8734416c8dc95e24501772c814278700eed6d1e4e8cf62d9c09db0fac349612b759e79a1,
     0x08ba738453bfed09cb546dbb0783dbb3a5f1f566ed67bb6be0e8c67e2e81a4cc68ee29813bb7994998f3eae0c9c6a265),
    (b"abc",
     0x03567bc5ef9c690c2ab2ecdf6a96ef1c139cc0b2f284dca0a9a7943388a49a3aee664ba5379a7655d3c68900be2f6903,
     0x0b9c15f3fe6e5cf4211f346271d7b01c8f3b28be689c8429c85b67af215533311f0b8dfaaa154fa6b88176c229f2885d),
]
R_ORDER = 0x73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001


@pytest.mark.parametrize("msg,want", XMD_VECTORS)
def test_expand_message_xmd_vectors(msg, want):
    assert h2c.expand_message_xmd(msg, XMD_DST, 32).hex() == want


def test_hash_to_field_vector():
    assert tuple(h2c.hash_to_field(b"", G1_DST, 2)) == G1_U_EMPTY


@pytest.mark.parametrize("msg,x,y", G1_VECTORS)
def test_hash_to_g1_vectors(msg, x, y):
    assert h2c.hash_to_g1(msg, G1_DST) == (x, y)


def _on_e(pt):
    x, y = pt
    return (y * y - x ** 3 - 4) % P == 0


def test_map_and_isogeny_properties():
    rnd = random.Random(3)
    for _ in range(6):
        u = rnd.randrange(P)
        q = h2c.map_to_curve_sswu(u)
        assert (q[1] ** 2 - q[0] ** 3 - h2c.A * q[0] - h2c.B) % P == 0
        assert q[1] % 2 == u % 2  # sgn0(y) = sgn0(u)
        e = h2c.iso_map(q)
        assert _on_e(e)
    # the isogeny is a group homomorphism
    a, b = h2c.map_to_curve_sswu(5), h2c.map_to_curve_sswu(7)
    assert h2c.iso_map(h2c.ec_add(a, b, h2c.A)) == h2c.ec_add(h2c.iso_map(a), h2c.iso_map(b), 0)
    # hash outputs are in the r-torsion
    pt = h2c.hash_to_g1(b"ic certification", h2c.IC_DST)
    assert _on_e(pt) and h2c.ec_mul(pt, R_ORDER, 0) is None


def test_device_isogeny_maps_equal_point_formula():
    import gen_constants as G
    xn, xd, yn, yd = G.iso11_maps()

    def ev(c, x):
        r = 0
        for a in reversed(c):
            r = (r * x + a) % P
        return r

    rnd = random.Random(4)
    for _ in range(4):
        x, y = h2c.map_to_curve_sswu(rnd.randrange(P))
        want = h2c.iso_map((x, y))
        got = (ev(xn, x) * h2c.inv(ev(xd, x)) % P, y * ev(yn, x) * h2c.inv(ev(yd, x)) % P)
        assert got == want


def test_constants_header_is_fresh(tmp_path):
    # field_constants.hpp is what gen_constants.py derives (its asserts also check the
    # reference's constants and the RFC's published isogeny coefficients)
    import gen_constants as G
    out = tmp_path / "field_constants.hpp"
    G.main(str(out))
    header = os.path.join(ROOT, "icp-gpu_amd", "csrc", "field_constants.hpp")
    assert out.read_text() == open(header).read()
