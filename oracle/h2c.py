"""Hash to G1 for BLS12-381, RFC 9380 suite BLS12381G1_XMD:SHA-256_SSWU_RO_ (the
ciphersuite ic_bls12_381's `HashToCurve<ExpandMsgXmd<Sha256>>` implements; the
reference calls it at ic-bench/src/main.rs:65-72 and its callers hash before the GPU
step, ic-integration/patches/ic-gpu-consensus.patch:454-457).

TEST INFRASTRUCTURE ONLY: the CPU restatement the GPU kernels are checked against.
Nothing in the library imports it.  Pure Python integers, for small cases.

Pinned by the RFC's own test vectors (tests/test_oracle_h2c.py): expand_message_xmd
(appendix K.1), hash_to_field and hash_to_curve outputs (appendix J.9.1).  The
isogeny is evaluated here by Velu's point formula, phi(P) = (x_P + sum (x_{P+Q} -
x_Q), y_P + sum (y_{P+Q} - y_Q)) over the ten non-zero points Q of the kernel, a
different computation from the rational maps the device evaluates.
"""
import hashlib

P = 0x1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f6241eabfffeb153ffffb9feffffffffaaab
BLS_X = 0xd201000000010000  # x = -BLS_X
H_EFF = BLS_X + 1  # RFC 9380 sec. 8.8.1: h_eff = 1 - x
# E': y^2 = x^3 + A x + B, the curve the simplified SWU map targets (sec. 8.8.1)
A = 0x144698a3b8e9433d693a02c96d4982b0ea985383ee66a8d8e8981aefd881ac98936f8da0e0f97f5cf428082d584c1d
B = 0x12e2908d11688030018b12e8753eee3b2016c1f0f24f4070a0b9c14fcef35ef55a23215a316ceaa5d1cc48e98e172be0
Z = 11
# the ciphersuite tag the reference's benchmark hashes with (ic-bench/src/main.rs:68)
IC_DST = b"BLS_SIG_BLS12381G2_XMD:SHA-256_SSWU_RO_NUL_"


def inv(a):
    return pow(a, P - 2, P)


def ec_add(p1, p2, a):
    """affine addition on y^2 = x^3 + a x + b; None is the identity"""
    if p1 is None:
        return p2
    if p2 is None:
        return p1
    (x1, y1), (x2, y2) = p1, p2
    if x1 == x2:
        if (y1 + y2) % P == 0:
            return None
        lam = (3 * x1 * x1 + a) * inv(2 * y1) % P
    else:
        lam = (y2 - y1) * inv(x2 - x1) % P
    x3 = (lam * lam - x1 - x2) % P
    return (x3, (lam * (x1 - x3) - y1) % P)


def ec_mul(pt, k, a):
    acc = None
    for bit in bin(k)[2:]:
        acc = ec_add(acc, acc, a)
        if bit == "1":
            acc = ec_add(acc, pt, a)
    return acc


def expand_message_xmd(msg, dst, n):
    """RFC 9380 sec. 5.3.1 with SHA-256 (b = 32, block 64 bytes)"""
    assert len(dst) <= 255 and n <= 255 * 32
    ell = (n + 31) // 32
    dst_prime = dst + bytes([len(dst)])
    b0 = hashlib.sha256(bytes(64) + msg + n.to_bytes(2, "big") + b"\x00" + dst_prime).digest()
    blocks = [hashlib.sha256(b0 + b"\x01" + dst_prime).digest()]
    for i in range(2, ell + 1):
        prev = bytes(x ^ y for x, y in zip(b0, blocks[-1]))
        blocks.append(hashlib.sha256(prev + bytes([i]) + dst_prime).digest())
    return b"".join(blocks)[:n]


def hash_to_field(msg, dst, count=2):
    """sec. 5.2: count elements of Fp from 64 bytes each (L = 64)"""
    ub = expand_message_xmd(msg, dst, 64 * count)
    return [int.from_bytes(ub[64 * i:64 * i + 64], "big") % P for i in range(count)]


def is_square(a):
    return a % P == 0 or pow(a, (P - 1) // 2, P) == 1


def sqrt(a):  # p = 3 mod 4
    return pow(a, (P + 1) // 4, P)


def map_to_curve_sswu(u):
    """sec. 6.6.2, simplified SWU for A B != 0; the point lies on E'"""
    tv1 = (Z * Z * pow(u, 4, P) + Z * u * u) % P
    if tv1 == 0:
        x1 = B * inv(Z * A) % P
    else:
        x1 = (-B) * inv(A) * (1 + inv(tv1)) % P
    gx1 = (x1 ** 3 + A * x1 + B) % P
    if is_square(gx1):
        x, y = x1, sqrt(gx1)
    else:
        x = Z * u * u * x1 % P
        y = sqrt((x ** 3 + A * x + B) % P)
    if u % 2 != y % 2:  # sgn0: parity of the canonical value
        y = (-y) % P
    return (x, y)


def _kernel():
    """the ten non-zero points of the unique order-11 subgroup of E'(Fp)"""
    n = P + 1 - (1 - BLS_X)  # #E'(Fp) = #E(Fp), trace x + 1
    x0 = 1
    while True:
        rhs = (x0 ** 3 + A * x0 + B) % P
        y0 = sqrt(rhs)
        if y0 * y0 % P == rhs:
            q = ec_mul((x0, y0), n // 11, A)
            if q is not None:
                break
        x0 += 1
    pts = [q]
    for _ in range(9):
        pts.append(ec_add(pts[-1], q, A))
    assert ec_add(pts[-1], q, A) is None
    return pts


_KERNEL = None


def iso_map(pt):
    """the 11-isogeny E' -> E: y^2 = x^3 + 4 (Velu's point formula, then the
    isomorphism (x, y) -> (x / 11^2, y / 11^3) from y^2 = x^3 + 4 * 11^6)"""
    global _KERNEL
    if _KERNEL is None:
        _KERNEL = _kernel()
    if pt is None:
        return None
    if any(pt == q for q in _KERNEL):
        return None
    x, y = pt
    for q in _KERNEL:
        s = ec_add(pt, q, A)
        x += s[0] - q[0]
        y += s[1] - q[1]
    return (x * inv(121) % P, y * inv(1331) % P)


def clear_cofactor(pt):
    return ec_mul(pt, H_EFF, 0)


def hash_to_g1(msg, dst):
    """sec. 3: hash_to_curve; the affine point (x, y) as integers, None = identity"""
    u0, u1 = hash_to_field(msg, dst, 2)
    q0, q1 = iso_map(map_to_curve_sswu(u0)), iso_map(map_to_curve_sswu(u1))
    return clear_cofactor(ec_add(q0, q1, 0))
