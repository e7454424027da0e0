#!/usr/bin/env python3
"""Generate field_constants.hpp: BLS12-381 constants in the device's internal
representation (14 signed 28-bit limbs, Montgomery radix R' = 2^392).

Every constant is derived from first principles (p, the curve equation, the
tower's non-residue 1+u, the BLS parameter x) with Python integers.  Where the
reference ships the same constant (in its R = 2^384 Montgomery form) the value
is cross-checked against that copy:
  Frobenius constants  gpu-crypto/cuda/field_sppark.cuh:117-119,162-163
  R mod p              gpu-crypto/cuda/api.cu:85-86
  R^2 mod p            ic-integration/gpu_crypto/cuda/decompress.cu:80-85
The expected reference values are read from tests/golden/fixtures.json.
"""
import json, os, sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

P = 0x1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f6241eabfffeb153ffffb9feffffffffaaab
R_ORDER = 0x73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001
BLS_X = 0xd201000000010000  # x = -BLS_X
NL, W = 14, 28
RI = 1 << (NL * W)       # internal Montgomery radix 2^392
RR = 1 << 384            # reference Montgomery radix


def limbs28(v):
    assert 0 <= v < (1 << (NL * W))
    return [(v >> (W * i)) & ((1 << W) - 1) for i in range(NL)]


def to_int(v):  # canonical value -> internal Montgomery form
    return v * RI % P


# Fp2 arithmetic on pairs (c0, c1) with u^2 = -1
def f2mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def f2pow(a, e):
    r = (1, 0)
    while e:
        if e & 1:
            r = f2mul(r, a)
        a = f2mul(a, a)
        e >>= 1
    return r


# ---- hash to G1 (RFC 9380, suite BLS12381G1_XMD:SHA-256_SSWU_RO_) ----
# The simplified SWU map runs on E': y^2 = x^3 + A'x + B' (RFC 9380 sec. 8.8.1), whose
# image reaches E: y^2 = x^3 + 4 through an isogeny of degree 11.  The isogeny's
# rational maps are derived here by Velu's formulas from the 11-torsion of E', not
# copied: E'(Fp) has the order of E(Fp) (checked) and a unique subgroup of order 11;
# the normalised isogeny with that kernel lands on y^2 = x^3 + 4 * 11^6, and
# (x, y) -> (x / 11^2, y / 11^3) maps that onto E.  Two coefficients the RFC publishes
# (appendix E.2: k_(1,0) and k_(3,0)) are checked against the derived maps.
SSWU_A = 0x144698a3b8e9433d693a02c96d4982b0ea985383ee66a8d8e8981aefd881ac98936f8da0e0f97f5cf428082d584c1d
SSWU_B = 0x12e2908d11688030018b12e8753eee3b2016c1f0f24f4070a0b9c14fcef35ef55a23215a316ceaa5d1cc48e98e172be0
SSWU_Z = 11
RFC_K10 = 0x11a05f2b1e833340b809101dd99815856b303e88a2d7005ff2627b56cdb4e2c85610c2d5f2e62d6eaeac1662734649b7
RFC_K30 = 0x90d97c81ba24ee0259d1f094980dcfa11ad138e48a869522b52af6c956543d3cd0c7aee9b3ba3c2be9845719707bb33


def _ec_add(p1, p2, a):  # affine points on y^2 = x^3 + a x + b; None = identity
    if p1 is None:
        return p2
    if p2 is None:
        return p1
    (x1, y1), (x2, y2) = p1, p2
    if x1 == x2:
        if (y1 + y2) % P == 0:
            return None
        lam = (3 * x1 * x1 + a) * pow(2 * y1, -1, P) % P
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, P) % P
    x3 = (lam * lam - x1 - x2) % P
    return (x3, (lam * (x1 - x3) - y1) % P)


def _ec_mul(pt, k, a):
    acc = None
    for bit in bin(k)[2:]:
        acc = _ec_add(acc, acc, a)
        if bit == "1":
            acc = _ec_add(acc, pt, a)
    return acc


def _poly_mul(f, g):  # coefficient lists, lowest degree first, mod P
    o = [0] * (len(f) + len(g) - 1)
    for i, x in enumerate(f):
        for j, y in enumerate(g):
            o[i + j] = (o[i + j] + x * y) % P
    return o


def _poly_add(f, g, sign=1):
    n = max(len(f), len(g))
    return [((f[i] if i < len(f) else 0) + sign * (g[i] if i < len(g) else 0)) % P for i in range(n)]


def _poly_der(f):
    return [i * f[i] % P for i in range(1, len(f))]


def _poly_mod(f, g):  # g monic
    f = f[:]
    while len(f) >= len(g):
        c, d = f[-1], len(f) - len(g)
        for i, y in enumerate(g):
            f[i + d] = (f[i + d] - c * y) % P
        f.pop()
    return f


def iso11_maps():
    """(x_num, x_den, y_num, y_den) of the 11-isogeny E' -> E, lowest degree first;
    x_den and y_den are monic of degrees 10 and 15."""
    a, b = SSWU_A, SSWU_B
    n = P + 1 - (1 - BLS_X)  # #E(Fp), trace t = x + 1 = 1 - BLS_X
    x0 = 1
    while True:  # a point of E' (deterministic), then its multiple of order 11
        rhs = (x0 ** 3 + a * x0 + b) % P
        y0 = pow(rhs, (P + 1) // 4, P)
        if y0 * y0 % P == rhs:
            pt = (x0, y0)
            assert _ec_mul(pt, n, a) is None  # #E'(Fp) = #E(Fp)
            q = _ec_mul(pt, n // 11, a)
            if q is not None:
                break
        x0 += 1
    assert _ec_mul(q, 11, a) is None
    kern = [_ec_mul(q, i, a) for i in range(1, 6)]  # one point of each +- pair
    dpoly = [1]
    for xq, _ in kern:
        dpoly = _poly_mul(dpoly, [(-xq) % P, 1])
    # Velu: with v_Q = 2(3 x_Q^2 + a), u_Q = 4 y_Q^2 over the five pairs,
    # X = x + sum v_Q / (x - x_Q) + u_Q / (x - x_Q)^2, codomain A = a - 5 sum v_Q,
    # B = b - 7 sum (u_Q + x_Q v_Q).  sum f(x_Q) / (x - x_Q) = ((f D') mod D) / D.
    v = sum(2 * (3 * xq * xq + a) for xq, _ in kern) % P
    w = sum(4 * (xq ** 3 + a * xq + b) + xq * 2 * (3 * xq * xq + a) for xq, _ in kern) % P
    assert (a - 5 * v) % P == 0 and (b - 7 * w) % P == 4 * 11 ** 6
    dd = _poly_der(dpoly)
    s1 = _poly_mod(_poly_mul([2 * a % P, 0, 6], dd), dpoly)
    s2 = _poly_mod(_poly_mul([4 * b % P, 4 * a % P, 0, 4], dd), dpoly)
    d2 = _poly_mul(dpoly, dpoly)
    # X = N / D^2 with N = x D^2 + s1 D - (s2' D - s2 D')
    num = _poly_add(_poly_add(_poly_mul([0, 1], d2), _poly_mul(s1, dpoly)),
                    _poly_add(_poly_mul(_poly_der(s2), dpoly), _poly_mul(s2, dd), -1), -1)
    # normalised: Y = y X'(x) = y (N' D - 2 N D') / D^3
    ynum = _poly_add(_poly_mul(_poly_der(num), dpoly), _poly_mul(num, dd), -2)
    i2, i3 = pow(121, -1, P), pow(1331, -1, P)
    xn, yn = [c * i2 % P for c in num], [c * i3 % P for c in ynum]
    xd, yd = d2, _poly_mul(d2, dpoly)
    assert len(xn) == 12 and len(xd) == 11 and len(yn) == 16 and len(yd) == 16 and xd[-1] == yd[-1] == 1
    assert xn[0] == RFC_K10 and yn[0] == RFC_K30
    return xn, xd, yn, yd


def window4(e):
    """e as a left-to-right sliding window of 4 bits: the first entry's odd digit
    starts the result; every later entry squares `sq` times, then multiplies by
    a^digit (digit odd, < 16); a final entry with digit 0 only squares"""
    bits = bin(e)[2:]
    win, i, pend = [], 0, 0
    while i < len(bits):
        if bits[i] == "0":
            pend += 1
            i += 1
            continue
        j = min(i + 4, len(bits))
        while bits[j - 1] == "0":
            j -= 1
        win.append((pend + (j - i), int(bits[i:j], 2)))
        pend, i = 0, j
    if pend:
        win.append((pend, 0))
    chk = 0
    for k, (sq, d) in enumerate(win):
        chk = d if k == 0 else (chk << sq) + d
    assert chk == e and all(d % 2 == 1 for _, d in win if d) and win[0][1] > 0
    return win


def main(out_path=None):
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "fixtures.json")))

    def ref_fp(hexlist):  # reference R=2^384 Montgomery limbs -> canonical value
        v = sum(int(h, 16) << (64 * i) for i, h in enumerate(hexlist))
        return v * pow(RR, -1, P) % P

    # -p^{-1} mod 2^28
    npr = (-pow(P, -1, 1 << W)) % (1 << W)
    beta = (1, 1)
    g1 = f2pow(beta, (P - 1) // 3)      # Fp6 Frobenius on c1
    g2 = f2pow(beta, 2 * (P - 1) // 3)  # Fp6 Frobenius on c2
    g12 = f2pow(beta, (P - 1) // 6)     # Fp12 Frobenius on w
    assert g1 == (0, ref_fp(fx["frob_fp6_c1"][1])) and ref_fp(fx["frob_fp6_c1"][0]) == 0
    assert g2 == (ref_fp(fx["frob_fp6_c2"][0]), 0)
    assert g12 == (ref_fp(fx["frob_fp12_c1"][0]), ref_fp(fx["frob_fp12_c1"][1]))
    assert ref_fp(fx["mont_one"]) == 1
    assert sum(int(h, 16) << (64 * i) for i, h in enumerate(fx["r2_mod_p"])) == RR * RR % P

    # cube root of unity beta with phi(x, y) = (beta x, y) acting as [-x^2] on G1,
    # selected by checking the generator (used by the G1 subgroup test).
    gx = int(fx["g1_gen_canonical"]["x"], 16)
    gy = int(fx["g1_gen_canonical"]["y"], 16)

    def g1add(p1, p2):
        if p1 is None:
            return p2
        if p2 is None:
            return p1
        (x1, y1), (x2, y2) = p1, p2
        if x1 == x2:
            if (y1 + y2) % P == 0:
                return None
            lam = 3 * x1 * x1 * pow(2 * y1, -1, P) % P
        else:
            lam = (y2 - y1) * pow(x2 - x1, -1, P) % P
        x3 = (lam * lam - x1 - x2) % P
        return (x3, (lam * (x1 - x3) - y1) % P)

    def g1mul(k, q):
        r = None
        while k:
            if k & 1:
                r = g1add(r, q)
            q = g1add(q, q)
            k >>= 1
        return r

    x2g = g1mul(BLS_X * BLS_X, (gx, gy))
    gen = 2
    while pow(gen, (P - 1) // 3, P) == 1:
        gen += 1
    betas = [b for b in (pow(gen, (P - 1) // 3, P), pow(gen, 2 * (P - 1) // 3, P))
             if x2g == (b * gx % P, (-gy) % P)]
    assert len(betas) == 1
    beta = betas[0]

    out = []
    emit = out.append
    emit("// Generated by gen_constants.py -- do not edit.")
    emit("// BLS12-381 constants, 14 x 28-bit limbs (little-endian), internal")
    emit("// Montgomery radix R' = 2^392.  'mont' values are canonical v*R' mod p.")
    emit("#pragma once")
    emit("#include <stdint.h>")
    emit("namespace bls {")

    def arr(name, v, comment=""):
        emit("// %s" % comment if comment else "//")
        emit("constexpr int32_t %s[14] = {%s};" % (name, ", ".join("0x%07x" % x for x in limbs28(v))))

    arr("P_L", P, "p")
    emit("constexpr uint32_t NPRIME = 0x%07x;  // -p^-1 mod 2^28" % npr)
    arr("ONE_M", to_int(1), "1 in internal Montgomery form")
    arr("B_G1_M", to_int(4), "curve constant b = 4 (G1)")
    arr("C_TO_INT", pow(2, 400, P), "2^400 mod p: mont(x, C_TO_INT) maps x*2^384 -> x*2^392")
    arr("C_TO_REF", pow(2, 384, P), "2^384 mod p: mont(x, C_TO_REF) maps x*2^392 -> x*2^384")
    arr("C_R2_INT", RI * RI % P, "R'^2 mod p: mont(v, C_R2_INT) maps plain v -> v*R'")
    arr("C_R3_INT", RI ** 3 % P, "R'^3 mod p: mont(x^-1, C_R3_INT) maps the inverse of x = aR' to a^-1 R'")
    arr("FROB6_C1_1", to_int(g1[1]), "(1+u)^((p-1)/3) = (0, this)")
    arr("FROB6_C2_0", to_int(g2[0]), "(1+u)^(2(p-1)/3) = (this, 0)")
    arr("FROB12_0", to_int(g12[0]), "(1+u)^((p-1)/6) = (this, .)")
    arr("FROB12_1", to_int(g12[1]), "(1+u)^((p-1)/6) = (., this)")
    arr("BETA_M", to_int(beta), "cube root of unity: (beta x, y) = [-x^2](x, y) on G1")
    arr("HALF_M", to_int((P + 1) // 2), "1/2 (the square root in Fp2 by the norm)")
    arr("G1_GEN_X_M", to_int(gx), "G1 generator x (the stand-in point of lanes with no item to test)")
    arr("G1_GEN_Y_M", to_int(gy), "G1 generator y")
    arr("B_G2_M", to_int(4), "G2 curve constant b' = 4(1+u) = (this, this)")
    # psi(x, y) = (x^p cx, y^p cy) on the twist, cx = (1+u)^-((p-1)/3), cy = (1+u)^-((p-1)/2)
    # (the endomorphism of the G2 membership test psi(Q) == [x]Q, eprint 2021/1130 sec. 4)
    cx = f2pow(f2pow((1, 1), (P - 1) // 3), P * P - 2)
    cy = f2pow(f2pow((1, 1), (P - 1) // 2), P * P - 2)
    assert f2mul(cx, f2pow((1, 1), (P - 1) // 3)) == (1, 0)
    arr("PSI_X0", to_int(cx[0]), "psi: cx.c0")
    arr("PSI_X1", to_int(cx[1]), "psi: cx.c1")
    arr("PSI_Y0", to_int(cy[0]), "psi: cy.c0")
    arr("PSI_Y1", to_int(cy[1]), "psi: cy.c1")
    # exponents as 32-bit words, little-endian
    def words(v, n):
        return ", ".join("0x%08xu" % ((v >> (32 * i)) & 0xffffffff) for i in range(n))
    emit("// (p+1)/4 for the square root, p-2 for the Fermat inverse, (p-1)/2 for Euler's criterion")
    emit("constexpr uint32_t SQRT_EXP[12] = {%s};" % words((P + 1) // 4, 12))
    emit("constexpr uint32_t INV_EXP[12] = {%s};" % words(P - 2, 12))
    emit("constexpr uint32_t HALF_P32[12] = {%s};" % words((P - 1) // 2, 12))
    emit("// (p-3)/4 for the square root in Fp2")
    emit("constexpr uint32_t P34_EXP[12] = {%s};" % words((P - 3) // 4, 12))
    emit("constexpr uint32_t P32[12] = {%s};" % words(P, 12))
    emit("constexpr uint32_t R_ORDER32[8] = {%s};" % words(R_ORDER, 8))
    emit("constexpr uint64_t BLS_X = 0x%016xull;  // x = -BLS_X" % BLS_X)
    # exponents as left-to-right sliding windows of 4 bits
    for name, e, what in (("SQRT_WIN", (P + 1) // 4, "(p+1)/4"), ("P34_WIN", (P - 3) // 4, "(p-3)/4")):
        win = window4(e)
        emit("// %s as a 4-bit sliding window: {squarings, odd digit} (see gen_constants.py)" % what)
        emit("constexpr int %s_LEN = %d;" % (name, len(win)))
        emit("constexpr uint8_t %s[%d][2] = {%s};" % (name, len(win), ", ".join("{%d, %d}" % w for w in win)))
    # hash to G1: SSWU on E' and the 11-isogeny to E
    arr("SSWU_A_M", to_int(SSWU_A), "A' of E' (RFC 9380 sec. 8.8.1)")
    arr("SSWU_B_M", to_int(SSWU_B), "B' of E'")
    arr("SSWU_Z_M", to_int(SSWU_Z), "Z = 11")
    c2 = pow((-SSWU_Z) % P, (P + 1) // 4, P)
    assert c2 * c2 % P == (-SSWU_Z) % P
    arr("SSWU_SQRT_MZ_M", to_int(c2), "a square root of -Z")
    arr("C_H2F_HI", pow(2, 384, P) * RI * RI % P, "2^384 R'^2 mod p: mont(hi, this) = hi 2^384 R' for the top of a 512-bit value")
    maps = iso11_maps()
    for name, poly in zip(("ISO_XNUM", "ISO_XDEN", "ISO_YNUM", "ISO_YDEN"), maps):
        coeffs = poly if name.endswith("NUM") else poly[:-1]  # denominators are monic
        emit("// %s of the 11-isogeny E' -> E, lowest degree first%s" %
             (name[4:].lower(), "" if name.endswith("NUM") else " (monic: the leading 1 is implied)"))
        emit("constexpr int32_t %s[%d][14] = {%s};" % (name, len(coeffs), ", ".join(
            "{%s}" % ", ".join("0x%07x" % x for x in limbs28(to_int(c))) for c in coeffs)))
    emit("}  // namespace bls")
    with open(out_path or os.path.join(HERE, "field_constants.hpp"), "w") as f:
        f.write("\n".join(out) + "\n")
    print("wrote", out_path or "field_constants.hpp")


if __name__ == "__main__":
    sys.exit(main())
