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


def main():
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
    # (p+1)/4 as a left-to-right sliding window of 4 bits: the first entry's odd digit
    # starts the result; every later entry squares `sq` times, then multiplies by
    # a^digit (digit odd, < 16); a final entry with digit 0 only squares
    e = (P + 1) // 4
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
    emit("// (p+1)/4 as a 4-bit sliding window: {squarings, odd digit} (see gen_constants.py)")
    emit("constexpr int SQRT_WIN_LEN = %d;" % len(win))
    emit("constexpr uint8_t SQRT_WIN[%d][2] = {%s};" % (len(win), ", ".join("{%d, %d}" % w for w in win)))
    emit("}  // namespace bls")
    with open(os.path.join(HERE, "field_constants.hpp"), "w") as f:
        f.write("\n".join(out) + "\n")
    print("wrote field_constants.hpp")


if __name__ == "__main__":
    sys.exit(main())
