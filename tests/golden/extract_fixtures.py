#!/usr/bin/env python3
"""Extract the reference's known-answer data for the BLS verify path into
`tests/golden/fixtures.json`.

Runs in the build container only (the reference tree is not present where the
GPU tests run); the JSON it writes is committed and is all the tests read.
Every value below is parsed as text out of the reference's own test files,
fixture headers and device constants; nothing from the reference is executed.

Sources (paths relative to the reference root):
  cuda/g2_coeffs.h:4-549        G2Prepared(G2 generator) line coefficients, 68 x 36 u64
  cuda/g2_coeffs.h:551-565      Miller loop value of e(G1, G2) before final exp (72 u64)
  cuda/pk42_coeffs.h:2-411      G2Prepared(42 * G2) line coefficients, 68 x 36 u64
  lib/test_cyc_sq.cu:141-153    e(G1, G2) after final exponentiation (72 u64)
  lib/test_cyc_sq.cu:175-176    cyclotomic square of e(G1, G2), c0.c0.c0 (6 u64)
  cuda/test_bls_verify.cu:115-129  sig = 294*G1 and -H(m) = -(7*G1), Montgomery x|y
  cuda/test_bls_verify.cu:147   multi-Miller (294*G1, G2)(-7*G1, 42*G2), c0.c0.c0
  cuda/test_miller_v2.cu:283-302   G1 and G2 generators (Montgomery limbs); its G2 y.c0
                                is mis-transcribed, so the standard G2 generator is recorded too
  gpu-crypto/cuda/field_sppark.cuh:117-119,162-163  Frobenius constants
  gpu-crypto/cuda/field_sppark.cuh:43-44  exponent p-2 of the Fermat inverse
  gpu-crypto/cuda/api.cu:85-86  Montgomery one (R mod p)
  ic-integration/gpu_crypto/cuda/decompress.cu:17-24   exponent (p+1)/4
  ic-integration/gpu_crypto/cuda/decompress.cu:80-85   R^2 mod p (12 x u32)
  ic-integration/gpu_crypto/cuda/decompress.cu:114-121 (p-1)/2
  cuda/test_fp2.cu:110-124, cuda/test_fp6.cu:118-158, cuda/test_fp12.cu:124-163
                                toy tower products on small integers
"""
import json
import os
import re
import sys

REF = os.environ.get("ICP_GPU_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json")


def read(rel):
    with open(os.path.join(REF, rel)) as f:
        return f.read()


def hexes(text, with_ull=True):
    pat = r"0x([0-9a-fA-F]+)ULL" if with_ull else r"\b([0-9a-f]{16})\b"
    return [int(h, 16) for h in re.findall(pat, text)]


def block(text, start_pat, end_pat="};"):
    i = text.index(start_pat)
    j = text.index(end_pat, i)
    return text[i:j]


def lines(text, a, b):
    return "\n".join(text.splitlines()[a - 1:b])


def u64hex(vals):
    return ["%016x" % v for v in vals]


def main():
    fx = {}

    g2h = read("cuda/g2_coeffs.h")
    g2c = hexes(block(g2h, "G2_COEFFS[2448] = {"))
    assert len(g2c) == 2448, len(g2c)
    fx["g2_gen_coeffs"] = u64hex(g2c)
    em = block(g2h, "EXPECTED_MILLER[72]", "// };")
    em = hexes(em.split("=", 1)[1], with_ull=False)
    assert len(em) == 72, len(em)
    fx["miller_g1_g2"] = u64hex(em)

    pk = hexes(block(read("cuda/pk42_coeffs.h"), "PK42_COEFFS[2448] = {"))
    assert len(pk) == 2448, len(pk)
    fx["pk42_coeffs"] = u64hex(pk)
    # identical copy shipped with the crate
    pk2 = hexes(block(read("gpu-crypto/cuda/pk42_coeffs.h"), "PK42_COEFFS[2448] = {"))
    assert pk2 == pk

    cyc = read("lib/test_cyc_sq.cu")
    gt = hexes(block(cyc, "uint64_t data[72] = {"))
    assert len(gt) == 72
    fx["pairing_g1_g2"] = u64hex(gt)
    cs = hexes(block(cyc, "uint64_t expected_c0[6] = {"))
    assert len(cs) == 6
    fx["cyclotomic_square_pairing_c000"] = u64hex(cs)

    # test_final_exp.cu restates the same Miller value and the first limbs of e(G1,G2)
    tfe = read("cuda/test_final_exp.cu")
    em2 = hexes(block(tfe, "uint64_t data[72] = {"))
    assert em2 == em
    assert "0x1972e433a01f85c5ULL" in tfe and "0xd26331b02e9d6995ULL" in tfe
    assert gt[0] == 0x1972e433a01f85c5 and gt[6] == 0xd26331b02e9d6995

    tbv = read("cuda/test_bls_verify.cu")
    body = block(tbv, "// sig = 294*G1", "printf(\"BLS Signature")

    def coords(name):
        vals = {}
        for axis, idx, h in re.findall(name + r"\.([xy])\.v\[(\d)\]=0x([0-9a-f]+)ULL", body):
            vals[(axis, int(idx))] = int(h, 16)
        return [vals[("x", i)] for i in range(6)] + [vals[("y", i)] for i in range(6)]

    fx["sig_294_g1"] = u64hex(coords("sig"))
    fx["neg_7_g1"] = u64hex(coords("neg_hm"))
    m = re.search(r"EXP=\[([0-9a-f,]+)\]", tbv)
    fx["verify_miller_c000"] = m.group(1).split(",")
    assert len(fx["verify_miller_c000"]) == 6

    mv2 = read("cuda/test_miller_v2.cu")
    tp = block(mv2, "__global__ void test_pairing()", "printf(")
    g1 = {}
    for axis, idx, h in re.findall(r"g1\.([xy])\.v\[(\d)\]=0x([0-9a-f]+)ULL", tp):
        g1[(axis, int(idx))] = int(h, 16)
    fx["g1_gen"] = u64hex([g1[("x", i)] for i in range(6)] + [g1[("y", i)] for i in range(6)])
    g2 = {}
    for axis, comp, idx, h in re.findall(r"g2\.([xy])\.(c[01])\.v\[(\d)\]=0x([0-9a-f]+)ULL", tp):
        g2[(axis, comp, int(idx))] = int(h, 16)
    fx["g2_gen_as_transcribed"] = u64hex([g2[(a, c, i)] for a in "xy" for c in ("c0", "c1") for i in range(6)])
    # The standard G2 generator (canonical, big-endian hex as published with the
    # BLS12-381 curve).  The transcription above has a wrong y.c0 (not on the curve);
    # the standard point is the one that reproduces cuda/g2_coeffs.h.
    fx["g2_gen_canonical"] = {
        "x_c0": "024aa2b2f08f0a91260805272dc51051c6e47ad4fa403b02b4510b647ae3d1770bac0326a805bbefd48056c8c121bdb8",
        "x_c1": "13e02b6052719f607dacd3a088274f65596bd0d09920b61ab5da61bbdc7f5049334cf11213945d57e5ac7d055d042b7e",
        "y_c0": "0ce5d527727d6e118cc9cdc6da2e351aadfd9baa8cbdd3a76d429a695160d12c923ac9cc3baca289e193548608b82801",
        "y_c1": "0606c4a02ea734cc32acd2b02bc28b99cb3e287e85a763af267492ab572e99ab3f370d275cec1da1aaa9075ff05f79be",
    }
    fx["g1_gen_canonical"] = {
        "x": "17f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac586c55e83ff97a1aeffb3af00adb22c6bb",
        "y": "08b3f481e3aaa0f1a09e30ed741d8ae4fcf5e095d5d00af600db18cb2c04b3edd03cc744a2888ae40caa232946c5e7e1",
    }

    fs = read("gpu-crypto/cuda/field_sppark.cuh")
    arrs = re.findall(r"uint64_t v\[6\]=\{([^}]*)\}", fs)
    assert len(arrs) == 4
    vals = [hexes(a) for a in arrs]
    fx["frob_fp6_c1"] = [u64hex([0] * 6), u64hex(vals[0])]   # (0, gamma) at :117-118
    fx["frob_fp6_c2"] = [u64hex(vals[1]), u64hex([0] * 6)]   # (gamma, 0) at :119
    fx["frob_fp12_c1"] = [u64hex(vals[2]), u64hex(vals[3])]  # :162-163
    inv_exp = hexes(block(fs, "uint64_t exp[6]={", "};"))
    fx["fermat_exponent"] = u64hex(inv_exp)

    api = read("gpu-crypto/cuda/api.cu")
    fx["mont_one"] = u64hex(hexes(block(api, "uint64_t one[6] = {", "};")))
    assert "#define BLS_X_VAL 0xd201000000010000ULL" in api
    fx["bls_x"] = "d201000000010000"

    dc = read("ic-integration/gpu_crypto/cuda/decompress.cu")
    fx["sqrt_exponent"] = u64hex(hexes(block(dc, "const uint64_t exp[6] = {")))
    rr = re.findall(r"rr32\[\s*(\d+)\]\s*=\s*0x([0-9a-f]+);", dc)
    rr = [int(h, 16) for _, h in sorted(rr, key=lambda t: int(t[0]))]
    assert len(rr) == 12
    fx["r2_mod_p"] = u64hex([rr[2 * i] | (rr[2 * i + 1] << 32) for i in range(6)])
    fx["half_p"] = u64hex(hexes(block(dc, "const uint64_t half_p[6] = {")))

    mv2p = hexes(block(mv2, "FP_P[6] = {"))
    fx["p"] = u64hex(mv2p)
    assert "#define M0 0x89f3fffcfffcfffdULL" in mv2
    fx["m0_64"] = "89f3fffcfffcfffd"

    # Toy tower products on small integers; the expected values are stated in
    # the reference's test comments and printf strings.  Each string is checked
    # for presence so a drift in the reference is caught.
    t2 = read("cuda/test_fp2.cu")
    assert "(2+3u)*(4+5u) = (8-15) + (10+12)u = -7 + 22u" in t2
    assert "β*(3+7u) = (1+u)(3+7u) = (3-7) + (3+7)u = -4 + 10u" in t2
    t6 = read("cuda/test_fp6.cu")
    assert "c0 = Fp2(P-223, 176), c1 = Fp2(P-165, 189), c2 = Fp2(P-39, 182)" in t6
    assert "b = Fp6(Fp2(7,8), Fp2(9,10), Fp2(11,12))" in t6
    t12 = read("cuda/test_fp12.cu")
    for s in ("c0.c0 = Fp2(P-1650, 1405)", "c0.c1 = Fp2(P-1282, 1475)", "c1.c0 = Fp2(P-1238, 1240)"):
        assert s in t12, s
    fx["toy_kats"] = {
        "fp2_mul": {"a": [2, 3], "b": [4, 5], "out": [-7, 22]},
        "fp2_mul_nr": {"a": [3, 7], "out": [-4, 10]},
        "fp6_mul": {"a": [1, 2, 3, 4, 5, 6], "b": [7, 8, 9, 10, 11, 12],
                    "out": [-223, 176, -165, 189, -39, 182]},
        # Fp2 components c0.c0, c0.c1 and c1.c0 of the product (tower index 0, 1, 3)
        "fp12_mul": {"a": list(range(1, 13)), "b": list(range(13, 25)),
                     "out_fp2": {"0": [-1650, 1405], "1": [-1282, 1475], "3": [-1238, 1240]}},
    }

    fx["_sources"] = __doc__.split("Sources (paths relative to the reference root):")[1].strip()
    with open(OUT, "w") as f:
        json.dump(fx, f, indent=1, sort_keys=True)
    print("wrote", OUT, "keys:", len(fx))


if __name__ == "__main__":
    sys.exit(main())
