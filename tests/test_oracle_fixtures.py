"""The CPU oracle against every known-answer value the reference's own tests hold
(tests/golden/fixtures.json, extracted by tests/golden/extract_fixtures.py)."""
import numpy as np

import oracle as O


def test_g2_generator_lines_match_reference():
    # cuda/g2_coeffs.h:4-549 (G2Prepared of the generator, from ic_bls12_381)
    fx = O.fixtures()
    assert (O.g2_prepare(O.g2_gen()) == O.u64(fx["g2_gen_coeffs"])).all()


def test_42_g2_lines_match_reference():
    # cuda/pk42_coeffs.h:2-411
    fx = O.fixtures()
    q42, inf = O.g2_mul(O.g2_gen(), 42)
    assert not inf
    assert (O.g2_prepare(q42) == O.u64(fx["pk42_coeffs"])).all()


def test_transcribed_g2_generator_is_off_curve():
    # cuda/test_miller_v2.cu:291-302 carries a wrong y.c0; the standard point is used
    fx = O.fixtures()
    bad = O.u64(fx["g2_gen_as_transcribed"])
    good = O.g2_gen()
    assert (bad[:12] == good[:12]).all() and (bad[18:] == good[18:]).all()
    assert not (bad[12:18] == good[12:18]).all()
    c = fx["g2_gen_canonical"]
    for k, sl in (("x_c0", slice(0, 6)), ("x_c1", slice(6, 12)), ("y_c0", slice(12, 18)), ("y_c1", slice(18, 24))):
        assert O.unmont(good[sl]) == int(c[k], 16)


def test_g1_generator_and_small_multiples():
    fx = O.fixtures()
    g = O.g1_gen()
    assert (g == O.u64(fx["g1_gen"])).all()
    assert O.unmont(g[:6]) == int(fx["g1_gen_canonical"]["x"], 16)
    assert O.unmont(g[6:]) == int(fx["g1_gen_canonical"]["y"], 16)
    s294, _ = O.g1_mul(g, 294)
    assert (s294 == O.u64(fx["sig_294_g1"])).all()  # cuda/test_bls_verify.cu:115-121
    h7, _ = O.g1_mul(g, 7)
    assert (O.g1_neg(h7) == O.u64(fx["neg_7_g1"])).all()  # :123-129




def test_verify_vector_miller_and_verdict():
    # sk = 42, H(m) = 7 G1, sig = 294 G1 (cuda/test_bls_verify.cu:112-172)
    fx = O.fixtures()
    g2c, pkc = O.u64(fx["g2_gen_coeffs"]), O.u64(fx["pk42_coeffs"])
    sig, nhm = O.u64(fx["sig_294_g1"]), O.u64(fx["neg_7_g1"])
    ml = np.zeros(72, np.uint64)
    O.lib().orc_miller2(O.ptr(sig), O.ptr(g2c), O.ptr(nhm), O.ptr(pkc), O.ptr(ml))
    assert ["%016x" % v for v in ml[:6]] == fx["verify_miller_c000"]
    ok, gt = O.verify(sig, nhm, g2c, pkc)
    assert ok
    one = np.zeros(72, np.uint64)
    one[:6] = O.u64(fx["mont_one"])
    assert (gt == one).all()
    # the reference's invalid vector: the G1 generator as signature (gpu-crypto/src/lib.rs:226-242)
    ok2, gt2 = O.verify(O.g1_gen(), nhm, g2c, pkc)
    assert not ok2 and not (gt2 == one).all()


def test_final_exponentiation_of_reference_miller_value():
    # cuda/g2_coeffs.h:551-565 -> lib/test_cyc_sq.cu:141-153
    fx = O.fixtures()
    fe = np.zeros(72, np.uint64)
    O.lib().orc_final_exp(O.ptr(O.u64(fx["miller_g1_g2"])), O.ptr(fe))
    assert (fe == O.u64(fx["pairing_g1_g2"])).all()


def test_cyclotomic_square_of_pairing():
    # lib/test_cyc_sq.cu:175-176
    fx = O.fixtures()
    cs = np.zeros(72, np.uint64)
    O.lib().orc_fp12_cyc_sqr(O.ptr(O.u64(fx["pairing_g1_g2"])), O.ptr(cs))
    assert ["%016x" % v for v in cs[:6]] == fx["cyclotomic_square_pairing_c000"]
    sq = np.zeros(72, np.uint64)
    O.lib().orc_fp12_sqr(O.ptr(O.u64(fx["pairing_g1_g2"])), O.ptr(sq))
    assert (sq == cs).all()


def test_constants():
    fx = O.fixtures()
    assert O.to_int(O.u64(fx["p"])) == O.P
    assert O.unmont(O.u64(fx["mont_one"])) == 1
    assert O.to_int(O.u64(fx["r2_mod_p"])) == O.R384 ** 2 % O.P
    assert O.to_int(O.u64(fx["sqrt_exponent"])) == (O.P + 1) // 4
    assert O.to_int(O.u64(fx["fermat_exponent"])) == O.P - 2
    assert O.to_int(O.u64(fx["half_p"])) == (O.P - 1) // 2
    assert (-pow(O.P, -1, 1 << 64)) % (1 << 64) == int(fx["m0_64"], 16)
