"""The per-G round tables that drive the lanes-per-item kernels (icp-gpu_amd/csrc/
gen_schedules.py -> schedules_gen.hpp), executed here with Python big integers exactly as
the kernel's round interpreter reads them: records per lane, uniform term counts, zero
padding, regions A/B/T/K, signed coefficients.  Every op at every lane count must give
the values of its symbolic definition, and the Fp12 product and square must equal the
independent schoolbook model (tests/pyfield.py), so a mistake in round assignment, record
layout or term encoding fails here, on the CPU."""
import os
import random
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "icp-gpu_amd", "csrc"))
import gen_schedules as GSCH  # noqa: E402
import pyfield as PF  # noqa: E402

P = GSCH.P
A, B, T, K = GSCH.A, GSCH.B, GSCH.T, GSCH.K


def decode(term):
    coef = (term >> 8) & 0xFF
    if coef >= 128:
        coef -= 256
    return (term >> 6) & 3, term & 63, coef


def run_op(rounds, a, b, t=None):
    """execute one op's rounds (as the kernel does) on input values a, b (lists of ints);
    t: values of T slots that an earlier op filled (the op's reserved slots)"""
    prounds, orounds = rounds
    regions = {A: a, B: b, K: [v % P for v in GSCH.CONSTS.vals], T: dict(t or {})}

    def val(term):
        reg, idx, coef = decode(term)
        if coef == 0:
            return 0
        return coef * regions[reg][idx]

    for na, nb, recs in prounds:
        for rec in recs:
            assert len(rec) == 8 * GSCH.PROD_Q
            if rec[0] == GSCH.IDLE:
                continue
            # the kernel reads exactly na and nb terms at fixed places
            assert all(decode(t)[2] == 0 for t in rec[1 + na:GSCH.B_AT] + rec[GSCH.B_AT + nb:])
            x = sum(val(t) for t in rec[1:1 + na])
            y = sum(val(t) for t in rec[GSCH.B_AT:GSCH.B_AT + nb])
            assert rec[0] not in regions[T]
            regions[T][rec[0]] = x * y % P
    out = {}
    for n, _, recs in orounds:
        for rec in recs:
            if rec[0] == GSCH.IDLE:
                continue
            assert all(decode(t)[2] == 0 for t in rec[1 + n:])
            assert rec[0] not in out
            out[rec[0]] = sum(val(t) for t in rec[1:1 + n]) % P
    return [out.get(i) for i in range(max(out) + 1)]


def reserved(prods, rnd):
    """random values for the T slots an op reserves (filled by an earlier op)"""
    return {i: rnd.randrange(P) for i, p in enumerate(prods) if p is None}


def symbolic(prods, outs, a, b, t=None):
    vals = {A: a, B: b, K: [v % P for v in GSCH.CONSTS.vals]}

    def ev(lin, tv):
        s = 0
        for (reg, idx), c in lin.t.items():
            s += c * (tv[idx] if reg == T else vals[reg][idx])
        return s % P

    tv = [t[i] if p is None else ev(p[0], None) * ev(p[1], None) % P for i, p in enumerate(prods)]
    return [None if o is None else ev(o, tv) for o in outs]


@pytest.fixture(scope="module")
def progs():
    return GSCH.op_programs()


@pytest.fixture(scope="module")
def variants(progs):
    """the programs the generator chooses from: plain Karatsuba and capped operands"""
    return [progs, GSCH.op_programs(cap=4)]


def test_copy_flags_of_the_emitted_tables(variants):
    """a round flagged as a plain copy of operand A (B) really holds one +1 term per busy
    lane: the kernel then loads that slot without accumulating"""
    path = os.path.join(os.path.dirname(HERE), "icp-gpu_amd", "csrc", "schedules_gen.hpp")
    text = open(path).read()
    for G in GSCH.GS:
        ops = GSCH.build_rounds_v(G, variants)
        rows = text.split("__constant__ const Round ROUNDS%d[" % G)[1].split("};")[0].split("= {", 1)[1]
        flags = [int(x.strip("{} ").split(",")[4]) for x in rows.split("}, {")]
        k = 0
        for name, prounds, orounds, _, _ in ops:
            for na, nb, recs in prounds:
                busy = [r for r in recs if r[0] != GSCH.IDLE]
                if flags[k] & 1:
                    assert na == 1 and all(decode(r[1])[2] == 1 for r in busy), (G, name)
                if flags[k] & 2:
                    assert nb == 1 and all(decode(r[GSCH.B_AT])[2] == 1 for r in busy), (G, name)
                k += 1
            for _ in orounds:
                assert flags[k] == 0
                k += 1
        assert k == len(flags)


@pytest.mark.parametrize("G", GSCH.GS)
def test_round_tables_execute_every_op(variants, G):
    """the tables the kernels use: per op the cheaper of the two programs"""
    rnd = random.Random(G)
    ops = GSCH.build_rounds_v(G, variants)
    for (name, prounds, orounds, prods, outs), (name2, _, _, _) in zip(ops, variants[0]):
        assert name == name2
        # every lane record of a round has the round's shape; idle lanes only at the end
        for na, nb, recs in prounds + orounds:
            assert len(recs) == G
            busy = [r[0] != GSCH.IDLE for r in recs]
            assert busy == sorted(busy, reverse=True)
        assert sum(1 for _, _, recs in prounds for r in recs if r[0] != GSCH.IDLE) == \
            sum(p is not None for p in prods)
        assert sum(1 for _, _, recs in orounds for r in recs if r[0] != GSCH.IDLE) == sum(o is not None for o in outs)
        for _ in range(3):
            a = [rnd.randrange(P) for _ in range(12)]
            b = [rnd.randrange(P) for _ in range(28)]
            t = reserved(prods, rnd)
            assert run_op((prounds, orounds), a, b, t) == symbolic(prods, outs, a, b, t), name


def tower(vals):
    return [(vals[2 * k], vals[2 * k + 1]) for k in range(6)]


def flat(t):
    return [x for pair in t for x in pair]


@pytest.mark.parametrize("G", [12, 13, 16, 18, 32])
def test_fp12_mul_and_square_against_schoolbook(variants, G):
    rnd = random.Random(100 + G)
    ops = {name: (p, o) for name, p, o, _, _ in GSCH.build_rounds_v(G, variants)}
    mul = ops["OP_MUL"]
    sqr = ops["OP_SQR"]
    for _ in range(4):
        a = [rnd.randrange(P) for _ in range(12)]
        b = [rnd.randrange(P) for _ in range(12)]
        want = PF.f12_to_tower(PF.f12mul_w(PF.f12_from_tower(tower(a)), PF.f12_from_tower(tower(b))))
        assert run_op(mul, a, b) == flat(want)
        want2 = PF.f12_to_tower(PF.f12mul_w(PF.f12_from_tower(tower(a)), PF.f12_from_tower(tower(a))))
        assert run_op(sqr, a, [0] * 12) == flat(want2)


@pytest.mark.parametrize("G", [12, 16, 32])
def test_split_fp12_mul_against_schoolbook(variants, G):
    """OP_MUL_S then OP_MUL_PQ (s passed in the first six T slots) equal the Fp12 product"""
    rnd = random.Random(400 + G)
    ops = {name: (p, o) for name, p, o, _, _ in GSCH.build_rounds_v(G, variants)}
    for _ in range(4):
        a = [rnd.randrange(P) for _ in range(12)]
        b = [rnd.randrange(P) for _ in range(12)]
        s = run_op(ops["OP_MUL_S"], a, b)
        assert len(s) == 6
        got = run_op(ops["OP_MUL_PQ"], a, b, dict(enumerate(s)))
        want = PF.f12_to_tower(PF.f12mul_w(PF.f12_from_tower(tower(a)), PF.f12_from_tower(tower(b))))
        assert got == flat(want)


@pytest.mark.parametrize("G", [12, 16, 20, 32])
def test_merged_line_step_against_schoolbook(variants, G):
    """LL = line1 * line2, f * LL (with the next step's scaled lines) and f^2 with the next
    LL equal the full Fp12 products of the sparse lines x0 + x1 v + y1 v w"""
    rnd = random.Random(200 + G)
    ops = {name: (p, o) for name, p, o, _, _ in GSCH.build_rounds_v(G, variants)}
    zero = (0, 0)
    for _ in range(3):
        b = [rnd.randrange(P) for _ in range(28)]
        s1 = [(b[16 + 2 * k], b[17 + 2 * k]) for k in range(3)]
        s2 = [(b[22 + 2 * k], b[23 + 2 * k]) for k in range(3)]
        l1 = [s1[0], s1[1], zero, zero, s1[2], zero]  # tower order c0.c0 c0.c1 c0.c2 c1.c0 c1.c1 c1.c2
        l2 = [s2[0], s2[1], zero, zero, s2[2], zero]
        want = PF.f12_to_tower(PF.f12mul_w(PF.f12_from_tower(l1), PF.f12_from_tower(l2)))
        got = run_op(ops["OP_LL"], [0] * 12, b)
        assert want[3] == zero
        assert got == flat([want[0], want[1], want[2], want[4], want[5]])
        # f * LL, LL written over the scaled lines (B 16..25)
        b2 = list(b)
        b2[16:26] = got
        f = [rnd.randrange(P) for _ in range(12)]
        prod = flat(PF.f12_to_tower(PF.f12mul_w(PF.f12_from_tower(tower(f)), PF.f12_from_tower(want))))
        out = run_op(ops["OP_FL_SCALE"], f, b2)
        assert out[:12] == prod and out[12:28] == [None] * 16
        # the scaled lines of the next step: (c2, c1 P.x, c0 P.y) per pair
        for pair in range(2):
            c = [(b2[6 * pair + 2 * k], b2[6 * pair + 2 * k + 1]) for k in range(3)]
            x, y = b2[12 + 2 * pair], b2[13 + 2 * pair]
            sc = [c[2], (c[1][0] * x % P, c[1][1] * x % P), (c[0][0] * y % P, c[0][1] * y % P)]
            assert out[28 + 6 * pair:34 + 6 * pair] == flat(sc)
        assert run_op(ops["OP_FL"], f, b2) == prod
        # f^2 and the line product in one op: f^2 at 0..11, LL over the scaled lines (28..37)
        sq = flat(PF.f12_to_tower(PF.f12mul_w(PF.f12_from_tower(tower(f)), PF.f12_from_tower(tower(f)))))
        out = run_op(ops["OP_SQR_LL"], f, b)
        assert out[:12] == sq and out[12:28] == [None] * 16
        assert out[28:38] == flat([want[0], want[1], want[2], want[4], want[5]])


def test_cyclotomic_square_on_a_cyclotomic_element(variants):
    """the cyclotomic-square schedule equals the full square on an element of the
    cyclotomic subgroup, f^((p^6 - 1)(p^2 + 1)) for a random f (and differs off it)"""
    rnd = random.Random(7)
    f = [(rnd.randrange(P), rnd.randrange(P)) for _ in range(6)]
    fw = PF.f12_from_tower(f)
    # f^(p^6 - 1) = conj(f) / f, then raised to p^2 + 1
    conj = [fw[i] if i % 2 == 0 else ((-fw[i][0]) % P, (-fw[i][1]) % P) for i in range(6)]
    inv = PF.f12pow_w(fw, P ** 12 - 2)
    g = PF.f12mul_w(conj, inv)
    g = PF.f12mul_w(PF.f12pow_w(g, P * P), g)
    gt = flat(PF.f12_to_tower(g))
    want = flat(PF.f12_to_tower(PF.f12mul_w(g, g)))
    for G in (12, 16, 32):
        ops = {name: (p, o) for name, p, o, _, _ in GSCH.build_rounds_v(G, variants)}
        assert run_op(ops["OP_CYC_SQR"], gt, [0] * 28) == want
    assert run_op(ops["OP_CYC_SQR"], flat(f), [0] * 28) != flat(PF.f12_to_tower(PF.f12mul_w(fw, fw)))


def _array(text, decl):
    body = text.split(decl, 1)[1].split("= {", 1)[1].split("};", 1)[0]
    return body


def _rows(body):
    return [[int(x, 0) for x in r.strip("{} ").replace("u", "").split(",")] for r in body.split("}, {")]


@pytest.mark.parametrize("G", GSCH.GS)
def test_static_tables_execute_the_straight_line_ops(variants, G):
    """DESC_S as the straight-line rounds read it: lane-major records per op, term lists
    whose busy lanes read one region encoded relative to it (slot * 14, 6-bit signed
    coefficient) and lists over two regions with a region bit (5-bit coefficient).
    Executed here, every such op gives the values of its symbolic definition."""
    path = os.path.join(os.path.dirname(HERE), "icp-gpu_amd", "csrc", "schedules_gen.hpp")
    text = open(path).read()
    tab = text.split("template <> struct Tab<%d> {" % G, 1)[1].split("\n};", 1)[0]
    ops_c = _rows(_array(tab, "OPS_C["))
    rounds_c = _rows(_array(tab, "ROUNDS_C["))
    srounds = _rows(_array(tab, "SROUNDS_C["))
    words = [int(x, 0) for x in _array(text, "DESC_S%d[" % G).split(",")]
    names = [n for n, _, _, _ in variants[0]]
    progs = {n: (prods, outs) for n, _, _, prods, outs in GSCH.build_rounds_v(G, variants)}
    rnd = random.Random(300 + G)
    for name in GSCH.static_ops(G):
        first, np_, no = ops_c[names.index(name)]
        prods, outs = progs[name]
        for _ in range(2):
            a = [rnd.randrange(P) for _ in range(12)]
            b = [rnd.randrange(P) for _ in range(28)]
            t0 = reserved(prods, rnd)
            regions = {A: a, B: b, T: dict(t0)}

            def term(t, reg):
                if reg == GSCH.MIXED:
                    r, idx, coef = decode(t)
                elif reg >= GSCH.PAIR:
                    off, coef = t & 1023, (t >> 11) & 31
                    coef -= 32 if coef >= 16 else 0
                    r = (reg - GSCH.PAIR) % 4 if t & 1024 else (reg - GSCH.PAIR) // 4
                    assert off % 14 == 0
                    idx = off // 14
                else:
                    off, coef = t & 1023, (t >> 10) & 63
                    coef -= 64 if coef >= 32 else 0
                    assert off % 14 == 0
                    r, idx = reg, off // 14
                return 0 if coef == 0 else coef * regions[r][idx]

            def weight(ts, reg):
                w = 0
                for t in ts:
                    if reg == GSCH.MIXED:
                        c = decode(t)[2]
                    elif reg >= GSCH.PAIR:
                        c = (t >> 11) & 31
                        c -= 32 if c >= 16 else 0
                    else:
                        c = (t >> 10) & 63
                        c -= 64 if c >= 32 else 0
                    w += abs(c)
                return w

            out = {}
            for k in range(np_ + no):
                desc, na, nb, q, flags = rounds_c[first + k]
                sdesc, stride, ra, rb, wa, wb = srounds[first + k]
                for lane in range(G):
                    at = 4 * (sdesc + lane * stride)
                    hw = [h for w in words[at:at + 4 * q] for h in (w & 0xFFFF, w >> 16)]
                    if hw[0] == GSCH.IDLE:
                        continue
                    if k < np_:
                        x = sum(term(t, ra) for t in hw[1:1 + na])
                        y = sum(term(t, rb) for t in hw[GSCH.B_AT:GSCH.B_AT + nb])
                        # the weights the kernel relies on (<= 2: the operand is not carried)
                        assert weight(hw[1:1 + na], ra) <= wa and weight(hw[GSCH.B_AT:GSCH.B_AT + nb], rb) <= wb
                        regions[T][hw[0]] = x * y % P
                    else:
                        out[hw[0]] = sum(term(t, ra) for t in hw[1:1 + na]) % P
            got = [out.get(i) for i in range(max(out) + 1)]
            assert got == symbolic(prods, outs, a, b, t0), (G, name)


def test_generated_header_is_current(variants):
    """schedules_gen.hpp is what the generator writes now (the build does not regenerate it)"""
    path = os.path.join(os.path.dirname(HERE), "icp-gpu_amd", "csrc", "schedules_gen.hpp")
    text = open(path).read()
    for G in GSCH.GS:
        tables, nbytes = GSCH.emit_tables(G, [(n, p, o) for n, p, o, _, _ in GSCH.build_rounds_v(G, variants)])
        for line in tables:
            assert line in text, "schedules_gen.hpp is stale: run gen_schedules.py (G=%d)" % G
