#!/usr/bin/env python3
"""Generate schedules_gen.hpp: Fp12-level operations as bilinear schedules.

Every extension-field operation used by the verifier is a set of independent Fp
products of small-integer linear combinations of its input components, followed by
small-integer linear combinations of those products (and inputs) for the outputs.  A
group of G lanes executes one such schedule for one verification item: the products
are distributed over the lanes (each lane gathers its operands from LDS, multiplies,
stores the product), then the outputs are distributed the same way.

The products and combinations are obtained by running the tower formulas below
symbolically.  They follow the reference's formulas (gpu-crypto/cuda/field_sppark.cuh:
52-197, api.cu:37 for the line evaluation); the Frobenius maps are expressed with
constants computed here from their closed forms ((1+u)^((p-1)/3) etc.).

Term encoding (one byte): bits 7..6 select the operand region (0 = A, 1 = B, 2 = T
product scratch, 3 = K constants), bits 5..0 the Fp index inside the region.  A
linear combination is a count of positive terms, a count of negative terms, then the
terms (a coefficient c appears as |c| copies of its term).  Product record: dst (T
index), a_pos, a_neg, b_pos, b_neg, then the terms of A and of B.  Output record: dst
index, n_pos, n_neg, then the terms.
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
P = 0x1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f6241eabfffeb153ffffb9feffffffffaaab
RI = 1 << 392

A, B, T, K = 0, 1, 2, 3


class Lin:
    """linear combination: {(region, index): coefficient}"""
    __slots__ = ("t",)

    def __init__(self, t=None):
        self.t = {k: v for k, v in (t or {}).items() if v}

    def __add__(self, o):
        r = dict(self.t)
        for k, v in o.t.items():
            r[k] = r.get(k, 0) + v
        return Lin(r)

    def __sub__(self, o):
        return self + o.scale(-1)

    def __neg__(self):
        return self.scale(-1)

    def scale(self, k):
        return Lin({a: v * k for a, v in self.t.items()})

    def regions(self):
        return {a[0] for a in self.t}


class Rec:
    """records the products of one schedule"""

    def __init__(self):
        self.products = []
        self.const = {}  # canonical Fp value -> K index (shared table, see ConstTable)

    def mul(self, a, b):
        assert T not in a.regions() and T not in b.regions(), "schedule must be single-layer"
        if not a.t or not b.t:
            return Lin()
        self.products.append((a, b))
        return Lin({(T, len(self.products) - 1): 1})


class ConstTable:
    def __init__(self):
        self.vals = []

    def idx(self, v):
        v %= P
        if v not in self.vals:
            self.vals.append(v)
        return self.vals.index(v)


CONSTS = ConstTable()


def kconst(v):
    return Lin({(K, CONSTS.idx(v)): 1}) if v % P else Lin()


# ---- symbolic tower (same shapes as tower.hpp) ----
class F2:
    def __init__(self, c0, c1):
        self.c0, self.c1 = c0, c1

    def __add__(self, o):
        return F2(self.c0 + o.c0, self.c1 + o.c1)

    def __sub__(self, o):
        return F2(self.c0 - o.c0, self.c1 - o.c1)

    def __neg__(self):
        return F2(-self.c0, -self.c1)

    def scale(self, k):
        return F2(self.c0.scale(k), self.c1.scale(k))

    def nr(self):
        return F2(self.c0 - self.c1, self.c0 + self.c1)

    def conj(self):
        return F2(self.c0, -self.c1)


def f2mul(R, a, b):
    t0 = R.mul(a.c0, b.c0)
    t1 = R.mul(a.c1, b.c1)
    s = R.mul(a.c0 + a.c1, b.c0 + b.c1)
    return F2(t0 - t1, s - t0 - t1)


def f2sqr(R, a):
    t = R.mul(a.c0, a.c1)
    c0 = R.mul(a.c0 + a.c1, a.c0 - a.c1)
    return F2(c0, t.scale(2))


def f2mul_fp(R, a, s):
    return F2(R.mul(a.c0, s), R.mul(a.c1, s))


def f2mul_const(R, a, c):
    """a * (c0 + c1 u) for a known constant; zero parts drop out"""
    c0, c1 = c[0] % P, c[1] % P
    if c1 == 0:
        return F2(R.mul(a.c0, kconst(c0)), R.mul(a.c1, kconst(c0)))
    if c0 == 0:
        return F2(-R.mul(a.c1, kconst(c1)), R.mul(a.c0, kconst(c1)))
    t0 = R.mul(a.c0, kconst(c0))
    t1 = R.mul(a.c1, kconst(c1))
    s = R.mul(a.c0 + a.c1, kconst(c0 + c1))
    return F2(t0 - t1, s - t0 - t1)


class F6:
    def __init__(self, c0, c1, c2):
        self.c = [c0, c1, c2]

    def __add__(self, o):
        return F6(*[x + y for x, y in zip(self.c, o.c)])

    def __sub__(self, o):
        return F6(*[x - y for x, y in zip(self.c, o.c)])

    def __neg__(self):
        return F6(*[-x for x in self.c])

    def mul_v(self):
        return F6(self.c[2].nr(), self.c[0], self.c[1])


def f6mul(R, a, b):
    a0, a1, a2 = a.c
    b0, b1, b2 = b.c
    aa, bb, cc = f2mul(R, a0, b0), f2mul(R, a1, b1), f2mul(R, a2, b2)
    t0 = f2mul(R, a1 + a2, b1 + b2)
    t1 = f2mul(R, a0 + a1, b0 + b1)
    t2 = f2mul(R, a0 + a2, b0 + b2)
    return F6(aa + (t0 - bb - cc).nr(), t1 - aa - bb + cc.nr(), t2 - aa - cc + bb)


def f6mul_by_01(R, s, c0, c1):
    aa, bb = f2mul(R, s.c[0], c0), f2mul(R, s.c[1], c1)
    o0 = f2mul(R, s.c[2], c1).nr() + aa
    o1 = f2mul(R, c0 + c1, s.c[0] + s.c[1]) - aa - bb
    o2 = f2mul(R, s.c[2], c0) + bb
    return F6(o0, o1, o2)


def f6mul_by_1(R, s, c1):
    return F6(f2mul(R, s.c[2], c1).nr(), f2mul(R, s.c[0], c1), f2mul(R, s.c[1], c1))


class F12:
    def __init__(self, c0, c1):
        self.c0, self.c1 = c0, c1

    def comps(self):
        return [self.c0.c[0], self.c0.c[1], self.c0.c[2], self.c1.c[0], self.c1.c[1], self.c1.c[2]]


def f12mul(R, a, b):
    aa, bb = f6mul(R, a.c0, b.c0), f6mul(R, a.c1, b.c1)
    s = f6mul(R, a.c0 + a.c1, b.c0 + b.c1)
    return F12(aa + bb.mul_v(), s - aa - bb)


def f12sqr(R, a):
    ab = f6mul(R, a.c0, a.c1)
    t = f6mul(R, a.c0 + a.c1.mul_v(), a.c0 + a.c1)
    return F12(t - ab - ab.mul_v(), ab + ab)


def f12mul_by_014(R, f, c0, c1, c4):
    aa = f6mul_by_01(R, f.c0, c0, c1)
    bb = f6mul_by_1(R, f.c1, c4)
    t = f6mul_by_01(R, f.c1 + f.c0, c0, c1 + c4)
    return F12(bb.mul_v() + aa, t - aa - bb)


def fp4sqr(R, a, b):
    t0, t1 = f2sqr(R, a), f2sqr(R, b)
    return t1.nr() + t0, f2sqr(R, a + b) - t0 - t1


def cyc_sqr(R, f):
    z0, z4, z3 = f.c0.c
    z2, z1, z5 = f.c1.c
    A0, A1 = fp4sqr(R, z0, z1)
    B0, B1 = fp4sqr(R, z2, z3)
    C0, C1 = fp4sqr(R, z4, z5)
    z0 = (A0 - z0).scale(2) + A0
    z1 = (A1 + z1).scale(2) + A1
    z4 = (B0 - z4).scale(2) + B0
    z5 = (B1 + z5).scale(2) + B1
    nC1 = C1.nr()
    z2 = (nC1 + z2).scale(2) + nC1
    z3 = (C0 - z3).scale(2) + C0
    return F12(F6(z0, z4, z3), F6(z2, z1, z5))


# Frobenius powers: frob^k maps tower component x_j to gamma_{k,j} * conj^k(x_j)
def f2pow(a, e):
    def m(x, y):
        return ((x[0] * y[0] - x[1] * y[1]) % P, (x[0] * y[1] + x[1] * y[0]) % P)
    r = (1, 0)
    while e:
        if e & 1:
            r = m(r, a)
        a = m(a, a)
        e >>= 1
    return r


def frob_consts(k):
    """gamma for the 6 tower components (order c0.c0, c0.c1, c0.c2, c1.c0, c1.c1, c1.c2):
    element = sum_j x_j w^(e_j) with e = (0, 2, 4, 1, 3, 5); frob^k(x w^e) =
    conj^k(x) * (w^e)^(p^k) and w^(p^k) = w * (1+u)^((p^k-1)/6)."""
    beta = (1, 1)
    exps = [0, 2, 4, 1, 3, 5]
    return [f2pow(beta, e * (P ** k - 1) // 6) for e in exps]


def frobk(R, f, k):
    comps = f.comps()
    g = frob_consts(k)
    out = []
    for x, c in zip(comps, g):
        if k % 2:
            x = x.conj()
        if c == (1, 0):
            out.append(x)
        else:
            out.append(f2mul_const(R, x, c))
    return F12(F6(*out[:3]), F6(*out[3:]))


def conj12(f):
    return F12(f.c0, -f.c1)


# ---- operands ----
def fp(region, i):
    return Lin({(region, i): 1})


def f2(region, i):
    return F2(fp(region, 2 * i), fp(region, 2 * i + 1))


def f12(region):
    return F12(F6(f2(region, 0), f2(region, 1), f2(region, 2)), F6(f2(region, 3), f2(region, 4), f2(region, 5)))


def outs_of(x):
    if isinstance(x, F12):
        return [c for comp in x.comps() for c in (comp.c0, comp.c1)]
    if isinstance(x, F2):
        return [x.c0, x.c1]
    return list(x)


# ---- schedule definitions ----
def op_mul():
    R = Rec()
    return R, outs_of(f12mul(R, f12(A), f12(B)))


def op_sqr():
    R = Rec()
    return R, outs_of(f12sqr(R, f12(A)))


def op_cyc_sqr():
    R = Rec()
    return R, outs_of(cyc_sqr(R, f12(A)))


def op_conj():
    R = Rec()
    return R, outs_of(conj12(f12(A)))


def op_frob(k):
    def build():
        R = Rec()
        return R, outs_of(frobk(R, f12(A), k))
    return build


def op_ell():
    # f (A, 12 Fp) times the line with scaled coefficients held in B:
    # B[0..1] = c2, B[2..3] = c1 * P.x, B[4..5] = c0 * P.y  (api.cu:37)
    R = Rec()
    return R, outs_of(f12mul_by_014(R, f12(A), f2(B, 0), f2(B, 1), f2(B, 2)))


def op_scale():
    # A = line (c0, c1, c2) as 6 Fp, B = (P.x, P.y); outputs in B-layout of op_ell:
    # (c2, c1 * P.x, c0 * P.y)
    R = Rec()
    c0, c1, c2 = f2(A, 0), f2(A, 1), f2(A, 2)
    x, y = fp(B, 0), fp(B, 1)
    return R, outs_of(c2) + outs_of(f2mul_fp(R, c1, x)) + outs_of(f2mul_fp(R, c0, y))


def op_copy():
    R = Rec()
    return R, outs_of(f12(A))


def op_mul_conja():
    # conj(a) * b: the easy part's f^(p^6) * f^-1
    R = Rec()
    return R, outs_of(f12mul(R, conj12(f12(A)), f12(B)))


def f6(region, off=0):
    return F6(f2(region, off), f2(region, off + 1), f2(region, off + 2))


def op_n12():
    # Fp12 -> Fp6 norm c0^2 - v c1^2 (field_sppark.cuh:152-156)
    R = Rec()
    a = f12(A)
    t = f6mul(R, a.c0, a.c0) - f6mul(R, a.c1, a.c1).mul_v()
    return R, [x for c in t.c for x in (c.c0, c.c1)]


def op_n6a():
    # Fp6 -> (t0, t1, t2) of the Fp6 inverse (field_sppark.cuh:102-107)
    R = Rec()
    c0, c1, c2 = f2(A, 0), f2(A, 1), f2(A, 2)
    t0 = f2sqr(R, c0) - f2mul(R, c1, c2).nr()
    t1 = f2sqr(R, c2).nr() - f2mul(R, c0, c1)
    t2 = f2sqr(R, c1) - f2mul(R, c0, c2)
    return R, outs_of(t0) + outs_of(t1) + outs_of(t2)


def op_n6b():
    # sc = c0 t0 + nr(c2 t1 + c1 t2), c in A, t in B (field_sppark.cuh:107-108)
    R = Rec()
    c0, c1, c2 = f2(A, 0), f2(A, 1), f2(A, 2)
    t0, t1, t2 = f2(B, 0), f2(B, 1), f2(B, 2)
    sc = f2mul(R, c0, t0) + (f2mul(R, c2, t1) + f2mul(R, c1, t2)).nr()
    return R, outs_of(sc)


def op_n2():
    # Fp2 -> Fp norm c0^2 + c1^2
    R = Rec()
    c = f2(A, 0)
    return R, [R.mul(c.c0, c.c0) + R.mul(c.c1, c.c1)]


def op_inv2():
    # (c0 n^-1, -c1 n^-1): Fp2 inverse from the norm's inverse in B[0]
    R = Rec()
    c = f2(A, 0)
    ni = fp(B, 0)
    return R, [R.mul(c.c0, ni), -R.mul(c.c1, ni)]


def op_inv6():
    # t * sc^-1 (3 Fp2 products): Fp6 inverse, t in A, sc^-1 in B
    R = Rec()
    si = f2(B, 0)
    return R, outs_of(f2mul(R, f2(A, 0), si)) + outs_of(f2mul(R, f2(A, 1), si)) + outs_of(f2mul(R, f2(A, 2), si))


def op_inv12():
    # (c0 t^-1, -c1 t^-1): Fp12 inverse, f in A, t^-1 in B
    R = Rec()
    a = f12(A)
    ti = f6(B)
    return R, outs_of(F12(f6mul(R, a.c0, ti), -f6mul(R, a.c1, ti)))


def op_scale2():
    # both pairs at once: A = (line1 (6), line2 (6)), B = (P1.x, P1.y, P2.x, P2.y);
    # outputs (c2, c1*Px, c0*Py) of pair 1 then of pair 2 (12 Fp)
    R = Rec()
    out = []
    for pair in range(2):
        c0, c1, c2 = f2(A, 3 * pair), f2(A, 3 * pair + 1), f2(A, 3 * pair + 2)
        x, y = fp(B, 2 * pair), fp(B, 2 * pair + 1)
        out += outs_of(c2) + outs_of(f2mul_fp(R, c1, x)) + outs_of(f2mul_fp(R, c0, y))
    return R, out


OPS = [
    ("OP_MUL", op_mul),
    ("OP_SQR", op_sqr),
    ("OP_CYC_SQR", op_cyc_sqr),
    ("OP_CONJ", op_conj),
    ("OP_FROB1", op_frob(1)),
    ("OP_FROB2", op_frob(2)),
    ("OP_FROB3", op_frob(3)),
    ("OP_ELL", op_ell),
    ("OP_SCALE", op_scale),
    ("OP_COPY", op_copy),
    ("OP_MUL_CONJA", op_mul_conja),
    ("OP_N12", op_n12),
    ("OP_N6A", op_n6a),
    ("OP_N6B", op_n6b),
    ("OP_N2", op_n2),
    ("OP_INV2", op_inv2),
    ("OP_INV6", op_inv6),
    ("OP_INV12", op_inv12),
    ("OP_SCALE2", op_scale2),
]


def term_byte(region, idx):
    assert 0 <= idx < 64
    return (region << 6) | idx


def encode_terms(lin):
    """(n_pos, n_neg, term bytes): positive terms then negative terms, a coefficient c
    expanded into |c| copies of its term"""
    pos, neg = [], []
    for (region, idx), c in sorted(lin.t.items()):
        assert -8 <= c <= 8, c
        (pos if c > 0 else neg).extend([term_byte(region, idx)] * abs(c))
    return len(pos), len(neg), pos + neg


def limbs28(v):
    return [(v >> (28 * i)) & ((1 << 28) - 1) for i in range(14)]


def main():
    code = []
    ops_meta = []
    max_t = 0
    max_terms = 0
    for name, fn in OPS:
        R, outs = fn()
        # products are independent: order them by operand term count so that the lanes
        # of one round (products j = r*G .. r*G+G-1) do similar amounts of gathering;
        # outputs carry their destination index, so they can be ordered the same way
        order = sorted(range(len(R.products)),
                       key=lambda j: sum(abs(c) for c in R.products[j][0].t.values()) +
                       sum(abs(c) for c in R.products[j][1].t.values()))
        remap = {old: new for new, old in enumerate(order)}
        R.products = [R.products[j] for j in order]

        def ren(lin):
            return Lin({((T, remap[i]) if r == T else (r, i)): c for (r, i), c in lin.t.items()})
        outs = [ren(o) for o in outs]
        out_order = sorted(range(len(outs)), key=lambda i: sum(abs(c) for c in outs[i].t.values()))
        p_start, o_start = [], []
        for j, (a, b) in enumerate(R.products):
            ap, an, ea = encode_terms(a)
            bp, bn, eb = encode_terms(b)
            max_terms = max(max_terms, len(ea), len(eb))
            p_start.append(len(code))
            code += [j, ap, an, bp, bn] + ea + eb
        for i in out_order:
            # output i is written to index i of the destination register
            o = outs[i]
            op_, on, eo = encode_terms(o)
            max_terms = max(max_terms, len(eo))
            o_start.append(len(code))
            code += [i, op_, on] + eo
        # hazard check: an output that reads input A index k other than its own index
        # would see a value overwritten in place; the kernels write outputs to a
        # distinct register whenever this is flagged.
        reads_other = any(any(r == A and idx != i for (r, idx) in o.t) for i, o in enumerate(outs))
        max_t = max(max_t, len(R.products))
        ops_meta.append((name, len(R.products), len(outs), p_start, o_start, reads_other))
    assert len(code) < 65536
    lines = ["// Generated by gen_schedules.py -- do not edit.",
             "// Bilinear schedules of the Fp12 operations (see gen_schedules.py for the encoding).",
             "#pragma once", "#include <stdint.h>", "namespace bls {", "namespace sched {"]
    for k, (name, np_, no, _, _, ro) in enumerate(ops_meta):
        lines.append("constexpr int %s = %d;  // %d products, %d outputs%s" %
                     (name, k, np_, no, ", outputs read other inputs" if ro else ""))
    lines.append("constexpr int NUM_OPS = %d;" % len(ops_meta))
    lines.append("constexpr int MAX_PRODUCTS = %d;" % max_t)
    lines.append("constexpr int MAX_TERMS = %d;" % max_terms)
    lines.append("constexpr int NUM_CONSTS = %d;" % len(CONSTS.vals))
    lines.append("constexpr uint16_t NPROD[NUM_OPS] = {%s};" % ", ".join(str(m[1]) for m in ops_meta))
    lines.append("constexpr uint16_t NOUT[NUM_OPS] = {%s};" % ", ".join(str(m[2]) for m in ops_meta))
    starts, offs = [], []
    for m in ops_meta:
        offs.append(len(starts))
        starts += m[3]
        starts += m[4]
    lines.append("constexpr uint16_t START_OFF[NUM_OPS] = {%s};  // index into START: products then outputs" %
                 ", ".join(str(o) for o in offs))
    lines.append("constexpr int START_LEN = %d;" % len(starts))
    starts_p = starts + [0] * (len(starts) % 2)
    lines.append("__device__ __align__(4) const uint16_t START[%d] = {%s};" % (len(starts_p), ", ".join(map(str, starts_p))))
    lines.append("constexpr int CODE_LEN = %d;" % len(code))
    body = ", ".join(str(c) for c in code)
    pad = (-len(code)) % 4
    body = ", ".join(str(c) for c in code + [0] * pad)
    lines.append("__device__ __align__(4) const uint8_t CODE[%d] = {%s};" % (len(code) + pad, body))
    lines.append("// constants (internal Montgomery form, 28-bit limbs)")
    kvals = []
    for v in CONSTS.vals:
        kvals.append("{%s}" % ", ".join("0x%07x" % x for x in limbs28(v * RI % P)))
    if not kvals:
        kvals = ["{0}"]
    lines.append("__device__ const int32_t KCONST[%d][14] = {%s};" % (max(1, len(CONSTS.vals)), ", ".join(kvals)))
    lines += ["}  // namespace sched", "}  // namespace bls"]
    with open(os.path.join(HERE, "schedules_gen.hpp"), "w") as f:
        f.write("\n".join(lines) + "\n")
    for m in ops_meta:
        print("%-12s products %3d outputs %3d%s" % (m[0], m[1], m[2], "  (reads other inputs)" if m[5] else ""))
    print("constants", len(CONSTS.vals), "max terms", max_terms, "code bytes", len(code))


if __name__ == "__main__":
    sys.exit(main())
