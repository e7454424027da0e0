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

    def reserve(self, n):
        """n T slots that an earlier op has filled: read as products, computed by no round"""
        start = len(self.products)
        self.products += [None] * n
        return [Lin({(T, start + i): 1}) for i in range(n)]


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


# When set, an Fp2 product whose Karatsuba sum operand would have more than this many
# terms is taken schoolbook (four products of the shorter operands).  The round tables
# of each lane count use whichever form of an op needs fewer rounds, then fewer terms.
CAP_TERMS = None


def f2mul(R, a, b):
    if CAP_TERMS is not None and (len((a.c0 + a.c1).t) > CAP_TERMS or len((b.c0 + b.c1).t) > CAP_TERMS):
        t00, t11 = R.mul(a.c0, b.c0), R.mul(a.c1, b.c1)
        return F2(t00 - t11, R.mul(a.c0, b.c1) + R.mul(a.c1, b.c0))
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
    """(a + b s)^2 with s^2 = 1 + u: (a^2 + (1+u) b^2, 2ab).  2ab is taken as a product
    (three Fp products whose operands have at most two terms) rather than as
    (a + b)^2 - a^2 - b^2 (two products of four-term operands): one more product, but
    half the operand gathering in the rounds of the lanes-per-item kernels"""
    t0, t1 = f2sqr(R, a), f2sqr(R, b)
    return t1.nr() + t0, f2mul(R, a, b).scale(2)


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


# ---- merged line steps of the Miller loop ----
# Region B of these ops is the item's line block: L1 (B 0..5), L2 (B 6..11) = next lines
# (c0, c1, c2) of the two pairs, P (B 12..15) = (P1.x, P1.y, P2.x, P2.y), S (B 16..27) =
# the scaled lines (c2, c1 P.x, c0 P.y) of pair 1 then pair 2, which the line product
# overwrites with LL (B 16..25).  A line is the sparse element x0 + x1 v + y1 v w.
ZERO2 = F2(Lin(), Lin())


def line_product(R, l1, l2):
    """(x0 + x1 v + y1 v w)(u0 + u1 v + z1 v w) with w^2 = v, v^3 = 1 + u: six Fp2 products
    (three of them Karatsuba sums); the w^0 part is full, the w^1 part has no 1-term"""
    x0, x1, y1 = l1
    u0, u1, z1 = l2
    t00, t11, tyy = f2mul(R, x0, u0), f2mul(R, x1, u1), f2mul(R, y1, z1)
    s01 = f2mul(R, x0 + x1, u0 + u1) - t00 - t11
    s0y = f2mul(R, x0 + y1, u0 + z1) - t00 - tyy
    s1y = f2mul(R, x1 + y1, u1 + z1) - t11 - tyy
    return F12(F6(t00 + tyy.nr(), s01, t11), F6(ZERO2, s0y, s1y))


def op_ll():
    # LL = line1 * line2 from the scaled lines S (B 16..27); outputs (c0.c0, c0.c1, c0.c2,
    # c1.c1, c1.c2), 10 Fp written over S
    R = Rec()
    l1 = [f2(B, 8), f2(B, 9), f2(B, 10)]
    l2 = [f2(B, 11), f2(B, 12), f2(B, 13)]
    ll = line_product(R, l1, l2)
    return R, outs_of(ll.c0.c[0]) + outs_of(ll.c0.c[1]) + outs_of(ll.c0.c[2]) + outs_of(ll.c1.c[1]) + \
        outs_of(ll.c1.c[2])


def op_sqr_ll():
    # f^2 (f in A) and the next step's line product LL from the scaled lines S
    # (B 16..27) in the same rounds: f^2 at outputs 0..11, LL at 28..37 (the S slots,
    # relative to f, written over S)
    R = Rec()
    sq = outs_of(f12sqr(R, f12(A)))
    l1 = [f2(B, 8), f2(B, 9), f2(B, 10)]
    l2 = [f2(B, 11), f2(B, 12), f2(B, 13)]
    ll = line_product(R, l1, l2)
    llo = outs_of(ll.c0.c[0]) + outs_of(ll.c0.c[1]) + outs_of(ll.c0.c[2]) + outs_of(ll.c1.c[1]) + \
        outs_of(ll.c1.c[2])
    return R, sq + [None] * 16 + llo


def ll_operand():
    return F12(F6(f2(B, 8), f2(B, 9), f2(B, 10)), F6(ZERO2, f2(B, 11), f2(B, 12)))


def scaled_lines(R):
    """(c2, c1 P.x, c0 P.y) of both pairs from L1, L2 and P (api.cu:37)"""
    out = []
    for pair in range(2):
        c0, c1, c2 = f2(B, 3 * pair), f2(B, 3 * pair + 1), f2(B, 3 * pair + 2)
        x, y = fp(B, 12 + 2 * pair), fp(B, 13 + 2 * pair)
        out += outs_of(c2) + outs_of(f2mul_fp(R, c1, x)) + outs_of(f2mul_fp(R, c0, y))
    return out


def op_fl_scale():
    # f (A) times LL, and the next step's scaled lines: outputs f at 0..11 and the
    # scaled lines at 28..39 (the item's S slots, relative to f)
    R = Rec()
    outs = outs_of(f12mul(R, f12(A), ll_operand()))
    return R, outs + [None] * 16 + scaled_lines(R)


def op_fl():
    R = Rec()
    return R, outs_of(f12mul(R, f12(A), ll_operand()))


def op_mul_s():
    # The Fp12 product in two ops, for kernels whose LDS holds fewer product slots:
    # OP_MUL_S forms s = (a0 + a1)(b0 + b1) (an Fp6, 6 outputs); OP_MUL_PQ reads it from
    # T slots 0..5, forms a0 b0 and a1 b1 in T slots 6.. and combines the three as
    # f12mul does.
    R = Rec()
    a, b = f12(A), f12(B)
    s = f6mul(R, a.c0 + a.c1, b.c0 + b.c1)
    return R, [x for c in s.c for x in (c.c0, c.c1)]


def op_mul_pq():
    R = Rec()
    t = R.reserve(6)
    s = F6(F2(t[0], t[1]), F2(t[2], t[3]), F2(t[4], t[5]))
    a, b = f12(A), f12(B)
    aa, bb = f6mul(R, a.c0, b.c0), f6mul(R, a.c1, b.c1)
    return R, outs_of(F12(aa + bb.mul_v(), s - aa - bb))


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
    ("OP_LL", op_ll),
    ("OP_FL_SCALE", op_fl_scale),
    ("OP_FL", op_fl),
    ("OP_SQR_LL", op_sqr_ll),
    ("OP_MUL_S", op_mul_s),
    ("OP_MUL_PQ", op_mul_pq),
]


# ---- per-G round tables ----
# Each op runs as rounds.  In a product round every lane of an item's group computes
# at most one product: operand A = sum of its A terms, B = sum of its B terms, each term
# a (slot, small signed coefficient) pair.  The number of terms is the same on every
# lane of a round (shorter lists are padded with zero-coefficient terms), so the kernel
# executes one instruction stream with no divergence.  Output rounds likewise compute
# one output per lane as a sum of terms over products, inputs and constants.
#
# Term (16 bits): bits 0..5 slot index in its region, bits 6..7 region (A, B, T, K),
# bits 8..15 the coefficient as a signed byte.  A lane's record of one round is a whole
# number of 16-byte units: product rounds [dst, A terms (8 places), B terms (8 places)]
# padded to 24 halfwords; output rounds [dst, terms] padded to a multiple of 8.
# dst = 0xffff marks an idle lane.  Lane s of the group reads the record at
# desc + s * q (in 16-byte units).
GS = [8, 9, 12, 13, 16, 18, 20, 24, 27, 32]
IDLE = 0xFFFF
PROD_Q = 3        # 24 halfwords per product record
B_AT = 9          # first B term of a product record
MAX_OPERAND_TERMS = 8
MAX_OUT_Q = 5     # output records of at most 39 terms


def term16(region, idx, coef):
    assert 0 <= idx < 64 and -128 <= coef <= 127 and coef != 0
    return (region << 6) | idx | ((coef & 0xFF) << 8)


def lin_terms(lin):
    return [term16(r, i, c) for (r, i), c in sorted(lin.t.items())]


def weight(lin):
    return sum(abs(c) for c in lin.t.values())


def op_programs(cap=None):
    """[(name, products [(A Lin, B Lin)], outputs [Lin])], products in T order, with
    the operand-weight bounds the kernel's arithmetic relies on checked; cap = the
    Karatsuba operand cap of f2mul (None: plain Karatsuba everywhere)"""
    global CAP_TERMS
    progs = []
    for name, fn in OPS:
        CAP_TERMS = cap
        try:
            R, outs = fn()
        finally:
            CAP_TERMS = None
        for a, b in (p for p in R.products if p is not None):
            # operands: at most 8 terms and weight 16 (|value| < 2^386.4 keeps the
            # Montgomery product N1); outputs: weight < 2^12 (64-bit accumulation)
            assert len(a.t) <= MAX_OPERAND_TERMS and len(b.t) <= MAX_OPERAND_TERMS, name
            assert weight(a) <= 16 and weight(b) <= 16, name
        for o in outs:
            assert o is None or (weight(o) < 4096 and len(o.t) <= 8 * MAX_OUT_Q - 1), name
        # hazard check: an output that reads input A index k other than its own index
        # would see a value overwritten in place; the kernels write such outputs to a
        # distinct register.
        reads_other = any(any(r == A and idx != i for (r, idx) in o.t) for i, o in enumerate(outs) if o is not None)
        progs.append((name, list(R.products), outs, reads_other))
    return progs


def build_rounds(G, progs):
    """per-op round lists for a group of G lanes:
    [(name, prod_rounds, out_rounds)], a round = (na, nb, [per-lane record halfwords])"""
    return [(n, p, o) for n, p, o, _, _ in build_rounds_v(G, [progs])]


def round_cost(prounds, orounds):
    """rounds first, then operand and output terms per lane"""
    return (len(prounds), len(orounds), sum(na + nb for na, nb, _ in prounds) + sum(n for n, _, _ in orounds))


def build_rounds_v(G, variants):
    """like build_rounds, choosing per op the cheapest of several programs (same op
    list, e.g. with and without the Karatsuba cap): [(name, prod_rounds, out_rounds,
    products, outputs)]"""
    best = None
    for progs in variants:
        ops = build_rounds_one(G, progs)
        if best is None:
            best = [(n, p, o, prods, outs) for (n, p, o), (_, prods, outs, _) in zip(ops, progs)]
            continue
        for k, ((n, p, o), (_, prods, outs, _)) in enumerate(zip(ops, progs)):
            if round_cost(p, o) < round_cost(best[k][1], best[k][2]):
                best[k] = (n, p, o, prods, outs)
    return best


def build_rounds_one(G, progs):
    ops = []
    for name, prods, outs, _ in progs:
        order = sorted((j for j in range(len(prods)) if prods[j] is not None),
                       key=lambda j: (len(prods[j][0].t) + len(prods[j][1].t), len(prods[j][0].t)))
        prounds = []
        for r0 in range(0, len(order), G):
            chunk = order[r0:r0 + G]
            na = max(len(prods[j][0].t) for j in chunk)
            nb = max(len(prods[j][1].t) for j in chunk)
            recs = []
            for s in range(G):
                rec = [0] * (8 * PROD_Q)
                rec[0] = IDLE
                if s < len(chunk):
                    j = chunk[s]
                    rec[0] = j
                    ta, tb = lin_terms(prods[j][0]), lin_terms(prods[j][1])
                    rec[1:1 + len(ta)] = ta
                    rec[B_AT:B_AT + len(tb)] = tb
                recs.append(rec)
            prounds.append((na, nb, recs))
        oorder = sorted((i for i in range(len(outs)) if outs[i] is not None), key=lambda i: len(outs[i].t))
        orounds = []
        for r0 in range(0, len(oorder), G):
            chunk = oorder[r0:r0 + G]
            n = max(len(outs[i].t) for i in chunk)
            q = (1 + n + 7) // 8
            recs = []
            for s in range(G):
                rec = [0] * (8 * q)
                rec[0] = IDLE
                if s < len(chunk):
                    i = chunk[s]
                    rec[0] = i
                    t = lin_terms(outs[i])
                    rec[1:1 + len(t)] = t
                recs.append(rec)
            orounds.append((n, 0, recs))
        ops.append((name, prounds, orounds))
    return ops


def limbs28(v):
    return [(v >> (28 * i)) & ((1 << 28) - 1) for i in range(14)]


# The final exponentiation's straight-line ops get a second copy of their records
# (DESC_S, lane-major per op): a term list (operand A, operand B, or an output's terms)
# whose busy lanes all read one region is encoded relative to that region, term = slot *
# 14 (the slot's word offset, bits 0..9) | coefficient (signed, bits 10..15), so that the
# kernel, knowing the region at compile time, forms the term's address with one
# instruction.  Mixed lists keep the encoding above.  Region codes per list: 0..3 = that
# region (A, B, T, K), 4 = mixed.  The Miller loop's straight-line ops use DESC_S only at
# the lane counts where that measured faster: at 8, 13 and 16 lanes the kernel, at the
# edge of its register budget, spills heavily with it and runs slower.
FE_STATIC_OPS = ("OP_MUL", "OP_CYC_SQR")
ML_STATIC_OPS = ("OP_LL", "OP_FL_SCALE", "OP_SQR_LL")
ML_STATIC_LANES = (12, 20, 24, 32)
# the split product of the compact final exponentiation (12 lanes), emitted after the
# others so that their records keep their places
FEC_STATIC_OPS = ("OP_MUL_S", "OP_MUL_PQ")
FEC_STATIC_LANES = (12,)


def static_ops(G):
    return FE_STATIC_OPS + (ML_STATIC_OPS if G in ML_STATIC_LANES else ()) + \
        (FEC_STATIC_OPS if G in FEC_STATIC_LANES else ())
MIXED = 4


# A list whose busy lanes read two regions r0 < r1 has code PAIR + 4 r0 + r1 and terms
# slot * 14 (bits 0..9) | (region == r1) << 10 | coefficient (signed, bits 11..15).
PAIR = 8


def list_region(recs, lo, n):
    regs = sorted(set((t >> 6) & 3 for r in recs if r[0] != IDLE for t in r[lo:lo + n] if (t >> 8) & 0xFF))
    if len(regs) == 1:
        return regs[0]
    if len(regs) == 2:
        return PAIR + 4 * regs[0] + regs[1]
    return MIXED


def term_rel(t, code):
    idx, reg, coef = t & 63, (t >> 6) & 3, (t >> 8) & 0xFF
    if coef >= 128:
        coef -= 256
    if coef == 0:
        return 0
    assert idx * 14 < 1024
    if code < MIXED:
        assert reg == code and -32 <= coef <= 31
        return (idx * 14) | ((coef & 0x3F) << 10)
    r0, r1 = (code - PAIR) // 4, (code - PAIR) % 4
    assert reg in (r0, r1) and -16 <= coef <= 15
    return (idx * 14) | ((reg == r1) << 10) | ((coef & 0x1F) << 11)


def static_record(rec, lists):
    """rec with each (lo, n, region code) list re-encoded unless its code is MIXED"""
    out = list(rec)
    for lo, n, code in lists:
        if code != MIXED:
            out[lo:lo + n] = [term_rel(t, code) for t in rec[lo:lo + n]]
    return out


def emit_tables(G, ops):
    """C++ arrays of one G: OPS<G>, ROUNDS<G>, DESC<G>, and DESC_S<G> with the static
    rounds' records (SROUNDS_C: offset into DESC_S and the region code of each list)"""
    op_rows, round_rows, words = [], [], []
    srows, swords = [], []
    for name, prounds, orounds in ops:
        op_rows.append("{%d, %d, %d}" % (len(round_rows), len(prounds), len(orounds)))
        for k, (na, nb, recs) in enumerate(prounds + orounds):
            if name not in static_ops(G):
                srows.append("{0xffffffffu, 0, %d, %d, 0, 0}" % (MIXED, MIXED))
            q = len(recs[0]) // 8
            flags = 0
            if k < len(prounds):
                busy = [r for r in recs if r[0] != IDLE]
                # an operand that is one slot with coefficient +1 on every busy lane is
                # loaded as it is (slot values are valid multiplier inputs)
                if na == 1 and all((r[1] >> 8) & 0xFF == 1 for r in busy):
                    flags |= 1
                if nb == 1 and all((r[B_AT] >> 8) & 0xFF == 1 for r in busy):
                    flags |= 2
            round_rows.append("{%d, %d, %d, %d, %d}" % (len(words) // 4, na, nb, q, flags))
            for rec in recs:
                for k in range(0, len(rec), 2):
                    words.append(rec[k] | (rec[k + 1] << 16))
        if name in static_ops(G):
            # lane-major: lane s's records of all the op's rounds are contiguous, so that a
            # lane addresses every round from one base with a small immediate offset
            rounds = prounds + orounds
            stride = sum(len(recs[0]) // 8 for _, _, recs in rounds)
            base = len(swords) // 4
            off = 0
            lane_words = [[] for _ in range(G)]
            for k, (na, nb, recs) in enumerate(rounds):
                def reg(lo, n):
                    return list_region(recs, lo, n)
                if k < len(prounds):
                    lists = [(1, na, reg(1, na)), (B_AT, nb, reg(B_AT, nb))]
                else:
                    lists = [(1, na, reg(1, na))]
                # largest sum of |coefficients| of each operand list over the busy lanes: a
                # sum of weight <= 2 of normalized slots is a valid multiplier operand as
                # it is (no carry pass)
                def weight(lo, n):
                    w = 0
                    for r in recs:
                        if r[0] != IDLE:
                            w = max(w, sum(abs(((t >> 8) & 0xFF) - (256 if (t >> 8) & 0x80 else 0))
                                           for t in r[lo:lo + n]))
                    return w
                wa = weight(1, na) if k < len(prounds) else 0
                wb = weight(B_AT, nb) if k < len(prounds) else 0
                srows.append("{%d, %d, %d, %d, %d, %d}" % (base + off, stride, lists[0][2],
                                                           lists[1][2] if len(lists) > 1 else MIXED, wa, wb))
                for sl, rec in enumerate(recs):
                    srec = static_record(rec, lists)
                    lane_words[sl] += [srec[j] | (srec[j + 1] << 16) for j in range(0, len(srec), 2)]
                off += len(recs[0]) // 8
            for sl in range(G):
                assert len(lane_words[sl]) == 4 * stride
                swords += lane_words[sl]
    assert len(round_rows) < 65536
    out = ["// lanes per item G = %d: %d rounds, %d bytes of records" % (G, len(round_rows), 4 * len(words)),
           "__constant__ const OpRounds OPS%d[NUM_OPS] = {%s};" % (G, ", ".join(op_rows)),
           "__constant__ const Round ROUNDS%d[%d] = {%s};" % (G, len(round_rows), ", ".join(round_rows)),
           "__device__ __align__(16) const uint32_t DESC%d[%d] = {%s};" % (
               G, len(words), ", ".join("0x%x" % w for w in words)),
           "__device__ __align__(16) const uint32_t DESC_S%d[%d] = {%s};" % (
               G, len(swords), ", ".join("0x%x" % w for w in swords)),
           "template <> struct Tab<%d> {" % G,
           "  // the same tables as compile-time constants (round shapes of the ops run as straight-line code)",
           "  static constexpr OpRounds OPS_C[NUM_OPS] = {%s};" % ", ".join(op_rows),
           "  static constexpr Round ROUNDS_C[%d] = {%s};" % (len(round_rows), ", ".join(round_rows)),
           "  static constexpr SRound SROUNDS_C[%d] = {%s};" % (len(srows), ", ".join(srows)),
           "  static __device__ __forceinline__ const uint4* sdesc() { return reinterpret_cast<const uint4*>(DESC_S%d); }"
           % G,
           "  static __device__ __forceinline__ const COpRounds* ops() { return (const COpRounds*)OPS%d; }" % G,
           "  static __device__ __forceinline__ const CRound* rounds() { return (const CRound*)ROUNDS%d; }" % G,
           "  static __device__ __forceinline__ const uint4* desc() { return reinterpret_cast<const uint4*>(DESC%d); }"
           % G,
           "};"]
    return out, 4 * len(words)


def main():
    progs = op_programs()
    variants = [progs, op_programs(cap=4)]
    lines = ["// Generated by gen_schedules.py -- do not edit.",
             "// Fp12-level operations as per-G round tables (see gen_schedules.py for the encoding).",
             "#pragma once", "#include <hip/hip_runtime.h>", "#include <stdint.h>", "namespace bls {",
             "namespace sched {"]
    for k, (name, prods, outs, ro) in enumerate(progs):
        lines.append("constexpr int %s = %d;  // %d products, %d outputs%s" %
                     (name, k, sum(p is not None for p in prods), sum(o is not None for o in outs),
                      ", outputs read other inputs" if ro else ""))
    lines.append("constexpr int NUM_OPS = %d;" % len(progs))
    max_t = max(len(prods) for G in GS for _, _, _, prods, _ in build_rounds_v(G, variants))
    lines.append("constexpr int MAX_PRODUCTS = %d;  // T slots an item needs (any op, any G)" % max_t)
    lines.append("constexpr int NUM_CONSTS = %d;" % len(CONSTS.vals))
    lines.append("constexpr int PROD_Q = %d;  // 16-byte units per product record" % PROD_Q)
    lines.append("constexpr int B_AT = %d;    // halfword index of the first B term" % B_AT)
    lines.append("constexpr int MAX_OUT_Q = %d;" % MAX_OUT_Q)
    lines.append("constexpr uint16_t IDLE = 0x%x;" % IDLE)
    lines.append("struct OpRounds { uint16_t first; uint8_t np, no; };")
    lines.append("// flags (product rounds): 1 = operand A is one slot with coefficient 1 on every")
    lines.append("// busy lane, 2 = the same for B")
    lines.append("struct Round { uint32_t desc; uint8_t na, nb, q, flags; };")
    lines.append("// the straight-line rounds' records in DESC_S, lane-major per op: lane s's record of")
    lines.append("// the round is at desc + s * stride (16-byte units); the region code of operand A (or")
    lines.append("// an output round's terms) and of operand B: 0..3 = all terms in region A, B, T, K,")
    lines.append("// encoded slot * 14 | coefficient << 10; 8 + 4 r0 + r1 = terms in regions r0 < r1,")
    lines.append("// slot * 14 | (region == r1) << 10 | coefficient << 11; 4 = terms as in DESC")
    lines.append("// wa, wb: the largest sum of |coefficients| of operand A / B over the busy lanes")
    lines.append("struct SRound { uint32_t desc; uint16_t stride; uint8_t ra, rb, wa, wb; };")
    lines.append("constexpr int MIXED = %d, PAIR = %d;" % (MIXED, PAIR))
    lines.append("// the op and round tables sit in the constant address space, so that loads at")
    lines.append("// uniform indices are scalar loads")
    lines.append("typedef const OpRounds __attribute__((address_space(4))) COpRounds;")
    lines.append("typedef const Round __attribute__((address_space(4))) CRound;")
    lines.append("template <int G> struct Tab;")
    lines.append("// constants (internal Montgomery form, 28-bit limbs)")
    kvals = ["{%s}" % ", ".join("0x%07x" % x for x in limbs28(v * RI % P)) for v in CONSTS.vals] or ["{0}"]
    lines.append("__device__ const int32_t KCONST[%d][14] = {%s};" % (max(1, len(CONSTS.vals)), ", ".join(kvals)))
    sizes = {}
    for G in GS:
        ops = [(n, p, o) for n, p, o, _, _ in build_rounds_v(G, variants)]
        t, nbytes = emit_tables(G, ops)
        lines += t
        sizes[G] = (nbytes, sum(len(p) + len(o) for _, p, o in ops))
    lines += ["}  // namespace sched", "}  // namespace bls"]
    with open(os.path.join(HERE, "schedules_gen.hpp"), "w") as f:
        f.write("\n".join(lines) + "\n")
    for name, prods, outs, ro in progs:
        print("%-12s products %3d outputs %3d%s" % (name, sum(p is not None for p in prods),
                                                   sum(o is not None for o in outs),
                                                   "  (reads other inputs)" if ro else ""))
    for G, (nb, nr) in sizes.items():
        print("G=%2d: %4d rounds, %6d bytes of records" % (G, nr, nb))
    print("constants", len(CONSTS.vals))


if __name__ == "__main__":
    sys.exit(main())
