"""Independent big-integer model of the BLS12-381 tower, written from the definitions
(Fp[u]/(u^2+1), then Fp12 = Fp2[w]/(w^6 - (1+u)) with v = w^2) with schoolbook
polynomial products.  It shares no formula with the oracle's Karatsuba code, so
agreement pins the oracle's tower arithmetic."""
P = 0x1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f6241eabfffeb153ffffb9feffffffffaaab
R384 = 1 << 384


def f2mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def f2add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


XI = (1, 1)  # w^6 = 1 + u


def f12_from_tower(c):
    """c: 6 Fp2 in the oracle's order (c0.c0, c0.c1, c0.c2, c1.c0, c1.c1, c1.c2) where
    element = (c00 + c01 v + c02 v^2) + (c10 + c11 v + c12 v^2) w, v = w^2.
    Returns coefficients of w^0..w^5."""
    c00, c01, c02, c10, c11, c12 = c
    return [c00, c10, c01, c11, c02, c12]


def f12_to_tower(w):
    return [w[0], w[2], w[4], w[1], w[3], w[5]]


def f12mul_w(a, b):
    r = [(0, 0)] * 11
    for i in range(6):
        for j in range(6):
            r[i + j] = f2add(r[i + j], f2mul(a[i], b[j]))
    out = r[:6]
    for k in range(6, 11):
        out[k - 6] = f2add(out[k - 6], f2mul(r[k], XI))
    return out


def f12pow_w(a, e):
    r = [(1, 0)] + [(0, 0)] * 5
    while e:
        if e & 1:
            r = f12mul_w(r, a)
        a = f12mul_w(a, a)
        e >>= 1
    return r


def limbs72_to_tower(limbs):
    """72 u64 (reference Montgomery form) -> 6 canonical Fp2 pairs in tower order"""
    ri = pow(R384, -1, P)
    fps = [sum(int(limbs[6 * k + i]) << (64 * i) for i in range(6)) * ri % P for k in range(12)]
    return [(fps[2 * k], fps[2 * k + 1]) for k in range(6)]


def tower_to_limbs72(t):
    out = []
    for c in t:
        for x in c:
            m = x * R384 % P
            out += [(m >> (64 * i)) & ((1 << 64) - 1) for i in range(6)]
    return out
