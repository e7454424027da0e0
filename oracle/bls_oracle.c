/*
 * CPU oracle for batch BLS12-381 signature verification -- TEST INFRASTRUCTURE.
 *
 * Loaded only by tests/, the smoke check and the benchmark's CPU-baseline leg,
 * as the checker.  The GPU product never links or calls this file.
 *
 * Restates, in plain C with 6 x 64-bit Montgomery limbs and unsigned __int128:
 *   Fp (sppark mont_t semantics: canonical [0,p) after every op)
 *       gpu-crypto/cuda/vendor/sppark/bls12-381.hpp:13-33, mont_t.cuh:221-347,402-422,632-668
 *   Fp2/Fp6/Fp12 tower, Frobenius, sparse 014 product, cyclotomic square
 *       gpu-crypto/cuda/field_sppark.cuh:40-197
 *   G2Prepared line coefficients (doubling/addition steps, eprint 2010/354 Alg. 26/27)
 *       cuda/test_miller_v2.cu:152-228, loop order :239-271
 *   two-pair Miller loop, cyclotomic exponentiation by x, final exponentiation,
 *   identity test  gpu-crypto/cuda/api.cu:22-91
 *   G1 decompression  ic-integration/gpu_crypto/cuda/decompress.cu:16-188, with the
 *       validity checks of ic_bls12_381's G1Affine::from_compressed added
 *   verify equation, H(m), keys  oracle/src/main.rs:376-420
 * Every result is a canonical field element, so values are bit-comparable with
 * any correct implementation of the same mathematical function.
 */
#include "bls_oracle.h"

#include <pthread.h>
#include <string.h>
#include <unistd.h>

typedef unsigned __int128 u128;
typedef uint64_t u64;

typedef struct { u64 l[6]; } fp;
typedef struct { fp c0, c1; } fp2;
typedef struct { fp2 c0, c1, c2; } fp6;
typedef struct { fp6 c0, c1; } fp12;

/* p, -p^-1 mod 2^64, R mod p, R^2 mod p: bls12-381.hpp:13-33 (32-bit limbs there) */
static const fp P = {{0xb9feffffffffaaabULL, 0x1eabfffeb153ffffULL, 0x6730d2a0f6b0f624ULL,
                      0x64774b84f38512bfULL, 0x4b1ba7b6434bacd7ULL, 0x1a0111ea397fe69aULL}};
static const u64 M0 = 0x89f3fffcfffcfffdULL;
static const fp ONE = {{0x760900000002fffdULL, 0xebf4000bc40c0002ULL, 0x5f48985753c758baULL,
                        0x77ce585370525745ULL, 0x5c071a97a256ec6dULL, 0x15f65ec3fa80e493ULL}};
static const fp R2 = {{0xf4df1f341c341746ULL, 0x0a76e6a609d104f1ULL, 0x8de5476c4c95b6d5ULL,
                       0x67eb88a9939d83c0ULL, 0x9a793e85b519952dULL, 0x11988fe592cae3aaULL}};
/* group order r, little-endian bytes */
static const u64 R_ORDER[4] = {0xffffffff00000001ULL, 0x53bda402fffe5bfeULL,
                               0x3339d80809a1d805ULL, 0x73eda753299d7d48ULL};
/* |x| of the BLS parameter, x = -0xd201000000010000 (api.cu:13) */
static const u64 BLS_X = 0xd201000000010000ULL;

/* ================= Fp ================= */
static int fp_is_zero(const fp* a) {
    u64 acc = 0;
    for (int i = 0; i < 6; i++) acc |= a->l[i];
    return acc == 0;
}
static int fp_eq(const fp* a, const fp* b) { return memcmp(a, b, sizeof(fp)) == 0; }

/* r = a - p if a >= p (a < 2p) */
static void fp_reduce_once(fp* r, const fp* a) {
    fp t; u64 borrow = 0;
    for (int i = 0; i < 6; i++) {
        u128 d = (u128)a->l[i] - P.l[i] - borrow;
        t.l[i] = (u64)d; borrow = (u64)(d >> 64) & 1;
    }
    *r = borrow ? *a : t;
}
static void fp_add(fp* r, const fp* a, const fp* b) {
    fp s; u64 c = 0;
    for (int i = 0; i < 6; i++) {
        u128 t = (u128)a->l[i] + b->l[i] + c;
        s.l[i] = (u64)t; c = (u64)(t >> 64);
    }
    fp_reduce_once(r, &s); /* p < 2^381 so no carry out of 384 bits */
}
static void fp_sub(fp* r, const fp* a, const fp* b) {
    fp d; u64 borrow = 0;
    for (int i = 0; i < 6; i++) {
        u128 t = (u128)a->l[i] - b->l[i] - borrow;
        d.l[i] = (u64)t; borrow = (u64)(t >> 64) & 1;
    }
    if (borrow) {
        u64 c = 0;
        for (int i = 0; i < 6; i++) {
            u128 t = (u128)d.l[i] + P.l[i] + c;
            d.l[i] = (u64)t; c = (u64)(t >> 64);
        }
    }
    *r = d;
}
/* zero stays zero (mont_t.cuh:299-347) */
static void fp_neg(fp* r, const fp* a) {
    if (fp_is_zero(a)) { *r = *a; return; }
    fp z = {{0}};
    fp_sub(r, &z, a);
}
/* CIOS Montgomery product a*b*R^-1 mod p */
static void fp_mul(fp* r, const fp* a, const fp* b) {
    u64 t[8] = {0};
    for (int i = 0; i < 6; i++) {
        u64 c = 0;
        for (int j = 0; j < 6; j++) {
            u128 s = (u128)a->l[j] * b->l[i] + t[j] + c;
            t[j] = (u64)s; c = (u64)(s >> 64);
        }
        u128 s = (u128)t[6] + c;
        t[6] = (u64)s; t[7] = (u64)(s >> 64);
        u64 m = t[0] * M0;
        s = (u128)m * P.l[0] + t[0];
        c = (u64)(s >> 64);
        for (int j = 1; j < 6; j++) {
            s = (u128)m * P.l[j] + t[j] + c;
            t[j - 1] = (u64)s; c = (u64)(s >> 64);
        }
        s = (u128)t[6] + c;
        t[5] = (u64)s;
        t[6] = t[7] + (u64)(s >> 64);
    }
    fp v; memcpy(v.l, t, 48);
    fp_reduce_once(r, &v);
}
static void fp_sqr(fp* r, const fp* a) { fp_mul(r, a, a); }
static void fp_pow(fp* r, const fp* a, const u64* e, int nlimbs) {
    fp acc = ONE, base = *a;
    for (int w = 0; w < nlimbs; w++)
        for (int b = 0; b < 64; b++) {
            if ((e[w] >> b) & 1) fp_mul(&acc, &acc, &base);
            fp_sqr(&base, &base);
        }
    *r = acc;
}
/* Fermat inverse a^(p-2): field_sppark.cuh:40-50 */
static void fp_inv(fp* r, const fp* a) {
    static const u64 PM2[6] = {0xb9feffffffffaaa9ULL, 0x1eabfffeb153ffffULL, 0x6730d2a0f6b0f624ULL,
                               0x64774b84f38512bfULL, 0x4b1ba7b6434bacd7ULL, 0x1a0111ea397fe69aULL};
    fp_pow(r, a, PM2, 6);
}
/* a^((p+1)/4): decompress.cu:16-40; returns 1 when the result squares back to a */
static int fp_sqrt(fp* r, const fp* a) {
    static const u64 E[6] = {0xee7fbfffffffeaabULL, 0x07aaffffac54ffffULL, 0xd9cc34a83dac3d89ULL,
                             0xd91dd2e13ce144afULL, 0x92c6e9ed90d2eb35ULL, 0x0680447a8e5ff9a6ULL};
    fp y, y2;
    fp_pow(&y, a, E, 6);
    fp_sqr(&y2, &y);
    *r = y;
    return fp_eq(&y2, a);
}
static void fp_from_mont(fp* r, const fp* a) {
    fp one = {{1, 0, 0, 0, 0, 0}};
    fp_mul(r, a, &one);
}
static void fp_to_mont(fp* r, const fp* a) { fp_mul(r, a, &R2); }
/* canonical value compared with (p-1)/2: decompress.cu:107-135 */
static int fp_lex_largest(const fp* a) {
    static const u64 HALF[6] = {0xdcff7fffffffd555ULL, 0x0f55ffff58a9ffffULL, 0xb39869507b587b12ULL,
                                0xb23ba5c279c2895fULL, 0x258dd3db21a5d66bULL, 0x0d0088f51cbff34dULL};
    fp n; fp_from_mont(&n, a);
    for (int i = 5; i >= 0; i--) {
        if (n.l[i] > HALF[i]) return 1;
        if (n.l[i] < HALF[i]) return 0;
    }
    return 0;
}
static int fp_lt_p(const fp* a) {
    for (int i = 5; i >= 0; i--) {
        if (a->l[i] < P.l[i]) return 1;
        if (a->l[i] > P.l[i]) return 0;
    }
    return 0;
}

/* ================= Fp2 = Fp[u]/(u^2+1) ================= */
static fp2 fp2_one(void) { fp2 r = {ONE, {{0}}}; return r; }
static int fp2_is_zero(const fp2* a) { return fp_is_zero(&a->c0) && fp_is_zero(&a->c1); }
static void fp2_add(fp2* r, const fp2* a, const fp2* b) { fp_add(&r->c0, &a->c0, &b->c0); fp_add(&r->c1, &a->c1, &b->c1); }
static void fp2_sub(fp2* r, const fp2* a, const fp2* b) { fp_sub(&r->c0, &a->c0, &b->c0); fp_sub(&r->c1, &a->c1, &b->c1); }
static void fp2_neg(fp2* r, const fp2* a) { fp_neg(&r->c0, &a->c0); fp_neg(&r->c1, &a->c1); }
static void fp2_conj(fp2* r, const fp2* a) { r->c0 = a->c0; fp_neg(&r->c1, &a->c1); }
/* times the non-residue 1+u */
static void fp2_mul_nr(fp2* r, const fp2* a) {
    fp t0, t1;
    fp_sub(&t0, &a->c0, &a->c1);
    fp_add(&t1, &a->c0, &a->c1);
    r->c0 = t0; r->c1 = t1;
}
static void fp2_mul_fp(fp2* r, const fp2* a, const fp* s) { fp_mul(&r->c0, &a->c0, s); fp_mul(&r->c1, &a->c1, s); }
/* Karatsuba, field_sppark.cuh:66-69 */
static void fp2_mul(fp2* r, const fp2* a, const fp2* b) {
    fp t0, t1, sa, sb, m;
    fp_mul(&t0, &a->c0, &b->c0);
    fp_mul(&t1, &a->c1, &b->c1);
    fp_add(&sa, &a->c0, &a->c1);
    fp_add(&sb, &b->c0, &b->c1);
    fp_mul(&m, &sa, &sb);
    fp_sub(&r->c0, &t0, &t1);
    fp_sub(&m, &m, &t0);
    fp_sub(&r->c1, &m, &t1);
}
/* (a0+a1)(a0-a1) + 2 a0 a1 u, field_sppark.cuh:70-79 */
static void fp2_sqr(fp2* r, const fp2* a) {
    fp t, s, d;
    fp_mul(&t, &a->c0, &a->c1);
    fp_add(&s, &a->c0, &a->c1);
    fp_sub(&d, &a->c0, &a->c1);
    fp_mul(&r->c0, &s, &d);
    fp_add(&r->c1, &t, &t);
}
static void fp2_inv(fp2* r, const fp2* a) {
    fp n, t, i;
    fp_sqr(&n, &a->c0);
    fp_sqr(&t, &a->c1);
    fp_add(&n, &n, &t);
    fp_inv(&i, &n);
    fp_mul(&r->c0, &a->c0, &i);
    fp_mul(&t, &a->c1, &i);
    fp_neg(&r->c1, &t);
}
static int fp2_eq(const fp2* a, const fp2* b) { return fp_eq(&a->c0, &b->c0) && fp_eq(&a->c1, &b->c1); }

/* ================= Fp6 = Fp2[v]/(v^3-(1+u)) ================= */
static void fp6_add(fp6* r, const fp6* a, const fp6* b) { fp2_add(&r->c0, &a->c0, &b->c0); fp2_add(&r->c1, &a->c1, &b->c1); fp2_add(&r->c2, &a->c2, &b->c2); }
static void fp6_sub(fp6* r, const fp6* a, const fp6* b) { fp2_sub(&r->c0, &a->c0, &b->c0); fp2_sub(&r->c1, &a->c1, &b->c1); fp2_sub(&r->c2, &a->c2, &b->c2); }
static void fp6_neg(fp6* r, const fp6* a) { fp2_neg(&r->c0, &a->c0); fp2_neg(&r->c1, &a->c1); fp2_neg(&r->c2, &a->c2); }
static void fp6_mul_by_v(fp6* r, const fp6* a) {
    fp2 t; fp2_mul_nr(&t, &a->c2);
    fp2 c0 = a->c0, c1 = a->c1;
    r->c0 = t; r->c1 = c0; r->c2 = c1;
}
/* field_sppark.cuh:93-100 */
static void fp6_mul(fp6* r, const fp6* a, const fp6* b) {
    fp2 aa, bb, cc, s1, s2, t, u;
    fp6 o;
    fp2_mul(&aa, &a->c0, &b->c0);
    fp2_mul(&bb, &a->c1, &b->c1);
    fp2_mul(&cc, &a->c2, &b->c2);
    fp2_add(&s1, &a->c1, &a->c2); fp2_add(&s2, &b->c1, &b->c2);
    fp2_mul(&t, &s1, &s2); fp2_sub(&t, &t, &bb); fp2_sub(&t, &t, &cc); fp2_mul_nr(&t, &t);
    fp2_add(&o.c0, &aa, &t);
    fp2_add(&s1, &a->c0, &a->c1); fp2_add(&s2, &b->c0, &b->c1);
    fp2_mul(&t, &s1, &s2); fp2_sub(&t, &t, &aa); fp2_sub(&t, &t, &bb);
    fp2_mul_nr(&u, &cc); fp2_add(&o.c1, &t, &u);
    fp2_add(&s1, &a->c0, &a->c2); fp2_add(&s2, &b->c0, &b->c2);
    fp2_mul(&t, &s1, &s2); fp2_sub(&t, &t, &aa); fp2_sub(&t, &t, &cc);
    fp2_add(&o.c2, &t, &bb);
    *r = o;
}
/* field_sppark.cuh:102-110 */
static void fp6_inv(fp6* r, const fp6* f) {
    fp2 c0s, c1s, c2s, c01, c02, c12, t0, t1, t2, sc, si, x, y;
    fp2_sqr(&c0s, &f->c0); fp2_sqr(&c1s, &f->c1); fp2_sqr(&c2s, &f->c2);
    fp2_mul(&c01, &f->c0, &f->c1); fp2_mul(&c02, &f->c0, &f->c2); fp2_mul(&c12, &f->c1, &f->c2);
    fp2_mul_nr(&x, &c12); fp2_sub(&t0, &c0s, &x);
    fp2_mul_nr(&x, &c2s); fp2_sub(&t1, &x, &c01);
    fp2_sub(&t2, &c1s, &c02);
    fp2_mul(&x, &f->c2, &t1); fp2_mul(&y, &f->c1, &t2); fp2_add(&x, &x, &y); fp2_mul_nr(&x, &x);
    fp2_mul(&y, &f->c0, &t0); fp2_add(&sc, &y, &x);
    fp2_inv(&si, &sc);
    fp2_mul(&r->c0, &t0, &si); fp2_mul(&r->c1, &t1, &si); fp2_mul(&r->c2, &t2, &si);
}
/* Frobenius constants, field_sppark.cuh:117-119 (Fp6) and :162-163 (Fp12) */
static const fp2 FROB6_C1 = {{{0}}, {{0xcd03c9e48671f071ULL, 0x5dab22461fcda5d2ULL, 0x587042afd3851b95ULL,
                                      0x8eb60ebe01bacb9eULL, 0x03f97d6e83d050d2ULL, 0x18f0206554638741ULL}}};
static const fp2 FROB6_C2 = {{{0x890dc9e4867545c3ULL, 0x2af322533285a5d5ULL, 0x50880866309b7e2cULL,
                               0xa20d1b8c7e881024ULL, 0x14e4f04fe2db9068ULL, 0x14e56d3f1564853aULL}}, {{0}}};
static const fp2 FROB12 = {{{0x07089552b319d465ULL, 0xc6695f92b50a8313ULL, 0x97e83cccd117228fULL,
                             0xa35baecab2dc29eeULL, 0x1ce393ea5daace4dULL, 0x08f2220fb0fb66ebULL}},
                           {{0xb2f66aad4ce5d646ULL, 0x5842a06bfc497cecULL, 0xcf4895d42599d394ULL,
                             0xc11b9cba40a8e8d0ULL, 0x2e3813cbe5a0de89ULL, 0x110eefda88847fafULL}}};
static void fp6_frob(fp6* r, const fp6* f) {
    fp2 c0, c1, c2;
    fp2_conj(&c0, &f->c0); fp2_conj(&c1, &f->c1); fp2_conj(&c2, &f->c2);
    fp2_mul(&c1, &c1, &FROB6_C1);
    fp2_mul(&c2, &c2, &FROB6_C2);
    r->c0 = c0; r->c1 = c1; r->c2 = c2;
}
/* sparse products, field_sppark.cuh:123-134 */
static void fp6_mul_by_01(fp6* r, const fp6* s, const fp2* c0, const fp2* c1) {
    fp2 aa, bb, t, x, y;
    fp6 o;
    fp2_mul(&aa, &s->c0, c0);
    fp2_mul(&bb, &s->c1, c1);
    fp2_mul(&t, &s->c2, c1); fp2_mul_nr(&t, &t); fp2_add(&o.c0, &t, &aa);
    fp2_add(&x, c0, c1); fp2_add(&y, &s->c0, &s->c1); fp2_mul(&t, &x, &y);
    fp2_sub(&t, &t, &aa); fp2_sub(&o.c1, &t, &bb);
    fp2_mul(&t, &s->c2, c0); fp2_add(&o.c2, &t, &bb);
    *r = o;
}
static void fp6_mul_by_1(fp6* r, const fp6* s, const fp2* c1) {
    fp6 o;
    fp2_mul(&o.c0, &s->c2, c1); fp2_mul_nr(&o.c0, &o.c0);
    fp2_mul(&o.c1, &s->c0, c1);
    fp2_mul(&o.c2, &s->c1, c1);
    *r = o;
}

/* ================= Fp12 = Fp6[w]/(w^2-v) ================= */
static fp12 fp12_one(void) { fp12 r; memset(&r, 0, sizeof r); r.c0.c0.c0 = ONE; return r; }
static void fp12_conj(fp12* r, const fp12* a) { r->c0 = a->c0; fp6_neg(&r->c1, &a->c1); }
/* field_sppark.cuh:140-145 */
static void fp12_mul(fp12* r, const fp12* a, const fp12* b) {
    fp6 aa, bb, s1, s2, t, vb;
    fp12 o;
    fp6_mul(&aa, &a->c0, &b->c0);
    fp6_mul(&bb, &a->c1, &b->c1);
    fp6_mul_by_v(&vb, &bb); fp6_add(&o.c0, &aa, &vb);
    fp6_add(&s1, &a->c0, &a->c1); fp6_add(&s2, &b->c0, &b->c1);
    fp6_mul(&t, &s1, &s2); fp6_sub(&t, &t, &aa); fp6_sub(&o.c1, &t, &bb);
    *r = o;
}
/* field_sppark.cuh:147-150 */
static void fp12_sqr(fp12* r, const fp12* a) {
    fp6 ab, s1, s2, t, v;
    fp12 o;
    fp6_mul(&ab, &a->c0, &a->c1);
    fp6_add(&s1, &a->c0, &a->c1);
    fp6_mul_by_v(&v, &a->c1); fp6_add(&s2, &a->c0, &v);
    fp6_mul(&t, &s2, &s1); fp6_sub(&t, &t, &ab); fp6_mul_by_v(&v, &ab); fp6_sub(&o.c0, &t, &v);
    fp6_add(&o.c1, &ab, &ab);
    *r = o;
}
/* field_sppark.cuh:152-156 */
static void fp12_inv(fp12* r, const fp12* f) {
    fp6 a, b, t, ti;
    fp6_mul(&a, &f->c0, &f->c0);
    fp6_mul(&b, &f->c1, &f->c1);
    fp6_mul_by_v(&b, &b);
    fp6_sub(&t, &a, &b);
    fp6_inv(&ti, &t);
    fp12 o;
    fp6_mul(&o.c0, &f->c0, &ti);
    fp6_mul(&t, &f->c1, &ti); fp6_neg(&o.c1, &t);
    *r = o;
}
/* field_sppark.cuh:158-166 */
static void fp12_frob(fp12* r, const fp12* f) {
    fp6 c0, c1;
    fp6_frob(&c0, &f->c0); fp6_frob(&c1, &f->c1);
    fp2_mul(&c1.c0, &c1.c0, &FROB12); fp2_mul(&c1.c1, &c1.c1, &FROB12); fp2_mul(&c1.c2, &c1.c2, &FROB12);
    r->c0 = c0; r->c1 = c1;
}
/* field_sppark.cuh:168-172 */
static void fp12_mul_by_014(fp12* r, const fp12* f, const fp2* c0, const fp2* c1, const fp2* c4) {
    fp6 aa, bb, s, t, vb;
    fp2 o;
    fp6_mul_by_01(&aa, &f->c0, c0, c1);
    fp6_mul_by_1(&bb, &f->c1, c4);
    fp2_add(&o, c1, c4);
    fp6_add(&s, &f->c1, &f->c0);
    fp6_mul_by_01(&t, &s, c0, &o);
    fp6_sub(&t, &t, &aa);
    fp12 out;
    fp6_sub(&out.c1, &t, &bb);
    fp6_mul_by_v(&vb, &bb);
    fp6_add(&out.c0, &vb, &aa);
    *r = out;
}
/* Granger-Scott squaring in the cyclotomic subgroup, field_sppark.cuh:174-197 */
static void fp4_sqr(fp2* o0, fp2* o1, const fp2* a, const fp2* b) {
    fp2 t0, t1, ab, x;
    fp2_sqr(&t0, a); fp2_sqr(&t1, b);
    fp2_mul_nr(&x, &t1); fp2_add(o0, &x, &t0);
    fp2_add(&ab, a, b); fp2_sqr(&x, &ab); fp2_sub(&x, &x, &t0); fp2_sub(o1, &x, &t1);
}
static void fp12_cyc_sqr(fp12* r, const fp12* f) {
    fp2 z0 = f->c0.c0, z4 = f->c0.c1, z3 = f->c0.c2, z2 = f->c1.c0, z1 = f->c1.c1, z5 = f->c1.c2;
    fp2 A0, A1, B0, B1, C0, C1, t;
    fp4_sqr(&A0, &A1, &z0, &z1);
    fp2_sub(&z0, &A0, &z0); fp2_add(&z0, &z0, &z0); fp2_add(&z0, &z0, &A0);
    fp2_add(&z1, &A1, &z1); fp2_add(&z1, &z1, &z1); fp2_add(&z1, &z1, &A1);
    fp4_sqr(&B0, &B1, &z2, &z3);
    fp4_sqr(&C0, &C1, &z4, &z5);
    fp2_sub(&z4, &B0, &z4); fp2_add(&z4, &z4, &z4); fp2_add(&z4, &z4, &B0);
    fp2_add(&z5, &B1, &z5); fp2_add(&z5, &z5, &z5); fp2_add(&z5, &z5, &B1);
    fp2_mul_nr(&t, &C1);
    fp2_add(&z2, &t, &z2); fp2_add(&z2, &z2, &z2); fp2_add(&z2, &z2, &t);
    fp2_sub(&z3, &C0, &z3); fp2_add(&z3, &z3, &z3); fp2_add(&z3, &z3, &C0);
    r->c0.c0 = z0; r->c0.c1 = z4; r->c0.c2 = z3;
    r->c1.c0 = z2; r->c1.c1 = z1; r->c1.c2 = z5;
}
static int fp12_is_one(const fp12* f) {
    fp12 one = fp12_one();
    return memcmp(f, &one, sizeof one) == 0;
}

/* ================= pairing ================= */
typedef struct { fp x, y; } g1a;
typedef struct { fp2 x, y; } g2a;

/* G2Prepared generation, cuda/test_miller_v2.cu:152-228 (ic_bls12_381 G2Prepared::from) */
typedef struct { fp2 x, y, z; } g2proj;
static void doubling_step(fp2 out[3], g2proj* r) {
    fp2 tmp0, tmp1, tmp2, tmp3, tmp4, tmp5, tmp6, zsq, t, u;
    fp2_sqr(&tmp0, &r->x);
    fp2_sqr(&tmp1, &r->y);
    fp2_sqr(&tmp2, &tmp1);
    fp2_add(&t, &tmp1, &r->x); fp2_sqr(&t, &t); fp2_add(&u, &tmp0, &tmp2); fp2_sub(&tmp3, &t, &u);
    fp2_add(&tmp3, &tmp3, &tmp3);
    fp2_add(&tmp4, &tmp0, &tmp0); fp2_add(&tmp4, &tmp4, &tmp0);
    fp2_add(&tmp6, &r->x, &tmp4);
    fp2_sqr(&tmp5, &tmp4);
    fp2_sqr(&zsq, &r->z);
    fp2_sub(&r->x, &tmp5, &tmp3); fp2_sub(&r->x, &r->x, &tmp3);
    fp2_add(&t, &r->z, &r->y); fp2_sqr(&t, &t); fp2_add(&u, &tmp1, &zsq); fp2_sub(&r->z, &t, &u);
    fp2_sub(&t, &tmp3, &r->x); fp2_mul(&r->y, &t, &tmp4);
    fp2_add(&tmp2, &tmp2, &tmp2); fp2_add(&tmp2, &tmp2, &tmp2); fp2_add(&tmp2, &tmp2, &tmp2);
    fp2_sub(&r->y, &r->y, &tmp2);
    fp2_mul(&tmp3, &tmp4, &zsq); fp2_add(&tmp3, &tmp3, &tmp3); fp2_neg(&tmp3, &tmp3);
    fp2_sqr(&t, &tmp6); fp2_add(&u, &tmp0, &tmp5); fp2_sub(&tmp6, &t, &u);
    fp2_add(&tmp1, &tmp1, &tmp1); fp2_add(&tmp1, &tmp1, &tmp1);
    fp2_sub(&tmp6, &tmp6, &tmp1);
    fp2_mul(&t, &r->z, &zsq); fp2_add(&t, &t, &t);
    out[0] = t; out[1] = tmp3; out[2] = tmp6;
}
static void addition_step(fp2 out[3], g2proj* r, const g2a* q) {
    fp2 zsq, ysq, t0, t1, t2, t3, t4, t5, t6, t7, t8, t9, t10, a, b;
    fp2_sqr(&zsq, &r->z);
    fp2_sqr(&ysq, &q->y);
    fp2_mul(&t0, &zsq, &q->x);
    fp2_add(&a, &q->y, &r->z); fp2_sqr(&a, &a); fp2_add(&b, &ysq, &zsq); fp2_sub(&a, &a, &b);
    fp2_mul(&t1, &a, &zsq);
    fp2_sub(&t2, &t0, &r->x);
    fp2_sqr(&t3, &t2);
    fp2_add(&t4, &t3, &t3); fp2_add(&t4, &t4, &t4);
    fp2_mul(&t5, &t4, &t2);
    fp2_sub(&t6, &t1, &r->y); fp2_sub(&t6, &t6, &r->y);
    fp2_mul(&t9, &t6, &q->x);
    fp2_mul(&t7, &t4, &r->x);
    fp2_sqr(&a, &t6); fp2_sub(&a, &a, &t5); fp2_sub(&a, &a, &t7); fp2_sub(&r->x, &a, &t7);
    fp2_add(&a, &r->z, &t2); fp2_sqr(&a, &a); fp2_add(&b, &zsq, &t3); fp2_sub(&r->z, &a, &b);
    fp2_add(&t10, &q->y, &r->z);
    fp2_sub(&a, &t7, &r->x); fp2_mul(&t8, &a, &t6);
    fp2_mul(&a, &r->y, &t5); fp2_add(&a, &a, &a);
    fp2_sub(&r->y, &t8, &a);
    fp2_sqr(&t10, &t10); fp2_sub(&t10, &t10, &ysq);
    fp2_sqr(&a, &r->z); fp2_sub(&t10, &t10, &a);
    fp2_add(&t9, &t9, &t9); fp2_sub(&t9, &t9, &t10);
    fp2_add(&a, &r->z, &r->z);
    fp2_neg(&t6, &t6); fp2_add(&b, &t6, &t6);
    out[0] = a; out[1] = b; out[2] = t9;
}
static void g2_prepare(fp2 (*coeffs)[3], const g2a* q) {
    g2proj r = {q->x, q->y, fp2_one()};
    int n = 0, found = 0;
    for (int b = 63; b >= 0; b--) {
        int bit = (int)(((BLS_X >> 1) >> b) & 1);
        if (!found) { found = bit; continue; }
        doubling_step(coeffs[n++], &r);
        if (bit) addition_step(coeffs[n++], &r, q);
    }
    doubling_step(coeffs[n++], &r);
}
/* ell(): api.cu:37 -- mul_by_014(c2, c1 * P.x, c0 * P.y) */
static void ell(fp12* f, const fp2 c[3], const g1a* p) {
    fp2 a, b;
    fp2_mul_fp(&a, &c[1], &p->x);
    fp2_mul_fp(&b, &c[0], &p->y);
    fp12_mul_by_014(f, f, &c[2], &a, &b);
}
/* api.cu:30-52 */
static void miller2(fp12* out, const g1a* p1, const fp2 (*q1)[3], const g1a* p2, const fp2 (*q2)[3]) {
    fp12 f = fp12_one();
    int ci = 0, found = 0;
    for (int b = 63; b >= 0; b--) {
        int bit = (int)(((BLS_X >> 1) >> b) & 1);
        if (!found) { found = bit; continue; }
        ell(&f, q1[ci], p1); ell(&f, q2[ci], p2); ci++;
        if (bit) { ell(&f, q1[ci], p1); ell(&f, q2[ci], p2); ci++; }
        fp12_sqr(&f, &f);
    }
    ell(&f, q1[ci], p1); ell(&f, q2[ci], p2);
    fp12_conj(out, &f);
}
/* f^x for x = -BLS_X, api.cu:54-63 */
static void cyc_exp(fp12* r, const fp12* f) {
    fp12 t = fp12_one();
    int found = 0;
    for (int i = 63; i >= 0; i--) {
        if (found) fp12_cyc_sqr(&t, &t);
        int bit = (int)((BLS_X >> i) & 1);
        if (!found) { found = bit; if (!bit) continue; }
        if (bit) fp12_mul(&t, &t, f);
    }
    fp12_conj(r, &t);
}
/* api.cu:65-78 */
static void final_exp(fp12* out, const fp12* fi) {
    fp12 f = *fi, t0, t1, t2, t3, t4, x;
    fp12_conj(&t0, &f);
    fp12_inv(&t1, &f);
    fp12_mul(&t2, &t0, &t1);
    t1 = t2;
    fp12_frob(&x, &t2); fp12_frob(&x, &x); fp12_mul(&t2, &x, &t1);
    f = t2;
    fp12_cyc_sqr(&x, &t2); fp12_conj(&t1, &x);
    cyc_exp(&t3, &t2);
    fp12_mul(&t4, &t1, &t3);
    cyc_exp(&t1, &t4);
    fp12_conj(&t4, &t4);
    fp12_mul(&f, &f, &t4);
    fp12_cyc_sqr(&t4, &t3);
    cyc_exp(&t0, &t1);
    fp12_mul(&t3, &t3, &t0);
    fp12_frob(&t3, &t3); fp12_frob(&t3, &t3);
    fp12_mul(&f, &f, &t3);
    cyc_exp(&x, &t0); fp12_mul(&t4, &t4, &x);
    cyc_exp(&x, &t4); fp12_mul(&f, &f, &x);
    fp12_conj(&x, &t2); fp12_mul(&t4, &t4, &x);
    fp12_mul(&t2, &t2, &t1);
    fp12_frob(&t2, &t2); fp12_frob(&t2, &t2); fp12_frob(&t2, &t2);
    fp12_mul(&f, &f, &t2);
    fp12_frob(&t4, &t4);
    fp12_mul(&f, &f, &t4);
    *out = f;
}
static int verify_one(const g1a* sig, const g1a* nhm, const fp2 (*g2c)[3], const fp2 (*pkc)[3], fp12* gt) {
    fp12 ml, r;
    miller2(&ml, sig, g2c, nhm, pkc);
    final_exp(&r, &ml);
    if (gt) *gt = r;
    return fp12_is_one(&r);
}

/* ================= curve arithmetic (Jacobian, a = 0) ================= */
typedef struct { fp x, y, z; } g1j;
typedef struct { fp2 x, y, z; } g2j;

#define DEFINE_JACOBIAN(PFX, F, J, A)                                                           \
    static void PFX##_dbl(J* r, const J* p) {                                                   \
        F a, b, c, d, e, f, t;                                                                  \
        if (F##_is_zero(&p->z)) { *r = *p; return; }                                            \
        F##_sqr(&a, &p->x); F##_sqr(&b, &p->y); F##_sqr(&c, &b);                                \
        F##_add(&t, &p->x, &b); F##_sqr(&t, &t); F##_sub(&t, &t, &a); F##_sub(&t, &t, &c);      \
        F##_add(&d, &t, &t);                                                                    \
        F##_add(&e, &a, &a); F##_add(&e, &e, &a);                                               \
        F##_sqr(&f, &e);                                                                        \
        J o;                                                                                    \
        F##_sub(&o.x, &f, &d); F##_sub(&o.x, &o.x, &d);                                         \
        F##_mul(&t, &p->y, &p->z); F##_add(&o.z, &t, &t);                                       \
        F##_sub(&t, &d, &o.x); F##_mul(&o.y, &e, &t);                                           \
        F##_add(&c, &c, &c); F##_add(&c, &c, &c); F##_add(&c, &c, &c);                          \
        F##_sub(&o.y, &o.y, &c);                                                                \
        *r = o;                                                                                 \
    }                                                                                           \
    static void PFX##_add(J* r, const J* p, const J* q) {                                       \
        if (F##_is_zero(&p->z)) { *r = *q; return; }                                            \
        if (F##_is_zero(&q->z)) { *r = *p; return; }                                            \
        F z1z1, z2z2, u1, u2, s1, s2, h, i, j, rr, v, t;                                        \
        F##_sqr(&z1z1, &p->z); F##_sqr(&z2z2, &q->z);                                           \
        F##_mul(&u1, &p->x, &z2z2); F##_mul(&u2, &q->x, &z1z1);                                 \
        F##_mul(&s1, &p->y, &q->z); F##_mul(&s1, &s1, &z2z2);                                   \
        F##_mul(&s2, &q->y, &p->z); F##_mul(&s2, &s2, &z1z1);                                   \
        F##_sub(&h, &u2, &u1); F##_sub(&rr, &s2, &s1);                                          \
        if (F##_is_zero(&h)) {                                                                  \
            if (F##_is_zero(&rr)) { PFX##_dbl(r, p); return; }                                  \
            memset(r, 0, sizeof *r); return;                                                    \
        }                                                                                       \
        F##_add(&rr, &rr, &rr);                                                                 \
        F##_add(&i, &h, &h); F##_sqr(&i, &i);                                                   \
        F##_mul(&j, &h, &i);                                                                    \
        F##_mul(&v, &u1, &i);                                                                   \
        J o;                                                                                    \
        F##_sqr(&o.x, &rr); F##_sub(&o.x, &o.x, &j); F##_sub(&o.x, &o.x, &v);                   \
        F##_sub(&o.x, &o.x, &v);                                                                \
        F##_sub(&t, &v, &o.x); F##_mul(&o.y, &rr, &t);                                          \
        F##_mul(&t, &s1, &j); F##_add(&t, &t, &t); F##_sub(&o.y, &o.y, &t);                     \
        F##_add(&t, &p->z, &q->z); F##_sqr(&t, &t); F##_sub(&t, &t, &z1z1);                     \
        F##_sub(&t, &t, &z2z2); F##_mul(&o.z, &t, &h);                                          \
        *r = o;                                                                                 \
    }                                                                                           \
    static void PFX##_from_affine(J* r, const A* a) { r->x = a->x; r->y = a->y; r->z = PFX##_fone(); } \
    static int PFX##_to_affine(A* r, const J* p) {                                              \
        if (F##_is_zero(&p->z)) { memset(r, 0, sizeof *r); return 1; }                         \
        F zi, zi2, zi3;                                                                         \
        F##_inv(&zi, &p->z); F##_sqr(&zi2, &zi); F##_mul(&zi3, &zi2, &zi);                      \
        F##_mul(&r->x, &p->x, &zi2); F##_mul(&r->y, &p->y, &zi3);                               \
        return 0;                                                                               \
    }                                                                                           \
    static void PFX##_mul(J* r, const J* p, const uint8_t k[32]) {                              \
        J acc; memset(&acc, 0, sizeof acc);                                                     \
        for (int i = 255; i >= 0; i--) {                                                        \
            PFX##_dbl(&acc, &acc);                                                              \
            if ((k[i >> 3] >> (i & 7)) & 1) PFX##_add(&acc, &acc, p);                           \
        }                                                                                       \
        *r = acc;                                                                               \
    }

static fp g1_fone(void) { return ONE; }
static fp2 g2_fone(void) { return fp2_one(); }
DEFINE_JACOBIAN(g1, fp, g1j, g1a)
DEFINE_JACOBIAN(g2, fp2, g2j, g2a)

static void r_bytes(uint8_t out[32]) {
    for (int i = 0; i < 32; i++) out[i] = (uint8_t)(R_ORDER[i / 8] >> (8 * (i % 8)));
}
static int g1a_on_curve(const g1a* p) {
    fp y2, x3, b;
    fp_sqr(&y2, &p->y);
    fp_sqr(&x3, &p->x); fp_mul(&x3, &x3, &p->x);
    fp_add(&b, &ONE, &ONE); fp_add(&b, &b, &b);
    fp_add(&x3, &x3, &b);
    return fp_eq(&y2, &x3);
}
static int g1a_in_subgroup(const g1a* p) {
    g1j j, o; uint8_t r[32];
    g1_from_affine(&j, p);
    r_bytes(r);
    g1_mul(&o, &j, r);
    return fp_is_zero(&o.z);
}
static int g2a_in_subgroup(const g2a* q) {
    g2j j, o; uint8_t r[32];
    g2_from_affine(&j, q);
    r_bytes(r);
    g2_mul(&o, &j, r);
    return fp2_is_zero(&o.z);
}

/* Generators in Montgomery limbs.  G1 as in cuda/test_miller_v2.cu:283-288.  G2 is the
 * standard generator; the copy at cuda/test_miller_v2.cu:291-302 carries a mis-transcribed
 * y.c0 (that point is not on the curve), so y.c0 here is the standard value
 * 0x0ce5d527...08b82801 times R.  Pinned by reproducing cuda/g2_coeffs.h. */
static const g1a G1_GEN = {
    {{0x5cb38790fd530c16ULL, 0x7817fc679976fff5ULL, 0x154f95c7143ba1c1ULL, 0xf0ae6acdf3d0e747ULL,
      0xedce6ecc21dbf440ULL, 0x120177419e0bfb75ULL}},
    {{0xbaac93d50ce72271ULL, 0x8c22631a7918fd8eULL, 0xdd595f13570725ceULL, 0x51ac582950405194ULL,
      0x0e1c8c3fad0059c0ULL, 0x0bbc3efc5008a26aULL}}};
static const g2a G2_GEN = {
    {{{0xf5f28fa202940a10ULL, 0xb3f5fb2687b4961aULL, 0xa1a893b53e2ae580ULL, 0x9894999d1a3caee9ULL,
       0x6f67b7631863366bULL, 0x058191924350bcd7ULL}},
     {{0xa5a9c0759e23f606ULL, 0xaaa0c59dbccd60c3ULL, 0x3bb17e18e2867806ULL, 0x1b1ab6cc8541b367ULL,
       0xc2b6ed0ef2158547ULL, 0x11922a097360edf3ULL}}},
    {{{0x4c730af860494c4aULL, 0x597cfa1f5e369c5aULL, 0xe7e6856caa0a635aULL, 0xbbefb5e96e0d495fULL,
       0x07d3a975f0ef25a2ULL, 0x0083fd8e7e80dae5ULL}},
     {{0xadc0fc92df64b05dULL, 0x18aa270a2b1461dcULL, 0x86adac6a3be4eba0ULL, 0x79495c4ec93da33aULL,
       0xe7175850a43ccaedULL, 0x0b2bc2a163de1bf2ULL}}}};

/* ================= SHA-256 (FIPS 180-4) ================= */
static const uint32_t K256[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
#define ROR(x, n) (((x) >> (n)) | ((x) << (32 - (n))))
static void sha_block(uint32_t h[8], const uint8_t* blk) {
    uint32_t w[64];
    for (int i = 0; i < 16; i++)
        w[i] = (uint32_t)blk[4 * i] << 24 | (uint32_t)blk[4 * i + 1] << 16 | (uint32_t)blk[4 * i + 2] << 8 | blk[4 * i + 3];
    for (int i = 16; i < 64; i++) {
        uint32_t s0 = ROR(w[i - 15], 7) ^ ROR(w[i - 15], 18) ^ (w[i - 15] >> 3);
        uint32_t s1 = ROR(w[i - 2], 17) ^ ROR(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
    for (int i = 0; i < 64; i++) {
        uint32_t S1 = ROR(e, 6) ^ ROR(e, 11) ^ ROR(e, 25);
        uint32_t ch = (e & f) ^ (~e & g);
        uint32_t t1 = hh + S1 + ch + K256[i] + w[i];
        uint32_t S0 = ROR(a, 2) ^ ROR(a, 13) ^ ROR(a, 22);
        uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
        uint32_t t2 = S0 + mj;
        hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}
void orc_sha256(const uint8_t* msg, size_t len, uint8_t out[32]) {
    uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    size_t i = 0;
    for (; i + 64 <= len; i += 64) sha_block(h, msg + i);
    uint8_t tail[128];
    size_t rem = len - i;
    memset(tail, 0, sizeof tail);
    memcpy(tail, msg + i, rem);
    tail[rem] = 0x80;
    size_t tl = (rem + 1 + 8 <= 64) ? 64 : 128;
    uint64_t bits = (uint64_t)len * 8;
    for (int k = 0; k < 8; k++) tail[tl - 1 - k] = (uint8_t)(bits >> (8 * k));
    sha_block(h, tail);
    if (tl == 128) sha_block(h, tail + 64);
    for (int k = 0; k < 8; k++) {
        out[4 * k] = (uint8_t)(h[k] >> 24); out[4 * k + 1] = (uint8_t)(h[k] >> 16);
        out[4 * k + 2] = (uint8_t)(h[k] >> 8); out[4 * k + 3] = (uint8_t)h[k];
    }
}

/* ================= exported wrappers ================= */
#define LD(T, v, src) T v; memcpy(&v, (src), sizeof(T))
#define ST(dst, v) memcpy((dst), &(v), sizeof(v))

void orc_fp_mul(const u64 a[6], const u64 b[6], u64 out[6]) { LD(fp, x, a); LD(fp, y, b); fp r; fp_mul(&r, &x, &y); ST(out, r); }
void orc_fp_add(const u64 a[6], const u64 b[6], u64 out[6]) { LD(fp, x, a); LD(fp, y, b); fp r; fp_add(&r, &x, &y); ST(out, r); }
void orc_fp_sub(const u64 a[6], const u64 b[6], u64 out[6]) { LD(fp, x, a); LD(fp, y, b); fp r; fp_sub(&r, &x, &y); ST(out, r); }
void orc_fp_neg(const u64 a[6], u64 out[6]) { LD(fp, x, a); fp r; fp_neg(&r, &x); ST(out, r); }
void orc_fp_inv(const u64 a[6], u64 out[6]) { LD(fp, x, a); fp r; fp_inv(&r, &x); ST(out, r); }
int orc_fp_sqrt(const u64 a[6], u64 out[6]) { LD(fp, x, a); fp r; int ok = fp_sqrt(&r, &x); ST(out, r); return ok; }
void orc_fp_to_mont(const u64 a[6], u64 out[6]) { LD(fp, x, a); fp r; fp_to_mont(&r, &x); ST(out, r); }
void orc_fp_from_mont(const u64 a[6], u64 out[6]) { LD(fp, x, a); fp r; fp_from_mont(&r, &x); ST(out, r); }
void orc_fp2_mul(const u64 a[12], const u64 b[12], u64 out[12]) { LD(fp2, x, a); LD(fp2, y, b); fp2 r; fp2_mul(&r, &x, &y); ST(out, r); }
void orc_fp2_sqr(const u64 a[12], u64 out[12]) { LD(fp2, x, a); fp2 r; fp2_sqr(&r, &x); ST(out, r); }
void orc_fp2_inv(const u64 a[12], u64 out[12]) { LD(fp2, x, a); fp2 r; fp2_inv(&r, &x); ST(out, r); }
void orc_fp6_mul(const u64 a[36], const u64 b[36], u64 out[36]) { LD(fp6, x, a); LD(fp6, y, b); fp6 r; fp6_mul(&r, &x, &y); ST(out, r); }
void orc_fp12_mul(const u64 a[72], const u64 b[72], u64 out[72]) { LD(fp12, x, a); LD(fp12, y, b); fp12 r; fp12_mul(&r, &x, &y); ST(out, r); }
void orc_fp12_sqr(const u64 a[72], u64 out[72]) { LD(fp12, x, a); fp12 r; fp12_sqr(&r, &x); ST(out, r); }
void orc_fp12_cyc_sqr(const u64 a[72], u64 out[72]) { LD(fp12, x, a); fp12 r; fp12_cyc_sqr(&r, &x); ST(out, r); }
void orc_fp12_inv(const u64 a[72], u64 out[72]) { LD(fp12, x, a); fp12 r; fp12_inv(&r, &x); ST(out, r); }
void orc_fp12_frob(const u64 a[72], u64 out[72]) { LD(fp12, x, a); fp12 r; fp12_frob(&r, &x); ST(out, r); }
void orc_fp12_conj(const u64 a[72], u64 out[72]) { LD(fp12, x, a); fp12 r; fp12_conj(&r, &x); ST(out, r); }
void orc_fp12_mul_by_014(const u64 f[72], const u64 c0[12], const u64 c1[12], const u64 c4[12], u64 out[72]) {
    LD(fp12, x, f); LD(fp2, a, c0); LD(fp2, b, c1); LD(fp2, c, c4); fp12 r;
    fp12_mul_by_014(&r, &x, &a, &b, &c); ST(out, r);
}
void orc_g2_prepare(const u64 q[24], u64 coeffs[2448]) {
    LD(g2a, Q, q);
    fp2 c[68][3];
    g2_prepare(c, &Q);
    memcpy(coeffs, c, sizeof c);
}
void orc_miller2(const u64 p1[12], const u64 q1c[2448], const u64 p2[12], const u64 q2c[2448], u64 out[72]) {
    LD(g1a, a, p1); LD(g1a, b, p2);
    fp12 r;
    miller2(&r, &a, (const fp2(*)[3])q1c, &b, (const fp2(*)[3])q2c);
    ST(out, r);
}
/* one-pair Miller loop over prepared lines (the two-pair loop with a single term) */
void orc_miller1(const u64 p1[12], const u64 q1c[2448], u64 out[72]) {
    LD(g1a, a, p1);
    const fp2(*q)[3] = (const fp2(*)[3])q1c;
    fp12 f = fp12_one();
    int ci = 0, found = 0;
    for (int b = 63; b >= 0; b--) {
        int bit = (int)(((BLS_X >> 1) >> b) & 1);
        if (!found) { found = bit; continue; }
        ell(&f, q[ci++], &a);
        if (bit) ell(&f, q[ci++], &a);
        fp12_sqr(&f, &f);
    }
    ell(&f, q[ci], &a);
    fp12 r; fp12_conj(&r, &f);
    ST(out, r);
}
/* e(P, Q) for affine P (12 u64) and affine Q (24 u64) */
void orc_pairing(const u64 p[12], const u64 q[24], u64 out[72]) {
    u64 c[2448], ml[72];
    orc_g2_prepare(q, c);
    orc_miller1(p, c, ml);
    orc_final_exp(ml, out);
}
void orc_final_exp(const u64 f[72], u64 out[72]) { LD(fp12, x, f); fp12 r; final_exp(&r, &x); ST(out, r); }
int orc_verify(const u64 sig[12], const u64 neg_hm[12], const u64 g2c[2448], const u64 pkc[2448], u64 gt_out[72]) {
    LD(g1a, s, sig); LD(g1a, n, neg_hm);
    fp12 gt;
    int ok = verify_one(&s, &n, (const fp2(*)[3])g2c, (const fp2(*)[3])pkc, &gt);
    if (gt_out) ST(gt_out, gt);
    return ok;
}

int orc_num_threads(void) {
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}
typedef struct {
    const u64 *sigs, *nhms, *g2c, *pkc, *key_coeffs;
    const int32_t* key_idx;
    int32_t* res;
    u64* gts;
    int m, stride, tid;
} job_t;
static void* verify_worker(void* arg) {
    job_t* j = (job_t*)arg;
    for (int i = j->tid; i < j->m; i += j->stride) {
        const u64* pk = j->key_idx ? j->key_coeffs + (size_t)j->key_idx[i] * 2448 : j->pkc + (size_t)i * 2448;
        j->res[i] = orc_verify(j->sigs + (size_t)i * 12, j->nhms + (size_t)i * 12, j->g2c, pk,
                               j->gts ? j->gts + (size_t)i * 72 : NULL);
    }
    return NULL;
}
static int run_batch(job_t base, int threads) {
    if (threads <= 0) threads = orc_num_threads();
    if (threads > 256) threads = 256;
    if (threads > base.m) threads = base.m > 0 ? base.m : 1;
    pthread_t th[256];
    job_t jobs[256];
    for (int t = 0; t < threads; t++) {
        jobs[t] = base; jobs[t].tid = t; jobs[t].stride = threads;
        if (t > 0) pthread_create(&th[t], NULL, verify_worker, &jobs[t]);
    }
    verify_worker(&jobs[0]);
    for (int t = 1; t < threads; t++) pthread_join(th[t], NULL);
    return 0;
}
int orc_verify_batch(const u64* sigs, const u64* nhms, const u64* g2c, const u64* pkc, int32_t* results,
                     u64* gts, int m, int threads) {
    job_t b = {sigs, nhms, g2c, pkc, NULL, NULL, results, gts, m, 1, 0};
    return run_batch(b, threads);
}
int orc_verify_batch_keyed(const u64* sigs, const u64* nhms, const u64* g2c, const u64* key_coeffs,
                           const int32_t* key_idx, int32_t* results, u64* gts, int m, int threads) {
    job_t b = {sigs, nhms, g2c, NULL, key_coeffs, key_idx, results, gts, m, 1, 0};
    return run_batch(b, threads);
}

/* fused CPU path: decompress (with all checks) + verify against keyed lines.
 * results as the GPU's fused path: 1 valid, 0 invalid, negative decode code; gts
 * (nullable) receives each item's GT element, zeros for a rejected encoding.
 * An identity signature (0xc0 encoding) is paired as the identity, as ic_bls12_381's
 * multi_miller_loop does with an identity term (its line evaluations leave f as it is,
 * oracle/src/main.rs:403-420): the Miller value is that of the (-H(m), pk) pair alone,
 * so GT = e(-H(m), pk) and the verdict is 0 for any pk other than the identity. */
typedef struct {
    const uint8_t* sig48; const u64 *nhms, *g2c, *key_coeffs; const int32_t* key_idx;
    int32_t* res; u64* gts; int m, stride, tid;
} cjob_t;
static void* cverify_worker(void* arg) {
    cjob_t* j = (cjob_t*)arg;
    for (int i = j->tid; i < j->m; i += j->stride) {
        u64 pt[12];
        u64* gt = j->gts ? j->gts + (size_t)i * 72 : NULL;
        const u64* pkc = j->key_coeffs + (size_t)j->key_idx[i] * 2448;
        int rc = orc_g1_decompress(j->sig48 + (size_t)i * 48, pt);
        if (rc == ORC_DEC_IDENTITY) {
            u64 ml[72], r[72];
            orc_miller1(j->nhms + (size_t)i * 12, pkc, ml);
            orc_final_exp(ml, r);
            LD(fp12, rv, r);
            j->res[i] = fp12_is_one(&rv);
            if (gt) memcpy(gt, r, sizeof r);
            continue;
        }
        if (rc != ORC_DEC_OK) {
            j->res[i] = rc;
            if (gt) memset(gt, 0, 72 * sizeof(u64));
            continue;
        }
        j->res[i] = orc_verify(pt, j->nhms + (size_t)i * 12, j->g2c, pkc, gt);
    }
    return NULL;
}
int orc_verify_compressed_batch(const uint8_t* sig48, const u64* nhms, const u64* g2c, const u64* key_coeffs,
                                const int32_t* key_idx, int32_t* results, u64* gts, int m, int threads) {
    if (threads <= 0) threads = orc_num_threads();
    if (threads > 256) threads = 256;
    if (threads > m) threads = m > 0 ? m : 1;
    pthread_t th[256];
    cjob_t jobs[256];
    for (int t = 0; t < threads; t++) {
        cjob_t b = {sig48, nhms, g2c, key_coeffs, key_idx, results, gts, m, threads, t};
        jobs[t] = b;
        if (t > 0) pthread_create(&th[t], NULL, cverify_worker, &jobs[t]);
    }
    cverify_worker(&jobs[0]);
    for (int t = 1; t < threads; t++) pthread_join(th[t], NULL);
    return 0;
}

/* ---- G1 encoding (ZCash format: bit7 compressed, bit6 infinity, bit5 sort) ---- */
static void fp_to_be(uint8_t out[48], const fp* a) {
    fp n; fp_from_mont(&n, a);
    for (int i = 0; i < 6; i++)
        for (int b = 0; b < 8; b++) out[i * 8 + b] = (uint8_t)(n.l[5 - i] >> (56 - 8 * b));
}
static void fp_from_be_raw(fp* r, const uint8_t in[48]) {
    for (int i = 0; i < 6; i++) {
        u64 v = 0;
        for (int b = 0; b < 8; b++) v = (v << 8) | in[i * 8 + b];
        r->l[5 - i] = v;
    }
}
void orc_g1_compress(const u64 p[12], int inf, uint8_t out[48]) {
    if (inf) { memset(out, 0, 48); out[0] = 0xc0; return; }
    LD(g1a, a, p);
    fp_to_be(out, &a.x);
    out[0] |= 0x80;
    if (fp_lex_largest(&a.y)) out[0] |= 0x20;
}
/* ic_bls12_381 G1Affine::from_compressed_unchecked semantics (published zkcrypto
 * algorithm): compression flag required; infinity requires sort=0 and x=0;
 * x must be canonical; x^3+4 must be a square; y chosen by the sort flag. */
int orc_g1_decompress_unchecked(const uint8_t in[48], u64 out[12]) {
    int cflag = (in[0] >> 7) & 1, iflag = (in[0] >> 6) & 1, sflag = (in[0] >> 5) & 1;
    uint8_t xb[48];
    memcpy(xb, in, 48);
    xb[0] &= 0x1f;
    fp raw;
    fp_from_be_raw(&raw, xb);
    memset(out, 0, 96);
    if (!cflag) return ORC_DEC_BAD_FLAGS;
    if (iflag) {
        if (sflag || !fp_is_zero(&raw)) return ORC_DEC_BAD_FLAGS;
        return ORC_DEC_IDENTITY;
    }
    if (!fp_lt_p(&raw)) return ORC_DEC_X_NOT_CANONICAL;
    g1a a;
    fp_to_mont(&a.x, &raw);
    fp y2, b;
    fp_sqr(&y2, &a.x); fp_mul(&y2, &y2, &a.x);
    fp_add(&b, &ONE, &ONE); fp_add(&b, &b, &b);
    fp_add(&y2, &y2, &b);
    if (!fp_sqrt(&a.y, &y2)) return ORC_DEC_NOT_ON_CURVE;
    if (fp_lex_largest(&a.y) != sflag) fp_neg(&a.y, &a.y);
    ST(out, a);
    return ORC_DEC_OK;
}
int orc_g1_decompress(const uint8_t in[48], u64 out[12]) {
    int rc = orc_g1_decompress_unchecked(in, out);
    if (rc != ORC_DEC_OK) return rc;
    LD(g1a, a, out);
    if (!g1a_in_subgroup(&a)) return ORC_DEC_NOT_IN_SUBGROUP;
    return ORC_DEC_OK;
}
/* ---- G2 compression (ic_bls12_381 G2Affine::to_compressed / from_compressed) ----
 * 96 bytes: x.c1 then x.c0, each 48 big-endian bytes; the flags sit in the top three
 * bits of byte 0 as for G1 (compression, infinity, sort).  The sort flag is set when y is
 * lexicographically largest in Fp2: y.c1 lexicographically largest, or y.c1 == 0 and
 * y.c0 lexicographically largest. */
static int fp2_lex_largest(const fp2* a) {
    return fp_lex_largest(&a->c1) || (fp_is_zero(&a->c1) && fp_lex_largest(&a->c0));
}
/* Square root in Fp2 for p = 3 mod 4 (Adj and Rodriguez-Henriquez, eprint 2012/685,
 * Algorithm 9): a1 = a^((p-3)/4), alpha = a1^2 a, x0 = a1 a; alpha == -1 gives the root
 * u x0, otherwise (1 + alpha)^((p-1)/2) x0.  Returns 1 when the candidate squares back
 * to a. */
static int fp2_sqrt(fp2* r, const fp2* a) {
    static const u64 P34[6] = {0xee7fbfffffffeaaaULL, 0x07aaffffac54ffffULL, 0xd9cc34a83dac3d89ULL,
                               0xd91dd2e13ce144afULL, 0x92c6e9ed90d2eb35ULL, 0x0680447a8e5ff9a6ULL};
    static const u64 P12[6] = {0xdcff7fffffffd555ULL, 0x0f55ffff58a9ffffULL, 0xb39869507b587b12ULL,
                               0xb23ba5c279c2895fULL, 0x258dd3db21a5d66bULL, 0x0d0088f51cbff34dULL};
    if (fp2_is_zero(a)) { *r = *a; return 1; }
    fp2 a1 = fp2_one(), base = *a;
    for (int w = 0; w < 6; w++)
        for (int b = 0; b < 64; b++) {
            if ((P34[w] >> b) & 1) fp2_mul(&a1, &a1, &base);
            fp2_sqr(&base, &base);
        }
    fp2 alpha, x0, m1, cand;
    fp2_sqr(&alpha, &a1);
    fp2_mul(&alpha, &alpha, a);
    fp2_mul(&x0, &a1, a);
    m1 = fp2_one();
    fp2_neg(&m1, &m1);
    if (fp2_eq(&alpha, &m1)) {
        fp nc1; fp_neg(&nc1, &x0.c1);
        cand.c0 = nc1; cand.c1 = x0.c0;
    } else {
        fp2 b = fp2_one(), t;
        fp2_add(&t, &alpha, &b);
        base = t;
        for (int w = 0; w < 6; w++)
            for (int bit = 0; bit < 64; bit++) {
                if ((P12[w] >> bit) & 1) fp2_mul(&b, &b, &base);
                fp2_sqr(&base, &base);
            }
        fp2_mul(&cand, &b, &x0);
    }
    fp2 c2; fp2_sqr(&c2, &cand);
    *r = cand;
    return fp2_eq(&c2, a);
}
void orc_g2_compress(const u64 q[24], int inf, uint8_t out[96]) {
    if (inf) { memset(out, 0, 96); out[0] = 0xc0; return; }
    LD(g2a, a, q);
    fp_to_be(out, &a.x.c1);
    fp_to_be(out + 48, &a.x.c0);
    out[0] |= 0x80;
    if (fp2_lex_largest(&a.y)) out[0] |= 0x20;
}
/* G2Affine::from_compressed_unchecked: the same flag rules as G1; both coordinates of x
 * canonical; y^2 = x^3 + 4(1 + u) solvable; y chosen by the sort flag. */
int orc_g2_decompress_unchecked(const uint8_t in[96], u64 out[24]) {
    int cflag = (in[0] >> 7) & 1, iflag = (in[0] >> 6) & 1, sflag = (in[0] >> 5) & 1;
    uint8_t xb[96];
    memcpy(xb, in, 96);
    xb[0] &= 0x1f;
    fp r1, r0;
    fp_from_be_raw(&r1, xb);
    fp_from_be_raw(&r0, xb + 48);
    memset(out, 0, 192);
    if (!cflag) return ORC_DEC_BAD_FLAGS;
    if (iflag) {
        if (sflag || !fp_is_zero(&r1) || !fp_is_zero(&r0)) return ORC_DEC_BAD_FLAGS;
        return ORC_DEC_IDENTITY;
    }
    if (!fp_lt_p(&r1) || !fp_lt_p(&r0)) return ORC_DEC_X_NOT_CANONICAL;
    g2a a;
    fp_to_mont(&a.x.c1, &r1);
    fp_to_mont(&a.x.c0, &r0);
    fp2 y2, b = fp2_one();
    fp_add(&b.c0, &b.c0, &b.c0); fp_add(&b.c0, &b.c0, &b.c0);  /* 4 */
    b.c1 = b.c0;                                                   /* 4 + 4u */
    fp2_sqr(&y2, &a.x);
    fp2_mul(&y2, &y2, &a.x);
    fp2_add(&y2, &y2, &b);
    if (!fp2_sqrt(&a.y, &y2)) return ORC_DEC_NOT_ON_CURVE;
    if (fp2_lex_largest(&a.y) != sflag) fp2_neg(&a.y, &a.y);
    ST(out, a);
    return ORC_DEC_OK;
}
int orc_g2_decompress(const uint8_t in[96], u64 out[24]) {
    int rc = orc_g2_decompress_unchecked(in, out);
    if (rc != ORC_DEC_OK) return rc;
    LD(g2a, a, out);
    if (!g2a_in_subgroup(&a)) return ORC_DEC_NOT_IN_SUBGROUP;
    return ORC_DEC_OK;
}
int orc_g2_is_on_curve(const u64 q[24]) {
    LD(g2a, a, q);
    fp2 y2, x3, b = fp2_one();
    fp_add(&b.c0, &b.c0, &b.c0); fp_add(&b.c0, &b.c0, &b.c0);
    b.c1 = b.c0;
    fp2_sqr(&y2, &a.y);
    fp2_sqr(&x3, &a.x); fp2_mul(&x3, &x3, &a.x);
    fp2_add(&x3, &x3, &b);
    return fp2_eq(&y2, &x3);
}
int orc_g1_is_on_curve(const u64 p[12]) { LD(g1a, a, p); return g1a_on_curve(&a); }
int orc_g1_in_subgroup(const u64 p[12]) { LD(g1a, a, p); return g1a_in_subgroup(&a); }
int orc_g2_in_subgroup(const u64 q[24]) { LD(g2a, a, q); return g2a_in_subgroup(&a); }
int orc_g1_mul(const u64 p[12], const uint8_t k[32], u64 out[12]) {
    LD(g1a, a, p); g1j j, o; g1a r;
    g1_from_affine(&j, &a); g1_mul(&o, &j, k);
    int inf = g1_to_affine(&r, &o);
    ST(out, r);
    return inf;
}
int orc_g2_mul(const u64 q[24], const uint8_t k[32], u64 out[24]) {
    LD(g2a, a, q); g2j j, o; g2a r;
    g2_from_affine(&j, &a); g2_mul(&o, &j, k);
    int inf = g2_to_affine(&r, &o);
    ST(out, r);
    return inf;
}
void orc_g1_neg(const u64 p[12], u64 out[12]) { LD(g1a, a, p); fp_neg(&a.y, &a.y); ST(out, a); }
void orc_g1_generator(u64 out[12]) { ST(out, G1_GEN); }
void orc_g2_generator(u64 out[24]) { ST(out, G2_GEN); }

/* oracle/src/main.rs:383-397 */
void orc_hash_to_g1(const uint8_t* msg, size_t len, u64 out[12]) {
    static const char DOM[] = "ic-bls-test-hash";
    uint8_t buf[16 + 4096];
    uint8_t h[32];
    if (len <= 4096) {
        memcpy(buf, DOM, 16); memcpy(buf + 16, msg, len);
        orc_sha256(buf, 16 + len, h);
    } else {
        memset(h, 0, 32); /* messages are <= 4 KiB in every workload here */
    }
    h[31] &= 0x0f;
    u64 g[12];
    orc_g1_generator(g);
    orc_g1_mul(g, h, out);
}

/* splitmix64 stream keyed by (seed, index) */
static u64 smix(u64* s) {
    u64 z = (*s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
void orc_rand_bytes(u64 seed, u64 index, uint8_t* out, size_t len) {
    u64 s = seed * 0xd1342543de82ef95ULL + index * 0x9e3779b97f4a7c15ULL + 1;
    for (size_t i = 0; i < len; i += 8) {
        u64 v = smix(&s);
        for (size_t b = 0; b < 8 && i + b < len; b++) out[i + b] = (uint8_t)(v >> (8 * b));
    }
}
void orc_rand_scalar(u64 seed, u64 index, uint8_t out[32]) {
    u64 s = seed * 0xd1342543de82ef95ULL + index * 0x9e3779b97f4a7c15ULL + 0x5ca1ab1eULL;
    for (;;) {
        u64 w[4];
        for (int i = 0; i < 4; i++) w[i] = smix(&s);
        w[3] &= 0x7fffffffffffffffULL;
        int lt = 0, nz = (w[0] | w[1] | w[2] | w[3]) != 0;
        for (int i = 3; i >= 0; i--) {
            if (w[i] < R_ORDER[i]) { lt = 1; break; }
            if (w[i] > R_ORDER[i]) break;
        }
        if (lt && nz) {
            for (int i = 0; i < 32; i++) out[i] = (uint8_t)(w[i / 8] >> (8 * (i % 8)));
            return;
        }
    }
}
void orc_rand_x(u64 seed, u64 index, int want_square, u64 x_mont[6]) {
    u64 s = seed * 0xd1342543de82ef95ULL + index * 0x9e3779b97f4a7c15ULL + 0xbadc0ffeeULL;
    for (;;) {
        fp raw;
        for (int i = 0; i < 6; i++) raw.l[i] = smix(&s);
        raw.l[5] &= 0x1fffffffffffffffULL;
        if (!fp_lt_p(&raw)) continue;
        fp x, y2, b, y;
        fp_to_mont(&x, &raw);
        fp_sqr(&y2, &x); fp_mul(&y2, &y2, &x);
        fp_add(&b, &ONE, &ONE); fp_add(&b, &b, &b);
        fp_add(&y2, &y2, &b);
        if (fp_sqrt(&y, &y2) == want_square) { ST(x_mont, x); return; }
    }
}
