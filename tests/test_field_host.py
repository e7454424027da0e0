"""The device field arithmetic (icp-gpu_amd/csrc/field.hpp) compiled for the host by
g++ with a stand-in HIP header (tests/hostshim) and checked against Python big-integer
arithmetic: Montgomery product and square, and the divsteps inversion against the
closed form x * x^-1 = 1 and against Fermat, including non-canonical limb inputs."""
import ctypes
import os
import random
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
P = 0x1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f6241eabfffeb153ffffb9feffffffffaaab
R = 1 << 384
MASK = (1 << 64) - 1


@pytest.fixture(scope="module")
def fh(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("fieldhost") / "field_host.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(HERE, "hostshim"),
                    "-I", os.path.join(ROOT, "icp-gpu_amd", "csrc"), os.path.join(HERE, "harness", "field_host.cpp"),
                    "-o", so], check=True)
    return ctypes.CDLL(so)


def ref(x):
    """canonical integer -> reference Montgomery limbs (6 x u64)"""
    m = x * R % P
    return np.array([(m >> (64 * i)) & MASK for i in range(6)], dtype=np.uint64)


def unref(a):
    m = sum(int(v) << (64 * i) for i, v in enumerate(a))
    assert m < P  # every exported value is canonical
    return m * pow(R, -1, P) % P


def call(fn, *args):
    out = np.zeros(6, dtype=np.uint64)
    fn(*[a.ctypes.data_as(ctypes.c_void_p) for a in args], out.ctypes.data_as(ctypes.c_void_p))
    return unref(out)


def samples(n, seed):
    rnd = random.Random(seed)
    edge = [0, 1, 2, P - 1, P - 2, (P - 1) // 2, (P + 1) // 2, 1 << 380, (1 << 381) % P]
    return edge + [rnd.randrange(P) for _ in range(n)]


def test_mul_and_sqr_closed_form(fh):
    xs = samples(200, 1)
    rnd = random.Random(2)
    for x in xs:
        y = rnd.randrange(P)
        assert call(fh.fh_mul, ref(x), ref(y)) == x * y % P
        assert call(fh.fh_sqr, ref(x)) == x * x % P


def test_divsteps_inverse(fh):
    for x in samples(500, 3):
        got = call(fh.fh_inv, ref(x))
        if x == 0:
            assert got == 0
        else:
            assert got * x % P == 1, hex(x)


def test_divsteps_inverse_equals_fermat(fh):
    for x in samples(40, 4):
        want = call(fh.fh_inv_fermat, ref(x))
        assert call(fh.fh_inv, ref(x)) == want
        assert call(fh.fh_inv_uniform, ref(x)) == want


def test_divsteps_inverse_of_noncanonical_limbs(fh):
    # internal limbs of values that are congruent to x but not reduced: x + k p with
    # signed limbs, the form lazy sums leave behind (the inverse canonicalizes first)
    rnd = random.Random(5)
    rp = pow(1 << 392, -1, P)  # internal radix R' = 2^392
    for _ in range(60):
        x = rnd.randrange(1, P)
        v = x * (1 << 392) % P + rnd.choice([-1, 0, 1, 2]) * P
        limbs = []
        t = v
        for i in range(13):
            limbs.append(t & ((1 << 28) - 1))
            t >>= 28
        limbs.append(t)
        # move some weight between neighbouring limbs (still the same value)
        k = rnd.randrange(13)
        limbs[k] += 1 << 28
        limbs[k + 1] -= 1
        arr = np.array(limbs, dtype=np.int32)
        got = call(fh.fh_inv_limbs, arr)
        assert got * x % P == 1
