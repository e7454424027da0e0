"""The certifier's share loop (ic-gpu-consensus.patch:98-120) restated in C++
(tests/harness/certifier_harness.cpp): one height's shares, all over the same certified
content, verified in one fused call with the committee's compressed keys registered on
the GPU, or on the CPU path when the library is unusable.  Both paths must give the
loop's actions: MoveToValidated (1), HandleInvalid (0), no action (-1)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import blsgpu
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import h2c  # noqa: E402

SRC = os.path.join(HERE, "harness", "certifier_harness.cpp")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("certifier") / "certifier_harness")
    subprocess.run(["g++", "-O1", "-std=c++17", SRC, "-o", exe, "-ldl"], check=True)
    return exe


def _height(tmp_path):
    """a committee of 5 (key 3's encoding corrupted), 9 shares over one content"""
    content = b"certified state: height 4711, hash " + bytes(range(32))
    sks, pks, _ = O.make_keys(71, 5)
    keys = [bytearray(O.g2_compress(pk)) for pk in pks]
    keys[3][0] &= 0x7F  # not a compressed encoding: the signer's key is unusable
    x, y = h2c.hash_to_g1(content, h2c.IC_DST)
    hm = np.concatenate([O.mont(x), O.mont(y)])
    xo, yo = h2c.hash_to_g1(content + b"?", h2c.IC_DST)
    other = np.concatenate([O.mont(xo), O.mont(yo)])
    shares, want = [], []

    def share(signer, point, sk_idx=None, action=1, mangle=False):
        s, _ = O.g1_mul(point, sks[signer if sk_idx is None else sk_idx])
        b = bytearray(O.g1_compress(s))
        if mangle:
            b[0] &= 0x7F
        shares.append((signer, bytes(b)))
        want.append(action)

    share(0, hm)
    share(1, hm)
    share(2, other, action=0)        # a share over other content
    share(3, hm, action=-1)          # the signer's key does not decode
    share(4, hm)
    share(1, hm, sk_idx=2, action=0)  # signed with another member's key
    share(2, hm, mangle=True, action=0)  # the signature does not decode
    share(9, hm, sk_idx=0, action=-1)    # no such committee member
    share(2, hm)
    inp = tmp_path / "height.bin"
    with open(inp, "wb") as f:
        f.write(np.array([len(shares), len(keys), len(content), len(h2c.IC_DST)], np.int32).tobytes())
        f.write(content + h2c.IC_DST + b"".join(bytes(k) for k in keys))
        for signer, sig in shares:
            f.write(np.int32(signer).tobytes() + sig)
        f.write(np.ascontiguousarray(O.g1_neg(hm), np.uint64).tobytes())
    return inp, want


def _run(harness, inp, out):
    subprocess.run([harness, blsgpu.LIB_PATH, O.ORACLE_PATH, str(inp), str(out)], check=True, timeout=120)
    v = np.fromfile(out, dtype=np.int32)
    return int(v[0]), v[1:].tolist()


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_cpu_path_without_gpu(harness, tmp_path):
    inp, want = _height(tmp_path)
    path, actions = _run(harness, inp, tmp_path / "out.bin")
    assert path == 0
    assert actions == want


@pytest.mark.gpu
def test_gpu_path_with_gpu(harness, tmp_path):
    print("case: certifier harness on the GPU", flush=True)
    inp, want = _height(tmp_path)
    path, actions = _run(harness, inp, tmp_path / "out.bin")
    assert path == 1
    assert actions == want
