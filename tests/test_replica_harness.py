"""The drop-in exercised the way the replica's Rust crate calls it (dlopen, the
three-symbol check, gpu_crypto_init, gpu_batch_bls_verify, CPU fallback on any
failure), through a C++ restatement of that caller sequence (tests/harness).  Without a
GPU the harness must take the fallback and still give the right verdicts; with one it
must take the GPU path and agree."""
import os
import subprocess

import numpy as np
import pytest

import blsgpu
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "harness", "replica_harness.cpp")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("harness") / "replica_harness")
    subprocess.run(["g++", "-O1", "-std=c++17", SRC, "-o", exe, "-ldl"], check=True)
    return exe


def _vector(tmp_path):
    fx = O.fixtures()
    sig = np.stack([O.u64(fx["sig_294_g1"]), O.g1_gen()])  # valid, then invalid
    nhm = np.stack([O.u64(fx["neg_7_g1"])] * 2)
    g2c = O.u64(fx["g2_gen_coeffs"])
    pkc = np.stack([O.u64(fx["pk42_coeffs"])] * 2)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(np.int32(2).tobytes())
        for a in (sig, nhm, g2c, pkc):
            f.write(np.ascontiguousarray(a, np.uint64).tobytes())
    return inp


def _run(harness, inp, out):
    subprocess.run([harness, blsgpu.LIB_PATH, O.ORACLE_PATH, str(inp), str(out)], check=True, timeout=120)
    v = np.fromfile(out, dtype=np.int32)
    return int(v[0]), v[1:].tolist()


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_cpu_fallback_without_gpu(harness, tmp_path):
    path, verdicts = _run(harness, _vector(tmp_path), tmp_path / "out.bin")
    assert path == 0
    assert verdicts == [1, 0]


@pytest.mark.gpu
def test_gpu_path_with_gpu(harness, tmp_path):
    print("case: replica harness on the GPU", flush=True)
    path, verdicts = _run(harness, _vector(tmp_path), tmp_path / "out.bin")
    assert path == 1
    assert verdicts == [1, 0]
