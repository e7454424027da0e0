"""The drop-in boundary on the CPU: libgpu_msm.so loads, exports every symbol that
include/gpu_msm.h declares, and keeps the reference's failure conventions.  No
compute entry point is called here (there is no GPU in this environment)."""
import ctypes
import os
import re
import subprocess

import pytest

import blsgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def declared_symbols():
    syms = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        syms |= set(re.findall(r"\b(gpu_\w+|blsgpu_\w+)\s*\(", text))
    return sorted(syms)


def test_header_declares_reference_abi():
    syms = declared_symbols()
    for s in blsgpu.REFERENCE_SYMBOLS:
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = blsgpu.load()
    for s in declared_symbols():
        assert hasattr(lib, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", blsgpu.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    for s in declared_symbols():
        assert s in exported, s


def test_dlopen_three_symbol_check_like_the_replica():
    # ic-integration/gpu_crypto/lib.rs:92-122: the library is used only if all three resolve
    h = ctypes.CDLL(blsgpu.LIB_PATH, mode=os.RTLD_NOW | os.RTLD_GLOBAL)
    for s in ("gpu_msm_g1", "gpu_jacobian_to_affine", "gpu_g1_decompress_batch"):
        assert getattr(h, s) is not None


def test_out_of_scope_entry_points_fail_so_callers_use_cpu():
    lib = blsgpu.load()
    assert lib.gpu_msm_g1(None, None, 0, None) == -1
    assert lib.gpu_jacobian_to_affine(None, None) == -1
    assert lib.gpu_batch_sha256(None, None, 0, 0) == -1


def test_invalid_sizes_rejected_before_any_device_work():
    lib = blsgpu.load()
    assert lib.gpu_batch_bls_verify(None, None, None, None, None, 0) == -1
    assert lib.gpu_g1_decompress_batch(None, None, 0) == -1
    assert lib.gpu_g1_decompress_batch(None, None, -3) == -1


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_init_reports_no_device():
    # gpu-crypto/cuda/api.cu:118-124: -1 when no device is found
    assert blsgpu.init() == -1


def test_batched_and_serving_entry_points_validate_arguments():
    lib = blsgpu.load()
    off = (ctypes.c_int32 * 3)(0, 5, 3)  # decreasing offsets
    assert lib.blsgpu_verify_batches(None, None, None, None, 0, None, None) == -1
    assert lib.blsgpu_verify_batches_device(None, None, None, ctypes.cast(off, ctypes.c_void_p), 2, None, None,
                                            None) == -1
    assert lib.blsgpu_set_batched_engine(11, 16, 0, 0, 1) == -1
    assert lib.blsgpu_set_batched_engine(12, 17, 0, 0, 1) == -1
    assert lib.blsgpu_set_batched_engine(12, 16, -2, 0, 1) == -1
    assert lib.blsgpu_set_batched_engine(13, 18, 1, -1, 0) == 0
    cfg = (ctypes.c_int * 5)()
    assert lib.blsgpu_get_batched_engine(100, ctypes.cast(cfg, ctypes.c_void_p)) == 0
    assert list(cfg) == [13, 18, 1, -1, 0]
    assert lib.blsgpu_set_batched_engine(12, 16, 0, 0, 1) == 0
    assert not lib.blsgpu_server_create(0, 16, 1 << 20, 0, 0)
    assert not lib.blsgpu_server_create(2, 0, 1 << 20, 0, 0)
    assert lib.blsgpu_server_submit(None, None, None, None, 1) == -1
    assert lib.blsgpu_server_poll(None, 0, None, None, None) == -1


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_server_needs_a_device():
    assert not blsgpu.load().blsgpu_server_create(2, 64, 1 << 20, 0, 0)
    assert "device" in blsgpu.last_error()


def test_static_archive_links_a_caller(tmp_path):
    # the same objects as libgpu_msm.a, for callers that link statically (the reference's
    # Bazel route, BUILD.bazel:22-26): every declared symbol defined, and a C caller links
    # against it with the HIP runtime and sees no device here
    ar = os.path.join(ROOT, "icp-gpu_amd", "lib", "libgpu_msm.a")
    out = subprocess.run(["nm", "-g", "--defined-only", ar], capture_output=True, text=True).stdout
    defined = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    for s in declared_symbols():
        assert s in defined, s
    src = tmp_path / "caller.cpp"
    src.write_text('#include <stdio.h>\n#include "gpu_msm.h"\n'
                   'int main(void) { printf("%d\\n", gpu_crypto_init()); return 0; }\n')
    exe = tmp_path / "caller"
    subprocess.run(["g++", str(src), "-I", os.path.join(ROOT, "include"), ar, "-L/opt/rocm/lib", "-lamdhip64",
                    "-Wl,-rpath,/opt/rocm/lib", "-lpthread", "-o", str(exe)], check=True, timeout=300)
    if os.path.exists("/dev/kfd"):
        return  # with a device the init would succeed; the link is what this test covers
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "-1"


def test_workload_generator_is_a_separate_library():
    """The synthetic workload generator (keys, signer) and the issue-rate probe live in
    libblsgpu_workload.so, which the benchmark and the tests load; libgpu_msm.so, the
    library a replica loads, exports none of them."""
    from blsgpu import workload
    hdr = os.path.join(ROOT, "icp-gpu_amd", "csrc", "blsgpu_workload.h")
    text = re.sub(r"/\*.*?\*/", "", open(hdr).read(), flags=re.S)
    wl_syms = set(re.findall(r"\b(blsgpu_\w+)\s*\(", text))
    assert "blsgpu_wl_items_msg" in wl_syms and "blsgpu_probe_mad_rate" in wl_syms
    nm = lambda p: set(l.split()[-1] for l in subprocess.run(["nm", "-D", "--defined-only", p], capture_output=True,
                                                               text=True).stdout.splitlines() if l.strip())
    main, wl = nm(blsgpu.LIB_PATH), nm(workload.WORKLOAD_LIB_PATH)
    assert wl_syms <= wl
    assert not (wl_syms & main)
    assert not any(s.startswith("blsgpu_wl_") for s in main)
