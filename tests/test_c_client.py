"""The C ABI from a C program: tests/c_abi_client.c includes only
include/voidray_hip.h and links libvoidray_hip.so, as a Rust -sys crate or
another native host binds the library.  On the CPU it must compile as C99 and
link; on the GPU it runs (a child process) and checks a trace with known hits,
a render and a tonemap."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "voidray_amd")
ROCM_LIB = "/opt/rocm/lib"


def _build(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    if not os.path.exists(os.path.join(PKG, "libvoidray_hip.so")):
        pytest.fail("libvoidray_hip.so is missing: build it first")
    exe = str(tmp_path / "c_abi_client")
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c_abi_client.c"), "-L", PKG, "-lvoidray_hip", f"-Wl,-rpath,{PKG}",
           "-L", ROCM_LIB, f"-Wl,-rpath,{ROCM_LIB}", "-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_client_compiles_and_links(tmp_path):
    assert os.path.exists(_build(tmp_path))


@pytest.mark.gpu
def test_c_client_runs(tmp_path):
    exe = _build(tmp_path)
    print("[gpu-case] C ABI client", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert r.stdout.strip().endswith("ok")
