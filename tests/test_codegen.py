"""Device code generation checks (CPU: hipcc compiles for gfx950 here).

No f64 instruction in any kernel: f64 is used on the host only (the
environment CDF's prefix sums, the oracle); the shading and tonemap code's
transcendentals are f32-only routines (f32math.h).

A reduction of per-lane bools that the optimiser vectorises into a compare
on <N x i1> and then bitcasts to an integer (bitcast <4 x i1> to i4, then
compared with 0) was lowered by the gfx950 backend through v_readfirstlane:
every lane of the wave received the first active lane's answer.  The
default-box test of the reference's slab test (box_is_default) compiled that
way until it was rewritten; this test keeps the pattern out of every kernel
source.
"""
import os
import re
import shutil
import subprocess

import pytest

from voidray_amd import build as vb

SOURCES = sorted(f for f in os.listdir(vb.CSRC) if f.endswith(".hip"))
PATTERN = re.compile(r"bitcast <\d+ x i1> .* to i\d+")


@pytest.mark.skipif(shutil.which(vb._hipcc()) is None and not os.path.exists(vb._hipcc()), reason="hipcc not found")
@pytest.mark.parametrize("src", SOURCES)
def test_no_i1_vector_bitcasts_in_device_code(src, tmp_path):
    out = tmp_path / (src + ".ll")
    cmd = [vb._hipcc(), "-x", "hip", *vb.HIP_FLAGS, "--cuda-device-only", "-S", "-emit-llvm",
           os.path.join(vb.CSRC, src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    hits = [l.strip() for l in out.read_text().splitlines() if PATTERN.search(l)]
    assert not hits, f"{src}: {len(hits)} i1-vector bitcasts, first: {hits[0]}"


@pytest.mark.skipif(shutil.which(vb._hipcc()) is None and not os.path.exists(vb._hipcc()), reason="hipcc not found")
@pytest.mark.parametrize("src", SOURCES)
def test_no_f64_instructions_in_device_code(src, tmp_path):
    out = tmp_path / (src + ".s")
    cmd = [vb._hipcc(), "-x", "hip", *vb.HIP_FLAGS, "--cuda-device-only", "-S", os.path.join(vb.CSRC, src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    f64 = re.compile(r"^\s+v_\w+_f64")
    hits = [l.strip() for l in out.read_text().splitlines() if f64.match(l)]
    assert not hits, f"{src}: {len(hits)} f64 instructions, first: {hits[0]}"
