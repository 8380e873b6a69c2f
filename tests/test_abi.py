"""The C ABI library loads and exports every function include/voidray_hip.h declares."""
import os
import re

from voidray_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "voidray_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vr_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = N.lib()
    names = _declared()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(N.SIGNATURES), set(names) ^ set(N.SIGNATURES)


def test_struct_sizes_match_header_layout(tmp_path):
    # sizes and field offsets as the C compiler lays out include/voidray_hip.h,
    # against the ctypes mirror
    import ctypes as C
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        import pytest
        pytest.skip("gcc not available")
    structs = [n for n in dir(N) if n.startswith("vr_") and isinstance(getattr(N, n), type)
               and issubclass(getattr(N, n), C.Structure)]
    assert len(structs) >= 14
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "voidray_hip.h"', "int main(void) {"]
    for n in structs:
        lines.append(f'printf("{n} %zu\\n", sizeof({n}));')
        for f, _ in getattr(N, n)._fields_:
            lines.append(f'printf("{n}.{f} %zu\\n", offsetof({n}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "abi.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                  text=True).stdout.splitlines())
    for n in structs:
        t = getattr(N, n)
        assert C.sizeof(t) == int(out[n]), n
        for f, _ in t._fields_:
            assert getattr(t, f).offset == int(out[f"{n}.{f}"]), f"{n}.{f}"


def test_errors_do_not_unwind():
    import ctypes as C
    lib = N.lib()
    assert lib.vr_scene_upload(None, None, None) != 0
    assert b"null" in lib.vr_last_error()
    assert lib.vr_trace(None, None, None, None, None, None, None, None, None) != 0
