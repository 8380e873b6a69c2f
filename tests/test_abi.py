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


def test_struct_sizes_match_header_layout():
    import ctypes as C
    assert C.sizeof(N.vr_mesh) == 8 + 4 * 8
    assert C.sizeof(N.vr_material) == 4 + 12 + 4 + 4 + 4
    assert C.sizeof(N.vr_trace_desc) == 16
    assert C.sizeof(N.vr_render_desc) == 4 * 10 + 8 + 8
    assert C.sizeof(N.vr_request) == 8 + C.sizeof(N.vr_render_desc) + 8
    assert C.sizeof(N.vr_tuning) == 8
    assert C.sizeof(N.vr_cache_desc) == 16 + 24 + 20 + 4 + 8
    assert C.sizeof(N.vr_cache_info) == 16 + 64 + 16 + 16
    assert C.sizeof(N.vr_cache_frame) == 24
    assert C.sizeof(N.vr_cache_stats) == 40


def test_errors_do_not_unwind():
    import ctypes as C
    lib = N.lib()
    assert lib.vr_scene_upload(None, None, None) != 0
    assert b"null" in lib.vr_last_error()
    assert lib.vr_trace(None, None, None, None, None, None, None, None, None) != 0
