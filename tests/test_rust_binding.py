"""The Rust -sys crate (rust/voidray_hip_sys/src/lib.rs) mirrors
include/voidray_hip.h: every constant with its value, every struct with its
fields in order and their types, every entry point with its parameter and
return types.  Struct layouts are computed from the Rust field types with the
C layout rules (#[repr(C)]) and compared with gcc's sizeof/offsetof of the
header.  No Rust toolchain is needed (there is none in this image)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "voidray_hip.h")
LIBRS = os.path.join(ROOT, "rust", "voidray_hip_sys", "src", "lib.rs")


def _strip_c(txt):
    return re.sub(r"/\*.*?\*/", "", txt, flags=re.S)


def _strip_rs(txt):
    return re.sub(r"//[^\n]*", "", txt)


def _c_int(v):
    v = v.strip().strip("()")
    v = re.sub(r"[uU]+$", "", v)
    return int(v, 0)


def header_consts():
    out = {}
    for name, val in re.findall(r"#define\s+(VR_[A-Z0-9_]+)\s+([^\s]+)", _strip_c(open(HEADER).read())):
        out[name] = _c_int(val)
    return out


def rust_consts():
    out = {}
    for name, val in re.findall(r"pub const\s+(VR_[A-Z0-9_]+)\s*:\s*\w+\s*=\s*([^;]+);", _strip_rs(open(LIBRS).read())):
        out[name] = int(val.replace("_", "").strip(), 0)
    return out


C2RS = {"uint32_t": "u32", "int32_t": "i32", "uint64_t": "u64", "uint8_t": "u8", "float": "f32",
        "double": "f64", "int": "c_int", "char": "c_char", "void": "c_void"}


def _c_type_to_rs(ctype):
    """'const float*' -> '*const f32'; 'vr_ctx**' -> '*mut *mut vr_ctx'."""
    t = ctype.strip()
    stars = t.count("*")
    const = t.startswith("const ")
    base = t.replace("const ", "").replace("*", "").strip()
    rs = C2RS.get(base, base)
    if stars == 0:
        return rs
    # the outermost pointer carries `const` only when the pointee is const
    # (const T* / const T** patterns used in the header)
    ptrs = ["*const " if (const and i == 0) else "*mut " for i in range(stars)]
    return "".join(reversed(ptrs)) + rs


def header_structs():
    txt = _strip_c(open(HEADER).read())
    out = {}
    for body, name in re.findall(r"typedef struct (?:\w+)\s*\{(.*?)\}\s*(\w+);", txt, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"^((?:const )?[\w]+\s*\**)\s*(.*)$", decl)
            ctype, names = m.group(1), m.group(2)
            for nm in names.split(","):
                nm = nm.strip()
                ptr = ""
                while nm.startswith("*"):
                    ptr += "*"
                    nm = nm[1:]
                arr = re.match(r"(\w+)\[(\w+)\]", nm)
                t = _c_type_to_rs(ctype + ptr)
                if arr:
                    fields.append((arr.group(1), f"[{t}; {arr.group(2)}]"))
                else:
                    fields.append((nm, t))
        out[name] = fields
    return out


def rust_structs():
    txt = _strip_rs(open(LIBRS).read())
    out = {}
    for name, body in re.findall(r"#\[repr\(C\)\][^{]*?pub struct (\w+)\s*\{(.*?)\}", txt, flags=re.S):
        fields = []
        for f in body.split(","):
            f = " ".join(f.split())
            if not f:
                continue
            m = re.match(r"^pub (\w+)\s*:\s*(.+)$", f)
            if m is None:  # opaque handle: _private
                fields = None
                break
            fields.append((m.group(1), m.group(2).replace(" ;", ";")))
        out[name] = fields
    return out


def header_functions():
    txt = _strip_c(open(HEADER).read())
    out = {}
    for ret, name, params in re.findall(r"^\s*((?:const )?\w+\s*\**)\s*(vr_\w+)\s*\(([^)]*)\);", txt, flags=re.M):
        ps = []
        params = " ".join(params.split())
        if params and params != "void":
            for p in params.split(","):
                p = p.strip()
                m = re.match(r"^(.*?)(\w+)$", p)
                ps.append(_c_type_to_rs(m.group(1)))
        out[name] = (_c_type_to_rs(ret), ps)
    return out


def rust_functions():
    txt = _strip_rs(open(LIBRS).read())
    block = txt[txt.index('extern "C"'):]
    out = {}
    for name, params, ret in re.findall(r"pub fn (vr_\w+)\s*\(([^)]*)\)\s*(?:->\s*([^;]+))?;", block, flags=re.S):
        ps = []
        params = " ".join(params.split())
        if params:
            for p in params.split(","):
                p = p.strip()
                if p:
                    ps.append(p.split(":", 1)[1].strip())
        out[name] = ((ret or "()").strip(), ps)
    return out


def test_constants_match_the_header():
    h, r = header_consts(), rust_consts()
    assert set(h) == set(r), (sorted(set(h) - set(r)), sorted(set(r) - set(h)))
    for k in h:
        assert h[k] == r[k], (k, h[k], r[k])


def test_struct_fields_match_the_header():
    h, r = header_structs(), rust_structs()
    assert len(h) >= 18
    for name, fields in h.items():
        assert name in r and r[name] is not None, name
        assert fields == r[name], (name, fields, r[name])
    opaque = {k for k, v in r.items() if v is None}
    assert opaque == {"vr_ctx", "vr_scene", "vr_cache"}
    assert set(r) - opaque == set(h)


def test_functions_match_the_header():
    h, r = header_functions(), rust_functions()
    assert len(h) >= 30
    assert set(h) == set(r), (sorted(set(h) - set(r)), sorted(set(r) - set(h)))
    for name, (ret, ps) in h.items():
        assert r[name] == (ret, ps), (name, (ret, ps), r[name])


# ---- layouts: the Rust field types under #[repr(C)] rules vs gcc --------------
PRIM = {"u8": 1, "i32": 4, "u32": 4, "f32": 4, "c_int": 4, "u64": 8, "f64": 8}


def _layout(ty, structs, consts):
    ty = ty.strip()
    if ty.startswith("*"):
        return 8, 8
    if ty in PRIM:
        return PRIM[ty], PRIM[ty]
    m = re.match(r"^\[(.+);\s*(\w+)\]$", ty)
    if m:
        n = int(m.group(2)) if m.group(2).isdigit() else consts[m.group(2)]
        s, a = _layout(m.group(1), structs, consts)
        return s * n, a
    off, align, _ = _struct_layout(ty, structs, consts)
    return off, align


def _struct_layout(name, structs, consts):
    off, align, offs = 0, 1, {}
    for f, t in structs[name]:
        s, a = _layout(t, structs, consts)
        off = (off + a - 1) // a * a
        offs[f] = off
        off += s
        align = max(align, a)
    return (off + align - 1) // align * align, align, offs


def test_rust_layouts_match_gcc(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    r = {k: v for k, v in rust_structs().items() if v is not None}
    consts = rust_consts()
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "voidray_hip.h"', "int main(void) {"]
    for n, fields in r.items():
        lines.append(f'printf("{n} %zu\\n", sizeof({n}));')
        for f, _ in fields:
            lines.append(f'printf("{n}.{f} %zu\\n", offsetof({n}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                  check=True).stdout.splitlines())
    for n in r:
        size, _, offs = _struct_layout(n, r, consts)
        assert int(got[n]) == size, (n, got[n], size)
        for f, o in offs.items():
            assert int(got[f"{n}.{f}"]) == o, (n, f, got[f"{n}.{f}"], o)
