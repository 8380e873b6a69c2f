"""ctypes wrapper of libf32math_probe.so (built by tests/native/build.py)."""
import ctypes
import os

import numpy as np

from tests.native import build as _build

FN = {"exp": 0, "ln": 1, "atan": 2, "sin": 3, "cos": 4}
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_build.LIB):
            _build.build()
        L = ctypes.CDLL(_build.LIB)
        P, U64 = ctypes.c_void_p, ctypes.c_uint64
        for name in ("f32m_eval_host", "f32m_eval_libm", "f32m_eval_f64", "f32m_eval_device"):
            getattr(L, name).argtypes = [ctypes.c_int, P, P, U64]
            getattr(L, name).restype = ctypes.c_int
        L.f32m_pow.argtypes = [ctypes.c_int, P, P, P, U64]
        L.f32m_pow.restype = ctypes.c_int
        L.f32m_sweep.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, P, P, U64]
        L.f32m_sweep.restype = ctypes.c_int
        _lib = L
    return _lib


def evaluate(where, fn, x):
    """where: "host" (the routines of f32math.h), "device", "libm" (f32 libm)
    or "f64" (the f64 libm function rounded once)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    rc = getattr(lib(), f"f32m_eval_{where}")(FN[fn], x.ctypes.data, y.ctypes.data, len(x))
    if rc != 0:
        raise RuntimeError(f"f32m_eval_{where}({fn}) returned {rc}")
    return y


def power(where, x, y):
    """pow over pairs; where: "host", "device", "f64" or "libm"."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(np.broadcast_to(np.asarray(y, np.float32), x.shape))
    z = np.empty_like(x)
    rc = lib().f32m_pow({"host": 0, "device": 1, "f64": 2, "libm": 3}[where], x.ctypes.data, y.ctypes.data,
                        z.ctypes.data, len(x))
    if rc != 0:
        raise RuntimeError(f"f32m_pow({where}) returned {rc}")
    return z


def sweep(fn, lo, hi, threads=None, cap=4096):
    """Every float with bit pattern in [lo, hi]: (inputs, differences from the
    f64 function rounded once, differences from the f32 libm function, the
    inputs of the first kind)."""
    threads = threads or max(1, min(8, os.cpu_count() or 1))
    counts = np.zeros(3, np.uint64)
    out = np.zeros(cap, np.uint32)
    rc = lib().f32m_sweep(FN[fn], lo, hi, threads, counts.ctypes.data, out.ctypes.data, cap)
    if rc != 0:
        raise RuntimeError(f"f32m_sweep({fn}) returned {rc}")
    n_bad = int(counts[1])
    return int(counts[0]), n_bad, int(counts[2]), out[:min(n_bad, cap)].view(np.float32).copy()


def bits(x):
    return int(np.float32(x).view(np.uint32))


def ulp_distance(a, b):
    """|a - b| in units in the last place (ordered bit patterns); NaN == NaN."""
    def ordered(v):
        i = np.asarray(v, np.float32).view(np.int32).astype(np.int64)
        return np.where(i < 0, -(i & 0x7FFFFFFF), i)
    d = np.abs(ordered(a) - ordered(b))
    return np.where(np.isnan(a) & np.isnan(b), 0, d)


def midpoint_distance(fn, x, got, other):
    """For inputs where the routine (`got`) and the f64 function rounded once
    (`other`) differ: the exact value's distance from the midpoint of the two
    candidates, relative to the value (mpmath, 120 bits), and whether `got`
    is the nearer one."""
    import mpmath as M
    M.mp.prec = 120
    f = {"exp": M.exp, "ln": M.log, "atan": M.atan, "sin": M.sin, "cos": M.cos}[fn]
    out = []
    for xi, gi, oi in zip(x, got, other):
        ex = f(M.mpf(float(xi)))
        mid = (M.mpf(float(gi)) + M.mpf(float(oi))) / 2
        out.append((float(abs(ex - mid) / abs(ex)), abs(M.mpf(float(gi)) - ex) < abs(M.mpf(float(oi)) - ex)))
    return out
