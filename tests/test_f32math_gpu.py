"""The microfacet material's f32 exp, ln, atan, sin and cos
(voidray_amd/csrc/f32math.h) on the device: the same bits as the host build
of the same header, which tests/test_f32math.py pins against the f64 libm
functions rounded once and against the f32 libm functions.  Inputs: 2^21
bit-uniform floats over each domain, 2^21 consecutive floats at the places
where the routines change path (denormal results of exp, denormal inputs of
ln, pi/2 for sin and cos, the thresholds of atan), and the special values.
"""
import numpy as np
import pytest

from tests.conftest import announce
from tests.native import f32math as F

pytestmark = pytest.mark.gpu
f32 = np.float32

DOMAIN = {"exp": (-104.0, 89.0), "ln": (0.0, 3.4028235e38), "atan": (-3.4028235e38, 3.4028235e38),
          "sin": (-2.0, 2.0), "cos": (-2.0, 2.0)}
RUN_START = {"exp": -87.0, "ln": 1e-45, "atan": 6.0e7, "sin": 1.55, "cos": 1.55}
SPECIAL = [0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 2.0, 88.72284, -103.97208, 1e-45, 3.4028235e38,
           f32(np.pi / 2), 2.0 ** 26, 2.0 ** -12, 2.5]


def _inputs(fn, rng):
    lo, hi = DOMAIN[fn]
    parts = []
    for a, b in ((lo, min(hi, -0.0)), (max(lo, 0.0), hi)):
        if a >= b:
            continue
        ia, ib = sorted((F.bits(f32(a)), F.bits(f32(b))))
        parts.append(rng.integers(ia, ib + 1, 1 << 20, dtype=np.int64).astype(np.uint32))
    s = F.bits(f32(RUN_START[fn]))
    parts.append(np.arange(s, s + (1 << 21), dtype=np.uint32))
    parts.append(np.array(SPECIAL, np.float32).view(np.uint32))
    return np.concatenate(parts).view(np.float32)


@pytest.mark.parametrize("fn", ["exp", "ln", "atan", "sin", "cos"])
def test_device_equals_host(fn):
    rng = np.random.default_rng(21 + F.FN[fn])
    x = _inputs(fn, rng)
    announce(f"f32math: {fn} over {len(x)} inputs on the device")
    dev = F.evaluate("device", fn, x)
    host = F.evaluate("host", fn, x)
    same = (dev.view(np.uint32) == host.view(np.uint32)) | (np.isnan(dev) & np.isnan(host))
    bad = np.flatnonzero(~same)
    assert bad.size == 0, [(float(x[i]).hex(), float(dev[i]).hex(), float(host[i]).hex()) for i in bad[:8]]


def test_pow_device_equals_host():
    rng = np.random.default_rng(41)
    x = np.concatenate([rng.integers(F.bits(f32(1e-38)), F.bits(f32(1e6)), 1 << 20, dtype=np.int64)
                        .astype(np.uint32).view(np.float32), np.array(SPECIAL, np.float32)])
    y = np.concatenate([rng.choice(np.array([1 / 2.2, 1 / 1.8, 2.2, 1.0, 0.5, -1.5, 3.0], np.float32), 1 << 20),
                        np.array(SPECIAL, np.float32)[::-1]])
    announce(f"f32math: pow over {len(x)} pairs on the device")
    dev, host = F.power("device", x, y), F.power("host", x, y)
    same = (dev.view(np.uint32) == host.view(np.uint32)) | (np.isnan(dev) & np.isnan(host))
    bad = np.flatnonzero(~same)
    assert bad.size == 0, [(float(x[i]), float(y[i]), float(dev[i]), float(host[i])) for i in bad[:8]]
