"""The f32-only exp, ln, atan, sin and cos of the microfacet material and
the pow and exp of the tonemap pass (voidray_amd/csrc/f32math.h), evaluated
on the host through the test probe.

The reference's MicrofacetBSDF calls Rust's f32::exp, ln, atan, sin, cos
(voidray_renderer/src/material/microfacet.rs:140-143, 183-186, 245-250 and the
Beckmann sampling at 215-312), i.e. the platform libm's expf, logf, atanf,
sinf, cosf.  Pins, each against an independent implementation:
* the f64 libm function rounded once (correctly rounded but for double
  rounding): equal on windows of 2^22 consecutive floats at the places where
  a mistake would show (range-reduction boundaries, underflow to denormals and
  zero, overflow, the last table entries, pi/2) and on 2^20 bit-uniform
  inputs over each domain; where they differ, mpmath at 120 bits decides, and
  the routine may only lose a case whose exact value lies within 2^-40
  relative of a rounding midpoint;
* the f32 libm function: within one ulp everywhere (glibc's are not all
  correctly rounded: measured on every f32 input of each domain, atanf
  differs from the correctly rounded result on 0.5% of them);
* special values: zeros, infinities, NaN, the overflow and underflow
  thresholds.
"""
import numpy as np
import pytest

from tests.native import f32math as F

f32 = np.float32
INF = f32(np.inf)


def _check_window(fn, start, limit):
    """2^22 consecutive floats from `start` outwards (away from zero), up to
    the float `limit`."""
    a = F.bits(start)
    b = min(a + (1 << 22) - 1, F.bits(limit))
    n, n64, nlibm, bad = F.sweep(fn, a, b)
    assert n == b - a + 1
    if n64:
        got = F.evaluate("host", fn, bad)
        ref = F.evaluate("f64", fn, bad)
        assert np.all(F.ulp_distance(got, ref) == 1)
        for rel, nearer in F.midpoint_distance(fn, bad, got, ref):
            # either the f64 path double-rounded (the routine is right) or the
            # exact value is a hard case within 2^-40 of a midpoint
            assert nearer or rel < 2.0 ** -40, (fn, rel)
    return n, n64, nlibm


# window starts, each followed by 2^22 floats: [start, end] below
WINDOWS = {
    # the first reduction boundary ln2/64; all of the denormal range and the
    # underflow to zero (-87 .. -119); the overflow (88 .. 120); mid-range
    "exp": [f32(0.0107), f32(-0.0107), f32(-87.0), f32(88.0), f32(-10.4), f32(1.0)],
    # the table ends at sqrt(1/2) and sqrt(2) (0.70 .. 0.95, 1.0 .. 1.5),
    # denormal inputs, small and the largest inputs
    "ln": [f32(0.70), f32(1.0), f32(1e-45), f32(1e-3), f32(3.0e38)],
    # the x-returning threshold 2^-12, the 1/x switch at 1 and the table's
    # last entry, the pi/2-returning threshold 2^26, |x| = 16, the table step
    # at 1/32
    "atan": [f32(2.0e-4), f32(0.97), f32(6.0e7), f32(15.9), f32(0.0305), f32(-0.97)],
    # pi/64 (the first table step), pi/2 and beyond, up to the domain's end
    "sin": [f32(0.045), f32(1.55), f32(1.75), f32(-1.55), f32(1e-4)],
    "cos": [f32(0.045), f32(1.55), f32(1.75), f32(-1.55), f32(1e-4)],
}


def _limit(fn, start):
    """The end of the domain on the side of `start`."""
    if fn == "exp":
        return f32(89.0) if start > 0 else f32(-104.0)
    top = f32(2.0) if fn in ("sin", "cos") else f32(3.4028235e38)
    return top if start > 0 else -top


@pytest.mark.parametrize("fn", list(WINDOWS))
def test_windows_are_correctly_rounded(fn):
    total = total64 = 0
    for start in WINDOWS[fn]:
        limit = _limit(fn, start)
        n, n64, _ = _check_window(fn, start, limit)
        total += n
        total64 += n64
    # the exhaustive sweep over each domain found 87 such cases in 7.6e9
    assert total64 <= 2, (fn, total64, total)
    print(f"[f32math] {fn}: {total} inputs in windows, {total64} differ from f64-rounded")


DOMAINS = {
    "exp": (f32(-104.0), f32(89.0)),
    "ln": (f32(1e-45), f32(3.4e38)),
    "atan": (f32(-3.4e38), f32(3.4e38)),
    "sin": (f32(-2.0), f32(2.0)),
    "cos": (f32(-2.0), f32(2.0)),
}


def _bit_uniform(lo, hi, n, rng):
    """Floats in [lo, hi] drawn uniformly over bit patterns, half from each
    sign's part of the range when it has both."""
    ranges = []
    if lo < 0:
        ranges.append((F.bits(f32(-0.0)), F.bits(lo)))
    if hi >= 0:
        ranges.append((F.bits(max(lo, f32(0.0))), F.bits(hi)))
    return np.concatenate([rng.integers(a, b + 1, n // len(ranges), dtype=np.int64).astype(np.uint32).view(np.float32)
                           for a, b in ranges])


@pytest.mark.parametrize("fn", list(DOMAINS))
def test_bit_uniform_inputs_match_f64_and_stay_within_an_ulp_of_libm(fn):
    rng = np.random.default_rng(11 + F.FN[fn])
    lo, hi = DOMAINS[fn]
    x = _bit_uniform(lo, hi, 1 << 20, rng)
    got = F.evaluate("host", fn, x)
    ref = F.evaluate("f64", fn, x)
    lm = F.evaluate("libm", fn, x)
    d64 = F.ulp_distance(got, ref)
    assert d64.max() <= 1 and np.count_nonzero(d64) <= 2, (fn, np.count_nonzero(d64))
    assert F.ulp_distance(got, lm).max() <= 1
    print(f"[f32math] {fn}: {len(x)} inputs, {np.count_nonzero(d64)} differ from f64-rounded, "
          f"{np.count_nonzero(F.ulp_distance(got, lm))} from libm (each by one ulp)")


def test_special_values():
    ev = lambda fn, v: F.evaluate("host", fn, np.array(v, np.float32))
    nan = f32(np.nan)
    e = ev("exp", [0.0, -0.0, -INF, INF, nan, 1.0, -1.0])
    assert e[0] == 1 and e[1] == 1 and e[2] == 0 and e[3] == INF and np.isnan(e[4])
    assert e[5] == f32(np.e) and e[6] == f32(1 / np.e)
    # overflow: the largest finite result and the first infinite one
    big = np.array([f32(88.72283), f32(88.72284), np.nextafter(f32(88.72284), INF)], np.float32)
    assert np.array_equal(ev("exp", big), F.evaluate("f64", "exp", big))
    assert np.isinf(ev("exp", [89.0])[0]) and not np.isinf(ev("exp", [88.72283])[0])
    # the smallest denormal and the last input that rounds to it
    tiny = ev("exp", [-103.27893, -103.97207, -103.97208, -104.0])
    assert tiny[0] > 0 and tiny[-1] == 0
    assert np.array_equal(tiny, F.evaluate("f64", "exp", np.array([-103.27893, -103.97207, -103.97208, -104.0], f32)))
    ln = ev("ln", [1.0, 0.0, -0.0, -1.0, INF, nan, 1e-45, 3.4028235e38, 2.0])
    assert ln[0] == 0 and ln[1] == -INF and ln[2] == -INF and np.isnan(ln[3]) and ln[4] == INF and np.isnan(ln[5])
    assert ln[6] == f32(np.log(np.float64(f32(1e-45)))) and ln[7] == f32(np.log(3.4028235e38))
    assert ln[8] == f32(np.log(2.0))
    at = ev("atan", [0.0, -0.0, INF, -INF, nan, 1.0, 2.0 ** 26, 2.0 ** -13])
    assert at[0] == 0 and np.signbit(at[1]) and at[2] == f32(np.pi / 2) and at[3] == -f32(np.pi / 2)
    assert np.isnan(at[4]) and at[5] == f32(np.pi / 4) and at[6] == f32(np.pi / 2) and at[7] == f32(2.0 ** -13)
    s = ev("sin", [0.0, -0.0, f32(np.pi / 2), nan, 2.5])
    c = ev("cos", [0.0, -0.0, f32(np.pi / 2), nan, 2.5])
    assert s[0] == 0 and np.signbit(s[1]) and s[2] == 1 and np.isnan(s[3])
    assert c[0] == 1 and c[1] == 1 and c[2] == f32(np.cos(np.float64(f32(np.pi / 2)))) and np.isnan(c[3])
    # outside the documented domain |x| <= 2: NaN, not a wrong number
    assert np.isnan(s[4]) and np.isnan(c[4])


def test_odd_and_even_symmetry():
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 2, 1 << 16).astype(np.float32)
    for fn, sign in (("sin", -1), ("cos", 1)):
        assert np.array_equal(F.evaluate("host", fn, -x), sign * F.evaluate("host", fn, x))
    y = rng.uniform(0, 100, 1 << 16).astype(np.float32)
    assert np.array_equal(F.evaluate("host", "atan", -y), -F.evaluate("host", "atan", y))


def test_microfacet_sampling_chain():
    """The chain of the Beckmann sample, theta = atan(sqrt(m^2 · -ln U)),
    sin theta, cos theta, with the roughness of the parity scenes and U from
    [0, 1): every step's result equals the f64-rounded one."""
    rng = np.random.default_rng(9)
    u = rng.random(1 << 18, dtype=np.float32)
    for rough in (0.05, 0.25, 0.5, 1.0):
        m2 = f32(rough) * f32(rough)
        l = F.evaluate("host", "ln", u)
        assert np.array_equal(l, F.evaluate("f64", "ln", u))
        a = np.sqrt(m2 * -l).astype(np.float32)
        th = F.evaluate("host", "atan", a)
        assert F.ulp_distance(th, F.evaluate("f64", "atan", a)).max() <= 1
        assert np.count_nonzero(th != F.evaluate("f64", "atan", a)) <= 1
        for fn in ("sin", "cos"):
            assert np.count_nonzero(F.evaluate("host", fn, th) != F.evaluate("f64", fn, th)) <= 1


def test_pow_matches_f64_and_libm():
    """pow_cr (the tonemap's GLSL pow: pow(c, 1/gamma), pow(2, exposure)) over
    bit-uniform bases with the gammas and exponents the tonemap uses and over
    random pairs: within one ulp of the f64 pow rounded once and of powf, and
    equal to the f64 result on all but a few inputs."""
    rng = np.random.default_rng(31)
    lo, hi = F.bits(f32(1e-30)), F.bits(f32(1e6))
    for y in (1 / 2.2, 1 / 1.8, 1 / 1.5, 2.2, 1.0, 0.5, 3.0, -1.5, 7.3):
        x = rng.integers(lo, hi, 1 << 18, dtype=np.int64).astype(np.uint32).view(np.float32)
        got = F.power("host", x, y)
        d64 = F.ulp_distance(got, F.power("f64", x, y))
        assert d64.max() <= 1 and np.count_nonzero(d64) <= 3, (y, np.count_nonzero(d64))
        assert F.ulp_distance(got, F.power("libm", x, y)).max() <= 1
    x = rng.integers(F.bits(f32(1e-3)), F.bits(f32(1e3)), 1 << 18, dtype=np.int64).astype(np.uint32).view(np.float32)
    y = rng.uniform(-12, 12, 1 << 18).astype(np.float32)
    got = F.power("host", x, y)
    d64 = F.ulp_distance(got, F.power("f64", x, y))
    assert d64.max() <= 1 and np.count_nonzero(d64) <= 3
    # exposure: pow(2, e) for the exposures of the tonemap tests and demos
    e = np.linspace(-8, 8, 4097, dtype=np.float32)
    assert np.array_equal(F.power("host", np.full_like(e, 2.0), e), F.power("f64", np.full_like(e, 2.0), e))


def test_pow_special_values_as_powf():
    """Zeros, ones, infinities, NaN, negative bases with integer and
    non-integer exponents, denormals: the same results as powf, sign of zero
    and infinity included."""
    sp = np.array([0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 0.5, -0.5, 3.0, -3.0, np.inf, -np.inf, np.nan,
                   1e-40, -1e-40, 1e30, 2.0 ** 24, 2.0 ** 24 + 2, 0.3], np.float32)
    X, Y = (a.ravel() for a in np.meshgrid(sp, sp))
    got, want = F.power("host", X, Y), F.power("libm", X, Y)
    same = (np.isnan(got) & np.isnan(want)) | ((got == want) & (np.signbit(got) == np.signbit(want)))
    assert same.all(), [(float(a), float(b), float(c), float(d)) for a, b, c, d, k in zip(X, Y, got, want, same) if not k][:8]
