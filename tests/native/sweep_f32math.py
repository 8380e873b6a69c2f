"""Every f32 input of each routine's domain through the host build of the
microfacet material's f32 math (voidray_amd/csrc/f32math.h): counts the
results that differ from the f64 libm function rounded once and from the f32
libm function, and has mpmath (120 bits) decide each difference of the first
kind.  About a minute on 8 cores.

    python -m tests.native.sweep_f32math
"""
import numpy as np

from tests.native import f32math as F

f32 = np.float32

DOMAINS = {
    "exp": [(f32(-0.0), f32(-104.0)), (f32(0.0), f32(89.0))],
    "ln": [(f32(1e-45), f32(3.4028235e38))],
    "atan": [(f32(0.0), f32(3.4028235e38))],
    "sin": [(f32(0.0), f32(2.0))],
    "cos": [(f32(0.0), f32(2.0))],
}


def main():
    print("| routine | inputs | differ from f64 rounded once | of which the routine is nearer | "
          "largest distance of such an exact value from a midpoint | differ from f32 libm |")
    print("|---|---|---|---|---|---|")
    for fn, ranges in DOMAINS.items():
        n = n64 = nlibm = 0
        bad = []
        for lo, hi in ranges:
            a, b = sorted((F.bits(lo), F.bits(hi)))
            c = F.sweep(fn, a, b, cap=1 << 16)
            n, n64, nlibm = n + c[0], n64 + c[1], nlibm + c[2]
            bad.append(c[3])
        bad = np.concatenate(bad)
        got, ref = F.evaluate("host", fn, bad), F.evaluate("f64", fn, bad)
        md = F.midpoint_distance(fn, bad, got, ref)
        nearer = sum(1 for _, ok in md if ok)
        worst = max((rel for rel, ok in md if not ok), default=0.0)
        print(f"| {fn} | {n:,} | {n64} | {nearer} | {worst:.2e} | {nlibm:,} ({nlibm / n:.3%}) |", flush=True)


if __name__ == "__main__":
    main()
