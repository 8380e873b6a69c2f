// f32 exp, ln, pow, sin, cos and atan computed in f32 arithmetic only, for
// the MicrofacetBSDF material, whose reference calls Rust's f32::exp, ln, sin,
// cos and atan (glibc's expf, logf, sinf, cosf, atanf on Linux), and for the
// tonemap pass (GLSL pow and exp).
//
// Each function carries the value as a pair of floats (hi + lo, about 48 bits)
// through range reduction, table lookup and polynomial, and rounds once at the
// end: the result is the correctly rounded f32 of the exact value except where
// that value lies within about 2^-40 relative of a rounding midpoint.  The
// device's own f32 expf returns 0 for results in the denormal range, where the
// reference's expf returns the denormal: exp_cr rounds those once, too.
//
// Operations: +, -, *, / and sqrt as IEEE f32 (no contraction), fmaf as one
// fused operation, rintf with ties to even, so that the host and the device
// compile this header to the same results bit for bit.
#pragma once
#include "vrmath.h"

namespace vr {
namespace f32m {

struct ff {
    float hi, lo;
};

// s + e == a + b exactly (Knuth).
VR_HD ff two_sum(float a, float b) {
    const float s = a + b;
    const float bb = s - a;
    return ff{s, (a - (s - bb)) + (b - bb)};
}
// s + e == a + b exactly, given |a| >= |b| or a == 0 (Dekker).
VR_HD ff fast_two_sum(float a, float b) {
    const float s = a + b;
    return ff{s, b - (s - a)};
}
// p + e == a * b exactly (barring underflow).
VR_HD ff two_prod(float a, float b) {
    const float p = a * b;
    return ff{p, fmaf(a, b, -p)};
}
// 2^n for n in [-126, 127].
VR_HD float pow2i(int n) { return u2f((uint32_t)(n + 127) << 23); }

// 2^(j/32), j = 0..31, as hi + lo.
constexpr float kExp2Tab[32][2] = {
    {0x1p+0f, 0.0f},
    {0x1.059b0ep+0f, -0x1.9d4f52p-25f},
    {0x1.0b5586p+0f, 0x1.9f3122p-25f},
    {0x1.11301ep+0f, -0x1.fdb496p-25f},
    {0x1.172b84p+0f, -0x1.c15742p-27f},
    {0x1.1d4874p+0f, -0x1.d2e8cap-25f},
    {0x1.2387a6p+0f, 0x1.ceac48p-25f},
    {0x1.29e9ep+0f, -0x1.5c0424p-25f},
    {0x1.306fep+0f, 0x1.4636e2p-25f},
    {0x1.371a74p+0f, -0x1.18aac6p-25f},
    {0x1.3dea64p+0f, 0x1.824684p-25f},
    {0x1.44e086p+0f, 0x1.8624b4p-30f},
    {0x1.4bfdaep+0f, -0x1.593abcp-25f},
    {0x1.5342b6p+0f, -0x1.2c561p-25f},
    {0x1.5ab07ep+0f, -0x1.5bd5ecp-27f},
    {0x1.6247ecp+0f, -0x1.f8b55p-25f},
    {0x1.6a09e6p+0f, 0x1.9fcef4p-26f},
    {0x1.71f75ep+0f, 0x1.1d8beep-25f},
    {0x1.7a1148p+0f, -0x1.829fdp-25f},
    {0x1.82589ap+0f, -0x1.accc7cp-26f},
    {0x1.8ace54p+0f, 0x1.15506ep-27f},
    {0x1.93737cp+0f, -0x1.e64744p-25f},
    {0x1.9c4918p+0f, 0x1.51f848p-27f},
    {0x1.a5503cp+0f, -0x1.b83b54p-25f},
    {0x1.ae89fap+0f, -0x1.a94b14p-26f},
    {0x1.b7f77p+0f, -0x1.a09438p-25f},
    {0x1.c199bep+0f, -0x1.3d56b2p-27f},
    {0x1.cb720ep+0f, -0x1.8837ccp-27f},
    {0x1.d5818ep+0f, -0x1.822dbcp-27f},
    {0x1.dfc974p+0f, -0x1.908c94p-25f},
    {0x1.ea4afap+0f, 0x1.52486cp-27f},
    {0x1.f50766p+0f, -0x1.246ebp-26f},
};

// e^x.  x = (32n + j) ln2/32 + r with |r| <= ln2/64, where ln2/32 is split
// into L1 + L2 + L3 (L1, L2 of 9 significant bits, so k·L1 and k·L2 are exact
// for |k| < 2^15); e^x = 2^n · 2^(j/32) · e^r, e^r by its Taylor series to r^5.
// e^(x + xlo) for x in [-104, 89] and |xlo| <= ulp(x)/2 (xlo = 0 for a
// plain f32 argument; pow_cr passes the low word of y·ln x).
VR_HD float exp_core(float x, float xlo) {
    constexpr float kInvL = 0x1.715476p+5f;  // 32/ln2
    constexpr float L1 = 0x1.63p-6f, L2 = -0x1.bdp-18f, L3 = -0x1.05c61p-34f;
    const float kf = rintf(x * kInvL);
    const int k = (int)kf;
    const float r1 = x - kf * L1;  // exact: kf·L1 within a factor 2 of x
    ff r = two_sum(r1, -(kf * L2));
    r.lo = (r.lo + xlo) - kf * L3;
    r = fast_two_sum(r.hi, r.lo);
    // e^r - 1 = r + r^2/2 + r^3 (1/6 + r/24 + r^2/120), r^2/2 kept exact
    const ff s = two_prod(r.hi, r.hi);
    const float c3 = (s.hi * r.hi) * (0x1.555556p-3f + r.hi * (0x1.555556p-5f + r.hi * 0x1.111112p-7f));
    const ff e1 = fast_two_sum(r.hi, 0.5f * s.hi);
    const float e1lo = e1.lo + (r.lo + (0.5f * s.lo + (r.hi * r.lo + c3)));
    // 2^(j/32) (1 + em1)
    const int j = k & 31, n = k >> 5;
    const float th = kExp2Tab[j][0], tl = kExp2Tab[j][1];
    const ff p = two_prod(th, e1.hi);
    const ff a = fast_two_sum(th, p.hi);
    const float tail = a.lo + (p.lo + (th * e1lo + (tl + tl * (e1.hi + e1lo))));
    const ff y = fast_two_sum(a.hi, tail);
    if (n >= -125) {  // normal result (or overflow): y.hi · 2^n rounds at most once
        const int n1 = n / 2;
        return (y.hi * pow2i(n1)) * pow2i(n - n1);
    }
    // denormal result: y.hi · 2^n rounds to the 2^-149 grid; where y.hi sits
    // exactly on a midpoint of that grid, y.lo decides the direction
    const float q = (y.hi * pow2i(n + 64)) * 0x1p-64f;
    const float back = (q * 0x1p+64f) * pow2i(-n - 64);  // exact
    const float diff = y.hi - back;                        // exact
    const float half = pow2i(-150 - n);
    if (diff == half && y.lo > 0.0f) return q + 0x1p-149f;
    if (diff == -half && y.lo < 0.0f) return q - 0x1p-149f;
    return q;
}

VR_HD float exp_cr(float x) {
    if (!(x == x)) return x + x;
    if (x > 89.0f) return __builtin_huge_valf();
    if (x < -104.0f) return 0.0f;
    return exp_core(x, 0.0f);
}

// ln(i/16), i = 11..23, as hi + lo.
constexpr float kLnTab[13][2] = {
    {-0x1.7fafa4p-2f, 0x1.09fabap-28f},
    {-0x1.269622p-2f, 0x1.d9648ep-27f},
    {-0x1.a93ed4p-3f, 0x1.ba930ep-30f},
    {-0x1.1178e8p-3f, -0x1.13f23ep-30f},
    {-0x1.08598cp-4f, 0x1.4c38cp-29f},
    {0.0f, 0.0f},
    {0x1.f0a30cp-5f, 0x1.162a66p-37f},
    {0x1.e27076p-4f, 0x1.c55e5cp-29f},
    {0x1.5ff308p-3f, -0x1.eb0d86p-28f},
    {0x1.c8ff7cp-3f, 0x1.e6a688p-29f},
    {0x1.1675cap-2f, 0x1.7574c2p-27f},
    {0x1.4618bcp-2f, 0x1.0e2f62p-29f},
    {0x1.739d8p-2f, -0x1.2886p-27f},
};

// ln x.  x = 2^e · m with m in [sqrt(1/2), sqrt(2)), c = i/16 the nearest
// sixteenth to m: ln x = e·ln2 + ln c + 2 atanh(f), f = (m - c)/(m + c),
// |f| <= 1/45; 2 atanh f = 2f + (2/3) f^3 + f^5 (2/5 + (2/7) f^2), the first
// two terms as hi + lo.  ln2 = LN2_1 + LN2_2 + LN2_3, the first two of 15
// significant bits so that e·LN2_1 and e·LN2_2 are exact.
// ln x as hi + lo (not renormalised) for finite x > 0.
VR_HD ff ln_ff(float x) {
    int e = 0;
    if (x < 0x1p-126f) {
        x = x * 0x1p+24f;
        e = -24;
    }
    uint32_t u = f2u(x);
    e += (int)(u >> 23) - 127;
    u = (u & 0x007FFFFFu) | 0x3F800000u;  // m in [1, 2)
    if (u > 0x3FB504F3u) {                 // m > sqrt(2): halve it
        u -= 0x00800000u;
        e += 1;
    }
    const float m = u2f(u);
    const int i = (int)rintf(m * 16.0f);  // 11..23
    const float c = (float)i * 0.0625f;
    const float num = m - c;  // exact: m within a factor 2 of c
    const ff den = two_sum(m, c);
    const float fh = num / den.hi;
    const float rho = fmaf(-fh, den.hi, num) - fh * den.lo;
    const float fl = rho / den.hi;
    // (2/3) f^3 as hi + lo
    const ff f2 = two_prod(fh, fh);
    ff f3 = two_prod(f2.hi, fh);
    f3.lo = f3.lo + (f2.lo * fh + 3.0f * f2.hi * fl);
    ff a = two_prod(f3.hi, 0x1.555556p-1f);
    a.lo = a.lo + (f3.lo * 0x1.555556p-1f + f3.hi * -0x1.555556p-26f);
    const float b = (f3.hi * f2.hi) * (0x1.99999ap-2f + f2.hi * 0x1.24924ap-2f);
    const float ch = kLnTab[i - 11][0], cl = kLnTab[i - 11][1];
    const ff s1 = two_sum(ch, 2.0f * fh);
    const ff s2 = two_sum(s1.hi, a.hi);
    const float lo = (s1.lo + s2.lo) + (cl + (2.0f * fl + (a.lo + b)));
    if (e == 0) return ff{s2.hi, lo};
    constexpr float LN2_1 = 0x1.62e4p-1f, LN2_2 = 0x1.7f7cp-20f, LN2_3 = 0x1.1cf79ap-36f;
    const float ef = (float)e;
    const ff s3 = two_sum(ef * LN2_1, s2.hi);
    const ff s4 = two_sum(s3.hi, ef * LN2_2);
    return ff{s4.hi, (s3.lo + s4.lo) + (lo + ef * LN2_3)};
}

VR_HD float ln_cr(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_huge_valf() : __builtin_nanf("");
    if (x == __builtin_huge_valf()) return x;
    const ff r = ln_ff(x);
    return r.hi + r.lo;
}

// x^y with C's powf special cases: pow(x, ±0) = 1 and pow(1, y) = 1 even for
// NaN; a negative finite x gives NaN unless y is an integer, when the sign
// follows y's parity; zeros and infinities as C99 Annex F.  Otherwise
// e^(y·ln|x|) with y·ln|x| carried as hi + lo (about 2^-40 relative, so the
// result is correctly rounded but within about 2^-40·|y ln x| of a midpoint).
VR_HD float pow_cr(float x, float y) {
    const float inf = __builtin_huge_valf();
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (!(x == x) || !(y == y)) return x + y;
    const float ay = fabsf(y);
    const bool yint = ay >= 0x1p+24f || truncf(y) == y;
    const bool yodd = ay < 0x1p+24f && yint && (((int64_t)y) & 1) != 0;
    if (ay == inf) {
        const float ax = fabsf(x);
        if (ax == 1.0f) return 1.0f;  // pow(-1, ±inf)
        return (ax < 1.0f) == (y > 0.0f) ? 0.0f : inf;
    }
    const bool neg = (f2u(x) >> 31) != 0;  // -0 included
    if (x == 0.0f) {
        if (y < 0.0f) return yodd && neg ? -inf : inf;
        return yodd && neg ? -0.0f : 0.0f;
    }
    if (fabsf(x) == inf) {
        const float mag = y < 0.0f ? 0.0f : inf;
        return yodd && neg ? -mag : mag;
    }
    if (neg && !yint) return __builtin_nanf("");
    const ff l = ln_ff(fabsf(x));
    ff t = two_prod(y, l.hi);
    t.lo = t.lo + y * l.lo;
    t = fast_two_sum(t.hi, t.lo);
    float m;
    if (t.hi > 89.0f) m = inf;
    else if (t.hi < -104.0f) m = 0.0f;
    else m = exp_core(t.hi, t.lo);
    return neg && yodd ? -m : m;
}

// atan(j/16), j = 0..16, as hi + lo.
constexpr float kAtanTab[17][2] = {
    {0.0f, 0.0f},
    {0x1.ff55bcp-5f, -0x1.1a6042p-30f},
    {0x1.fd5baap-4f, -0x1.54f424p-30f},
    {0x1.7b97b4p-3f, 0x1.79cb6p-28f},
    {0x1.f5b76p-3f, -0x1.b4dfc8p-29f},
    {0x1.362774p-2f, -0x1.1f0286p-27f},
    {0x1.6f6194p-2f, 0x1.e4defp-30f},
    {0x1.a64eecp-2f, 0x1.e611fep-29f},
    {0x1.dac67p-2f, 0x1.586ed4p-28f},
    {0x1.0657eap-1f, -0x1.6499e6p-26f},
    {0x1.1e00bap-1f, 0x1.7bdfd6p-26f},
    {0x1.345f02p-1f, -0x1.98e422p-28f},
    {0x1.4978fap-1f, 0x1.934f7p-28f},
    {0x1.5d5898p-1f, 0x1.c5a6c6p-27f},
    {0x1.700a7cp-1f, 0x1.5e118cp-27f},
    {0x1.819d0cp-1f, -0x1.1d4eb6p-26f},
    {0x1.921fb6p-1f, -0x1.777a5cp-26f},
};

// atan x.  For |x| > 1, atan |x| = pi/2 - atan(1/|x|); y in [0, 1] is then
// reduced by c = j/16: atan y = atan c + atan r, r = (y - c)/(1 + y c),
// |r| <= 1/32, atan r = r - r^3/3 + r^5/5 - r^7/7 + r^9/9.
VR_HD float atan_cr(float x) {
    if (!(x == x)) return x + x;
    const bool neg = (f2u(x) >> 31) != 0;  // -0 too: atan(-0) = -0
    const float ax = fabsf(x);
    float res;
    if (ax > 0x1p+26f) {
        res = 0x1.921fb6p+0f;  // pi/2 rounded; pi/2 - 1/x rounds to it here
    } else if (ax < 0x1p-12f) {
        res = ax;  // x - x^3/3 rounds to x here
    } else {
        const bool inv = ax > 1.0f;
        ff y{ax, 0.0f};
        if (inv) {
            y.hi = 1.0f / ax;
            y.lo = fmaf(-y.hi, ax, 1.0f) / ax;
        }
        const int j = (int)rintf(y.hi * 16.0f);
        const float c = (float)j * 0.0625f;
        const ff num = two_sum(y.hi - c, y.lo);  // y.hi - c exact
        const ff yc = two_prod(y.hi, c);
        ff den = fast_two_sum(1.0f, yc.hi);
        den.lo = den.lo + (yc.lo + y.lo * c);
        const float rh = num.hi / den.hi;
        const float rho = fmaf(-rh, den.hi, num.hi) + (num.lo - rh * den.lo);
        const float rl = rho / den.hi;
        // r^3/3 as hi + lo
        const ff s = two_prod(rh, rh);
        ff r3 = two_prod(s.hi, rh);
        r3.lo = r3.lo + (s.lo * rh + 3.0f * s.hi * rl);
        ff t3 = two_prod(r3.hi, 0x1.555556p-2f);
        t3.lo = t3.lo + (r3.lo * 0x1.555556p-2f + r3.hi * -0x1.555556p-27f);
        const float z = s.hi;
        const float t5 = (r3.hi * z) * (0x1.99999ap-3f - z * (0x1.24924ap-3f - z * 0x1.c71c72p-4f));
        const ff h1 = two_sum(kAtanTab[j][0], rh);
        const ff h2 = two_sum(h1.hi, -t3.hi);
        const float rest = h1.lo + (h2.lo + (kAtanTab[j][1] + (rl + (t5 - t3.lo))));
        if (inv) {
            const ff d = two_sum(0x1.921fb6p+0f, -h2.hi);
            res = d.hi + (d.lo + (-0x1.777a5cp-25f - rest));
        } else {
            res = h2.hi + rest;
        }
    }
    return neg ? -res : res;
}

// sin and cos of N·pi/32, N = 0..21, as hi + lo: {sin hi, sin lo, cos hi, cos lo}.
constexpr float kSinCosTab[22][4] = {
    {0.0f, 0.0f, 0x1p+0f, 0.0f},
    {0x1.917a6cp-4f, -0x1.eb25eap-31f, 0x1.fd88dap-1f, 0x1.e89292p-28f},
    {0x1.8f8b84p-3f, -0x1.cb2cfap-30f, 0x1.f6297cp-1f, 0x1.feeb96p-26f},
    {0x1.294062p-2f, 0x1.dab3ep-27f, 0x1.e9f416p-1f, -0x1.273a44p-26f},
    {0x1.87de2ap-2f, 0x1.abaa58p-28f, 0x1.d906bcp-1f, 0x1.e651a8p-26f},
    {0x1.e2b5d4p-2f, -0x1.fe4272p-28f, 0x1.c38b3p-1f, -0x1.cfe84ap-26f},
    {0x1.1c73b4p-1f, -0x1.9465cep-27f, 0x1.a9b662p-1f, 0x1.21d434p-26f},
    {0x1.44cf32p-1f, 0x1.424776p-27f, 0x1.8bc806p-1f, 0x1.62a2e8p-26f},
    {0x1.6a09e6p-1f, 0x1.9fcef4p-27f, 0x1.6a09e6p-1f, 0x1.9fcef4p-27f},
    {0x1.8bc806p-1f, 0x1.62a2e8p-26f, 0x1.44cf32p-1f, 0x1.424776p-27f},
    {0x1.a9b662p-1f, 0x1.21d434p-26f, 0x1.1c73b4p-1f, -0x1.9465cep-27f},
    {0x1.c38b3p-1f, -0x1.cfe84ap-26f, 0x1.e2b5d4p-2f, -0x1.fe4272p-28f},
    {0x1.d906bcp-1f, 0x1.e651a8p-26f, 0x1.87de2ap-2f, 0x1.abaa58p-28f},
    {0x1.e9f416p-1f, -0x1.273a44p-26f, 0x1.294062p-2f, 0x1.dab3ep-27f},
    {0x1.f6297cp-1f, 0x1.feeb96p-26f, 0x1.8f8b84p-3f, -0x1.cb2cfap-30f},
    {0x1.fd88dap-1f, 0x1.e89292p-28f, 0x1.917a6cp-4f, -0x1.eb25eap-31f},
    {0x1p+0f, 0.0f, 0.0f, 0.0f},
    {0x1.fd88dap-1f, 0x1.e89292p-28f, -0x1.917a6cp-4f, 0x1.eb25eap-31f},
    {0x1.f6297cp-1f, 0x1.feeb96p-26f, -0x1.8f8b84p-3f, 0x1.cb2cfap-30f},
    {0x1.e9f416p-1f, -0x1.273a44p-26f, -0x1.294062p-2f, -0x1.dab3ep-27f},
    {0x1.d906bcp-1f, 0x1.e651a8p-26f, -0x1.87de2ap-2f, -0x1.abaa58p-28f},
    {0x1.c38b3p-1f, -0x1.cfe84ap-26f, -0x1.e2b5d4p-2f, 0x1.fe4272p-28f},
};

// sin and cos of x for |x| <= 2 (the callers pass atan results in
// [0, pi/2]; outside that range both are NaN).  |x| = N·pi/32 + r with
// pi/32 = P1 + P2 + P3 + P4 (P1, P2 of 19 significant bits, so N·P1 and N·P2
// are exact for N < 32; 86 bits in all, enough for x near pi/2, the closest
// f32 to which is 4.4e-8 away); then
//   sin x = S + S (cos r - 1) + C r + C (sin r - r),
//   cos x = C + C (cos r - 1) - S r - S (sin r - r),
// with S, C = sin, cos (N·pi/32), r^2/2 and r^3/6 kept as hi + lo.
VR_HD void sincos_cr(float x, float& sn, float& cs) {
    const float ax = fabsf(x);
    if (!(ax <= 2.0f)) {
        sn = cs = __builtin_nanf("");
        return;
    }
    constexpr float P1 = 0x1.921fcp-4f, P2 = -0x1.5777cp-25f, P3 = 0x1.a308d4p-45f, P4 = -0x1.d9ccecp-70f;
    const float nf = rintf(ax * 0x1.45f306p+3f);  // 32/pi
    const int n = (int)nf;
    const float r1 = ax - nf * P1;  // exact
    ff r = two_sum(r1, -(nf * P2));
    const ff r3p = two_sum(r.hi, -(nf * P3));
    r = fast_two_sum(r3p.hi, (r.lo + r3p.lo) - nf * P4);
    // cos r - 1 = -r^2/2 + r^4 (1/24 - r^2/720 + r^4/40320)
    const ff s = two_prod(r.hi, r.hi);
    const float z = s.hi;
    const float cmh = -0.5f * s.hi;
    const float cml = -0.5f * s.lo - r.hi * r.lo + (z * z) * (0x1.555556p-5f - z * (0x1.6c16c2p-10f - z * 0x1.a01a02p-16f));
    // sin r - r = -r^3/6 + r^5 (1/120 - r^2/5040)
    ff r3 = two_prod(s.hi, r.hi);
    r3.lo = r3.lo + (s.lo * r.hi + 3.0f * s.hi * r.lo);
    ff t3 = two_prod(r3.hi, 0x1.555556p-3f);
    t3.lo = t3.lo + (r3.lo * 0x1.555556p-3f + r3.hi * -0x1.555556p-28f);
    const float smh = -t3.hi;
    const float sml = -t3.lo + (r3.hi * z) * (0x1.111112p-7f - z * 0x1.a01a02p-13f);
    const float Sh = kSinCosTab[n][0], Sl = kSinCosTab[n][1], Ch = kSinCosTab[n][2], Cl = kSinCosTab[n][3];
    {
        // sin: S + C r + S cm + C sm
        const ff b = two_prod(Ch, r.hi);
        const ff d = two_prod(Sh, cmh);
        const ff e = two_prod(Ch, smh);
        const ff x1 = two_sum(Sh, b.hi);
        const ff x2 = two_sum(x1.hi, d.hi);
        const ff x3 = two_sum(x2.hi, e.hi);
        const float lo = (x1.lo + x2.lo + x3.lo) + (Sl + (b.lo + Ch * r.lo + Cl * r.hi) +
                                                     (d.lo + Sh * cml + Sl * cmh) + (e.lo + Ch * sml + Cl * smh));
        sn = x3.hi + lo;
    }
    {
        // cos: C - S r + C cm - S sm
        const ff b = two_prod(-Sh, r.hi);
        const ff d = two_prod(Ch, cmh);
        const ff e = two_prod(-Sh, smh);
        const ff x1 = two_sum(Ch, b.hi);
        const ff x2 = two_sum(x1.hi, d.hi);
        const ff x3 = two_sum(x2.hi, e.hi);
        const float lo = (x1.lo + x2.lo + x3.lo) + (Cl + (b.lo - Sh * r.lo - Sl * r.hi) +
                                                     (d.lo + Ch * cml + Cl * cmh) + (e.lo - Sh * sml - Sl * smh));
        cs = x3.hi + lo;
    }
    if (f2u(x) >> 31) sn = -sn;  // sin(-0) = -0
}

}  // namespace f32m
}  // namespace vr
