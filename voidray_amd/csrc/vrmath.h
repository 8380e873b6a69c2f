// Exact-order f32 vector arithmetic shared by host set-up code and gfx950
// kernels.  Every decision-bearing expression follows the operation order of
// the reference (cgmath 0.18 as used by voidray_renderer/src/util/vector.rs):
// dot = (x*x' + y*y') + z*z', cross = (y z' - z y', z x' - x z', x y' - y x'),
// normalize = v * (1 / sqrt(dot(v, v))).  The library is compiled with
// -ffp-contract=off so none of these products is fused; the few places that
// want an FMA (box slabs, which only need to be conservative) call fmaf.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define VR_HD __host__ __device__ __forceinline__

struct vf3 {
    float x, y, z;
};
struct vf2 {
    float x, y;
};

VR_HD vf3 mk3(float x, float y, float z) { return vf3{x, y, z}; }
VR_HD vf3 operator+(vf3 a, vf3 b) { return vf3{a.x + b.x, a.y + b.y, a.z + b.z}; }
VR_HD vf3 operator-(vf3 a, vf3 b) { return vf3{a.x - b.x, a.y - b.y, a.z - b.z}; }
VR_HD vf3 operator-(vf3 a) { return vf3{-a.x, -a.y, -a.z}; }
VR_HD vf3 operator*(vf3 a, float s) { return vf3{a.x * s, a.y * s, a.z * s}; }
VR_HD vf3 operator*(float s, vf3 a) { return vf3{s * a.x, s * a.y, s * a.z}; }
VR_HD vf3 operator/(vf3 a, float s) { return vf3{a.x / s, a.y / s, a.z / s}; }
VR_HD vf3 mulv(vf3 a, vf3 b) { return vf3{a.x * b.x, a.y * b.y, a.z * b.z}; }
VR_HD float dot3(vf3 a, vf3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
VR_HD vf3 cross3(vf3 a, vf3 b) {
    return vf3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
VR_HD float mag2(vf3 a) { return dot3(a, a); }
VR_HD float mag(vf3 a) { return sqrtf(mag2(a)); }
VR_HD vf3 normalize3(vf3 a) { return a * (1.0f / mag(a)); }
VR_HD vf3 vminc(vf3 a, float c) { return vf3{fminf(a.x, c), fminf(a.y, c), fminf(a.z, c)}; }
VR_HD float comp(vf3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Clang native vectors for wide loads/stores (plain .x/.y/.z/.w element
// access, no wrapper classes).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

VR_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
VR_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

static constexpr float kPi = 3.14159265358979323846f;
static constexpr float kInf = __builtin_huge_valf();

// ---- Philox4x32-10 (Salmon et al., SC'11) --------------------------------
VR_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                         uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Per-path counter-based stream: word k is lane k%4 of
// Philox(counter = (k/4, sample, pixel, request), key = seed).  The state that
// survives between kernels is just the word counter `n`.
struct PathRng {
    uint32_t k0, k1, sample, pixel, request, n, cached;
    uint32_t block[4];
    VR_HD void init(uint64_t seed, uint32_t request_, uint32_t pixel_, uint32_t sample_, uint32_t n_) {
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32);
        request = request_;
        pixel = pixel_;
        sample = sample_;
        n = n_;
        cached = 0xFFFFFFFFu;
    }
    VR_HD uint32_t next_u32() {
        uint32_t b = n >> 2;
        if (b != cached) {
            philox4x32_10(b, sample, pixel, request, k0, k1, block);
            cached = b;
        }
        uint32_t w = n & 3u;
        ++n;
        return w == 0 ? block[0] : (w == 1 ? block[1] : (w == 2 ? block[2] : block[3]));
    }
    VR_HD float uniform() { return (float)(next_u32() >> 8) * 5.9604644775390625e-08f; }
};

// ---- axis-aligned boxes ----------------------------------------------------
struct Box3 {
    vf3 lo, hi;
};

// AABB::default() (inf, -inf), compared as bit patterns (each infinity has
// one encoding, so this equals comparing the floats).  Written as six float
// compares, the optimiser vectorised them into a <4 x i1> compare whose
// bitcast to an integer the gfx950 backend lowered through v_readfirstlane:
// one lane's answer for the whole wave.  The empty asm keeps the OR of the
// differences opaque, so it stays one integer test per lane (not volatile:
// the volatile form cost the closest-hit kernel 1.3%, this one 0.3%).
VR_HD bool box_is_default(const Box3& b) {
    const uint32_t pinf = 0x7F800000u, ninf = 0xFF800000u;
    uint32_t diff = (f2u(b.lo.x) ^ pinf) | (f2u(b.lo.y) ^ pinf) | (f2u(b.lo.z) ^ pinf) | (f2u(b.hi.x) ^ ninf) |
                    (f2u(b.hi.y) ^ ninf) | (f2u(b.hi.z) ^ ninf);
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(diff));
#endif
    return diff == 0u;
}

// AABB::hit (voidray_renderer/src/util/aabb.rs:86-148), same operations.
VR_HD bool ref_slab_hit(const Box3& b, vf3 o, vf3 d, float t_min, float t_max) {
    if (box_is_default(b)) return true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float inv_d = 1.0f / comp(d, a);
        float t0 = (comp(b.lo, a) - comp(o, a)) * inv_d;
        float t1 = (comp(b.hi, a) - comp(o, a)) * inv_d;
        if (inv_d < 0.0f) {
            float tmp = t0;
            t0 = t1;
            t1 = tmp;
        }
        if (t0 > t_min) t_min = t0;
        if (t1 < t_max) t_max = t1;
        if (t_max <= t_min) return false;
    }
    return true;
}

// AABB::epsilon_expand (aabb.rs:62-83) of the triangle box (mesh.rs:193-205).
VR_HD Box3 tri_box_eps(vf3 p0, vf3 p1, vf3 p2) {
    Box3 b;
    b.lo = mk3(fminf(p0.x, fminf(p1.x, p2.x)), fminf(p0.y, fminf(p1.y, p2.y)), fminf(p0.z, fminf(p1.z, p2.z)));
    b.hi = mk3(fmaxf(p0.x, fmaxf(p1.x, p2.x)), fmaxf(p0.y, fmaxf(p1.y, p2.y)), fmaxf(p0.z, fmaxf(p1.z, p2.z)));
    const float eps = 0.001f;
    vf3 dim = b.hi - b.lo;
    vf3 c = (b.lo + b.hi) / 2.0f;
    if (dim.x < eps) { b.lo.x = c.x - eps; b.hi.x = c.x + eps; }
    if (dim.y < eps) { b.lo.y = c.y - eps; b.hi.y = c.y + eps; }
    if (dim.z < eps) { b.lo.z = c.z - eps; b.hi.z = c.z + eps; }
    return b;
}

// Möller–Trumbore exactly as Triangle::hit (core/mesh.rs:144-188): the
// determinant epsilon is tmin and t_max is not consulted.
struct MTResult {
    float t, u, v;
};
// det_eps: the determinant threshold, t_min as the reference uses it
// (negative: t_min).  Inside an instance the triple product a = e1 . (d x e2)
// is the world one divided by det M, so the instance passes t_min / |det M|
// to take the same decision as the placed triangle would.
VR_HD bool mt_hit(vf3 o, vf3 d, vf3 p0, vf3 p1, vf3 p2, float tmin, MTResult& r, float det_eps = -1.0f) {
    vf3 e1 = p1 - p0;
    vf3 e2 = p2 - p0;
    vf3 h = cross3(d, e2);
    float a = dot3(e1, h);
    const float eps = det_eps < 0.0f ? tmin : det_eps;
    if (a > -eps && a < eps) return false;
    float f = 1.0f / a;
    vf3 s = o - p0;
    float u = f * dot3(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    vf3 q = cross3(s, e1);
    float v = f * dot3(d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = f * dot3(e2, q);
    if (!(t > tmin)) return false;
    r.t = t;
    r.u = u;
    r.v = v;
    return true;
}

// ---- watertight ray/triangle test (Woop, Benthin & Wald, "Watertight
// Ray/Triangle Intersection", JCGT 2(1), 2013) ----------------------------
// The ray is sheared so that it runs along +z of a permuted frame; each edge
// function depends only on the ray and the edge's two vertices, and is
// computed as fl(fl(a*b) - fl(c*d)), which negates exactly when the edge is
// walked the other way round (no contraction: -ffp-contract=off).  Two
// triangles sharing an edge therefore see the same value with opposite signs,
// and a ray through the edge (or a shared vertex) is inside at least one of
// them: no ray slips through a closed mesh.  An edge value of exactly zero
// counts as inside on both sides.  No counterpart in the reference, which
// uses Moller-Trumbore (core/mesh.rs:144-188); used by the extended estimator
// only.
struct WTRay {
    int kx, ky, kz;
    float sx, sy, sz;
};

VR_HD WTRay wt_prepare(vf3 d) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    WTRay w;
    w.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    w.kx = w.kz == 2 ? 0 : w.kz + 1;
    w.ky = w.kx == 2 ? 0 : w.kx + 1;
    const float dz = comp(d, w.kz);
    if (dz < 0.0f) {  // keep the winding of the permuted frame
        const int t = w.kx;
        w.kx = w.ky;
        w.ky = t;
    }
    w.sx = comp(d, w.kx) / dz;
    w.sy = comp(d, w.ky) / dz;
    w.sz = 1.0f / dz;
    return w;
}

// Closest-hit candidate with t > tmin; u, v = barycentric weights of p1, p2
// (the same roles as Moller-Trumbore's u, v).
VR_HD bool wt_hit(const WTRay& w, vf3 o, vf3 p0, vf3 p1, vf3 p2, float tmin, MTResult& r) {
    const vf3 A = p0 - o, B = p1 - o, C = p2 - o;
    const float Az = comp(A, w.kz), Bz = comp(B, w.kz), Cz = comp(C, w.kz);
    const float Ax = comp(A, w.kx) - w.sx * Az, Ay = comp(A, w.ky) - w.sy * Az;
    const float Bx = comp(B, w.kx) - w.sx * Bz, By = comp(B, w.ky) - w.sy * Bz;
    const float Cx = comp(C, w.kx) - w.sx * Cz, Cy = comp(C, w.ky) - w.sy * Cz;
    const float U = Cx * By - Cy * Bx;
    const float V = Ax * Cy - Ay * Cx;
    const float W = Bx * Ay - By * Ax;
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float T = U * (w.sz * Az) + V * (w.sz * Bz) + W * (w.sz * Cz);
    const float inv = 1.0f / det;
    const float t = T * inv;
    if (!(t > tmin)) return false;
    r.t = t;
    r.u = V * inv;
    r.v = W * inv;
    return true;
}

// Slab test that never rejects a ray crossing the box, for the watertight
// engine's exact fallback: the box grows by 2^-20 of its coordinates' size
// and an empty interval is tolerated up to 2^-17 relative.
VR_HD bool wt_slab_hit(const Box3& b, vf3 o, vf3 d, float t_min, float t_max) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float lo = comp(b.lo, a), hi = comp(b.hi, a);
        const float pad = (fabsf(lo) + fabsf(hi)) * (1.0f / 1048576.0f) + 1e-30f;
        const float inv_d = 1.0f / comp(d, a);
        float t0 = (lo - pad - comp(o, a)) * inv_d;
        float t1 = (hi + pad - comp(o, a)) * inv_d;
        if (inv_d < 0.0f) {
            const float tmp = t0;
            t0 = t1;
            t1 = tmp;
        }
        if (t0 > t_min) t_min = t0;
        if (t1 < t_max) t_max = t1;
    }
    return t_min <= t_max * (1.0f + 1.0f / 131072.0f) || !(t_min == t_min) || !(t_max == t_max);
}
