// ORACLE — test infrastructure only.  Nothing in the product library may
// include, link or call this code; only tests/, __graft_entry__.smoke() and
// bench.py's CPU-baseline leg use it, as the checker.
//
// Vector/colour arithmetic restated from the reference's use of cgmath 0.18
// (voidray_renderer/src/util/vector.rs:1-34) with the association order of
// cgmath's published source: dot = (x*x' + y*y') + z*z', cross as below,
// normalize = v * (1 / |v|).  cgmath itself cannot be imported here, so that
// order is "parity unpinned" against cgmath; it is pinned against closed forms
// in tests/test_oracle.py.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace orc {

template <class R> struct V3 {
    R x, y, z;
};
template <class R> struct V2 {
    R x, y;
};

template <class R> inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
// vector * scalar and scalar * vector are both component-wise products.
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline V3<R> mulv(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R magnitude2(V3<R> a) { return dot(a, a); }
template <class R> inline R magnitude(V3<R> a) { return std::sqrt(magnitude2(a)); }
template <class R> inline V3<R> normalize(V3<R> a) { return a * (R(1) / magnitude(a)); }
// cgmath InnerSpace::angle for Vector3: atan2(|a x b|, a . b).
template <class R> inline R angle(V3<R> a, V3<R> b) { return std::atan2(magnitude(cross(a, b)), dot(a, b)); }
// Rust f32::min / f32::max (IEEE minNum/maxNum: a NaN operand yields the other).
template <class R> inline R rmin(R a, R b) { return std::fmin(a, b); }
template <class R> inline R rmax(R a, R b) { return std::fmax(a, b); }
template <class R> inline V3<R> vmin(V3<R> a, R c) { return {rmin(a.x, c), rmin(a.y, c), rmin(a.z, c)}; }

template <class R> inline R pi_v();
template <> inline float pi_v<float>() { return 3.14159265358979323846f; }
template <> inline double pi_v<double>() { return 3.14159265358979323846; }
template <class R> inline R inf_v() { return std::numeric_limits<R>::infinity(); }

// Rust `x as usize` for floats: saturating, NaN -> 0.
template <class R> inline uint64_t as_usize(R x) {
    if (!(x >= R(0))) return 0;
    if (x >= R(18446744073709551616.0)) return UINT64_MAX;
    return (uint64_t)x;
}

// f32::total_cmp key (Rust core::f32::total_cmp).
inline int32_t total_key(float f) {
    int32_t k;
    std::memcpy(&k, &f, 4);
    k ^= (int32_t)(((uint32_t)(k >> 31)) >> 1);
    return k;
}
inline int64_t total_key(double f) {
    int64_t k;
    std::memcpy(&k, &f, 8);
    k ^= (int64_t)(((uint64_t)(k >> 63)) >> 1);
    return k;
}

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the counter-based
// stream that stands in for rand::thread_rng (iterative.rs:29). ------------
inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
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

// Per-path word stream: word k comes from block k/4 =
// Philox(counter = (k/4, sample, pixel, request), key = (seed_lo, seed_hi)).
struct PathRng {
    uint32_t key[2];
    uint32_t sample, pixel, request;
    uint32_t n = 0;
    uint32_t cached = 0xFFFFFFFFu;
    uint32_t block[4];
    PathRng(uint64_t seed, uint32_t request_, uint32_t pixel_, uint32_t sample_)
        : sample(sample_), pixel(pixel_), request(request_) {
        key[0] = (uint32_t)seed;
        key[1] = (uint32_t)(seed >> 32);
    }
    uint32_t next_u32() {
        uint32_t b = n >> 2;
        if (b != cached) {
            uint32_t ctr[4] = {b, sample, pixel, request};
            philox4x32_10(ctr, key, block);
            cached = b;
        }
        return block[(n++) & 3u];
    }
    // 24-bit uniform in [0, 1), exact in f32.
    float uniform() { return (float)(next_u32() >> 8) * 5.9604644775390625e-08f; }
};

}  // namespace orc
