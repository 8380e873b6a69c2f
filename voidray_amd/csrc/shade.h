// Shading routines of the wavefront renderer: hit records, textures, the
// environment, random distributions and the materials.  Every f32 expression
// follows the operation order of the reference so that, with the same random
// words, a path makes the same decisions as voidray's CPU integrator.
#pragma once
#include "f32math.h"
#include "internal.h"

namespace vr {

// Rust `x as usize` (saturating; NaN -> 0).
VR_HD uint64_t as_usize(float x) {
    if (!(x >= 0.0f)) return 0;
    if (x >= 18446744073709551616.0f) return ~0ull;
    return (uint64_t)x;
}

VR_HD vf3 texel(const SceneView& sv, const GpuTexture& t, uint64_t i) {
    if (t.format == VR_TEX_RGBA8) {
        const uint8_t* p = sv.tex_data + t.offset + 4 * i;
        return mk3((float)p[0] / 255.0f, (float)p[1] / 255.0f, (float)p[2] / 255.0f);
    }
    const float* p = (const float*)(sv.tex_data + t.offset) + 3 * i;
    return mk3(p[0], p[1], p[2]);
}

VR_HD vf3 lerp3(vf3 a, vf3 b, float t) { return a * (1.0f - t) + b * t; }

// ImageTexture::bilinear_sample (texture.rs:59-78); HDRIEnvironment uses the
// same arithmetic (environments.rs:57-76).
template <class Fetch>
VR_HD vf3 bilinear(uint32_t W, uint32_t H, float x, float y, Fetch fetch) {
    uint64_t len = (uint64_t)W * H;
    uint64_t x0 = as_usize(x);
    if (x0 > W - 1) x0 = W - 1;
    uint64_t y0 = as_usize(y);
    if (y0 > H - 1) y0 = H - 1;
    float ax = x - (float)x0;
    float ay = y - (float)y0;
    return lerp3(lerp3(fetch((y0 * W + x0) % len), fetch((y0 * W + x0 + 1) % len), ax),
                 lerp3(fetch(((y0 + 1) * W + x0) % len), fetch(((y0 + 1) * W + x0 + 1) % len), ax), ay);
}

// Rust's x % 1.0 (fmod(x, 1)): x - trunc(x), which is exact, with the sign
// of x (fmod(-2, 1) = -0); inf and NaN give NaN.
VR_HD float fmod1(float x) { return copysignf(x - truncf(x), x); }

// ImageTexture::sample (texture.rs:82-98)
VR_HD vf3 tex_sample(const SceneView& sv, int32_t ti, float u, float v) {
    const GpuTexture t = sv.textures[ti];
    if (u < 0.0f) u -= truncf(u) - 1.0f;
    if (v < 0.0f) v -= truncf(v) - 1.0f;
    float x = fmod1(u) * (float)t.w;
    float y = (1.0f - fmod1(v)) * (float)t.h;
    if (t.sample_type == VR_SAMPLE_NEAREST) {
        uint64_t xi = as_usize(x), yi = as_usize(y);
        if (xi > t.w - 1) xi = t.w - 1;
        if (yi > t.h - 1) yi = t.h - 1;
        return texel(sv, t, (yi * t.w + xi) % ((uint64_t)t.w * t.h));
    }
    return bilinear(t.w, t.h, x, y, [&](uint64_t i) { return texel(sv, t, i); });
}

VR_HD vf3 env_texel(const SceneView& sv, uint64_t i) { return mk3(sv.env[3 * i], sv.env[3 * i + 1], sv.env[3 * i + 2]); }

// Environment::sample: UniformEnvironment (environments.rs:29-33) and
// HDRIEnvironment::sample with to_spherical_coords (environments.rs:80-90,
// util/math.rs:24-29).
VR_HD vf3 env_lookup(const SceneView& sv, vf3 dir) {
    if (sv.env_kind == VR_ENV_UNIFORM) return mk3(sv.env_color[0], sv.env_color[1], sv.env_color[2]);
    if (sv.env_kind != VR_ENV_HDRI) return mk3(0.f, 0.f, 0.f);
    vf3 d = normalize3(dir);
    float th = acosf(-d.y);
    float ph = atan2f(-d.z, d.x) + kPi;
    float u = th / kPi;
    float v = ph / (2.0f * kPi);
    float x = v * (float)sv.env_w;
    float y = (float)(sv.env_h - 1) - (u * (float)sv.env_h);
    return bilinear(sv.env_w, sv.env_h, x, y, [&](uint64_t i) { return env_texel(sv, i); });
}

// ---- random distributions (rand_distr 0.4.3 algorithms, capped) ------------
constexpr int kRejectCap = 32;

// UnitSphere: Marsaglia (1972).
VR_HD vf3 unit_sphere(PathRng& rng, uint32_t& capped) {
    for (int i = 0; i < kRejectCap; ++i) {
        float x1 = rng.uniform() * 2.0f - 1.0f;
        float x2 = rng.uniform() * 2.0f - 1.0f;
        float sum = x1 * x1 + x2 * x2;
        if (sum >= 1.0f) continue;
        float factor = 2.0f * sqrtf(1.0f - sum);
        return mk3(x1 * factor, x2 * factor, 1.0f - 2.0f * sum);
    }
    ++capped;
    return mk3(0.f, 0.f, 1.f);
}

// UnitDisc: square rejection.
VR_HD vf2 unit_disc(PathRng& rng, uint32_t& capped) {
    for (int i = 0; i < kRejectCap; ++i) {
        float x1 = rng.uniform() * 2.0f - 1.0f;
        float x2 = rng.uniform() * 2.0f - 1.0f;
        if (x1 * x1 + x2 * x2 <= 1.0f) return vf2{x1, x2};
    }
    ++capped;
    return vf2{0.f, 0.f};
}

VR_HD bool near_zero(vf3 v) {
    const float e = 1.0e-8f;
    return fabsf(v.x) < e && fabsf(v.y) < e && fabsf(v.z) < e;
}
// reflect / refract (util/math.rs:12-21)
VR_HD vf3 reflect3(vf3 v, vf3 n) { return v - (2.0f * dot3(v, n)) * n; }
VR_HD vf3 refract3(vf3 uv, vf3 n, float eta) {
    float cos_t = fminf(dot3(n, -uv), 1.0f);
    vf3 perp = eta * (uv + cos_t * n);
    vf3 par = (-sqrtf(fabsf(1.0f - mag2(perp)))) * n;
    return perp + par;
}
// ---- MicrofacetBSDF (voidray_common/src/microfacet.rs) ----------------------
// powi as LLVM expands it (binary exponentiation): x^3 = x * x^2, x^5 = x * (x^2)^2.
VR_HD float powi2(float x) { return x * x; }
VR_HD float powi3(float x) { return x * (x * x); }
VR_HD float powi5(float x) {
    const float x2 = x * x;
    return x * (x2 * x2);
}
// f32::is_normal: neither zero, subnormal, infinite nor NaN.
VR_HD bool is_normal_f(float x) { return fabsf(x) >= 1.17549435e-38f && fabsf(x) <= 3.40282347e38f; }
// f32::is_sign_positive
VR_HD bool sign_positive(float x) { return !(f2u(x) >> 31); }

// local_to_world(n) * h (util/math.rs:50-59).  The reference fills the
// matrix column by column with (ns.x, nss.x, n.x), (ns.y, nss.y, n.y), ...,
// so the product is (ns·h, nss·h, n·h); reproduced as written.
VR_HD vf3 local_to_world_mul(vf3 n, vf3 h) {
    const vf3 ns = is_normal_f(n.x) ? normalize3(mk3(n.y, -n.x, 0.0f)) : normalize3(mk3(0.0f, -n.z, n.y));
    const vf3 nss = cross3(n, ns);
    return mk3(dot3(ns, h), dot3(nss, h), dot3(n, h));
}

// UnitCircle (rand_distr 0.4.3): von Neumann's rejection method.
VR_HD vf2 unit_circle(PathRng& rng, uint32_t& capped) {
    for (int i = 0; i < kRejectCap; ++i) {
        const float x1 = rng.uniform() * 2.0f - 1.0f;
        const float x2 = rng.uniform() * 2.0f - 1.0f;
        const float sum = x1 * x1 + x2 * x2;
        if (sum >= 1.0f) continue;
        const float diff = x1 * x1 - x2 * x2;
        return vf2{diff / sum, (2.0f * x1) * x2 / sum};
    }
    ++capped;
    return vf2{1.f, 0.f};
}

// f32 exp, ln, atan, sin and cos of the microfacet material (Rust's f32::exp
// etc., glibc's expf, logf, ... on the reference's platform): the f32-only
// routines of f32math.h, which return the correctly rounded result but in
// cases within about 2^-43 relative of a rounding midpoint, denormal results
// of exp included (the device's own expf returns 0 there, and 0 turns a sample
// the reference keeps, with a tiny pdf and a clamped firefly, into "no
// scatter").  Not inlined: only microfacet hits run them, and inlined into the
// shading kernel their registers would count against every path.
__noinline__ VR_HD float exp_mf(float x) { return f32m::exp_cr(x); }
__noinline__ VR_HD float ln_mf(float x) { return f32m::ln_cr(x); }
__noinline__ VR_HD float atan_mf(float x) { return f32m::atan_cr(x); }
__noinline__ VR_HD vf2 sincos_mf(float x) {
    vf2 r;
    f32m::sincos_cr(x, r.x, r.y);
    return r;
}

// Beckmann distribution term D (microfacet.rs:140-143, 183-186).
VR_HD float beckmann_d(float m2, float nh2) { return exp_mf((nh2 - 1.0f) / (m2 * nh2)) / (((m2 * kPi) * nh2) * nh2); }

// MicrofacetBSDF::bsdf (microfacet.rs:120-213); n, wo, wi as the caller passes them.
VR_HD vf3 microfacet_bsdf(const GpuMaterial& m, vf3 n, vf3 wo, vf3 wi) {
    const vf3 color = mk3(m.color[0], m.color[1], m.color[2]);
    const float n_dot_wi = dot3(n, wi), n_dot_wo = dot3(n, wo);
    const bool wi_outside = sign_positive(n_dot_wi), wo_outside = sign_positive(n_dot_wo);
    if (!m.transparent && (!wi_outside || !wo_outside)) return mk3(0.f, 0.f, 0.f);
    const float m2 = m.roughness * m.roughness;
    const float f0s = powi2((m.index - 1.0f) / (m.index + 1.0f));
    const vf3 f0 = mk3(f0s, f0s, f0s) * (1.0f - m.metallic) + color * m.metallic;  // lerp (math.rs:40-45)
    if (wi_outside == wo_outside) {
        const vf3 h = normalize3(wi + wo);
        const float wo_dot_h = dot3(wo, h), n_dot_h = dot3(n, h);
        const float nh2 = powi2(n_dot_h);
        const float d = beckmann_d(m2, nh2);
        vf3 f;
        if (!wi_outside && sqrtf(1.0f - wo_dot_h * wo_dot_h) * m.index > 1.0f) {
            f = mk3(1.f, 1.f, 1.f);  // total internal reflection
        } else {
            const float p5 = powi5(1.0f - wo_dot_h);
            f = f0 + (mk3(1.f, 1.f, 1.f) - f0) * p5;
        }
        float g = fminf(n_dot_wi * n_dot_h, n_dot_wo * n_dot_h);
        g = (2.0f * g) / wo_dot_h;
        g = fminf(g, 1.0f);
        const vf3 specular = ((d * f) * g) / ((4.0f * n_dot_wo) * n_dot_wi);
        if (m.transparent) return specular;
        const vf3 diffuse = mulv(mk3(1.f, 1.f, 1.f) - f, color) / kPi;
        return specular + diffuse;
    }
    const float eta_t = wo_outside ? m.index : 1.0f / m.index;
    const vf3 h = normalize3(wi * eta_t + wo);
    const float wi_dot_h = dot3(wi, h), wo_dot_h = dot3(wo, h), n_dot_h = dot3(n, h);
    const float nh2 = powi2(n_dot_h);
    const float d = beckmann_d(m2, nh2);
    const float p5 = powi5(1.0f - fabsf(wi_dot_h));
    const vf3 f = f0 + (mk3(1.f, 1.f, 1.f) - f0) * p5;
    float g = fminf(fabsf(n_dot_wi * n_dot_h), fabsf(n_dot_wo * n_dot_h));
    g = (2.0f * g) / fabsf(wo_dot_h);
    g = fminf(g, 1.0f);
    const float a = fabsf((wi_dot_h * wo_dot_h) / (n_dot_wi * n_dot_wo));
    const vf3 btdf = a * (((d * (mk3(1.f, 1.f, 1.f) - f)) * g) / powi2(eta_t * wi_dot_h + wo_dot_h));
    return mulv(btdf, color);
}

// Beckmann pdf of a half vector (microfacet.rs:245-250).
VR_HD float beckmann_pdf(float m2, vf3 n, vf3 h) {
    const float cos_t = fabsf(dot3(h, n));
    const float sin_t = sqrtf(1.0f - cos_t * cos_t);
    return (1.0f / ((kPi * m2) * powi3(cos_t))) * exp_mf((-powi2(sin_t / cos_t)) / m2);
}

// MicrofacetBSDF::sample (microfacet.rs:215-312); false = None.
VR_HD bool microfacet_sample(const GpuMaterial& m, vf3 n, vf3 wo, PathRng& rng, uint32_t& capped, vf3& wi,
                             float& pdf) {
    const vf3 color = mk3(m.color[0], m.color[1], m.color[2]);
    const float m2 = m.roughness * m.roughness;
    const float f0 = powi2((m.index - 1.0f) / (m.index + 1.0f));
    float f = (1.0f - m.metallic) * f0 + m.metallic * (((color.x + color.y) + color.z) / 3.0f);  // Color::mean
    f = f * (1.0f - 0.2f) + 1.0f * 0.2f;                                                       // lerp(f, 1, 0.2)
    const float eta_t = dot3(wo, n) > 0.0f ? m.index : 1.0f / m.index;
    auto beckmann = [&]() {
        // θ = arctan √(-m^2 ln U), azimuth from UnitCircle
        const float theta = atan_mf(sqrtf(m2 * -ln_mf(rng.uniform())));
        const vf2 sc = sincos_mf(theta);
        const float sin_t = sc.x, cos_t = sc.y;
        const vf2 c = unit_circle(rng, capped);
        return local_to_world_mul(n, mk3(c.x * sin_t, c.y * sin_t, cos_t));
    };
    if (rng.uniform() < f) {  // gen_bool(f)
        wi = -reflect3(wo, beckmann());
    } else if (!m.transparent) {
        const vf2 q = unit_disc(rng, capped);  // Malley's method
        const float z = sqrtf((1.0f - q.x * q.x) - q.y * q.y);
        wi = local_to_world_mul(n, mk3(q.x, q.y, z));
    } else {
        const vf3 h = beckmann();
        const float cos_to = dot3(h, wo);
        const vf3 wo_perp = wo - h * cos_to;
        const vf3 wi_perp = (-wo_perp) / eta_t;
        const float sin2_ti = mag2(wi_perp);
        if (sin2_ti > 1.0f) return false;  // total internal reflection
        const float cos_ti = sqrtf(1.0f - sin2_ti);
        wi = ((-copysignf(1.0f, cos_to)) * cos_ti) * h + wi_perp;
    }
    float p = 0.0f;
    {
        const vf3 h = normalize3(wi + wo);
        p += f * beckmann_pdf(m2, n, h) / (4.0f * fabsf(dot3(h, wo)));
    }
    if (!m.transparent) {
        p += (1.0f - f) * fmaxf(dot3(wi, n), 0.0f) / kPi;
    } else if (sign_positive(dot3(wo, n)) != sign_positive(dot3(wi, n))) {
        const vf3 h = normalize3(wi * eta_t + wo);
        const float p_h = beckmann_pdf(m2, n, h);
        const float h_dot_wo = dot3(h, wo), h_dot_wi = dot3(h, wi);
        const float jacobian = fabsf(h_dot_wo) / powi2(eta_t * h_dot_wi + h_dot_wo);
        p += ((1.0f - f) * p_h) * jacobian;
    }
    if (p == 0.0f) return false;
    pdf = p;
    return true;
}

// Dielectric::reflectance (simple.rs:192-197); powi(x, 5) = x * (x*x)^2
VR_HD float schlick(float c, float idx) {
    float r0 = (1.0f - idx) / (1.0f + idx);
    r0 = r0 * r0;
    float x = 1.0f - c;
    float x2 = x * x;
    return r0 + (1.0f - r0) * (x * (x2 * x2));
}

// HitRecord (util/ray.rs:27-49) for a closest hit (prim, t, u, v) of ray (o, d).
struct HitRec {
    vf3 point, normal;
    vf2 uv;
    bool front;
    uint32_t surface;
};

VR_HD HitRec make_hit(const SceneView& sv, vf3 o, vf3 d, uint32_t prim, float t, float u, float v) {
    HitRec h;
    uint32_t surf = sv.prim_surface[prim];
    SurfaceInfo si = sv.surfaces[surf];
    if (si.analytic >= 0) {
        // Sphere::hit / GroundPlane::hit hit records (surfaces.rs:71-77, 96-103):
        // outward normal (point - centre) / radius or +y; u, v carry the uv
        const AnalyticDev& a = sv.analytic[si.analytic];
        h.point = o + d * t;
        vf3 n = a.kind == VR_ANALYTIC_SPHERE ? (h.point - mk3(a.c[0], a.c[1], a.c[2])) / a.r : mk3(0.f, 1.f, 0.f);
        h.uv.x = u;
        h.uv.y = v;
        h.front = dot3(d, n) < 0.0f;
        h.normal = h.front ? n : -n;
        h.surface = surf;
        return h;
    }
    MeshInfo mi = sv.meshes[si.mesh];
    uint32_t tri = prim - si.prim_base;
    const uint32_t* ix = sv.mesh_idx + 3 * (size_t)(mi.tri_base + tri);
    size_t v0 = mi.vert_base + ix[0], v1 = mi.vert_base + ix[1], v2 = mi.vert_base + ix[2];
    vf3 n0 = mk3(sv.vnrm[3 * v0], sv.vnrm[3 * v0 + 1], sv.vnrm[3 * v0 + 2]);
    vf3 n1 = mk3(sv.vnrm[3 * v1], sv.vnrm[3 * v1 + 1], sv.vnrm[3 * v1 + 2]);
    vf3 n2 = mk3(sv.vnrm[3 * v2], sv.vnrm[3 * v2 + 1], sv.vnrm[3 * v2 + 2]);
    float w = (1.0f - u) - v;
    // Triangle::hit shading (mesh.rs:176-181)
    vf3 n = (u * n1 + v * n2) + w * n0;
    h.uv.x = (u * sv.vuv[2 * v1] + v * sv.vuv[2 * v2]) + w * sv.vuv[2 * v0];
    h.uv.y = (u * sv.vuv[2 * v1 + 1] + v * sv.vuv[2 * v2 + 1]) + w * sv.vuv[2 * v0 + 1];
    const float* fp = sv.face_n + 3 * (size_t)(mi.tri_base + tri);
    vf3 fn = mk3(fp[0], fp[1], fp[2]);
    if (si.instance >= 0) {
        // an instance's mesh holds object-space normals: n' = (M^-1)^T n; the
        // face normal of the placed triangle is that of its object-space
        // normal, flipped when M reverses the winding
        const InstanceDev& x = sv.instances[si.instance];
        const float* q = x.nrm;
        n = mk3((q[0] * n.x + q[1] * n.y) + q[2] * n.z, (q[3] * n.x + q[4] * n.y) + q[5] * n.z,
                (q[6] * n.x + q[7] * n.y) + q[8] * n.z);
        fn = mk3((q[0] * fn.x + q[1] * fn.y) + q[2] * fn.z, (q[3] * fn.x + q[4] * fn.y) + q[5] * fn.z,
                 (q[6] * fn.x + q[7] * fn.y) + q[8] * fn.z);
        fn = fn * (x.face_sign / mag(fn));
    }
    const float thirty = (30.0f * kPi) / 180.0f;
    if (atan2f(mag(cross3(n, fn)), dot3(n, fn)) > thirty) n = fn;
    // HitRecord::new
    h.point = o + d * t;
    h.front = dot3(d, n) < 0.0f;
    h.normal = h.front ? n : -n;
    h.surface = surf;
    return h;
}

struct ScatterOut {
    vf3 att;
    vf3 dir;   // normalised (Ray::new)
    bool has_ray;
    bool diffuse;  // Lambertian (cosine lobe about `lobe`) for MIS
    vf3 lobe;      // unit lobe axis of a diffuse scatter
};

// Material::scatter (voidray_common/src/simple.rs, core/traits.rs:19-41).
// `cosine_lobe`: the extended estimator's Lambertian, which scatters about the
// normalised shading normal so that its pdf is cos/pi exactly.  MICRO: the
// scene has MicrofacetBSDF materials (compiled out otherwise: the caller
// only passes false for scenes without them).
template <bool MICRO = true>
VR_HD ScatterOut scatter(const SceneView& sv, const GpuMaterial& m, vf3 d, const HitRec& hit, PathRng& rng,
                         uint32_t& capped, bool cosine_lobe) {
    ScatterOut s;
    s.diffuse = false;
    s.lobe = mk3(0.f, 0.f, 0.f);
    switch (m.kind) {
        case VR_MAT_LAMBERTIAN: {
            vf3 n = m.normal_tex >= 0 ? tex_sample(sv, m.normal_tex, hit.uv.x, hit.uv.y) : hit.normal;
            if (cosine_lobe) n = normalize3(n);
            vf3 dir = n + unit_sphere(rng, capped);
            if (near_zero(dir)) dir = n;
            s.dir = normalize3(dir);
            s.has_ray = true;
            s.att = m.albedo_tex >= 0 ? tex_sample(sv, m.albedo_tex, hit.uv.x, hit.uv.y)
                                      : mk3(m.color[0], m.color[1], m.color[2]);
            s.diffuse = true;
            s.lobe = n;
            return s;
        }
        case VR_MAT_METAL: {
            vf3 refl = normalize3(reflect3(d, hit.normal));
            for (int i = 0; i < kRejectCap; ++i) {
                vf3 sd = normalize3(refl + m.param * unit_sphere(rng, capped));
                if (dot3(sd, hit.normal) > 0.0f) {
                    s.att = mk3(m.color[0], m.color[1], m.color[2]);
                    s.dir = sd;
                    s.has_ray = true;
                    return s;
                }
            }
            ++capped;  // the reference loops forever here; the path is absorbed
            s.att = mk3(0.f, 0.f, 0.f);
            s.has_ray = false;
            return s;
        }
        case VR_MAT_DIELECTRIC: {
            float ratio = hit.front ? 1.0f / m.param : m.param;
            vf3 ud = normalize3(d);
            float cos_t = fminf(dot3(hit.normal, -ud), 1.0f);
            float sin_t = sqrtf(1.0f - cos_t * cos_t);
            bool cannot = (ratio * sin_t) > 1.0f;
            vf3 dir;
            if (cannot || schlick(cos_t, ratio) > rng.uniform()) dir = reflect3(ud, hit.normal);
            else dir = refract3(ud, hit.normal, ratio);
            s.att = mk3(1.f, 1.f, 1.f);
            s.dir = normalize3(dir);
            s.has_ray = true;
            return s;
        }
        case VR_MAT_EMISSION:
            s.att = mk3(m.color[0], m.color[1], m.color[2]);
            s.has_ray = false;
            return s;
        case VR_MAT_MICROFACET: {
            if (!MICRO) {
                s.att = mk3(0.f, 0.f, 0.f);
                s.has_ray = false;
                return s;
            }
            // BSDFMaterial::scatter (traits.rs:19-41): wo is the ray's own
            // direction, normalised, as the reference passes it
            const vf3 wo = normalize3(d);
            vf3 wi;
            float pdf;
            if (!microfacet_sample(m, hit.normal, wo, rng, capped, wi, pdf)) {
                s.att = mk3(0.f, 0.f, 0.f);
                s.has_ray = false;
                return s;
            }
            const vf3 f = microfacet_bsdf(m, hit.normal, wo, wi);
            s.dir = normalize3(wi);
            s.att = (f * fabsf(dot3(wi, hit.normal))) * (1.0f / pdf);
            s.has_ray = true;
            return s;
        }
        case VR_MAT_LAMBERTIAN_BSDF: {
            if (mag2(hit.normal) == 0.0f) {  // the reference asserts here
                ++capped;
                s.att = mk3(0.f, 0.f, 0.f);
                s.has_ray = false;
                return s;
            }
            vf3 wi = normalize3(unit_sphere(rng, capped));
            float pdf = 1.0f;
            vf3 f = mk3(m.color[0], m.color[1], m.color[2]) / kPi;
            s.dir = normalize3(wi);
            s.att = (f * fabsf(dot3(wi, hit.normal))) * (1.0f / pdf);
            s.has_ray = true;
            return s;
        }
    }
    s.att = mk3(0.f, 0.f, 0.f);
    s.has_ray = false;
    return s;
}

// ---- environment importance sampling (extended estimator) -------------------
// Continuous lookup coordinates (x, y) of HDRIEnvironment::sample map to
// theta = pi (H - 1 - y) / H, phi = 2 pi x / W; a cell has unit area in (x, y),
// so pdf_omega = pdf_cell * W * H / (2 pi^2 sin theta).
VR_HD uint32_t cdf_search(const float* cdf, uint32_t n, float u) {
    // largest k in [0, n) with cdf[k] <= u  (cdf[0] = 0, cdf[n] = 1)
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The same cell as cdf_search(cdf, n, u), searched only between the guide
// table's bounds for u (G a power of two, so floor(u * G) is exact).
VR_HD uint32_t cdf_search_guided(const float* cdf, const uint32_t* guide, uint32_t n, uint32_t G, float u) {
    uint32_t j = (uint32_t)(u * (float)G);
    j = j < G ? j : G - 1u;
    uint32_t lo = guide[j], hi = guide[j + 1] + 1u;
    hi = hi < n ? hi : n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid;
    }
    return lo;
}

VR_HD float env_pdf_omega(const SceneView& sv, float x, float y, float sin_t) {
    if (!(sin_t > 0.f)) return 0.f;
    int64_t i = (int64_t)floorf(x);
    int64_t r = (int64_t)floorf(y) + 1;
    if (i < 0) i = 0;
    if (i > (int64_t)sv.env_w - 1) i = sv.env_w - 1;
    if (r < 0) r = 0;
    if (r > (int64_t)sv.env_h - 1) r = sv.env_h - 1;
    float pc = sv.env_pdf[(size_t)r * sv.env_w + (size_t)i];
    return pc * ((float)sv.env_w * (float)sv.env_h) / (2.0f * kPi * kPi * sin_t);
}

// pdf (solid angle) of sampling direction d by the environment tables.
VR_HD float env_pdf_dir(const SceneView& sv, vf3 dir) {
    vf3 d = normalize3(dir);
    float th = acosf(-d.y);
    float ph = atan2f(-d.z, d.x) + kPi;
    float x = (ph / (2.0f * kPi)) * (float)sv.env_w;
    float y = (float)(sv.env_h - 1) - (th / kPi) * (float)sv.env_h;
    return env_pdf_omega(sv, x, y, sinf(th));
}

VR_HD vf3 env_sample_dir(const SceneView& sv, float u1, float u2, float& pdf) {
    uint32_t H = sv.env_h, W = sv.env_w;
    uint32_t r = cdf_search_guided(sv.env_marg_cdf, sv.env_marg_guide, H, sv.env_gm, u1);
    float m0 = sv.env_marg_cdf[r], m1 = sv.env_marg_cdf[r + 1];
    const float* c = sv.env_cond_cdf + (size_t)r * (W + 1);
    uint32_t i = cdf_search_guided(c, sv.env_cond_guide + (size_t)r * (sv.env_gc + 1), W, sv.env_gc, u2);
    float c0 = c[i], c1 = c[i + 1];
    float fx = c1 > c0 ? (u2 - c0) / (c1 - c0) : 0.5f;
    float fy = m1 > m0 ? (u1 - m0) / (m1 - m0) : 0.5f;
    fx = fminf(fmaxf(fx, 0.f), 0.99999994f);
    fy = fminf(fmaxf(fy, 0.f), 0.99999994f);
    float x = (float)i + fx;
    float y = ((float)r - 1.0f) + fy;
    float th = kPi * ((float)H - 1.0f - y) / (float)H;
    float ph = 2.0f * kPi * x / (float)W;
    float st = sinf(th), ct = cosf(th);
    float sp = sinf(ph - kPi), cp = cosf(ph - kPi);
    pdf = env_pdf_omega(sv, x, y, st);
    return mk3(st * cp, -ct, -st * sp);
}

}  // namespace vr
