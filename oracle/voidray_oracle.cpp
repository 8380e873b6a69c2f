// ORACLE — test infrastructure only (see oracle_math.h).  A CPU restatement of
// voidray's CPU path tracer, operation for operation, instantiated in f32
// (parity) and f64 (convergence oracle).  Every function cites the reference
// item it follows (paths relative to the voidray repository).
//
// Exported as extern "C" for ctypes; built by oracle/build.py with
// g++ -O2 -ffp-contract=off (no fast-math, no FMA contraction).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "../include/voidray_hip.h"
#include "oracle_math.h"

using namespace orc;

namespace {

template <class R> struct AABB {
    V3<R> min, max;
    // AABB::default (util/aabb.rs:151-158)
    static AABB def() { return {v3(inf_v<R>(), inf_v<R>(), inf_v<R>()), v3(-inf_v<R>(), -inf_v<R>(), -inf_v<R>())}; }
    static AABB from_point(V3<R> p) { return {p, p}; }
    bool is_default() const {
        return min.x == inf_v<R>() && min.y == inf_v<R>() && min.z == inf_v<R>() && max.x == -inf_v<R>() &&
               max.y == -inf_v<R>() && max.z == -inf_v<R>();
    }
    // AABB::centroid (aabb.rs:33-35)
    V3<R> centroid() const { return (min + max) / R(2); }
};

// AABB::surround (aabb.rs:46-59)
template <class R> AABB<R> surround(const AABB<R>& a, const AABB<R>& b) {
    return {v3(rmin(a.min.x, b.min.x), rmin(a.min.y, b.min.y), rmin(a.min.z, b.min.z)),
            v3(rmax(a.max.x, b.max.x), rmax(a.max.y, b.max.y), rmax(a.max.z, b.max.z))};
}

// AABB::epsilon_expand (aabb.rs:62-83)
template <class R> AABB<R> epsilon_expand(AABB<R> b, R eps) {
    V3<R> dim = v3(b.max.x - b.min.x, b.max.y - b.min.y, b.max.z - b.min.z);
    V3<R> c = b.centroid();
    if (dim.x < eps) { b.min.x = c.x - eps; b.max.x = c.x + eps; }
    if (dim.y < eps) { b.min.y = c.y - eps; b.max.y = c.y + eps; }
    if (dim.z < eps) { b.min.z = c.z - eps; b.max.z = c.z + eps; }
    return b;
}

// Ray (util/ray.rs:5-21); the direction is stored as given (Ray::new
// normalises before construction, callers do that explicitly).
template <class R> struct Ray {
    V3<R> o, d;
    V3<R> at(R t) const { return o + d * t; }
};
template <class R> Ray<R> ray_new(V3<R> o, V3<R> d) { return {o, normalize(d)}; }

// AABB::hit (aabb.rs:86-148): slab test, inclusive reject `t_max <= t_min`.
template <class R> bool aabb_hit(const AABB<R>& b, const Ray<R>& r, R t_min, R t_max) {
    if (b.is_default()) return true;
    const R* mn = &b.min.x;
    const R* mx = &b.max.x;
    const R* o = &r.o.x;
    const R* d = &r.d.x;
    for (int a = 0; a < 3; ++a) {
        R inv_d = R(1) / d[a];
        R t0 = (mn[a] - o[a]) * inv_d;
        R t1 = (mx[a] - o[a]) * inv_d;
        if (inv_d < R(0)) std::swap(t0, t1);
        if (t0 > t_min) t_min = t0;
        if (t1 < t_max) t_max = t1;
        if (t_max <= t_min) return false;
    }
    return true;
}

template <class R> struct Hit {
    V3<R> point, normal;
    R t;
    V2<R> uv;
    bool front_face;
    R u, v;             // Möller–Trumbore barycentrics (local in Triangle::hit)
    uint32_t surface;   // surface index
    uint32_t tri;       // triangle index within the mesh
};

// Reference BVH node (core/bvh.rs:40-45), flattened: leaf => obj >= 0.
template <class R> struct Node {
    AABB<R> box;
    int32_t left = -1, right = -1;
    int64_t obj = -1;
};

// BvhNode::from_list (bvh.rs:48-130): node box = union of the object boxes,
// axis = largest centroid spread (strict >, ties -> Z), stable sort with
// total_cmp on that centroid coordinate, left = first n/2.
template <class R> int32_t build_ref(std::vector<uint32_t>& objs, const std::vector<AABB<R>>& bounds,
                                     std::vector<Node<R>>& nodes) {
    size_t n = objs.size();
    if (n == 0) return -1;
    if (n == 1) {
        Node<R> leaf;
        leaf.obj = objs[0];
        nodes.push_back(leaf);
        return (int32_t)nodes.size() - 1;
    }
    AABB<R> box = bounds[objs[0]];
    for (size_t i = 1; i < n; ++i) box = surround(box, bounds[objs[i]]);
    AABB<R> cb = AABB<R>::from_point(bounds[objs[0]].centroid());
    for (size_t i = 1; i < n; ++i) cb = surround(cb, AABB<R>::from_point(bounds[objs[i]].centroid()));
    V3<R> spread = cb.max - cb.min;
    int axis;
    if (spread.x > spread.y && spread.x > spread.z) axis = 0;
    else if (spread.y > spread.x && spread.y > spread.z) axis = 1;
    else axis = 2;
    std::vector<std::pair<decltype(total_key(R(0))), uint32_t>> keyed(n);
    for (size_t i = 0; i < n; ++i) {
        V3<R> c = bounds[objs[i]].centroid();
        R k = axis == 0 ? c.x : (axis == 1 ? c.y : c.z);
        keyed[i] = {total_key(k), objs[i]};
    }
    std::stable_sort(keyed.begin(), keyed.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<uint32_t> left, right;
    left.reserve(n / 2);
    right.reserve(n - n / 2);
    for (size_t i = 0; i < n; ++i) (i < n / 2 ? left : right).push_back(keyed[i].second);
    int32_t me = (int32_t)nodes.size();
    nodes.push_back(Node<R>());
    nodes[me].box = box;
    std::vector<uint32_t>().swap(objs);  // free memory early on deep trees
    int32_t l = build_ref(left, bounds, nodes);
    int32_t r = build_ref(right, bounds, nodes);
    nodes[me].left = l;
    nodes[me].right = r;
    return me;
}

template <class R> struct Texture {
    uint32_t w = 0, h = 0;
    int sample_type = VR_SAMPLE_NEAREST;
    std::vector<V3<R>> img;  // image.to_rgb32f() as Vec<Color>
};

template <class R> struct Mesh {
    std::vector<V3<R>> pos, nrm;
    std::vector<V2<R>> uv;
    std::vector<uint32_t> idx;
    std::vector<V3<R>> face_n;
    AABB<R> bounds;
    std::vector<Node<R>> tree;  // empty: brute-force small mesh (mesh.rs:125-138)
    int32_t root = -1;
    uint32_t prim_base = 0;
};

template <class R> struct Material {
    int kind;
    V3<R> color;
    R param;
    int albedo_tex, normal_tex;
    R index = R(1.5), roughness = R(1), metallic = R(0);  // MicrofacetBSDF (microfacet.rs:9-27)
    bool transparent = false;
};

template <class R> struct Scene {
    std::vector<Mesh<R>> meshes;
    std::vector<uint32_t> surface_mesh;  // VR_SURFACE_ANALYTIC | index for analytic surfaces
    std::vector<vr_analytic> analytic;
    std::vector<vr_object> objects;
    std::vector<Material<R>> materials;
    std::vector<Texture<R>> textures;
    int env_kind = VR_ENV_NONE;
    V3<R> env_color;
    Texture<R> env;
    // CameraAcceleration (camera.rs:38-67)
    V3<R> cam_o, cam_dir, cam_right, cam_up;
    R cam_d;
    bool dof = false;
    R aperture, focal_length;
    // top-level tree over surfaces (scene.rs:163-179)
    std::vector<Node<R>> top;
    int32_t top_root = -1;
    uint32_t num_prims = 0;
    bool is_analytic(uint32_t surf) const { return (surface_mesh[surf] & VR_SURFACE_ANALYTIC) != 0; }
    const vr_analytic& analytic_of(uint32_t surf) const { return analytic[surface_mesh[surf] & ~VR_SURFACE_ANALYTIC]; }
    // prims of a surface: its triangles, or 1 for an analytic surface
    uint32_t surface_prims(uint32_t surf) const {
        return is_analytic(surf) ? 1u : (uint32_t)(meshes[surface_mesh[surf]].idx.size() / 3);
    }
};

// Triangle bounds (mesh.rs:193-205): eps-expanded to 0.001.
template <class R> AABB<R> tri_bounds(const Mesh<R>& m, uint32_t t) {
    const uint32_t* i = &m.idx[3 * t];
    AABB<R> b = surround(AABB<R>::from_point(m.pos[i[0]]),
                         surround(AABB<R>::from_point(m.pos[i[1]]), AABB<R>::from_point(m.pos[i[2]])));
    return epsilon_expand(b, R(0.001f));
}

template <class R> V3<R> load3(const float* p, size_t i) { return v3<R>(R(p[3 * i]), R(p[3 * i + 1]), R(p[3 * i + 2])); }

// Mesh::from_buffers (mesh.rs:76-116).
template <class R> void make_mesh(Mesh<R>& m, const vr_mesh& src) {
    m.pos.resize(src.num_vertices);
    m.nrm.assign(src.num_vertices, v3<R>(0, 0, 0));
    m.uv.assign(src.num_vertices, V2<R>{0, 0});
    for (uint32_t i = 0; i < src.num_vertices; ++i) {
        m.pos[i] = load3<R>(src.positions, i);
        if (src.normals) m.nrm[i] = load3<R>(src.normals, i);
        if (src.uvs) m.uv[i] = V2<R>{R(src.uvs[2 * i]), R(src.uvs[2 * i + 1])};
    }
    m.idx.assign(src.indices, src.indices + 3 * (size_t)src.num_triangles);
    m.face_n.resize(src.num_triangles);
    for (uint32_t t = 0; t < src.num_triangles; ++t) {
        const uint32_t* i = &m.idx[3 * t];
        V3<R> e1 = m.pos[i[0]] - m.pos[i[1]];
        V3<R> e2 = m.pos[i[2]] - m.pos[i[1]];
        m.face_n[t] = normalize(cross(e2, e1));
    }
    m.bounds = AABB<R>::def();
    for (auto& p : m.pos) m.bounds = surround(m.bounds, AABB<R>{p, p});
    if (3 * (size_t)src.num_triangles > 4 * 3) {
        std::vector<AABB<R>> b(src.num_triangles);
        for (uint32_t t = 0; t < src.num_triangles; ++t) b[t] = tri_bounds(m, t);
        std::vector<uint32_t> objs(src.num_triangles);
        for (uint32_t t = 0; t < src.num_triangles; ++t) objs[t] = t;
        m.tree.reserve(2 * (size_t)src.num_triangles);
        m.root = build_ref(objs, b, m.tree);
    }
}

template <class R> void make_texture(Texture<R>& t, const vr_texture& s) {
    t.w = s.width;
    t.h = s.height;
    t.sample_type = s.sample_type;
    t.img.resize((size_t)s.width * s.height);
    for (size_t i = 0; i < t.img.size(); ++i) {
        if (s.format == VR_TEX_RGBA8) {
            const uint8_t* p = (const uint8_t*)s.data + 4 * i;
            t.img[i] = v3<R>(R((float)p[0] / 255.0f), R((float)p[1] / 255.0f), R((float)p[2] / 255.0f));
        } else {
            const float* p = (const float*)s.data + 3 * i;
            t.img[i] = v3<R>(R(p[0]), R(p[1]), R(p[2]));
        }
    }
}

template <class R> Scene<R>* make_scene(const vr_scene_desc* d) {
    auto* s = new Scene<R>();
    s->meshes.resize(d->num_meshes);
    for (uint32_t i = 0; i < d->num_meshes; ++i) make_mesh(s->meshes[i], d->meshes[i]);
    s->surface_mesh.assign(d->surface_mesh, d->surface_mesh + d->num_surfaces);
    if (d->num_analytic) s->analytic.assign(d->analytic, d->analytic + d->num_analytic);
    s->objects.assign(d->objects, d->objects + d->num_objects);
    for (uint32_t i = 0; i < d->num_materials; ++i) {
        const vr_material& m = d->materials[i];
        s->materials.push_back(Material<R>{m.kind, load3<R>(m.color, 0), R(m.param), m.albedo_texture, m.normal_texture,
                                           R(m.index), R(m.roughness), R(m.metallic), m.transparent != 0});
    }
    s->textures.resize(d->num_textures);
    for (uint32_t i = 0; i < d->num_textures; ++i) make_texture(s->textures[i], d->textures[i]);
    s->env_kind = d->environment.kind;
    s->env_color = load3<R>(d->environment.color, 0);
    if (s->env_kind == VR_ENV_HDRI) {
        vr_texture t{d->environment.width, d->environment.height, VR_TEX_RGB32F, VR_SAMPLE_BILINEAR, d->environment.data};
        make_texture(s->env, t);
    }
    // Camera::build_acceleration (camera.rs:38-54); tan is computed in f32 on
    // both sides and shared.
    const vr_camera& c = d->camera;
    s->cam_o = load3<R>(c.eye, 0);
    s->cam_dir = load3<R>(c.direction, 0);
    s->cam_up = load3<R>(c.up, 0);
    s->cam_d = R(1) / R(std::tan((float)c.fov / 2.0f));
    if (sizeof(R) == 8) s->cam_d = R(1) / std::tan(R(c.fov) / R(2));
    s->cam_right = normalize(cross(s->cam_dir, s->cam_up));
    s->dof = c.has_dof != 0;
    s->aperture = R(c.aperture);
    s->focal_length = dot(load3<R>(c.focal_point, 0) - s->cam_o, s->cam_dir);
    // global prim numbering: surfaces in order, triangles in face order
    uint32_t base = 0;
    std::vector<uint32_t> surf_base(d->num_surfaces);
    for (uint32_t i = 0; i < d->num_surfaces; ++i) {
        surf_base[i] = base;
        base += s->surface_prims(i);
    }
    s->num_prims = base;
    // top-level BVH over surfaces, unexpanded mesh bounds (scene.rs:72-80, 176)
    // and the analytic surfaces' own bounds (surfaces.rs:36-43, 106-113)
    std::vector<AABB<R>> sb(d->num_surfaces);
    for (uint32_t i = 0; i < d->num_surfaces; ++i) {
        if (!s->is_analytic(i)) {
            sb[i] = s->meshes[s->surface_mesh[i]].bounds;
            continue;
        }
        const vr_analytic& a = s->analytic_of(i);
        if (a.kind == VR_ANALYTIC_SPHERE) {
            V3<R> c = load3<R>(a.center, 0), rr = v3<R>(R(a.radius), R(a.radius), R(a.radius));
            sb[i] = AABB<R>{c - rr, c + rr};
        } else {
            sb[i] = AABB<R>{v3<R>(-inf_v<R>(), R(a.height) - R(0.0001f), -inf_v<R>()),
                            v3<R>(inf_v<R>(), R(a.height) + R(0.0001f), inf_v<R>())};
        }
    }
    std::vector<uint32_t> objs(d->num_surfaces);
    for (uint32_t i = 0; i < d->num_surfaces; ++i) objs[i] = i;
    s->top_root = build_ref(objs, sb, s->top);
    (void)surf_base;
    return s;
}

// Triangle::hit (mesh.rs:144-188): Möller–Trumbore as written; tmin is the
// determinant epsilon, t_max is ignored.
template <class R> bool tri_hit(const Mesh<R>& m, uint32_t tri, const Ray<R>& ray, R tmin, Hit<R>& out) {
    const uint32_t* i = &m.idx[3 * tri];
    V3<R> p0 = m.pos[i[0]], p1 = m.pos[i[1]], p2 = m.pos[i[2]];
    V3<R> e1 = p1 - p0;
    V3<R> e2 = p2 - p0;
    V3<R> h = cross(ray.d, e2);
    R a = dot(e1, h);
    if (a > -tmin && a < tmin) return false;
    R f = R(1) / a;
    V3<R> s = ray.o - p0;
    R u = f * dot(s, h);
    if (u < R(0) || u > R(1)) return false;
    V3<R> q = cross(s, e1);
    R v = f * dot(ray.d, q);
    if (v < R(0) || u + v > R(1)) return false;
    R t = f * dot(e2, q);
    R w = (R(1) - u) - v;
    V3<R> n = (u * m.nrm[i[1]] + v * m.nrm[i[2]]) + w * m.nrm[i[0]];
    V2<R> uv{(u * m.uv[i[1]].x + v * m.uv[i[2]].x) + w * m.uv[i[0]].x,
             (u * m.uv[i[1]].y + v * m.uv[i[2]].y) + w * m.uv[i[0]].y};
    R thirty = (R(30) * pi_v<R>()) / R(180);
    if (angle(n, m.face_n[tri]) > thirty) n = m.face_n[tri];
    if (!(t > tmin)) return false;
    // HitRecord::new (ray.rs:34-49)
    out.point = ray.at(t);
    out.front_face = dot(ray.d, n) < R(0);
    out.normal = out.front_face ? n : -n;
    out.t = t;
    out.uv = uv;
    out.u = u;
    out.v = v;
    out.tri = tri;
    return true;
}

// Ray-query counters (per thread).
struct Counters {
    uint64_t boxes = 0, tris = 0;
};

// BvhNode::hit + merge_optionals (bvh.rs:132-181) for a mesh tree: both
// children, unshrunk t_max.  A leaf yields (HitRecord, handle) (bvh.rs:143-147)
// and merge_optionals keeps `left` iff left < right under Rust's lexicographic
// tuple order, where HitRecord compares by t alone (ray.rs:52-62): on equal t
// the smaller handle (triangle index) wins.  The merge is a minimum over a
// total order, so a running minimum over the reached leaves in any visiting
// order gives the recursion's answer.  With prune = true the walk skips boxes
// beyond the current best (used only for the FP64 convergence oracle).
template <class R>
bool mesh_tree_hit(const Mesh<R>& m, const Ray<R>& ray, R tmin, Hit<R>& best, bool have, bool prune, Counters& ctr) {
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = m.root;
    while (sp) {
        const Node<R>& nd = m.tree[stack[--sp]];
        if (nd.obj >= 0) {
            Hit<R> h;
            ++ctr.tris;
            if (tri_hit(m, (uint32_t)nd.obj, ray, tmin, h)) {
                if (!have || h.t < best.t || (h.t == best.t && h.tri < best.tri)) {
                    best = h;
                    have = true;
                }
            }
            continue;
        }
        ++ctr.boxes;
        R tmax = (prune && have) ? best.t * R(1.0000001) : inf_v<R>();
        if (!aabb_hit(nd.box, ray, tmin, tmax)) continue;
        stack[sp++] = nd.right;  // popped after the whole left subtree
        stack[sp++] = nd.left;
    }
    return have;
}

// Mesh::hit (mesh.rs:122-140).
template <class R> bool mesh_hit(const Mesh<R>& m, const Ray<R>& ray, R tmin, Hit<R>& out, bool prune, Counters& ctr) {
    if (m.root >= 0) return mesh_tree_hit(m, ray, tmin, out, false, prune, ctr);
    bool have = false;
    R closest = inf_v<R>();
    for (uint32_t t = 0; t < m.idx.size() / 3; ++t) {
        Hit<R> h;
        ++ctr.tris;
        if (tri_hit(m, t, ray, tmin, h)) {
            if (closest > h.t) {
                closest = h.t;
                out = h;
                have = true;
            }
        }
    }
    return have;
}

// AnalyticSurface::hit with t_max = inf: Sphere::hit (surfaces.rs:46-78) and
// GroundPlane::hit (surfaces.rs:88-104); u, v carry the hit record's uv.  A
// ray lying exactly in a ground plane (t = 0/0) is taken as a miss.
template <class R> bool analytic_hit(const vr_analytic& a, const Ray<R>& ray, R tmin, Hit<R>& out) {
    R t;
    V3<R> n;
    V2<R> uv{R(0), R(0)};
    if (a.kind == VR_ANALYTIC_SPHERE) {
        V3<R> c = load3<R>(a.center, 0);
        V3<R> oc = ray.o - c;
        R aa = dot(ray.d, ray.d);
        R half_b = dot(oc, ray.d);
        R cc = dot(oc, oc) - R(a.radius) * R(a.radius);
        R disc = half_b * half_b - aa * cc;
        if (disc < R(0)) return false;
        R sqrtd = std::sqrt(disc);
        R root = (-half_b - sqrtd) / aa;
        if (root < tmin || inf_v<R>() < root) {
            root = (-half_b + sqrtd) / aa;
            if (root < tmin || inf_v<R>() < root) return false;
        }
        t = root;
        n = (ray.at(t) - c) / R(a.radius);
    } else {
        t = (R(a.height) - ray.o.y) / ray.d.y;
        if (t <= tmin || t >= inf_v<R>() || t != t) return false;
        V3<R> p = ray.at(t);
        uv = V2<R>{p.x, p.z};
        n = v3<R>(R(0), R(1), R(0));
    }
    // HitRecord::new (ray.rs:34-49)
    out.point = ray.at(t);
    out.front_face = dot(ray.d, n) < R(0);
    out.normal = out.front_face ? n : -n;
    out.t = t;
    out.uv = uv;
    out.u = uv.x;
    out.v = uv.y;
    out.tri = 0;
    return true;
}

// SceneAcceleration::hit (scene.rs:182-185) via the top-level tree: the same
// merge over (HitRecord, surface index) tuples, so on equal t the smaller
// surface index wins (the mesh's own tie was settled inside mesh_hit).
template <class R> bool scene_hit(const Scene<R>& s, const Ray<R>& ray, R tmin, Hit<R>& best, bool prune, Counters& ctr) {
    if (s.top_root < 0) return false;
    bool have = false;
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = s.top_root;
    while (sp) {
        const Node<R>& nd = s.top[stack[--sp]];
        if (nd.obj >= 0) {
            uint32_t surf = (uint32_t)nd.obj;
            Hit<R> h;
            if (s.is_analytic(surf) ? analytic_hit(s.analytic_of(surf), ray, tmin, h)
                                    : mesh_hit(s.meshes[s.surface_mesh[surf]], ray, tmin, h, prune, ctr)) {
                if (!have || h.t < best.t || (h.t == best.t && surf < best.surface)) {
                    h.surface = surf;
                    best = h;
                    have = true;
                }
            }
            continue;
        }
        ++ctr.boxes;
        R tmax = (prune && have) ? best.t * R(1.0000001) : inf_v<R>();
        if (!aabb_hit(nd.box, ray, tmin, tmax)) continue;
        stack[sp++] = nd.right;
        stack[sp++] = nd.left;
    }
    return have;
}

// ---- textures and environments -----------------------------------------

template <class R> V3<R> lerp(V3<R> a, V3<R> b, R t) { return a * (R(1) - t) + b * t; }

// ImageTexture::bilinear_sample (texture.rs:59-78) / HDRIEnvironment (environments.rs:57-76)
template <class R> V3<R> bilinear(const Texture<R>& t, R x, R y) {
    uint64_t len = t.img.size();
    uint64_t x0 = std::min<uint64_t>(as_usize(x), t.w - 1);
    uint64_t y0 = std::min<uint64_t>(as_usize(y), t.h - 1);
    R ax = x - R(x0);
    R ay = y - R(y0);
    return lerp(lerp(t.img[(y0 * t.w + x0) % len], t.img[(y0 * t.w + x0 + 1) % len], ax),
                lerp(t.img[((y0 + 1) * t.w + x0) % len], t.img[((y0 + 1) * t.w + x0 + 1) % len], ax), ay);
}

// ImageTexture::sample (texture.rs:82-98), nearest_sample (52-57)
template <class R> V3<R> tex_sample(const Texture<R>& t, R u, R v) {
    if (u < R(0)) u -= std::trunc(u) - R(1);
    if (v < R(0)) v -= std::trunc(v) - R(1);
    R x = std::fmod(u, R(1)) * R(t.w);
    R y = (R(1) - std::fmod(v, R(1))) * R(t.h);
    if (t.sample_type == VR_SAMPLE_NEAREST) {
        uint64_t xi = std::min<uint64_t>(as_usize(x), t.w - 1);
        uint64_t yi = std::min<uint64_t>(as_usize(y), t.h - 1);
        return t.img[(yi * t.w + xi) % t.img.size()];
    }
    return bilinear(t, x, y);
}

// Environment::sample (environments.rs:29-33, 80-90) with to_spherical_coords (math.rs:24-29)
template <class R> V3<R> env_sample(const Scene<R>& s, const Ray<R>& ray) {
    if (s.env_kind == VR_ENV_UNIFORM) return s.env_color;
    V3<R> d = normalize(ray.d);
    R th = std::acos(-d.y);
    R ph = std::atan2(-d.z, d.x) + pi_v<R>();
    R u = th / pi_v<R>();
    R v = ph / (R(2) * pi_v<R>());
    R x = v * R(s.env.w);
    R y = R(s.env.h - 1) - (u * R(s.env.h));
    return bilinear(s.env, x, y);
}

// ---- random distributions (rand_distr 0.4.3 algorithms) ------------------

constexpr int kRejectCap = 32;

// UnitSphere: Marsaglia (1972) rejection.
template <class R> V3<R> unit_sphere(PathRng& rng, uint64_t& capped) {
    for (int i = 0; i < kRejectCap; ++i) {
        R x1 = R(rng.uniform() * 2.0f - 1.0f);
        R x2 = R(rng.uniform() * 2.0f - 1.0f);
        R sum = x1 * x1 + x2 * x2;
        if (sum >= R(1)) continue;
        R factor = R(2) * std::sqrt(R(1) - sum);
        return v3(x1 * factor, x2 * factor, R(1) - R(2) * sum);
    }
    ++capped;
    return v3<R>(0, 0, 1);
}

// UnitDisc: square rejection.
template <class R> V2<R> unit_disc(PathRng& rng, uint64_t& capped) {
    for (int i = 0; i < kRejectCap; ++i) {
        R x1 = R(rng.uniform() * 2.0f - 1.0f);
        R x2 = R(rng.uniform() * 2.0f - 1.0f);
        if (x1 * x1 + x2 * x2 <= R(1)) return V2<R>{x1, x2};
    }
    ++capped;
    return V2<R>{0, 0};
}

template <class R> bool near_zero(V3<R> v) {
    const R e = R(1.0e-8f);
    return std::fabs(v.x) < e && std::fabs(v.y) < e && std::fabs(v.z) < e;
}
// reflect / refract (math.rs:12-21)
template <class R> V3<R> reflect(V3<R> v, V3<R> n) { return v - (R(2) * dot(v, n)) * n; }

// ---- MicrofacetBSDF (voidray_common/src/microfacet.rs) ----------------------
// powi as LLVM expands it: x^3 = x * x^2, x^5 = x * (x^2)^2.
template <class R> R powi2(R x) { return x * x; }
template <class R> R powi3(R x) { return x * (x * x); }
template <class R> R powi5(R x) {
    R x2 = x * x;
    return x * (x2 * x2);
}
// local_to_world(n) * h (util/math.rs:50-59): the matrix is filled column by
// column with (ns.x, nss.x, n.x), ..., so the product is (ns.h, nss.h, n.h).
template <class R> V3<R> local_to_world_mul(V3<R> n, V3<R> h) {
    V3<R> ns = std::isnormal(n.x) ? normalize(v3<R>(n.y, -n.x, R(0))) : normalize(v3<R>(R(0), -n.z, n.y));
    V3<R> nss = cross(n, ns);
    return v3<R>(dot(ns, h), dot(nss, h), dot(n, h));
}
// UnitCircle (rand_distr 0.4.3): von Neumann's rejection method.
template <class R> V2<R> unit_circle(PathRng& rng, uint64_t& capped) {
    for (int i = 0; i < kRejectCap; ++i) {
        R x1 = R(rng.uniform() * 2.0f - 1.0f);
        R x2 = R(rng.uniform() * 2.0f - 1.0f);
        R sum = x1 * x1 + x2 * x2;
        if (sum >= R(1)) continue;
        R diff = x1 * x1 - x2 * x2;
        return V2<R>{diff / sum, (R(2) * x1) * x2 / sum};
    }
    ++capped;
    return V2<R>{R(1), R(0)};
}
template <class R> R beckmann_d(R m2, R nh2) {
    return std::exp((nh2 - R(1)) / (m2 * nh2)) / (((m2 * pi_v<R>()) * nh2) * nh2);
}
template <class R> bool sign_positive(R x) { return !std::signbit(x); }

// MicrofacetBSDF::bsdf (microfacet.rs:120-213)
template <class R> V3<R> microfacet_bsdf(const Material<R>& m, V3<R> n, V3<R> wo, V3<R> wi) {
    const V3<R> one = v3<R>(R(1), R(1), R(1));
    R n_dot_wi = dot(n, wi), n_dot_wo = dot(n, wo);
    bool wi_outside = sign_positive(n_dot_wi), wo_outside = sign_positive(n_dot_wo);
    if (!m.transparent && (!wi_outside || !wo_outside)) return v3<R>(0, 0, 0);
    R m2 = m.roughness * m.roughness;
    R f0s = powi2((m.index - R(1)) / (m.index + R(1)));
    V3<R> f0 = v3<R>(f0s, f0s, f0s) * (R(1) - m.metallic) + m.color * m.metallic;  // lerp (math.rs:40-45)
    if (wi_outside == wo_outside) {
        V3<R> h = normalize(wi + wo);
        R wo_dot_h = dot(wo, h), n_dot_h = dot(n, h);
        R nh2 = powi2(n_dot_h);
        R d = beckmann_d(m2, nh2);
        V3<R> f = (!wi_outside && std::sqrt(R(1) - wo_dot_h * wo_dot_h) * m.index > R(1))
                      ? one
                      : f0 + (one - f0) * powi5(R(1) - wo_dot_h);
        R g = rmin(n_dot_wi * n_dot_h, n_dot_wo * n_dot_h);
        g = (R(2) * g) / wo_dot_h;
        g = rmin(g, R(1));
        V3<R> specular = ((d * f) * g) / ((R(4) * n_dot_wo) * n_dot_wi);
        if (m.transparent) return specular;
        V3<R> diffuse = mulv(one - f, m.color) / pi_v<R>();
        return specular + diffuse;
    }
    R eta_t = wo_outside ? m.index : R(1) / m.index;
    V3<R> h = normalize(wi * eta_t + wo);
    R wi_dot_h = dot(wi, h), wo_dot_h = dot(wo, h), n_dot_h = dot(n, h);
    R nh2 = powi2(n_dot_h);
    R d = beckmann_d(m2, nh2);
    V3<R> f = f0 + (one - f0) * powi5(R(1) - std::fabs(wi_dot_h));
    R g = rmin(std::fabs(n_dot_wi * n_dot_h), std::fabs(n_dot_wo * n_dot_h));
    g = (R(2) * g) / std::fabs(wo_dot_h);
    g = rmin(g, R(1));
    R a = std::fabs((wi_dot_h * wo_dot_h) / (n_dot_wi * n_dot_wo));
    V3<R> btdf = a * (((d * (one - f)) * g) / powi2(eta_t * wi_dot_h + wo_dot_h));
    return mulv(btdf, m.color);
}

template <class R> R beckmann_pdf(R m2, V3<R> n, V3<R> h) {
    R cos_t = std::fabs(dot(h, n));
    R sin_t = std::sqrt(R(1) - cos_t * cos_t);
    return (R(1) / ((pi_v<R>() * m2) * powi3(cos_t))) * std::exp((-powi2(sin_t / cos_t)) / m2);
}

// MicrofacetBSDF::sample (microfacet.rs:215-312); false = None.
template <class R>
bool microfacet_sample(const Material<R>& m, V3<R> n, V3<R> wo, PathRng& rng, uint64_t& capped, V3<R>& wi, R& pdf) {
    R m2 = m.roughness * m.roughness;
    R f0 = powi2((m.index - R(1)) / (m.index + R(1)));
    R f = (R(1) - m.metallic) * f0 + m.metallic * (((m.color.x + m.color.y) + m.color.z) / R(3));
    f = f * (R(1) - R(0.2f)) + R(1) * R(0.2f);
    R eta_t = dot(wo, n) > R(0) ? m.index : R(1) / m.index;
    auto beckmann = [&]() {
        R theta = std::atan(std::sqrt(m2 * -std::log(R(rng.uniform()))));
        R sin_t = std::sin(theta), cos_t = std::cos(theta);
        V2<R> c = unit_circle<R>(rng, capped);
        return local_to_world_mul(n, v3<R>(c.x * sin_t, c.y * sin_t, cos_t));
    };
    if (R(rng.uniform()) < f) {
        wi = -reflect(wo, beckmann());
    } else if (!m.transparent) {
        V2<R> q = unit_disc<R>(rng, capped);
        R z = std::sqrt((R(1) - q.x * q.x) - q.y * q.y);
        wi = local_to_world_mul(n, v3<R>(q.x, q.y, z));
    } else {
        V3<R> h = beckmann();
        R cos_to = dot(h, wo);
        V3<R> wo_perp = wo - h * cos_to;
        V3<R> wi_perp = (-wo_perp) / eta_t;
        R sin2_ti = dot(wi_perp, wi_perp);
        if (sin2_ti > R(1)) return false;
        R cos_ti = std::sqrt(R(1) - sin2_ti);
        wi = ((-std::copysign(R(1), cos_to)) * cos_ti) * h + wi_perp;
    }
    R p = R(0);
    {
        V3<R> h = normalize(wi + wo);
        p += f * beckmann_pdf(m2, n, h) / (R(4) * std::fabs(dot(h, wo)));
    }
    if (!m.transparent) {
        p += (R(1) - f) * rmax(dot(wi, n), R(0)) / pi_v<R>();
    } else if (sign_positive(dot(wo, n)) != sign_positive(dot(wi, n))) {
        V3<R> h = normalize(wi * eta_t + wo);
        R p_h = beckmann_pdf(m2, n, h);
        R h_dot_wo = dot(h, wo), h_dot_wi = dot(h, wi);
        R jacobian = std::fabs(h_dot_wo) / powi2(eta_t * h_dot_wi + h_dot_wo);
        p += ((R(1) - f) * p_h) * jacobian;
    }
    if (p == R(0)) return false;
    pdf = p;
    return true;
}
template <class R> V3<R> refract(V3<R> uv, V3<R> n, R eta) {
    R cos_t = rmin(dot(n, -uv), R(1));
    V3<R> perp = eta * (uv + cos_t * n);
    V3<R> par = (-std::sqrt(std::fabs(R(1) - magnitude2(perp)))) * n;
    return perp + par;
}
// Dielectric::reflectance (simple.rs:192-197); powi(x, 5) = x * (x*x)^2.
template <class R> R schlick(R c, R idx) {
    R r0 = (R(1) - idx) / (R(1) + idx);
    r0 = r0 * r0;
    R x = R(1) - c;
    R x2 = x * x;
    return r0 + (R(1) - r0) * (x * (x2 * x2));
}

struct RenderCounters {
    uint64_t segments = 0, misses = 0, capped = 0;
    Counters q;
};

template <class R> struct Scatter {
    V3<R> att;
    bool has_ray;
    Ray<R> ray;
};

// Material::scatter implementations (voidray_common/src/simple.rs, core/traits.rs:19-41)
template <class R>
Scatter<R> scatter(const Scene<R>& s, const Material<R>& m, const Ray<R>& ray, const Hit<R>& hit, PathRng& rng,
                   RenderCounters& rc, bool cosine_lobe = false, V3<R>* lobe = nullptr) {
    Scatter<R> out;
    switch (m.kind) {
        case VR_MAT_LAMBERTIAN: {
            V3<R> n = m.normal_tex >= 0 ? tex_sample(s.textures[m.normal_tex], hit.uv.x, hit.uv.y) : hit.normal;
            // extended estimator: scatter about the normalised shading normal (pdf = cos/pi)
            if (cosine_lobe) n = normalize(n);
            if (lobe) *lobe = n;
            V3<R> dir = n + unit_sphere<R>(rng, rc.capped);
            if (near_zero(dir)) dir = n;
            out.ray = ray_new(hit.point, dir);
            out.has_ray = true;
            out.att = m.albedo_tex >= 0 ? tex_sample(s.textures[m.albedo_tex], hit.uv.x, hit.uv.y) : m.color;
            return out;
        }
        case VR_MAT_METAL: {
            V3<R> refl = normalize(reflect(ray.d, hit.normal));
            for (int i = 0; i < kRejectCap; ++i) {
                Ray<R> sc = ray_new(hit.point, refl + m.param * unit_sphere<R>(rng, rc.capped));
                if (dot(sc.d, hit.normal) > R(0)) {
                    out.att = m.color;
                    out.has_ray = true;
                    out.ray = sc;
                    return out;
                }
            }
            ++rc.capped;  // the reference loops forever here; the path is absorbed
            out.att = v3<R>(0, 0, 0);
            out.has_ray = false;
            return out;
        }
        case VR_MAT_DIELECTRIC: {
            R ratio = hit.front_face ? R(1) / m.param : m.param;
            V3<R> ud = normalize(ray.d);
            R cos_t = rmin(dot(hit.normal, -ud), R(1));
            R sin_t = std::sqrt(R(1) - cos_t * cos_t);
            bool cannot = (ratio * sin_t) > R(1);
            V3<R> dir;
            if (cannot || schlick(cos_t, ratio) > R(rng.uniform())) dir = reflect(ud, hit.normal);
            else dir = refract(ud, hit.normal, ratio);
            out.att = v3<R>(1, 1, 1);
            out.has_ray = true;
            out.ray = ray_new(hit.point, dir);
            return out;
        }
        case VR_MAT_EMISSION:
            out.att = m.color;
            out.has_ray = false;
            return out;
        case VR_MAT_MICROFACET: {
            // blanket impl (traits.rs:19-41) + MicrofacetBSDF: wo = the ray's direction, normalised
            V3<R> wo = normalize(ray.d);
            V3<R> wi;
            R pdf;
            if (!microfacet_sample(m, hit.normal, wo, rng, rc.capped, wi, pdf)) {
                out.att = v3<R>(0, 0, 0);
                out.has_ray = false;
                return out;
            }
            V3<R> f = microfacet_bsdf(m, hit.normal, wo, wi);
            out.ray = ray_new(ray.at(hit.t), wi);
            out.att = (f * std::fabs(dot(wi, hit.normal))) * (R(1) / pdf);
            out.has_ray = true;
            return out;
        }
        case VR_MAT_LAMBERTIAN_BSDF: {
            // blanket impl (traits.rs:19-41) + LambertianBSDF (simple.rs:60-81)
            if (magnitude2(hit.normal) == R(0)) {  // the reference asserts here
                ++rc.capped;
                out.att = v3<R>(0, 0, 0);
                out.has_ray = false;
                return out;
            }
            V3<R> wi = normalize(unit_sphere<R>(rng, rc.capped));
            R pdf = R(1);
            V3<R> f = m.color / pi_v<R>();
            out.ray = ray_new(ray.at(hit.t), wi);
            out.att = (f * std::fabs(dot(wi, hit.normal))) * (R(1) / pdf);
            out.has_ray = true;
            return out;
        }
    }
    out.att = v3<R>(0, 0, 0);
    out.has_ray = false;
    return out;
}

struct RenderParams {
    uint32_t max_bounces;
    float clamp;
    int mode;
};

// trace_ray_internal (core/tracer.rs:19-56): recursive, per-level clamp.
template <class R>
V3<R> trace_rec(const Scene<R>& s, const RenderParams& p, const Ray<R>& ray, uint32_t depth, PathRng& rng,
                RenderCounters& rc, bool prune) {
    V3<R> color = v3<R>(0, 0, 0);
    if (depth < p.max_bounces) {
        Hit<R> hit;
        ++rc.segments;
        if (!scene_hit(s, ray, R(0.00001f), hit, prune, rc.q)) {
            ++rc.misses;
            if (s.env_kind == VR_ENV_NONE) return v3<R>(0, 0, 0);
            return env_sample(s, ray);
        }
        const Material<R>& m = s.materials[s.objects[hit.surface].material];
        Scatter<R> sc;
        if (p.mode == VR_MODE_FULL) {
            sc = scatter(s, m, ray, hit, rng, rc);
        } else {
            V3<R> nn = normalize(hit.normal);
            sc.att = v3(R(0.5) * nn.x + R(0.5), R(0.5) * nn.y + R(0.5), R(0.5) * nn.z + R(0.5));
            sc.has_ray = false;
        }
        V3<R> delta = sc.has_ray ? mulv(sc.att, trace_rec(s, p, sc.ray, depth + 1, rng, rc, prune)) : sc.att;
        color = color + vmin(delta, R(p.clamp));
    }
    return color;
}

// CameraAcceleration::cast_ray (camera.rs:69-82)
template <class R> Ray<R> cast_ray(const Scene<R>& s, R x, R y, PathRng& rng, RenderCounters& rc) {
    V3<R> origin = s.cam_o;
    V3<R> nd = (s.cam_d * s.cam_dir + x * s.cam_right) + y * s.cam_up;
    if (s.dof) {
        V3<R> fp = origin + normalize(nd) * s.focal_length;
        V2<R> dk = unit_disc<R>(rng, rc.capped);
        origin = origin + (dk.x * s.cam_right + dk.y * s.cam_up) * s.aperture;
        nd = fp - origin;
    }
    return ray_new(origin, normalize(nd));
}

// ---- extended estimator (no counterpart in the reference) ---------------------
// Environment importance sampling over the continuous lookup coordinates of
// HDRIEnvironment::sample: cell (i, r) covers x in [i, i+1), y in [r-1, r), its
// weight is luminance(lookup at the cell centre) * sin(theta_centre) + a floor
// of 1e-3 * mean; theta = pi (H - 1 - y) / H, phi = 2 pi x / W, so
// pdf_omega = pdf_cell * W * H / (2 pi^2 sin theta).  Prefix sums in FP64.
template <class R> struct EnvTables {
    std::vector<R> marg, cond, pdf;  // H+1, H*(W+1), H*W
    uint32_t W = 0, H = 0;
};

template <class R> EnvTables<R> make_env_tables(const Texture<R>& env) {
    EnvTables<R> T;
    uint32_t W = env.w, H = env.h;
    T.W = W;
    T.H = H;
    size_t len = (size_t)W * H;
    auto lum = [&](size_t i) {
        const V3<R>& c = env.img[i % len];
        return 0.2126 * (double)c.x + 0.7152 * (double)c.y + 0.0722 * (double)c.z;
    };
    std::vector<double> w(len);
    double mean = 0;
    for (uint32_t r = 0; r < H; ++r) {
        double yc = (double)r - 0.5;
        double th = M_PI * ((double)H - 1.0 - yc) / (double)H;
        double st = std::sin(th);
        for (uint32_t i = 0; i < W; ++i) {
            double x = i + 0.5, y = yc;
            size_t x0 = std::min<size_t>(x >= 0 ? (size_t)x : 0, W - 1);
            size_t y0 = std::min<size_t>(y >= 0 ? (size_t)y : 0, H - 1);
            double ax = x - (double)x0, ay = y - (double)y0;
            double a = lum(y0 * W + x0) * (1 - ax) + lum(y0 * W + x0 + 1) * ax;
            double b = lum((y0 + 1) * W + x0) * (1 - ax) + lum((y0 + 1) * W + x0 + 1) * ax;
            double l = std::max(0.0, a * (1 - ay) + b * ay);
            if (!std::isfinite(l)) l = 0;
            w[(size_t)r * W + i] = l * st;
            mean += l * st;
        }
    }
    mean /= (double)len;
    double fl = std::max(mean * 1e-3, 1e-12);
    for (auto& x : w) x += fl;
    std::vector<double> m(H + 1, 0.0), c((size_t)H * (W + 1), 0.0);
    for (uint32_t r = 0; r < H; ++r) {
        double* cr = &c[(size_t)r * (W + 1)];
        for (uint32_t i = 0; i < W; ++i) cr[i + 1] = cr[i] + w[(size_t)r * W + i];
        m[r + 1] = m[r] + cr[W];
    }
    double total = m[H];
    T.marg.resize(H + 1);
    T.cond.resize((size_t)H * (W + 1));
    T.pdf.resize(len);
    for (uint32_t r = 0; r <= H; ++r) T.marg[r] = R(m[r] / total);
    T.marg[H] = R(1);
    for (uint32_t r = 0; r < H; ++r) {
        for (uint32_t i = 0; i <= W; ++i) T.cond[(size_t)r * (W + 1) + i] = R(c[(size_t)r * (W + 1) + i] / c[(size_t)r * (W + 1) + W]);
        T.cond[(size_t)r * (W + 1) + W] = R(1);
        for (uint32_t i = 0; i < W; ++i) T.pdf[(size_t)r * W + i] = R(w[(size_t)r * W + i] / total);
    }
    return T;
}

template <class R> uint32_t cdf_find(const R* cdf, uint32_t n, R u) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <class R> R env_pdf_omega(const EnvTables<R>& T, R x, R y, R sin_t) {
    if (!(sin_t > R(0))) return R(0);
    int64_t i = (int64_t)std::floor(x), r = (int64_t)std::floor(y) + 1;
    i = std::min<int64_t>(std::max<int64_t>(i, 0), T.W - 1);
    r = std::min<int64_t>(std::max<int64_t>(r, 0), T.H - 1);
    R pc = T.pdf[(size_t)r * T.W + (size_t)i];
    return pc * (R(T.W) * R(T.H)) / (R(2) * pi_v<R>() * pi_v<R>() * sin_t);
}

template <class R> R env_pdf_dir(const EnvTables<R>& T, V3<R> dir) {
    V3<R> d = normalize(dir);
    R th = std::acos(-d.y);
    R ph = std::atan2(-d.z, d.x) + pi_v<R>();
    R x = (ph / (R(2) * pi_v<R>())) * R(T.W);
    R y = R(T.H - 1) - (th / pi_v<R>()) * R(T.H);
    return env_pdf_omega(T, x, y, std::sin(th));
}

template <class R> V3<R> env_sample_dir(const EnvTables<R>& T, R u1, R u2, R& pdf) {
    uint32_t r = cdf_find(T.marg.data(), T.H, u1);
    R m0 = T.marg[r], m1 = T.marg[r + 1];
    const R* c = &T.cond[(size_t)r * (T.W + 1)];
    uint32_t i = cdf_find(c, T.W, u2);
    R c0 = c[i], c1 = c[i + 1];
    R fx = c1 > c0 ? (u2 - c0) / (c1 - c0) : R(0.5);
    R fy = m1 > m0 ? (u1 - m0) / (m1 - m0) : R(0.5);
    R top = R(0.99999994f);
    fx = std::min(std::max(fx, R(0)), top);
    fy = std::min(std::max(fy, R(0)), top);
    R x = R(i) + fx, y = (R(r) - R(1)) + fy;
    R th = pi_v<R>() * (R(T.H) - R(1) - y) / R(T.H);
    R ph = R(2) * pi_v<R>() * x / R(T.W);
    R st = std::sin(th), ct = std::cos(th);
    R sp = std::sin(ph - pi_v<R>()), cp = std::cos(ph - pi_v<R>());
    pdf = env_pdf_omega(T, x, y, st);
    return v3(st * cp, -ct, -st * sp);
}

// Extended estimator: forward throughput, environment NEE with MIS (power
// heuristic) at Lambertian vertices, Russian roulette; firefly clamp off.
// Same order of random draws per bounce as the GPU: [NEE: 2] [material] [RR: 1].
template <class R>
V3<R> trace_ext(const Scene<R>& s, const EnvTables<R>* T, const vr_render_desc& d, Ray<R> ray, PathRng& rng,
                RenderCounters& rc, bool prune) {
    V3<R> L = v3<R>(0, 0, 0), A = v3<R>(1, 1, 1);
    R pdf_prev = R(0);
    bool env_is = (d.flags & VR_RENDER_ENV_IS) && T && s.env_kind == VR_ENV_HDRI;
    for (uint32_t depth = 0; depth < d.max_bounces; ++depth) {
        Hit<R> hit;
        ++rc.segments;
        if (!scene_hit(s, ray, R(0.00001f), hit, prune, rc.q)) {
            ++rc.misses;
            if (s.env_kind == VR_ENV_NONE) break;
            V3<R> env = env_sample(s, ray);
            R w = R(1);
            if (env_is && pdf_prev > R(0)) {
                R pe = env_pdf_dir(*T, ray.d);
                w = (pdf_prev * pdf_prev) / (pdf_prev * pdf_prev + pe * pe);
            }
            L = L + mulv(A, env) * w;
            break;
        }
        const Material<R>& m = s.materials[s.objects[hit.surface].material];
        if (m.kind == VR_MAT_EMISSION) {
            L = L + mulv(A, m.color);
            break;
        }
        if (m.kind == VR_MAT_LAMBERTIAN && env_is) {
            R u1 = R(rng.uniform()), u2 = R(rng.uniform());
            V3<R> ns = m.normal_tex >= 0 ? tex_sample(s.textures[m.normal_tex], hit.uv.x, hit.uv.y) : hit.normal;
            ns = normalize(ns);
            R pe = 0;
            V3<R> wi = env_sample_dir(*T, u1, u2, pe);
            R cw = dot(ns, wi);
            if (cw > R(0) && pe > R(0)) {
                R pb = cw / pi_v<R>();
                R w = (pe * pe) / (pe * pe + pb * pb);
                V3<R> albedo = m.albedo_tex >= 0 ? tex_sample(s.textures[m.albedo_tex], hit.uv.x, hit.uv.y) : m.color;
                V3<R> contrib = mulv(mulv(A, albedo), env_sample(s, Ray<R>{hit.point, wi})) * (pb * w / pe);
                Hit<R> sh;
                ++rc.segments;
                if (!scene_hit(s, Ray<R>{hit.point, wi}, R(0.00001f), sh, true, rc.q)) L = L + contrib;
            }
        }
        V3<R> lobe = v3<R>(0, 0, 0);
        Scatter<R> sc = scatter(s, m, ray, hit, rng, rc, true, &lobe);
        if (!sc.has_ray || depth + 1 >= d.max_bounces) break;
        A = mulv(A, sc.att);
        // a non-finite throughput (0 * inf: a zero BSDF value over an underflowed pdf) ends the path
        if (!std::isfinite(A.x) || !std::isfinite(A.y) || !std::isfinite(A.z)) break;
        R pb = m.kind == VR_MAT_LAMBERTIAN ? std::max(R(0), dot(lobe, sc.ray.d)) / pi_v<R>() : R(0);
        if ((d.flags & VR_RENDER_RUSSIAN_ROULETTE) && depth >= d.rr_depth) {
            R q = std::min(R(0.95f), std::max(A.x, std::max(A.y, A.z)));
            if (R(rng.uniform()) >= q) break;
            A = A / q;
        }
        pdf_prev = pb;
        ray = sc.ray;
    }
    return L;
}

template <class R> struct RenderJob {
    const Scene<R>* s;
    const vr_render_desc* d;
    R* rgba;
    uint32_t row_step;
    bool prune;
    const EnvTables<R>* env_tables;
};

// iterative_render per-pixel body (render/iterative.rs:25-51)
template <class R> void render_pixel(const RenderJob<R>& job, uint32_t index, RenderCounters& rc) {
    const vr_render_desc& d = *job.d;
    uint32_t W = d.width, H = d.height;
    uint32_t x = index % W;
    uint32_t y = (d.flags & VR_RENDER_ROWS_BY_HEIGHT) ? index / H : index / W;
    R dd = R(std::max(W, H));
    R xf = (R(2 * x + 1) - R(W)) / dd;
    R yf = (R((uint32_t)(2u * (uint32_t)(H - y) - 1u)) - R(H)) / dd;
    RenderParams p{d.max_bounces, d.firefly_clamp, d.render_mode};
    V3<R> color = v3<R>(0, 0, 0);
    R lo = R(-1) / dd, hi = R(1) / dd;
    bool ext = (d.flags & (VR_RENDER_ENV_IS | VR_RENDER_RUSSIAN_ROULETTE)) != 0;
    for (uint32_t k = 0; k < d.samples; ++k) {
        PathRng rng(d.seed, d.request, index, d.sample_offset + k);
        R dx = R(rng.uniform()) * (hi - lo) + lo;
        R dy = R(rng.uniform()) * (hi - lo) + lo;
        Ray<R> r = cast_ray(*job.s, xf + dx, yf + dy, rng, rc);
        if (ext) color = color + trace_ext(*job.s, job.env_tables, d, r, rng, rc, job.prune);
        else color = color + trace_rec(*job.s, p, r, 0, rng, rc, job.prune);
    }
    color = color * (R(1) / R(d.total_samples));
    R* px = job.rgba + 4 * (size_t)index;
    px[0] += color.x;
    px[1] += color.y;
    px[2] += color.z;
    if (d.flags & VR_RENDER_ALPHA_PER_CALL) px[3] += R(1);
}

template <class R> void render_all(const RenderJob<R>& job, int nthreads, RenderCounters& total) {
    uint32_t n = job.d->width * job.d->height;
    std::atomic<uint32_t> next{0};
    std::vector<RenderCounters> per(std::max(1, nthreads));
    auto worker = [&](int tid) {
        const uint32_t chunk = 16;
        for (;;) {
            uint32_t b = next.fetch_add(chunk);
            if (b >= n) break;
            for (uint32_t i = b; i < std::min(n, b + chunk); ++i) {
                uint32_t y = (job.d->flags & VR_RENDER_ROWS_BY_HEIGHT) ? i / job.d->height : i / job.d->width;
                if (job.row_step > 1 && (y % job.row_step) != 0) continue;
                render_pixel(job, i, per[tid]);
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nthreads; ++t) th.emplace_back(worker, t);
    worker(0);
    for (auto& t : th) t.join();
    for (auto& p : per) {
        total.segments += p.segments;
        total.misses += p.misses;
        total.capped += p.capped;
        total.q.boxes += p.q.boxes;
        total.q.tris += p.q.tris;
    }
}

template <class R> void trace_all(const Scene<R>* s, uint32_t n, const float* o, const float* d, float tmin, int any,
                                  uint32_t* prim, R* t, R* u, R* v, uint8_t* hit, int nthreads, bool prune,
                                  Counters* tot) {
    // global prim numbering
    std::vector<uint32_t> surf_base(s->surface_mesh.size());
    uint32_t base = 0;
    for (size_t i = 0; i < surf_base.size(); ++i) {
        surf_base[i] = base;
        base += s->surface_prims((uint32_t)i);
    }
    std::atomic<uint32_t> next{0};
    std::vector<Counters> per(std::max(1, nthreads));
    auto worker = [&](int tid) {
        const uint32_t chunk = 64;
        for (;;) {
            uint32_t b = next.fetch_add(chunk);
            if (b >= n) break;
            for (uint32_t i = b; i < std::min(n, b + chunk); ++i) {
                Ray<R> r{load3<R>(o, i), load3<R>(d, i)};
                Hit<R> h;
                bool ok = scene_hit(*s, r, R(tmin), h, prune, per[tid]);
                if (hit) hit[i] = ok ? 1 : 0;
                (void)any;
                if (prim) prim[i] = ok ? surf_base[h.surface] + h.tri : 0xFFFFFFFFu;
                if (t) t[i] = ok ? h.t : R(0);
                if (u) u[i] = ok ? h.u : R(0);
                if (v) v[i] = ok ? h.v : R(0);
            }
        }
    };
    std::vector<std::thread> th;
    for (int k = 1; k < nthreads; ++k) th.emplace_back(worker, k);
    worker(0);
    for (auto& x : th) x.join();
    if (tot)
        for (auto& p : per) {
            tot->boxes += p.boxes;
            tot->tris += p.tris;
        }
}

int hw_threads(int n) {
    if (n > 0) return n;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

}  // namespace

extern "C" {

void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox4x32_10(ctr, key, out); }

// First `count` words of the path stream (seed, request, pixel, sample).
void orc_path_words(uint64_t seed, uint32_t request, uint32_t pixel, uint32_t sample, uint32_t count, uint32_t* out) {
    PathRng r(seed, request, pixel, sample);
    for (uint32_t i = 0; i < count; ++i) out[i] = r.next_u32();
}

void orc_unit_sphere_f32(uint64_t seed, uint32_t count, float* out) {
    uint64_t capped = 0;
    for (uint32_t i = 0; i < count; ++i) {
        PathRng r(seed, 0, i, 0);
        V3<float> p = unit_sphere<float>(r, capped);
        out[3 * i] = p.x;
        out[3 * i + 1] = p.y;
        out[3 * i + 2] = p.z;
    }
}

// MicrofacetBSDF pieces for the oracle's own tests.  mat: index, roughness,
// metallic, transparent, r, g, b.
static Material<double> microfacet_material(const double* mat) {
    Material<double> m{};
    m.kind = VR_MAT_MICROFACET;
    m.index = mat[0];
    m.roughness = mat[1];
    m.metallic = mat[2];
    m.transparent = mat[3] != 0.0;
    m.color = v3<double>(mat[4], mat[5], mat[6]);
    return m;
}
void orc_microfacet_bsdf_f64(const double* mat, const double* n, const double* wo, const double* wi, double* out) {
    Material<double> m = microfacet_material(mat);
    V3<double> f = microfacet_bsdf(m, v3<double>(n[0], n[1], n[2]), v3<double>(wo[0], wo[1], wo[2]),
                                   v3<double>(wi[0], wi[1], wi[2]));
    out[0] = f.x;
    out[1] = f.y;
    out[2] = f.z;
}
// count samples (stream: pixel i of the given seed); pdf[i] = 0 for None
void orc_microfacet_sample_f64(const double* mat, const double* n, const double* wo, uint64_t seed, uint32_t count,
                               double* wi, double* pdf) {
    Material<double> m = microfacet_material(mat);
    uint64_t capped = 0;
    for (uint32_t i = 0; i < count; ++i) {
        PathRng r(seed, 0, i, 0);
        V3<double> w;
        double p = 0.0;
        if (!microfacet_sample(m, v3<double>(n[0], n[1], n[2]), v3<double>(wo[0], wo[1], wo[2]), r, capped, w, p)) {
            w = v3<double>(0, 0, 0);
            p = 0.0;
        }
        wi[3 * i] = w.x;
        wi[3 * i + 1] = w.y;
        wi[3 * i + 2] = w.z;
        pdf[i] = p;
    }
}

void* orc_scene_new_f32(const vr_scene_desc* d) { return make_scene<float>(d); }
void* orc_scene_new_f64(const vr_scene_desc* d) { return make_scene<double>(d); }
void orc_scene_free_f32(void* s) { delete (Scene<float>*)s; }
void orc_scene_free_f64(void* s) { delete (Scene<double>*)s; }

uint32_t orc_scene_num_prims(void* s) { return ((Scene<float>*)s)->num_prims; }

// In-order leaf sequence of mesh m's reference tree (empty for small meshes).
uint32_t orc_mesh_leaf_order(void* sp, uint32_t m, uint32_t* out) {
    auto* s = (Scene<float>*)sp;
    const Mesh<float>& mesh = s->meshes[m];
    if (mesh.root < 0) return 0;
    uint32_t n = 0;
    std::vector<int32_t> st{mesh.root};
    while (!st.empty()) {
        int32_t i = st.back();
        st.pop_back();
        const auto& nd = mesh.tree[i];
        if (nd.obj >= 0) {
            if (out) out[n] = (uint32_t)nd.obj;
            ++n;
            continue;
        }
        st.push_back(nd.right);
        st.push_back(nd.left);
    }
    return n;
}

// Flat dump of mesh m's reference tree: per node (min xyz, max xyz) and
// (left, right, obj).  Returns the node count.
uint32_t orc_mesh_tree(void* sp, uint32_t m, float* boxes, int64_t* links) {
    auto* s = (Scene<float>*)sp;
    const Mesh<float>& mesh = s->meshes[m];
    for (size_t i = 0; i < mesh.tree.size(); ++i) {
        const auto& nd = mesh.tree[i];
        if (boxes) {
            float* b = boxes + 6 * i;
            b[0] = nd.box.min.x; b[1] = nd.box.min.y; b[2] = nd.box.min.z;
            b[3] = nd.box.max.x; b[4] = nd.box.max.y; b[5] = nd.box.max.z;
        }
        if (links) {
            links[3 * i] = nd.left;
            links[3 * i + 1] = nd.right;
            links[3 * i + 2] = nd.obj;
        }
    }
    return (uint32_t)mesh.tree.size();
}

void orc_trace_f32(void* s, uint32_t n, const float* o, const float* d, float tmin, int any, uint32_t* prim, float* t,
                   float* u, float* v, uint8_t* hit, int nthreads, uint64_t* boxes_tris) {
    Counters c;
    trace_all<float>((Scene<float>*)s, n, o, d, tmin, any, prim, t, u, v, hit, hw_threads(nthreads), false, &c);
    if (boxes_tris) {
        boxes_tris[0] = c.boxes;
        boxes_tris[1] = c.tris;
    }
}

// Brute force over every triangle, no boxes at all, with the reference tie
// rules resolved by the given rank (higher rank wins equal t).
void orc_trace_bruteforce_f32(void* sp, uint32_t n, const float* o, const float* d, float tmin, const uint32_t* rank,
                              uint32_t* prim, float* t, float* u, float* v, int nthreads) {
    auto* s = (Scene<float>*)sp;
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        for (;;) {
            uint32_t i = next.fetch_add(1);
            if (i >= n) break;
            Ray<float> r{load3<float>(o, i), load3<float>(d, i)};
            bool have = false;
            Hit<float> best;
            uint32_t best_p = 0, base = 0;
            for (size_t sf = 0; sf < s->surface_mesh.size(); ++sf) {
                if (s->is_analytic((uint32_t)sf)) {
                    Hit<float> h;
                    if (analytic_hit(s->analytic_of((uint32_t)sf), r, tmin, h) &&
                        (!have || h.t < best.t || (h.t == best.t && rank[base] > rank[best_p]))) {
                        best = h;
                        best_p = base;
                        have = true;
                    }
                    base += 1;
                    continue;
                }
                const Mesh<float>& m = s->meshes[s->surface_mesh[sf]];
                uint32_t nt = (uint32_t)(m.idx.size() / 3);
                for (uint32_t k = 0; k < nt; ++k) {
                    Hit<float> h;
                    if (!tri_hit(m, k, r, tmin, h)) continue;
                    uint32_t p = base + k;
                    if (!have || h.t < best.t || (h.t == best.t && rank[p] > rank[best_p])) {
                        best = h;
                        best_p = p;
                        have = true;
                    }
                }
                base += nt;
            }
            prim[i] = have ? best_p : 0xFFFFFFFFu;
            t[i] = have ? best.t : 0.f;
            u[i] = have ? best.u : 0.f;
            v[i] = have ? best.v : 0.f;
        }
    };
    std::vector<std::thread> th;
    int nt = hw_threads(nthreads);
    for (int k = 1; k < nt; ++k) th.emplace_back(worker);
    worker();
    for (auto& x : th) x.join();
}

// iterative_render with reference semantics (f32).  row_step > 1 renders only
// rows y % row_step == 0 (bounded CPU-baseline samples).  stats (optional):
// segments, misses, capped, boxes, tris.
double orc_render_f32(void* s, const vr_render_desc* d, float* rgba, int nthreads, uint32_t row_step, uint64_t* stats) {
    auto* sc = (Scene<float>*)s;
    std::unique_ptr<EnvTables<float>> T;
    if (sc->env_kind == VR_ENV_HDRI) T.reset(new EnvTables<float>(make_env_tables(sc->env)));
    RenderJob<float> job{sc, d, rgba, row_step, false, T.get()};
    RenderCounters rc;
    auto t0 = std::chrono::steady_clock::now();
    render_all(job, hw_threads(nthreads), rc);
    double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (stats) {
        stats[0] = rc.segments;
        stats[1] = rc.misses;
        stats[2] = rc.capped;
        stats[3] = rc.q.boxes;
        stats[4] = rc.q.tris;
    }
    return sec;
}

double orc_render_f64(void* s, const vr_render_desc* d, double* rgba, int nthreads, uint32_t row_step, int prune,
                      uint64_t* stats) {
    auto* sc = (Scene<double>*)s;
    std::unique_ptr<EnvTables<double>> T;
    if (sc->env_kind == VR_ENV_HDRI) T.reset(new EnvTables<double>(make_env_tables(sc->env)));
    RenderJob<double> job{sc, d, rgba, row_step, prune != 0, T.get()};
    RenderCounters rc;
    auto t0 = std::chrono::steady_clock::now();
    render_all(job, hw_threads(nthreads), rc);
    double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (stats) {
        stats[0] = rc.segments;
        stats[1] = rc.misses;
        stats[2] = rc.capped;
        stats[3] = rc.q.boxes;
        stats[4] = rc.q.tris;
    }
    return sec;
}

// ---- tonemapping (shaders/post_process.glsl:23-49, tonemapping.glsl:2-40) --

static void aces(float c[3]) {
    // GLSL mat3 constructors are column-major.
    float r = c[0], g = c[1], b = c[2];
    float x = 0.59719f * r + 0.35458f * g + 0.04823f * b;
    float y = 0.076f * r + 0.90834f * g + 0.01566f * b;
    float z = 0.0284f * r + 0.13383f * g + 0.83777f * b;
    float v[3] = {x, y, z};
    for (int i = 0; i < 3; ++i) v[i] = (v[i] * (v[i] + 0.0245786f) - 0.000090537f) / (v[i] * (0.983729f * v[i] + 0.432951f) + 0.238081f);
    c[0] = 1.60475f * v[0] + -0.53108f * v[1] + -0.07367f * v[2];
    c[1] = -0.10208f * v[0] + 1.10813f * v[1] + -0.00605f * v[2];
    c[2] = -0.00327f * v[0] + -0.07276f * v[1] + 1.07602f * v[2];
}
static void reinhard(float c[3]) {
    float white = 2.f;
    float luma = c[0] * 0.2126f + c[1] * 0.7152f + c[2] * 0.0722f;
    float tl = luma * (1.f + luma / (white * white)) / (1.f + luma);
    for (int i = 0; i < 3; ++i) c[i] *= tl / luma;
}
static void filmic(float c[3]) {
    for (int i = 0; i < 3; ++i) {
        float x = std::fmax(0.f, c[i] - 0.004f);
        c[i] = (x * (6.2f * x + .5f)) / (x * (6.2f * x + 1.7f) + 0.06f);
    }
}
static void uncharted2(float c[3]) {
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f, W = 11.2f, ex = 2.f;
    float white = ((W * (A * W + C * B) + D * E) / (W * (A * W + B) + D * F)) - E / F;
    for (int i = 0; i < 3; ++i) {
        float x = c[i] * ex;
        x = ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
        c[i] = x / white;
    }
}

// The legacy viewer's curves (voidray/src/shaders/tonemapping.glsl:17-45).
static void simple_reinhard(float c[3]) {
    const float exposure = 1.5f;
    for (int i = 0; i < 3; ++i) c[i] *= exposure / (1.f + c[i] / exposure);
}
static void luma_reinhard(float c[3]) {
    float luma = c[0] * 0.2126f + c[1] * 0.7152f + c[2] * 0.0722f;
    float tl = luma / (1.f + luma);
    for (int i = 0; i < 3; ++i) c[i] *= tl / luma;
}
static void rombindahouse(float c[3]) {
    for (int i = 0; i < 3; ++i) c[i] = std::exp(-1.0f / (2.72f * c[i] + 0.15f));
}

// voidray/src/shaders/viewport_fragment.glsl:27-67.
static void legacy_post(const vr_post_desc* p, const float* s, float* d) {
    float c[3];
    for (int k = 0; k < 3; ++k) c[k] = (s[k] * p->scale) * p->exposure;
    float gamma = p->gamma;
    switch (p->tonemap) {
        case VR_TONEMAP_LEGACY_SIMPLE_ACES: aces(c); break;
        case VR_TONEMAP_LEGACY_SIMPLE_REINHARD: simple_reinhard(c); break;
        case VR_TONEMAP_LEGACY_LUMA_REINHARD: luma_reinhard(c); break;
        case VR_TONEMAP_LEGACY_LUMA_WHITE_PRESERVING_REINHARD: reinhard(c); break;
        case VR_TONEMAP_LEGACY_ROM_BIN_DA_HOUSE: rombindahouse(c); break;
        case VR_TONEMAP_LEGACY_FILMIC:
            filmic(c);
            gamma = 1.f;
            break;
        case VR_TONEMAP_LEGACY_UNCHARTED2: uncharted2(c); break;
        default: break;
    }
    for (int k = 0; k < 3; ++k) d[k] = std::pow(c[k], 1.f / gamma);
    d[3] = p->keep_alpha ? s[3] * p->scale : 1.f;
}

// f32 evaluation of the published formulas (the tests pin it against FP64
// closed forms).
void orc_tonemap(const vr_post_desc* p, const float* src, float* dst) {
    uint32_t W = p->width, H = p->height;
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            size_t i = (size_t)y * W + x;
            if (p->clip_to_1024 && (x >= 1024 || y >= 1024)) continue;
            if (p->tonemap >= VR_TONEMAP_LEGACY_NONE) {
                legacy_post(p, src + 4 * i, dst + 4 * i);
                continue;
            }
            float c[3];
            float ex = std::pow(2.f, p->exposure);
            for (int k = 0; k < 3; ++k) c[k] = (src[4 * i + k] * p->scale) * ex;
            switch (p->tonemap) {
                case VR_TONEMAP_ACES: aces(c); break;
                case VR_TONEMAP_REINHARD: reinhard(c); break;
                case VR_TONEMAP_FILMIC: filmic(c); break;
                case VR_TONEMAP_UNCHARTED2: uncharted2(c); break;
                default: break;
            }
            for (int k = 0; k < 3; ++k) dst[4 * i + k] = std::pow(c[k], 1.f / p->gamma);
            dst[4 * i + 3] = 1.f;
        }
}

int orc_num_threads(void) { return hw_threads(0); }

// ---- pieces of the path exported for their closed-form pins (tests only) ----

// Dielectric::reflectance (simple.rs:192-197), reflect / refract (math.rs:12-21).
float orc_schlick_f32(float c, float idx) { return schlick<float>(c, idx); }
void orc_reflect_f32(const float* v, const float* n, float* out) {
    V3<float> r = reflect(load3<float>(v, 0), load3<float>(n, 0));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
void orc_refract_f32(const float* uv, const float* n, float eta, float* out) {
    V3<float> r = refract(load3<float>(uv, 0), load3<float>(n, 0), eta);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
// AABB::hit (aabb.rs:86-148) on one box.
int orc_aabb_hit_f32(const float* lo, const float* hi, const float* o, const float* d, float tmin, float tmax) {
    AABB<float> b{load3<float>(lo, 0), load3<float>(hi, 0)};
    Ray<float> r{load3<float>(o, 0), load3<float>(d, 0)};
    return aabb_hit(b, r, tmin, tmax) ? 1 : 0;
}
// Material::scatter for one hit record, random stream (seed, 0, pixel,
// sample) as a render path draws it.  hit: point[3], normal[3] (the record's,
// already facing the ray), uv[2], t; returns 1 when a scattered ray exists.
int orc_scatter_f32(const vr_material* mat, const float* ro, const float* rd, const float* hit_point,
                    const float* hit_normal, int front_face, const float* uv, float t, uint64_t seed, uint32_t pixel,
                    uint32_t sample, float* att, float* out_o, float* out_d) {
    Scene<float> sc;
    Material<float> m{mat->kind, load3<float>(mat->color, 0), mat->param, -1, -1, mat->index, mat->roughness,
                      mat->metallic, mat->transparent != 0};
    Ray<float> ray{load3<float>(ro, 0), load3<float>(rd, 0)};
    Hit<float> h{};
    h.point = load3<float>(hit_point, 0);
    h.normal = load3<float>(hit_normal, 0);
    h.front_face = front_face != 0;
    h.uv = V2<float>{uv[0], uv[1]};
    h.t = t;
    PathRng rng(seed, 0, pixel, sample);
    RenderCounters rc;
    Scatter<float> out = scatter(sc, m, ray, h, rng, rc);
    for (int a = 0; a < 3; ++a) {
        att[a] = (&out.att.x)[a];
        out_o[a] = out.has_ray ? (&out.ray.o.x)[a] : 0.f;
        out_d[a] = out.has_ray ? (&out.ray.d.x)[a] : 0.f;
    }
    return out.has_ray ? 1 : 0;
}
// CameraAcceleration::cast_ray (camera.rs:69-82) at NDC (x, y); the thin-lens
// disc sample, if any, is the first draw of stream (seed, 0, pixel, 0).
void orc_cast_ray_f32(void* sp, float x, float y, uint64_t seed, uint32_t pixel, float* o, float* d) {
    PathRng rng(seed, 0, pixel, 0);
    RenderCounters rc;
    Ray<float> r = cast_ray(*(Scene<float>*)sp, x, y, rng, rc);
    for (int a = 0; a < 3; ++a) {
        o[a] = (&r.o.x)[a];
        d[a] = (&r.d.x)[a];
    }
}
// Pixel -> NDC of iterative.rs:25-33 (the pixel centre, before jitter).
void orc_pixel_ndc_f32(uint32_t W, uint32_t H, uint32_t index, uint32_t flags, float* xy) {
    uint32_t x = index % W;
    uint32_t y = (flags & VR_RENDER_ROWS_BY_HEIGHT) ? index / H : index / W;
    float dd = (float)std::max(W, H);
    xy[0] = ((float)(2 * x + 1) - (float)W) / dd;
    xy[1] = ((float)((uint32_t)(2u * (uint32_t)(H - y) - 1u)) - (float)H) / dd;
}
// Environment importance sampling in FP64: per-cell pdf (H*W), the solid-angle
// pdf of directions, and directions drawn from (u1, u2).
void orc_env_tables_f64(uint32_t w, uint32_t h, const float* data, double* pdf_cell) {
    Texture<double> env;
    vr_texture t{w, h, VR_TEX_RGB32F, VR_SAMPLE_BILINEAR, data};
    make_texture(env, t);
    EnvTables<double> T = make_env_tables(env);
    for (size_t i = 0; i < T.pdf.size(); ++i) pdf_cell[i] = T.pdf[i];
}
void orc_env_pdf_dir_f64(uint32_t w, uint32_t h, const float* data, uint32_t n, const double* dirs, double* pdf) {
    Texture<double> env;
    vr_texture t{w, h, VR_TEX_RGB32F, VR_SAMPLE_BILINEAR, data};
    make_texture(env, t);
    EnvTables<double> T = make_env_tables(env);
    for (uint32_t i = 0; i < n; ++i) pdf[i] = env_pdf_dir(T, v3<double>(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}
void orc_env_sample_dir_f64(uint32_t w, uint32_t h, const float* data, uint32_t n, const double* u, double* dirs,
                            double* pdf) {
    Texture<double> env;
    vr_texture t{w, h, VR_TEX_RGB32F, VR_SAMPLE_BILINEAR, data};
    make_texture(env, t);
    EnvTables<double> T = make_env_tables(env);
    for (uint32_t i = 0; i < n; ++i) {
        V3<double> dv = env_sample_dir(T, u[2 * i], u[2 * i + 1], pdf[i]);
        dirs[3 * i] = dv.x;
        dirs[3 * i + 1] = dv.y;
        dirs[3 * i + 2] = dv.z;
    }
}
// UnitDisc (square rejection), first draw of stream (seed, 0, i, 0).
void orc_unit_disc_f32(uint64_t seed, uint32_t count, float* out) {
    uint64_t capped = 0;
    for (uint32_t i = 0; i < count; ++i) {
        PathRng r(seed, 0, i, 0);
        V2<float> p = unit_disc<float>(r, capped);
        out[2 * i] = p.x;
        out[2 * i + 1] = p.y;
    }
}

}  // extern "C"

extern "C" {
// Direct lookups for known-answer tests of texture.rs:82-98 and environments.rs:80-90.
void orc_tex_sample_f32(const vr_texture* t, float u, float v, float* out) {
    Texture<float> tex;
    make_texture(tex, *t);
    V3<float> c = tex_sample(tex, u, v);
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}
void orc_env_sample_f32(uint32_t w, uint32_t h, const float* data, const float* dir, float* out) {
    Scene<float> s;
    s.env_kind = VR_ENV_HDRI;
    vr_texture t{w, h, VR_TEX_RGB32F, VR_SAMPLE_BILINEAR, data};
    make_texture(s.env, t);
    Ray<float> r{v3<float>(0, 0, 0), v3<float>(dir[0], dir[1], dir[2])};
    V3<float> c = env_sample(s, r);
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}
}
