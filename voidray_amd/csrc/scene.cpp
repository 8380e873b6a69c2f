// Scene flattening + upload: the MI355X counterpart of
// Scene::build_acceleration (voidray_renderer/src/core/scene.rs:163-179) and
// the per-mesh BVH build of Mesh::from_buffers (core/mesh.rs:76-116).
//
// HBM layout (all arrays 256-byte aligned, one allocation each):
//   qnodes   QNode[num_qnodes]      64 B  SAH 4-wide with spatial splits, 8-bit outward-quantised
//   tris     TriRec[R]              48 B  world vertices in leaf order + prim id (R >= T:
//                                         one record per leaf holding a piece of a split triangle)
//   ref/top  RefNode[]              32 B  reference median-split trees (exact engine)
//   ref_parent, prim_anc            4 B   reference-tree parents, per prim its leaf's parent
//   per mesh triangle: idx (12 B), face normal (12 B)
//   per vertex: position, normal (12 B each), uv (8 B)
//   materials, textures (RGBA8 or RGB f32 texels), HDRI (RGB f32) + IS tables
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "bvh_build.h"
#include "envmap.h"
#include "internal.h"

namespace vr {

namespace {
template <class T> int upload(vr_scene* s, const std::vector<T>& host, const T** dev) {
    *dev = nullptr;
    if (host.empty()) return 0;
    void* p = nullptr;
    size_t bytes = host.size() * sizeof(T);
    if (hipMalloc(&p, bytes) != hipSuccess) return fail(VR_ERR_OOM, "hipMalloc failed for scene array");
    s->allocations.push_back(p);
    s->stats.device_bytes += bytes;
    VR_HIP_CHECK(hipMemcpy(p, host.data(), bytes, hipMemcpyHostToDevice));
    *dev = (const T*)p;
    return 0;
}

vf3 ld3(const float* p, size_t i) { return mk3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
}  // namespace

int build_host_scene(const vr_scene_desc* d, HostScene& H) {
    using clk = std::chrono::steady_clock;
    SceneView& v = H.view;
    if (!d) return fail(VR_ERR_INVALID, "null scene description");
    if (d->num_surfaces > 0 && !d->surface_mesh) return fail(VR_ERR_INVALID, "surfaces without surface_mesh");
    if (d->num_objects < d->num_surfaces)
        return fail(VR_ERR_INVALID, "objects[] must cover every surface index (scene.rs:184 looks up objects[surface])");
    for (uint32_t i = 0; i < d->num_surfaces; ++i) {
        const uint32_t sm = d->surface_mesh[i];
        if (!(sm & VR_SURFACE_ANALYTIC) && (sm & VR_SURFACE_INSTANCE)) {
            const uint32_t k = sm & ~VR_SURFACE_INSTANCE;
            if (d->num_instances >= (1u << 29)) return fail(VR_ERR_INVALID, "too many instances (2^29)");
            if (!d->instances || k >= d->num_instances) return fail(VR_ERR_INVALID, "instance index out of range");
            if (!d->meshes || d->instances[k].mesh >= d->num_meshes)
                return fail(VR_ERR_INVALID, "instance mesh index out of range");
            for (int e = 0; e < 12; ++e)
                if (!std::isfinite(d->instances[k].transform[e]))
                    return fail(VR_ERR_INVALID, "instance transform is not finite");
            continue;
        }
        if (sm & VR_SURFACE_ANALYTIC) {
            const uint32_t k = sm & ~VR_SURFACE_ANALYTIC;
            if (!d->analytic || k >= d->num_analytic) return fail(VR_ERR_INVALID, "analytic surface index out of range");
            const int32_t kind = d->analytic[k].kind;
            if (kind != VR_ANALYTIC_SPHERE && kind != VR_ANALYTIC_GROUND_PLANE)
                return fail(VR_ERR_INVALID, "unknown analytic surface kind");
        } else if (!d->meshes || sm >= d->num_meshes) {
            return fail(VR_ERR_INVALID, "surface_mesh index out of range");
        }
    }
    for (uint32_t i = 0; i < d->num_surfaces; ++i)
        if (d->objects[i].material >= d->num_materials) return fail(VR_ERR_INVALID, "object material out of range");
    for (uint32_t i = 0; i < d->num_materials; ++i) {
        const vr_material& m = d->materials[i];
        if (m.kind < 0 || m.kind > VR_MAT_MICROFACET) return fail(VR_ERR_INVALID, "unknown material kind");
        if (m.albedo_texture >= (int32_t)d->num_textures || m.normal_texture >= (int32_t)d->num_textures)
            return fail(VR_ERR_INVALID, "material texture index out of range");
    }
    for (uint32_t i = 0; i < d->num_textures; ++i) {
        const vr_texture& t = d->textures[i];
        if (!t.data || t.width == 0 || t.height == 0) return fail(VR_ERR_INVALID, "empty texture");
        if (t.format != VR_TEX_RGB32F && t.format != VR_TEX_RGBA8) return fail(VR_ERR_INVALID, "unknown texture format");
    }
    auto t_ref0 = clk::now();

    // ---- meshes: concatenated vertices / triangles, reference trees -------
    uint32_t M = d->num_meshes;
    std::vector<MeshInfo>& minfo = H.meshes;
    minfo.assign(M, MeshInfo{});
    std::vector<float>&vpos = H.vpos, &vnrm = H.vnrm, &vuv = H.vuv, &face_n = H.face_n;
    std::vector<uint32_t>& midx = H.mesh_idx;
    std::vector<RefNode>& ref_nodes = H.ref_nodes;
    std::vector<int32_t>&ref_parent = H.ref_parent, &leaf_parent = H.mesh_leaf_parent;
    std::vector<Box3> mesh_bounds(M);
    uint64_t tv = 0, tt = 0;
    for (uint32_t m = 0; m < M; ++m) {
        tv += d->meshes[m].num_vertices;
        tt += d->meshes[m].num_triangles;
    }
    if (tv >= (1ull << 32) || tt >= (1ull << 31)) return fail(VR_ERR_INVALID, "scene too large");
    vpos.reserve(3 * tv);
    vnrm.reserve(3 * tv);
    vuv.reserve(2 * tv);
    midx.reserve(3 * tt);
    face_n.reserve(3 * tt);
    for (uint32_t m = 0; m < M; ++m) {
        const vr_mesh& src = d->meshes[m];
        if (src.num_vertices && !src.positions) return fail(VR_ERR_INVALID, "mesh without positions");
        if (src.num_triangles && !src.indices) return fail(VR_ERR_INVALID, "mesh without indices");
        MeshInfo& mi = minfo[m];
        mi.vert_base = (uint32_t)(vpos.size() / 3);
        mi.tri_base = (uint32_t)(midx.size() / 3);
        mi.num_tris = src.num_triangles;
        mi.root = -1;
        Box3 mb{mk3(kInf, kInf, kInf), mk3(-kInf, -kInf, -kInf)};  // AABB::default
        for (uint32_t i = 0; i < src.num_vertices; ++i) {
            for (int a = 0; a < 3; ++a) {
                vpos.push_back(src.positions[3 * i + a]);
                vnrm.push_back(src.normals ? src.normals[3 * i + a] : 0.0f);
            }
            vuv.push_back(src.uvs ? src.uvs[2 * i] : 0.0f);
            vuv.push_back(src.uvs ? src.uvs[2 * i + 1] : 0.0f);
            vf3 p = ld3(src.positions, i);
            mb.lo = mk3(fminf(mb.lo.x, p.x), fminf(mb.lo.y, p.y), fminf(mb.lo.z, p.z));
            mb.hi = mk3(fmaxf(mb.hi.x, p.x), fmaxf(mb.hi.y, p.y), fmaxf(mb.hi.z, p.z));
        }
        mesh_bounds[m] = mb;
        std::vector<Box3> tb(src.num_triangles);
        for (uint32_t t = 0; t < src.num_triangles; ++t) {
            uint32_t i0 = src.indices[3 * t], i1 = src.indices[3 * t + 1], i2 = src.indices[3 * t + 2];
            if (i0 >= src.num_vertices || i1 >= src.num_vertices || i2 >= src.num_vertices)
                return fail(VR_ERR_INVALID, "triangle index out of range");
            midx.push_back(i0);
            midx.push_back(i1);
            midx.push_back(i2);
            vf3 p0 = ld3(src.positions, i0), p1 = ld3(src.positions, i1), p2 = ld3(src.positions, i2);
            // face normal (mesh.rs:80-84)
            vf3 fn = normalize3(cross3(p2 - p1, p0 - p1));
            face_n.push_back(fn.x);
            face_n.push_back(fn.y);
            face_n.push_back(fn.z);
            tb[t] = tri_box_eps(p0, p1, p2);
        }
        if (src.num_triangles > 4) {  // SMALL_MESH (mesh.rs:43, 111-114)
            std::vector<RefNode> tree;
            build_reference_tree(tb, tree);
            int32_t off = (int32_t)ref_nodes.size();
            for (auto& nd : tree)
                if (nd.obj < 0) nd.right += off;
            mi.root = off;
            ref_nodes.insert(ref_nodes.end(), tree.begin(), tree.end());
            // parents (depth-first layout: left child = k + 1) and each
            // triangle's leaf parent, for the exact ancestor walk
            ref_parent.resize(ref_nodes.size(), -1);
            leaf_parent.resize(midx.size() / 3, -1);
            for (size_t k = 0; k < tree.size(); ++k) {
                const RefNode& nd = tree[k];
                int32_t g = off + (int32_t)k;
                if (nd.obj < 0) {
                    ref_parent[(size_t)g + 1] = g;
                    ref_parent[(size_t)nd.right] = g;
                }
            }
            for (size_t k = 0; k < tree.size(); ++k)
                if (tree[k].obj >= 0) leaf_parent[mi.tri_base + (uint32_t)tree[k].obj] = ref_parent[off + k];
        }
    }
    ref_parent.resize(ref_nodes.size(), -1);
    leaf_parent.resize(midx.size() / 3, -1);

    // ---- top-level tree over surfaces (scene.rs:72-92, 176) ---------------
    uint32_t S = d->num_surfaces;
    std::vector<Box3> sb(S);
    auto is_analytic = [&](uint32_t i) { return (d->surface_mesh[i] & VR_SURFACE_ANALYTIC) != 0; };
    auto is_instance = [&](uint32_t i) { return !is_analytic(i) && (d->surface_mesh[i] & VR_SURFACE_INSTANCE) != 0; };
    auto surface_mesh_of = [&](uint32_t i) {
        return is_instance(i) ? d->instances[d->surface_mesh[i] & ~VR_SURFACE_INSTANCE].mesh : d->surface_mesh[i];
    };
    auto surface_prims = [&](uint32_t i) { return is_analytic(i) ? 1u : minfo[surface_mesh_of(i)].num_tris; };
    // world-space vertices of an instance, with the flattening's f32 operation order
    auto xform = [](const float* xf, vf3 p) {
        return mk3(((xf[0] * p.x + xf[1] * p.y) + xf[2] * p.z) + xf[3],
                   ((xf[4] * p.x + xf[5] * p.y) + xf[6] * p.z) + xf[7],
                   ((xf[8] * p.x + xf[9] * p.y) + xf[10] * p.z) + xf[11]);
    };
    for (uint32_t i = 0; i < S; ++i) {
        if (is_instance(i)) {  // bounds of the transformed vertices, as the flattened mesh has them
            const vr_instance& in = d->instances[d->surface_mesh[i] & ~VR_SURFACE_INSTANCE];
            const MeshInfo& mi = minfo[in.mesh];
            const uint32_t nv = d->meshes[in.mesh].num_vertices;
            Box3 b{mk3(kInf, kInf, kInf), mk3(-kInf, -kInf, -kInf)};
            for (uint32_t k = 0; k < nv; ++k) {
                vf3 p = xform(in.transform, ld3(vpos.data(), mi.vert_base + k));
                b.lo = mk3(fminf(b.lo.x, p.x), fminf(b.lo.y, p.y), fminf(b.lo.z, p.z));
                b.hi = mk3(fmaxf(b.hi.x, p.x), fmaxf(b.hi.y, p.y), fmaxf(b.hi.z, p.z));
            }
            sb[i] = b;
            continue;
        }
        if (!is_analytic(i)) {
            sb[i] = mesh_bounds[d->surface_mesh[i]];
            continue;
        }
        const vr_analytic& an = d->analytic[d->surface_mesh[i] & ~VR_SURFACE_ANALYTIC];
        if (an.kind == VR_ANALYTIC_SPHERE) {  // Sphere::bounds (surfaces.rs:36-43)
            const float r = an.radius;
            sb[i].lo = mk3(an.center[0] - r, an.center[1] - r, an.center[2] - r);
            sb[i].hi = mk3(an.center[0] + r, an.center[1] + r, an.center[2] + r);
        } else {  // GroundPlane::bounds (surfaces.rs:106-113)
            sb[i].lo = mk3(-kInf, an.height - 0.0001f, -kInf);
            sb[i].hi = mk3(kInf, an.height + 0.0001f, kInf);
        }
    }
    std::vector<RefNode>& top = H.top_nodes;
    build_reference_tree(sb, top);
    std::vector<SurfaceInfo>& sinfo = H.surfaces;
    sinfo.assign(S, SurfaceInfo{});
    // Global prim ids run surface by surface, and within a mesh in triangle
    // order: on equal t the reference keeps the smaller handle (merge_optionals
    // compares (HitRecord, handle) tuples whose HitRecord part orders by t
    // alone, bvh.rs:143-175 and ray.rs:52-62; the brute-force mesh loop keeps
    // the earlier triangle, mesh.rs:131), first among surfaces (scene.rs:183),
    // then among the triangles of the winning mesh, so the winner of a tie is
    // the smallest prim id.
    uint32_t pb = 0;
    std::vector<AnalyticDev>& analytic = H.analytic;
    std::vector<InstanceDev>& instances = H.instances;
    for (uint32_t i = 0; i < S; ++i) {
        sinfo[i].mesh = is_analytic(i) ? kNoMesh : surface_mesh_of(i);
        sinfo[i].prim_base = pb;
        sinfo[i].material = d->objects[i].material;
        sinfo[i].analytic = -1;
        sinfo[i].instance = -1;
        if (is_instance(i)) {
            const vr_instance& in = d->instances[d->surface_mesh[i] & ~VR_SURFACE_INSTANCE];
            InstanceDev x{};
            for (int e = 0; e < 12; ++e) x.xf[e] = in.transform[e];
            // inverse of the 3x3 part and of the translation, in f64, rounded once
            double m[3][3], iv[3][3];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) m[r][c] = in.transform[4 * r + c];
            const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                               m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
            // (scale-free test: an invertible map of any size passes, one
            // whose inverse does not fit a float does not)
            bool invertible = det != 0.0 && std::isfinite(det);
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    const int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
                    iv[r][c] = invertible ? (m[r1][c1] * m[r2][c2] - m[r1][c2] * m[r2][c1]) / det : 0.0;
                    invertible = invertible && std::isfinite((float)iv[r][c]);
                }
            if (!invertible) return fail(VR_ERR_INVALID, "singular instance transform");
            for (int r = 0; r < 3; ++r) {
                double t = 0;
                for (int c = 0; c < 3; ++c) {
                    x.inv[4 * r + c] = (float)iv[r][c];
                    t -= iv[r][c] * (double)in.transform[4 * c + 3];
                    x.nrm[3 * r + c] = (float)iv[c][r];
                }
                x.inv[4 * r + 3] = (float)t;
            }
            x.face_sign = det < 0 ? -1.0f : 1.0f;
            x.det_inv = (float)(1.0 / std::fabs(det));
            x.mesh = in.mesh;
            x.surface = i;
            x.prim_base = pb;
            x.blas_root = kInvalidChild;
            sinfo[i].instance = (int32_t)instances.size();
            instances.push_back(x);
        }
        if (is_analytic(i)) {
            const vr_analytic& an = d->analytic[d->surface_mesh[i] & ~VR_SURFACE_ANALYTIC];
            sinfo[i].analytic = (int32_t)analytic.size();
            AnalyticDev a{};
            a.kind = an.kind;
            for (int k = 0; k < 3; ++k) a.c[k] = an.center[k];
            a.r = an.radius;
            a.h = an.height;
            a.surface = i;
            a.prim = pb;
            analytic.push_back(a);
        }
        pb += surface_prims(i);
    }
    uint32_t T = pb;
    // ancestor lists of every surface leaf (root -> parent)
    std::vector<uint32_t>& top_path = H.top_path;
    {
        std::vector<std::pair<int32_t, std::vector<uint32_t>>> st;
        if (!top.empty()) st.push_back({0, {}});
        std::vector<std::vector<uint32_t>> paths(S);
        while (!st.empty()) {
            auto [ni, path] = st.back();
            st.pop_back();
            const RefNode& nd = top[ni];
            if (nd.obj >= 0) {
                paths[nd.obj] = path;
                continue;
            }
            auto p2 = path;
            p2.push_back((uint32_t)ni);
            st.push_back({nd.right, p2});
            st.push_back({ni + 1, p2});
        }
        for (uint32_t i = 0; i < S; ++i) {
            sinfo[i].path_off = (uint32_t)top_path.size();
            sinfo[i].path_len = (uint32_t)paths[i].size();
            top_path.insert(top_path.end(), paths[i].begin(), paths[i].end());
        }
    }
    auto t_ref1 = clk::now();

    // ---- fast structure: triangle records + SAH quantised BVH4 ------------
    std::vector<TriRec> recs;
    recs.reserve(T);
    std::vector<uint32_t>& prim_surface = H.prim_surface;
    prim_surface.assign(T, 0);
    std::vector<int32_t>& prim_anc = H.prim_anc;
    prim_anc.assign(T, -1);
    for (uint32_t i = 0; i < S; ++i) {
        if (sinfo[i].analytic >= 0) {  // not in the wide BVH: tested per ray (traverse.h)
            prim_surface[sinfo[i].prim_base] = i;
            continue;
        }
        if (sinfo[i].instance >= 0) {  // one top-level entry: the instance's world box
            const MeshInfo& mi = minfo[sinfo[i].mesh];
            for (uint32_t t = 0; t < mi.num_tris; ++t) prim_surface[sinfo[i].prim_base + t] = i;
            if (mi.num_tris == 0) continue;
            recs.emplace_back();
            TriRec& r = recs.back();
            const Box3& b = sb[i];
            for (int a = 0; a < 3; ++a) {
                // room for the rounding of the ray's map into object space
                const float lo = comp(b.lo, a), hi = comp(b.hi, a);
                const float pad = (std::fabs(lo) + std::fabs(hi) + (hi - lo)) * (1.0f / 1048576.0f) + 1e-30f;
                r.p0[a] = lo - pad;
                r.p1[a] = hi + pad;
            }
            const uint32_t k = (uint32_t)sinfo[i].instance, marker = kInstanceMarker;
            std::memcpy(&r.p2[0], &marker, 4);
            std::memcpy(&r.p2[1], &k, 4);
            r.p2[2] = 0.f;
            r.prim = sinfo[i].prim_base;
            r.reserved = 0;
            r.flags = i << 1;
            continue;
        }
        const MeshInfo& mi = minfo[sinfo[i].mesh];
        for (uint32_t t = 0; t < mi.num_tris; ++t) {
            recs.emplace_back();
            TriRec& r = recs.back();
            const uint32_t* ix = &midx[3 * (size_t)(mi.tri_base + t)];
            for (int a = 0; a < 3; ++a) {
                r.p0[a] = vpos[3 * (size_t)(mi.vert_base + ix[0]) + a];
                r.p1[a] = vpos[3 * (size_t)(mi.vert_base + ix[1]) + a];
                r.p2[a] = vpos[3 * (size_t)(mi.vert_base + ix[2]) + a];
            }
            r.prim = sinfo[i].prim_base + t;
            r.reserved = 0;
            r.flags = (mi.root >= 0 ? 1u : 0u) | (i << 1);
            uint32_t xb;
            std::memcpy(&xb, &r.p2[0], 4);
            if (xb == kInstanceMarker) {  // a vertex may not spell the instance marker: it is a NaN anyway
                xb = 0x7FC00000u;
                std::memcpy(&r.p2[0], &xb, 4);
            }
            prim_surface[sinfo[i].prim_base + t] = i;
            prim_anc[sinfo[i].prim_base + t] = leaf_parent[mi.tri_base + t];
        }
    }
    std::vector<QNode>& qn = H.qnodes;
    std::vector<TriRec>& tris = H.tris;
    uint32_t depth = 0;
    std::string err;
    QbvhCheck qcheck;
    if (build_qbvh(recs, qn, tris, depth, err, &qcheck) != 0) return fail(VR_ERR_INVALID, err);
    // one object-space wide BVH per instanced mesh, appended to the node and
    // record arrays (its child links shifted accordingly)
    if (!instances.empty()) {
        std::vector<uint32_t> blas_of(M, kInvalidChild);
        for (InstanceDev& x : instances) {
            const MeshInfo& mi = minfo[x.mesh];
            if (mi.num_tris == 0) continue;
            if (blas_of[x.mesh] == kInvalidChild) {
                std::vector<TriRec> brec(mi.num_tris);
                for (uint32_t t = 0; t < mi.num_tris; ++t) {
                    TriRec& r = brec[t];
                    const uint32_t* ix = &midx[3 * (size_t)(mi.tri_base + t)];
                    for (int a = 0; a < 3; ++a) {
                        r.p0[a] = vpos[3 * (size_t)(mi.vert_base + ix[0]) + a];
                        r.p1[a] = vpos[3 * (size_t)(mi.vert_base + ix[1]) + a];
                        r.p2[a] = vpos[3 * (size_t)(mi.vert_base + ix[2]) + a];
                    }
                    uint32_t xb;
                    std::memcpy(&xb, &r.p2[0], 4);
                    if (xb == kInstanceMarker) {
                        xb = 0x7FC00000u;
                        std::memcpy(&r.p2[0], &xb, 4);
                    }
                    r.prim = t;  // within the mesh: the instance adds its prim base
                    r.reserved = 0;
                    r.flags = 0;
                }
                std::vector<QNode> bq;
                std::vector<TriRec> bt;
                uint32_t bdepth = 0;
                if (build_qbvh(brec, bq, bt, bdepth, err, &qcheck) != 0) return fail(VR_ERR_INVALID, err);
                const uint32_t qoff = (uint32_t)qn.size(), toff = (uint32_t)tris.size();
                if ((uint64_t)toff + bt.size() >= (1ull << 27))
                    return fail(VR_ERR_INVALID, "too many triangle references for the 27-bit leaf encoding");
                for (QNode& q : bq)
                    for (int c = 0; c < 4; ++c) {
                        if (q.child[c] == kInvalidChild) continue;
                        q.child[c] = (q.child[c] & kLeafBit) ? q.child[c] + toff : q.child[c] + qoff;
                    }
                qn.insert(qn.end(), bq.begin(), bq.end());
                tris.insert(tris.end(), bt.begin(), bt.end());
                blas_of[x.mesh] = qoff;
                depth = std::max(depth, bdepth + 1);
            }
            x.blas_root = blas_of[x.mesh];
        }
    }
    if (qn.size() >= (1u << 29)) return fail(VR_ERR_INVALID, "too many wide BVH nodes (2^29)");
    const size_t num_leaf_entries = tris.size();
    // Spatial splits copy a triangle into every leaf that holds a piece of it.
    // On request leaves hold 4-byte references into one record per triangle
    // (first-use order, so neighbouring leaves find their triangles close
    // together): less memory, one more dependent load per leaf entry.
    const bool refs = (d->build_flags & VR_BUILD_LEAF_REFS) && !(d->build_flags & VR_BUILD_LEAF_RECORDS) &&
                      instances.empty();
    H.tri_ref.clear();
    if (refs) {
        std::vector<uint32_t> uniq_of_prim(T, 0xFFFFFFFFu);
        std::vector<TriRec> uniq;
        uniq.reserve(recs.size());
        H.tri_ref.resize(tris.size());
        for (size_t i = 0; i < tris.size(); ++i) {
            uint32_t& u = uniq_of_prim[tris[i].prim];
            if (u == 0xFFFFFFFFu) {
                u = (uint32_t)uniq.size();
                uniq.push_back(tris[i]);
            }
            H.tri_ref[i] = u;
        }
        tris.swap(uniq);
    }
    std::vector<TriRec>().swap(recs);
    auto t_sah1 = clk::now();

    // ---- materials / textures / environment / camera ----------------------
    std::vector<GpuMaterial>& mats = H.materials;
    mats.assign(d->num_materials, GpuMaterial{});
    for (uint32_t i = 0; i < d->num_materials; ++i) {
        const vr_material& m = d->materials[i];
        mats[i] = GpuMaterial{m.kind, {m.color[0], m.color[1], m.color[2]}, m.param, m.albedo_texture,
                              m.normal_texture, m.transparent ? 1 : 0, m.index, m.roughness, m.metallic, 0};
    }
    std::vector<GpuTexture>& texs = H.textures;
    texs.assign(d->num_textures, GpuTexture{});
    std::vector<uint8_t>& texdata = H.tex_data;
    for (uint32_t i = 0; i < d->num_textures; ++i) {
        const vr_texture& t = d->textures[i];
        size_t texel = t.format == VR_TEX_RGBA8 ? 4 : 12;
        size_t bytes = texel * t.width * (size_t)t.height;
        size_t off = (texdata.size() + 255) & ~(size_t)255;
        texdata.resize(off + bytes);
        std::memcpy(texdata.data() + off, t.data, bytes);
        texs[i] = GpuTexture{t.width, t.height, t.format, t.sample_type, off};
    }
    v.env_kind = d->environment.kind;
    for (int a = 0; a < 3; ++a) v.env_color[a] = d->environment.color[a];
    std::vector<float>&env = H.env, &env_marg = H.env_marg, &env_cond = H.env_cond, &env_pdf = H.env_pdf;
    if (v.env_kind == VR_ENV_HDRI) {
        const vr_environment& e = d->environment;
        if (!e.data || e.width == 0 || e.height == 0) return fail(VR_ERR_INVALID, "empty HDRI environment");
        env.assign(e.data, e.data + 3 * (size_t)e.width * e.height);
        v.env_w = e.width;
        v.env_h = e.height;
        if (!(d->build_flags & VR_BUILD_NO_ENV_CDF)) {
            double total = build_env_tables(env.data(), e.width, e.height, env_marg, env_cond, env_pdf);
            v.env_total = (float)total;
            v.env_gm = guide_size(e.height);
            v.env_gc = guide_size(e.width);
            H.env_marg_guide.resize(v.env_gm + 1);
            build_guide(env_marg.data(), e.height, v.env_gm, H.env_marg_guide.data());
            H.env_cond_guide.resize((size_t)e.height * (v.env_gc + 1));
            for (uint32_t r = 0; r < e.height; ++r)
                build_guide(env_cond.data() + (size_t)r * (e.width + 1), e.width, v.env_gc,
                            H.env_cond_guide.data() + (size_t)r * (v.env_gc + 1));
        }
    } else if (v.env_kind != VR_ENV_NONE && v.env_kind != VR_ENV_UNIFORM) {
        return fail(VR_ERR_INVALID, "unknown environment kind");
    }
    const vr_camera& c = d->camera;
    vf3 dir = mk3(c.direction[0], c.direction[1], c.direction[2]);
    vf3 up = mk3(c.up[0], c.up[1], c.up[2]);
    vf3 eye = mk3(c.eye[0], c.eye[1], c.eye[2]);
    vf3 right = normalize3(cross3(dir, up));
    for (int a = 0; a < 3; ++a) {
        v.cam_o[a] = c.eye[a];
        v.cam_dir[a] = c.direction[a];
        v.cam_up[a] = c.up[a];
        v.cam_right[a] = comp(right, a);
    }
    v.cam_d = 1.0f / std::tan(c.fov / 2.0f);  // (fov / 2).tan().recip()  (camera.rs:39)
    v.dof = c.has_dof;
    v.aperture = c.aperture;
    v.focal_length = dot3(mk3(c.focal_point[0], c.focal_point[1], c.focal_point[2]) - eye, dir);

    v.num_qnodes = (uint32_t)qn.size();
    v.num_tris = T;
    v.num_surfaces = S;
    v.top_root_valid = top.empty() ? 0 : 1;
    // host view (the traversal code can run over it on the CPU)
    v.qnodes = qn.data();
    v.tris = tris.data();
    v.tri_ref = H.tri_ref.empty() ? nullptr : H.tri_ref.data();
    v.ref_nodes = ref_nodes.data();
    v.top_nodes = top.data();
    v.top_path = top_path.data();
    v.meshes = minfo.data();
    v.surfaces = sinfo.data();
    v.mesh_idx = midx.data();
    v.face_n = face_n.data();
    v.vpos = vpos.data();
    v.vnrm = vnrm.data();
    v.vuv = vuv.data();
    v.prim_surface = prim_surface.data();
    v.ref_parent = ref_parent.data();
    v.prim_anc = prim_anc.data();
    v.analytic = analytic.data();
    v.num_analytic = (uint32_t)analytic.size();
    v.instances = instances.data();
    v.num_instances = (uint32_t)instances.size();
    v.materials = mats.data();
    v.textures = texs.data();
    v.tex_data = texdata.data();
    v.env = env.data();
    v.env_marg_cdf = env_marg.data();
    v.env_cond_cdf = env_cond.data();
    v.env_marg_guide = H.env_marg_guide.empty() ? nullptr : H.env_marg_guide.data();
    v.env_cond_guide = H.env_cond_guide.empty() ? nullptr : H.env_cond_guide.data();
    v.env_pdf = env_pdf.data();

    vr_scene_stats& st = H.stats;
    st.num_triangles = T - (uint32_t)analytic.size();  // mesh triangles
    st.num_surfaces = S;
    st.num_qnodes = (uint32_t)qn.size();
    st.num_ref_nodes = (uint32_t)ref_nodes.size();
    st.qbvh_depth = depth;
    st.max_leaf_size = kMaxLeaf;
    st.num_tri_records = (uint32_t)num_leaf_entries;
    st.qbox_checked = qcheck.boxes_checked;
    st.qbox_not_containing = qcheck.boxes_not_containing;
    st.leaf_refs = H.tri_ref.empty() ? 0u : 1u;
    st.build_ref_seconds = std::chrono::duration<double>(t_ref1 - t_ref0).count();
    st.build_sah_seconds = std::chrono::duration<double>(t_sah1 - t_ref1).count();
    return 0;
}

int build_scene(vr_ctx* ctx, const vr_scene_desc* d, vr_scene* s) {
    s->ctx = ctx;
    HostScene H;
    int rc = build_host_scene(d, H);
    if (rc) return rc;
    auto t0 = std::chrono::steady_clock::now();
    SceneView& v = s->view;
    v = H.view;
    s->stats = H.stats;
    s->stats.device_bytes = 0;
    rc = 0;
    rc |= upload(s, H.qnodes, &v.qnodes);
    rc |= upload(s, H.tris, &v.tris);
    rc |= upload(s, H.tri_ref, &v.tri_ref);
    rc |= upload(s, H.ref_nodes, &v.ref_nodes);
    rc |= upload(s, H.top_nodes, &v.top_nodes);
    rc |= upload(s, H.top_path, &v.top_path);
    rc |= upload(s, H.meshes, &v.meshes);
    rc |= upload(s, H.surfaces, &v.surfaces);
    rc |= upload(s, H.mesh_idx, &v.mesh_idx);
    rc |= upload(s, H.face_n, &v.face_n);
    rc |= upload(s, H.vpos, &v.vpos);
    rc |= upload(s, H.vnrm, &v.vnrm);
    rc |= upload(s, H.vuv, &v.vuv);
    rc |= upload(s, H.prim_surface, &v.prim_surface);
    rc |= upload(s, H.ref_parent, &v.ref_parent);
    rc |= upload(s, H.prim_anc, &v.prim_anc);
    rc |= upload(s, H.analytic, &v.analytic);
    rc |= upload(s, H.instances, &v.instances);
    rc |= upload(s, H.materials, &v.materials);
    rc |= upload(s, H.textures, &v.textures);
    rc |= upload(s, H.tex_data, &v.tex_data);
    rc |= upload(s, H.env, &v.env);
    rc |= upload(s, H.env_marg, &v.env_marg_cdf);
    rc |= upload(s, H.env_cond, &v.env_cond_cdf);
    rc |= upload(s, H.env_marg_guide, &v.env_marg_guide);
    rc |= upload(s, H.env_cond_guide, &v.env_cond_guide);
    rc |= upload(s, H.env_pdf, &v.env_pdf);
    if (rc) return VR_ERR_OOM;
    s->num_prims = v.num_tris;
    s->has_microfacet = false;
    for (const GpuMaterial& m : H.materials) s->has_microfacet |= m.kind == VR_MAT_MICROFACET;
    s->stats.upload_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

void free_scene(vr_scene* s) {
    for (void* p : s->allocations) (void)hipFree(p);
    s->allocations.clear();
    if (s->fallback_list) (void)hipFree(s->fallback_list);
    if (s->counters) (void)hipFree(s->counters);
    if (s->spill) (void)hipFree(s->spill);
    for (auto& e : s->trace_ev)
        if (e) (void)hipEventDestroy(e);
    s->trace_ev[0] = s->trace_ev[1] = nullptr;
    s->spill = nullptr;
    s->fallback_list = nullptr;
    s->counters = nullptr;
}

}  // namespace vr
