// Device traversal routines shared by the ray-query kernels and the
// wavefront renderer's extension kernel.
//
// Two engines answer SceneAcceleration::hit (core/scene.rs:182-185):
//
//  * wide  — SAH 4-wide BVH with 8-bit outward-rounded child boxes, one ray
//            per lane, a short LDS stack, nearest-child-first order and t_max
//            shrinking.  Box tests are conservative (padded boxes, a relative
//            slack on the slab exit, a relative slack on the cull distance), so
//            every triangle the reference could return is reached.  Equal-t
//            hits are resolved by the reference's in-order rank.  The winner is
//            then confirmed: its hit point must lie strictly inside the
//            eps-expanded triangle box by a margin that dominates the rounding
//            of the reference's slab test (so every reference ancestor box
//            passes), and its top-level ancestors must pass the exact reference
//            slab test.  Unconfirmed winners and stack overflows go to the
//            exact engine.
//  * exact — the flattened reference trees walked exactly as BvhNode::hit
//            (core/bvh.rs:132-181): both children, unshrunk t_max, no box test
//            at leaves, brute force for meshes of <= 4 triangles
//            (core/mesh.rs:125-138), with the same slab arithmetic
//            (util/aabb.rs:86-148).  Equal t is resolved by rank, which equals
//            merge_optionals' right-wins rule and the brute-force first-wins rule.
#pragma once
#include "internal.h"

namespace vr {

struct HitOut {
    float t, u, v;
    uint32_t prim;  // 0xFFFFFFFF: miss
};

constexpr float kSlabSlack = 1.0f + 1.0f / 131072.0f;  // 2^-17 relative on the slab exit
constexpr float kCullSlack = 1.0f + 1.0f / 65536.0f;   // 2^-16 relative on t_best
constexpr float kRobustRel = 1.0f / 1048576.0f;        // 2^-20 interior margin


// Geometry of a triangle record: two 16-byte loads + one 4-byte load, all
// float.  The integer words (prim = 9, rank = 10, flags = 11) are read with
// their own scalar loads, only when a hit needs them.
VR_HD void load_tri_geom(const TriRec* tris, uint32_t slot, vf3& p0, vf3& p1, vf3& p2) {
    const f32x4* q = (const f32x4*)(tris + slot);
    f32x4 a = q[0], b = q[1];
    float c = ((const float*)(tris + slot))[8];
    p0 = mk3(a.x, a.y, a.z);
    p1 = mk3(a.w, b.x, b.y);
    p2 = mk3(b.z, b.w, c);
}
VR_HD uint32_t tri_word(const TriRec* tris, uint32_t slot, int w) { return ((const uint32_t*)(tris + slot))[w]; }

// Confirms that the reference would consider this hit (see header comment).
VR_HD bool confirm_winner(const SceneView& sv, vf3 o, vf3 d, float tmin, float t, vf3 p0, vf3 p1,
                                               vf3 p2, uint32_t flags) {
    uint32_t surf = flags >> 1;
    if (sv.num_surfaces > 1) {
        SurfaceInfo si = sv.surfaces[surf];
        for (uint32_t k = 0; k < si.path_len; ++k) {
            const RefNode& nd = sv.top_nodes[sv.top_path[si.path_off + k]];
            Box3 b{mk3(nd.lo[0], nd.lo[1], nd.lo[2]), mk3(nd.hi[0], nd.hi[1], nd.hi[2])};
            if (!ref_slab_hit(b, o, d, tmin, kInf)) return false;
        }
    }
    if (flags & 1u) {
        vf3 P = o + d * t;  // Ray::at (ray.rs:19-21)
        Box3 bb = tri_box_eps(p0, p1, p2);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float pa = comp(P, a);
            float delta = kRobustRel * (t + fabsf(pa) + fabsf(comp(o, a)));
            if (!(pa - comp(bb.lo, a) > delta && comp(bb.hi, a) - pa > delta)) return false;
        }
    }
    return true;
}

// Result codes of the wide engine.
enum : int { kWideDone = 0, kWideFallback = 1, kWideOverflow = 2 };

// Wide-engine query for one ray.  `stk` is this lane's LDS stack, entries
// strided by `stride`.  ANY = true: stop at the first confirmed hit with
// t < tmax_any.  `out` is fully written on kWideDone.
template <int STACK, bool ANY>
VR_HD int trace_wide(const SceneView& sv, vf3 o, vf3 d, float tmin, float tmax_any, u32x2* stk, int stride,
                     HitOut& out) {
    out.t = 0.f;
    out.u = 0.f;
    out.v = 0.f;
    out.prim = 0xFFFFFFFFu;
    const float kTiny = 1e-20f;
    float idx = 1.0f / (fabsf(d.x) > kTiny ? d.x : copysignf(kTiny, d.x));
    float idy = 1.0f / (fabsf(d.y) > kTiny ? d.y : copysignf(kTiny, d.y));
    float idz = 1.0f / (fabsf(d.z) > kTiny ? d.z : copysignf(kTiny, d.z));
    float tbest = kInf;
    uint32_t best_rank = 0, best_slot = 0xFFFFFFFFu;
    float bu = 0.f, bv = 0.f;
    bool unconfirmed = false, found = false;
    int sp = 0;
    uint32_t cur = 0;
    bool have = sv.num_qnodes > 0;
    for (;;) {
        if (!have) {
            if (sp == 0) break;
            --sp;
            u32x2 e = stk[sp * stride];
            if (!ANY && u2f(e.y) > tbest * kCullSlack) continue;
            cur = e.x;
        }
        have = false;
        if (cur & kLeafBit) {
            uint32_t first = cur & 0x07FFFFFFu;
            uint32_t cnt = ((cur >> 27) & 7u) + 1u;
            for (uint32_t k = 0; k < cnt; ++k) {
                vf3 p0, p1, p2;
                load_tri_geom(sv.tris, first + k, p0, p1, p2);
                MTResult r;
                if (!mt_hit(o, d, p0, p1, p2, tmin, r)) continue;
                if (ANY) {
                    if (r.t < tmax_any) {
                        if (confirm_winner(sv, o, d, tmin, r.t, p0, p1, p2, tri_word(sv.tris, first + k, 11))) {
                            found = true;
                            tbest = r.t;
                            best_slot = first + k;
                            bu = r.u;
                            bv = r.v;
                            break;
                        }
                        unconfirmed = true;
                    }
                    continue;
                }
                if (r.t > tbest) continue;
                uint32_t rank = tri_word(sv.tris, first + k, 10);
                if (r.t < tbest || rank > best_rank) {
                    tbest = r.t;
                    best_rank = rank;
                    best_slot = first + k;
                    bu = r.u;
                    bv = r.v;
                }
            }
            if (ANY && found) break;
            continue;
        }
        const u32x4* qp = (const u32x4*)(sv.qnodes + cur);
        u32x4 w0 = qp[0], w1 = qp[1], w2 = qp[2], w3 = qp[3];
        uint32_t e = w0.w;
        float sx = u2f((e & 0xFFu) << 23) * idx;
        float sy = u2f(((e >> 8) & 0xFFu) << 23) * idy;
        float sz = u2f(((e >> 16) & 0xFFu) << 23) * idz;
        float bx = (u2f(w0.x) - o.x) * idx;
        float by = (u2f(w0.y) - o.y) * idy;
        float bz = (u2f(w0.z) - o.z) * idz;
        uint32_t ch[4] = {w1.x, w1.y, w1.z, w1.w};
        uint32_t qlx = w2.x, qly = w2.y, qlz = w2.z, qhx = w2.w, qhy = w3.x, qhz = w3.y;
        float tcull = ANY ? tmax_any : tbest * kCullSlack;
        float tn[4];
        uint32_t cid[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float lx = (float)((qlx >> (8 * c)) & 0xFFu), hx = (float)((qhx >> (8 * c)) & 0xFFu);
            float ly = (float)((qly >> (8 * c)) & 0xFFu), hy = (float)((qhy >> (8 * c)) & 0xFFu);
            float lz = (float)((qlz >> (8 * c)) & 0xFFu), hz = (float)((qhz >> (8 * c)) & 0xFFu);
            float t0x = fmaf(lx, sx, bx), t1x = fmaf(hx, sx, bx);
            float t0y = fmaf(ly, sy, by), t1y = fmaf(hy, sy, by);
            float t0z = fmaf(lz, sz, bz), t1z = fmaf(hz, sz, bz);
            float nx = fminf(t0x, t1x), fx = fmaxf(t0x, t1x);
            float ny = fminf(t0y, t1y), fy = fmaxf(t0y, t1y);
            float nz = fminf(t0z, t1z), fz = fmaxf(t0z, t1z);
            float tnear = fmaxf(fmaxf(nx, ny), fmaxf(nz, tmin));
            float tfar = fminf(fminf(fx, fy), fminf(fz, tcull));
            bool hit = ch[c] != kInvalidChild && tnear <= tfar * kSlabSlack;
            tn[c] = hit ? tnear : kInf;
            cid[c] = ch[c];
        }
        // sort (tnear, child) ascending; misses carry +inf and sink to the end
#define VR_CSWAP(a, b)                                   \
    if (tn[b] < tn[a]) {                                 \
        float tt = tn[a];                                \
        tn[a] = tn[b];                                   \
        tn[b] = tt;                                      \
        uint32_t cc = cid[a];                            \
        cid[a] = cid[b];                                 \
        cid[b] = cc;                                     \
    }
        VR_CSWAP(0, 1)
        VR_CSWAP(2, 3)
        VR_CSWAP(0, 2)
        VR_CSWAP(1, 3)
        VR_CSWAP(1, 2)
#undef VR_CSWAP
        if (tn[0] == kInf) continue;
        cur = cid[0];
        have = true;
        bool overflow = false;
#pragma unroll
        for (int c = 3; c >= 1; --c) {
            if (tn[c] == kInf) continue;
            if (sp >= STACK) {
                overflow = true;
                continue;
            }
            stk[sp * stride] = u32x2{cid[c], f2u(tn[c])};
            ++sp;
        }
        if (overflow) return kWideOverflow;  // the exact engine answers this ray
    }
    if (ANY) {
        if (found) {
            out.t = tbest;
            out.u = bu;
            out.v = bv;
            out.prim = tri_word(sv.tris, best_slot, 9);
            return kWideDone;
        }
        return unconfirmed ? kWideFallback : kWideDone;
    }
    if (best_slot == 0xFFFFFFFFu) return kWideDone;
    vf3 p0, p1, p2;
    load_tri_geom(sv.tris, best_slot, p0, p1, p2);
    if (!confirm_winner(sv, o, d, tmin, tbest, p0, p1, p2, tri_word(sv.tris, best_slot, 11))) return kWideFallback;
    out.t = tbest;
    out.u = bu;
    out.v = bv;
    out.prim = tri_word(sv.tris, best_slot, 9);
    return kWideDone;
}

// Exact reference engine for one ray (closest hit).
__host__ __device__ inline void trace_exact(const SceneView& sv, vf3 o, vf3 d, float tmin, HitOut& out) {
    out.prim = 0xFFFFFFFFu;
    out.t = out.u = out.v = 0.f;
    if (!sv.top_root_valid) return;
    bool have = false;
    float bt = 0.f;
    uint32_t brank = 0;
    int32_t tstack[64];
    int tsp = 0;
    tstack[tsp++] = 0;
    while (tsp) {
        int32_t ni = tstack[--tsp];
        RefNode tn = sv.top_nodes[ni];
        if (tn.obj < 0) {
            Box3 b{mk3(tn.lo[0], tn.lo[1], tn.lo[2]), mk3(tn.hi[0], tn.hi[1], tn.hi[2])};
            if (ref_slab_hit(b, o, d, tmin, kInf)) {
                tstack[tsp++] = tn.right;
                tstack[tsp++] = ni + 1;
            }
            continue;
        }
        uint32_t surf = (uint32_t)tn.obj;
        SurfaceInfo si = sv.surfaces[surf];
        MeshInfo mi = sv.meshes[si.mesh];
        auto test = [&](uint32_t tri) {
            const uint32_t* ix = sv.mesh_idx + 3 * (size_t)(mi.tri_base + tri);
            const float* vp = sv.vpos + 3 * (size_t)mi.vert_base;
            vf3 p0 = mk3(vp[3 * ix[0]], vp[3 * ix[0] + 1], vp[3 * ix[0] + 2]);
            vf3 p1 = mk3(vp[3 * ix[1]], vp[3 * ix[1] + 1], vp[3 * ix[1] + 2]);
            vf3 p2 = mk3(vp[3 * ix[2]], vp[3 * ix[2] + 1], vp[3 * ix[2] + 2]);
            MTResult r;
            if (!mt_hit(o, d, p0, p1, p2, tmin, r)) return;
            uint32_t rank = si.rank_base + sv.mesh_rank[mi.tri_base + tri];
            if (!have || r.t < bt || (r.t == bt && rank > brank)) {
                have = true;
                bt = r.t;
                brank = rank;
                out.t = r.t;
                out.u = r.u;
                out.v = r.v;
                out.prim = si.prim_base + tri;
            }
        };
        if (mi.root < 0) {
            for (uint32_t k = 0; k < mi.num_tris; ++k) test(k);
            continue;
        }
        int32_t mstack[64];
        int msp = 0;
        mstack[msp++] = mi.root;
        while (msp) {
            int32_t mj = mstack[--msp];
            RefNode nd = sv.ref_nodes[mj];
            if (nd.obj >= 0) {
                test((uint32_t)nd.obj);
                continue;
            }
            Box3 b{mk3(nd.lo[0], nd.lo[1], nd.lo[2]), mk3(nd.hi[0], nd.hi[1], nd.hi[2])};
            if (!ref_slab_hit(b, o, d, tmin, kInf)) continue;
            mstack[msp++] = nd.right;
            mstack[msp++] = mj + 1;
        }
    }
}

}  // namespace vr
