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
//            hits go to the smaller prim id (see below).  The winner is
//            then confirmed to be a triangle the reference reaches: its
//            top-level ancestors must pass the exact reference slab test, and
//            so must its mesh-tree ancestors — taken as passed when the hit
//            point lies strictly inside the eps-expanded triangle box by a
//            margin that dominates the slab test's rounding, walked exactly
//            otherwise.  Winners the reference does not reach and stack
//            overflows go to the exact engine.
//  * exact — the flattened reference trees walked exactly as BvhNode::hit
//            (core/bvh.rs:132-181): both children, unshrunk t_max, no box test
//            at leaves, brute force for meshes of <= 4 triangles
//            (core/mesh.rs:125-138), with the same slab arithmetic
//            (util/aabb.rs:86-148).
//
// Ties.  merge_optionals (bvh.rs:164-181) compares Option<(HitRecord, usize)>
// tuples, and HitRecord orders by t alone (ray.rs:52-62), so on equal t the
// smaller handle wins: the smaller surface index at the top level
// (scene.rs:182-185), then the smaller triangle index inside the mesh
// (mesh.rs:269-279; the brute-force loop of mesh.rs:125-138 keeps the earlier
// triangle, the same rule).  Prim ids are numbered surface by surface in
// triangle order, so both engines keep the lexicographic minimum of (t, prim).
//
// Instances (template flag INST, scenes with vr_instance surfaces; no
// counterpart in the reference): an instance is one record of the top-level
// wide BVH holding its world box (kInstanceMarker in p2.x).  A leaf that meets
// one pushes an "enter" entry at the box's entry distance; popping it pushes an
// "exit" entry, maps the ray into the instance's object space (t is unchanged:
// the direction is mapped, not renormalised) and walks the shared object-space
// BVH of its mesh; popping the exit entry restores the world ray.  Hits inside
// an instance carry the instance's prim base and need no confirmation.
//
// Watertight mode (template flag WT, engine VR_ENGINE_WATERTIGHT): the same
// traversal with the watertight ray/triangle test of vrmath.h in place of
// Moller-Trumbore and no reference semantics: every candidate the
// conservative boxes reach counts (no ancestor confirmation), and the exact
// fallback walks the median-split trees with slab tests that never reject a
// ray crossing the box.  For the extended estimator only.
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
// float.  The integer words (prim = 9, flags = 11) are read with their own
// loads, only when a hit needs them.
VR_HD void load_tri_geom(const TriRec* tris, uint32_t slot, vf3& p0, vf3& p1, vf3& p2) {
    const f32x4* q = (const f32x4*)(tris + slot);
    f32x4 a = q[0], b = q[1];
    float c = ((const float*)(tris + slot))[8];
    p0 = mk3(a.x, a.y, a.z);
    p1 = mk3(a.w, b.x, b.y);
    p2 = mk3(b.z, b.w, c);
}
VR_HD uint32_t tri_word(const TriRec* tris, uint32_t slot, int w) { return ((const uint32_t*)(tris + slot))[w]; }

// The ray (o, d) in an instance's object space: o' = M^-1 o + c', d' = M^-1 d
// (not renormalised, so a hit's t is the world t), f32, the same operation
// order in every engine.
VR_HD void instance_ray(const InstanceDev& x, vf3 o, vf3 d, vf3& oo, vf3& dd) {
    const float* m = x.inv;
    oo = mk3(((m[0] * o.x + m[1] * o.y) + m[2] * o.z) + m[3], ((m[4] * o.x + m[5] * o.y) + m[6] * o.z) + m[7],
             ((m[8] * o.x + m[9] * o.y) + m[10] * o.z) + m[11]);
    dd = mk3((m[0] * d.x + m[1] * d.y) + m[2] * d.z, (m[4] * d.x + m[5] * d.y) + m[6] * d.z,
             (m[8] * d.x + m[9] * d.y) + m[10] * d.z);
}

// Exact membership of a triangle in the reference's candidate set within its
// mesh: BvhNode::hit (core/bvh.rs:132-181) descends into a child only after
// the parent's box passes the slab test, and tests leaves without a box, so
// the triangle is reached iff every inner ancestor of its leaf passes
// AABB::hit with the unshrunk t_max.  Walks leaf parent -> mesh root with the
// reference's own slab arithmetic, so the answer is the reference's.
VR_HD bool mesh_ancestors_pass(const SceneView& sv, vf3 o, vf3 d, float tmin, uint32_t prim) {
    int32_t nd = sv.prim_anc[prim];
    while (nd >= 0) {
        const RefNode& n = sv.ref_nodes[nd];
        int32_t up = sv.ref_parent[nd];
        Box3 b{mk3(n.lo[0], n.lo[1], n.lo[2]), mk3(n.hi[0], n.hi[1], n.hi[2])};
        if (!ref_slab_hit(b, o, d, tmin, kInf)) return false;
        nd = up;
    }
    return true;
}

// Confirms that the reference would consider this hit (see header comment).
VR_HD bool confirm_winner(const SceneView& sv, vf3 o, vf3 d, float tmin, float t, vf3 p0, vf3 p1,
                                               vf3 p2, uint32_t flags, uint32_t prim) {
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
        // fast path: a hit point inside the eps-expanded triangle box by a
        // margin that dominates the slab test's rounding passes every
        // ancestor box; otherwise walk the ancestors exactly
        vf3 P = o + d * t;  // Ray::at (ray.rs:19-21)
        Box3 bb = tri_box_eps(p0, p1, p2);
        bool inside = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float pa = comp(P, a);
            float delta = kRobustRel * (t + fabsf(pa) + fabsf(comp(o, a)));
            inside = inside && pa - comp(bb.lo, a) > delta && comp(bb.hi, a) - pa > delta;
        }
        if (!inside) return mesh_ancestors_pass(sv, o, d, tmin, prim);
    }
    return true;
}


// ---- analytic surfaces -------------------------------------------------------
// AnalyticSurface::hit with t_max = inf, as SceneAcceleration::hit passes it:
// Sphere::hit (voidray_common/src/surfaces.rs:46-78) and GroundPlane::hit
// (surfaces.rs:88-104), same f32 operations.  Out: t and the hit record's uv.
// A ray lying exactly in a ground plane (t = 0/0) counts as a miss.
VR_HD bool analytic_hit(const AnalyticDev& a, vf3 o, vf3 d, float tmin, float& t, float& u, float& v) {
    if (a.kind == VR_ANALYTIC_SPHERE) {
        const vf3 oc = o - mk3(a.c[0], a.c[1], a.c[2]);
        const float aa = mag2(d);
        const float half_b = dot3(oc, d);
        const float c = mag2(oc) - a.r * a.r;
        const float disc = half_b * half_b - aa * c;
        if (disc < 0.0f) return false;
        const float sqrtd = sqrtf(disc);
        float root = (-half_b - sqrtd) / aa;
        if (root < tmin || kInf < root) {
            root = (-half_b + sqrtd) / aa;
            if (root < tmin || kInf < root) return false;
        }
        t = root;
        u = v = 0.0f;
        return true;
    }
    const float tt = (a.h - o.y) / d.y;
    if (tt <= tmin || tt >= kInf || tt != tt) return false;
    const vf3 p = o + d * tt;  // Ray::at
    t = tt;
    u = p.x;
    v = p.z;
    return true;
}

// Every top-level ancestor of `surf` passes the reference slab test.
VR_HD bool top_path_pass(const SceneView& sv, uint32_t surf, vf3 o, vf3 d, float tmin) {
    const SurfaceInfo& si = sv.surfaces[surf];
    for (uint32_t k = 0; k < si.path_len; ++k) {
        const RefNode& nd = sv.top_nodes[sv.top_path[si.path_off + k]];
        Box3 b{mk3(nd.lo[0], nd.lo[1], nd.lo[2]), mk3(nd.hi[0], nd.hi[1], nd.hi[2])};
        if (!ref_slab_hit(b, o, d, tmin, kInf)) return false;
    }
    return true;
}

// Closest analytic surface the reference reaches (lexicographic min of (t,
// prim), like the triangles); with `first_below` set, the first one found
// with t < first_below.  Returns its index or -1.  WT: every analytic
// surface counts (no top-level reach test).
template <bool WT = false>
VR_HD int analytic_best(const SceneView& sv, vf3 o, vf3 d, float tmin, float first_below, float& bt, uint32_t& bprim,
                        float& bu, float& bv) {
    int best = -1;
    for (uint32_t k = 0; k < sv.num_analytic; ++k) {
        const AnalyticDev& a = sv.analytic[k];
        float t, u, v;
        if (!analytic_hit(a, o, d, tmin, t, u, v)) continue;
        if (best >= 0 && !(t < bt || (t == bt && a.prim < bprim))) continue;
        if (!WT && !top_path_pass(sv, a.surface, o, d, tmin)) continue;
        best = (int)k;
        bt = t;
        bprim = a.prim;
        bu = u;
        bv = v;
        if (t < first_below) break;
    }
    return best;
}

constexpr uint32_t kAnalyticSlot = 0x40000000u;  // best_slot of an analytic winner (triangle slots < 2^27)

// ---- wide engine as a resumable per-lane state machine ----------------------
// The stack lives in LDS (LDS_STACK entries per lane, strided by `stride`) and
// spills its oldest entries to a per-lane global region (`spill`, strided by
// `spill_stride`, `spill_cap` entries) when it overflows.  Each entry is
// (node or leaf code, bits of the entry distance) so a popped subtree that
// lies beyond the current best hit is skipped without touching its node.
struct WideState {
    vf3 o, d;  // the ray as the current BVH sees it (object space inside an instance)
    WTRay wt;  // watertight mode: the ray's shear (unused otherwise)
    float idx, idy, idz;
    float tbest, bu, bv;
    uint32_t best_prim, best_slot;
    uint32_t cur;
    int sp, lo;
    int top;  // LDS ring slot of entry sp (sp modulo the ring size, kept incrementally)
    // 0 or 1.  As words they live in vector registers; as bools the compiler
    // keeps each as a lane mask in scalar registers, which the persistent
    // kernel runs out of (spilled masks cost readlanes in every node step).
    uint32_t have, unconfirmed, found;
    // INST only: the instance being walked (kNoInstance: none), its prim
    // base and triangle determinant threshold, the instance of the best hit
    // (the world ray is the caller's)
    uint32_t inst, inst_base, best_inst;
    float inst_eps;
};
constexpr uint32_t kNoInstance = 0xFFFFFFFFu;
// Stack codes of instance entries (INST): bit 30 set, below the leaf bit.
// Node indices stay below 2^30 (the scene build checks it).
constexpr uint32_t kInstEntry = 0x40000000u;            // | instance index: enter
constexpr uint32_t kInstExit = kInstEntry | 0x3FFFFFFFu;  // leave the current instance

struct WideStack {
    u32x2* lds;
    int stride;
    u32x2* spill;
    int spill_stride, spill_cap;
};

enum : int { kStepContinue = 0, kStepDone = 1, kStepFallback = 2 };

template <bool WT = false>
VR_HD void wide_set_ray(WideState& st, vf3 o, vf3 d) {
    const float kTiny = 1e-20f;
    st.o = o;
    st.d = d;
    if (WT) st.wt = wt_prepare(d);
    st.idx = 1.0f / (fabsf(d.x) > kTiny ? d.x : copysignf(kTiny, d.x));
    st.idy = 1.0f / (fabsf(d.y) > kTiny ? d.y : copysignf(kTiny, d.y));
    st.idz = 1.0f / (fabsf(d.z) > kTiny ? d.z : copysignf(kTiny, d.z));
}

template <bool WT = false>
VR_HD void wide_init(WideState& st, vf3 o, vf3 d, bool nonempty) {
    wide_set_ray<WT>(st, o, d);
    st.inst = kNoInstance;
    st.inst_base = 0;
    st.best_inst = kNoInstance;
    st.tbest = kInf;
    st.bu = st.bv = 0.f;
    st.best_prim = 0xFFFFFFFFu;
    st.best_slot = 0xFFFFFFFFu;
    st.cur = 0;
    st.sp = 0;
    st.lo = 0;
    st.top = 0;
    st.have = nonempty;
    st.unconfirmed = false;
    st.found = false;
}

// Analytic surfaces are few and not in the wide BVH: each ray tests them all
// up front, and the best one seeds the closest hit (so the triangle walk culls
// against it and ties resolve by prim id as usual).  An any-hit ray that hits
// one below tmax_any is finished.
template <bool ANY, bool WT = false>
VR_HD void wide_seed_analytic(const SceneView& sv, WideState& st, float tmin, float tmax_any) {
    float t = kInf, u = 0.f, v = 0.f;
    uint32_t prim = 0xFFFFFFFFu;
    const int k = analytic_best<WT>(sv, st.o, st.d, tmin, ANY ? tmax_any : -kInf, t, prim, u, v);
    if (k < 0) return;
    if (ANY) {
        if (!(t < tmax_any)) return;
        st.found = true;
        st.have = false;
    }
    st.tbest = t;
    st.best_prim = prim;
    st.best_slot = kAnalyticSlot | (uint32_t)k;
    st.bu = u;
    st.bv = v;
}

// The LDS part of the stack is a ring of LDS_STACK entries: entry i lives in
// slot i mod LDS_STACK, and st.top tracks the slot of entry st.sp without a
// division.  When the ring is full (sp - lo == LDS_STACK) the slot of entry sp
// is that of the oldest LDS entry lo, which moves to the spill region first.
template <int LDS_STACK>
VR_HD bool wide_push(WideState& st, const WideStack& ws, u32x2 e) {
    if (st.sp - st.lo == LDS_STACK) {
        if (st.lo >= ws.spill_cap) return false;
        ws.spill[st.lo * ws.spill_stride] = ws.lds[st.top * ws.stride];
        ++st.lo;
    }
    ws.lds[st.top * ws.stride] = e;
    ++st.sp;
    st.top = st.top + 1 == LDS_STACK ? 0 : st.top + 1;
    return true;
}

template <int LDS_STACK>
VR_HD u32x2 wide_pop(WideState& st, const WideStack& ws) {
    --st.sp;
    st.top = st.top == 0 ? LDS_STACK - 1 : st.top - 1;
    if (st.sp < st.lo) {
        st.lo = st.sp;
        return ws.spill[st.sp * ws.spill_stride];
    }
    return ws.lds[st.top * ws.stride];
}

// Pops the nearest pending entry that survives culling into st.cur; false
// when the stack is empty.
template <int LDS_STACK, bool ANY>
VR_HD bool wide_pop_next(WideState& st, const WideStack& ws) {
    while (st.sp > 0) {
        u32x2 e = wide_pop<LDS_STACK>(st, ws);
        if (!ANY && u2f(e.y) > st.tbest * kCullSlack) continue;
        st.cur = e.x;
        st.have = true;
        return true;
    }
    return false;
}

// Tests the triangles of leaf `code`.  Returns kStepContinue, kStepDone when
// an any-hit query has its confirmed hit, or kStepFallback when an instance
// entry finds the stack exhausted (INST).
template <int LDS_STACK, bool ANY, bool WT = false, bool INST = false>
VR_HD int wide_leaf(const SceneView& sv, WideState& st, const WideStack& ws, uint32_t code, float tmin,
                    float tmax_any) {
    const vf3 o = st.o, d = st.d;
    const bool in_inst = INST && st.inst != kNoInstance;
    uint32_t first = code & 0x07FFFFFFu;
    uint32_t cnt = ((code >> 27) & 7u) + 1u;
    for (uint32_t k = 0; k < cnt; ++k) {
        // the record's index (best_slot keeps it, not the leaf entry)
        const uint32_t slot = sv.tri_ref ? sv.tri_ref[first + k] : first + k;
        vf3 p0, p1, p2;
        load_tri_geom(sv.tris, slot, p0, p1, p2);
        if (INST && f2u(p2.x) == kInstanceMarker) {
            // an instance: its world box is [p0, p1]; enter it later, in
            // near-to-far order with the other pending entries
            const float tcull = ANY ? tmax_any : st.tbest * kCullSlack;
            const float ax = (p0.x - o.x) * st.idx, bx = (p1.x - o.x) * st.idx;
            const float ay = (p0.y - o.y) * st.idy, by = (p1.y - o.y) * st.idy;
            const float az = (p0.z - o.z) * st.idz, bz = (p1.z - o.z) * st.idz;
            const float tnear = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tmin));
            const float tfar = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tcull));
            if (tnear <= tfar * kSlabSlack &&
                !wide_push<LDS_STACK>(st, ws, u32x2{kInstEntry | f2u(p2.y), f2u(tnear)}))
                return kStepFallback;
            continue;
        }
        MTResult r;
        if (WT ? !wt_hit(st.wt, o, p0, p1, p2, tmin, r)
               : !mt_hit(o, d, p0, p1, p2, tmin, r, in_inst ? st.inst_eps : -1.0f))
            continue;
        if (ANY) {
            if (r.t < tmax_any) {
                if (WT || in_inst ||
                    confirm_winner(sv, o, d, tmin, r.t, p0, p1, p2, tri_word(sv.tris, slot, 11),
                                   tri_word(sv.tris, slot, 9))) {
                    st.found = true;
                    st.tbest = r.t;
                    st.best_slot = slot;
                    st.bu = r.u;
                    st.bv = r.v;
                    if (INST) {
                        st.best_inst = st.inst;
                        st.best_prim = st.inst_base + tri_word(sv.tris, slot, 9);
                    }
                    return kStepDone;
                }
                st.unconfirmed = true;
            }
            continue;
        }
        if (r.t > st.tbest) continue;
        const uint32_t prim = (in_inst ? st.inst_base : 0u) + tri_word(sv.tris, slot, 9);
        if (r.t < st.tbest || prim < st.best_prim) {
            st.tbest = r.t;
            st.best_prim = prim;
            st.best_slot = slot;
            st.bu = r.u;
            st.bv = r.v;
            if (INST) st.best_inst = st.inst;
        }
    }
    return kStepContinue;
}

// Pops of instance entries (INST).  Enter (code kInstEntry | k, popped in
// world space): push the exit entry, map the ray into object space
// (instance_ray) and continue at the mesh BVH's root; returns kStepFallback
// when the exit entry does not fit on the stack.  Exit (kInstExit): the caller
// hands back the world ray (the state does not keep it, to save registers).
template <int LDS_STACK, bool WT = false>
VR_HD int wide_instance_enter(const SceneView& sv, WideState& st, const WideStack& ws, uint32_t code, float tmin) {
    const uint32_t k = code & ~kInstEntry;
    if (!wide_push<LDS_STACK>(st, ws, u32x2{kInstExit, 0u})) return kStepFallback;
    const InstanceDev& x = sv.instances[k];
    vf3 oo, dd;
    instance_ray(x, st.o, st.d, oo, dd);
    wide_set_ray<WT>(st, oo, dd);
    st.inst = k;
    st.inst_base = x.prim_base;
    st.inst_eps = tmin * x.det_inv;
    st.cur = x.blas_root;
    st.have = true;
    return kStepContinue;
}

template <bool WT = false>
VR_HD void wide_instance_exit(WideState& st, vf3 world_o, vf3 world_d) {
    wide_set_ray<WT>(st, world_o, world_d);
    st.inst = kNoInstance;
}

// Visits inner node `cur`: children in near-to-far order, the nearest becomes
// st.cur (st.have), the others are pushed.  Returns kStepContinue, or
// kStepFallback when the stack is exhausted (the exact engine answers).
template <int LDS_STACK, bool ANY>
VR_HD int wide_node(const SceneView& sv, WideState& st, const WideStack& ws, uint32_t cur, float tmin,
                    float tmax_any) {
    const vf3 o = st.o, d = st.d;
    const u32x4* qp = (const u32x4*)(sv.qnodes + cur);
    u32x4 w0 = qp[0], w1 = qp[1], w2 = qp[2], w3 = qp[3];
    uint32_t e = w0.w;
    float sx = u2f((e & 0xFFu) << 23) * st.idx;
    float sy = u2f(((e >> 8) & 0xFFu) << 23) * st.idy;
    float sz = u2f(((e >> 16) & 0xFFu) << 23) * st.idz;
    float bx = (u2f(w0.x) - o.x) * st.idx;
    float by = (u2f(w0.y) - o.y) * st.idy;
    float bz = (u2f(w0.z) - o.z) * st.idz;
    uint32_t ch[4] = {w1.x, w1.y, w1.z, w1.w};
    // Near and far planes chosen once per node from the ray's direction signs
    // (all four children's bytes at once): with s > 0, fma(lo, s, b) <=
    // fma(hi, s, b), so this equals taking min/max per child and axis.
    const bool negx = st.idx < 0.f, negy = st.idy < 0.f, negz = st.idz < 0.f;
    uint32_t qnx = negx ? w2.w : w2.x, qfx = negx ? w2.x : w2.w;
    uint32_t qny = negy ? w3.x : w2.y, qfy = negy ? w2.y : w3.x;
    uint32_t qnz = negz ? w3.y : w2.z, qfz = negz ? w2.z : w3.y;
    float tcull = ANY ? tmax_any : st.tbest * kCullSlack;
    float tn[4];
    uint32_t cid[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float nx = fmaf((float)((qnx >> (8 * c)) & 0xFFu), sx, bx), fx = fmaf((float)((qfx >> (8 * c)) & 0xFFu), sx, bx);
        float ny = fmaf((float)((qny >> (8 * c)) & 0xFFu), sy, by), fy = fmaf((float)((qfy >> (8 * c)) & 0xFFu), sy, by);
        float nz = fmaf((float)((qnz >> (8 * c)) & 0xFFu), sz, bz), fz = fmaf((float)((qfz >> (8 * c)) & 0xFFu), sz, bz);
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
    if (tn[0] == kInf) return kStepContinue;
    st.cur = cid[0];
    st.have = true;
    bool ok = true;
#pragma unroll
    for (int c = 3; c >= 1; --c) {
        if (tn[c] == kInf) continue;
        ok = ok && wide_push<LDS_STACK>(st, ws, u32x2{cid[c], f2u(tn[c])});
    }
    return ok ? kStepContinue : kStepFallback;
}

// One traversal step (one node or one leaf).  Returns kStepContinue,
// kStepDone (result ready for wide_finish) or kStepFallback.
// (wo, wd): the query's world ray, for leaving an instance (INST).
template <int LDS_STACK, bool ANY, bool WT = false, bool INST = false>
VR_HD int wide_step(const SceneView& sv, WideState& st, const WideStack& ws, float tmin, float tmax_any, vf3 wo,
                    vf3 wd) {
    if (!st.have && !wide_pop_next<LDS_STACK, ANY>(st, ws)) return kStepDone;
    st.have = false;
    const uint32_t cur = st.cur;
    if (cur & kLeafBit) return wide_leaf<LDS_STACK, ANY, WT, INST>(sv, st, ws, cur, tmin, tmax_any);
    if (INST && cur == kInstExit) {
        wide_instance_exit<WT>(st, wo, wd);
        return kStepContinue;
    }
    if (INST && (cur & kInstEntry)) return wide_instance_enter<LDS_STACK, WT>(sv, st, ws, cur, tmin);
    return wide_node<LDS_STACK, ANY>(sv, st, ws, cur, tmin, tmax_any);
}

// Result of a finished query (after kStepDone).  Returns kStepDone or
// kStepFallback (winner not confirmed / any-hit candidates unconfirmed).
template <bool ANY, bool WT = false, bool INST = false>
VR_HD int wide_finish(const SceneView& sv, const WideState& st, float tmin, HitOut& out) {
    out.t = 0.f;
    out.u = 0.f;
    out.v = 0.f;
    out.prim = 0xFFFFFFFFu;
    if (ANY) {
        if (st.found) {
            out.t = st.tbest;
            out.u = st.bu;
            out.v = st.bv;
            out.prim = (st.best_slot & kAnalyticSlot) ? sv.analytic[st.best_slot & ~kAnalyticSlot].prim
                       : (INST && st.best_inst != kNoInstance) ? st.best_prim
                                                                : tri_word(sv.tris, st.best_slot, 9);
            return kStepDone;
        }
        return st.unconfirmed ? kStepFallback : kStepDone;
    }
    if (st.best_slot == 0xFFFFFFFFu) return kStepDone;
    if ((st.best_slot & kAnalyticSlot) || (INST && st.best_inst != kNoInstance)) {
        // analytic: confirmed when it was seeded; instance: nothing to confirm
        out.t = st.tbest;
        out.u = st.bu;
        out.v = st.bv;
        out.prim = st.best_prim;
        return kStepDone;
    }
    vf3 p0, p1, p2;
    load_tri_geom(sv.tris, st.best_slot, p0, p1, p2);
    // (a finished query is back in world space: every exit entry was popped)
    if (!WT && !confirm_winner(sv, st.o, st.d, tmin, st.tbest, p0, p1, p2, tri_word(sv.tris, st.best_slot, 11),
                               tri_word(sv.tris, st.best_slot, 9)))
        return kStepFallback;
    out.t = st.tbest;
    out.u = st.bu;
    out.v = st.bv;
    out.prim = tri_word(sv.tris, st.best_slot, 9);
    return kStepDone;
}

// Whole query for one ray (used by the host-side checks and simple kernels).
template <int LDS_STACK, bool ANY, bool WT = false, bool INST = false>
VR_HD int trace_wide(const SceneView& sv, vf3 o, vf3 d, float tmin, float tmax_any, const WideStack& ws,
                     HitOut& out) {
    WideState st;
    wide_init<WT>(st, o, d, sv.num_qnodes > 0);
    if (sv.num_analytic) wide_seed_analytic<ANY, WT>(sv, st, tmin, tmax_any);
    int rc;
    while ((rc = wide_step<LDS_STACK, ANY, WT, INST>(sv, st, ws, tmin, tmax_any, o, d)) == kStepContinue) {
    }
    if (rc == kStepFallback) {
        out.prim = 0xFFFFFFFFu;
        return kStepFallback;
    }
    return wide_finish<ANY, WT, INST>(sv, st, tmin, out);
}

// Exact reference engine for one ray (closest hit).  The two traversal
// stacks (64 entries each, element k at [k * stride]) are the caller's: local
// arrays, or a lane's slice of LDS.  WT: the watertight engine's fallback
// (conservative slab tests, watertight triangle test).
template <bool WT = false>
__host__ __device__ inline void trace_exact_stk(const SceneView& sv, vf3 o, vf3 d, float tmin, HitOut& out,
                                                int32_t* tstack, int32_t* mstack, int stride) {
    out.prim = 0xFFFFFFFFu;
    out.t = out.u = out.v = 0.f;
    if (!sv.top_root_valid) return;
    WTRay wt{};
    if (WT) wt = wt_prepare(d);
    bool have = false;
    float bt = 0.f;
    int tsp = 0;
    tstack[tsp++ * stride] = 0;
    while (tsp) {
        int32_t ni = tstack[--tsp * stride];
        RefNode tn = sv.top_nodes[ni];
        if (tn.obj < 0) {
            Box3 b{mk3(tn.lo[0], tn.lo[1], tn.lo[2]), mk3(tn.hi[0], tn.hi[1], tn.hi[2])};
            if (WT ? wt_slab_hit(b, o, d, tmin, kInf) : ref_slab_hit(b, o, d, tmin, kInf)) {
                tstack[tsp++ * stride] = tn.right;
                tstack[tsp++ * stride] = ni + 1;
            }
            continue;
        }
        uint32_t surf = (uint32_t)tn.obj;
        SurfaceInfo si = sv.surfaces[surf];
        if (si.analytic >= 0) {
            float t, u, v;
            if (analytic_hit(sv.analytic[si.analytic], o, d, tmin, t, u, v) &&
                (!have || t < bt || (t == bt && si.prim_base < out.prim))) {
                have = true;
                bt = t;
                out.t = t;
                out.u = u;
                out.v = v;
                out.prim = si.prim_base;
            }
            continue;
        }
        MeshInfo mi = sv.meshes[si.mesh];
        // an instance's mesh is walked with the ray in its object space
        vf3 mo = o, md = d;
        WTRay mwt = wt;
        float meps = -1.0f;
        if (si.instance >= 0) {
            instance_ray(sv.instances[si.instance], o, d, mo, md);
            if (WT) mwt = wt_prepare(md);
            meps = tmin * sv.instances[si.instance].det_inv;
        }
        auto test = [&](uint32_t tri) {
            const uint32_t* ix = sv.mesh_idx + 3 * (size_t)(mi.tri_base + tri);
            const float* vp = sv.vpos + 3 * (size_t)mi.vert_base;
            vf3 p0 = mk3(vp[3 * ix[0]], vp[3 * ix[0] + 1], vp[3 * ix[0] + 2]);
            vf3 p1 = mk3(vp[3 * ix[1]], vp[3 * ix[1] + 1], vp[3 * ix[1] + 2]);
            vf3 p2 = mk3(vp[3 * ix[2]], vp[3 * ix[2] + 1], vp[3 * ix[2] + 2]);
            MTResult r;
            if (WT ? !wt_hit(mwt, mo, p0, p1, p2, tmin, r) : !mt_hit(mo, md, p0, p1, p2, tmin, r, meps)) return;
            const uint32_t prim = si.prim_base + tri;
            if (!have || r.t < bt || (r.t == bt && prim < out.prim)) {
                have = true;
                bt = r.t;
                out.t = r.t;
                out.u = r.u;
                out.v = r.v;
                out.prim = prim;
            }
        };
        if (mi.root < 0) {
            for (uint32_t k = 0; k < mi.num_tris; ++k) test(k);
            continue;
        }
        int msp = 0;
        mstack[msp++ * stride] = mi.root;
        while (msp) {
            int32_t mj = mstack[--msp * stride];
            RefNode nd = sv.ref_nodes[mj];
            if (nd.obj >= 0) {
                test((uint32_t)nd.obj);
                continue;
            }
            Box3 b{mk3(nd.lo[0], nd.lo[1], nd.lo[2]), mk3(nd.hi[0], nd.hi[1], nd.hi[2])};
            if (!(WT ? wt_slab_hit(b, mo, md, tmin, kInf) : ref_slab_hit(b, mo, md, tmin, kInf))) continue;
            mstack[msp++ * stride] = nd.right;
            mstack[msp++ * stride] = mj + 1;
        }
    }
}

template <bool WT = false>
__host__ __device__ inline void trace_exact(const SceneView& sv, vf3 o, vf3 d, float tmin, HitOut& out) {
    int32_t tstack[64], mstack[64];
    trace_exact_stk<WT>(sv, o, d, tmin, out, tstack, mstack, 1);
}

}  // namespace vr
