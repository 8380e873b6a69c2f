// Host-side acceleration-structure builders.
//
//  * build_reference_tree: the median-split binary tree of
//    voidray_renderer/src/core/bvh.rs:48-130, flattened in pre-order.  Its
//    boxes decide which triangles the reference reaches: the fast engine
//    confirms its winner against the ancestors' slab tests, and the exact
//    reference engine walks it for the rare rays whose closest hit cannot be
//    confirmed that way.
//  * build_qbvh: a binned-SAH binary BVH with spatial splits collapsed to
//    4-wide nodes with 8-bit outward-rounded child boxes (QNode, 64 B) and
//    triangle records in leaf order.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <memory>
#include <mutex>
#include <thread>

#include "bvh_build.h"

namespace vr {

// ---------------------------------------------------------------------------
// reference median-split tree
// ---------------------------------------------------------------------------

static inline int32_t total_key(float f) {
    int32_t k;
    std::memcpy(&k, &f, 4);
    k ^= (int32_t)(((uint32_t)(k >> 31)) >> 1);  // f32::total_cmp ordering
    return k;
}

namespace {
struct RefBuilder {
    const std::vector<Box3>& bounds;
    std::vector<int32_t> ckey[3];  // total_cmp keys of the box centroids
    explicit RefBuilder(const std::vector<Box3>& b) : bounds(b) {
        size_t n = b.size();
        for (int a = 0; a < 3; ++a) ckey[a].resize(n);
        for (size_t i = 0; i < n; ++i) {
            vf3 c = (b[i].lo + b[i].hi) / 2.0f;  // AABB::centroid (aabb.rs:33-35)
            ckey[0][i] = total_key(c.x);
            ckey[1][i] = total_key(c.y);
            ckey[2][i] = total_key(c.z);
        }
    }
    // Builds the subtree of `objs` into `out` (pre-order, local indices).
    void build(std::vector<uint32_t>& objs, std::vector<RefNode>& out, int par_depth) {
        size_t n = objs.size();
        if (n == 0) return;
        if (n == 1) {
            RefNode leaf{};
            leaf.right = -1;
            leaf.obj = (int32_t)objs[0];
            out.push_back(leaf);
            return;
        }
        Box3 box = bounds[objs[0]];
        vf3 c0 = (bounds[objs[0]].lo + bounds[objs[0]].hi) / 2.0f;
        Box3 cb{c0, c0};
        for (size_t i = 1; i < n; ++i) {
            const Box3& b = bounds[objs[i]];
            box.lo = mk3(fminf(box.lo.x, b.lo.x), fminf(box.lo.y, b.lo.y), fminf(box.lo.z, b.lo.z));
            box.hi = mk3(fmaxf(box.hi.x, b.hi.x), fmaxf(box.hi.y, b.hi.y), fmaxf(box.hi.z, b.hi.z));
            vf3 c = (b.lo + b.hi) / 2.0f;
            cb.lo = mk3(fminf(cb.lo.x, c.x), fminf(cb.lo.y, c.y), fminf(cb.lo.z, c.z));
            cb.hi = mk3(fmaxf(cb.hi.x, c.x), fmaxf(cb.hi.y, c.y), fmaxf(cb.hi.z, c.z));
        }
        vf3 sp = cb.hi - cb.lo;
        int axis = (sp.x > sp.y && sp.x > sp.z) ? 0 : ((sp.y > sp.x && sp.y > sp.z) ? 1 : 2);
        const std::vector<int32_t>& key = ckey[axis];
        // stable sort by key == sort by (key, position in the current list)
        std::vector<uint64_t> kp(n);
        for (size_t i = 0; i < n; ++i) {
            uint32_t k = (uint32_t)key[objs[i]] ^ 0x80000000u;  // signed -> unsigned order
            kp[i] = ((uint64_t)k << 32) | (uint64_t)i;
        }
        std::sort(kp.begin(), kp.end());
        size_t half = n / 2;
        std::vector<uint32_t> left(half), right(n - half);
        for (size_t i = 0; i < n; ++i) {
            uint32_t o = objs[(uint32_t)kp[i]];
            if (i < half) left[i] = o;
            else right[i - half] = o;
        }
        std::vector<uint32_t>().swap(objs);
        std::vector<uint64_t>().swap(kp);
        size_t me = out.size();
        RefNode node{};
        node.lo[0] = box.lo.x; node.lo[1] = box.lo.y; node.lo[2] = box.lo.z;
        node.hi[0] = box.hi.x; node.hi[1] = box.hi.y; node.hi[2] = box.hi.z;
        node.obj = -1;
        out.push_back(node);
        if (par_depth > 0 && n > 65536) {
            std::vector<RefNode> rsub;
            rsub.reserve(2 * right.size());
            auto fut = std::async(std::launch::async, [&] { build(right, rsub, par_depth - 1); });
            build(left, out, par_depth - 1);
            fut.get();
            int32_t roff = (int32_t)out.size();
            for (auto& r : rsub)
                if (r.obj < 0) r.right += roff;
            out[me].right = roff;
            out.insert(out.end(), rsub.begin(), rsub.end());
        } else {
            build(left, out, 0);
            out[me].right = (int32_t)out.size();
            // children's `right` links are local to `out` already
            size_t before = out.size();
            (void)before;
            build(right, out, 0);
        }
    }
};
}  // namespace

// Rebase the `right` links of a subtree that was built into its own vector
// starting at index 0 (done by the parallel branch above).
void build_reference_tree(const std::vector<Box3>& bounds, std::vector<RefNode>& out) {
    out.clear();
    size_t n = bounds.size();
    if (n == 0) return;
    RefBuilder b(bounds);
    std::vector<uint32_t> objs(n);
    for (size_t i = 0; i < n; ++i) objs[i] = (uint32_t)i;
    out.reserve(2 * n);
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int par = 0;
    while ((1u << par) < hw && par < 6) ++par;
    b.build(objs, out, par);
}

// ---------------------------------------------------------------------------
// SAH binary BVH -> 4-wide quantised BVH
// ---------------------------------------------------------------------------

namespace {

struct BBox {
    float lo[3], hi[3];
    void empty() {
        for (int a = 0; a < 3; ++a) {
            lo[a] = kInf;
            hi[a] = -kInf;
        }
    }
    void grow(const BBox& b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    void grow(const float* p) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0) || !(dy >= 0) || !(dz >= 0)) return 0.f;
        return 2.f * (dx * dy + dy * dz + dz * dx);
    }
};

struct BNode {
    BBox box;
    int32_t child[2];  // -1 for leaves
    uint32_t start, count;
};

// Outward padding of a triangle (or piece) box by 8 ulps of the coordinate
// magnitude: absorbs the rounding of the slab test's subtraction.
static void pad_box(BBox& b) {
    for (int a = 0; a < 3; ++a) {
        float m = std::max(std::fabs(b.lo[a]), std::fabs(b.hi[a]));
        float pad = std::max(m * 1.1920929e-7f * 8.f, 1e-30f);
        b.lo[a] -= pad;
        b.hi[a] += pad;
    }
}

static float overlap_area(const BBox& a, const BBox& b) {
    BBox c;
    for (int k = 0; k < 3; ++k) {
        c.lo[k] = std::max(a.lo[k], b.lo[k]);
        c.hi[k] = std::min(a.hi[k], b.hi[k]);
        if (!(c.lo[k] <= c.hi[k])) return 0.f;
    }
    return c.area();
}

// ---- triangle clipping (double) ---------------------------------------------
struct DV3 {
    double v[3];
};
constexpr int kMaxPoly = 32;

// Sutherland-Hodgman against the closed half-space x[axis] <= plane (below) or
// x[axis] >= plane; returns the vertex count of the clipped polygon.
static int clip_half(const DV3* in, int n, int axis, double plane, bool below, DV3* out) {
    int m = 0;
    for (int i = 0; i < n && m < kMaxPoly - 2; ++i) {
        const DV3& a = in[i];
        const DV3& b = in[(i + 1) % n];
        double da = below ? plane - a.v[axis] : a.v[axis] - plane;
        double db = below ? plane - b.v[axis] : b.v[axis] - plane;
        if (da >= 0) out[m++] = a;
        if ((da >= 0) != (db >= 0)) {
            double t = da / (da - db);
            DV3 p;
            for (int k = 0; k < 3; ++k) p.v[k] = a.v[k] + t * (b.v[k] - a.v[k]);
            p.v[axis] = plane;
            out[m++] = p;
        }
    }
    return m;
}

// A top-level entry standing for an instance: p0/p1 hold its world box.
static inline bool is_instance_rec(const TriRec& t) {
    uint32_t m;
    std::memcpy(&m, &t.p2[0], 4);
    return m == kInstanceMarker;
}

// Float box (rounded outward, padded) of the part of triangle `t` inside `box`
// (closed); false when that part is empty.  Never larger than `box`.  An
// instance entry's part is its box clipped to `box` (its geometry lies inside
// its box).
static bool clipped_box(const TriRec& t, const BBox& box, BBox& out) {
    if (is_instance_rec(t)) {
        for (int k = 0; k < 3; ++k) {
            out.lo[k] = std::max(t.p0[k], box.lo[k]);
            out.hi[k] = std::min(t.p1[k], box.hi[k]);
            if (!(out.lo[k] <= out.hi[k])) return false;
        }
        return true;
    }
    DV3 a[kMaxPoly], b[kMaxPoly];
    for (int k = 0; k < 3; ++k) {
        a[0].v[k] = t.p0[k];
        a[1].v[k] = t.p1[k];
        a[2].v[k] = t.p2[k];
    }
    int n = 3;
    for (int ax = 0; ax < 3 && n > 0; ++ax) {
        n = clip_half(a, n, ax, (double)box.lo[ax], false, b);
        if (n > 0) n = clip_half(b, n, ax, (double)box.hi[ax], true, a);
    }
    if (n == 0) return false;
    for (int k = 0; k < 3; ++k) {
        double lo = INFINITY, hi = -INFINITY;
        for (int i = 0; i < n; ++i) {
            lo = std::min(lo, a[i].v[k]);
            hi = std::max(hi, a[i].v[k]);
        }
        float fl = (float)lo, fh = (float)hi;
        if ((double)fl > lo) fl = std::nextafter(fl, -INFINITY);
        if ((double)fh < hi) fh = std::nextafter(fh, INFINITY);
        out.lo[k] = fl;
        out.hi[k] = fh;
    }
    pad_box(out);
    for (int k = 0; k < 3; ++k) {
        out.lo[k] = std::max(out.lo[k], box.lo[k]);
        out.hi[k] = std::min(out.hi[k], box.hi[k]);
    }
    return true;
}

// ---- binned SAH with spatial splits (SBVH: Stich, Friedrich & Dietrich 2009)
// A node is split by the cheaper of an object split (binned on reference
// centroids) and, where the object split's children overlap by more than
// kAlpha of the root's area, a spatial split: a plane cuts the node and every
// reference straddling it is clipped into a piece on each side.  A long thin
// diagonal triangle therefore ends up in several leaves, each with the box of
// its clipped piece, instead of one leaf whose box is mostly empty space.
// Every point of a triangle lies in the piece on its side of each cut, and
// each piece box is rounded outward and padded, so a ray that hits the
// triangle reaches at least one leaf holding it; the copies of a triangle give
// the same t and prim id, so the closest hit does not change.
struct SRef {
    BBox box;
    uint32_t tri;
};

struct SNode {
    BBox box;
    SNode* child[2];
    uint32_t* tris;  // leaf: triangle indices (arena storage)
    uint32_t count;
};

// Bump allocator owned by one build task.
struct Arena {
    static constexpr size_t kNodes = 1 << 14, kIds = 1 << 16;
    std::vector<std::unique_ptr<SNode[]>> nodes;
    std::vector<std::unique_ptr<uint32_t[]>> ids;
    size_t nused = kNodes, iused = kIds;
    SNode* node() {
        if (nused == kNodes) {
            nodes.emplace_back(new SNode[kNodes]);
            nused = 0;
        }
        return &nodes.back()[nused++];
    }
    uint32_t* take_ids(uint32_t c) {
        if (iused + c > kIds) {
            ids.emplace_back(new uint32_t[kIds]);
            iused = 0;
        }
        uint32_t* p = &ids.back()[iused];
        iused += c;
        return p;
    }
};

struct SpatialBuilder {
    const std::vector<TriRec>& tris;
    float root_area = 0.f;
    std::mutex mu;
    std::vector<std::unique_ptr<Arena>> arenas;
    static constexpr int kObjBins = 16;
    static constexpr int kSpatialBins = 16;
    static constexpr float kAlpha = 1e-4f;  // overlap (fraction of the root area) that enables spatial splits
    static constexpr int kMaxDepth = 64;    // spatial splits stop below this depth
    static constexpr uint64_t kMaxDup = 8;  // references per triangle at most, over the whole tree

    explicit SpatialBuilder(const std::vector<TriRec>& t) : tris(t) {}

    Arena* new_arena() {
        std::lock_guard<std::mutex> g(mu);
        arenas.emplace_back(new Arena);
        return arenas.back().get();
    }

    SNode* make_leaf(const std::vector<SRef>& refs, const BBox& box, Arena& ar) {
        SNode* nd = ar.node();
        nd->box = box;
        nd->child[0] = nd->child[1] = nullptr;
        nd->count = (uint32_t)refs.size();
        nd->tris = ar.take_ids(nd->count);
        for (size_t i = 0; i < refs.size(); ++i) nd->tris[i] = refs[i].tri;
        return nd;
    }

    // `budget`: most references this subtree may hold (>= refs.size()).  A
    // node's budget is shared by its children in proportion to their
    // references, so the tree never holds more than kMaxDup references per
    // triangle whatever the geometry, and the split decisions stay independent
    // of thread timing.
    SNode* build(std::vector<SRef>& refs, Arena& ar, int depth, int par, uint64_t budget) {
        const uint32_t n = (uint32_t)refs.size();
        BBox box, cb;
        box.empty();
        cb.empty();
        for (const SRef& r : refs) {
            box.grow(r.box);
            float c[3];
            for (int a = 0; a < 3; ++a) c[a] = 0.5f * (r.box.lo[a] + r.box.hi[a]);
            cb.grow(c);
        }
        if (depth == 0) root_area = box.area();
        if (n <= 1) return make_leaf(refs, box, ar);

        // object split: binned SAH over reference centroids
        float obj_cost = kInf;
        int obj_axis = -1, obj_split = -1;
        BBox obj_l, obj_r;
        for (int a = 0; a < 3; ++a) {
            float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0.f)) continue;
            BBox bb[kObjBins];
            uint32_t bc[kObjBins] = {0};
            for (auto& x : bb) x.empty();
            float k = kObjBins * (1.f - 1e-6f) / ext;
            for (const SRef& r : refs) {
                int bin = (int)((0.5f * (r.box.lo[a] + r.box.hi[a]) - cb.lo[a]) * k);
                bin = std::min(std::max(bin, 0), kObjBins - 1);
                bb[bin].grow(r.box);
                ++bc[bin];
            }
            BBox rbox[kObjBins];
            uint32_t rcnt[kObjBins];
            BBox acc;
            acc.empty();
            uint32_t c = 0;
            for (int i = kObjBins - 1; i > 0; --i) {
                acc.grow(bb[i]);
                c += bc[i];
                rbox[i] = acc;
                rcnt[i] = c;
            }
            acc.empty();
            c = 0;
            for (int i = 0; i < kObjBins - 1; ++i) {
                acc.grow(bb[i]);
                c += bc[i];
                if (c == 0 || rcnt[i + 1] == 0) continue;
                float cost = acc.area() * (float)c + rbox[i + 1].area() * (float)rcnt[i + 1];
                if (cost < obj_cost) {
                    obj_cost = cost;
                    obj_axis = a;
                    obj_split = i;
                    obj_l = acc;
                    obj_r = rbox[i + 1];
                }
            }
        }

        // spatial split where the object split leaves heavy overlap
        float sp_cost = kInf;
        int sp_axis = -1;
        float sp_plane = 0.f;
        bool try_spatial = depth < kMaxDepth && budget > n &&
                           (obj_axis < 0 || overlap_area(obj_l, obj_r) > kAlpha * root_area);
        if (try_spatial) {
            for (int a = 0; a < 3; ++a) {
                float lo = box.lo[a], hi = box.hi[a];
                float ext = hi - lo;
                if (!(ext > 0.f)) continue;
                float planes[kSpatialBins + 1];
                for (int i = 0; i <= kSpatialBins; ++i) planes[i] = lo + ext * ((float)i / kSpatialBins);
                planes[kSpatialBins] = hi;
                BBox bb[kSpatialBins];
                uint32_t enter[kSpatialBins] = {0}, leave[kSpatialBins] = {0};
                for (auto& x : bb) x.empty();
                float k = kSpatialBins * (1.f - 1e-6f) / ext;
                for (const SRef& r : refs) {
                    int b0 = std::min(std::max((int)((r.box.lo[a] - lo) * k), 0), kSpatialBins - 1);
                    int b1 = std::min(std::max((int)((r.box.hi[a] - lo) * k), 0), kSpatialBins - 1);
                    ++enter[b0];
                    ++leave[b1];
                    if (b0 == b1) {
                        bb[b0].grow(r.box);
                        continue;
                    }
                    for (int bi = b0; bi <= b1; ++bi) {
                        BBox slab = r.box;
                        slab.lo[a] = std::max(slab.lo[a], planes[bi]);
                        slab.hi[a] = std::min(slab.hi[a], planes[bi + 1]);
                        BBox piece;
                        if (slab.lo[a] <= slab.hi[a] && clipped_box(tris[r.tri], slab, piece)) bb[bi].grow(piece);
                    }
                }
                BBox rbox[kSpatialBins];
                uint32_t rcnt[kSpatialBins];
                BBox acc;
                acc.empty();
                uint32_t c = 0;
                for (int i = kSpatialBins - 1; i > 0; --i) {
                    acc.grow(bb[i]);
                    c += leave[i];
                    rbox[i] = acc;
                    rcnt[i] = c;
                }
                acc.empty();
                c = 0;
                for (int i = 0; i < kSpatialBins - 1; ++i) {
                    acc.grow(bb[i]);
                    c += enter[i];
                    if (c == 0 || rcnt[i + 1] == 0 || c >= n || rcnt[i + 1] >= n) continue;
                    float cost = acc.area() * (float)c + rbox[i + 1].area() * (float)rcnt[i + 1];
                    if (cost < sp_cost) {
                        sp_cost = cost;
                        sp_axis = a;
                        sp_plane = planes[i + 1];
                    }
                }
            }
        }

        const bool use_spatial = sp_axis >= 0 && sp_cost < obj_cost;
        const float best = use_spatial ? sp_cost : obj_cost;
        if (n <= kMaxLeaf) {
            // SAH: traversal cost ~1 box test vs 1 triangle test per reference
            if (!(best < kInf) || box.area() * 1.0f + best >= box.area() * (float)n) return make_leaf(refs, box, ar);
        }
        std::vector<SRef> left, right;
        if (use_spatial) {
            const int a = sp_axis;
            left.reserve(n);
            right.reserve(n);
            for (const SRef& r : refs) {
                if (r.box.hi[a] <= sp_plane) {
                    left.push_back(r);
                } else if (r.box.lo[a] >= sp_plane) {
                    right.push_back(r);
                } else if (is_instance_rec(tris[r.tri])) {
                    // an instance is not cut: a copy in two leaves would walk
                    // its whole object-space BVH twice
                    (0.5f * (r.box.lo[a] + r.box.hi[a]) < sp_plane ? left : right).push_back(r);
                } else {
                    BBox lb = r.box, rb = r.box, lp, rp;
                    lb.hi[a] = sp_plane;
                    rb.lo[a] = sp_plane;
                    bool hl = clipped_box(tris[r.tri], lb, lp), hr = clipped_box(tris[r.tri], rb, rp);
                    if (hl) left.push_back(SRef{lp, r.tri});
                    if (hr) right.push_back(SRef{rp, r.tri});
                    if (!hl && !hr) left.push_back(r);  // (numerically empty on both sides)
                }
            }
            if (left.empty() || right.empty() || left.size() + right.size() > budget) {
                left.clear();
                right.clear();
            }
        }
        if (left.empty() && right.empty()) {
            if (obj_axis >= 0) {
                const int a = obj_axis;
                float k = kObjBins * (1.f - 1e-6f) / (cb.hi[a] - cb.lo[a]);
                for (const SRef& r : refs) {
                    int bin = (int)((0.5f * (r.box.lo[a] + r.box.hi[a]) - cb.lo[a]) * k);
                    bin = std::min(std::max(bin, 0), kObjBins - 1);
                    (bin <= obj_split ? left : right).push_back(r);
                }
            }
            if (left.empty() || right.empty()) {
                if (n <= kMaxLeaf) return make_leaf(refs, box, ar);
                left.assign(refs.begin(), refs.begin() + n / 2);  // all centroids equal
                right.assign(refs.begin() + n / 2, refs.end());
            }
        }
        std::vector<SRef>().swap(refs);
        const uint64_t nl = left.size(), nr = right.size();
        const uint64_t bl = budget * nl / (nl + nr), br = budget - bl;  // bl >= nl, br >= nr
        SNode* nd = ar.node();
        nd->box = box;
        nd->tris = nullptr;
        nd->count = 0;
        if (par > 0 && n > 32768) {
            Arena* ra = new_arena();
            auto fut = std::async(std::launch::async, [&, ra] { return build(right, *ra, depth + 1, par - 1, br); });
            nd->child[0] = build(left, ar, depth + 1, par - 1, bl);
            nd->child[1] = fut.get();
        } else {
            nd->child[0] = build(left, ar, depth + 1, 0, bl);
            nd->child[1] = build(right, ar, depth + 1, 0, br);
        }
        return nd;
    }
};

// Per-axis power-of-two grid for a node: smallest e with 255 * 2^e >= extent.
static int pick_exponent(float lo, float hi) {
    double ext = (double)hi - (double)lo;
    if (!(ext > 0)) return -100;
    int e = (int)std::ceil(std::log2(ext / 255.0));
    while (std::ldexp(255.0, e) < ext) ++e;
    return std::max(e, -100);
}

}  // namespace

int build_qbvh(const std::vector<TriRec>& tris_in, std::vector<QNode>& qnodes, std::vector<TriRec>& tris_out,
               uint32_t& depth_out, std::string& err, QbvhCheck* check) {
    size_t n = tris_in.size();
    qnodes.clear();
    tris_out.clear();
    depth_out = 0;
    if (n == 0) return 0;
    if (n >= (1u << 27)) {
        err = "too many triangles for the 27-bit leaf encoding";
        return -1;
    }
    // references: every triangle with its padded box
    std::vector<SRef> refs(n);
    for (size_t i = 0; i < n; ++i) {
        BBox b;
        b.empty();
        b.grow(tris_in[i].p0);
        b.grow(tris_in[i].p1);
        if (!is_instance_rec(tris_in[i])) b.grow(tris_in[i].p2);
        pad_box(b);
        refs[i] = SRef{b, (uint32_t)i};
    }
    SpatialBuilder sb(tris_in);
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int par = 0;
    while ((1u << par) < hw && par < 6) ++par;
    Arena* ra = sb.new_arena();
    SNode* sroot = sb.build(refs, *ra, 0, par, SpatialBuilder::kMaxDup * (uint64_t)n);

    // flatten in pre-order; triangle records in leaf order (a triangle split
    // by spatial cuts appears once per leaf holding a piece of it)
    std::vector<BNode> bn;
    std::vector<uint32_t> order;
    bn.reserve(2 * n);
    order.reserve(n + n / 2);
    {
        std::vector<std::pair<SNode*, int32_t>> st;  // node, slot of the parent's child link (-1: root)
        st.push_back({sroot, -1});
        while (!st.empty()) {
            auto [sn, link] = st.back();
            st.pop_back();
            int32_t me = (int32_t)bn.size();
            if (link >= 0) bn[link >> 1].child[link & 1] = me;
            BNode b;
            b.box = sn->box;
            b.child[0] = b.child[1] = -1;
            b.start = b.count = 0;
            if (!sn->child[0]) {
                b.start = (uint32_t)order.size();
                b.count = sn->count;
                order.insert(order.end(), sn->tris, sn->tris + sn->count);
            }
            bn.push_back(b);
            if (sn->child[0]) {
                st.push_back({sn->child[1], 2 * me + 1});
                st.push_back({sn->child[0], 2 * me});
            }
        }
    }
    if (order.size() >= (1u << 27)) {
        err = "too many triangle references for the 27-bit leaf encoding";
        return -1;
    }
    const int32_t root = 0;
    tris_out.resize(order.size());
    for (size_t i = 0; i < order.size(); ++i) tris_out[i] = tris_in[order[i]];

    // collapse to 4-wide: children of a wide node = up to 4 binary nodes
    struct Pending {
        int32_t bnode;
        uint32_t qindex;
        uint32_t depth;
    };
    std::vector<Pending> work;
    qnodes.reserve(n / 2 + 1);
    auto is_leaf = [&](int32_t b) { return bn[b].child[0] < 0; };
    auto leaf_code = [&](int32_t b) -> uint32_t {
        return kLeafBit | ((bn[b].count - 1) << 27) | bn[b].start;
    };
    // A leaf-only tree gets a single wide node with one leaf child.
    qnodes.push_back(QNode{});
    work.push_back({root, 0, 1});
    size_t wi = 0;
    while (wi < work.size()) {
        Pending pw = work[wi++];
        depth_out = std::max(depth_out, pw.depth);
        int32_t kids[4];
        int nk = 0;
        if (is_leaf(pw.bnode)) {
            kids[nk++] = pw.bnode;
        } else {
            kids[nk++] = bn[pw.bnode].child[0];
            kids[nk++] = bn[pw.bnode].child[1];
            while (nk < 4) {
                // open the internal child with the largest surface area
                int best = -1;
                float ba = -1.f;
                for (int i = 0; i < nk; ++i)
                    if (!is_leaf(kids[i]) && bn[kids[i]].box.area() > ba) {
                        ba = bn[kids[i]].box.area();
                        best = i;
                    }
                if (best < 0) break;
                int32_t b = kids[best];
                kids[best] = bn[b].child[0];
                kids[nk++] = bn[b].child[1];
            }
        }
        // quantise
        BBox nb;
        nb.empty();
        for (int i = 0; i < nk; ++i) nb.grow(bn[kids[i]].box);
        QNode q{};
        q.ox = nb.lo[0];
        q.oy = nb.lo[1];
        q.oz = nb.lo[2];
        q.nchild = (uint8_t)nk;
        float org[3] = {nb.lo[0], nb.lo[1], nb.lo[2]};
        uint8_t ex[3];
        for (int a = 0; a < 3; ++a) {
            int e = pick_exponent(nb.lo[a], nb.hi[a]);
            for (;;) {
                bool ok = true;
                double s = std::ldexp(1.0, e);
                for (int i = 0; i < nk && ok; ++i) {
                    const BBox& cbx = bn[kids[i]].box;
                    double ql = std::floor(((double)cbx.lo[a] - (double)org[a]) / s);
                    double qh = std::ceil(((double)cbx.hi[a] - (double)org[a]) / s);
                    if (ql < 0) ql = 0;
                    // exact decode must contain the child box
                    while ((double)org[a] + ql * s > (double)cbx.lo[a] && ql > 0) ql -= 1;
                    while ((double)org[a] + qh * s < (double)cbx.hi[a]) qh += 1;
                    if (qh > 255 || (double)org[a] + ql * s > (double)cbx.lo[a]) {
                        ok = false;
                        break;
                    }
                    q.qlo[a][i] = (uint8_t)ql;
                    q.qhi[a][i] = (uint8_t)qh;
                }
                if (ok) break;
                ++e;
            }
            ex[a] = (uint8_t)(e + 127);
        }
        q.ex = ex[0];
        q.ey = ex[1];
        q.ez = ex[2];
        for (int i = 0; i < 4; ++i) q.child[i] = kInvalidChild;
        for (int i = nk; i < 4; ++i)
            for (int a = 0; a < 3; ++a) {
                q.qlo[a][i] = 255;
                q.qhi[a][i] = 0;
            }
        for (int i = 0; i < nk; ++i) {
            if (is_leaf(kids[i])) {
                q.child[i] = leaf_code(kids[i]);
            } else {
                uint32_t ci = (uint32_t)qnodes.size();
                qnodes.push_back(QNode{});
                q.child[i] = ci;
                work.push_back({kids[i], ci, pw.depth + 1});
            }
        }
        qnodes[pw.qindex] = q;
        if (check) {
            // read back what was stored, decoded as the traversal decodes it
            const QNode& w = qnodes[pw.qindex];
            const double wo[3] = {w.ox, w.oy, w.oz};
            const int we[3] = {w.ex, w.ey, w.ez};
            for (int i = 0; i < nk; ++i)
                for (int a = 0; a < 3; ++a) {
                    const double s = std::ldexp(1.0, we[a] - 127);
                    const double lo = wo[a] + w.qlo[a][i] * s, hi = wo[a] + w.qhi[a][i] * s;
                    const BBox& ex = bn[kids[i]].box;
                    ++check->boxes_checked;
                    if (!(lo <= (double)ex.lo[a] && hi >= (double)ex.hi[a])) ++check->boxes_not_containing;
                }
        }
    }
    return 0;
}

}  // namespace vr
