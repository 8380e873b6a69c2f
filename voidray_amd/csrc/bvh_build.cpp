// Host-side acceleration-structure builders.
//
//  * build_reference_tree: the median-split binary tree of
//    voidray_renderer/src/core/bvh.rs:48-130, flattened in pre-order.  It is
//    needed for two things: the in-order leaf rank that resolves equal-t ties
//    exactly as merge_optionals (bvh.rs:164-181) does, and the exact
//    reference engine that re-traces the rare rays whose closest hit cannot be
//    confirmed by the fast engine.
//  * build_qbvh: a binned-SAH binary BVH collapsed to 4-wide nodes with 8-bit
//    outward-rounded child boxes (QNode, 64 B) and triangles in leaf order.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
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
void build_reference_tree(const std::vector<Box3>& bounds, std::vector<RefNode>& out, std::vector<uint32_t>& rank) {
    out.clear();
    size_t n = bounds.size();
    rank.assign(n, 0);
    if (n == 0) return;
    RefBuilder b(bounds);
    std::vector<uint32_t> objs(n);
    for (size_t i = 0; i < n; ++i) objs[i] = (uint32_t)i;
    out.reserve(2 * n);
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int par = 0;
    while ((1u << par) < hw && par < 6) ++par;
    b.build(objs, out, par);
    // in-order leaf rank (pre-order layout: leaves appear in in-order order)
    uint32_t r = 0;
    for (auto& nd : out)
        if (nd.obj >= 0) rank[nd.obj] = r++;
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

struct SahBuilder {
    const std::vector<BBox>& pb;   // prim boxes
    std::vector<float> cen;        // 3 per prim
    std::vector<uint32_t> idx;
    std::vector<BNode> nodes;
    std::atomic<int32_t> next{0};
    static constexpr int kBins = 16;

    explicit SahBuilder(const std::vector<BBox>& b) : pb(b) {
        size_t n = b.size();
        cen.resize(3 * n);
        idx.resize(n);
        for (size_t i = 0; i < n; ++i) {
            for (int a = 0; a < 3; ++a) cen[3 * i + a] = 0.5f * (b[i].lo[a] + b[i].hi[a]);
            idx[i] = (uint32_t)i;
        }
        nodes.resize(std::max<size_t>(1, 2 * n));
    }

    int32_t alloc() { return next.fetch_add(1); }

    void build(int32_t ni, uint32_t start, uint32_t count, int par) {
        BNode& nd = nodes[ni];
        BBox box, cb;
        box.empty();
        cb.empty();
        for (uint32_t i = start; i < start + count; ++i) {
            box.grow(pb[idx[i]]);
            cb.grow(&cen[3 * idx[i]]);
        }
        nd.box = box;
        nd.child[0] = nd.child[1] = -1;
        nd.start = start;
        nd.count = count;
        if (count <= 1) return;
        // binned SAH
        float best_cost = kInf;
        int best_axis = -1, best_split = -1;
        for (int a = 0; a < 3; ++a) {
            float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0.f)) continue;
            BBox bb[kBins];
            uint32_t bc[kBins] = {0};
            for (auto& x : bb) x.empty();
            float k = kBins * (1.f - 1e-6f) / ext;
            for (uint32_t i = start; i < start + count; ++i) {
                uint32_t p = idx[i];
                int bin = (int)((cen[3 * p + a] - cb.lo[a]) * k);
                bin = std::min(std::max(bin, 0), kBins - 1);
                bb[bin].grow(pb[p]);
                ++bc[bin];
            }
            float rarea[kBins];
            uint32_t rcnt[kBins];
            BBox acc;
            acc.empty();
            uint32_t c = 0;
            for (int i = kBins - 1; i > 0; --i) {
                acc.grow(bb[i]);
                c += bc[i];
                rarea[i] = acc.area();
                rcnt[i] = c;
            }
            acc.empty();
            c = 0;
            for (int i = 0; i < kBins - 1; ++i) {
                acc.grow(bb[i]);
                c += bc[i];
                if (c == 0 || rcnt[i + 1] == 0) continue;
                float cost = acc.area() * (float)c + rarea[i + 1] * (float)rcnt[i + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_split = i;
                }
            }
        }
        float leaf_cost = box.area() * (float)count;
        float parent_area = box.area();
        bool split_ok = best_axis >= 0;
        if (split_ok && count <= kMaxLeaf) {
            // SAH: traversal cost ~1 box test vs 1 triangle test per prim
            float split_cost = parent_area * 1.0f + best_cost;
            if (split_cost >= leaf_cost) return;
        }
        uint32_t mid;
        if (split_ok) {
            int a = best_axis;
            float ext = cb.hi[a] - cb.lo[a];
            float k = kBins * (1.f - 1e-6f) / ext;
            auto it = std::partition(idx.begin() + start, idx.begin() + start + count, [&](uint32_t p) {
                int bin = (int)((cen[3 * p + a] - cb.lo[a]) * k);
                bin = std::min(std::max(bin, 0), kBins - 1);
                return bin <= best_split;
            });
            mid = (uint32_t)(it - idx.begin());
        } else {
            if (count <= kMaxLeaf) return;
            mid = start + count / 2;  // all centroids equal: split in the middle
        }
        if (mid == start || mid == start + count) mid = start + count / 2;
        int32_t l = alloc(), r = alloc();
        nd.child[0] = l;
        nd.child[1] = r;
        if (par > 0 && count > 32768) {
            auto fut = std::async(std::launch::async, [&] { build(r, mid, start + count - mid, par - 1); });
            build(l, start, mid - start, par - 1);
            fut.get();
        } else {
            build(l, start, mid - start, 0);
            build(r, mid, start + count - mid, 0);
        }
    }
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

// ---- early split clipping (Ernst & Greiner 2007) ---------------------------
// A long, thin triangle lying diagonally has a box far larger than itself, and
// every ray through that empty space visits its leaf.  Such a triangle is
// referenced from several leaves instead: its box is cut at the midpoint of
// the longest axis, the triangle is clipped to each half (in double), and each
// piece gets the box of its clipped polygon, recursively while a cut removes
// enough box area.  The pieces' boxes cover the whole triangle (each point of
// it lies in one half, and each piece box is rounded outward and padded), so a
// ray that hits the triangle reaches at least one leaf holding it; duplicate
// records of a triangle give the same t and rank, so the result is unchanged.
struct DV3 {
    double v[3];
};

static int clip_half(const DV3* in, int n, int axis, double plane, bool keep_below, DV3* out) {
    int m = 0;
    for (int i = 0; i < n; ++i) {
        const DV3& a = in[i];
        const DV3& b = in[(i + 1) % n];
        double da = keep_below ? plane - a.v[axis] : a.v[axis] - plane;
        double db = keep_below ? plane - b.v[axis] : b.v[axis] - plane;
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

static double dbox_area(const double lo[3], const double hi[3]) {
    double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

static void poly_bounds(const DV3* p, int n, double lo[3], double hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = INFINITY;
        hi[a] = -INFINITY;
    }
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[i].v[a]);
            hi[a] = std::max(hi[a], p[i].v[a]);
        }
}

constexpr int kSplitDepth = 2;          // at most 4 pieces per triangle
constexpr double kSplitKeep = 0.7;      // cut only if the halves keep < 70% of the box area
constexpr double kSplitRatio = 14.0;    // split only triangles whose box area exceeds 14x their area

static double tri_area(const DV3* p) {
    double e1[3], e2[3];
    for (int a = 0; a < 3; ++a) {
        e1[a] = p[1].v[a] - p[0].v[a];
        e2[a] = p[2].v[a] - p[0].v[a];
    }
    double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
    return 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
}

static void split_piece(const DV3* poly, int n, int depth, const BBox& tri_box, std::vector<BBox>& out) {
    double lo[3], hi[3];
    poly_bounds(poly, n, lo, hi);
    // An axis-aligned right triangle has a box of 4x its area and a tilted
    // grid triangle (a heightfield's) about 6-12x: those are left whole.
    bool emit = depth >= kSplitDepth || (depth == 0 && !(dbox_area(lo, hi) > kSplitRatio * tri_area(poly)));
    DV3 pl[9], pr[9];
    int nl = 0, nr = 0;
    if (!emit) {
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (hi[a] - lo[a] > hi[axis] - lo[axis]) axis = a;
        double mid = 0.5 * (lo[axis] + hi[axis]);
        nl = clip_half(poly, n, axis, mid, true, pl);
        nr = clip_half(poly, n, axis, mid, false, pr);
        if (nl < 3 || nr < 3) {
            emit = true;
        } else {
            double ll[3], lh[3], rl[3], rh[3];
            poly_bounds(pl, nl, ll, lh);
            poly_bounds(pr, nr, rl, rh);
            if (dbox_area(ll, lh) + dbox_area(rl, rh) > kSplitKeep * dbox_area(lo, hi)) emit = true;
        }
    }
    if (emit) {
        BBox b;
        for (int a = 0; a < 3; ++a) {
            float fl = (float)lo[a], fh = (float)hi[a];
            if ((double)fl > lo[a]) fl = std::nextafter(fl, -INFINITY);
            if ((double)fh < hi[a]) fh = std::nextafter(fh, INFINITY);
            b.lo[a] = fl;
            b.hi[a] = fh;
        }
        pad_box(b);
        for (int a = 0; a < 3; ++a) {  // never larger than the whole triangle's box
            b.lo[a] = std::max(b.lo[a], tri_box.lo[a]);
            b.hi[a] = std::min(b.hi[a], tri_box.hi[a]);
        }
        out.push_back(b);
        return;
    }
    split_piece(pl, nl, depth + 1, tri_box, out);
    split_piece(pr, nr, depth + 1, tri_box, out);
}

static void split_references(const std::vector<TriRec>& tris, const std::vector<BBox>& pb, std::vector<BBox>& rb,
                             std::vector<uint32_t>& rtri) {
    const size_t n = tris.size();
    unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    size_t chunk = (n + hw - 1) / hw;
    std::vector<std::vector<BBox>> boxes(hw);
    std::vector<std::vector<uint32_t>> owners(hw);
    std::vector<std::future<void>> jobs;
    for (unsigned w = 0; w < hw; ++w) {
        jobs.push_back(std::async(std::launch::async, [&, w] {
            size_t b0 = w * chunk, b1 = std::min(n, b0 + chunk);
            std::vector<BBox> tmp;
            for (size_t i = b0; i < b1; ++i) {
                DV3 poly[3];
                for (int a = 0; a < 3; ++a) {
                    poly[0].v[a] = tris[i].p0[a];
                    poly[1].v[a] = tris[i].p1[a];
                    poly[2].v[a] = tris[i].p2[a];
                }
                tmp.clear();
                split_piece(poly, 3, 0, pb[i], tmp);
                if (tmp.size() <= 1) {
                    boxes[w].push_back(pb[i]);
                    owners[w].push_back((uint32_t)i);
                } else {
                    for (auto& b : tmp) {
                        boxes[w].push_back(b);
                        owners[w].push_back((uint32_t)i);
                    }
                }
            }
        }));
    }
    for (auto& j : jobs) j.get();
    rb.clear();
    rtri.clear();
    for (unsigned w = 0; w < hw; ++w) {
        rb.insert(rb.end(), boxes[w].begin(), boxes[w].end());
        rtri.insert(rtri.end(), owners[w].begin(), owners[w].end());
    }
}

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
               uint32_t& depth_out, std::string& err) {
    size_t n = tris_in.size();
    qnodes.clear();
    tris_out.clear();
    depth_out = 0;
    if (n == 0) return 0;
    if (n >= (1u << 27)) {
        err = "too many triangles for the 27-bit leaf encoding";
        return -1;
    }
    // padded triangle boxes: outward by a few ulps of the coordinate magnitude
    std::vector<BBox> pb(n);
    for (size_t i = 0; i < n; ++i) {
        BBox b;
        b.empty();
        b.grow(tris_in[i].p0);
        b.grow(tris_in[i].p1);
        b.grow(tris_in[i].p2);
        pad_box(b);
        pb[i] = b;
    }
    // references: a triangle, or a clipped piece of one (split_references)
    std::vector<BBox> rb;
    std::vector<uint32_t> rtri;
    split_references(tris_in, pb, rb, rtri);
    const size_t nr = rb.size();
    if (nr >= (1u << 27)) {
        err = "too many triangle references for the 27-bit leaf encoding";
        return -1;
    }
    SahBuilder sb(rb);
    int32_t root = sb.alloc();
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int par = 0;
    while ((1u << par) < hw && par < 6) ++par;
    sb.build(root, 0, (uint32_t)nr, par);
    std::vector<BNode>& bn = sb.nodes;

    // triangle records in leaf order (a split triangle appears once per piece)
    tris_out.resize(nr);
    for (size_t i = 0; i < nr; ++i) tris_out[i] = tris_in[rtri[sb.idx[i]]];

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
    }
    return 0;
}

}  // namespace vr
