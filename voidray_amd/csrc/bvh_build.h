#pragma once
#include <string>
#include <vector>

#include "internal.h"

namespace vr {
// Median-split tree of bvh.rs:48-130 over `bounds`, flattened in pre-order.
void build_reference_tree(const std::vector<Box3>& bounds, std::vector<RefNode>& out);
// Self-check of a wide BVH's stored encoding: child boxes decoded as the
// kernels decode them (origin + q * 2^(e - 127), here in double) against the
// exact boxes they stand for.
struct QbvhCheck {
    uint64_t boxes_checked = 0;     // (child, axis) pairs
    uint64_t boxes_not_containing = 0;
};
// SAH-built 4-wide quantised BVH; tris_out = tris_in in leaf order.
int build_qbvh(const std::vector<TriRec>& tris_in, std::vector<QNode>& qnodes, std::vector<TriRec>& tris_out,
               uint32_t& depth, std::string& err, QbvhCheck* check = nullptr);
}  // namespace vr
