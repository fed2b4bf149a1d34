// crx_cull.cpp — the node cull over many boxes: a hierarchy of index ranges
// whose answer equals each box's own positive-vertex test (SNTriangles.h:29-105).
#include <cmath>
#include <vector>

#include "crx_host.h"

namespace crx {

namespace {

bool box_finite(const crx_box3& b) {
    return std::isfinite(b.min.x) && std::isfinite(b.min.y) && std::isfinite(b.min.z) && std::isfinite(b.max.x) &&
           std::isfinite(b.max.y) && std::isfinite(b.max.z);
}

// Every corner of the box passes every plane (the corner nearest each plane
// passes): then every box inside it passes box_in_frustum.
bool box_inside_frustum(const float planes[16], const crx_box3& wb) {
    for (int p = 0; p < 4; p++) {
        const float* pl = planes + 4 * p;
        const f3 n{pl[0], pl[1], pl[2]};
        const f3 c{n.x >= 0.0f ? wb.min.x : wb.max.x, n.y >= 0.0f ? wb.min.y : wb.max.y,
                   n.z >= 0.0f ? wb.min.z : wb.max.z};
        if (dot3(n, c) + pl[3] < 0.0f) return false;
    }
    return true;
}

constexpr uint32_t kNodeLeaf = 4;

}  // namespace

void NodeTree::build(const crx_box3* boxes, uint32_t n) {
    ranges.clear();
    if (n == 0) return;
    ranges.reserve(2 * ((n + kNodeLeaf - 1) / kNodeLeaf));
    // ranges are made parent first; children are filled in after
    ranges.push_back(Range{crx_box3{}, 0, n, 0, 0, true});
    for (size_t k = 0; k < ranges.size(); k++) {
        const uint32_t lo = ranges[k].lo, hi = ranges[k].hi;
        if (hi - lo > kNodeLeaf) {
            const uint32_t mid = lo + (hi - lo) / 2;
            ranges[k].left = (uint32_t)ranges.size();
            ranges.push_back(Range{crx_box3{}, lo, mid, 0, 0, true});
            ranges[k].right = (uint32_t)ranges.size();
            ranges.push_back(Range{crx_box3{}, mid, hi, 0, 0, true});
        }
    }
    // boxes bottom-up: each range's box covers every coordinate of its nodes'
    // boxes (min and max of both corners, so an inverted box is covered too)
    for (size_t k = ranges.size(); k-- > 0;) {
        Range& r = ranges[k];
        crx_box3 b{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
        bool finite = true;
        auto cover = [&](const crx_box3& c) {
            for (const crx_vec3& p : {c.min, c.max}) {
                b.min.x = std::fmin(b.min.x, p.x); b.min.y = std::fmin(b.min.y, p.y); b.min.z = std::fmin(b.min.z, p.z);
                b.max.x = std::fmax(b.max.x, p.x); b.max.y = std::fmax(b.max.y, p.y); b.max.z = std::fmax(b.max.z, p.z);
            }
        };
        if (r.left) {
            cover(ranges[r.left].box);
            cover(ranges[r.right].box);
            finite = ranges[r.left].finite && ranges[r.right].finite;
        } else {
            for (uint32_t i = r.lo; i < r.hi; i++) {
                cover(boxes[i]);
                finite = finite && box_finite(boxes[i]);
            }
        }
        r.box = b;
        r.finite = finite;
    }
}

void NodeTree::visible(const float planes[16], const crx_box3* boxes, std::vector<uint32_t>& out) const {
    out.clear();
    if (ranges.empty()) return;
    // the shortcuts rest on every product of the plane test being finite
    bool finitePlanes = true;
    for (int k = 0; k < 16; k++) finitePlanes = finitePlanes && std::isfinite(planes[k]);
    if (!finitePlanes) {
        for (uint32_t i = 0; i < ranges[0].hi; i++)
            if (box_in_frustum(planes, boxes[i])) out.push_back(i);
        return;
    }
    uint32_t stack[64];
    int top = 0;
    stack[top++] = 0;
    while (top > 0) {
        const Range& r = ranges[stack[--top]];
        if (r.finite) {
            if (!box_in_frustum(planes, r.box)) continue;
            if (box_inside_frustum(planes, r.box)) {
                for (uint32_t i = r.lo; i < r.hi; i++) out.push_back(i);
                continue;
            }
        }
        if (r.left) {
            // right first on the stack: the left range's nodes come out first
            stack[top++] = r.right;
            stack[top++] = r.left;
        } else {
            for (uint32_t i = r.lo; i < r.hi; i++)
                if (box_in_frustum(planes, boxes[i])) out.push_back(i);
        }
    }
}

}  // namespace crx

using namespace crx;

extern "C" {

crx_status crx_frustum_cull(const crx_view* view, const crx_box3* boxes, uint32_t n, uint32_t* ids,
                            uint32_t capacity, uint32_t* count) {
    if (!view || !count || (n > 0 && !boxes)) return fail(CRX_INVALID_ARG, "NULL argument");
    float planes[16];
    frustum_planes(*view, planes);
    NodeTree tree;
    tree.build(boxes, n);
    std::vector<uint32_t> vis;
    tree.visible(planes, boxes, vis);
    for (uint32_t k = 0; k < (uint32_t)vis.size() && ids && k < capacity; k++) ids[k] = vis[k];
    *count = (uint32_t)vis.size();
    return CRX_OK;
}

}  // extern "C"
