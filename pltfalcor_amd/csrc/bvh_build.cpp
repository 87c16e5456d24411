// Binned-SAH BVH2 builder (host, multithreaded). Replaces the reference's DXR BLAS/TLAS build
// (Scene/Scene.cpp:3092-3400, :3624-3700) with a software BVH whose closest-hit results are
// independent of the tree (watertight intersection with a deterministic tie-break), so the tree
// only affects speed.
//
// Build: per-triangle boxes and centroids; recursive 128-bin SAH over all three axes, with an
// object-median split wherever the SAH split would leave a child that cannot be finished within
// kMaxBvh2Depth levels; subtrees of more than 32k triangles are built on their own threads
// (disjoint index ranges, partitioned in place). The binary tree is then collapsed into 4-wide
// nodes (each node adopts the children of its largest-area inner child until it has four) and
// flattened depth-first so that a parent and its first child are adjacent.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <future>
#include <limits>
#include <memory>
#include <thread>
#include "scene_host.hpp"

namespace plt {
namespace {

constexpr int kBins = 128;
constexpr uint32_t kParallelThreshold = 32768;

struct BNode {
    Aabb box;
    std::unique_ptr<BNode> l, r;
    uint32_t start = 0, count = 0;  // leaf range (count > 0 for leaves)
};

struct Builder {
    const std::vector<float>& pos;
    const std::vector<uint32_t>& idx;
    std::vector<Aabb> tbox;
    std::vector<float> cent;  // 3 per triangle
    std::vector<uint32_t> order;
    int max_leaf;
    std::atomic<int> threads_left;

    Builder(const std::vector<float>& p, const std::vector<uint32_t>& i, int ml)
        : pos(p), idx(i), max_leaf(ml), threads_left(int(std::max(1u, std::thread::hardware_concurrency())) * 2) {
        const size_t nt = idx.size() / 3;
        tbox.resize(nt);
        cent.resize(3 * nt);
        order.resize(nt);
        for (size_t t = 0; t < nt; ++t) {
            Aabb b;
            for (int v = 0; v < 3; ++v) b.grow(&pos[3 * size_t(idx[3 * t + v])]);
            tbox[t] = b;
            for (int a = 0; a < 3; ++a) cent[3 * t + a] = 0.5f * (b.lo[a] + b.hi[a]);
            order[t] = uint32_t(t);
        }
    }

    // Inner levels a subtree of n triangles needs at least: median splits halve it down to leaves.
    int levels_needed(uint32_t n) const {
        int l = 0;
        while (n > uint32_t(max_leaf)) {
            n = (n + 1) / 2;
            ++l;
        }
        return l;
    }

    // depth: this node's level (root = 1); invariant depth - 1 + levels_needed(count) <= kMaxBvh2Depth
    std::unique_ptr<BNode> build(uint32_t start, uint32_t count, int depth) {
        auto n = std::make_unique<BNode>();
        Aabb cb;
        for (uint32_t i = start; i < start + count; ++i) {
            n->box.grow(tbox[order[i]]);
            cb.grow(&cent[3 * size_t(order[i])]);
        }
        if (count <= uint32_t(max_leaf)) {
            n->start = start;
            n->count = count;
            return n;
        }
        // SAH over bins on each axis
        float best_cost = 3.4e38f;
        int best_axis = -1, best_split = -1;
        for (int a = 0; a < 3; ++a) {
            const float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0.f)) continue;
            const float scale = float(kBins) / ext;
            Aabb bb[kBins];
            uint32_t bc[kBins] = {};
            for (uint32_t i = start; i < start + count; ++i) {
                const uint32_t t = order[i];
                int b = int((cent[3 * size_t(t) + a] - cb.lo[a]) * scale);
                b = std::min(std::max(b, 0), kBins - 1);
                bb[b].grow(tbox[t]);
                ++bc[b];
            }
            float rarea[kBins];
            uint32_t rcount[kBins];
            Aabb acc;
            uint32_t c = 0;
            for (int b = kBins - 1; b > 0; --b) {
                acc.grow(bb[b]);
                c += bc[b];
                rarea[b] = acc.area();
                rcount[b] = c;
            }
            Aabb lacc;
            uint32_t lc = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                lacc.grow(bb[b]);
                lc += bc[b];
                if (lc == 0 || rcount[b + 1] == 0) continue;
                const float cost = lacc.area() * float(lc) + rarea[b + 1] * float(rcount[b + 1]);
                if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = b; }
            }
        }
        const float parea = n->box.area();
        const float leaf_cost = float(count);
        const float split_cost = parea > 0.f ? 1.f + best_cost / parea : 3.4e38f;
        uint32_t mid;
        if (best_axis < 0) {
            if (count <= 8u) { n->start = start; n->count = count; return n; }
            mid = start + count / 2;  // degenerate centroids: split by position in the list
        } else {
            if (count <= 8u && leaf_cost <= split_cost) { n->start = start; n->count = count; return n; }
            const float scale = float(kBins) / (cb.hi[best_axis] - cb.lo[best_axis]);
            auto it = std::partition(order.begin() + start, order.begin() + start + count, [&](uint32_t t) {
                int b = int((cent[3 * size_t(t) + best_axis] - cb.lo[best_axis]) * scale);
                b = std::min(std::max(b, 0), kBins - 1);
                return b <= best_split;
            });
            mid = uint32_t(it - order.begin());
            if (mid == start || mid == start + count) mid = start + count / 2;
        }
        if (depth + std::max(levels_needed(mid - start), levels_needed(start + count - mid)) > kMaxBvh2Depth) {
            // the depth bound: split at the centroid median of the widest axis (children of
            // ceil(count / 2) and floor(count / 2) triangles need one level less than this node)
            int axis = 0;
            for (int a = 1; a < 3; ++a)
                if (cb.hi[a] - cb.lo[a] > cb.hi[axis] - cb.lo[axis]) axis = a;
            mid = start + (count + 1) / 2;
            std::nth_element(order.begin() + start, order.begin() + mid, order.begin() + start + count,
                             [&](uint32_t x, uint32_t y) { return cent[3 * size_t(x) + axis] < cent[3 * size_t(y) + axis]; });
        }
        const uint32_t lcnt = mid - start, rcnt = count - lcnt;
        if (count > kParallelThreshold && threads_left.fetch_sub(1) > 0) {
            auto fut = std::async(std::launch::async, [this, start, lcnt, depth] { return build(start, lcnt, depth + 1); });
            n->r = build(mid, rcnt, depth + 1);
            n->l = fut.get();
        } else {
            n->l = build(start, lcnt, depth + 1);
            n->r = build(mid, rcnt, depth + 1);
        }
        return n;
    }
};

int32_t leaf_ref(uint32_t start, uint32_t count) { return ~int32_t((start << 3) | (count - 1)); }

BvhNode empty_node() {
    BvhNode n{};
    const float inf = std::numeric_limits<float>::infinity();
    for (int c = 0; c < 4; ++c) {
        n.lox[c] = n.hix[c] = n.loy[c] = n.hiy[c] = n.loz[c] = n.hiz[c] = inf;
        n.child[c] = 0;
    }
    return n;
}

void set_child_box(BvhNode& n, int c, const Aabb& b) {
    n.lox[c] = b.lo[0]; n.hix[c] = b.hi[0];
    n.loy[c] = b.lo[1]; n.hiy[c] = b.hi[1];
    n.loz[c] = b.lo[2]; n.hiz[c] = b.hi[2];
}

// Depth-first flattening into 4-wide nodes; returns the index of the node created for `b` (an
// inner node of the binary tree).
int32_t flatten(const BNode* b, std::vector<BvhNode>& out, int depth, int& max_depth) {
    max_depth = std::max(max_depth, depth);
    const BNode* ch[4] = {b->l.get(), b->r.get(), nullptr, nullptr};
    int n = 2;
    while (n < 4) {  // adopt the children of the largest inner child
        int best = -1;
        float best_area = -1.f;
        for (int i = 0; i < n; ++i)
            if (ch[i]->count == 0 && ch[i]->box.area() > best_area) {
                best_area = ch[i]->box.area();
                best = i;
            }
        if (best < 0) break;
        const BNode* x = ch[best];
        ch[best] = x->l.get();
        ch[n++] = x->r.get();
    }
    const int32_t me = int32_t(out.size());
    out.push_back(empty_node());
    for (int c = 0; c < n; ++c) {
        set_child_box(out[me], c, ch[c]->box);
        const int32_t ref = ch[c]->count > 0 ? leaf_ref(ch[c]->start, ch[c]->count) : flatten(ch[c], out, depth + 1, max_depth);
        out[me].child[c] = ref;
    }
    return me;
}
}  // namespace

HostBvh build_bvh(const std::vector<float>& pos, const std::vector<uint32_t>& idx, int max_leaf) {
    const auto t0 = std::chrono::steady_clock::now();
    HostBvh out;
    const uint32_t nt = uint32_t(idx.size() / 3);
    if (nt == 0) return out;
    Builder b(pos, idx, std::min(max_leaf, 8));
    std::unique_ptr<BNode> root = b.build(0, nt, 1);
    if (root->count > 0) {  // a single leaf: a root with one used slot
        BvhNode n = empty_node();
        set_child_box(n, 0, root->box);
        n.child[0] = leaf_ref(root->start, root->count);
        out.nodes.push_back(n);
        out.max_depth = 1;
    } else {
        out.nodes.reserve(size_t(nt) / 2 + 1);
        flatten(root.get(), out.nodes, 1, out.max_depth);
    }
    out.tri_order = std::move(b.order);
    out.build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

}  // namespace plt
