// Binned-SAH BVH2 builder (host, multithreaded). Replaces the reference's DXR BLAS/TLAS build
// (Scene/Scene.cpp:3092-3400, :3624-3700) with a software BVH whose closest-hit results are
// independent of the tree (watertight intersection with a deterministic tie-break), so the tree
// only affects speed.
//
// Build: per-triangle boxes and centroids; recursive 32-bin SAH over all three axes; subtrees of
// more than 32k triangles are built on their own threads (disjoint index ranges, partitioned in
// place). The result is flattened depth-first so that a parent and its first child are adjacent.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <future>
#include <memory>
#include <thread>
#include "scene_host.hpp"

namespace plt {
namespace {

constexpr int kBins = 32;
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

    std::unique_ptr<BNode> build(uint32_t start, uint32_t count) {
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
        const uint32_t lcnt = mid - start, rcnt = count - lcnt;
        if (count > kParallelThreshold && threads_left.fetch_sub(1) > 0) {
            auto fut = std::async(std::launch::async, [this, start, lcnt] { return build(start, lcnt); });
            n->r = build(mid, rcnt);
            n->l = fut.get();
        } else {
            n->l = build(start, lcnt);
            n->r = build(mid, rcnt);
        }
        return n;
    }
};

int32_t leaf_ref(uint32_t start, uint32_t count) { return ~int32_t((start << 3) | (count - 1)); }

void set_child_box(BvhNode& n, int c, const Aabb& b) {
    float* bx = c == 0 ? n.b0x : n.b1x;
    float* by = c == 0 ? n.b0y : n.b1y;
    bx[0] = b.lo[0]; bx[1] = b.hi[0];
    by[0] = b.lo[1]; by[1] = b.hi[1];
    n.bz[2 * c] = b.lo[2]; n.bz[2 * c + 1] = b.hi[2];
}

// Depth-first flattening; returns the index of the node created for `b` (an inner node).
int32_t flatten(const BNode* b, std::vector<BvhNode>& out, int depth, int& max_depth) {
    max_depth = std::max(max_depth, depth);
    const int32_t me = int32_t(out.size());
    out.emplace_back();
    const BNode* ch[2] = {b->l.get(), b->r.get()};
    for (int c = 0; c < 2; ++c) {
        set_child_box(out[me], c, ch[c]->box);
        if (ch[c]->count > 0) {
            out[me].child[c] = leaf_ref(ch[c]->start, ch[c]->count);
        } else {
            const int32_t ci = flatten(ch[c], out, depth + 1, max_depth);
            out[me].child[c] = ci;
        }
    }
    return me;
}
}  // namespace

HostBvh build_bvh(const std::vector<float>& pos, const std::vector<uint32_t>& idx, int max_leaf) {
    HostBvh out;
    const uint32_t nt = uint32_t(idx.size() / 3);
    if (nt == 0) return out;
    Builder b(pos, idx, std::min(max_leaf, 8));
    std::unique_ptr<BNode> root = b.build(0, nt);
    if (root->count > 0) {
        // Single leaf: a root whose two children are the same leaf (a duplicate test is harmless;
        // an inverted "empty" box would pass a min/max slab test and must not be used).
        BvhNode n{};
        set_child_box(n, 0, root->box);
        set_child_box(n, 1, root->box);
        n.child[0] = leaf_ref(root->start, root->count);
        n.child[1] = leaf_ref(root->start, root->count);
        out.nodes.push_back(n);
        out.max_depth = 1;
    } else {
        out.nodes.reserve(2 * size_t(nt) / 2 + 1);
        flatten(root.get(), out.nodes, 1, out.max_depth);
    }
    out.tri_order = std::move(b.order);
    return out;
}

}  // namespace plt
