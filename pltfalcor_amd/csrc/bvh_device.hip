// Device BVH builder: parallel locally-ordered clustering (PLOC; Meister and Bittner 2018,
// "Parallel Locally-Ordered Clustering for Bounding Volume Hierarchy Construction", IEEE TVCG
// 24(3)) of triangles sorted along a 63-bit Morton curve, collapsed top-down into the 4-wide nodes
// the traversal reads (BvhNode, 128 B) with leaves of at most four triangles. It stands where the
// reference builds its acceleration structures on the device (Scene/Scene.cpp:3092-3400 BLAS,
// :3624-3700 TLAS). Closest hits do not depend on the tree (traverse.hpp), so a scene renders
// bit-identically with this tree and with the host builder's.
//
// Stages, on one stream (the host reads one word back per clustering iteration and per tree
// level, to size the next launch, and the last clusters for the top of the tree):
//   1. triangle boxes, and the centroid bounds (block partials, then one reducing block);
//   2. Morton keys of the centroids (21 bits per axis) and a radix sort of (key, triangle) pairs
//      (rocPRIM; stable, so equal keys stay in triangle order and the build is deterministic);
//   3. clustering: every cluster picks, among the clusters at most kRadius places away along the
//      curve, the one whose union box has the smallest surface area (ties: pair_rank);
//      mutual pairs merge into a new node at the lower place, and an inclusive scan of the
//      (kept, created) flags compacts the list in curve order and numbers the new nodes. The last
//      kTopClusters clusters are joined top-down on the host by a full-sweep SAH (greedy clustering
//      makes its last merges between large clusters, where a split that sees the whole set does
//      better: the 1080p render of the 5.24M-triangle scene on the device tree gained 3.6 %);
//   4. collapse, one launch per level: a 4-wide node starts from its binary node and replaces its
//      inner child of largest area by that child's two children until it has four; a child of at
//      most kLeafMax triangles becomes a leaf, whose triangles it writes in leaf order (3 x float4,
//      the triangle id in the first record's w), and the others become the next level's nodes;
//   5. a depth-first renumbering of the nodes (subtree sizes bottom-up, positions top-down).
// Every merge and every level sees only data written by earlier launches, so the build needs no
// device-wide fences between workgroups.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <utility>
#include <vector>
#include "plt_internal.hpp"
#include "scene_device.hpp"
#include "traverse.hpp"

namespace plt {
namespace {

constexpr int kBlock = 256;
constexpr int kRadius = 8;                // clustering search radius (places along the curve)
constexpr uint32_t kBoxBlocks = 2048;     // blocks of the triangle-box pass (grid-stride beyond)
constexpr uint32_t kTopClusters = 1024;   // the top of the tree by SAH over this many last clusters
constexpr uint32_t kLeafMax = 4;          // triangles per leaf at most
// Clustering iterations above kTopClusters clusters at most (each merges at least one pair;
// geometry that merges only a few per iteration, such as sizes growing geometrically, would also
// make a tree too deep for the traversal stack).
constexpr int kMaxIterations = 4096;
// Levels of 4-wide nodes the traversal stack holds (three entries pushed per level).
constexpr int kMaxLevels = 1 + (kTraversalStackDepth + kTraversalSpill) / 3;

// Spreads the low 21 bits of v to every third bit.
__device__ __forceinline__ uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | x << 32) & 0x1F00000000FFFFull;
    x = (x | x << 16) & 0x1F0000FF0000FFull;
    x = (x | x << 8) & 0x100F00F00F00F00Full;
    x = (x | x << 4) & 0x10C30C30C30C30C3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

// Surface area of the union of two boxes, up to the factor 2 (symmetric in the two boxes).
__device__ __forceinline__ float union_area(float4 a, float4 A, float4 b, float4 B) {
    const float dx = fmaxf(A.x, B.x) - fminf(a.x, b.x);
    const float dy = fmaxf(A.y, B.y) - fminf(a.y, b.y);
    const float dz = fmaxf(A.z, B.z) - fminf(a.z, b.z);
    return dx * dy + dy * dz + dz * dx;
}
// Order among the candidates j = i + d of equal union area, so that the preferred pair under
// (area, |i - j|, whether {i, j} = {2k, 2k + 1}, min(i, j)) -- a key symmetric in i and j -- is
// also the pair each of the two prefers: the globally least pair is then mutual and every
// iteration merges, and equal boxes (duplicated triangles) pair off as (2k, 2k + 1) into a
// balanced subtree instead of a chain.
__device__ __forceinline__ int pair_rank(int i, int d) {
    const int ad = d < 0 ? -d : d;
    return 4 * ad + ((i + d) != (i ^ 1) ? 2 : 0) + (d > 0 ? 1 : 0);
}
__device__ __forceinline__ float box_area(float4 a, float4 A) {
    const float dx = A.x - a.x, dy = A.y - a.y, dz = A.z - a.z;
    return dx * dy + dy * dz + dz * dx;
}

// Triangle boxes (xyz; w unused) and, per block, the bounds of their centroids (a grid-stride loop
// over the triangles, then a block reduction: one partial per block, no atomics).
__global__ void __launch_bounds__(kBlock) tri_box_kernel(uint32_t n, const float* __restrict__ pos,
                                                         const uint32_t* __restrict__ idx, float4* __restrict__ tlo,
                                                         float4* __restrict__ thi, float* __restrict__ partial) {
    constexpr float kInf = __builtin_inff();
    float cmin[3] = {kInf, kInf, kInf}, cmax[3] = {-kInf, -kInf, -kInf};
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
        float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            const float* p = pos + 3 * size_t(idx[3 * size_t(t) + v]);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = fminf(lo[a], p[a]);
                hi[a] = fmaxf(hi[a], p[a]);
            }
        }
        tlo[t] = make_float4(lo[0], lo[1], lo[2], 0.f);
        thi[t] = make_float4(hi[0], hi[1], hi[2], 0.f);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float c = 0.5f * (lo[a] + hi[a]);
            cmin[a] = fminf(cmin[a], c);
            cmax[a] = fmaxf(cmax[a], c);
        }
    }
    __shared__ float red[kBlock / 64][6];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        for (int o = 32; o > 0; o >>= 1) {
            cmin[a] = fminf(cmin[a], __shfl_xor(cmin[a], o));
            cmax[a] = fmaxf(cmax[a], __shfl_xor(cmax[a], o));
        }
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            red[wave][a] = cmin[a];
            red[wave][3 + a] = cmax[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        float v = red[0][threadIdx.x];
        for (int w = 1; w < kBlock / 64; ++w) v = threadIdx.x < 3 ? fminf(v, red[w][threadIdx.x]) : fmaxf(v, red[w][threadIdx.x]);
        partial[6 * blockIdx.x + threadIdx.x] = v;
    }
}

// The centroid bounds from the blocks' partials (one block): cb = (min xyz, max xyz).
__global__ void __launch_bounds__(kBlock) bounds_kernel(uint32_t blocks, const float* __restrict__ partial,
                                                        float* __restrict__ cb) {
    __shared__ float red[kBlock / 64][6];
    constexpr float kInf = __builtin_inff();
    float v[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
    for (uint32_t b = threadIdx.x; b < blocks; b += kBlock) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            v[a] = fminf(v[a], partial[6 * b + a]);
            v[3 + a] = fmaxf(v[3 + a], partial[6 * b + 3 + a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        for (int o = 32; o > 0; o >>= 1) {
            v[a] = fminf(v[a], __shfl_xor(v[a], o));
            v[3 + a] = fmaxf(v[3 + a], __shfl_xor(v[3 + a], o));
        }
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 6; ++a) red[wave][a] = v[a];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        float r = red[0][threadIdx.x];
        for (int w = 1; w < kBlock / 64; ++w) r = threadIdx.x < 3 ? fminf(r, red[w][threadIdx.x]) : fmaxf(r, red[w][threadIdx.x]);
        cb[threadIdx.x] = r;
    }
}

// 63-bit Morton key of each centroid in the centroid bounds; the value is the triangle id.
__global__ void __launch_bounds__(kBlock) morton_kernel(uint32_t n, const float4* __restrict__ tlo,
                                                        const float4* __restrict__ thi, const float* __restrict__ cb,
                                                        uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const float4 lo = tlo[t], hi = thi[t];
    const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float b0 = cb[a], b1 = cb[3 + a];
        const float ext = b1 - b0;
        const float s = ext > 0.f ? 2097151.f / ext : 0.f;
        q[a] = uint32_t(fminf(fmaxf((c[a] - b0) * s, 0.f), 2097151.f));
    }
    keys[t] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
    vals[t] = t;
}

// Node p < n is the leaf of the triangle at place p of the curve. Node records: lo = (box lo, left
// child or triangle id), hi = (box hi, right child); cnt = triangles below. Cluster records:
// lo = (box lo, node), hi = (box hi, triangles below).
__global__ void __launch_bounds__(kBlock) leaf_init_kernel(uint32_t n, const uint32_t* __restrict__ ids,
                                                           const float4* __restrict__ tlo, const float4* __restrict__ thi,
                                                           float4* __restrict__ nlo, float4* __restrict__ nhi,
                                                           uint32_t* __restrict__ ncnt, float4* __restrict__ clo,
                                                           float4* __restrict__ chi) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const uint32_t t = ids[p];
    const float4 lo = tlo[t], hi = thi[t];
    nlo[p] = make_float4(lo.x, lo.y, lo.z, __uint_as_float(t));
    nhi[p] = make_float4(hi.x, hi.y, hi.z, __uint_as_float(0xFFFFFFFFu));
    ncnt[p] = 1u;
    clo[p] = make_float4(lo.x, lo.y, lo.z, __uint_as_float(p));
    chi[p] = make_float4(hi.x, hi.y, hi.z, __uint_as_float(1u));
}

// Nearest neighbour of every cluster within kRadius places (the window staged in LDS).
__global__ void __launch_bounds__(kBlock) nn_kernel(uint32_t n, const float4* __restrict__ clo,
                                                    const float4* __restrict__ chi, uint32_t* __restrict__ nn) {
    __shared__ float4 slo[kBlock + 2 * kRadius], shi[kBlock + 2 * kRadius];
    const int base = int(blockIdx.x) * kBlock;
    for (int k = threadIdx.x; k < kBlock + 2 * kRadius; k += kBlock) {
        const int j = base - kRadius + k;
        if (j >= 0 && j < int(n)) {
            slo[k] = clo[j];
            shi[k] = chi[j];
        }
    }
    __syncthreads();
    const int i = base + int(threadIdx.x);
    if (i >= int(n)) return;
    const float4 a = slo[threadIdx.x + kRadius], A = shi[threadIdx.x + kRadius];
    float best = __builtin_inff();
    int bj = i, br = 1 << 30;
    for (int d = -kRadius; d <= kRadius; ++d) {
        const int j = i + d;
        if (d == 0 || j < 0 || j >= int(n)) continue;
        const float s = union_area(a, A, slo[threadIdx.x + kRadius + d], shi[threadIdx.x + kRadius + d]);
        const int r = pair_rank(i, d);
        if (s < best || (s == best && r < br)) {
            best = s;
            bj = j;
            br = r;
        }
    }
    nn[i] = uint32_t(bj);
}

// (kept << 32) | created per cluster: the higher place of a mutual pair is dropped, the lower one
// creates the merged node.
__global__ void __launch_bounds__(kBlock) flag_kernel(uint32_t n, const uint32_t* __restrict__ nn,
                                                      uint64_t* __restrict__ flags) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = nn[i];
    const bool mutual = nn[j] == i;
    const uint64_t kept = (mutual && j < i) ? 0u : 1u;
    const uint64_t created = (mutual && i < j) ? 1u : 0u;
    flags[i] = (kept << 32) | created;
}

__global__ void __launch_bounds__(kBlock) merge_kernel(uint32_t n, const float4* __restrict__ clo,
                                                       const float4* __restrict__ chi, const uint32_t* __restrict__ nn,
                                                       const uint64_t* __restrict__ scan, uint32_t node_base,
                                                       float4* __restrict__ nlo, float4* __restrict__ nhi,
                                                       uint32_t* __restrict__ ncnt, float4* __restrict__ clo2,
                                                       float4* __restrict__ chi2) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = nn[i];
    const bool mutual = nn[j] == i;
    if (mutual && j < i) return;
    const uint64_t s = scan[i];
    const uint32_t at = uint32_t(s >> 32) - 1u;
    if (!mutual) {
        clo2[at] = clo[i];
        chi2[at] = chi[i];
        return;
    }
    const uint32_t id = node_base + uint32_t(s) - 1u;
    const float4 a = clo[i], A = chi[i], b = clo[j], B = chi[j];
    const float4 lo = make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.f);
    const float4 hi = make_float4(fmaxf(A.x, B.x), fmaxf(A.y, B.y), fmaxf(A.z, B.z), 0.f);
    const uint32_t cnt = __float_as_uint(A.w) + __float_as_uint(B.w);
    nlo[id] = make_float4(lo.x, lo.y, lo.z, a.w);
    nhi[id] = make_float4(hi.x, hi.y, hi.z, b.w);
    ncnt[id] = cnt;
    clo2[at] = make_float4(lo.x, lo.y, lo.z, __uint_as_float(id));
    chi2[at] = make_float4(hi.x, hi.y, hi.z, __uint_as_float(cnt));
}

// One level of the collapse: items (binary node, first leaf-order place) become the 4-wide nodes
// level_base + k; inner children are appended to `next` (their node index is level_base + m plus
// their slot there).
__global__ void __launch_bounds__(kBlock) collapse_kernel(uint32_t m, const uint2* __restrict__ items,
                                                          uint32_t level_base, uint2* __restrict__ next,
                                                          uint32_t* __restrict__ next_count, uint32_t n_tris,
                                                          const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                                                          const uint32_t* __restrict__ ncnt, const float* __restrict__ pos,
                                                          const uint32_t* __restrict__ idx, float4* __restrict__ nodes,
                                                          float4* __restrict__ tris) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= m) return;
    const uint2 it = items[k];
    uint32_t ch[4] = {it.x, 0u, 0u, 0u};
    uint32_t cnt[4] = {ncnt[it.x], 0u, 0u, 0u};
    int nc = 1;
    while (nc < 4) {
        int best = -1;
        float best_area = -1.f;
        for (int c = 0; c < nc; ++c) {
            if (cnt[c] <= kLeafMax) continue;
            const float s = box_area(nlo[ch[c]], nhi[ch[c]]);
            if (s > best_area) {
                best_area = s;
                best = c;
            }
        }
        if (best < 0) break;
        const uint32_t x = ch[best];
        const uint32_t l = __float_as_uint(nlo[x].w), r = __float_as_uint(nhi[x].w);
        ch[best] = l;
        cnt[best] = ncnt[l];
        ch[nc] = r;
        cnt[nc] = ncnt[r];
        ++nc;
    }
    constexpr float kInf = __builtin_inff();
    float lx[4] = {kInf, kInf, kInf, kInf}, hx[4] = {kInf, kInf, kInf, kInf};
    float ly[4] = {kInf, kInf, kInf, kInf}, hy[4] = {kInf, kInf, kInf, kInf};
    float lz[4] = {kInf, kInf, kInf, kInf}, hz[4] = {kInf, kInf, kInf, kInf};
    int32_t ref[4] = {0, 0, 0, 0};
    uint32_t off = it.y;
    for (int c = 0; c < nc; ++c) {
        const float4 lo = nlo[ch[c]], hi = nhi[ch[c]];
        lx[c] = lo.x; ly[c] = lo.y; lz[c] = lo.z;
        hx[c] = hi.x; hy[c] = hi.y; hz[c] = hi.z;
        if (cnt[c] <= kLeafMax) {
            // a subtree of at most kLeafMax leaves: depth-first, at most kLeafMax subtrees pending
            uint32_t st[kLeafMax];
            int sp = 0;
            st[sp++] = ch[c];
            uint32_t o = off;
            while (sp > 0) {
                const uint32_t x = st[--sp];
                if (x < n_tris) {
                    const uint32_t t = __float_as_uint(nlo[x].w);
#pragma unroll
                    for (int v = 0; v < 3; ++v) {
                        const float* p = pos + 3 * size_t(idx[3 * size_t(t) + v]);
                        tris[3 * size_t(o) + v] = make_float4(p[0], p[1], p[2], v == 0 ? __uint_as_float(t) : 0.f);
                    }
                    ++o;
                } else if (sp + 2 <= int(kLeafMax)) {
                    st[sp++] = __float_as_uint(nhi[x].w);
                    st[sp++] = __float_as_uint(nlo[x].w);
                }
            }
            ref[c] = ~int32_t((off << 3) | (cnt[c] - 1u));
        } else {
            const uint32_t slot = atomicAdd(next_count, 1u);
            next[slot] = make_uint2(ch[c], off);
            ref[c] = int32_t(level_base + m + slot);
        }
        off += cnt[c];
    }
    float4* q = nodes + 8 * size_t(level_base + k);
    q[0] = make_float4(lx[0], lx[1], lx[2], lx[3]);
    q[1] = make_float4(hx[0], hx[1], hx[2], hx[3]);
    q[2] = make_float4(ly[0], ly[1], ly[2], ly[3]);
    q[3] = make_float4(hy[0], hy[1], hy[2], hy[3]);
    q[4] = make_float4(lz[0], lz[1], lz[2], lz[3]);
    q[5] = make_float4(hz[0], hz[1], hz[2], hz[3]);
    q[6] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), __int_as_float(ref[2]), __int_as_float(ref[3]));
    q[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Depth-first layout of the collapsed tree (a node's first inner child right after it, each subtree
// contiguous), computed level by level: the collapse numbers nodes breadth-first in the order its
// threads append them, which scatters a subtree over a whole level. Inner child references are > 0
// (the root, 0, is nobody's child), leaf references < 0 and empty slots 0.
// Nodes in each subtree (itself included), from the deepest level up.
__global__ void __launch_bounds__(kBlock) subtree_size_kernel(uint32_t first, uint32_t count,
                                                              const float4* __restrict__ nodes,
                                                              uint32_t* __restrict__ size) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    const float4 cf = nodes[8 * size_t(first + k) + 6];
    const int32_t r[4] = {__float_as_int(cf.x), __float_as_int(cf.y), __float_as_int(cf.z), __float_as_int(cf.w)};
    uint32_t s = 1;
#pragma unroll
    for (int c = 0; c < 4; ++c)
        if (r[c] > 0) s += size[r[c]];
    size[first + k] = s;
}
// Depth-first positions of the next level's nodes from this level's.
__global__ void __launch_bounds__(kBlock) dfs_index_kernel(uint32_t first, uint32_t count, const float4* __restrict__ nodes,
                                                           const uint32_t* __restrict__ size, uint32_t* __restrict__ at) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    const float4 cf = nodes[8 * size_t(first + k) + 6];
    const int32_t r[4] = {__float_as_int(cf.x), __float_as_int(cf.y), __float_as_int(cf.z), __float_as_int(cf.w)};
    uint32_t q = at[first + k] + 1;
#pragma unroll
    for (int c = 0; c < 4; ++c)
        if (r[c] > 0) {
            at[r[c]] = q;
            q += size[r[c]];
        }
}
__global__ void __launch_bounds__(kBlock) permute_kernel(uint32_t n_nodes, const float4* __restrict__ nodes,
                                                         const uint32_t* __restrict__ at, float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_nodes) return;
    const float4* q = nodes + 8 * size_t(i);
    float4* o = out + 8 * size_t(at[i]);
#pragma unroll
    for (int row = 0; row < 8; ++row) {
        float4 v = q[row];
        if (row == 6) {
            const int32_t r[4] = {__float_as_int(v.x), __float_as_int(v.y), __float_as_int(v.z), __float_as_int(v.w)};
            v = make_float4(__int_as_float(r[0] > 0 ? int32_t(at[r[0]]) : r[0]),
                            __int_as_float(r[1] > 0 ? int32_t(at[r[1]]) : r[1]),
                            __int_as_float(r[2] > 0 ? int32_t(at[r[2]]) : r[2]),
                            __int_as_float(r[3] > 0 ? int32_t(at[r[3]]) : r[3]));
        }
        o[row] = v;
    }
}

// The top of the binary tree over the last clusters, built on the host by a full-sweep SAH (the
// clusters' centroids sorted along each axis, ties by place; a median split once a branch is
// deeper than kTopDepth): greedy clustering makes its last merges between large clusters, where a
// top-down split sees the whole set. New nodes take ids from `next` on; returns the root.
struct TopCluster {
    float lo[3], hi[3];
    uint32_t node, cnt;
};
constexpr int kTopDepth = 24;

float top_area(const float* lo, const float* hi) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

uint32_t build_top(std::vector<TopCluster>& c, size_t b, size_t e, int depth, uint32_t& next, std::vector<float4>& olo,
                   std::vector<float4>& ohi, std::vector<uint32_t>& ocnt) {
    if (e - b == 1) return c[b].node;
    const size_t n = e - b;
    float best = __builtin_inff();
    int axis = -1;
    size_t split = 0;
    std::vector<float> rarea(n);
    for (int a = 0; a < 3 && depth < kTopDepth; ++a) {
        std::stable_sort(c.begin() + b, c.begin() + e, [a](const TopCluster& x, const TopCluster& y) {
            return x.lo[a] + x.hi[a] < y.lo[a] + y.hi[a];
        });
        float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
        float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
        uint32_t cnt = 0;
        for (size_t i = n; i-- > 1;) {
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], c[b + i].lo[k]);
                hi[k] = std::max(hi[k], c[b + i].hi[k]);
            }
            cnt += c[b + i].cnt;
            rarea[i] = top_area(lo, hi) * float(cnt);
        }
        float llo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
        float lhi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
        uint32_t lcnt = 0;
        for (size_t i = 1; i < n; ++i) {
            for (int k = 0; k < 3; ++k) {
                llo[k] = std::min(llo[k], c[b + i - 1].lo[k]);
                lhi[k] = std::max(lhi[k], c[b + i - 1].hi[k]);
            }
            lcnt += c[b + i - 1].cnt;
            const float cost = top_area(llo, lhi) * float(lcnt) + rarea[i];
            if (cost < best) {
                best = cost;
                axis = a;
                split = i;
            }
        }
    }
    if (axis < 0) {  // too deep (or no extent): the median of the current order
        split = n / 2;
    } else {
        std::stable_sort(c.begin() + b, c.begin() + e, [axis](const TopCluster& x, const TopCluster& y) {
            return x.lo[axis] + x.hi[axis] < y.lo[axis] + y.hi[axis];
        });
    }
    const uint32_t l = build_top(c, b, b + split, depth + 1, next, olo, ohi, ocnt);
    const uint32_t r = build_top(c, b + split, e, depth + 1, next, olo, ohi, ocnt);
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    uint32_t cnt = 0;
    for (size_t i = b; i < e; ++i) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], c[i].lo[k]);
            hi[k] = std::max(hi[k], c[i].hi[k]);
        }
        cnt += c[i].cnt;
    }
    const uint32_t id = next++;
    float lw, rw;
    std::memcpy(&lw, &l, 4);
    std::memcpy(&rw, &r, 4);
    olo.push_back(make_float4(lo[0], lo[1], lo[2], lw));
    ohi.push_back(make_float4(hi[0], hi[1], hi[2], rw));
    ocnt.push_back(cnt);
    return id;
}

// Device allocations of one build, released together unless kept.
struct Allocs {
    std::vector<void*> ptrs;
    ~Allocs() {
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
    }
    template <typename T>
    T* get(size_t count) {
        void* p = nullptr;
        if (hipMalloc(&p, count * sizeof(T) > 0 ? count * sizeof(T) : 16) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    void keep(void* p) {
        for (void*& q : ptrs)
            if (q == p) q = nullptr;
    }
};

int hip_fail(const char* what) {
    (void)hipGetLastError();
    return set_error(PLT_EHIP, std::string("device BVH build: ") + what + " failed");
}

inline unsigned grid_of(uint32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

int build_bvh_device(const DeviceSceneBuffers& d, hipStream_t st, DeviceBvh& out) {
    const uint32_t n = d.n_tris;
    if (n == 0) return set_error(PLT_EINVAL, "device BVH build: the scene has no triangles");
    // leaf references hold (first << 3) | (count - 1) in 31 bits
    if (n >= (1u << 28)) return set_error(PLT_EINVAL, "device BVH build: more than 2^28 triangles");
    const auto t0 = std::chrono::steady_clock::now();
    Allocs A;
    const size_t nn2 = 2 * size_t(n);
    const uint32_t box_blocks = grid_of(n) < kBoxBlocks ? grid_of(n) : kBoxBlocks;
    uint32_t* words = A.get<uint32_t>(8);  // [7]: the next level's width
    float* cb = A.get<float>(6);
    float* partial = A.get<float>(6 * size_t(box_blocks));
    float4* tlo = A.get<float4>(n);  // triangle boxes, later the second cluster list
    float4* thi = A.get<float4>(n);
    uint64_t* bufA = A.get<uint64_t>(n);  // keys, then flags, then collapse items
    uint64_t* bufB = A.get<uint64_t>(n);  // sorted keys, then scanned flags, then next items
    uint32_t* vals = A.get<uint32_t>(n);
    uint32_t* ids = A.get<uint32_t>(n);
    float4* nlo = A.get<float4>(nn2);
    float4* nhi = A.get<float4>(nn2);
    uint32_t* ncnt = A.get<uint32_t>(nn2);
    float4* clo = A.get<float4>(n);
    float4* chi = A.get<float4>(n);
    uint32_t* nn = A.get<uint32_t>(n);
    float4* nodes = A.get<float4>(8 * size_t(n));
    float4* tris = A.get<float4>(3 * size_t(n));
    if (!words || !cb || !partial || !tlo || !thi || !bufA || !bufB || !vals || !ids || !nlo || !nhi || !ncnt || !clo || !chi || !nn || !nodes || !tris)
        return set_error(PLT_ENOMEM, "device BVH build: hipMalloc failed");
    size_t sort_bytes = 0, scan_bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, sort_bytes, bufA, bufB, vals, ids, n, 0, 63, st) != hipSuccess ||
        rocprim::inclusive_scan(nullptr, scan_bytes, bufA, bufB, size_t(n), rocprim::plus<uint64_t>(), st) != hipSuccess)
        return hip_fail("temporary storage query");
    void* tmp = A.get<uint8_t>(sort_bytes > scan_bytes ? sort_bytes : scan_bytes);
    if (!tmp) return set_error(PLT_ENOMEM, "device BVH build: hipMalloc failed");
    const size_t tmp_bytes = sort_bytes > scan_bytes ? sort_bytes : scan_bytes;

    hipLaunchKernelGGL(tri_box_kernel, dim3(box_blocks), dim3(kBlock), 0, st, n, d.positions, d.indices, tlo, thi, partial);
    hipLaunchKernelGGL(bounds_kernel, dim3(1), dim3(kBlock), 0, st, box_blocks, partial, cb);
    hipLaunchKernelGGL(morton_kernel, dim3(grid_of(n)), dim3(kBlock), 0, st, n, tlo, thi, cb, bufA, vals);
    size_t sb = tmp_bytes;
    if (rocprim::radix_sort_pairs(tmp, sb, bufA, bufB, vals, ids, n, 0, 63, st) != hipSuccess) return hip_fail("radix sort");
    hipLaunchKernelGGL(leaf_init_kernel, dim3(grid_of(n)), dim3(kBlock), 0, st, n, ids, tlo, thi, nlo, nhi, ncnt, clo, chi);
    if (hipGetLastError() != hipSuccess) return hip_fail("launch");

    // clustering
    uint32_t m = n, node_base = n;
    float4 *c_lo = clo, *c_hi = chi, *c_lo2 = tlo, *c_hi2 = thi;
    uint64_t* flags = bufA;
    uint64_t* scan = bufB;
        int iterations = 0;
    while (m > kTopClusters) {
        hipLaunchKernelGGL(nn_kernel, dim3(grid_of(m)), dim3(kBlock), 0, st, m, c_lo, c_hi, nn);
        hipLaunchKernelGGL(flag_kernel, dim3(grid_of(m)), dim3(kBlock), 0, st, m, nn, flags);
        size_t b = tmp_bytes;
        if (rocprim::inclusive_scan(tmp, b, flags, scan, size_t(m), rocprim::plus<uint64_t>(), st) != hipSuccess)
            return hip_fail("scan");
        hipLaunchKernelGGL(merge_kernel, dim3(grid_of(m)), dim3(kBlock), 0, st, m, c_lo, c_hi, nn, scan, node_base, nlo, nhi,
                           ncnt, c_lo2, c_hi2);
        uint64_t tot = 0;
        if (hipMemcpyAsync(&tot, scan + (m - 1), 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return hip_fail("clustering");
        const uint32_t kept = uint32_t(tot >> 32), made = uint32_t(tot);
        // the globally closest pair under (area, lower place) is always mutual, so every
        // iteration merges at least one pair
        if (made == 0 || kept + made != m) return set_error(PLT_EHIP, "device BVH build: clustering made no progress");
        if (iterations == kMaxIterations)
            return set_error(PLT_EINVAL, "device BVH build: the clustering did not converge in " +
                                             std::to_string(kMaxIterations) + " iterations (chain-like geometry)");
        m = kept;
        node_base += made;
        std::swap(c_lo, c_lo2);
        std::swap(c_hi, c_hi2);
        ++iterations;
    }
    uint32_t root = 0;
    {
        std::vector<float4> hlo(m), hhi(m);
        if (hipMemcpyAsync(hlo.data(), c_lo, m * sizeof(float4), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(hhi.data(), c_hi, m * sizeof(float4), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return hip_fail("top clusters");
        std::vector<TopCluster> top(m);
        for (uint32_t i = 0; i < m; ++i) {
            uint32_t node, cnt;
            std::memcpy(&node, &hlo[i].w, 4);
            std::memcpy(&cnt, &hhi[i].w, 4);
            top[i] = TopCluster{{hlo[i].x, hlo[i].y, hlo[i].z}, {hhi[i].x, hhi[i].y, hhi[i].z}, node, cnt};
        }
        std::vector<float4> olo, ohi;
        std::vector<uint32_t> ocnt;
        uint32_t next_id = node_base;
        root = build_top(top, 0, m, 0, next_id, olo, ohi, ocnt);
        if (!olo.empty() &&
            (hipMemcpyAsync(nlo + node_base, olo.data(), olo.size() * sizeof(float4), hipMemcpyHostToDevice, st) != hipSuccess ||
             hipMemcpyAsync(nhi + node_base, ohi.data(), ohi.size() * sizeof(float4), hipMemcpyHostToDevice, st) != hipSuccess ||
             hipMemcpyAsync(ncnt + node_base, ocnt.data(), ocnt.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
             hipStreamSynchronize(st) != hipSuccess))
            return hip_fail("top nodes");
    }
    if (root >= 2 * n - 1) return set_error(PLT_EHIP, "device BVH build: bad root");

    // collapse, level by level
    uint2* items = reinterpret_cast<uint2*>(bufA);
    uint2* next = reinterpret_cast<uint2*>(bufB);
    uint32_t* next_count = words + 7;
    const uint2 first = make_uint2(root, 0u);
    if (hipMemcpyAsync(items, &first, sizeof(first), hipMemcpyHostToDevice, st) != hipSuccess) return hip_fail("collapse");
    uint32_t level_base = 0, width = 1;
    int levels = 0;
    std::vector<uint32_t> level_first, level_width;
    while (width > 0) {
        if (levels == kMaxLevels)
            return set_error(PLT_EINVAL, "device BVH build: the tree needs more than " + std::to_string(kMaxLevels) +
                                             " levels of 4-wide nodes, more than the traversal stack holds");
        if (hipMemsetAsync(next_count, 0, 4, st) != hipSuccess) return hip_fail("collapse");
        hipLaunchKernelGGL(collapse_kernel, dim3(grid_of(width)), dim3(kBlock), 0, st, width, items, level_base, next,
                           next_count, n, nlo, nhi, ncnt, d.positions, d.indices, nodes, tris);
        uint32_t w2 = 0;
        if (hipGetLastError() != hipSuccess || hipMemcpyAsync(&w2, next_count, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return hip_fail("collapse");
        level_first.push_back(level_base);
        level_width.push_back(width);
        level_base += width;
        width = w2;
        std::swap(items, next);
        ++levels;
    }
    // depth-first layout, into an allocation of the exact size (the clustering's counts and
    // neighbour lists are free again: sizes in ncnt, positions in nn, both at least n_nodes long)
    float4* exact = A.get<float4>(8 * size_t(level_base));
    if (!exact) return set_error(PLT_ENOMEM, "device BVH build: hipMalloc failed");
    uint32_t* size = ncnt;
    uint32_t* at = nn;
    for (int l = levels - 1; l >= 0; --l)
        hipLaunchKernelGGL(subtree_size_kernel, dim3(grid_of(level_width[l])), dim3(kBlock), 0, st, level_first[l],
                           level_width[l], nodes, size);
    if (hipMemsetAsync(at, 0, 4, st) != hipSuccess) return hip_fail("layout");
    for (int l = 0; l + 1 < levels; ++l)
        hipLaunchKernelGGL(dfs_index_kernel, dim3(grid_of(level_width[l])), dim3(kBlock), 0, st, level_first[l],
                           level_width[l], nodes, size, at);
    hipLaunchKernelGGL(permute_kernel, dim3(grid_of(level_base)), dim3(kBlock), 0, st, level_base, nodes, at, exact);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) return hip_fail("layout");
    A.keep(exact);
    A.keep(tris);
    out.nodes = exact;
    out.tris = tris;
    out.n_nodes = level_base;
    out.max_depth = levels;
    out.iterations = iterations;
    out.ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PLT_OK;
}

}  // namespace plt
