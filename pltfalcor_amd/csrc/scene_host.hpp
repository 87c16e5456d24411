// Host-side scene representation shared by the scene builder, the BVH builder and the device
// upload. Device layouts are documented next to the GPU structs in scene_device.hpp.
#pragma once
#include <cstdint>
#include <vector>
#include "../../include/plt_abi.h"

namespace plt {

struct Aabb {
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f};
    float hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    void grow(const float* p) {
        for (int a = 0; a < 3; ++a) { lo[a] = lo[a] < p[a] ? lo[a] : p[a]; hi[a] = hi[a] > p[a] ? hi[a] : p[a]; }
    }
    void grow(const Aabb& b) {
        for (int a = 0; a < 3; ++a) { lo[a] = lo[a] < b.lo[a] ? lo[a] : b.lo[a]; hi[a] = hi[a] > b.hi[a] ? hi[a] : b.hi[a]; }
    }
    float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return (dx < 0.f || dy < 0.f || dz < 0.f) ? 0.f : 2.f * (dx * dy + dy * dz + dz * dx);
    }
};

// BVH4 node as stored in HBM (128 B, two cache lines fetched together): the four children's boxes
// as SoA rows (lo x, hi x, lo y, hi y, lo z, hi z; one float per child) and the child references.
// Child reference >= 0: inner node index; < 0: leaf, ~ref = (first << 3) | (count - 1). An unused
// slot has the box [+inf, +inf] on every axis, which no ray enters.
struct BvhNode {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t child[4];
    int32_t pad[4];
};
static_assert(sizeof(BvhNode) == 128, "BvhNode must be 128 B");

// Inner levels of the binary SAH tree at most (root = 1): the builder splits at the object median
// wherever a SAH split would need more, so every tree meets it. Collapsing to four children never
// deepens the tree, so the BVH4 has at most as many levels.
constexpr int kMaxBvh2Depth = 32;

struct HostBvh {
    std::vector<BvhNode> nodes;
    std::vector<uint32_t> tri_order;  // triangle ids in leaf order
    int max_depth = 0;                // BVH4 levels (root = 1)
    float build_ms = 0.f;             // wall-clock time of the build
};

// Builds a binned-SAH BVH over triangles (positions as 3 floats per vertex, indices 3 per tri).
HostBvh build_bvh(const std::vector<float>& pos, const std::vector<uint32_t>& idx, int max_leaf = 4);

// Environment map light: RGBA32F lat-long image, local-to-world rotation (row-major 3x3) and its
// inverse, radiance scale (intensity x tint), and the importance map's mip chain (level k of
// (512 >> k)^2 texels at mip_offset[k]).
constexpr uint32_t kEnvImportanceDim = 512;
struct HostEnvMap {
    uint32_t width = 0, height = 0;
    std::vector<float> rgba;
    float transform[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float inv_transform[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float scale[3] = {1, 1, 1};
    std::vector<float> importance;
    std::vector<uint32_t> mip_offset;
    uint32_t base_mip = 0;
};

struct HostScene {
    std::vector<float> positions;       // 3 per vertex
    std::vector<float> normals;         // 3 per vertex (binary16-rounded)
    std::vector<uint32_t> indices;      // 3 per triangle
    std::vector<uint32_t> tri_mesh;
    std::vector<uint32_t> mesh_first_tri;
    std::vector<uint32_t> mesh_material;
    std::vector<plt_material> materials;
    std::vector<plt_spectral_profile> profiles;
    plt_camera camera{};
    HostBvh bvh;
    std::vector<plt_light> lights;  // analytic lights
    HostEnvMap env;                 // environment map (width 0: none)
};

// Procedural workload scene (plt_procedural_scene_desc).
int build_procedural_scene(const plt_procedural_scene_desc& d, HostScene& out);

// binary16 round-trip helpers for host-side parameter storage.
uint16_t f32_to_h(float f);
float h_to_f32(uint16_t h);

}  // namespace plt
