// Device-resident scene buffers (HBM). Owned by plt_scene; raw pointers are passed to kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>
#include "../../include/plt_abi.h"
#include "scene_host.hpp"

namespace plt {

// Traversal stack entries kept in LDS per lane; deeper entries go to private memory (traverse.hpp).
// 32 slots x 4 B x 64 lanes = 8 KB per one-wave block, which leaves LDS room for 20 blocks per CU.
constexpr int kTraversalStackDepth = 32;

struct DeviceSceneBuffers {
    float4* nodes = nullptr;          // 8 x float4 per BVH node (BvhNode, 128 B)
    float4* tris = nullptr;           // 3 x float4 per triangle, leaf order (48 B)
    float* positions = nullptr;       // 3 floats per vertex
    float* normals = nullptr;         // 3 floats per vertex
    uint32_t* indices = nullptr;      // 3 per triangle
    uint32_t* tri_mesh = nullptr;     // mesh per triangle
    uint32_t* mesh_material = nullptr;
    plt_spectral_profile* profiles = nullptr;
    uint32_t n_tris = 0, n_nodes = 0;
    std::vector<void*> owned;
};

// A tree built on the device (bvh_device.hip): nodes and leaf-order triangles in allocations the
// caller takes over.
struct DeviceBvh {
    float4* nodes = nullptr;
    float4* tris = nullptr;
    uint32_t n_nodes = 0;
    int max_depth = 0;   // levels of 4-wide nodes (root = 1)
    int iterations = 0;  // clustering iterations
    float ms = 0.f;      // wall-clock build time
};
// Builds the BVH of the scene's device triangles (positions, indices, n_tris) on `st`; returns
// PLT_EINVAL, keeping no allocation, when the tree would be deeper than the traversal stack holds.
int build_bvh_device(const DeviceSceneBuffers& d, hipStream_t st, DeviceBvh& out);

}  // namespace plt

// Definition of the ABI's opaque scene handle.
struct plt_scene {
    plt_context* ctx = nullptr;
    plt::HostScene host;
    plt::DeviceSceneBuffers dev;
    plt_bvh_info bvh{};  // the tree the scene holds: builder, nodes, levels, build time
};
