// Scene objects of the C ABI: procedural construction, upload of the BVH / triangle / shading
// arrays to HBM, host views for callers, and the batched ray-cast entry point.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include "plt_internal.hpp"
#include "scene_host.hpp"
#include "scene_device.hpp"
#include "traverse.hpp"

using plt::set_error;

namespace plt {
namespace {

constexpr int kTraceBlock = 128;

template <typename T>
int upload(const std::vector<T>& v, T** dptr, std::vector<void*>& owned) {
    *dptr = nullptr;
    if (v.empty()) return PLT_OK;
    void* p = nullptr;
    if (hipMalloc(&p, v.size() * sizeof(T)) != hipSuccess) return set_error(PLT_ENOMEM, "scene upload: hipMalloc failed");
    owned.push_back(p);
    if (hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return set_error(PLT_EHIP, "scene upload: hipMemcpy failed");
    *dptr = static_cast<T*>(p);
    return PLT_OK;
}

template <bool kAny>
__global__ void __launch_bounds__(kTraceBlock) trace_kernel(TraceScene sc, uint32_t n, const float* __restrict__ rays,
                                                            uint32_t* __restrict__ hits) {
    __shared__ int stack[kTraversalStackDepth * kTraceBlock];
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    const size_t N = n;
    Ray r;
    r.ox = rays[0 * N + i]; r.oy = rays[1 * N + i]; r.oz = rays[2 * N + i];
    r.dx = rays[3 * N + i]; r.dy = rays[4 * N + i]; r.dz = rays[5 * N + i];
    r.tmin = rays[6 * N + i]; r.tmax = rays[7 * N + i];
    const Hit h = traverse<kAny, kTraversalStackDepth>(sc, r, stack + threadIdx.x, kTraceBlock);
    if (kAny) {
        hits[i] = h.tri != kNoHit ? 1u : 0u;
    } else {
        hits[0 * N + i] = h.tri;
        hits[1 * N + i] = __float_as_uint(h.b1);
        hits[2 * N + i] = __float_as_uint(h.b2);
        hits[3 * N + i] = __float_as_uint(h.t);
    }
}
}  // namespace

int upload_scene(const HostScene& hs, DeviceSceneBuffers& d) {
    // a BVH4 of L levels pushes at most 3 (L - 1) entries; the builder's depth bound keeps L <= 32
    if (3 * (hs.bvh.max_depth - 1) > kTraversalStackDepth + kTraversalSpill)
        return set_error(PLT_EINVAL, "BVH depth " + std::to_string(hs.bvh.max_depth) + " exceeds the traversal stack (" +
                                         std::to_string(kTraversalStackDepth + kTraversalSpill) + " entries)");
    // Nodes as float4 x 8; triangles in leaf order as (v0, id), (v1, 0), (v2, 0).
    std::vector<float4> nodes(hs.bvh.nodes.size() * 8);
    std::memcpy(nodes.data(), hs.bvh.nodes.data(), hs.bvh.nodes.size() * sizeof(BvhNode));
    const size_t nt = hs.bvh.tri_order.size();
    std::vector<float4> tris(3 * nt);
    for (size_t i = 0; i < nt; ++i) {
        const uint32_t t = hs.bvh.tri_order[i];
        for (int v = 0; v < 3; ++v) {
            const float* p = &hs.positions[3 * size_t(hs.indices[3 * size_t(t) + v])];
            float w = 0.f;
            if (v == 0) std::memcpy(&w, &t, 4);
            tris[3 * i + v] = make_float4(p[0], p[1], p[2], w);
        }
    }
    int rc;
    if ((rc = upload(nodes, &d.nodes, d.owned)) < 0) return rc;
    if ((rc = upload(tris, &d.tris, d.owned)) < 0) return rc;
    if ((rc = upload(hs.positions, &d.positions, d.owned)) < 0) return rc;
    if ((rc = upload(hs.normals, &d.normals, d.owned)) < 0) return rc;
    if ((rc = upload(hs.indices, &d.indices, d.owned)) < 0) return rc;
    if ((rc = upload(hs.tri_mesh, &d.tri_mesh, d.owned)) < 0) return rc;
    if ((rc = upload(hs.mesh_material, &d.mesh_material, d.owned)) < 0) return rc;
    if ((rc = upload(hs.profiles, &d.profiles, d.owned)) < 0) return rc;
    d.n_tris = uint32_t(nt);
    d.n_nodes = uint32_t(hs.bvh.nodes.size());
    return PLT_OK;
}

void free_scene(DeviceSceneBuffers& d) {
    for (void* p : d.owned) (void)hipFree(p);
    d.owned.clear();
}

}  // namespace plt

extern "C" {

int plt_scene_create_procedural(plt_context* ctx, const plt_procedural_scene_desc* desc, plt_scene** out) {
    if (!desc || !out) return set_error(PLT_EINVAL, "plt_scene_create_procedural: bad arguments");
    plt_scene* s = new (std::nothrow) plt_scene();
    if (!s) return set_error(PLT_ENOMEM, "plt_scene_create_procedural: out of host memory");
    s->ctx = ctx;
    if (plt::build_procedural_scene(*desc, s->host) != PLT_OK) {
        delete s;
        return set_error(PLT_EINVAL, "plt_scene_create_procedural: invalid description");
    }
    if (!ctx) {  // host-only scene: arrays and BVH, nothing on the device
        *out = s;
        return PLT_OK;
    }
    const int rc = plt::upload_scene(s->host, s->dev);
    if (rc < 0) {
        plt::free_scene(s->dev);
        delete s;
        return rc;
    }
    *out = s;
    return PLT_OK;
}

int plt_scene_create(plt_context* ctx, const plt_scene_desc* d, plt_scene** out) {
    if (!d || !out) return set_error(PLT_EINVAL, "plt_scene_create: bad arguments");
    if (d->n_vertices == 0 || d->n_triangles == 0 || d->n_meshes == 0 || d->n_materials == 0 || !d->positions ||
        !d->normals || !d->indices || !d->tri_mesh || !d->mesh_material || !d->materials || (d->n_profiles && !d->profiles))
        return set_error(PLT_EINVAL, "plt_scene_create: empty scene or missing array");
    for (size_t i = 0; i < size_t(d->n_triangles) * 3; ++i)
        if (d->indices[i] >= d->n_vertices) return set_error(PLT_EINVAL, "plt_scene_create: vertex index out of range");
    for (uint32_t t = 0; t < d->n_triangles; ++t)
        if (d->tri_mesh[t] >= d->n_meshes) return set_error(PLT_EINVAL, "plt_scene_create: mesh id out of range");
    for (uint32_t m = 0; m < d->n_meshes; ++m)
        if (d->mesh_material[m] >= d->n_materials) return set_error(PLT_EINVAL, "plt_scene_create: material id out of range");
    for (uint32_t m = 0; m < d->n_materials; ++m) {
        const plt_material& mt = d->materials[m];
        if (mt.type < PLT_MAT_DIFFUSE || mt.type > PLT_MAT_MULTILAYER_STACK)
            return set_error(PLT_EINVAL, "plt_scene_create: unknown material type (the reference's MaterialType values 2-11)");
        // radial gratings take their direction from texture coordinates (DiffractionGrating.slang:52-57)
        if ((mt.type == PLT_MAT_GRATED_CONDUCTOR || mt.type == PLT_MAT_GRATED_OPAQUE_DIELECTRIC) && ((mt.data1[3] >> 8) & 0x10u))
            return set_error(PLT_ENOTIMPL, "plt_scene_create: radial diffraction gratings need texture coordinates, which are not built");
        for (uint32_t id : {mt.emission_profile, mt.ior_n_profile, mt.ior_k_profile})
            if (id != PLT_NO_PROFILE && id >= d->n_profiles) return set_error(PLT_EINVAL, "plt_scene_create: profile id out of range");
    }
    for (uint32_t p = 0; p < d->n_profiles; ++p)
        if (d->profiles[p].bins < 1 || d->profiles[p].bins > PLT_SPECTRAL_PROFILE_MAX_BINS)
            return set_error(PLT_EINVAL, "plt_scene_create: profile bin count out of range");
    plt_scene* s = new (std::nothrow) plt_scene();
    if (!s) return set_error(PLT_ENOMEM, "plt_scene_create: out of host memory");
    s->ctx = ctx;
    plt::HostScene& h = s->host;
    h.positions.assign(d->positions, d->positions + 3 * size_t(d->n_vertices));
    h.normals.resize(3 * size_t(d->n_vertices));
    for (size_t i = 0; i < h.normals.size(); ++i) h.normals[i] = plt::h_to_f32(plt::f32_to_h(d->normals[i]));
    h.indices.assign(d->indices, d->indices + 3 * size_t(d->n_triangles));
    h.tri_mesh.assign(d->tri_mesh, d->tri_mesh + d->n_triangles);
    h.mesh_first_tri.assign(d->n_meshes, 0xFFFFFFFFu);
    for (uint32_t t = d->n_triangles; t-- > 0;) h.mesh_first_tri[d->tri_mesh[t]] = t;
    h.mesh_material.assign(d->mesh_material, d->mesh_material + d->n_meshes);
    h.materials.assign(d->materials, d->materials + d->n_materials);
    if (d->n_profiles) h.profiles.assign(d->profiles, d->profiles + d->n_profiles);
    h.camera = d->camera;
    h.bvh = plt::build_bvh(h.positions, h.indices, 4);
    if (!ctx) {
        *out = s;
        return PLT_OK;
    }
    const int rc = plt::upload_scene(s->host, s->dev);
    if (rc < 0) {
        plt::free_scene(s->dev);
        delete s;
        return rc;
    }
    *out = s;
    return PLT_OK;
}

int plt_scene_set_lights(plt_scene* scene, const plt_light* lights, uint32_t count) {
    if (!scene || (count && !lights)) return set_error(PLT_EINVAL, "plt_scene_set_lights: bad arguments");
    for (uint32_t i = 0; i < count; ++i) {
        if (lights[i].type > PLT_LIGHT_SPHERE) return set_error(PLT_EINVAL, "plt_scene_set_lights: unknown light type");
        if (lights[i].intensity_profile >= scene->host.profiles.size())
            return set_error(PLT_EINVAL, "plt_scene_set_lights: intensity profile id out of range");
    }
    scene->host.lights.assign(lights, lights + count);
    return PLT_OK;
}

int plt_scene_set_camera(plt_scene* scene, const plt_camera* camera) {
    if (!scene || !camera) return set_error(PLT_EINVAL, "plt_scene_set_camera: bad arguments");
    const plt_camera& c = *camera;
    const float f[3] = {c.target[0] - c.pos[0], c.target[1] - c.pos[1], c.target[2] - c.pos[2]};
    if (!(f[0] * f[0] + f[1] * f[1] + f[2] * f[2] > 0.f) || !(c.fov_y > 0.f && c.fov_y < 3.1415926f) || !(c.aspect > 0.f) ||
        !(c.focal_distance > 0.f) || !(c.polarizer >= 0.f && c.polarizer <= 1.f) || !std::isfinite(c.polarizer_theta))
        return set_error(PLT_EINVAL, "plt_scene_set_camera: degenerate camera (target at the position, fov outside (0, pi), "
                                     "aspect or focal distance not positive, polarizer outside [0, 1])");
    scene->host.camera = c;
    return PLT_OK;
}

int plt_scene_destroy(plt_scene* scene) {
    if (!scene) return PLT_OK;
    plt::free_scene(scene->dev);
    delete scene;
    return PLT_OK;
}

int plt_scene_get_arrays(const plt_scene* scene, plt_scene_arrays* out) {
    if (!scene || !out) return set_error(PLT_EINVAL, "plt_scene_get_arrays: bad arguments");
    const plt::HostScene& h = scene->host;
    out->n_vertices = uint32_t(h.positions.size() / 3);
    out->n_triangles = uint32_t(h.indices.size() / 3);
    out->n_meshes = uint32_t(h.mesh_first_tri.size());
    out->n_materials = uint32_t(h.materials.size());
    out->n_profiles = uint32_t(h.profiles.size());
    out->n_bvh_nodes = uint32_t(h.bvh.nodes.size());
    out->bvh_max_depth = uint32_t(h.bvh.max_depth);
    out->positions = h.positions.data();
    out->normals = h.normals.data();
    out->indices = h.indices.data();
    out->tri_mesh = h.tri_mesh.data();
    out->mesh_first_tri = h.mesh_first_tri.data();
    out->mesh_material = h.mesh_material.data();
    out->materials = h.materials.data();
    out->profiles = h.profiles.data();
    out->camera = h.camera;
    return PLT_OK;
}

int plt_scene_trace(plt_scene* scene, uint32_t n, const float* rays, uint32_t* hits, int mode, plt_stream stream) {
    if (!scene || (n && (!rays || !hits)) || (mode != 0 && mode != 1))
        return set_error(PLT_EINVAL, "plt_scene_trace: bad arguments");
    if (!scene->ctx) return set_error(PLT_ESTATE, "plt_scene_trace: host-only scene (created without a context)");
    if (n == 0) return PLT_OK;
    plt::TraceScene sc{scene->dev.nodes, scene->dev.tris};
    const dim3 grid((n + plt::kTraceBlock - 1) / plt::kTraceBlock), block(plt::kTraceBlock);
    if (mode == 0)
        hipLaunchKernelGGL(plt::trace_kernel<false>, grid, block, 0, static_cast<hipStream_t>(stream), sc, n, rays, hits);
    else
        hipLaunchKernelGGL(plt::trace_kernel<true>, grid, block, 0, static_cast<hipStream_t>(stream), sc, n, rays, hits);
    return plt::check_launch("trace_kernel");
}

}  // extern "C"
