// Scene objects of the C ABI: procedural construction, upload of the BVH / triangle / shading
// arrays to HBM, host views for callers, and the batched ray-cast entry point.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <utility>
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

// Replaces the owned allocation `old` (if any) by `now` in the scene's list.
void replace_owned(DeviceSceneBuffers& d, void* old, void* now) {
    for (void*& p : d.owned)
        if (old && p == old) {
            (void)hipFree(p);
            p = now;
            return;
        }
    d.owned.push_back(now);
}

// Uploads the host tree: nodes as float4 x 8 and triangles in leaf order as (v0, id), (v1, 0),
// (v2, 0), replacing the tree the scene holds.
int upload_bvh(const HostScene& hs, DeviceSceneBuffers& d) {
    // a BVH4 of L levels pushes at most 3 (L - 1) entries; the builder's depth bound keeps L <= 32
    if (3 * (hs.bvh.max_depth - 1) > kTraversalStackDepth + kTraversalSpill)
        return set_error(PLT_EINVAL, "BVH depth " + std::to_string(hs.bvh.max_depth) + " exceeds the traversal stack (" +
                                         std::to_string(kTraversalStackDepth + kTraversalSpill) + " entries)");
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
    std::vector<void*> fresh;
    float4 *dn = nullptr, *dt = nullptr;
    int rc;
    if ((rc = upload(nodes, &dn, fresh)) < 0 || (rc = upload(tris, &dt, fresh)) < 0) {
        for (void* p : fresh) (void)hipFree(p);
        return rc;
    }
    replace_owned(d, d.nodes, dn);
    replace_owned(d, d.tris, dt);
    d.nodes = dn;
    d.tris = dt;
    d.n_tris = uint32_t(nt);
    d.n_nodes = uint32_t(hs.bvh.nodes.size());
    return PLT_OK;
}

// Uploads the geometry and shading arrays (everything but the tree).
int upload_arrays(const HostScene& hs, DeviceSceneBuffers& d) {
    int rc;
    if ((rc = upload(hs.positions, &d.positions, d.owned)) < 0) return rc;
    if ((rc = upload(hs.normals, &d.normals, d.owned)) < 0) return rc;
    if ((rc = upload(hs.indices, &d.indices, d.owned)) < 0) return rc;
    if ((rc = upload(hs.tri_mesh, &d.tri_mesh, d.owned)) < 0) return rc;
    if ((rc = upload(hs.mesh_material, &d.mesh_material, d.owned)) < 0) return rc;
    if ((rc = upload(hs.profiles, &d.profiles, d.owned)) < 0) return rc;
    d.n_tris = uint32_t(hs.indices.size() / 3);
    return PLT_OK;
}

int upload_scene(const HostScene& hs, DeviceSceneBuffers& d) {
    int rc;
    if ((rc = upload_arrays(hs, d)) < 0) return rc;
    return upload_bvh(hs, d);
}

void host_bvh_info(const HostScene& hs, plt_bvh_info& info) {
    info = plt_bvh_info{};
    info.builder = PLT_BVH_HOST_SAH;
    info.n_nodes = uint32_t(hs.bvh.nodes.size());
    info.max_depth = uint32_t(hs.bvh.max_depth);
    info.build_ms = hs.bvh.build_ms;
}

// Builds the scene's tree on the device and installs it; the scene keeps its old tree on failure.
int device_bvh(plt_scene* s, hipStream_t st) {
    DeviceBvh b;
    const int rc = build_bvh_device(s->dev, st, b);
    if (rc < 0) return rc;
    if (hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(b.nodes);
        (void)hipFree(b.tris);
        return set_error(PLT_EHIP, "device BVH build: device synchronization failed");
    }
    replace_owned(s->dev, s->dev.nodes, b.nodes);
    replace_owned(s->dev, s->dev.tris, b.tris);
    s->dev.nodes = b.nodes;
    s->dev.tris = b.tris;
    s->dev.n_nodes = b.n_nodes;
    // the host tree no longer describes the device's
    s->host.bvh = HostBvh{};
    s->bvh = plt_bvh_info{};
    s->bvh.builder = PLT_BVH_DEVICE_PLOC;
    s->bvh.n_nodes = b.n_nodes;
    s->bvh.max_depth = uint32_t(b.max_depth);
    s->bvh.iterations = uint32_t(b.iterations);
    s->bvh.build_ms = b.ms;
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
    plt::host_bvh_info(s->host, s->bvh);
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
    return plt_scene_create_ex(ctx, d, PLT_BVH_HOST_SAH, out, nullptr);
}

int plt_scene_create_ex(plt_context* ctx, const plt_scene_desc* d, int builder, plt_scene** out, plt_bvh_info* info) {
    if (!d || !out || (builder != PLT_BVH_HOST_SAH && builder != PLT_BVH_DEVICE_PLOC))
        return set_error(PLT_EINVAL, "plt_scene_create: bad arguments");
    if (builder == PLT_BVH_DEVICE_PLOC && !ctx)
        return set_error(PLT_EINVAL, "plt_scene_create: the device BVH builder needs a context");
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
    int rc = PLT_OK;
    if (builder == PLT_BVH_DEVICE_PLOC) {
        rc = plt::upload_arrays(h, s->dev);
        if (rc >= 0) rc = plt::device_bvh(s, nullptr);
        if (rc == PLT_EINVAL) {  // too deep for the traversal stack: the host builder bounds the depth
            h.bvh = plt::build_bvh(h.positions, h.indices, 4);
            plt::host_bvh_info(h, s->bvh);
            rc = plt::upload_bvh(h, s->dev);
        }
    } else {
        h.bvh = plt::build_bvh(h.positions, h.indices, 4);
        plt::host_bvh_info(h, s->bvh);
        if (ctx) rc = plt::upload_scene(h, s->dev);
    }
    if (rc < 0) {
        plt::free_scene(s->dev);
        delete s;
        return rc;
    }
    if (info) *info = s->bvh;
    *out = s;
    return PLT_OK;
}

int plt_scene_build_bvh(plt_scene* scene, int builder, plt_stream stream, plt_bvh_info* info) {
    if (!scene || (builder != PLT_BVH_HOST_SAH && builder != PLT_BVH_DEVICE_PLOC))
        return set_error(PLT_EINVAL, "plt_scene_build_bvh: bad arguments");
    plt::HostScene& h = scene->host;
    if (builder == PLT_BVH_DEVICE_PLOC) {
        if (!scene->ctx) return set_error(PLT_ESTATE, "plt_scene_build_bvh: host-only scene (created without a context)");
        const int rc = plt::device_bvh(scene, static_cast<hipStream_t>(stream));
        if (rc < 0) return rc;
    } else {
        plt::HostBvh b = plt::build_bvh(h.positions, h.indices, 4);
        if (scene->ctx) {
            if (hipDeviceSynchronize() != hipSuccess) return set_error(PLT_EHIP, "plt_scene_build_bvh: device synchronization failed");
            std::swap(h.bvh, b);
            const int rc = plt::upload_bvh(h, scene->dev);
            if (rc < 0) {
                std::swap(h.bvh, b);
                return rc;
            }
        } else {
            std::swap(h.bvh, b);
        }
        plt::host_bvh_info(h, scene->bvh);
    }
    if (info) *info = scene->bvh;
    return PLT_OK;
}

int plt_scene_get_bvh(const plt_scene* scene, plt_bvh_info* info, void* nodes, size_t nodes_bytes, void* tris,
                      size_t tris_bytes) {
    if (!scene) return set_error(PLT_EINVAL, "plt_scene_get_bvh: NULL scene");
    const size_t nb = size_t(scene->bvh.n_nodes) * sizeof(plt::BvhNode);
    const size_t nt = scene->host.indices.size() / 3;
    const size_t tb = nt * 3 * sizeof(float4);
    if ((nodes && nodes_bytes < nb) || (tris && tris_bytes < tb))
        return set_error(PLT_EINVAL, "plt_scene_get_bvh: output smaller than the tree");
    if (info) *info = scene->bvh;
    if (scene->ctx) {
        if ((nodes && hipMemcpy(nodes, scene->dev.nodes, nb, hipMemcpyDeviceToHost) != hipSuccess) ||
            (tris && hipMemcpy(tris, scene->dev.tris, tb, hipMemcpyDeviceToHost) != hipSuccess))
            return set_error(PLT_EHIP, "plt_scene_get_bvh: hipMemcpy failed");
        return PLT_OK;
    }
    // host-only scene: the host tree, in the device layout
    const plt::HostScene& h = scene->host;
    if (nodes) std::memcpy(nodes, h.bvh.nodes.data(), nb);
    if (tris) {
        float* out = static_cast<float*>(tris);
        for (size_t i = 0; i < h.bvh.tri_order.size(); ++i) {
            const uint32_t t = h.bvh.tri_order[i];
            for (int v = 0; v < 3; ++v) {
                const float* p = &h.positions[3 * size_t(h.indices[3 * size_t(t) + v])];
                float* o = out + 12 * i + 4 * v;
                o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
                o[3] = 0.f;
                if (v == 0) std::memcpy(&o[3], &t, 4);
            }
        }
    }
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
    out->n_bvh_nodes = scene->bvh.n_nodes;
    out->bvh_max_depth = scene->bvh.max_depth;
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
