// Path-tracer launch parameters and bounce-record layout shared by host and device code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/plt_abi.h"
#include "traverse.hpp"

namespace plt {

// Bounce record (the reference's 40 B BounceData, Path.slang:113-142, with the hit packed as
// triangle id + barycentrics): 9 words per (bounce, path), stored as planes
// rec[(bounce * kRecWords + field) * n_paths + path] so every field access of a wave is one
// contiguous 256 B segment.
enum RecField { kRecTri = 0, kRecB1, kRecB2, kRecWiX, kRecWiY, kRecLobe, kRecFlags, kRecContrib, kRecLastPdf, kRecWords };

// RayFlags (Path.slang:47-109)
enum : uint32_t {
    kFlagLastDelta = 0x1, kFlagDispersive = 0x2, kFlagOnlyDelta = 0x4, kFlagMiss = 0x10, kFlagTerminated = 0x20
};

enum : uint32_t {  // option bits
    kOptNEE = 0x1, kOptMIS = 0x2, kOptRR = 0x4, kOptHwssMis = 0x8, kOptISEmitters = 0x10, kOptDirectLights = 0x20,
    kOptEmissiveLights = 0x40, kOptMNEE = 0x80, kOptUniformEmissive = 0x100, kOptAnalyticLights = 0x200,
    kOptEnvLight = 0x400
};

// Largest number of transmissive occluders a manifold walk may pass (MNEE_MAX_MS_OCCLUDERS).
constexpr uint32_t kMaxMsOccluders = 8;

struct PtScene {
    TraceScene tsc;
    const float* positions;
    const float* normals;
    const uint32_t* indices;
    const uint32_t* tri_mesh;
    const uint32_t* mesh_material;
    const plt_material* materials;
    const plt_spectral_profile* profiles;
    uint32_t n_profiles;
    // camera (Camera.cpp:187-193); rows of the view rotation (the right-handed look-at view matrix's
    // upper 3x3); polarizer strength and angle (CameraData.slang:67-68)
    float cam_pos[3], cam_u[3], cam_v[3], cam_w[3], jitter[2];
    float view_rot[9];
    float polarizer, polarizer_theta;
    // emissive triangles: 4 x float4 per light (p0+area, p1+material, p2, normal), alias table,
    // per-light selection pdf, triangle -> light index
    const float4* lights;
    const uint2* alias;
    const float* light_pdf;
    const uint32_t* tri_light;
    uint32_t n_lights;
    // lights with non-zero flux, for the uniform emissive sampler
    const uint32_t* active_lights;
    uint32_t n_active;
    // analytic lights (point, directional, distant, rect, disc, sphere)
    const plt_light* analytic;
    uint32_t n_analytic;
    // environment map: RGBA32F lat-long image, rotation and inverse (row-major 3x3), radiance
    // scale, importance map mip chain (level k at env_mip[k], (512 >> k)^2 texels)
    const float4* env_rgba;
    uint32_t env_w, env_h;
    float env_m[9], env_mi[9], env_scale[3];
    const float* env_imp;
    uint32_t env_base_mip;
    uint32_t env_mip[12];
};

struct PtFrame {
    uint32_t width, height, frame, max_bounces, hwss, options;
    uint32_t path_base, n_paths;  // paths [path_base, path_base + n_paths) of this launch (row-major pixels)
    float sourcing_area, max_omega;
    // manifold NEE: occluder limit, Newton iteration limit, constraint threshold (Manifolds.slangh:42-54)
    uint32_t mnee_max_occluders, mnee_max_iterations;
    float mnee_threshold;
    // sample generator (SampleGeneratorType.slangh:32-36): 1 = uniform (xoshiro128**), 0 = tiny (LCG)
    uint32_t sample_generator;
    // NEEUsePerTileSG: the NEE light-type draws of the per-tile generator, tile_draws[tile * (max_bounces
    // + 1) + k] = the k-th draw of tile `tile` (row-major tiles of tile_size^2 pixels); NULL = off
    const float* tile_draws;
    uint32_t tile_size, tiles_x;
    uint32_t* records;
    float4* color;
    float4* albedo;
    float4* normal;
};

// Wavefront sample pass: per-path state (SoA, n_paths entries each), two path queues of
// n_paths entries, the material-sorted queue and its keys, and counters.
constexpr uint32_t kWfKeys = 16;  // material types 0..14, misses last
enum : uint32_t { kWfCount0 = 0, kWfHist = 2, kWfCursor = kWfHist + kWfKeys, kWfCounters = kWfCursor + kWfKeys };
struct WfBuffers {
    uint4* rng;
    float4* org;   // origin, hero wavelength
    float4* dir;   // direction, path weight
    float4* pdfs;  // pdf, thp, last non-delta pdf, flags (bits)
    uint32_t* len;
    float4* hit;   // triangle (bits), b1, b2, t
    uint32_t* queue;   // 2 x n_paths
    uint32_t* sorted;  // n_paths
    uint32_t* keys;    // n_paths
    uint32_t* counters;
};

// Launch parameters of the sample pass (the autotuner's search space): the megakernel's block size
// and the wavefront schedule's persistent grid, in 64-thread blocks per CU.
struct LaunchConfig {
    uint32_t sample_block = 64;
    uint32_t wf_blocks_per_cu = 16;
};

int launch_sample_pass(const PtScene& sc, const PtFrame& fr, const LaunchConfig& lc, hipStream_t st);
int launch_sample_pass_wavefront(const PtScene& sc, const PtFrame& fr, const WfBuffers& w, const LaunchConfig& lc, hipStream_t st);
int launch_solve_pass(const PtScene& sc, const PtFrame& fr, hipStream_t st);
int launch_tile_draws(const PtFrame& fr, uint32_t tiles, float* draws, hipStream_t st);

}  // namespace plt
