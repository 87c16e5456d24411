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
    uint32_t path_base, n_paths;  // paths [path_base, path_base + n_paths) of this launch (pixels: path_pixel)
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
    // split solve pass: the region table of the light-sampling kernel; when set, the megakernel
    // solves only the paths whose region overflowed its event buffer (NULL = every path)
    const uint32_t* solve_overflow;
};

// The row-major pixel of a frame's path g (g = path_base + li). Paths run in 8-row bands, each band
// column by column (a column of the band's 8 rows, fewer in a last, shorter band), so the 64
// consecutive paths of a wave, and of a split-pass region, are an 8 x 8 pixel block: their camera
// rays, hits and shadow rays are coherent, which shortens the wave's traversals. Bounce records
// and the split pass's regions stay in path order; colour, albedo, normal and the random streams
// are the pixel's.
__host__ __device__ inline uint32_t path_pixel(uint32_t width, uint32_t height, uint32_t g) {
    const uint32_t bandPaths = 8u * width;
    const uint32_t band = g / bandPaths, r = g - band * bandPaths;
    const uint32_t rows = height - band * 8u < 8u ? height - band * 8u : 8u;
    const uint32_t col = r / rows, row = r - col * rows;
    return (band * 8u + row) * width + col;
}


// Split solve pass. The light-sampling kernel (one 64-lane wave per region of 64 consecutive paths)
// replays each path's vertices with the path's random stream -- light samples, shadow rays and
// manifold walks -- and appends one light event per sourced beam that reaches the vertex: NEE,
// emissive hit or environment miss. The unwind kernel (one wave per region) rebuilds each beam,
// carries it through the stored vertices to the camera and sums every path's events in the order
// the megakernel adds them. Events are SoA planes of kEvWords words; a region's events occupy
// [region * 64 * ev_per_path, +64 * ev_per_path) in the order they were appended, which is each
// path's own vertex order. Walk vertices of manifold NEE events live in a second area, kHitWords
// planes per vertex, 64 * hits_per_path per region.
enum EvField {
    kEvMeta = 0,  // lane (6 bits) | vertex (7 bits) << 6 | kind (2 bits) << 13 | kEvBit* | walk vertices << 21
    kEvLe,        // 4 planes: sourced spectral radiance per HWSS component
    kEvLam = kEvLe + 4,  // 4 planes: wavelengths
    kEvPos = kEvLam + 4, // 3 planes: beam position (area sources)
    kEvDir = kEvPos + 3, // 3 planes: beam direction as sourced
    kEvSize = kEvDir + 3,  // beam area (mm^2) or solid angle
    kEvDist,      // source distance (area sources)
    kEvTypePdf,   // NEE: light-type selection probability
    kEvLsPdf,     // NEE: light-sample pdf (with the type selection)
    kEvMsPdf,     // manifold NEE: walk pdf
    kEvWo,        // 3 planes: manifold NEE: last walk direction towards the light
    kEvOrg = kEvWo + 3,  // 3 planes: manifold NEE: walk origin
    kEvHitBase = kEvOrg + 3,  // manifold NEE: first walk vertex in the region's vertex area
    kEvWords
};
// kEvKindVoid: a manifold NEE event whose deferred walk failed; it adds nothing to its path.
enum : uint32_t { kEvKindNee = 0, kEvKindEmissive = 1, kEvKindMiss = 2, kEvKindVoid = 3 };
enum : uint32_t {
    kEvBitComps1 = 1u << 15, kEvBitDistant = 1u << 16, kEvBitMnee = 1u << 17, kEvBitMisL = 1u << 18,
    kEvBitCurrentMs = 1u << 19, kEvBitDispersive = 1u << 20
};
enum HitField { kHitTri = 0, kHitB1, kHitB2, kHitPos, kHitLobe = kHitPos + 3, kHitWords };
constexpr uint32_t kSolveRegion = 64;  // paths per region: one wave of the split solve kernels
constexpr size_t kEventsPerExecute = 6;  // kernel-timing events per execute of the pass
// per region: event count, walk-vertex count, overflow mask (2 words: bit = lane whose path is left
// to the megakernel)
enum { kRegEvents = 0, kRegHits, kRegOverflowLo, kRegOverflowHi, kRegWords };
// The learned BSDF surrogate in the split solve pass: a query pass (the unwind kernel appending the
// network's inputs for every unwind step whose lobe the surrogate stands for), one batched
// inference, and an apply pass (the unwind kernel using the network's Mueller matrices at those
// steps). Steps whose inputs leave the training box, or that find no room, are evaluated exactly.
struct SurPath {
    const float* stats;  // outputs' mean [16] and std [16], inputs' training box lo [22] and hi [22]
    float* X;            // query rows [qcap x 22]
    const float* Y;      // network outputs of the rows [qcap x 16]
    uint32_t* counter;   // [0] rows appended this frame, [1] eligible steps evaluated exactly
    uint32_t qcap;
    uint32_t* qidx;      // [k * ev_cap + e]: first row of event e's k-th unwind step, ~0 = exact
    uint32_t steps_max;  // unwind steps per event at most: walk vertices + current vertex + forward steps
    uint32_t teacher;    // 1: the query pass writes the exact lobe's matrix as the outputs (no network)
};
enum : int { kUnwindExact = 0, kUnwindQuery = 1, kUnwindApply = 2 };
// Deferred manifold walks: the light-sampling kernel stores a walk's starting vertices (everything
// that draws random numbers: the current vertex's and the occluders' delta-lobe samples) with its
// event and appends the event's index here; the walk kernel, a persistent grid whose lanes each take
// the next walk as soon as theirs ends, runs the Newton iterations and the light's visibility, writes
// the converged vertices and the last direction into the event, or turns it into kEvKindVoid.
struct WalkQueue {
    uint32_t* items;     // event indices, ev_cap entries
    uint32_t* counters;  // [0] walks appended this frame, [1] walks taken by the walk kernel, [2] rays cast
};
struct SolveBuffers {
    uint32_t* ev;
    uint32_t* hits;
    uint32_t* region;
    uint32_t ev_per_path, hits_per_path;
    size_t ev_cap, hit_cap;  // plane lengths: regions * 64 * per-path capacity
    WalkQueue walk;
    // Staged unwind: the unwind kernel with `stage` set writes the RGB contribution of every event
    // that is not a manifold walk into `contrib` (3 planes of ev_cap) and sums nothing; the
    // walk-unwind kernel adds the walks' contributions, and the sum kernel adds each path's
    // contributions in slot order (the megakernel's order) and writes the colour.
    float* contrib;
    uint32_t stage;
    SurPath sur;             // used by the query and apply unwind modes only
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
    uint32_t solve_schedule = 0;  // PLT_SOLVE_*
    uint32_t solve_events_per_path = 8;
    uint32_t solve_hits_per_path = 4;
};

int launch_sample_pass(const PtScene& sc, const PtFrame& fr, const LaunchConfig& lc, hipStream_t st);
int launch_sample_pass_wavefront(const PtScene& sc, const PtFrame& fr, const WfBuffers& w, const LaunchConfig& lc, hipStream_t st);
int launch_solve_pass(const PtScene& sc, const PtFrame& fr, hipStream_t st);
// A second stream for the split pass: with it, the unwind of the events that are not manifold walks
// runs beside the walk kernel (fork and join through the two events); without it (aux = NULL) the
// launches are sequential on `st`.
struct SplitFork {
    hipStream_t aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
// after_light / after_walk / after_unwind (may be NULL): events recorded on `st` after the
// light-sampling kernel, after the walk kernel, and after the unwind (when forked: the walks'
// unwind, the join with the other events' unwind on the second stream, and the per-path sums);
// kernel timing outside graph capture
int launch_solve_pass_split(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, hipStream_t st,
                            hipEvent_t after_light = nullptr, hipEvent_t after_walk = nullptr, hipEvent_t after_unwind = nullptr,
                            plt_surrogate* surrogate = nullptr, const SplitFork& fork = SplitFork{});
int launch_tile_draws(const PtFrame& fr, uint32_t tiles, float* draws, hipStream_t st);

}  // namespace plt
