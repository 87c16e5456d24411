// The PLTPT render pass over the C ABI (Source/RenderPasses/PLTPT/PLTPT.{h,cpp}): option
// dictionary with the reference's keys and defaults (PLTPT.h:74-108, PLTPT.cpp:102-204),
// scene binding with emissive-light tables (LightCollection + EmissivePowerSampler), and execute =
// sample pass + solve pass over the whole frame.
//
// The reference splits the frame into 512^2 tiles to bound its bounce buffer; here one dispatch
// covers the frame (the 36 B x (maxBounces+1) records of a 1080p frame are < 1 GB of HBM), so
// `tileSize` is accepted and reported but does not change the result or the launch shape.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <new>
#include <vector>
#include "pathtracer.hpp"
#include "plt_internal.hpp"
#include "sample_pattern.hpp"
#include "scene_device.hpp"
#include "scene_host.hpp"
#include "surrogate.hpp"

using plt::set_error;

namespace plt {
namespace {

struct OptionSpec {
    const char* key;
    double def;
};
// Keys and defaults of PLTPT (PLTPT.h:74-108). sampleGenerator 1 = SAMPLE_GENERATOR_DEFAULT =
// uniform (xoshiro128**), 0 = tiny (SampleGeneratorType.slangh:32-36); emissiveSampler 2 = power
// (EmissiveLightSamplerType: Uniform 0, LightBVH 1, Power 2; EmissiveLightSamplerType.slangh:36-42).
const OptionSpec kOptions[] = {
    {"maxBounces", 32}, {"tileSize", 512}, {"HWSS", 4}, {"HWSSDoMIS", 1}, {"sampleGenerator", 1},
    {"debugView", 0}, {"debugViewIntensity", 1}, {"sourcingAreaFromEmissiveGeometry", 1},
    {"sourcingMaxBeamOmega", .025}, {"useDirectLights", 1}, {"useEnvLights", 1}, {"useEmissiveLights", 1},
    {"useAnalyticLights", 1}, {"alphaMasking", 1}, {"doNEE", 1}, {"doMIS", 1}, {"emissiveSampler", 2},
    {"doRussianRoulette", 1}, {"NEEUsePerTileSG", 0}, {"doImportanceSampleEmitters", 1}, {"doMNEE", 1},
    {"MNEEMaxOccluders", 2}, {"MNEEMaxIterations", 60}, {"MNEESolverThreshold", 5e-5},
};

struct LightTables {
    std::vector<float4> lights;
    std::vector<uint2> alias;
    std::vector<float> pdf;          // power sampler selection pdf
    std::vector<uint32_t> tri_light;
    std::vector<float> pdf_uniform;  // uniform sampler selection pdf: 1 / active count, 0 if inactive
    std::vector<uint32_t> active;    // lights with non-zero flux (LightCollection::updateActiveTriangleList)
};

// Emissive triangles in mesh order (LightCollection), flux = pi * area (the single emissive
// profile's luminance cancels in the selection probabilities), alias table as
// EmissivePowerSampler::generateAliasTable (EmissivePowerSampler.cpp:86-176) with its
// default-seeded std::mt19937 shuffle.
LightTables build_lights(const HostScene& s) {
    LightTables t;
    const uint32_t nt = uint32_t(s.indices.size() / 3);
    t.tri_light.assign(nt, 0xFFFFFFFFu);
    std::vector<float> flux;
    for (uint32_t tri = 0; tri < nt; ++tri) {
        const uint32_t mat = s.mesh_material[s.tri_mesh[tri]];
        const plt_material& m = s.materials[mat];
        if (m.type != PLT_MAT_DIFFUSE || m.emission_profile == PLT_NO_PROFILE) continue;
        const float* p0 = &s.positions[3 * size_t(s.indices[3 * size_t(tri)])];
        const float* p1 = &s.positions[3 * size_t(s.indices[3 * size_t(tri) + 1])];
        const float* p2 = &s.positions[3 * size_t(s.indices[3 * size_t(tri) + 2])];
        const float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const float e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        const float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const float len = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        const float area = 0.5f * len;
        t.tri_light[tri] = uint32_t(t.lights.size() / 4);
        float mbits;
        std::memcpy(&mbits, &mat, 4);
        t.lights.push_back(make_float4(p0[0], p0[1], p0[2], area));
        t.lights.push_back(make_float4(p1[0], p1[1], p1[2], mbits));
        t.lights.push_back(make_float4(p2[0], p2[1], p2[2], 0.f));
        t.lights.push_back(make_float4(c[0] / len, c[1] / len, c[2] / len, 0.f));
        flux.push_back(area * float(M_PI));
    }
    const uint32_t N = uint32_t(flux.size());
    if (N == 0) return t;
    // active triangles: flux > 0 (LightCollection.cpp:497-523)
    for (uint32_t i = 0; i < N; ++i)
        if (flux[i] > 0.f) t.active.push_back(i);
    const float uni = t.active.empty() ? 0.f : 1.f / float(t.active.size());
    for (uint32_t i = 0; i < N; ++i) t.pdf_uniform.push_back(flux[i] > 0.f ? uni : 0.f);
    double sum = 0.0;
    for (float f : flux) sum += f;
    const float invSum = 1.f / float(sum);
    for (float f : flux) t.pdf.push_back(f * invSum);
    std::vector<float> w = flux;
    for (float& f : w) f *= float(N) / float(sum);
    std::vector<uint32_t> perm(N), redirect(N);
    std::vector<float> thr(N);
    for (uint32_t i = 0; i < N; ++i) perm[i] = i;
    std::sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return w[a] < w[b]; });
    uint32_t head = 0, tail = N - 1;
    if (N == 1) { thr[0] = 1.f; redirect[0] = 0; }
    while (head != tail) {
        const uint32_t i = perm[head], j = perm[tail];
        thr[i] = w[i];
        redirect[i] = j;
        w[j] -= 1.0f - w[i];
        if (head == tail - 1) { thr[j] = 1.0f; redirect[j] = j; break; }
        else if (w[j] < 1.0f) { std::swap(perm[head], perm[tail]); tail--; }
        else head++;
    }
    for (uint32_t i = 0; i < N; ++i) perm[i] = i;
    std::mt19937 rng;
    for (uint32_t i = 0; i < N; ++i) {
        const uint32_t dst = i + (uint32_t(rng()) % (N - i));
        std::swap(thr[i], thr[dst]);
        std::swap(redirect[i], redirect[dst]);
        std::swap(perm[i], perm[dst]);
    }
    for (uint32_t i = 0; i < N; ++i) {
        const uint32_t prob = uint32_t(f32_to_h(thr[i])) << 16u;
        const uint32_t lx = redirect[i] & 0xFFFFFFu, ly = perm[i] & 0xFFFFFFu;
        t.alias.push_back(make_uint2(prob | ((lx >> 8u) & 0xFFFFu), ((lx & 0xFFu) << 24u) | ly));
    }
    return t;
}

// Host-side fault injection for tests of the scene-binding error path: when >= 0, the upload that
// finds it at 0 fails as an allocation failure would (it counts down once per upload).
int g_fail_upload_countdown = -1;

template <typename T>
int upload_vec(const std::vector<T>& v, T** d, std::vector<void*>& owned) {
    *d = nullptr;
    if (g_fail_upload_countdown >= 0 && g_fail_upload_countdown-- == 0)
        return set_error(PLT_ENOMEM, "pass: scene upload failed (injected)");
    if (v.empty()) return PLT_OK;
    void* p = nullptr;
    if (hipMalloc(&p, v.size() * sizeof(T)) != hipSuccess) return set_error(PLT_ENOMEM, "pass: hipMalloc failed");
    owned.push_back(p);
    if (hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return set_error(PLT_EHIP, "pass: hipMemcpy failed");
    *d = static_cast<T*>(p);
    return PLT_OK;
}
}  // namespace
}  // namespace plt

struct plt_pass {
    plt_context* ctx = nullptr;
    std::map<std::string, double> opt;
    plt_scene* scene = nullptr;
    uint32_t frame = 0;
    // device state
    std::vector<void*> owned;  // scene-dependent buffers
    plt_material* d_materials = nullptr;
    float4* d_lights = nullptr;
    uint2* d_alias = nullptr;
    float* d_light_pdf = nullptr;
    uint32_t* d_tri_light = nullptr;
    uint32_t n_lights = 0;
    float* d_light_pdf_uniform = nullptr;
    uint32_t* d_active = nullptr;
    uint32_t n_active = 0;
    plt_light* d_analytic = nullptr;
    uint32_t n_analytic = 0;
    float* d_env_rgba = nullptr;
    float* d_env_imp = nullptr;
    uint32_t* d_records = nullptr;
    size_t records_words = 0;
    float* d_tile_draws = nullptr;  // NEEUsePerTileSG draws (tiles x (maxBounces + 1))
    size_t tile_draws_cap = 0;
    size_t last_words = 0;  // record words written by the last execute
    uint32_t last_width = 0, last_height = 0;  // its frame size (records are in path order)
    float cam[12] = {};
    float view_rot[9] = {};
    // kernel timing: kEventsPerExecute events per execute: before the sample pass, after it, after the
    // split pass's light-sampling, walk and unwind kernels (all three at the solve pass's start
    // otherwise), after the solve pass
    bool profiling = false;
    std::vector<hipEvent_t> events;
    size_t events_used = 0;
    // sample-pass schedule (PLT_SCHEDULE_*), its launch parameters, and the wavefront schedule's buffers
    int schedule = PLT_SCHEDULE_MEGAKERNEL;
    plt::LaunchConfig launch;
    // graph replay: captured on capture_stream, instantiated once and updated per frame
    bool graph = false;
    hipStream_t capture_stream = nullptr;
    // split solve pass: the second stream and its fork/join events (created at the first split execute)
    plt::SplitFork fork;
    hipGraphExec_t graph_sample = nullptr, graph_solve = nullptr;
    void* d_wf = nullptr;
    size_t wf_paths = 0;
    // split solve pass: event planes, walk-vertex planes and region table in one allocation, and the
    // layout the last execute used
    void* d_solve = nullptr;
    size_t solve_bytes = 0;
    plt::SolveBuffers last_solve{};
    size_t last_regions = 0;
    // learned BSDF surrogate in the split solve pass (not owned), and its query buffers
    plt_surrogate* surrogate = nullptr;
    bool surrogate_teacher = false;
    void* d_sur = nullptr;
    size_t sur_bytes = 0;
    // camera jitter pattern of the G-buffer pass (Center: the camera's own jitter)
    plt::SamplePattern pattern;
    float last_jitter[2] = {0.f, 0.f};
};

namespace plt {
namespace {
void free_scene_state(plt_pass* p) {
    for (void* q : p->owned) (void)hipFree(q);
    p->owned.clear();
    p->d_materials = nullptr;
    p->d_lights = nullptr;
    p->d_alias = nullptr;
    p->d_light_pdf = nullptr;
    p->d_tri_light = nullptr;
    p->n_lights = 0;
    p->d_light_pdf_uniform = nullptr;
    p->d_active = nullptr;
    p->n_active = 0;
    p->d_analytic = nullptr;
    p->n_analytic = 0;
    p->d_env_rgba = nullptr;
    p->d_env_imp = nullptr;
}
// The rotation of the camera's view matrix, rmcv::lookAt = glm::lookAt, right-handed
// (Camera.cpp:135, Utils/Math/Matrix/Matrix.h): rows s = normalize(f x up), u = s x f, -f with
// f = normalize(target - pos).
void view_rotation(const plt_camera& c, float out[9]) {
    float f[3] = {c.target[0] - c.pos[0], c.target[1] - c.pos[1], c.target[2] - c.pos[2]};
    const float fl = 1.f / std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (float& x : f) x *= fl;
    float sv[3] = {f[1] * c.up[2] - f[2] * c.up[1], f[2] * c.up[0] - f[0] * c.up[2], f[0] * c.up[1] - f[1] * c.up[0]};
    const float sl = 1.f / std::sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
    for (float& x : sv) x *= sl;
    const float u[3] = {sv[1] * f[2] - sv[2] * f[1], sv[2] * f[0] - sv[0] * f[2], sv[0] * f[1] - sv[1] * f[0]};
    for (int a = 0; a < 3; ++a) {
        out[a] = sv[a];
        out[3 + a] = u[a];
        out[6 + a] = -f[a];
    }
}
// Camera frame (Camera.cpp:187-193).
void camera_frame(const plt_camera& c, float out[12]) {
    auto norm = [](float v[3]) { const float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); v[0] /= l; v[1] /= l; v[2] /= l; };
    float W[3] = {c.target[0] - c.pos[0], c.target[1] - c.pos[1], c.target[2] - c.pos[2]};
    norm(W);
    for (float& x : W) x *= c.focal_distance;
    float U[3] = {W[1] * c.up[2] - W[2] * c.up[1], W[2] * c.up[0] - W[0] * c.up[2], W[0] * c.up[1] - W[1] * c.up[0]};
    norm(U);
    float V[3] = {U[1] * W[2] - U[2] * W[1], U[2] * W[0] - U[0] * W[2], U[0] * W[1] - U[1] * W[0]};
    norm(V);
    const float tn = std::tan(c.fov_y * 0.5f);
    const float ulen = c.focal_distance * tn * c.aspect, vlen = c.focal_distance * tn;
    for (int a = 0; a < 3; ++a) {
        out[a] = c.pos[a];
        out[3 + a] = U[a] * ulen;
        out[6 + a] = V[a] * vlen;
        out[9 + a] = W[a];
    }
}
int bind_scene(plt_pass* p, plt_scene* scene);

void drop_graphs(plt_pass* p) {
    if (p->graph_sample) (void)hipGraphExecDestroy(p->graph_sample);
    if (p->graph_solve) (void)hipGraphExecDestroy(p->graph_solve);
    p->graph_sample = p->graph_solve = nullptr;
}

// Captures the launches `seq` enqueues into a graph on the pass's capture stream, updates the
// instantiated graph `exec` with it (kernel arguments change every frame: frame number, jitter),
// instantiating anew when the update is refused (a different launch sequence), and launches it
// on `st`.
template <class Seq>
int run_captured(plt_pass* p, hipGraphExec_t& exec, Seq&& seq, hipStream_t st) {
    if (!p->capture_stream && hipStreamCreateWithFlags(&p->capture_stream, hipStreamNonBlocking) != hipSuccess)
        return set_error(PLT_EHIP, "plt_pass_execute: capture stream creation failed");
    if (hipStreamBeginCapture(p->capture_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return set_error(PLT_EHIP, "plt_pass_execute: hipStreamBeginCapture failed");
    const int rc = seq(p->capture_stream);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(p->capture_stream, &g);
    if (rc < 0 || ec != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return rc < 0 ? rc : set_error(PLT_EHIP, "plt_pass_execute: hipStreamEndCapture failed");
    }
    if (exec) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res;
        if (hipGraphExecUpdate(exec, g, &err_node, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(exec);
            exec = nullptr;
        }
    }
    if (!exec && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        exec = nullptr;
        return set_error(PLT_EHIP, "plt_pass_execute: hipGraphInstantiate failed");
    }
    (void)hipGraphDestroy(g);
    if (hipGraphLaunch(exec, st) != hipSuccess) return set_error(PLT_EHIP, "plt_pass_execute: hipGraphLaunch failed");
    return PLT_OK;
}
}  // namespace
}  // namespace plt

extern "C" {

int plt_debug_fail_upload(int n) {
    plt::g_fail_upload_countdown = n < 0 ? -1 : n;
    return PLT_OK;
}

int plt_pass_create(plt_context* ctx, const char* type, plt_pass** out) {
    if (!out) return set_error(PLT_EINVAL, "plt_pass_create: bad arguments");
    if (!type || std::strcmp(type, "PLTPT") != 0) return set_error(PLT_EINVAL, "plt_pass_create: unknown pass type (only \"PLTPT\")");
    plt_pass* p = new (std::nothrow) plt_pass();
    if (!p) return set_error(PLT_ENOMEM, "plt_pass_create: out of host memory");
    p->ctx = ctx;
    for (const auto& o : plt::kOptions) p->opt[o.key] = o.def;
    *out = p;
    return PLT_OK;
}

int plt_pass_destroy(plt_pass* p) {
    if (!p) return PLT_OK;
    plt::free_scene_state(p);
    if (p->d_records) (void)hipFree(p->d_records);
    if (p->d_tile_draws) (void)hipFree(p->d_tile_draws);
    if (p->d_wf) (void)hipFree(p->d_wf);
    if (p->d_solve) (void)hipFree(p->d_solve);
    if (p->d_sur) (void)hipFree(p->d_sur);
    for (hipEvent_t e : p->events) (void)hipEventDestroy(e);
    plt::drop_graphs(p);
    if (p->capture_stream) (void)hipStreamDestroy(p->capture_stream);
    if (p->fork.aux) (void)hipStreamDestroy(p->fork.aux);
    if (p->fork.fork) (void)hipEventDestroy(p->fork.fork);
    if (p->fork.join) (void)hipEventDestroy(p->fork.join);
    delete p;
    return PLT_OK;
}

int plt_pass_set_graph(plt_pass* p, int enable) {
    if (!p) return set_error(PLT_EINVAL, "plt_pass_set_graph: NULL pass");
    p->graph = enable != 0;
    if (!p->graph) plt::drop_graphs(p);
    return PLT_OK;
}

int plt_pass_set_launch_config(plt_pass* p, const plt_launch_config* c) {
    if (!p || !c) return set_error(PLT_EINVAL, "plt_pass_set_launch_config: bad arguments");
    if (c->schedule != PLT_SCHEDULE_MEGAKERNEL && c->schedule != PLT_SCHEDULE_WAVEFRONT)
        return set_error(PLT_EINVAL, "plt_pass_set_launch_config: unknown schedule");
    if (c->sample_block != 64 && c->sample_block != 128 && c->sample_block != 256)
        return set_error(PLT_EINVAL, "plt_pass_set_launch_config: sample_block must be 64, 128 or 256");
    if (c->wf_blocks_per_cu < 1 || c->wf_blocks_per_cu > 64)
        return set_error(PLT_EINVAL, "plt_pass_set_launch_config: wf_blocks_per_cu must be in [1, 64]");
    if (c->solve_schedule != PLT_SOLVE_MEGAKERNEL && c->solve_schedule != PLT_SOLVE_SPLIT)
        return set_error(PLT_EINVAL, "plt_pass_set_launch_config: unknown solve schedule");
    if (c->solve_events_per_path < 1 || c->solve_events_per_path > 64 || c->solve_hits_per_path > 64)
        return set_error(PLT_EINVAL, "plt_pass_set_launch_config: solve_events_per_path must be in [1, 64], "
                                     "solve_hits_per_path in [0, 64]");
    p->schedule = int(c->schedule);
    p->launch.sample_block = c->sample_block;
    p->launch.wf_blocks_per_cu = c->wf_blocks_per_cu;
    p->launch.solve_schedule = c->solve_schedule;
    p->launch.solve_events_per_path = c->solve_events_per_path;
    p->launch.solve_hits_per_path = c->solve_hits_per_path;
    p->graph = c->graph != 0;
    if (!p->graph) plt::drop_graphs(p);
    return PLT_OK;
}

int plt_pass_get_launch_config(const plt_pass* p, plt_launch_config* c) {
    if (!p || !c) return set_error(PLT_EINVAL, "plt_pass_get_launch_config: bad arguments");
    c->schedule = uint32_t(p->schedule);
    c->sample_block = p->launch.sample_block;
    c->wf_blocks_per_cu = p->launch.wf_blocks_per_cu;
    c->graph = p->graph ? 1u : 0u;
    c->solve_schedule = p->launch.solve_schedule;
    c->solve_events_per_path = p->launch.solve_events_per_path;
    c->solve_hits_per_path = p->launch.solve_hits_per_path;
    return PLT_OK;
}

int plt_pass_set_schedule(plt_pass* p, int schedule) {
    if (!p || (schedule != PLT_SCHEDULE_MEGAKERNEL && schedule != PLT_SCHEDULE_WAVEFRONT))
        return set_error(PLT_EINVAL, "plt_pass_set_schedule: bad arguments");
    p->schedule = schedule;
    return PLT_OK;
}

int plt_pass_set_option(plt_pass* p, const char* key, double value) {
    if (!p || !key) return set_error(PLT_EINVAL, "plt_pass_set_option: bad arguments");
    auto it = p->opt.find(key);
    if (it == p->opt.end()) return set_error(PLT_WARN_UNKNOWN_KEY, std::string("Unknown field '") + key + "' in PLTPathTracer dictionary.");
    if (!std::isfinite(value)) return set_error(PLT_EINVAL, std::string("plt_pass_set_option: non-finite value for ") + key);
    const std::string k = key;
    if (k == "HWSS") value = std::max(1.0, std::min(4.0, std::floor(value)));  // PLTPT.cpp:291-292
    it->second = value;
    return PLT_OK;
}

int plt_pass_get_option(const plt_pass* p, const char* key, double* value) {
    if (!p || !key || !value) return set_error(PLT_EINVAL, "plt_pass_get_option: bad arguments");
    auto it = p->opt.find(key);
    if (it == p->opt.end()) return set_error(PLT_WARN_UNKNOWN_KEY, std::string("Unknown field '") + key + "'");
    *value = it->second;
    return PLT_OK;
}

const char* plt_pass_option_name(uint32_t i) {
    return i < sizeof(plt::kOptions) / sizeof(plt::kOptions[0]) ? plt::kOptions[i].key : nullptr;
}

int plt_pass_set_scene(plt_pass* p, plt_scene* scene) {
    if (!p) return set_error(PLT_EINVAL, "plt_pass_set_scene: bad arguments");
    plt::free_scene_state(p);
    p->scene = nullptr;
    p->frame = 0;  // PLTPT.cpp:510
    p->pattern.configure(p->pattern.type, p->pattern.count);  // a new generator (GBufferBase.cpp:185-189)
    if (!scene) return PLT_OK;
    if (!scene->ctx || !p->ctx) return set_error(PLT_ESTATE, "plt_pass_set_scene: host-only scene or pass");
    // the scene is bound only once every buffer is on the device: a failed upload leaves the pass
    // unbound (execute then returns PLT_ESTATE) with nothing allocated
    const int rc = plt::bind_scene(p, scene);
    if (rc < 0) {
        plt::free_scene_state(p);
        return rc;
    }
    p->scene = scene;
    return PLT_OK;
}

}  // extern "C"

namespace plt {
namespace {
int bind_scene(plt_pass* p, plt_scene* scene) {
    const plt::LightTables lt = plt::build_lights(scene->host);
    int rc;
    if ((rc = plt::upload_vec(scene->host.materials, &p->d_materials, p->owned)) < 0) return rc;
    if ((rc = plt::upload_vec(lt.lights, &p->d_lights, p->owned)) < 0) return rc;
    if ((rc = plt::upload_vec(lt.alias, &p->d_alias, p->owned)) < 0) return rc;
    if ((rc = plt::upload_vec(lt.pdf, &p->d_light_pdf, p->owned)) < 0) return rc;
    if ((rc = plt::upload_vec(lt.tri_light, &p->d_tri_light, p->owned)) < 0) return rc;
    if ((rc = plt::upload_vec(lt.pdf_uniform, &p->d_light_pdf_uniform, p->owned)) < 0) return rc;
    if ((rc = plt::upload_vec(lt.active, &p->d_active, p->owned)) < 0) return rc;
    p->n_active = uint32_t(lt.active.size());
    if ((rc = plt::upload_vec(scene->host.lights, &p->d_analytic, p->owned)) < 0) return rc;
    p->n_analytic = uint32_t(scene->host.lights.size());
    if ((rc = plt::upload_vec(scene->host.env.rgba, &p->d_env_rgba, p->owned)) < 0) return rc;
    if ((rc = plt::upload_vec(scene->host.env.importance, &p->d_env_imp, p->owned)) < 0) return rc;
    // Scene::useEmissiveLights requires an active triangle; otherwise emissive lighting is off
    p->n_lights = p->n_active > 0 ? uint32_t(lt.pdf.size()) : 0;
    plt::camera_frame(scene->host.camera, p->cam);
    plt::view_rotation(scene->host.camera, p->view_rot);
    return PLT_OK;
}
}  // namespace
}  // namespace plt

extern "C" {

int plt_pass_set_sample_pattern(plt_pass* p, int pattern, uint32_t sample_count) {
    if (!p) return set_error(PLT_EINVAL, "plt_pass_set_sample_pattern: NULL pass");
    return p->pattern.configure(pattern, sample_count);
}

int plt_pass_get_jitter(const plt_pass* p, float* xy) {
    if (!p || !xy) return set_error(PLT_EINVAL, "plt_pass_get_jitter: bad arguments");
    xy[0] = p->last_jitter[0];
    xy[1] = p->last_jitter[1];
    return PLT_OK;
}

int plt_pass_set_frame(plt_pass* p, uint32_t frame) {
    if (!p) return set_error(PLT_EINVAL, "plt_pass_set_frame: bad arguments");
    p->frame = frame;
    return PLT_OK;
}

int plt_pass_execute(plt_pass* p, const plt_render_data* rd, plt_stream stream) {
    if (!p || !rd) return set_error(PLT_EINVAL, "plt_pass_execute: bad arguments");
    if (!p->ctx) return set_error(PLT_ESTATE, "plt_pass_execute: pass created without a device context");
    if (!p->scene) return set_error(PLT_ESTATE, "plt_pass_execute: no scene (the reference clears its outputs, PLTPT.cpp:375-384)");
    if (!rd->color || !rd->albedo || !rd->normal || rd->width == 0 || rd->height == 0)
        return set_error(PLT_EINVAL, "plt_pass_execute: missing output buffers or empty frame");
    auto o = [&](const char* k) { return p->opt[k]; };
    const double mneeOcc = o("MNEEMaxOccluders"), mneeIt = o("MNEEMaxIterations"), mneeThr = o("MNEESolverThreshold");
    if (!(mneeOcc >= 0) || !(mneeIt >= 0) || mneeIt > 1e6 || !std::isfinite(mneeThr))
        return set_error(PLT_EINVAL, "MNEEMaxOccluders and MNEEMaxIterations must be non-negative (iterations <= 1e6), "
                                     "MNEESolverThreshold finite");
    if (mneeOcc > plt::kMaxMsOccluders) return set_error(PLT_ENOTIMPL, "MNEEMaxOccluders above 8 is not built");
    // EmissiveLightSamplerType: Uniform 0, Power 2; the pass creates no other sampler (PLTPT.cpp:399-412)
    const double emissiveSampler = o("emissiveSampler");
    if (emissiveSampler != 0 && emissiveSampler != 2)
        return set_error(PLT_ENOTIMPL, "emissiveSampler must be 0 (Uniform) or 2 (Power); the reference's pass "
                                       "creates no other sampler");
    // SampleGeneratorType.slangh:32-36: 0 = tiny uniform (LCG), 1 = uniform (xoshiro128**)
    const double sgType = o("sampleGenerator");
    if (sgType != 0 && sgType != 1) return set_error(PLT_EINVAL, "sampleGenerator must be 0 (tiny uniform) or 1 (uniform)");
    const double tileSize = o("tileSize");
    if (!(tileSize >= 1) || tileSize > 65536) return set_error(PLT_EINVAL, "tileSize must be in [1, 65536]");
    if (o("debugView") != 0) return set_error(PLT_ENOTIMPL, "debug views are not built");
    const uint32_t hwss = uint32_t(o("HWSS"));
    if (hwss != 1 && hwss != 4) return set_error(PLT_ENOTIMPL, "HWSS must be 1 or 4");
    const uint32_t maxb = uint32_t(o("maxBounces"));
    if (maxb > 64) return set_error(PLT_EINVAL, "maxBounces > 64");
    const size_t npaths = size_t(rd->width) * rd->height;
    if (npaths >= (size_t(1) << 31)) return set_error(PLT_EINVAL, "frame too large");
    const size_t words = size_t(maxb + 1) * plt::kRecWords * npaths;
    if (words > p->records_words) {
        if (p->d_records) (void)hipFree(p->d_records);
        p->d_records = nullptr;
        p->records_words = 0;
        void* q = nullptr;
        if (hipMalloc(&q, words * 4) != hipSuccess) return set_error(PLT_ENOMEM, "plt_pass_execute: bounce records allocation failed");
        p->d_records = static_cast<uint32_t*>(q);
        p->records_words = words;
    }
    const plt_scene* s = p->scene;
    // the camera as the scene holds it now (plt_scene_set_camera may have changed it since binding)
    plt::camera_frame(s->host.camera, p->cam);
    plt::view_rotation(s->host.camera, p->view_rot);
    plt::PtScene sc{};
    sc.tsc = plt::TraceScene{s->dev.nodes, s->dev.tris};
    sc.positions = s->dev.positions;
    sc.normals = s->dev.normals;
    sc.indices = s->dev.indices;
    sc.tri_mesh = s->dev.tri_mesh;
    sc.mesh_material = s->dev.mesh_material;
    sc.materials = p->d_materials;
    sc.profiles = s->dev.profiles;
    sc.n_profiles = uint32_t(s->host.profiles.size());
    for (int a = 0; a < 3; ++a) {
        sc.cam_pos[a] = p->cam[a];
        sc.cam_u[a] = p->cam[3 + a];
        sc.cam_v[a] = p->cam[6 + a];
        sc.cam_w[a] = p->cam[9 + a];
    }
    sc.jitter[0] = s->host.camera.jitter[0];
    sc.jitter[1] = s->host.camera.jitter[1];
    for (int a = 0; a < 9; ++a) sc.view_rot[a] = p->view_rot[a];
    sc.polarizer = s->host.camera.polarizer;
    sc.polarizer_theta = s->host.camera.polarizer_theta;
    if (p->pattern.type != PLT_SAMPLE_PATTERN_CENTER) {
        // Camera::beginFrame: jitter = next() * (1 / frame size) (Camera.cpp:63-70, GBufferBase.cpp:221-229)
        float j[2];
        p->pattern.next(j);
        const float inv_w = 1.f / float(rd->width), inv_h = 1.f / float(rd->height);
        sc.jitter[0] = j[0] * inv_w;
        sc.jitter[1] = j[1] * inv_h;
    }
    p->last_jitter[0] = sc.jitter[0];
    p->last_jitter[1] = sc.jitter[1];
    sc.lights = p->d_lights;
    sc.alias = p->d_alias;
    sc.light_pdf = emissiveSampler == 0 ? p->d_light_pdf_uniform : p->d_light_pdf;
    sc.active_lights = p->d_active;
    sc.n_active = p->n_active;
    sc.analytic = p->d_analytic;
    sc.n_analytic = p->n_analytic;
    {
        const plt::HostEnvMap& e = s->host.env;
        sc.env_rgba = reinterpret_cast<const float4*>(p->d_env_rgba);
        sc.env_w = p->d_env_rgba ? e.width : 0;
        sc.env_h = e.height;
        for (int i = 0; i < 9; ++i) {
            sc.env_m[i] = e.transform[i];
            sc.env_mi[i] = e.inv_transform[i];
        }
        for (int c = 0; c < 3; ++c) sc.env_scale[c] = e.scale[c];
        sc.env_imp = p->d_env_imp;
        sc.env_base_mip = e.base_mip;
        for (size_t k = 0; k < 12; ++k) sc.env_mip[k] = k < e.mip_offset.size() ? e.mip_offset[k] : 0u;
    }
    sc.tri_light = p->d_tri_light;
    sc.n_lights = p->n_lights;

    plt::PtFrame fr{};
    fr.width = rd->width;
    fr.height = rd->height;
    fr.frame = p->frame;
    fr.max_bounces = maxb;
    fr.hwss = hwss;
    uint32_t opts = 0;
    if (o("doNEE") != 0) opts |= plt::kOptNEE;
    if (o("doMIS") != 0) opts |= plt::kOptMIS;
    if (o("doRussianRoulette") != 0) opts |= plt::kOptRR;
    if (o("HWSSDoMIS") != 0) opts |= plt::kOptHwssMis;
    if (o("doImportanceSampleEmitters") != 0) opts |= plt::kOptISEmitters;
    if (o("useDirectLights") != 0) opts |= plt::kOptDirectLights;
    if (o("useEmissiveLights") != 0) opts |= plt::kOptEmissiveLights;
    if (o("doMNEE") != 0) opts |= plt::kOptMNEE;
    if (emissiveSampler == 0) opts |= plt::kOptUniformEmissive;
    if (o("useAnalyticLights") != 0) opts |= plt::kOptAnalyticLights;
    if (o("useEnvLights") != 0) opts |= plt::kOptEnvLight;
    fr.options = opts;
    fr.mnee_max_occluders = uint32_t(mneeOcc);
    fr.mnee_max_iterations = uint32_t(mneeIt);
    fr.mnee_threshold = float(mneeThr);
    fr.path_base = 0;
    fr.n_paths = uint32_t(npaths);
    fr.sample_generator = uint32_t(sgType);
    fr.tile_draws = nullptr;
    fr.tile_size = uint32_t(tileSize);
    fr.tiles_x = (rd->width + fr.tile_size - 1) / fr.tile_size;
    const uint32_t tiles = fr.tiles_x * ((rd->height + fr.tile_size - 1) / fr.tile_size);
    if (o("NEEUsePerTileSG") != 0 && (opts & plt::kOptNEE)) {
        const size_t need = size_t(tiles) * (maxb + 1);
        if (need > p->tile_draws_cap) {
            if (p->d_tile_draws) (void)hipFree(p->d_tile_draws);
            p->d_tile_draws = nullptr;
            p->tile_draws_cap = 0;
            void* q = nullptr;
            if (hipMalloc(&q, need * sizeof(float)) != hipSuccess)
                return set_error(PLT_ENOMEM, "plt_pass_execute: per-tile generator buffer allocation failed");
            p->d_tile_draws = static_cast<float*>(q);
            p->tile_draws_cap = need;
        }
        fr.tile_draws = p->d_tile_draws;
    }
    fr.sourcing_area = float(o("sourcingAreaFromEmissiveGeometry"));
    fr.max_omega = float(o("sourcingMaxBeamOmega"));
    fr.records = p->d_records;
    fr.color = reinterpret_cast<float4*>(rd->color);
    fr.albedo = reinterpret_cast<float4*>(rd->albedo);
    fr.normal = reinterpret_cast<float4*>(rd->normal);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    // device buffers first: nothing is allocated once launches are enqueued (or captured)
    plt::WfBuffers w{};
    if (p->schedule == PLT_SCHEDULE_WAVEFRONT) {
        if (npaths > p->wf_paths) {  // per path: 5 x 16 B state + 4 B length + 4 x 4 B queues and keys
            if (p->d_wf) (void)hipFree(p->d_wf);
            p->d_wf = nullptr;
            p->wf_paths = 0;
            void* q = nullptr;
            const size_t bytes = npaths * (5 * 16 + 4 + 4 * 4) + plt::kWfCounters * 4 + 256;
            if (hipMalloc(&q, bytes) != hipSuccess) return set_error(PLT_ENOMEM, "plt_pass_execute: wavefront buffers allocation failed");
            p->d_wf = q;
            p->wf_paths = npaths;
        }
        char* b = static_cast<char*>(p->d_wf);
        const size_t cap = p->wf_paths;
        w.rng = reinterpret_cast<uint4*>(b);
        w.org = reinterpret_cast<float4*>(b + 16 * cap);
        w.dir = reinterpret_cast<float4*>(b + 32 * cap);
        w.pdfs = reinterpret_cast<float4*>(b + 48 * cap);
        w.hit = reinterpret_cast<float4*>(b + 64 * cap);
        w.len = reinterpret_cast<uint32_t*>(b + 80 * cap);
        w.queue = reinterpret_cast<uint32_t*>(b + 84 * cap);
        w.sorted = reinterpret_cast<uint32_t*>(b + 92 * cap);
        w.keys = reinterpret_cast<uint32_t*>(b + 96 * cap);
        w.counters = reinterpret_cast<uint32_t*>(b + 100 * cap);
    }
    plt::SolveBuffers sbuf{};
    const bool useSplit = p->launch.solve_schedule == PLT_SOLVE_SPLIT || p->surrogate;
    if (useSplit) {
        const size_t regions = (npaths + plt::kSolveRegion - 1) / plt::kSolveRegion;
        const size_t evCap = regions * plt::kSolveRegion * p->launch.solve_events_per_path;
        const size_t hitCap = regions * plt::kSolveRegion * p->launch.solve_hits_per_path;
        if (evCap > size_t(UINT32_MAX)) return set_error(PLT_EINVAL, "plt_pass_execute: frame too large for the split solve pass's event area");
        const size_t evBytes = evCap * plt::kEvWords * 4, hitBytes = hitCap * plt::kHitWords * 4;
        const size_t regBytes = regions * plt::kRegWords * 4, walkBytes = evCap * 4, contribBytes = evCap * 3 * 4;
        const size_t bytes = evBytes + hitBytes + regBytes + walkBytes + contribBytes + 256;
        if (bytes > p->solve_bytes) {
            if (p->d_solve) (void)hipFree(p->d_solve);
            p->d_solve = nullptr;
            p->solve_bytes = 0;
            void* q = nullptr;
            if (hipMalloc(&q, bytes) != hipSuccess) return set_error(PLT_ENOMEM, "plt_pass_execute: solve event area allocation failed");
            p->d_solve = q;
            p->solve_bytes = bytes;
        }
        char* b = static_cast<char*>(p->d_solve);
        sbuf.ev = reinterpret_cast<uint32_t*>(b);
        sbuf.hits = reinterpret_cast<uint32_t*>(b + evBytes);
        sbuf.region = reinterpret_cast<uint32_t*>(b + evBytes + hitBytes);
        sbuf.walk.items = reinterpret_cast<uint32_t*>(b + evBytes + hitBytes + regBytes);
        sbuf.contrib = reinterpret_cast<float*>(b + evBytes + hitBytes + regBytes + walkBytes);
        sbuf.walk.counters = reinterpret_cast<uint32_t*>(b + evBytes + hitBytes + regBytes + walkBytes + contribBytes);
        sbuf.ev_per_path = p->launch.solve_events_per_path;
        sbuf.hits_per_path = p->launch.solve_hits_per_path;
        sbuf.ev_cap = evCap;
        sbuf.hit_cap = hitCap;
    }
    if (p->surrogate) {
        // query rows: four per path (the 1080p depth-8 workload appends about 1.4 per path); steps
        // beyond are evaluated exactly and counted
        const uint32_t stepsMax = maxb + 2u + uint32_t(mneeOcc);
        const size_t qcap = std::min(size_t(4) * npaths, size_t(1) << 28);
        const size_t xBytes = qcap * PLT_SURROGATE_INPUTS * 4, yBytes = qcap * PLT_SURROGATE_OUTPUTS * 4;
        const size_t qBytes = size_t(stepsMax) * sbuf.ev_cap * 4;
        const size_t bytes = 256 + xBytes + yBytes + qBytes;
        if (bytes > p->sur_bytes) {
            if (p->d_sur) (void)hipFree(p->d_sur);
            p->d_sur = nullptr;
            p->sur_bytes = 0;
            void* q = nullptr;
            if (hipMalloc(&q, bytes) != hipSuccess) return set_error(PLT_ENOMEM, "plt_pass_execute: surrogate query buffers allocation failed");
            p->d_sur = q;
            p->sur_bytes = bytes;
        }
        char* b = static_cast<char*>(p->d_sur);
        sbuf.sur.counter = reinterpret_cast<uint32_t*>(b);
        sbuf.sur.X = reinterpret_cast<float*>(b + 256);
        sbuf.sur.Y = reinterpret_cast<const float*>(b + 256 + xBytes);
        sbuf.sur.qidx = reinterpret_cast<uint32_t*>(b + 256 + xBytes + yBytes);
        sbuf.sur.qcap = uint32_t(qcap);
        sbuf.sur.steps_max = stepsMax;
        sbuf.sur.stats = p->surrogate->stats;
        sbuf.sur.teacher = p->surrogate_teacher ? 1u : 0u;
    }
    constexpr size_t kEv = plt::kEventsPerExecute;
    hipEvent_t ev[kEv] = {};
    p->last_solve = sbuf;
    p->last_regions = sbuf.ev ? (npaths + plt::kSolveRegion - 1) / plt::kSolveRegion : 0;
    if (p->profiling) {
        while (p->events.size() < p->events_used + kEv) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return set_error(PLT_EHIP, "plt_pass_execute: hipEventCreate failed");
            p->events.push_back(e);
        }
        for (size_t i = 0; i < kEv; ++i) ev[i] = p->events[p->events_used + i];
        p->events_used += kEv;
    }
    const int schedule = p->schedule;
    const plt::LaunchConfig lc = p->launch;
    float* tile_draws = p->d_tile_draws;
    // the sample pass (with the per-tile generator's draws) and the solve pass, each a sequence of
    // launches on one stream
    auto sample_seq = [&](hipStream_t s) -> int {
        int rc;
        if (fr.tile_draws && (rc = plt::launch_tile_draws(fr, tiles, tile_draws, s)) < 0) return rc;
        return schedule == PLT_SCHEDULE_WAVEFRONT ? plt::launch_sample_pass_wavefront(sc, fr, w, lc, s)
                                                   : plt::launch_sample_pass(sc, fr, lc, s);
    };
    const bool split = useSplit;
    plt_surrogate* surrogate = p->surrogate;
    if (split && !surrogate && !p->fork.aux) {
        if (hipStreamCreateWithFlags(&p->fork.aux, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&p->fork.fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&p->fork.join, hipEventDisableTiming) != hipSuccess)
            return set_error(PLT_EHIP, "plt_pass_execute: second stream of the split solve pass could not be created");
    }
    const plt::SplitFork fork = p->fork;
    // the split pass's own timing events are recorded between its launches when it runs eagerly
    const bool split_marks = split && !p->graph && ev[2];
    auto solve_seq = [&](hipStream_t s) -> int {
        return split ? plt::launch_solve_pass_split(sc, fr, sbuf, s, split_marks ? ev[2] : nullptr, split_marks ? ev[3] : nullptr,
                                                    split_marks ? ev[4] : nullptr, surrogate, fork)
                     : plt::launch_solve_pass(sc, fr, s);
    };
    int rc;
    if (ev[0]) (void)hipEventRecord(ev[0], st);
    auto mark_solve_start = [&]() {
        if (ev[1]) (void)hipEventRecord(ev[1], st);
        if (ev[2] && !split_marks) {
            (void)hipEventRecord(ev[2], st);
            (void)hipEventRecord(ev[3], st);
            (void)hipEventRecord(ev[4], st);
        }
    };
    if (p->graph) {
        // each pass replayed from a graph captured on the pass's own stream (a null caller stream
        // cannot be captured); the graph launches on the caller's stream, so ordering is the caller's
        if ((rc = plt::run_captured(p, p->graph_sample, sample_seq, st)) < 0) return rc;
        mark_solve_start();
        if ((rc = plt::run_captured(p, p->graph_solve, solve_seq, st)) < 0) return rc;
    } else {
        if ((rc = sample_seq(st)) < 0) return rc;
        mark_solve_start();
        if ((rc = solve_seq(st)) < 0) return rc;
    }
    if (ev[5]) (void)hipEventRecord(ev[5], st);
    ++p->frame;  // PLTPT.cpp:498
    p->last_words = words;
    p->last_width = rd->width;
    p->last_height = rd->height;
    return PLT_OK;
}

int plt_pass_copy_bounce_records(plt_pass* p, void* dst, size_t bytes) {
    if (!p || !dst) return set_error(PLT_EINVAL, "plt_pass_copy_bounce_records: bad arguments");
    if (!p->d_records || p->last_words == 0) return set_error(PLT_ESTATE, "plt_pass_copy_bounce_records: nothing rendered yet");
    if (bytes < p->last_words * 4) return set_error(PLT_EINVAL, "plt_pass_copy_bounce_records: destination too small");
    std::vector<uint32_t> paths;
    try {
        paths.resize(p->last_words);
    } catch (const std::bad_alloc&) {
        return set_error(PLT_ENOMEM, "plt_pass_copy_bounce_records: out of host memory");
    }
    if (hipMemcpy(paths.data(), p->d_records, p->last_words * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return set_error(PLT_EHIP, "plt_pass_copy_bounce_records: hipMemcpy failed");
    // path order (8 x 8 pixel blocks) to the row-major pixel order of the layout
    const uint32_t w = p->last_width, h = p->last_height;
    const size_t n = size_t(w) * h, planes = p->last_words / n;
    uint32_t* out = static_cast<uint32_t*>(dst);
    std::vector<uint32_t> pixel(n);
    for (uint32_t g = 0; g < uint32_t(n); ++g) pixel[g] = plt::path_pixel(w, h, g);
    for (size_t k = 0; k < planes; ++k)
        for (size_t g = 0; g < n; ++g) out[k * n + pixel[g]] = paths[k * n + g];
    return PLT_OK;
}

int plt_pass_set_profiling(plt_pass* p, int enable) {
    if (!p) return set_error(PLT_EINVAL, "plt_pass_set_profiling: bad arguments");
    p->profiling = enable != 0;
    p->events_used = 0;
    return PLT_OK;
}

int plt_pass_autotune(plt_pass* p, const plt_render_data* rd, uint32_t frames, plt_stream stream, plt_launch_config* best,
                      double* default_ms, double* best_ms) {
    if (!p || !rd || frames == 0 || frames > 64) return set_error(PLT_EINVAL, "plt_pass_autotune: bad arguments (frames 1..64)");
    if (!p->ctx || !p->scene) return set_error(PLT_ESTATE, "plt_pass_autotune: no device context or scene");
    // the state the candidates' executes advance, restored afterwards
    const uint32_t frame0 = p->frame;
    const plt::SamplePattern pattern0 = p->pattern;
    const bool prof0 = p->profiling;
    p->profiling = false;
    plt_launch_config start{};
    (void)plt_pass_get_launch_config(p, &start);
    std::vector<plt_launch_config> cands;
    for (uint32_t b : {64u, 128u, 256u})
        cands.push_back({PLT_SCHEDULE_MEGAKERNEL, b, start.wf_blocks_per_cu, 0u, start.solve_schedule, start.solve_events_per_path,
                         start.solve_hits_per_path});
    for (uint32_t g : {4u, 8u, 16u, 32u})
        cands.push_back({PLT_SCHEDULE_WAVEFRONT, start.sample_block, g, 0u, start.solve_schedule, start.solve_events_per_path,
                         start.solve_hits_per_path});
    const hipStream_t st = static_cast<hipStream_t>(stream);
    // mean frame time of a configuration, and of its sample pass alone (the pass's kernel timing)
    auto time_config = [&](const plt_launch_config& c, double& ms, double* sample_ms = nullptr) -> int {
        int rc = plt_pass_set_launch_config(p, &c);
        if (rc < 0) return rc;
        p->frame = frame0;
        p->pattern = pattern0;
        p->profiling = false;
        if ((rc = plt_pass_execute(p, rd, stream)) < 0) return rc;  // warm-up (and capture)
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return set_error(PLT_EHIP, "plt_pass_autotune: hipEventCreate failed");
        p->frame = frame0;
        p->pattern = pattern0;
        p->profiling = sample_ms != nullptr;
        p->events_used = 0;
        (void)hipEventRecord(e0, st);
        for (uint32_t f = 0; f < frames && rc >= 0; ++f) rc = plt_pass_execute(p, rd, stream);
        (void)hipEventRecord(e1, st);
        float t = 0.f;
        if (rc >= 0 && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess))
            rc = set_error(PLT_EHIP, "plt_pass_autotune: event timing failed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ms = double(t) / frames;
        if (rc >= 0 && sample_ms) {
            double smp = 0.0, solve = 0.0;
            uint32_t ex = 0;
            rc = plt_pass_get_kernel_times(p, &smp, &solve, &ex);
            *sample_ms = ex ? smp / ex : 0.0;
        }
        p->profiling = false;
        return rc;
    };
    // the sample pass first, by its own time (the solve pass's run-to-run spread would hide its
    // differences), then the solve pass and graph replay by the frame time; the choice is never
    // slower than the starting configuration as measured
    double dms = 0.0, bms = 0.0, dsm = 0.0, bsm = 0.0;
    plt_launch_config bs = start;
    int rc = time_config(start, dms, &dsm);
    bsm = dsm;
    for (const auto& c : cands) {
        if (rc < 0) break;
        double ms = 0.0, sm = 0.0;
        rc = time_config(c, ms, &sm);
        if (rc >= 0 && sm < bsm) { bsm = sm; bs = c; }
    }
    bms = dms;
    plt_launch_config bc = start;
    // the solve pass's schedule and event-area size, with the fastest sample pass
    if (rc >= 0) {
        std::vector<plt_launch_config> solve;
        plt_launch_config c = bs;
        c.solve_schedule = PLT_SOLVE_MEGAKERNEL;
        solve.push_back(c);
        for (uint32_t e : {4u, 8u, 16u}) {
            c.solve_schedule = PLT_SOLVE_SPLIT;
            c.solve_events_per_path = e;
            c.solve_hits_per_path = e / 2;
            solve.push_back(c);
        }
        for (const auto& sc : solve) {
            if (rc < 0) break;
            double ms = 0.0;
            rc = time_config(sc, ms);
            if (rc >= 0 && ms < bms) { bms = ms; bc = sc; }
        }
    }
    if (rc >= 0) {  // graph replay of the best launch sequence
        // Replay saves the launch overhead of about ten launches per frame, well inside the frame
        // time's run-to-run spread, so it is taken only when it is more than 1 % faster: otherwise
        // timing noise alone would decide it.
        plt_launch_config g = bc;
        g.graph = 1u;
        double ms = 0.0;
        rc = time_config(g, ms);
        if (rc >= 0 && ms < 0.99 * bms) { bms = ms; bc = g; }
    }
    p->frame = frame0;
    p->pattern = pattern0;
    p->profiling = prof0;
    p->events_used = 0;
    const int rc2 = plt_pass_set_launch_config(p, rc >= 0 ? &bc : &start);
    if (rc < 0) return rc;
    if (best) *best = bc;
    if (default_ms) *default_ms = dms;
    if (best_ms) *best_ms = bms;
    return rc2;
}

int plt_pass_set_surrogate(plt_pass* p, plt_surrogate* s, int mode) {
    if (!p || (mode != PLT_SURROGATE_NETWORK && mode != PLT_SURROGATE_TEACHER)) return set_error(PLT_EINVAL, "plt_pass_set_surrogate: bad arguments");
    if (s && !s->has_stats)
        return set_error(PLT_ESTATE, "plt_pass_set_surrogate: the surrogate has no training statistics (plt_surrogate_set_stats)");
    p->surrogate = s;
    p->surrogate_teacher = mode == PLT_SURROGATE_TEACHER;
    plt::drop_graphs(p);
    return PLT_OK;
}

int plt_pass_get_surrogate_stats(plt_pass* p, uint64_t* rows, uint64_t* exact_steps) {
    if (!p || !rows || !exact_steps) return set_error(PLT_EINVAL, "plt_pass_get_surrogate_stats: bad arguments");
    if (!p->last_solve.sur.counter) return set_error(PLT_ESTATE, "plt_pass_get_surrogate_stats: no frame rendered with the surrogate");
    uint32_t c[2] = {0u, 0u};
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(c, p->last_solve.sur.counter, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess)
        return set_error(PLT_EHIP, "plt_pass_get_surrogate_stats: read-back failed");
    *rows = c[0];
    *exact_steps = c[1];
    return PLT_OK;
}

int plt_pass_copy_surrogate_queries(plt_pass* p, float* dst, uint64_t max_rows, uint64_t* rows) {
    if (!p || !rows || (max_rows && !dst)) return set_error(PLT_EINVAL, "plt_pass_copy_surrogate_queries: bad arguments");
    const plt::SurPath& sp = p->last_solve.sur;
    if (!sp.counter) return set_error(PLT_ESTATE, "plt_pass_copy_surrogate_queries: no frame rendered with the surrogate");
    uint32_t c = 0;
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(&c, sp.counter, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess)
        return set_error(PLT_EHIP, "plt_pass_copy_surrogate_queries: read-back failed");
    const uint64_t n = std::min<uint64_t>(std::min<uint64_t>(c, sp.qcap), max_rows);
    if (n && hipMemcpy(dst, sp.X, n * PLT_SURROGATE_INPUTS * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return set_error(PLT_EHIP, "plt_pass_copy_surrogate_queries: read-back failed");
    *rows = n;
    return PLT_OK;
}

int plt_pass_get_solve_stats(plt_pass* p, uint64_t* events, uint64_t* walk_vertices, uint64_t* event_vertices,
                              uint64_t* overflowed_paths, uint64_t* walks, uint64_t* failed_walks, uint64_t* walk_rays) {
    if (!p || !events || !walk_vertices || !event_vertices || !overflowed_paths || !walks || !failed_walks || !walk_rays)
        return set_error(PLT_EINVAL, "plt_pass_get_solve_stats: bad arguments");
    const plt::SolveBuffers& sb = p->last_solve;
    if (!sb.ev || p->last_regions == 0) return set_error(PLT_ESTATE, "plt_pass_get_solve_stats: no split solve pass rendered yet");
    std::vector<uint32_t> reg(p->last_regions * plt::kRegWords), meta(sb.ev_cap);
    uint32_t wc[3] = {0u, 0u, 0u};
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(reg.data(), sb.region, reg.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(meta.data(), sb.ev + size_t(plt::kEvMeta) * sb.ev_cap, meta.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(wc, sb.walk.counters, sizeof(wc), hipMemcpyDeviceToHost) != hipSuccess)
        return set_error(PLT_EHIP, "plt_pass_get_solve_stats: read-back failed");
    uint64_t ne = 0, nh = 0, nv = 0, no = 0, nf = 0;
    for (size_t r = 0; r < p->last_regions; ++r) {
        const uint32_t* rg = &reg[r * plt::kRegWords];
        ne += rg[plt::kRegEvents];
        nh += rg[plt::kRegHits];
        no += uint64_t(__builtin_popcount(rg[plt::kRegOverflowLo])) + uint64_t(__builtin_popcount(rg[plt::kRegOverflowHi]));
        const size_t base = r * plt::kSolveRegion * sb.ev_per_path;
        for (uint32_t k = 0; k < rg[plt::kRegEvents]; ++k) {
            nv += ((meta[base + k] >> 6) & 127u) + 1u;
            nf += ((meta[base + k] >> 13) & 3u) == plt::kEvKindVoid;
        }
    }
    *events = ne;
    *walk_vertices = nh;
    *event_vertices = nv;
    *overflowed_paths = no;
    *walks = wc[0];
    *failed_walks = nf;
    *walk_rays = wc[2];
    return PLT_OK;
}

int plt_pass_get_kernel_times_split(plt_pass* p, double* sample_ms, double* light_ms, double* walk_ms, double* unwind_ms,
                                    double* megakernel_ms, uint32_t* executes) {
    if (!p || !sample_ms || !light_ms || !walk_ms || !unwind_ms || !megakernel_ms || !executes)
        return set_error(PLT_EINVAL, "plt_pass_get_kernel_times: bad arguments");
    constexpr size_t kEv = plt::kEventsPerExecute;
    double sums[kEv - 1] = {};
    const size_t n = p->events_used / kEv;
    for (size_t i = 0; i < n; ++i) {
        hipEvent_t* e = &p->events[kEv * i];
        if (hipEventSynchronize(e[kEv - 1]) != hipSuccess) return set_error(PLT_EHIP, "plt_pass_get_kernel_times: hipEventSynchronize failed");
        for (size_t k = 0; k + 1 < kEv; ++k) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, e[k], e[k + 1]) != hipSuccess)
                return set_error(PLT_EHIP, "plt_pass_get_kernel_times: hipEventElapsedTime failed");
            sums[k] += t;
        }
    }
    *sample_ms = sums[0];
    *light_ms = sums[1];
    *walk_ms = sums[2];
    *unwind_ms = sums[3];
    *megakernel_ms = sums[4];
    *executes = uint32_t(n);
    p->events_used = 0;
    return PLT_OK;
}

int plt_pass_get_kernel_times(plt_pass* p, double* sample_ms, double* solve_ms, uint32_t* executes) {
    if (!p || !sample_ms || !solve_ms || !executes) return set_error(PLT_EINVAL, "plt_pass_get_kernel_times: bad arguments");
    double light = 0.0, walk = 0.0, unwind = 0.0, mega = 0.0;
    const int rc = plt_pass_get_kernel_times_split(p, sample_ms, &light, &walk, &unwind, &mega, executes);
    if (rc < 0) return rc;
    *solve_ms = light + walk + unwind + mega;
    return PLT_OK;
}

}  // extern "C"
