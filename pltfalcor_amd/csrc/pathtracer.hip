// Sample pass of the PLT path tracer: the megakernel and the wavefront schedule, and their launchers.
#include "pathtracer_kernels.hpp"

namespace plt {
namespace {


// Megakernel schedule: one lane follows its path to the end. B threads per block (64, 128 or 256:
// a launch parameter the pass's autotuner picks), each with a traversal stack of
// kTraversalStackDepth entries interleaved in LDS.
template <int B>
__global__ void __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(4, 4))) sample_kernel(PtScene sc, PtFrame fr) {
    __shared__ int stack[kTraversalStackDepth * B];
    const uint32_t li = blockIdx.x * B + threadIdx.x;
    if (li >= fr.n_paths) return;
    const uint32_t pid = path_pixel(fr.width, fr.height, fr.path_base + li);
    const uint32_t px = pid % fr.width, py = pid / fr.width;
    int* stk = stack + threadIdx.x;
    Rng sg = Rng::make(px, py, fr.frame, fr.sample_generator == 0);
    PathState ps = start_path(sc, fr, px, py, pid, sg);
    while (ps.len < fr.max_bounces + 1) {
        const Ray ray{ps.org.x, ps.org.y, ps.org.z, ps.dir.x, ps.dir.y, ps.dir.z, 0.f, kFltMax};
        const Hit h = traverse<false, kTraversalStackDepth>(sc.tsc, ray, stk, B);
        if (!sample_vertex(sc, fr, h, li, pid, ps, sg)) break;
    }
}

// Clears the wavefront counters (a kernel, not a memset node, so the sequence can be replayed from
// a captured graph).
__global__ void wf_clear_kernel(uint32_t* counters) {
    if (threadIdx.x < kWfCounters) counters[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------------------------------
// Wavefront schedule of the sample pass: per bounce, a traversal-only kernel over the queue of
// live paths, a counting sort of the hits by material type (the sort key; misses last), and a
// shading kernel over the material-sorted queue that runs the same per-vertex code and appends the
// surviving paths to the next queue. Each path's random stream lives in its state, so every path
// draws the same numbers in the same order as in the megakernel: the records are identical.
// Every kernel is launched for the largest possible queue and reads the live count from device
// memory, so the host never waits between bounces.
// ---------------------------------------------------------------------------------------
constexpr int kWfBlock = 256;
constexpr uint32_t kMissKey = kWfKeys - 1;

PLT_DEV void store_path(const WfBuffers& w, uint32_t li, const PathState& ps, const Rng& sg) {
    w.rng[li] = make_uint4(sg.s0, sg.s1, sg.s2, sg.s3);
    w.org[li] = make_float4(ps.org.x, ps.org.y, ps.org.z, ps.lambda);
    w.dir[li] = make_float4(ps.dir.x, ps.dir.y, ps.dir.z, ps.weight);
    w.pdfs[li] = make_float4(ps.pdf, ps.thp, ps.lastPdf, __uint_as_float(ps.flags));
    w.len[li] = ps.len;
}
PLT_DEV void load_path(const WfBuffers& w, uint32_t li, bool tiny, PathState& ps, Rng& sg) {
    const uint4 r = w.rng[li];
    sg = Rng::from_state(r, tiny);
    const float4 o = w.org[li], d = w.dir[li], p = w.pdfs[li];
    ps.org = v3(o.x, o.y, o.z);
    ps.lambda = o.w;
    ps.dir = v3(d.x, d.y, d.z);
    ps.weight = d.w;
    ps.pdf = p.x;
    ps.thp = p.y;
    ps.lastPdf = p.z;
    ps.flags = __float_as_uint(p.w);
    ps.len = w.len[li];
}

__global__ void __launch_bounds__(kWfBlock) wf_start_kernel(PtScene sc, PtFrame fr, WfBuffers w) {
    const uint32_t li = blockIdx.x * kWfBlock + threadIdx.x;
    if (li == 0) w.counters[kWfCount0] = fr.n_paths;
    if (li >= fr.n_paths) return;
    const uint32_t pid = path_pixel(fr.width, fr.height, fr.path_base + li);
    const uint32_t px = pid % fr.width, py = pid / fr.width;
    Rng sg = Rng::make(px, py, fr.frame, fr.sample_generator == 0);
    const PathState ps = start_path(sc, fr, px, py, pid, sg);
    store_path(w, li, ps, sg);
    w.queue[li] = li;
}

__global__ void __launch_bounds__(kPtBlock) wf_trace_kernel(PtScene sc, PtFrame fr, WfBuffers w, int cur) {
    __shared__ int stack[kTraversalStackDepth * kPtBlock];
    __shared__ uint32_t hist[kWfKeys];
    if (threadIdx.x < kWfKeys) hist[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t count = w.counters[kWfCount0 + cur];
    const uint32_t* q = w.queue + size_t(cur) * fr.n_paths;
    for (uint32_t i = blockIdx.x * kPtBlock + threadIdx.x; i < count; i += gridDim.x * kPtBlock) {
        const uint32_t li = q[i];
        const float4 o = w.org[li], d = w.dir[li];
        const Ray ray{o.x, o.y, o.z, d.x, d.y, d.z, 0.f, kFltMax};
        const Hit h = traverse<false, kTraversalStackDepth>(sc.tsc, ray, stack + threadIdx.x, kPtBlock);
        w.hit[li] = make_float4(__uint_as_float(h.tri), h.b1, h.b2, h.t);
        const uint32_t key = h.tri == kNoHit ? kMissKey : min(sc.materials[sc.mesh_material[sc.tri_mesh[h.tri]]].type, kMissKey - 1);
        w.keys[i] = key;
        atomicAdd(&hist[key], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kWfKeys && hist[threadIdx.x]) atomicAdd(&w.counters[kWfHist + threadIdx.x], hist[threadIdx.x]);
}

// Exclusive prefix sum of the key histogram into the scatter cursors; clears the histogram and the
// next queue's count.
__global__ void wf_scan_kernel(WfBuffers w, int cur) {
    if (threadIdx.x != 0) return;
    uint32_t acc = 0;
    for (uint32_t k = 0; k < kWfKeys; ++k) {
        w.counters[kWfCursor + k] = acc;
        acc += w.counters[kWfHist + k];
        w.counters[kWfHist + k] = 0u;
    }
    w.counters[kWfCount0 + (cur ^ 1)] = 0u;
}

__global__ void __launch_bounds__(kWfBlock) wf_scatter_kernel(PtFrame fr, WfBuffers w, int cur) {
    __shared__ uint32_t local[kWfKeys], base[kWfKeys];
    const uint32_t count = w.counters[kWfCount0 + cur];
    // block-uniform loop: every thread of the block takes part in each round's barriers
    for (uint32_t b0 = blockIdx.x * kWfBlock; b0 < count; b0 += gridDim.x * kWfBlock) {
        if (threadIdx.x < kWfKeys) local[threadIdx.x] = 0u;
        __syncthreads();
        const uint32_t i = b0 + threadIdx.x;
        uint32_t key = 0, rank = 0;
        const bool live = i < count;
        if (live) {
            key = w.keys[i];
            rank = atomicAdd(&local[key], 1u);
        }
        __syncthreads();
        if (threadIdx.x < kWfKeys) base[threadIdx.x] = local[threadIdx.x] ? atomicAdd(&w.counters[kWfCursor + threadIdx.x], local[threadIdx.x]) : 0u;
        __syncthreads();
        if (live) w.sorted[base[key] + rank] = w.queue[size_t(cur) * fr.n_paths + i];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(4, 4))) wf_shade_kernel(PtScene sc, PtFrame fr,
                                                                                                       WfBuffers w, int cur) {
    const uint32_t count = w.counters[kWfCount0 + cur];
    const int lane = int(threadIdx.x & 63u);
    // wave-uniform loop (one wave per block), so the ballot below sees every lane
    for (uint32_t b0 = blockIdx.x * kPtBlock; b0 < count; b0 += gridDim.x * kPtBlock) {
        const uint32_t i = b0 + threadIdx.x;
        bool cont = false;
        uint32_t li = 0;
        if (i < count) {
            li = w.sorted[i];
            PathState ps;
            Rng sg;
            load_path(w, li, fr.sample_generator == 0, ps, sg);
            const float4 hv = w.hit[li];
            const Hit h{__float_as_uint(hv.x), hv.y, hv.z, hv.w};
            cont = sample_vertex(sc, fr, h, li, path_pixel(fr.width, fr.height, fr.path_base + li), ps, sg) && ps.len < fr.max_bounces + 1;
            if (cont) store_path(w, li, ps, sg);
        }
        // append the surviving paths to the next queue, one atomic per wave
        const uint64_t mask = __ballot(cont);
        const int leader = mask ? __ffsll(static_cast<unsigned long long>(mask)) - 1 : 0;
        uint32_t wbase = 0;
        if (mask && lane == leader) wbase = atomicAdd(&w.counters[kWfCount0 + (cur ^ 1)], uint32_t(__popcll(mask)));
        wbase = __shfl(wbase, leader);
        if (cont) {
            const uint32_t pos = wbase + uint32_t(__popcll(mask & ((uint64_t(1) << lane) - 1u)));
            w.queue[size_t(cur ^ 1) * fr.n_paths + pos] = li;
        }
    }
}

}  // namespace

int launch_sample_pass(const PtScene& sc, const PtFrame& fr, const LaunchConfig& lc, hipStream_t st) {
    if (fr.n_paths == 0) return PLT_OK;
    const uint32_t n = fr.n_paths;
    switch (lc.sample_block) {
        case 64: hipLaunchKernelGGL(sample_kernel<64>, dim3((n + 63) / 64), dim3(64), 0, st, sc, fr); break;
        case 128: hipLaunchKernelGGL(sample_kernel<128>, dim3((n + 127) / 128), dim3(128), 0, st, sc, fr); break;
        case 256: hipLaunchKernelGGL(sample_kernel<256>, dim3((n + 255) / 256), dim3(256), 0, st, sc, fr); break;
        default: return set_error(PLT_EINVAL, "sample pass: block size must be 64, 128 or 256");
    }
    return check_launch("sample_kernel");
}
int launch_sample_pass_wavefront(const PtScene& sc, const PtFrame& fr, const WfBuffers& w, const LaunchConfig& lc, hipStream_t st) {
    if (fr.n_paths == 0) return PLT_OK;
    if (lc.wf_blocks_per_cu < 1 || lc.wf_blocks_per_cu > 64) return set_error(PLT_EINVAL, "wavefront sample pass: blocks per CU in [1, 64]");
    hipLaunchKernelGGL(wf_clear_kernel, dim3(1), dim3(64), 0, st, w.counters);
    const uint32_t n = fr.n_paths;
    hipLaunchKernelGGL(wf_start_kernel, dim3((n + kWfBlock - 1) / kWfBlock), dim3(kWfBlock), 0, st, sc, fr, w);
    // persistent grids sized to the resident capacity (the kernels stride over the live queue):
    // wf_blocks_per_cu 64-thread blocks per CU for traversal and shading, a quarter as many
    // 256-thread blocks for the scatter
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint32_t g64 = std::min((n + kPtBlock - 1) / kPtBlock, uint32_t(lc.wf_blocks_per_cu * cus));
    const uint32_t g256 = std::min((n + kWfBlock - 1) / kWfBlock, uint32_t(std::max(1u, lc.wf_blocks_per_cu / 2) * cus));
    for (uint32_t b = 0; b <= fr.max_bounces; ++b) {
        const int cur = int(b & 1u);
        hipLaunchKernelGGL(wf_trace_kernel, dim3(g64), dim3(kPtBlock), 0, st, sc, fr, w, cur);
        hipLaunchKernelGGL(wf_scan_kernel, dim3(1), dim3(64), 0, st, w, cur);
        hipLaunchKernelGGL(wf_scatter_kernel, dim3(g256), dim3(kWfBlock), 0, st, fr, w, cur);
        hipLaunchKernelGGL(wf_shade_kernel, dim3(g64), dim3(kPtBlock), 0, st, sc, fr, w, cur);
    }
    return check_launch("wavefront sample pass");
}

// The per-tile generator of NEEUsePerTileSG (pltpt_solve.rt.slang:412): SampleGenerator(kTile,
// gFrameCount << 2) of each tile, its first max_bounces + 1 draws (at most one light-type choice per
// path vertex), so the solve kernel reads the k-th draw instead of carrying a second generator.
__global__ void tile_draws_kernel(uint32_t tiles_x, uint32_t tiles, uint32_t frame, uint32_t per_tile, bool tiny,
                                  float* draws) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tiles) return;
    Rng g = Rng::make(t % tiles_x, t / tiles_x, frame << 2, tiny);
    for (uint32_t k = 0; k < per_tile; ++k) draws[size_t(t) * per_tile + k] = g.next1D();
}

int launch_tile_draws(const PtFrame& fr, uint32_t tiles, float* draws, hipStream_t st) {
    hipLaunchKernelGGL(tile_draws_kernel, dim3((tiles + 63) / 64), dim3(64), 0, st, fr.tiles_x, tiles, fr.frame, fr.max_bounces + 1,
                       fr.sample_generator == 0, draws);
    return check_launch("tile_draws_kernel");
}

int launch_solve_4(const PtScene& sc, const PtFrame& fr, hipStream_t st);
int launch_solve_4_lights(const PtScene& sc, const PtFrame& fr, hipStream_t st);
int launch_solve_1(const PtScene& sc, const PtFrame& fr, hipStream_t st, bool lights);
int launch_solve_sample_4(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, hipStream_t st);
int launch_solve_sample_4_lights(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, hipStream_t st);
int launch_solve_sample_1(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, hipStream_t st, bool lights);
int launch_solve_unwind(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, hipStream_t st);
int launch_walk_clear(const SolveBuffers& sb, hipStream_t st);
int launch_solve_walk_unwind(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, hipStream_t st);
int launch_solve_sum(const PtFrame& fr, const SolveBuffers& sb, hipStream_t st);
int launch_solve_walks(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, hipStream_t st);
int launch_solve_unwind_surrogate(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, plt_surrogate* s, hipStream_t st);

int launch_solve_pass(const PtScene& sc, const PtFrame& fr, hipStream_t st) {
    if (fr.n_paths == 0) return PLT_OK;
    const bool analytic = ((fr.options & kOptAnalyticLights) != 0 && sc.n_analytic > 0) ||
                          ((fr.options & kOptEnvLight) != 0 && sc.env_w > 0);
    if (fr.hwss == 4) return analytic ? launch_solve_4_lights(sc, fr, st) : launch_solve_4(sc, fr, st);
    if (fr.hwss == 1) return launch_solve_1(sc, fr, st, analytic);
    return set_error(PLT_ENOTIMPL, "HWSS sample counts other than 1 and 4 are not built");
}

// Split solve pass: light sampling, unwind, then the megakernel over the paths whose region's event
// area overflowed (its other blocks return at once). Three launches on one stream, no host waits.
int launch_solve_pass_split(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, hipStream_t st, hipEvent_t after_light,
                            hipEvent_t after_walk, hipEvent_t after_unwind, plt_surrogate* surrogate, const SplitFork& fork) {
    if (fr.n_paths == 0) {
        if (after_light) (void)hipEventRecord(after_light, st);
        if (after_walk) (void)hipEventRecord(after_walk, st);
        if (after_unwind) (void)hipEventRecord(after_unwind, st);
        return PLT_OK;
    }
    if (fr.hwss != 4 && fr.hwss != 1) return set_error(PLT_ENOTIMPL, "HWSS sample counts other than 1 and 4 are not built");
    const bool analytic = ((fr.options & kOptAnalyticLights) != 0 && sc.n_analytic > 0) ||
                          ((fr.options & kOptEnvLight) != 0 && sc.env_w > 0);
    int rc = launch_walk_clear(sb, st);
    if (rc < 0) return rc;
    rc = fr.hwss == 4 ? (analytic ? launch_solve_sample_4_lights(sc, fr, sb, st) : launch_solve_sample_4(sc, fr, sb, st))
                      : launch_solve_sample_1(sc, fr, sb, st, analytic);
    if (rc < 0) return rc;
    if (after_light) (void)hipEventRecord(after_light, st);
    // manifold walks: only frames with manifold NEE queue any, but the kernel is cheap to launch
    // empty and a captured graph then holds the same launches whatever the options
    if (surrogate || !fork.aux) {
        if ((rc = launch_solve_walks(sc, fr, sb, st)) < 0) return rc;
        if (after_walk) (void)hipEventRecord(after_walk, st);
        if ((rc = surrogate ? launch_solve_unwind_surrogate(sc, fr, sb, surrogate, st) : launch_solve_unwind(sc, fr, sb, st)) < 0)
            return rc;
    } else {
        // The walk kernel's last walks keep a few waves busy for milliseconds: the events that are
        // not walks are unwound beside it on the second stream (their contributions stored), the
        // walks that succeeded after it, and then every path's contributions are summed in order.
        SolveBuffers staged = sb;
        staged.stage = 1u;
        if (hipEventRecord(fork.fork, st) != hipSuccess || hipStreamWaitEvent(fork.aux, fork.fork, 0) != hipSuccess)
            return set_error(PLT_EHIP, "split solve pass: fork to the second stream failed");
        if ((rc = launch_solve_unwind(sc, fr, staged, fork.aux)) < 0) return rc;
        if (hipEventRecord(fork.join, fork.aux) != hipSuccess) return set_error(PLT_EHIP, "split solve pass: join event failed");
        if ((rc = launch_solve_walks(sc, fr, sb, st)) < 0) return rc;
        if (after_walk) (void)hipEventRecord(after_walk, st);
        if ((rc = launch_solve_walk_unwind(sc, fr, sb, st)) < 0) return rc;
        if (hipStreamWaitEvent(st, fork.join, 0) != hipSuccess) return set_error(PLT_EHIP, "split solve pass: join failed");
        if ((rc = launch_solve_sum(fr, sb, st)) < 0) return rc;
    }
    if (after_unwind) (void)hipEventRecord(after_unwind, st);
    PtFrame fb = fr;
    fb.solve_overflow = sb.region;
    return launch_solve_pass(sc, fb, st);
}

}  // namespace plt
