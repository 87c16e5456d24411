// The split solve pass's deferred manifold walks: the persistent walk kernel over the frame's walk
// queue, sized to fill every CU (12 one-wave blocks per CU: three waves per SIMD, each with up to
// 168 VGPRs), and the clear of the queue's counters. The kernel's time is the latency chain of its
// longest walks, so it trades occupancy for registers: at four waves per SIMD (128 VGPRs) it
// spilled, and every Newton iteration paid the spill reloads' round trips (1080p frame, 5.24M triangles: 3.85 ms at
// four waves, 3.48 ms at three, 3.59 ms at two).
#include "pathtracer_split.hpp"

namespace plt {
namespace {

// ---------------------------------------------------------------------------------------
// Deferred manifold walks: manifoldNEESample's Newton iterations and the light's visibility
// (Manifolds.slangh:183-281), the same steps as ms_nee_sample after its occluder trace.
// ---------------------------------------------------------------------------------------
constexpr int kWalkBlock = 64;
enum : int { kWalkIdle = 0, kWalkNewton, kWalkLine, kWalkVis };

PLT_DEV V3 hit_pos(const SolveBuffers& sb, const uint32_t* h) {
    return v3(__uint_as_float(h[kHitPos * sb.hit_cap]), __uint_as_float(h[(kHitPos + 1) * sb.hit_cap]),
              __uint_as_float(h[(kHitPos + 2) * sb.hit_cap]));
}

// A persistent grid: each lane runs one walk at a time as a sequence of ray casts, and every trip
// through the loop casts at most one ray per lane. A lane whose walk ends takes the next one from
// the queue at the top of the next trip, so a wave's lanes stay busy while its longest walk
// iterates (running whole walks per lane would hold every lane of a wave until its longest walk,
// up to MNEEMaxIterations casts, had ended). A walk is deterministic given its starting vertices,
// so which lane runs it, and when, changes nothing in its result.
//
// Every lane ends: a lane retires once the queue is empty, and a walk ends after at most
// MNEEMaxIterations line-search casts, one visibility cast and, between casts, at most two passes
// over its vertices (a pass that casts nothing either meets the threshold at every vertex, which
// ends the walk, or follows a pass that cast).
__global__ void __launch_bounds__(kWalkBlock) __attribute__((amdgpu_waves_per_eu(3, 3)))
solve_walk_kernel(PtScene sc, PtFrame fr, SolveBuffers sb) {
    __shared__ int stack[kTraversalStackDepth * kWalkBlock];
    __shared__ float betaS[kMaxMsOccluders * kWalkBlock];  // per-vertex step sizes, [vertex][lane]
    const uint32_t lane = threadIdx.x;
    int* stk = stack + lane;
    float* beta = betaS + lane;
    const uint32_t total = sb.walk.counters[0];
    const size_t regionEvents = size_t(kPtBlock) * sb.ev_per_path, regionHits = size_t(kPtBlock) * sb.hits_per_path;
    const int maxIt = int(fr.mnee_max_iterations);
    int state = kWalkIdle;
    bool retired = false;
    size_t e = 0;
    uint32_t meta = 0;
    uint32_t* hb = nullptr;  // the walk's first vertex in the vertex planes
    int count = 0, idx = 0, iterations = 0;
    bool passOk = true, distant = false;
    float wavelength = 0.f;
    V3 origin = v3(0.f, 0.f, 0.f), lightPos = v3(0.f, 0.f, 0.f), lightDir = v3(0.f, 0.f, 0.f);
    V3 hpos = v3(0.f, 0.f, 0.f), stepDir = v3(0.f, 0.f, 0.f);  // line search: the vertex and its Newton step
    uint32_t hmesh = 0;
    Ray ray{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    uint32_t casts = 0;
    auto finish = [&](bool ok) {
        if (ok) ev_put3(sb, e, kEvWo, v3(ray.dx, ray.dy, ray.dz));
        else ev_put(sb, e, kEvMeta, (meta & ~(3u << 13)) | (kEvKindVoid << 13));
        state = kWalkIdle;
    };
    for (;;) {
        if (state == kWalkIdle && !retired) {
            const uint32_t k = wave_append(sb.walk.counters + 1, 1u);
            if (k >= total) {
                retired = true;
            } else {
                e = sb.walk.items[k];
                meta = ev_get(sb, e, kEvMeta);
                count = int((meta >> 21) & 15u);
                distant = (meta & kEvBitDistant) != 0;
                const size_t region = e / regionEvents;
                hb = sb.hits + region * regionHits + ev_get(sb, e, kEvHitBase);
                origin = ev_get3(sb, e, kEvOrg);
                lightDir = -ev_get3(sb, e, kEvDir);  // the event stores the beam's direction, -lsDir
                lightPos = distant ? v3(0.f, 0.f, 0.f) : ev_get3(sb, e, kEvPos);
                wavelength = ev_getf(sb, e, kEvLam);
                for (int i = 0; i < count; ++i) beta[i * kWalkBlock] = 1.f;
                idx = 0;
                iterations = 0;
                passOk = true;
                state = kWalkNewton;
            }
        }
        if (__ballot(state != kWalkIdle) == 0) break;  // every lane has retired (the whole wave leaves together)
        // advance each walk to its next ray
        bool cast = false;
        if (state == kWalkNewton) {
            for (;;) {
                if (idx == count) {
                    if (!passOk) {  // another pass over the vertices
                        passOk = true;
                        idx = 0;
                        continue;
                    }
                    // every vertex within the threshold: the light's visibility from the last one
                    const V3 p = hit_pos(sb, hb + (count - 1));
                    V3 dir;
                    float dist;
                    if (!distant) {
                        dir = lightPos - p;
                        dist = wlength(dir);
                        dir = wdiv(dir, dist);
                    } else {
                        dir = lightDir;
                        dist = kFltMax;
                    }
                    ray = Ray{p.x, p.y, p.z, dir.x, dir.y, dir.z, 1e-4f, dist};
                    state = kWalkVis;
                    cast = true;
                    break;
                }
                const uint32_t* h = hb + idx;
                const uint32_t tri = h[kHitTri * sb.hit_cap];
                const DSG dsg = dsg_at(sc, tri, __uint_as_float(h[kHitB1 * sb.hit_cap]), __uint_as_float(h[kHitB2 * sb.hit_cap]));
                const V3 p0 = idx == 0 ? origin : hit_pos(sb, h - 1);
                const PDV pdv0 = dsg.pdw(p0);
                const bool last = idx + 1 == count;
                const V3 p2 = last ? lightPos : hit_pos(sb, h + 1);
                const PDV pdv2 = last && distant ? PDV{lightDir, v3(0.f, 0.f, 0.f), v3(0.f, 0.f, 0.f)} : dsg.pdw(p2);
                const bool woIsIn = wdot(pdv2.v, dsg.n.v) < 0.f;
                const PDV wi = woIsIn ? pdv0 : pdv2, wo = woIsIn ? pdv2 : pdv0;
                hmesh = sc.tri_mesh[tri];
                const Mat m = material_at(sc, sc.mesh_material[hmesh]);
                const PDV wos = mat_partial_derivatives(m, wi, dsg, h[kHitLobe * sb.hit_cap], wavelength);
                float lenC;
                if ((wos.v.x == 0.f && wos.v.y == 0.f && wos.v.z == 0.f) || !manifold_one_step(wo, wos, dsg.p, lenC, stepDir)) {
                    finish(false);
                    break;
                }
                if (lenC < fr.mnee_threshold) {
                    ++idx;
                    continue;
                }
                passOk = false;
                hpos = hit_pos(sb, h);
                ray.ox = p0.x;
                ray.oy = p0.y;
                ray.oz = p0.z;
                state = kWalkLine;
                break;
            }
        }
        if (state == kWalkLine) {  // one step of the line search from p0 = the ray's origin
            if (++iterations > maxIt) {
                finish(false);
            } else {
                const V3 p0 = v3(ray.ox, ray.oy, ray.oz);
                V3 proposal;
                {
#pragma clang fp contract(off)
                    proposal = hpos - wscale(stepDir, beta[idx * kWalkBlock]);
                }
                V3 dirTo = proposal - p0;
                const float dist = wlength(dirTo);
                dirTo = wdiv(dirTo, dist);
                ray = Ray{p0.x, p0.y, p0.z, dirTo.x, dirTo.y, dirTo.z, fmaxf(1e-4f, .99f * dist), 1.01f * dist};
                cast = true;
            }
        }
        if (cast) {
            ++casts;
            const Hit ch = traverse_rays<kTraversalStackDepth>(sc.tsc, ray, stk, kWalkBlock, state == kWalkVis);
            if (state == kWalkVis) {
                finish(ch.tri == kNoHit);
            } else if (ch.tri == kNoHit || sc.tri_mesh[ch.tri] != hmesh) {
                beta[idx * kWalkBlock] /= 4.f;
            } else {
                uint32_t* h = hb + idx;
                const V3 pos = walk_point(v3(ray.ox, ray.oy, ray.oz), v3(ray.dx, ray.dy, ray.dz), ch.t);
                h[kHitTri * sb.hit_cap] = ch.tri;
                h[kHitB1 * sb.hit_cap] = __float_as_uint(ch.b1);
                h[kHitB2 * sb.hit_cap] = __float_as_uint(ch.b2);
                h[kHitPos * sb.hit_cap] = __float_as_uint(pos.x);
                h[(kHitPos + 1) * sb.hit_cap] = __float_as_uint(pos.y);
                h[(kHitPos + 2) * sb.hit_cap] = __float_as_uint(pos.z);
                beta[idx * kWalkBlock] = fminf(1.f, 2.f * beta[idx * kWalkBlock]);
                ++idx;
                state = kWalkNewton;
            }
        }
    }
    // the wave's ray casts, one atomic per wave
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) casts += __shfl_xor(casts, o);
    if (lane == 0) atomicAdd(sb.walk.counters + 2, casts);
}

// Clears the walk queue's counters (a kernel, so a captured graph replays it).
__global__ void walk_clear_kernel(uint32_t* counters) {
    if (threadIdx.x < 3) counters[threadIdx.x] = 0u;
}

}  // namespace

int launch_walk_clear(const SolveBuffers& sb, hipStream_t st) {
    hipLaunchKernelGGL(walk_clear_kernel, dim3(1), dim3(64), 0, st, sb.walk.counters);
    return check_launch("walk_clear_kernel");
}

int launch_solve_walks(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, hipStream_t st) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        return set_error(PLT_EHIP, "solve_walk_kernel: device query failed");
    hipLaunchKernelGGL(solve_walk_kernel, dim3(uint32_t(cus) * 12u), dim3(kWalkBlock), 0, st, sc, fr, sb);
    return check_launch("solve_walk_kernel");
}

}  // namespace plt
