// Split solve pass of the PLT path tracer (pltpt_solve.rt.slang:37-476 in two kernels).
//
// solve_sample_kernel replays each path's vertices with its random stream exactly as the
// megakernel does -- light-type choice, light sample, shadow ray or manifold walk (NEE,
// pltpt_solve.rt.slang:245-366), emissive-hit and environment-miss sourcing (:152-200, :371-396) --
// but instead of carrying each sourced beam through the stored vertices it appends a light event:
// what the beam is (spectral radiance, wavelengths, position, direction, size), the scale factors
// the megakernel applies to it in order (type selection, light pdf, walk pdf) and, for a manifold
// walk, its vertices. Of a manifold walk it runs only the part that draws random numbers (the
// delta-lobe samples at the current vertex and at the occluders of the shadow ray) and queues the
// event for the walk kernel (pathtracer_walk.hip), which runs the Newton iterations and the light's
// visibility and completes the event or voids it. It holds the traversal stack but no beam and no
// BSDF evaluation.
//
// solve_unwind_kernel rebuilds each event's beam and runs the rest of the megakernel's sequence
// for it -- MIS weight, walk unwind (manifoldForwardUnwind, :80-93), evalNEE (:219-243),
// forwardUnwind (:48-79), measure (:96-147) -- and sums each path's contributions in the order the
// megakernel adds them (the events of a path are stored in its own vertex order), so the colour is
// the megakernel's bit for bit. It holds the beam and the material evaluations but no traversal.
//
// A region (64 consecutive paths, one wave) has room for 64 * ev_per_path events and 64 *
// hits_per_path walk vertices. A path whose events do not fit is flagged in the region's overflow
// mask; the unwind kernel skips it and the megakernel, launched after it with the mask, solves it.
#pragma once
#include "pathtracer_kernels.hpp"
#include "surrogate_features.hpp"

namespace plt {
namespace {

static_assert(kPtBlock == int(kSolveRegion), "a region of the split solve pass is one block of the path kernels");

PLT_DEV void ev_put(const SolveBuffers& sb, size_t e, int f, uint32_t v) { sb.ev[size_t(f) * sb.ev_cap + e] = v; }
PLT_DEV void ev_putf(const SolveBuffers& sb, size_t e, int f, float v) { ev_put(sb, e, f, __float_as_uint(v)); }
PLT_DEV void ev_put3(const SolveBuffers& sb, size_t e, int f, V3 v) {
    ev_putf(sb, e, f, v.x);
    ev_putf(sb, e, f + 1, v.y);
    ev_putf(sb, e, f + 2, v.z);
}
// Event and bounce-record reads are non-temporal: each word is read once, by the unwind kernels,
// which run beside the walk kernel on a second stream. Measured per 1080p frame of the
// 5.24M-triangle scene (mean of three interleaved runs): walk kernel 3.54 -> 3.47 ms with its L2
// fetch unchanged; the staged unwind's fetch rose 7.6 % and its writes fell 10 %.
PLT_DEV uint32_t ev_get(const SolveBuffers& sb, size_t e, int f) { return __builtin_nontemporal_load(sb.ev + size_t(f) * sb.ev_cap + e); }
PLT_DEV float ev_getf(const SolveBuffers& sb, size_t e, int f) { return __uint_as_float(ev_get(sb, e, f)); }
PLT_DEV V3 ev_get3(const SolveBuffers& sb, size_t e, int f) { return v3(ev_getf(sb, e, f), ev_getf(sb, e, f + 1), ev_getf(sb, e, f + 2)); }

// The beam every event kind sources: spectral radiance and wavelengths, position, direction, size.
template <int N>
PLT_DEV void ev_put_source(const SolveBuffers& sb, size_t e, const float* Le, const float* lam, V3 pos, V3 dir, float size, float dist) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
        ev_putf(sb, e, kEvLe + j, Le[j]);
        ev_putf(sb, e, kEvLam + j, lam[j]);
    }
    ev_put3(sb, e, kEvPos, pos);
    ev_put3(sb, e, kEvDir, dir);
    ev_putf(sb, e, kEvSize, size);
    ev_putf(sb, e, kEvDist, dist);
}

// `rows` consecutive entries of a counted area for each active lane, one atomic per wave.
PLT_DEV uint32_t wave_append(uint32_t* counter, uint32_t rows) {
    const uint64_t act = __ballot(1);
    const int lane = int(__lane_id());
    const int leader = __ffsll(static_cast<unsigned long long>(act)) - 1;
    const uint32_t rank = uint32_t(__popcll(act & ((uint64_t(1) << lane) - 1u)));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, rows * uint32_t(__popcll(act)));
    base = __shfl(base, leader);
    return base + rank * rows;
}

template <int N, bool Lights>
__global__ void __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(4, 4)))
solve_sample_kernel(PtScene sc, PtFrame fr, SolveBuffers sb) {
    __shared__ int stack[kTraversalStackDepth * kPtBlock];
    __shared__ uint32_t counters[2];  // events and walk vertices appended by this region
    const uint32_t lane = threadIdx.x, region = blockIdx.x;
    if (lane == 0) {
        counters[0] = 0u;
        counters[1] = 0u;
    }
    __syncthreads();
    const uint32_t li = region * kPtBlock + lane;
    const size_t evBase = size_t(region) * kPtBlock * sb.ev_per_path;
    const uint32_t evCap = kPtBlock * sb.ev_per_path;
    const uint32_t hitCap = kPtBlock * sb.hits_per_path;
    bool overflow = false;
    // A slot of the region's event area; one LDS atomic per lane, so each lane's events take
    // increasing slots in its own program order (the order the unwind kernel sums them in).
    auto reserve = [&](size_t& e) -> bool {
        const uint32_t k = atomicAdd(&counters[0], 1u);
        if (k >= evCap) {
            overflow = true;
            return false;
        }
        e = evBase + k;
        return true;
    };
    if (li < fr.n_paths) {
        const uint32_t pid = path_pixel(fr.width, fr.height, fr.path_base + li);
        const uint32_t px = pid % fr.width, py = pid / fr.width;
        const size_t n = fr.n_paths;
        int* stk = stack + lane;
        const bool doNEE = (fr.options & kOptNEE) != 0, doMIS = doNEE && (fr.options & kOptMIS) != 0;
        const bool useEmissive = (fr.options & kOptEmissiveLights) != 0 && sc.n_lights > 0;
        const bool useAnalytic = Lights && (fr.options & kOptAnalyticLights) != 0 && sc.n_analytic > 0;
        const bool useEnv = Lights && (fr.options & kOptEnvLight) != 0 && sc.env_w > 0;
        const bool importance = (fr.options & kOptISEmitters) != 0;
        const bool doMNEE = doNEE && (fr.options & kOptMNEE) != 0;
        // getLightTypeSelectionProbabilities (LightSampling.slangh:37-55)
        float pEnv = useEnv ? 1.f : 0.f, pEmissive = useEmissive ? 1.f : 0.f, pAnalytic = useAnalytic ? 1.f : 0.f;
        {
            const float psum = pEnv + pEmissive + pAnalytic;
            if (psum != 0.f) {
                const float inv = 1.f / psum;
                pEnv *= inv;
                pEmissive *= inv;
                pAnalytic *= inv;
            }
        }
        Rng sg = Rng::make(px, py, fr.frame, fr.sample_generator == 0);
        const float lambda = kMinWavelength + sg.next1D() * (kMaxWavelength - kMinWavelength);
        const size_t tileBase = fr.tile_draws ? size_t((py / fr.tile_size) * fr.tiles_x + px / fr.tile_size) * (fr.max_bounces + 1) : 0;
        uint32_t neeDraws = 0;
        V3 prevPos = v3(sc.cam_pos[0], sc.cam_pos[1], sc.cam_pos[2]);  // the megakernel's sv.origin

        for (int p = 0;; ++p) {
            const Rec bounce = read_rec(fr.records, n, li, p);
            const uint32_t metaBase = lane | (uint32_t(p) << 6);
            const int comps = (!(fr.options & kOptHwssMis) && (bounce.flags & kFlagDispersive)) ? 1 : N;
            const uint32_t compsBit = comps == 1 ? kEvBitComps1 : 0u;
            if (bounce.flags & kFlagMiss) {
                // onMiss (pltpt_solve.rt.slang:371-396)
                if (Lights && useEnv) {
                    const bool didNEE = doNEE && p > 0 &&
                                        (!(bounce.flags & kFlagLastDelta) ||
                                         (doMNEE && (bounce.flags >> 16) <= fr.mnee_max_occluders && p > 1));
                    if (doMIS || !didNEE) {
                        const bool needMIS = doMIS && didNEE;
                        const V3 wo = bounce.wi;
                        const float lightPdf = pEnv * env_pdf(sc, -wo);
                        const float mis = needMIS ? bounce.lastPdf / (bounce.lastPdf + lightPdf) : 1.f;
                        float lam[N], Le[N], rgb[3];
                        hwss_lambdas<N>(lambda, lam);
                        env_eval(sc, -wo, rgb);
#pragma unroll
                        for (int j = 0; j < N; ++j) Le[j] = mis * rgb_to_spectrum(rgb[0], rgb[1], rgb[2], lam[j]);
                        size_t e;
                        if (!reserve(e)) break;
                        ev_put(sb, e, kEvMeta, metaBase | (kEvKindMiss << 13) | compsBit | kEvBitDistant);
                        ev_put_source<N>(sb, e, Le, lam, v3(0.f, 0.f, 0.f), wo, fr.max_omega, 0.f);
                    }
                }
                break;
            }
            const Surf s = surface_at(sc, bounce.tri, bounce.b1, bounce.b2, bounce.wi);
            const Mat m = material_at(sc, s.material);
            const uint32_t props = mat_props(m);

            // ---- NEE (pltpt_solve.rt.slang:245-366) up to the beam's transport ----
            if (doNEE) {
                const uint32_t lobes = mat_lobes(m);
                const bool allowMS = doMNEE && p > 0 && !(bounce.flags & kFlagLastDelta) && (props & kPropMWReflective);
                const bool sampleRefl = (lobes & kLobeNonDeltaReflection) != 0 || allowMS;
                const bool sampleTrans = (lobes & kLobeNonDeltaTransmission) != 0;
                if (sampleRefl || sampleTrans) {
                    float uType = fr.tile_draws ? fr.tile_draws[tileBase + neeDraws++] : sg.next1D();
                    int ltype = -1;
                    float typePdf = 0.f;
                    if (uType < pEnv) {
                        ltype = 0;
                        typePdf = pEnv;
                    } else {
                        uType -= pEnv;
                        if (uType < pEmissive) {
                            ltype = 1;
                            typePdf = pEmissive;
                        } else {
                            uType -= pEmissive;
                            if (uType < pAnalytic) {
                                ltype = 2;
                                typePdf = pAnalytic;
                            }
                        }
                    }
                    bool valid = ltype == 1;
                    float lsPdf = 0.f, lsDist = 0.f;
                    V3 lsDir = v3(0.f, 0.f, 0.f), lsOrigin = v3(0.f, 0.f, 0.f), tPos = v3(0.f, 0.f, 0.f);
                    V3 bPos = v3(0.f, 0.f, 0.f);
                    float bSize = fr.sourcing_area;
                    bool bDistant = false;
                    float Le[N], lam[N];
#pragma unroll
                    for (int j = 0; j < N; ++j) { Le[j] = 0.f; lam[j] = lambda; }
                    if (valid) {
#pragma clang fp contract(off)
                        const float uL = sg.next1D();
                        uint32_t ti;
                        if (fr.options & kOptUniformEmissive) {
                            ti = sc.active_lights[min(uint32_t(uL * float(sc.n_active)), sc.n_active - 1)];
                        } else {
                            ti = min(uint32_t(uL * float(sc.n_lights)), sc.n_lights - 1);
                            const uint2 pk = sc.alias[ti];
                            const float thr = float(__builtin_bit_cast(_Float16, uint16_t(pk.x >> 16)));
                            const uint32_t above = ((pk.x & 0xFFFFu) << 8) | ((pk.y >> 24) & 0xFFu);
                            const uint32_t below = pk.y & 0xFFFFFFu;
                            ti = sg.next1D() >= thr ? above : below;
                        }
                        const float selPdf = sc.light_pdf[ti];
                        const float4 q0 = sc.lights[4 * ti], q1 = sc.lights[4 * ti + 1], q2 = sc.lights[4 * ti + 2], qn = sc.lights[4 * ti + 3];
                        const float u0 = sg.next1D(), u1 = sg.next1D();
                        const float su = sqrtf(u0);
                        const float bb0 = 1.f - su, bb1 = u1 * su;
                        const float w0 = 1.f - bb0 - bb1;
                        tPos = v3(q0.x * w0 + q1.x * bb0 + q2.x * bb1, q0.y * w0 + q1.y * bb0 + q2.y * bb1, q0.z * w0 + q1.z * bb0 + q2.z * bb1);
                        const V3 toL = tPos - s.pos;
                        const float d2 = fmaxf(kFltMin, toL.x * toL.x + toL.y * toL.y + toL.z * toL.z);
                        const float dist = sqrtf(d2);
                        const V3 ldir = v3(toL.x / dist, toL.y / dist, toL.z / dist);
                        const V3 ln = v3(qn.x, qn.y, qn.z);
                        const float cosT = ln.x * -ldir.x + ln.y * -ldir.y + ln.z * -ldir.z;
                        if (cosT <= 0.f) {
                            valid = false;
                        } else {
                            const uint32_t lm = __float_as_uint(q1.w);
                            const plt_material& em = sc.materials[lm];
                            const plt_spectral_profile* ep = em.emission_profile < sc.n_profiles ? sc.profiles + em.emission_profile : nullptr;
                            sample_spectrum<N>(ep, lambda, sg, importance, Le, lam);
#pragma unroll
                            for (int j = 0; j < N; ++j) Le[j] *= rgb_to_spectrum(1.f, 1.f, 1.f, lam[j]);
                            const float tpdf = d2 / fmaxf(kFltMin, cosT * q0.w) * selPdf;
                            lsPdf = tpdf;
#pragma unroll
                            for (int j = 0; j < N; ++j) Le[j] = tpdf > 0.f ? Le[j] / tpdf : 0.f;
                            const V3 lightPos = ray_origin(tPos, ln);
                            const V3 rd = lightPos - s.pos;
                            const float side = s.faceN.x * rd.x + s.faceN.y * rd.y + s.faceN.z * rd.z;
                            lsOrigin = ray_origin(s.pos, side >= 0.f ? s.faceN : -s.faceN);
                            const V3 tl = lightPos - lsOrigin;
                            lsDist = sqrtf(tl.x * tl.x + tl.y * tl.y + tl.z * tl.z);
                            lsDir = v3(tl.x / lsDist, tl.y / lsDist, tl.z / lsDist);
                            bPos = tPos;
                            bool anyL = false;
#pragma unroll
                            for (int j = 0; j < N; ++j) anyL |= Le[j] > 0.f;
                            valid = anyL;
                        }
                    }
                    if (Lights && ltype == 0) {
                        const float u0 = sg.next1D(), u1 = sg.next1D();
                        float epdf, rgb[3];
                        env_sample(sc, u0, u1, lsDir, epdf);
                        hwss_lambdas<N>(lambda, lam);
                        env_eval(sc, lsDir, rgb);
                        bool anyL = false;
#pragma unroll
                        for (int j = 0; j < N; ++j) {
                            const float le = rgb_to_spectrum(rgb[0], rgb[1], rgb[2], lam[j]);
                            Le[j] = epdf > 0.f ? le / epdf : 0.f;
                            anyL |= Le[j] > 0.f;
                        }
                        lsPdf = epdf;
                        const float side = s.faceN.x * lsDir.x + s.faceN.y * lsDir.y + s.faceN.z * lsDir.z;
                        lsOrigin = ray_origin(s.pos, side >= 0.f ? s.faceN : -s.faceN);
                        lsDist = kFltMax;
                        bDistant = true;
                        bSize = fr.max_omega;
                        valid = anyL;
                    }
                    if (Lights && ltype == 2) {
                        AnalyticSample as;
                        valid = sample_analytic<N>(sc, s, lambda, importance, sg, as, Le, lam);
                        if (valid) {
                            lsDir = as.dir;
                            lsOrigin = as.origin;
                            lsDist = as.dist;
                            lsPdf = as.pdf;
                            bPos = as.beam_pos;
                            bSize = as.beam_size;
                            bDistant = as.distant;
                            bool anyL = false;
#pragma unroll
                            for (int j = 0; j < N; ++j) anyL |= Le[j] > 0.f;
                            valid = anyL;
                        }
                    }
                    lsPdf *= typePdf;
                    const bool isRefl = dot(s.faceN, lsDir) * dot(s.faceN, bounce.wi) > 0.f;
                    valid = valid && lsPdf > 1e-6f && !(!sampleRefl && isRefl) && !(!sampleTrans && !isRefl);
                    bool anyLe = false;
#pragma unroll
                    for (int j = 0; j < N; ++j) anyLe |= Le[j] > 0.f;
                    if (valid && anyLe) {
                        const bool misL = doMIS && ltype != 2;
                        const int maxMSBounces = int(fr.max_bounces) - p;
                        const bool mnee = doMNEE && maxMSBounces > 0;
                        bool ok = true, currentVertexMS = false;
                        V3 origin = s.pos;
                        MsData ms;
                        ms.count = 0;
                        ms.iterations = 0;
                        ms.wavelength = lam[0];
                        ms.pdf = 1.f;
                        ms.dispersive = false;
                        ms.woLast = v3(0.f, 0.f, 0.f);
                        if (!mnee) {
                            const Ray vr{lsOrigin.x, lsOrigin.y, lsOrigin.z, lsDir.x, lsDir.y, lsDir.z, 0.f, lsDist};
                            ok = traverse<true, kTraversalStackDepth>(sc.tsc, vr, stk, kPtBlock).tri == kNoHit;
                        } else {  // manifold sampling on NEE (pltpt_solve.rt.slang:296-366); the beam arrives along -lsDir
                            const V3 lightDir = lsDir;
                            const bool isRefl2 = dot(s.f.N, lightDir) * dot(s.f.N, bounce.wi) > 0.f;
                            currentVertexMS = isRefl2 && allowMS;
                            if (currentVertexMS)
                                ok = ms_sample_vertex(bounce.tri, bounce.b1, bounce.b2, s.pos, s, m, bounce.wi, lightDir, sg, ms) &&
                                     (ms.hits[0].lobe & kLobeReflection) != 0;
                            origin = !currentVertexMS ? s.pos : prevPos;
                            // the walk's random part (manifoldNEESample up to its Newton iterations,
                            // Manifolds.slangh:165-182); the iterations and the light's visibility draw
                            // no random numbers and are left to the walk kernel
                            ok = ok && ms_trace_occluders(sc, lsOrigin, lsDir, lsDist, int(fr.mnee_max_occluders), maxMSBounces, sg, ms, stk);
                            ok = ok && !(ms.pdf < 1e-6f);
                        }
                        if (ok) {
                            uint32_t hk = 0;
                            if (ms.count > 0) {
                                hk = atomicAdd(&counters[1], uint32_t(ms.count));
                                if (hk + uint32_t(ms.count) > hitCap) {
                                    overflow = true;
                                    break;
                                }
                            }
                            size_t e;
                            if (!reserve(e)) break;
                            uint32_t meta = metaBase | (kEvKindNee << 13) | compsBit | (bDistant ? kEvBitDistant : 0u) |
                                            (misL ? kEvBitMisL : 0u);
                            if (mnee)
                                meta |= kEvBitMnee | (currentVertexMS ? kEvBitCurrentMs : 0u) | (ms.dispersive ? kEvBitDispersive : 0u) |
                                        (uint32_t(ms.count) << 21);
                            ev_put(sb, e, kEvMeta, meta);
                            ev_put_source<N>(sb, e, Le, lam, bPos, -lsDir, bSize, lsDist);
                            ev_putf(sb, e, kEvTypePdf, typePdf);
                            ev_putf(sb, e, kEvLsPdf, lsPdf);
                            if (mnee) {
                                ev_putf(sb, e, kEvMsPdf, ms.pdf);
                                if (ms.count > 0) {
                                    ev_put3(sb, e, kEvOrg, origin);
                                    ev_put(sb, e, kEvHitBase, hk);
                                    const size_t hb = size_t(region) * kPtBlock * sb.hits_per_path + hk;
                                    for (int k = 0; k < ms.count; ++k) {
                                        uint32_t* hp = sb.hits + hb + k;
                                        hp[kHitTri * sb.hit_cap] = ms.hits[k].tri;
                                        hp[kHitB1 * sb.hit_cap] = __float_as_uint(ms.hits[k].b1);
                                        hp[kHitB2 * sb.hit_cap] = __float_as_uint(ms.hits[k].b2);
                                        hp[kHitPos * sb.hit_cap] = __float_as_uint(ms.hits[k].pos.x);
                                        hp[(kHitPos + 1) * sb.hit_cap] = __float_as_uint(ms.hits[k].pos.y);
                                        hp[(kHitPos + 2) * sb.hit_cap] = __float_as_uint(ms.hits[k].pos.z);
                                        hp[kHitLobe * sb.hit_cap] = ms.hits[k].lobe;
                                    }
                                    sb.walk.items[wave_append(sb.walk.counters, 1u)] = uint32_t(e);
                                }
                            }
                        }
                    }
                }
            }

            // ---- emissive hit (evalEmissive, pltpt_solve.rt.slang:152-200), sourcing only ----
            const bool didNEE = doNEE && p > 0 &&
                                (!(bounce.flags & kFlagLastDelta) ||
                                 (doMNEE && (bounce.flags >> 16) <= fr.mnee_max_occluders && p > 1));
            if (useEmissive && (props & kPropEmissive) && (doMIS || !didNEE)) {
                const bool needMIS = doMIS && didNEE;
                float lightPdf = 0.f;
                const uint32_t lt = sc.tri_light[bounce.tri];
                if (lt != kNoHit) {
                    const V3 hn = s.front ? s.faceN : -s.faceN;
                    const V3 toL = s.pos - prevPos;
                    const float d2 = dot(toL, toL);
                    if (d2 > kFltMin) {
                        const V3 L = toL * (1.f / sqrtf(d2));
                        const float ct = dot(hn, -L);
                        if (ct > 0.f) lightPdf = pEmissive * (sc.light_pdf[lt] * (d2 / fmaxf(kFltMin, ct * sc.lights[4 * lt].w)));
                    }
                }
                const float mis = needMIS ? bounce.lastPdf / (bounce.lastPdf + lightPdf) : 1.f;
                const plt_material& em = sc.materials[s.material];
                const plt_spectral_profile* ep = em.emission_profile < sc.n_profiles ? sc.profiles + em.emission_profile : nullptr;
                float Le[N], lam[N];
                sample_spectrum<N>(ep, lambda, sg, importance, Le, lam);
                bool anyLe = false;
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    Le[j] *= mis * rgb_to_spectrum(1.f, 1.f, 1.f, lam[j]);
                    anyLe |= Le[j] > 0.f;
                }
                if (anyLe) {
                    const V3 dd = prevPos - s.pos;
                    const float dist = sqrtf(dot(dd, dd));
                    size_t e;
                    if (!reserve(e)) break;
                    ev_put(sb, e, kEvMeta, metaBase | (kEvKindEmissive << 13) | compsBit);
                    ev_put_source<N>(sb, e, Le, lam, s.pos, bounce.wi, fr.sourcing_area, dist);
                }
            }

            if ((bounce.flags & kFlagTerminated) || p == int(fr.max_bounces)) break;
            prevPos = s.pos;
        }
    }
    __syncthreads();
    const uint64_t ovf = __ballot(overflow);
    if (lane == 0) {
        uint32_t* rg = sb.region + size_t(region) * kRegWords;
        rg[kRegEvents] = min(counters[0], evCap);
        rg[kRegHits] = min(counters[1], hitCap);
        rg[kRegOverflowLo] = uint32_t(ovf);
        rg[kRegOverflowHi] = uint32_t(ovf >> 32);
    }
}

// ---------------------------------------------------------------------------------------
// The learned surrogate in the unwind (query and apply modes).
// ---------------------------------------------------------------------------------------
// Adds the number of active lanes with `pred` to *counter, one atomic per wave.
PLT_DEV void sur_count(uint32_t* counter, bool pred) {
    const uint64_t m = __ballot(pred);
    if (pred && int(__lane_id()) == __ffsll(static_cast<unsigned long long>(m)) - 1) atomicAdd(counter, uint32_t(__popcll(m)));
}

// Query mode at an eligible step: the network's inputs for each HWSS component, as one row each.
// The coherence area is the one the coated opaque dielectric's interference term sees: its
// coherence exp(-(k/2pi)^2 tau^2 Z^T (D/r^2) Z / 2) (Z the normal in the exitant beam frame) equals
// that of an isotropic beam of area pi (lambda/1000)^2 |Z_t|^2 / (Z^T (D/r^2) Z) at 1 mm, the beams
// the network was trained on (clamped near the top of its range, where the factor is 1 to 1e-6); the
// other two families' lobes do not depend on coherence and take the middle of the range.
template <int N>
PLT_DEV void sur_query(const Mat& m, int fam, const Frame& sd, V3 V, V3 woW, uint32_t lobe, const Beam<N>& beam, const SolveBuffers& sb,
                       size_t e, uint32_t k) {
    const SurPath& sp = sb.sur;
    const V3 wiS = sd.to_local(V), woS = sd.to_local(woW);
    float areaScale = -1.f;  // area = areaScale * (lambda/1000)^2; < 0: the neutral area
    if (fam == kSurCoatedOpaque) {
        Beam<1> g;  // the beam's frame and coherence, taken to the exitant frame as the exact evaluation does
        g.S[0][0] = g.S[0][1] = g.S[0][2] = g.S[0][3] = 0.f;
        g.lambda[0] = beam.lambda[0];
        g.spdf[0] = 1.f;
        g.comps = 1;
        g.D00 = beam.D00; g.D01 = beam.D01; g.D11 = beam.D11;
        g.r_m = beam.r_m;
        g.pos = beam.pos;
        g.r = sd.to_local(beam.r);
        g.t = sd.to_local(beam.t);
        g.distant = beam.distant;
        g.construct_local_frame(woS);
        const V3 Z = g.to_local(v3(0.f, 0.f, 1.f));
        const float zz = Z.x * Z.x + Z.y * Z.y;
        if (g.r_m == 0.f) {
            areaScale = 0.f;  // not propagated: incoherent across any offset (outside the training range)
        } else if (zz > 1e-12f) {
            const float ir = 1.f / g.r_mm(), sc2 = ir * ir;
            const float q = Z.x * (sc2 * g.D00 * Z.x + sc2 * g.D01 * Z.y) + Z.y * (sc2 * g.D01 * Z.x + sc2 * g.D11 * Z.y);
            areaScale = q > 0.f ? kPi * zz / q : 1e30f;
        }
    }
    const uint32_t row0 = wave_append(sp.counter, uint32_t(N));
    bool inside = row0 + uint32_t(N) <= sp.qcap;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (!inside || j >= beam.comps) continue;
        const float l0 = 1e-3f * beam.lambda[j];
        const float area = areaScale < 0.f ? kSurAreaNeutral : fminf(areaScale * (l0 * l0), kSurAreaClamp);
        float x[kSurIn];
        sur_features(m, fam, wiS, woS, beam.lambda[j], area, x);
        float* xr = sp.X + size_t(row0 + j) * kSurIn;
#pragma unroll
        for (int f = 0; f < kSurIn; ++f) {
            inside &= x[f] >= sp.stats[2 * kSurOut + f] && x[f] <= sp.stats[2 * kSurOut + kSurIn + f];
            xr[f] = x[f];
        }
        if (sp.teacher) {
            // the exact matrix of the same evaluation in the network's frames, in its output encoding
            // (standardised ln M00 and asinh(Mk / M00))
            float M[16];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float ec[4] = {c == 0 ? 1.f : 0.f, c == 1 ? 1.f : 0.f, c == 2 ? 1.f : 0.f, c == 3 ? 1.f : 0.f};
                float col[4];
                exact_stokes(m, wiS, woS, lobe, beam.lambda[j], area, ec, col);
#pragma unroll
                for (int r = 0; r < 4; ++r) M[4 * r + c] = col[r];
            }
            float* yr = const_cast<float*>(sp.Y) + size_t(row0 + j) * kSurOut;
            const bool pos = M[0] > 0.f;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const float t = q == 0 ? (pos ? logf(M[0]) : -1e30f) : (pos ? asinhf(M[q] / M[0]) : 0.f);
                yr[q] = (t - sp.stats[q]) / sp.stats[kSurOut + q];
            }
        }
    }
    sp.qidx[size_t(k) * sb.ev_cap + e] = inside ? row0 : ~0u;
    sur_count(sp.counter + 1, !inside);
}

// Apply mode at an eligible step: the network's Mueller matrices (M00 = exp(y0), Mk = M00 sinh(yk)
// after de-standardising) on the Stokes vectors taken to the network's incident frame; the beam's
// frame and coherence change as in the exact evaluation. Returns false for a step to evaluate exactly.
template <int N>
PLT_DEV bool sur_apply(const Mat& m, const Frame& sd, V3 V, V3 woW, uint32_t lobe, Beam<N>& beam, const SolveBuffers& sb, size_t e,
                       uint32_t k) {
    const SurPath& sp = sb.sur;
    const uint32_t row0 = sp.qidx[size_t(k) * sb.ev_cap + e];
    if (row0 == ~0u) return false;
    const V3 wiS = sd.to_local(V);
    beam.rotate_frame(sd.from_local(perp_stark(-wiS)), beam.r);
    float Sc[N][4];
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
        for (int c = 0; c < 4; ++c) Sc[j][c] = beam.S[j][c];
    const Frame f = mat_frame(m, sd);
    beam.r = f.to_local(beam.r);
    beam.t = f.to_local(beam.t);
    const V3 wo = f.to_local(woW);
    beam.construct_local_frame(wo);
    const float* st = sp.stats;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (j >= beam.comps) continue;
        const float* y = sp.Y + size_t(row0 + j) * kSurOut;
        const float m00 = expf(y[0] * st[kSurOut] + st[0]);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float acc = 0.f;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int q = 4 * r + c;
                const float v = q == 0 ? 1.f : sinhf(y[q] * st[kSurOut + q] + st[q]);
                acc += v * Sc[j][c];
            }
            beam.S[j][r] = m00 * acc;
        }
    }
    mat_coherence_transform<N>(m, wo, lobe, beam);
    beam.r = f.from_local(beam.r);
    beam.t = f.from_local(beam.t);
    return true;
}

// unwind_step in the given mode; `k` is the step's index within its event (the query table's row).
template <int N, int Mode>
PLT_DEV void unwind_step_m(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, size_t e, uint32_t k, uint32_t tri, float b1,
                           float b2, V3 wo, uint32_t lobe, Beam<N>& beam, const V3* posW = nullptr) {
    if (Mode == kUnwindExact) {
        unwind_step<N>(sc, fr, tri, b1, b2, wo, lobe, beam, posW);
        return;
    }
    const V3 V = -beam.r;
    Surf s = surface_at(sc, tri, b1, b2, V);
    if (posW) s.pos = *posW;
    beam.propagate(s.pos);
    const Mat m = material_at(sc, s.material);
    const int fam = sur_family(m, lobe);
    bool done = false;
    if (fam >= 0 && k < sb.sur.steps_max) {
        if (Mode == kUnwindQuery) sur_query<N>(m, fam, s.f, V, wo, lobe, beam, sb, e, k);
        else done = sur_apply<N>(m, s.f, V, wo, lobe, beam, sb, e, k);
    }
    if (!done) mat_eval<N>(m, s.f, V, wo, lobe, beam);
    if (fr.options & kOptHwssMis) {
        float p[N];
        mat_lambda_pdf<N>(m, s.f, V, wo, lobe, beam, p);
#pragma unroll
        for (int j = 0; j < N; ++j) beam.spdf[j] *= p[j];
    }
}

// One event's transport: the megakernel's sequence from the sourced beam to measure. Writes the
// RGB contribution the megakernel adds to the path's colour (zeros where it adds nothing). Its
// unwind steps are numbered in order (walk vertices, current vertex, stored vertices) for the
// surrogate's query table.
template <int N, int Mode>
PLT_DEV void unwind_event(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, size_t e, uint32_t meta, uint32_t region,
                          float* out) {
    const uint32_t lane = meta & 63u, kind = (meta >> 13) & 3u;
    const int p = int((meta >> 6) & 127u);
    const size_t n = fr.n_paths;
    const uint32_t li = region * kPtBlock + lane;
    const Rec bounce = read_rec(fr.records, n, li, p);
    float Le[N], lam[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        Le[j] = ev_getf(sb, e, kEvLe + j);
        lam[j] = ev_getf(sb, e, kEvLam + j);
    }
    const int comps = (meta & kEvBitComps1) ? 1 : N;
    const V3 dir = ev_get3(sb, e, kEvDir);
    const float size = ev_getf(sb, e, kEvSize);
    Beam<N> beam = (meta & kEvBitDistant) ? source_beam_distant<N>(dir, size, Le, lam, fr.max_omega, comps)
                                          : source_beam_area<N>(ev_get3(sb, e, kEvPos), dir, size, ev_getf(sb, e, kEvDist), Le, lam,
                                                                fr.max_omega, comps);
    uint32_t k = 0;  // unwind step index
    if (kind == kEvKindNee) {
        const float typePdf = ev_getf(sb, e, kEvTypePdf), lsPdf = ev_getf(sb, e, kEvLsPdf);
        beam.scale(1.f / typePdf);
        const V3 lsDir = -dir;
        bool evalCurrent = true;
        if (!(meta & kEvBitMnee)) {
            if (meta & kEvBitMisL) {
                const Surf s = surface_at(sc, bounce.tri, bounce.b1, bounce.b2, bounce.wi);
                const Mat m = material_at(sc, s.material);
                const float bp = mat_pdf(m, s.f, bounce.wi, lsDir, lam[0]);
                beam.scale(lsPdf / (lsPdf + bp));
            }
        } else {
            const int cnt = int((meta >> 21) & 15u);
            const float msPdf = ev_getf(sb, e, kEvMsPdf);
            const bool currentVertexMS = (meta & kEvBitCurrentMs) != 0;
            V3 dirL = lsDir;
            float misBsdfPdf = 0.f;
            if (cnt > 0) {
                if (!(fr.options & kOptHwssMis) && (meta & kEvBitDispersive)) beam.comps = 1;
                if (!beam.distant) {
                    const V3 wo = ev_get3(sb, e, kEvWo);
                    dirL = wo;
                    beam.r = -wo;
                    beam.t = perp_stark(beam.r);
                }
                beam.scale(1.f / msPdf);
                // manifoldForwardUnwind (pltpt_solve.rt.slang:80-93): the walk's vertices, last first
                const uint32_t* hb = sb.hits + size_t(region) * kPtBlock * sb.hits_per_path + ev_get(sb, e, kEvHitBase);
                V3 pos = v3(__uint_as_float(hb[(cnt - 1) + kHitPos * sb.hit_cap]), __uint_as_float(hb[(cnt - 1) + (kHitPos + 1) * sb.hit_cap]),
                            __uint_as_float(hb[(cnt - 1) + (kHitPos + 2) * sb.hit_cap]));
                for (int idx = cnt - 1; idx >= 0; --idx) {
                    const uint32_t* h = hb + idx;
                    const V3 p2 = idx == 0 ? ev_get3(sb, e, kEvOrg)
                                           : v3(__uint_as_float(h[-1 + kHitPos * sb.hit_cap]), __uint_as_float(h[-1 + (kHitPos + 1) * sb.hit_cap]),
                                                __uint_as_float(h[-1 + (kHitPos + 2) * sb.hit_cap]));
                    unwind_step_m<N, Mode>(sc, fr, sb, e, k++, h[kHitTri * sb.hit_cap], __uint_as_float(h[kHitB1 * sb.hit_cap]),
                                           __uint_as_float(h[kHitB2 * sb.hit_cap]), wnormalize(p2 - pos), h[kHitLobe * sb.hit_cap], beam, &pos);
                    pos = p2;
                }
                if (currentVertexMS) {
                    evalCurrent = false;
                    misBsdfPdf = bounce.lastPdf;
                }
            }
            if (meta & kEvBitMisL) {
                if (cnt == 0 || !currentVertexMS) {
                    const Surf s = surface_at(sc, bounce.tri, bounce.b1, bounce.b2, bounce.wi);
                    const Mat m = material_at(sc, s.material);
                    misBsdfPdf = mat_pdf(m, s.f, bounce.wi, dirL, beam.lambda[0]);
                }
                beam.scale(lsPdf / (lsPdf + misBsdfPdf * msPdf));
            }
        }
        // evalNEE (pltpt_solve.rt.slang:219-243): the current vertex's non-delta lobes
        if (evalCurrent) unwind_step_m<N, Mode>(sc, fr, sb, e, k, bounce.tri, bounce.b1, bounce.b2, bounce.wi, kLobeNonDelta, beam);
        ++k;
    }
    // forwardUnwind (pltpt_solve.rt.slang:68-79): the stored vertices, last first
    for (int q = p - 1; q >= 0; --q) {
        const Rec r = read_rec(fr.records, n, li, q);
        unwind_step_m<N, Mode>(sc, fr, sb, e, k++, r.tri, r.b1, r.b2, r.wi, r.lobe, beam);
    }
    if (Mode == kUnwindQuery) return;  // the query pass needs no colour
    Solve sv;
    sv.r = sv.g = sv.b = 0.f;
    measure<N>(sc, fr, sv, p, bounce.contrib, beam);
    out[0] = sv.r;
    out[1] = sv.g;
    out[2] = sv.b;
}

// A manifold NEE event whose walk the walk kernel ran (its contribution is the walk-unwind
// kernel's in a staged unwind).
PLT_DEV bool is_walk_event(uint32_t meta) {
    return ((meta >> 13) & 3u) == kEvKindNee && (meta & kEvBitMnee) && ((meta >> 21) & 15u) != 0u;
}
PLT_DEV void contrib_put(const SolveBuffers& sb, size_t e, const float* c) {
    sb.contrib[e] = c[0];
    sb.contrib[sb.ev_cap + e] = c[1];
    sb.contrib[2 * sb.ev_cap + e] = c[2];
}

// Each lane's path's contributions in slot order (the megakernel's order): one block-uniform round
// per 64 slots; `owner` and `part` hold the round's lanes' paths and contributions.
struct PathSums {
    float r = 0.f, g = 0.f, b = 0.f;
    PLT_DEV void add_round(const uint32_t* owner, const float (*part)[kPtBlock], uint32_t m, uint32_t lane) {
        for (uint32_t j = 0; j < m; ++j) {
            if (owner[j] == lane) {
                r += part[0][j];
                g += part[1][j];
                b += part[2][j];
            }
        }
    }
};

// Two waves per SIMD: the beam, the material instance and the evaluation temporaries fit the 256
// registers without spilling, and this kernel (no traversal, few dependent fetches) loses less to
// the lower occupancy than it gains (1080p depth 8: 1.30 ms per frame at two waves, 2.14 ms at four
// with 612 B of scratch per lane, 2.59 ms at six). Mode: exact unwind, or the surrogate's query or
// apply pass (the query pass writes no colour). With sb.stage set (exact mode) it writes the
// contributions of the events that are not manifold walks and sums nothing.
template <int N, int Mode>
__global__ void __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(2, 2)))
solve_unwind_kernel(PtScene sc, PtFrame fr, SolveBuffers sb) {
    __shared__ float part[3][kPtBlock];
    __shared__ uint32_t owner[kPtBlock];
    const uint32_t lane = threadIdx.x, region = blockIdx.x;
    const uint32_t* rg = sb.region + size_t(region) * kRegWords;
    const uint32_t count = rg[kRegEvents];
    const uint64_t ovf = uint64_t(rg[kRegOverflowLo]) | (uint64_t(rg[kRegOverflowHi]) << 32);
    const size_t evBase = size_t(region) * kPtBlock * sb.ev_per_path;
    const bool stage = Mode == kUnwindExact && sb.stage != 0u;
    PathSums sum;
    // block-uniform loop (one wave per block): 64 events per round, then each lane adds its own
    // path's contributions in slot order, which is the order the megakernel adds them in
    for (uint32_t base = 0; base < count; base += kPtBlock) {
        const uint32_t k = base + lane;
        uint32_t who = kPtBlock;
        float c[3] = {0.f, 0.f, 0.f};
        if (k < count) {
            const size_t e = evBase + k;
            const uint32_t meta = ev_get(sb, e, kEvMeta);
            who = meta & 63u;
            if (((ovf >> who) & 1u) || ((meta >> 13) & 3u) == kEvKindVoid || (stage && is_walk_event(meta))) {
                who = kPtBlock;  // left to the megakernel, void, or the walk-unwind kernel's
            } else {
                unwind_event<N, Mode>(sc, fr, sb, e, meta, region, c);
                if (stage) contrib_put(sb, e, c);
            }
        }
        if (Mode == kUnwindQuery || stage) continue;
        owner[lane] = who;
        part[0][lane] = c[0];
        part[1][lane] = c[1];
        part[2][lane] = c[2];
        __syncthreads();
        sum.add_round(owner, part, min(count - base, uint32_t(kPtBlock)), lane);
        __syncthreads();
    }
    if (Mode == kUnwindQuery || stage) return;
    const uint32_t li = region * kPtBlock + lane;
    if (li < fr.n_paths && !((ovf >> lane) & 1u)) fr.color[path_pixel(fr.width, fr.height, fr.path_base + li)] = make_float4(sum.r, sum.g, sum.b, 1.f);
}

// Staged unwind, after the walk kernel: the contributions of the walks that succeeded, one lane per
// queued walk (the void ones add nothing).
template <int N>
__global__ void __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(2, 2)))
solve_walk_unwind_kernel(PtScene sc, PtFrame fr, SolveBuffers sb) {
    const uint32_t total = sb.walk.counters[0];
    const size_t regionEvents = size_t(kPtBlock) * sb.ev_per_path;
    for (uint32_t k = blockIdx.x * kPtBlock + threadIdx.x; k < total; k += gridDim.x * kPtBlock) {
        const size_t e = sb.walk.items[k];
        const uint32_t meta = ev_get(sb, e, kEvMeta);
        if (((meta >> 13) & 3u) == kEvKindVoid) continue;
        const uint32_t region = uint32_t(e / regionEvents), who = meta & 63u;
        const uint32_t* rg = sb.region + size_t(region) * kRegWords;
        if ((rg[who < 32u ? kRegOverflowLo : kRegOverflowHi] >> (who & 31u)) & 1u) continue;  // the megakernel's
        float c[3] = {0.f, 0.f, 0.f};
        unwind_event<N, kUnwindExact>(sc, fr, sb, e, meta, region, c);
        contrib_put(sb, e, c);
    }
}

// Staged unwind, last: each path's contributions summed in slot order, the colour written.
__global__ void __launch_bounds__(kPtBlock) solve_sum_kernel(PtFrame fr, SolveBuffers sb) {
    __shared__ float part[3][kPtBlock];
    __shared__ uint32_t owner[kPtBlock];
    const uint32_t lane = threadIdx.x, region = blockIdx.x;
    const uint32_t* rg = sb.region + size_t(region) * kRegWords;
    const uint32_t count = rg[kRegEvents];
    const uint64_t ovf = uint64_t(rg[kRegOverflowLo]) | (uint64_t(rg[kRegOverflowHi]) << 32);
    const size_t evBase = size_t(region) * kPtBlock * sb.ev_per_path;
    PathSums sum;
    for (uint32_t base = 0; base < count; base += kPtBlock) {
        const uint32_t k = base + lane;
        uint32_t who = kPtBlock;
        float c[3] = {0.f, 0.f, 0.f};
        if (k < count) {
            const size_t e = evBase + k;
            const uint32_t meta = ev_get(sb, e, kEvMeta);
            who = meta & 63u;
            if (((ovf >> who) & 1u) || ((meta >> 13) & 3u) == kEvKindVoid) {
                who = kPtBlock;
            } else {
                c[0] = sb.contrib[e];
                c[1] = sb.contrib[sb.ev_cap + e];
                c[2] = sb.contrib[2 * sb.ev_cap + e];
            }
        }
        owner[lane] = who;
        part[0][lane] = c[0];
        part[1][lane] = c[1];
        part[2][lane] = c[2];
        __syncthreads();
        sum.add_round(owner, part, min(count - base, uint32_t(kPtBlock)), lane);
        __syncthreads();
    }
    const uint32_t li = region * kPtBlock + lane;
    if (li < fr.n_paths && !((ovf >> lane) & 1u)) fr.color[path_pixel(fr.width, fr.height, fr.path_base + li)] = make_float4(sum.r, sum.g, sum.b, 1.f);
}

// Clears the surrogate's query counters (a kernel, so a captured graph replays it).
__global__ void sur_clear_kernel(uint32_t* counter) {
    if (threadIdx.x < 2) counter[threadIdx.x] = 0u;
}

}  // namespace
}  // namespace plt
