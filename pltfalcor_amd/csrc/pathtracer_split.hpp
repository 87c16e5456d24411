// Split solve pass of the PLT path tracer (pltpt_solve.rt.slang:37-476 in two kernels).
//
// solve_sample_kernel replays each path's vertices with its random stream exactly as the
// megakernel does -- light-type choice, light sample, shadow ray or manifold walk (NEE,
// pltpt_solve.rt.slang:245-366), emissive-hit and environment-miss sourcing (:152-200, :371-396) --
// but instead of carrying each sourced beam through the stored vertices it appends a light event:
// what the beam is (spectral radiance, wavelengths, position, direction, size), the scale factors
// the megakernel applies to it in order (type selection, light pdf, walk pdf) and, for a manifold
// walk, its vertices. It holds the traversal stack and the walk state but no beam and no BSDF
// evaluation.
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
PLT_DEV uint32_t ev_get(const SolveBuffers& sb, size_t e, int f) { return sb.ev[size_t(f) * sb.ev_cap + e]; }
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
        const uint32_t pid = fr.path_base + li;
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
                            ok = ok && ms_nee_sample(sc, fr, lsOrigin, lsDir, lsDist, origin, bDistant ? v3(0.f, 0.f, 0.f) : bPos,
                                                     bDistant, lightDir, maxMSBounces, sg, ms, stk);
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
                                    ev_put3(sb, e, kEvWo, ms.woLast);
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

// One event's transport: the megakernel's sequence from the sourced beam to measure. Writes the
// RGB contribution the megakernel adds to the path's colour (zeros where it adds nothing).
template <int N>
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
                    unwind_step<N>(sc, fr, h[kHitTri * sb.hit_cap], __uint_as_float(h[kHitB1 * sb.hit_cap]),
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
        if (evalCurrent) unwind_step<N>(sc, fr, bounce.tri, bounce.b1, bounce.b2, bounce.wi, kLobeNonDelta, beam);
    }
    forward_unwind<N>(sc, fr, n, li, p, beam);
    Solve sv;
    sv.r = sv.g = sv.b = 0.f;
    measure<N>(sc, fr, sv, p, bounce.contrib, beam);
    out[0] = sv.r;
    out[1] = sv.g;
    out[2] = sv.b;
}

template <int N>
__global__ void __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(4, 4)))
solve_unwind_kernel(PtScene sc, PtFrame fr, SolveBuffers sb) {
    __shared__ float part[3][kPtBlock];
    __shared__ uint32_t owner[kPtBlock];
    const uint32_t lane = threadIdx.x, region = blockIdx.x;
    const uint32_t* rg = sb.region + size_t(region) * kRegWords;
    const uint32_t count = rg[kRegEvents];
    const uint64_t ovf = uint64_t(rg[kRegOverflowLo]) | (uint64_t(rg[kRegOverflowHi]) << 32);
    const size_t evBase = size_t(region) * kPtBlock * sb.ev_per_path;
    float sr = 0.f, sgr = 0.f, sbl = 0.f;
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
            if ((ovf >> who) & 1u) who = kPtBlock;  // left to the megakernel
            else unwind_event<N>(sc, fr, sb, e, meta, region, c);
        }
        owner[lane] = who;
        part[0][lane] = c[0];
        part[1][lane] = c[1];
        part[2][lane] = c[2];
        __syncthreads();
        const uint32_t m = min(count - base, uint32_t(kPtBlock));
        for (uint32_t j = 0; j < m; ++j) {
            if (owner[j] == lane) {
                sr += part[0][j];
                sgr += part[1][j];
                sbl += part[2][j];
            }
        }
        __syncthreads();
    }
    const uint32_t li = region * kPtBlock + lane;
    if (li < fr.n_paths && !((ovf >> lane) & 1u)) fr.color[fr.path_base + li] = make_float4(sr, sgr, sbl, 1.f);
}

}  // namespace
}  // namespace plt
