// PLT oracle: the PLTPT sample and solve passes on the CPU (C entry point oracle_render).
//
// TEST INFRASTRUCTURE ONLY (see plt_oracle.hpp). Restates, per pixel:
//   sample pass  Source/RenderPasses/PLTPT/pltpt_sample.rt.slang:36-185, Path.slang:146-196,
//                pltpt.slangh:132-176
//   solve pass   Source/RenderPasses/PLTPT/pltpt_solve.rt.slang:37-476 (manifold NEE off,
//                no environment or analytic lights), LightSampling.slangh:40-206,
//                Rendering/Lights/EmissivePowerSampler.slang:62-121,
//                EmissiveLightSamplerHelpers.slang:56-150, Scene/Scene.slang:251-327.
// Camera, shading frames and emissive tables are built here from the scene arrays, independently
// of the product.
#include <omp.h>
#include <limits>
#include <vector>
#include "plt_oracle_pt.hpp"

using namespace pltoracle;

namespace {

using F = float;
constexpr int HW = 4;
using Beam = PLTBeam<F, HW>;
constexpr float kFLT_MIN = 1.175494351e-38f;

struct OScene {
    const float* pos;
    const float* nrm;
    const uint32_t* idx;
    const uint32_t* triMesh;
    const uint32_t* meshMat;
    const MatData* mats;
    const SpectralProfile* prof;
    uint32_t nprof, ntri;
    OMesh mesh;
    OBvh bvh;
    // camera (Camera.cpp:187-193), view matrix rotation rows, polarizer (CameraData.slang:67-68)
    V3<F> camPos, camU, camV, camW;
    float jitterX, jitterY;
    V3<F> viewRow[3];
    float polarizer = 0.f, polarizerTheta = 0.f;
    // emissive triangles
    std::vector<uint32_t> lightTri;
    std::vector<float> lightArea, lightPdf;
    std::vector<V3<F>> lightN;
    std::vector<uint32_t> triToLight;
    std::vector<uint32_t> aliasX, aliasY;
    // EmissiveUniformSampler: lights with flux > 0 and their selection pdf 1 / count
    std::vector<uint32_t> activeLights;
    std::vector<float> lightPdfUniform;
    // analytic lights (LightData, Scene/Lights/LightData.slang:47-66)
    struct Light {
        float pos[3];
        uint32_t type;
        float dir[3];
        float openingAngle, cosOpeningAngle, cosSubtendedAngle, penumbraAngle;
        uint32_t profile;
        float surfaceArea, pad[3];
        float M[16], MI[16];
    };
    std::vector<Light> analytic;
    // environment map (EnvMap.slang) and its importance map mip chain (EnvMapSampler)
    uint32_t envW = 0, envH = 0;
    std::vector<float> envRGBA;
    float envM[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, envMI[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, envScale[3] = {1, 1, 1};
    std::vector<float> envImp;
    std::vector<uint32_t> envMip;
    uint32_t envBaseMip = 0;
};
constexpr uint32_t kEnvDim = 512;
enum { LT_Point = 0, LT_Directional, LT_Distant, LT_Rect, LT_Disc, LT_Sphere };

struct Options {
    uint32_t width, height, frame, maxBounces;
    int doNEE, doMIS, doRR, hwssMIS, isEmitters, directLights, emissiveLights;
    float sourcingArea, maxOmega;
    int doMNEE = 0, mneeMaxOccluders = 2, mneeMaxIterations = 60;
    float mneeThreshold = 5e-5f;
    int uniformEmissive = 0;  // emissiveSampler Uniform (else Power)
    int analyticLights = 1;   // useAnalyticLights
    int envLight = 1;         // useEnvLights
    int sampleGenerator = 1;  // SampleGeneratorType.slangh:32-36: 0 tiny, 1 uniform
    int perTileSG = 0;        // NEEUsePerTileSG
    uint32_t tileSize = 512;  // tileSize (the tiles kTile of the per-tile generator)
};
constexpr int kMaxMSOccluders = 8;

// ndir_to_oct_equal_area_snorm / oct_to_ndir_equal_area_snorm (MathHelpers.slang:199-240)
V2<F> ndirToOct(V3<F> n) {
    const F r = std::sqrt(F(1) - std::fabs(n.z));
    const F phi = std::atan2(std::fabs(n.y), std::fabs(n.x));
    V2<F> p;
    p.y = r * phi * K<F>::_2_PI;
    p.x = r - p.y;
    if (n.z < F(0)) p = V2<F>{F(1) - p.y, F(1) - p.x};
    return V2<F>{p.x * sgn(n.x), p.y * sgn(n.y)};
}
V3<F> octToNdir(V2<F> p) {
    const F d = F(1) - (std::fabs(p.x) + std::fabs(p.y));
    const F r = F(1) - std::fabs(d);
    const F phi = r > F(0) ? ((std::fabs(p.y) - std::fabs(p.x)) / r + F(1)) * K<F>::PI_4 : F(0);
    const F f = r * std::sqrt(F(2) - r * r);
    return V3<F>{f * sgn(p.x) * std::cos(phi), f * sgn(p.y) * std::sin(phi), sgn(d) * (F(1) - r * r)};
}

struct Bounce {  // BounceData (Path.slang:113-142)
    uint32_t tri = 0xFFFFFFFFu;
    float b1 = 0, b2 = 0;
    V2<F> encWi{0, 0};
    uint32_t lobe = 0, flags = 0;
    float pathContribution = 0, lastNonDeltaPdf = 0;
    V3<F> wi() const { return octToNdir(encWi); }
};
enum : uint32_t { RF_LastDelta = 0x1, RF_Dispersive = 0x2, RF_OnlyDelta = 0x4, RF_Miss = 0x10, RF_Terminated = 0x20 };

// getVertexData + prepareShadingData (Scene.slang:530-567, MaterialFactory.slang:79-130,
// ShadingUtils.slang:126-153). No tangents -> perp_stark frame.
SD<F> shadingData(const OScene& s, uint32_t tri, float b1, float b2, V3<F> viewDir, const V3<F>* posOverride = nullptr) {
    const V3<F> bary = {F(1) - b1 - b2, b1, b2};
    V3<F> p[3], n[3];
    for (int k = 0; k < 3; ++k) {
        const uint32_t v = s.idx[3 * size_t(tri) + k];
        p[k] = V3<F>{s.pos[3 * size_t(v)], s.pos[3 * size_t(v) + 1], s.pos[3 * size_t(v) + 2]};
        n[k] = V3<F>{s.nrm[3 * size_t(v)], s.nrm[3 * size_t(v) + 1], s.nrm[3 * size_t(v) + 2]};
    }
    SD<F> sd;
    sd.posW = p[0] * bary.x;
    sd.posW = sd.posW + p[1] * bary.y;
    sd.posW = sd.posW + p[2] * bary.z;
    if (posOverride) sd.posW = *posOverride;
    V3<F> nw = n[0] * bary.x;
    nw = nw + n[1] * bary.y;
    nw = nw + n[2] * bary.z;
    sd.f.N = normalize(nw);
    sd.faceN = normalize(cross(p[1] - p[0], p[2] - p[0]));
    sd.f.T_ = perp_stark(sd.f.N);
    sd.f.B = cross(sd.f.N, sd.f.T_);
    sd.V = viewDir;
    sd.frontFacing = dot(sd.V, sd.faceN) >= F(0);
    sd.materialID = s.meshMat[s.triMesh[tri]];
    return sd;
}
V3<F> rayOrigin(V3<F> p, V3<F> n) {
    const float pp[3] = {p.x, p.y, p.z}, nn[3] = {n.x, n.y, n.z};
    float o[3];
    compute_ray_origin(pp, nn, o);
    return V3<F>{o[0], o[1], o[2]};
}
Instance<F> instanceAt(const OScene& s, uint32_t m) { return setupMaterialInstance<F>(s.mats[m], s.prof, s.nprof); }

// Scene::sampleSpectrum (Scene.slang:296-327) with HW = 4 HWSS samples.
template <class S> void sampleSpectrum(const SpectralProfile* sp, float hero, S& sg, bool importance, float* Le, float* lam) {
    if (!sp) { for (int j = 0; j < HW; ++j) { Le[j] = 0; lam[j] = hero; } return; }
    if (importance) {
        lam[0] = hero;
        Le[0] = evalSpectrum<float>(sp, hero);
        float x[3];
        for (int k = 0; k < 3; ++k) x[k] = float(sp->bins) * sampleNext1D<float>(sg);
        for (int j = 1; j < HW; ++j) {
            const uint32_t id = uint32_t(x[j - 1]);
            const float a = id < sp->bins ? sp->icdf[id] : 1.f;
            const float b = id + 1 < sp->bins ? sp->icdf[id + 1] : 1.f;
            const float cdf = hlerp(a, b, x[j - 1] - std::floor(x[j - 1]));
            lam[j] = hlerp(sp->minWavelength, sp->maxWavelength, cdf);
            Le[j] = sp->intensity / float(sp->bins);
        }
    } else {
        Spectrum<float>::hwssWavelengths(hero, HW, lam);
        for (int j = 0; j < HW; ++j) Le[j] = evalSpectrum<float>(sp, lam[j]);
    }
}
// pltSourceBeamDistant (PLT.slang:253-269)
Beam sourceBeamDistant(V3<F> dir, float solidAngle, const float* Le, const float* lam, float maxOmega, int comps) {
    solidAngle = std::fmin(solidAngle, maxOmega);
    Beam b;
    for (int j = 0; j < HW; ++j) { b.L[j] = Le[j]; b.L1[j] = b.L2[j] = b.L3[j] = 0; b.wavelengths[j] = lam[j]; b.spectralPDFs[j] = 1; }
    b.hwss_comps = comps;
    b.Theta.D[0] = solidAngle; b.Theta.D[1] = 0; b.Theta.D[2] = solidAngle;
    b.Theta.r_m = 0;
    b.posW = V3<F>{0, 0, 0};
    b.r = dir;
    b.t = perp_stark(dir);
    b.distant = true;
    return b;
}
// An analytic light sample: generateAnalyticLightSample (LightSampling.slangh:122-163) over sampleLight
// (Rendering/Lights/LightHelpers.slang:70-290). Returns false where the reference's sampler does.
template <class S>
bool sampleAnalytic(const OScene& s, const Options& o, const SD<F>& sd, float hero, S& sg, V3<F>& dir, V3<F>& origin,
                    float& dist, float& pdf, V3<F>& bPos, float& bSize, bool& distant, float* Li, float* lam) {
    const uint32_t cnt = uint32_t(s.analytic.size());
    const uint32_t idx = std::min(uint32_t(sampleNext1D<float>(sg) * float(cnt)), cnt - 1);
    const OScene::Light& L = s.analytic[idx];
    const SpectralProfile* sp = L.profile < s.nprof ? s.prof + L.profile : nullptr;
    const float* M = L.M;
    const float* MI = L.MI;
    const float k2Pi = 6.28318530717958647692f;
    float lpdf = 0.f;
    distant = false;
    if (L.type == LT_Rect || L.type == LT_Disc || L.type == LT_Sphere) {
        const float u0 = sampleNext1D<float>(sg), u1 = sampleNext1D<float>(sg);
        V3<F> pl, nl{0, 0, 1};
        if (L.type == LT_Rect) {
            pl = V3<F>{u0 * 2.f - 1.f, u1 * 2.f - 1.f, 0.f};
        } else if (L.type == LT_Disc) {  // sample_disk (MathHelpers.slang:255-263)
            const float r = std::sqrt(u0), phi = k2Pi * u1;
            pl = V3<F>{r * std::cos(phi), r * std::sin(phi), 0.f};
        } else {  // sample_sphere (MathHelpers.slang:282-288)
            const float phi = k2Pi * u1, ct = 1.f - 2.f * u0;
            const float st = std::sqrt(std::fmax(0.f, 1.f - ct * ct));
            pl = V3<F>{st * std::cos(phi), st * std::sin(phi), ct};
            nl = pl;
        }
        const V3<F> pw{M[0] * pl.x + M[1] * pl.y + M[2] * pl.z + M[3], M[4] * pl.x + M[5] * pl.y + M[6] * pl.z + M[7],
                       M[8] * pl.x + M[9] * pl.y + M[10] * pl.z + M[11]};
        const V3<F> nw = normalize(V3<F>{MI[0] * nl.x + MI[1] * nl.y + MI[2] * nl.z, MI[4] * nl.x + MI[5] * nl.y + MI[6] * nl.z,
                                         MI[8] * nl.x + MI[9] * nl.y + MI[10] * nl.z});
        // finalizeAreaLightSample (LightHelpers.slang:70-93)
        const V3<F> toL = pw - sd.posW;
        const float d2 = std::fmax(dot(toL, toL), 1e-9f);
        dist = std::sqrt(d2);
        dir = toL / dist;
        const float ct = dot(nw, -dir);
        if (ct <= 0.f) return false;
        sampleSpectrum(sp, hero, sg, o.isEmitters != 0, Li, lam);
        const float f = L.surfaceArea * ct / d2;
        for (int j = 0; j < HW; ++j) Li[j] *= f;
        lpdf = d2 / (ct * L.surfaceArea);
    } else if (L.type == LT_Distant) {  // sampleDistantLight (:170-198), sample_cone (MathHelpers.slang:270-276)
        const float u0 = sampleNext1D<float>(sg), u1 = sampleNext1D<float>(sg);
        const float c = L.cosSubtendedAngle;
        const float z = u0 * (1.f - c) + c;
        const float r = std::sqrt(1.f - z * z), phi = k2Pi * u1;
        const V3<F> d{r * std::cos(phi), r * std::sin(phi), z};
        dir = normalize(V3<F>{M[0] * d.x + M[1] * d.y + M[2] * d.z, M[4] * d.x + M[5] * d.y + M[6] * d.z,
                              M[8] * d.x + M[9] * d.y + M[10] * d.z});
        dist = std::numeric_limits<float>::max();
        sampleSpectrum(sp, hero, sg, o.isEmitters != 0, Li, lam);
        lpdf = 1.f / (k2Pi * (1.f - c));
        distant = true;
    } else if (L.type == LT_Directional) {  // sampleDirectionalLight (:200-225)
        dir = V3<F>{-L.dir[0], -L.dir[1], -L.dir[2]};
        dist = std::numeric_limits<float>::max();
        sampleSpectrum(sp, hero, sg, o.isEmitters != 0, Li, lam);
        lpdf = 0.f;
        distant = true;
    } else {  // samplePointLight (:227-269); pdf 0, so the sample is rejected
        const V3<F> toL = V3<F>{L.pos[0], L.pos[1], L.pos[2]} - sd.posW;
        const float d2 = std::fmax(dot(toL, toL), 1e-9f);
        dist = std::sqrt(d2);
        dir = toL / dist;
        sampleSpectrum(sp, hero, sg, o.isEmitters != 0, Li, lam);
        lpdf = 0.f;
    }
    for (int j = 0; j < HW; ++j) Li[j] *= float(cnt);
    pdf = lpdf / float(cnt);
    origin = rayOrigin(sd.posW, dot(sd.faceN, dir) >= 0.f ? sd.faceN : -sd.faceN);  // computeNewRayOrigin
    bPos = V3<F>{L.pos[0], L.pos[1], L.pos[2]};
    bSize = distant ? 4.f * 3.14159265358979323846f * (1.f - L.cosSubtendedAngle * L.cosSubtendedAngle) : L.surfaceArea;
    return true;
}

// ---- environment map ----
V3<F> mul3(const float* M, V3<F> d) {
    return V3<F>{M[0] * d.x + M[1] * d.y + M[2] * d.z, M[3] * d.x + M[4] * d.y + M[5] * d.z, M[6] * d.x + M[7] * d.y + M[8] * d.z};
}
// world_to_latlong_map (MathHelpers.slang:106-113); acos argument clamped to [-1, 1]
void latlong(V3<F> d, float& u, float& v) {
    const float l = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
    const V3<F> p{d.x / l, d.y / l, d.z / l};
    u = std::atan2(p.x, -p.z) * 0.159154943091895335769f + 0.5f;
    v = std::acos(std::fmin(std::fmax(p.y, -1.f), 1.f)) * 0.318309886183790671538f;
}
// bilinear lookup, u wrapping, v clamped (EnvMap.cpp:123-126)
void envTexel(const OScene& s, float u, float v, float rgb[3]) {
    const float ss = u * float(s.envW) - 0.5f, t = v * float(s.envH) - 0.5f;
    const float fs = std::floor(ss), ft = std::floor(t);
    const float a = ss - fs, b = t - ft;
    const int W = int(s.envW), H = int(s.envH);
    int x0 = int(fs) % W;
    if (x0 < 0) x0 += W;
    const int x1 = (x0 + 1) % W;
    const int y0 = std::min(std::max(int(ft), 0), H - 1), y1 = std::min(std::max(int(ft) + 1, 0), H - 1);
    for (int c = 0; c < 3; ++c) {
        const float c00 = s.envRGBA[4 * (size_t(y0) * W + x0) + c], c10 = s.envRGBA[4 * (size_t(y0) * W + x1) + c];
        const float c01 = s.envRGBA[4 * (size_t(y1) * W + x0) + c], c11 = s.envRGBA[4 * (size_t(y1) * W + x1) + c];
        const float top = c00 + (c10 - c00) * a, bot = c01 + (c11 - c01) * a;
        rgb[c] = top + (bot - top) * b;
    }
}
// EnvMap::eval(dir): radiance from world direction dir
void envEval(const OScene& s, V3<F> dir, float rgb[3]) {
    float u, v;
    latlong(mul3(s.envMI, dir), u, v);
    envTexel(s, u, v, rgb);
    for (int c = 0; c < 3; ++c) rgb[c] *= s.envScale[c];
}
// oct_to_ndir_equal_area_unorm / ndir_to_oct_equal_area_unorm (MathHelpers.slang:199-243)
V3<F> octUnormToDir(float u, float v) { return octToNdir(V2<F>{u * 2.f - 1.f, v * 2.f - 1.f}); }
// EnvMapSamplerSetup.cs.slang + Texture::generateMips: 512^2 texels of mean luminance over 8x8
// samples of the octahedral map, then 2x2 box levels down to 1x1
void buildEnvImportance(OScene& s) {
    const uint32_t spp = 8;
    std::vector<float> level(size_t(kEnvDim) * kEnvDim);
    const float inv = 1.f / float(spp * spp), dimS = float(kEnvDim * spp);
    for (uint32_t y = 0; y < kEnvDim; ++y)
        for (uint32_t x = 0; x < kEnvDim; ++x) {
            float L = 0.f;
            for (uint32_t sy = 0; sy < spp; ++sy)
                for (uint32_t sx = 0; sx < spp; ++sx) {
                    const float px = (float(x * spp + sx) + 0.5f) / dimS, py = (float(y * spp + sy) + 0.5f) / dimS;
                    float u, v, rgb[3];
                    latlong(octUnormToDir(px, py), u, v);
                    envTexel(s, u, v, rgb);
                    L += rgb[0] * 0.2126f + rgb[1] * 0.7152f + rgb[2] * 0.0722f;
                }
            level[size_t(y) * kEnvDim + x] = L * inv;
        }
    s.envImp.clear();
    s.envMip.clear();
    uint32_t d = kEnvDim;
    for (;;) {
        s.envMip.push_back(uint32_t(s.envImp.size()));
        s.envImp.insert(s.envImp.end(), level.begin(), level.end());
        if (d == 1) break;
        const uint32_t h = d / 2;
        std::vector<float> next(size_t(h) * h);
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < h; ++x)
                next[size_t(y) * h + x] = ((level[size_t(2 * y) * d + 2 * x] + level[size_t(2 * y) * d + 2 * x + 1]) +
                                           (level[size_t(2 * y + 1) * d + 2 * x] + level[size_t(2 * y + 1) * d + 2 * x + 1])) * 0.25f;
        level.swap(next);
        d = h;
    }
    s.envBaseMip = uint32_t(s.envMip.size() - 1);
}
// EnvMapSampler.evalPdf (:150-163)
float envPdf(const OScene& s, V3<F> dir) {
    const V2<F> o = ndirToOct(mul3(s.envMI, dir));
    const float u = o.x * 0.5f + 0.5f, v = o.y * 0.5f + 0.5f;
    const int dim = int(kEnvDim);
    const int x = std::min(std::max(int(std::floor(u * float(dim))), 0), dim - 1);
    const int y = std::min(std::max(int(std::floor(v * float(dim))), 0), dim - 1);
    return s.envImp[size_t(y) * dim + x] / s.envImp[s.envMip[s.envBaseMip]] * (1.f / 12.5663706143591729539f);
}
// EnvMapSampler.sample (:56-148)
void envSample(const OScene& s, float px, float py, V3<F>& dir, float& pdf) {
    uint32_t x = 0, y = 0;
    for (int mip = int(s.envBaseMip) - 1; mip >= 0; --mip) {
        x *= 2; y *= 2;
        const uint32_t dim = kEnvDim >> mip;
        const float* L = s.envImp.data() + s.envMip[mip];
        const float w0 = L[size_t(y) * dim + x], w1 = L[size_t(y) * dim + x + 1];
        const float w2 = L[size_t(y + 1) * dim + x], w3 = L[size_t(y + 1) * dim + x + 1];
        const float q0 = w0 + w2, q1 = w1 + w3;
        const float dd = q0 / (q0 + q1);
        uint32_t ox, oy;
        if (px < dd) { ox = 0; px = px / dd; } else { ox = 1; px = (px - dd) / (1.f - dd); }
        const float e = ox == 0 ? (w0 / q0) : (w1 / q1);
        if (py < e) { oy = 0; py = py / e; } else { oy = 1; py = (py - e) / (1.f - e); }
        x += ox; y += oy;
    }
    const float inv = 1.f / float(kEnvDim);
    const V3<F> dl = octUnormToDir((float(x) + px) * inv, (float(y) + py) * inv);
    pdf = s.envImp[size_t(y) * kEnvDim + x] / s.envImp[s.envMip[s.envBaseMip]] * 0.0795774715459476678844f;
    dir = mul3(s.envM, dl);
}
void hwssLambdas(float hero, float* lam) { Spectrum<float>::hwssWavelengths(hero, HW, lam); }

// pltSourceBeamArea (PLT.slang:275-290)
Beam sourceBeamArea(V3<F> posW, V3<F> dir, float A, float dist, const float* Le, const float* lam, float maxOmega, int comps) {
    const float r2 = sqr(dist * 1e+3f);
    A = std::fmin(A, maxOmega * r2);
    Beam b;
    for (int j = 0; j < HW; ++j) { b.L[j] = Le[j]; b.L1[j] = b.L2[j] = b.L3[j] = 0; b.wavelengths[j] = lam[j]; b.spectralPDFs[j] = 1; }
    b.hwss_comps = comps;
    b.Theta.D[0] = A; b.Theta.D[1] = 0; b.Theta.D[2] = A;
    b.Theta.r_m = 0;
    b.posW = posW;
    b.r = dir;
    b.t = perp_stark(dir);
    b.distant = false;
    return b;
}

// ---------------------------------------------------------------- sample pass
// aov (optional): albedo and view-space normal of the first non-delta vertex (RGBA each), zero when
// there is none (pltpt_sample.rt.slang:108-112, :162-163).
template <class S>
void samplePass(const OScene& s, const Options& o, uint32_t px, uint32_t py, std::vector<Bounce>& rec, float* aov) {
    S sg(px, py, o.frame);
    if (aov)
        for (int c = 0; c < 8; ++c) aov[c] = 0.f;
    const float lambda = Spectrum<float>::sampleOne(sampleNext1D<float>(sg));
    // Camera.computeRayPinhole (Camera.slang:46-78)
    float sx = (float(px) + 0.5f) / float(o.width), sy = (float(py) + 0.5f) / float(o.height);
    sx += -s.jitterX;
    sy += s.jitterY;
    const float nx = 2.f * sx + -1.f, ny = -2.f * sy + 1.f;
    V3<F> org = s.camPos;
    V3<F> dir = normalize(s.camU * nx + s.camV * ny + s.camW);
    float weight = 1, pdf = 1, thp = 1, lastNonDeltaPdf = 0;
    uint32_t flags = RF_OnlyDelta;
    for (uint32_t len = 0; len < o.maxBounces + 1;) {
        ORay ray;
        ray.o[0] = org.x; ray.o[1] = org.y; ray.o[2] = org.z;
        ray.d[0] = dir.x; ray.d[1] = dir.y; ray.d[2] = dir.z;
        ray.tmin = 0.f; ray.tmax = 3.402823466e+38f;
        const OHit h = s.bvh.closest(s.mesh, ray, false);
        const V3<F> wi = -dir;
        Bounce& b = rec[len];
        if (h.tri == 0xFFFFFFFFu) {
            flags |= RF_Miss;
            b = Bounce{};
            b.flags = flags; b.encWi = ndirToOct(wi); b.lastNonDeltaPdf = lastNonDeltaPdf; b.pathContribution = 1.f / (pdf * thp);
            break;
        }
        const SD<F> sd = shadingData(s, h.tri, h.b1, h.b2, wi);
        const Instance<F> in = instanceAt(s, sd.materialID);
        const uint32_t props = getPropertyFlags(in);
        float rcp = 1.f;
        if (o.doRR) {
            const float u = sampleNext1D<float>(sg);
            const float p = std::pow(1.f - weight, 2.f);
            rcp = u < p ? 0.f : 1.f - p;
        }
        bool end = len == o.maxBounces || (rcp == 0.f && len > 0);
        BSDFSample bs;
        if (!end) {
            end = !materialSample(in, sd, sg, bs, lambda) || bs.pdf == 0.f;
            end |= pdf * bs.pdf * thp * rcp < 1e-6f;
        }
        const bool isDelta = (bs.lobe & Lobe_Delta) != 0, isDisp = (bs.lobe & Lobe_DispersiveDelta) != 0, isRefl = (bs.lobe & Lobe_Reflection) != 0;
        const bool isMS = (isRefl && (props & Prop_MWR)) || (!isRefl && (props & Prop_MWT));
        if (aov && !isDelta && (flags & RF_OnlyDelta)) {
            const V3<F> alb = materialAlbedo(in);
            const V3<F> N = sd.f.N;
            aov[0] = alb.x; aov[1] = alb.y; aov[2] = alb.z; aov[3] = 1.f;
            aov[4] = dot(s.viewRow[0], N); aov[5] = dot(s.viewRow[1], N); aov[6] = dot(s.viewRow[2], N); aov[7] = 1.f;
        }
        b.tri = h.tri; b.b1 = h.b1; b.b2 = h.b2; b.encWi = ndirToOct(wi); b.lastNonDeltaPdf = lastNonDeltaPdf;
        if (end) {
            flags |= RF_Terminated;
            b.flags = flags; b.lobe = 0; b.pathContribution = 1.f / (pdf * thp);
            break;
        }
        weight *= bs.weight;
        thp *= rcp;
        pdf *= bs.pdf;
        if (isDisp) flags |= RF_Dispersive;
        b.flags = flags; b.lobe = bs.lobe; b.pathContribution = 1.f / (pdf * thp);
        ++len;
        lastNonDeltaPdf = !isDelta ? bs.pdf : lastNonDeltaPdf;
        const uint32_t msc = isMS ? (flags >> 16) + 1u : 0u;
        flags = (flags & 0xffffu) | (msc << 16);
        flags = isDelta ? (flags | RF_LastDelta) : (flags & ~RF_LastDelta);
        if (!isDelta) flags &= ~RF_OnlyDelta;
        org = rayOrigin(sd.posW, (sd.frontFacing == isRefl) ? sd.faceN : -sd.faceN);
        dir = bs.wo;
    }
}

// ---------------------------------------------------------------- solve pass
struct SolveData { V3<F> Lrgb{0, 0, 0}; V3<F> pathOrigin, pathNormal; float wavelength; };

// forwardUnwindStep (pltpt_solve.rt.slang:48-66); posW optionally overridden (manifold vertices).
void unwindStep(const OScene& s, const Options& o, uint32_t tri, float b1, float b2, const V3<F>* posW, V3<F> wo, uint32_t lobe, Beam& beam) {
    const SD<F> sd = shadingData(s, tri, b1, b2, -beam.r, posW);
    beam.propagate(sd.posW);
    const Instance<F> in = instanceAt(s, sd.materialID);
    materialEval<F, HW>(in, sd, wo, lobe, beam);
    if (o.hwssMIS) {
        float pd[HW];
        materialLambdaPdf<F, HW>(in, sd, wo, lobe, beam, pd);
        for (int j = 0; j < HW; ++j) beam.spectralPDFs[j] *= pd[j];
    }
}
void forwardUnwind(const OScene& s, const Options& o, const std::vector<Bounce>& rec, int bounceIdx, Beam& beam) {
    for (int p = bounceIdx - 1; p >= 0; --p) {
        const Bounce& b = rec[size_t(p)];
        unwindStep(s, o, b.tri, b.b1, b.b2, nullptr, b.wi(), b.lobe, beam);
    }
}

// ---------------------------------------------------------------- manifold NEE (Manifolds.slangh)
struct MSHit { uint32_t tri; float b1, b2; V3<F> posW; uint32_t lobe; };
struct MSData {  // ManifoldSampleData (:57-81)
    MSHit hits[kMaxMSOccluders];
    V3<F> woLastHitToLight{0, 0, 0};
    int hitCount = 0;
    int iterations = 0;
    float wavelength = 0, pdf = 1;
    bool dispersive = false;
};
OHit shadowRayClosest(const OScene& s, V3<F> org, V3<F> dir, float tmin, float tmax) {
    ORay r;
    r.o[0] = org.x; r.o[1] = org.y; r.o[2] = org.z;
    r.d[0] = dir.x; r.d[1] = dir.y; r.d[2] = dir.z;
    r.tmin = tmin; r.tmax = tmax;
    return s.bvh.closest(s.mesh, r, false);
}
bool visible(const OScene& s, V3<F> org, V3<F> dir, float tmin, float tmax) {
    ORay r;
    r.o[0] = org.x; r.o[1] = org.y; r.o[2] = org.z;
    r.d[0] = dir.x; r.d[1] = dir.y; r.d[2] = dir.z;
    r.tmin = tmin; r.tmax = tmax;
    return s.bvh.closest(s.mesh, r, true).tri == 0xFFFFFFFFu;
}
DSG<F> dsgAt(const OScene& s, uint32_t tri, float b1, float b2) {
    V3<F> P[3], Nv[3];
    for (int k = 0; k < 3; ++k) {
        const uint32_t v = s.idx[3 * size_t(tri) + k];
        P[k] = V3<F>{s.pos[3 * size_t(v)], s.pos[3 * size_t(v) + 1], s.pos[3 * size_t(v) + 2]};
        Nv[k] = V3<F>{s.nrm[3 * size_t(v)], s.nrm[3 * size_t(v) + 1], s.nrm[3 * size_t(v) + 2]};
    }
    return makeDSG<F>(P, Nv, b1, b2);
}
template <class S>
bool manifoldSampleVertex(uint32_t tri, float b1, float b2, const SD<F>& sd, const Instance<F>& in, V3<F> lightDir, S& sg, MSData& d) {  // :84-105
    BSDFSample smp;
    if (!materialSampleHinted(in, sd, sg, smp, d.wavelength, Lobe_Delta, lightDir)) return false;
    d.pdf *= smp.pdf;
    d.hits[d.hitCount] = MSHit{tri, b1, b2, sd.posW, smp.lobe};
    d.dispersive |= (smp.lobe & Lobe_DispersiveDelta) != 0;
    ++d.hitCount;
    return true;
}
template <class S>
bool traceMNEEOccluders(const OScene& s, const Options& o, V3<F> org, V3<F> dir, float tmax, int maxMSBounces, S& sg, MSData& d) {  // :107-150
    const V3<F> wi = -dir;
    for (;;) {
        const OHit h = shadowRayClosest(s, org, dir, 0.f, tmax);
        if (h.tri == 0xFFFFFFFFu) return true;
        if (d.hitCount >= o.mneeMaxOccluders || d.hitCount >= maxMSBounces) return false;
        const V3<F> posW = org + dir * h.t;
        const SD<F> sd = shadingData(s, h.tri, h.b1, h.b2, wi, &posW);
        const Instance<F> in = instanceAt(s, sd.materialID);
        if (!(getPropertyFlags(in) & Prop_MWT)) return false;
        if (!manifoldSampleVertex(h.tri, h.b1, h.b2, sd, in, -wi, sg, d) || (d.hits[d.hitCount - 1].lobe & Lobe_Transmission) == 0) return false;
        org = rayOrigin(sd.posW, dot(sd.faceN, dir) >= 0.f ? sd.faceN : -sd.faceN);
        tmax -= h.t;
    }
}
// sphcoords (:314-319) in FP64 of the FP32 direction.
V2<double> sphcoords(V3<F> w) {
    const double theta = walkAcos(std::fmin(std::fmax(double(w.z), -1.0), 1.0));
    double phi = walkAtan2(double(w.y), double(w.x));
    phi += phi < 0.0 ? 6.28318530717958647692 : 0.0;
    return {theta, phi};
}
// Newton step on the angle-difference constraint (:283-312), in FP64 (the north-star precision of
// the step; the reference's is FP32). Matrices as rows (dTheta, dPhi); the step is rounded back
// to FP32 where the walk moves its vertex.
bool manifoldOneStep(const PDV<F>& w, const KPDV<F>& ws, const PDV<F>& p, float& lenC, V3<F>& stepDir) {
    const V2<double> so = sphcoords(w.v), si = sphcoords(ws.v);
    const double dt = so.x - si.x;
    double dp = so.y - si.y;
    dp = std::fabs(dp) > 3.14159265358979323846 ? dp - (dp > 0.0 ? 6.28318530717958647692 : -6.28318530717958647692) : dp;
    lenC = float(std::sqrt(dt * dt + dp * dp));
    auto derivs = [](V3<F> v, V3<F> ddu, V3<F> ddv, double out[2][2]) {  // :321-333
        const double vx = v.x, vy = v.y, vz = v.z;
        const double dAcos = -1.0 / std::sqrt(std::fmax(0.0, 1.0 - vz * vz));
        const double invNorm = 1.0 / (vx * vx + vy * vy);
        out[0][0] = dAcos * double(ddu.z); out[0][1] = dAcos * double(ddv.z);
        out[1][0] = invNorm * (vx * double(ddu.y) - vy * double(ddu.x));
        out[1][1] = invNorm * (vx * double(ddv.y) - vy * double(ddv.x));
    };
    double a[2][2], b[2][2];
    derivs(w.v, w.ddu, w.ddv, a);
    derivs(ws.v, ws.ddu, ws.ddv, b);  // the wavelength column of the 2x3 form is not used
    const double j00 = a[0][0] - b[0][0], j01 = a[0][1] - b[0][1], j10 = a[1][0] - b[1][0], j11 = a[1][1] - b[1][1];
    // tryInvert (MathHelpers.slang:484-496)
    const double det = j00 * j11 - j01 * j10;
    if (std::fabs(det) < 4e-6) return false;
    const double id = 1.0 / det;
    const double i00 = j11 * id, i11 = j00 * id, i01 = -j01 * id, i10 = -j10 * id;
    const double dx = i00 * dt + i01 * dp, dy = i10 * dt + i11 * dp;
    stepDir = V3<F>{float(double(p.ddu.x) * dx + double(p.ddv.x) * dy), float(double(p.ddu.y) * dx + double(p.ddv.y) * dy),
                    float(double(p.ddu.z) * dx + double(p.ddv.z) * dy)};
    return true;
}
template <class S>
bool manifoldNEESample(const OScene& s, const Options& o, V3<F> visOrg, V3<F> visDir, float visTmax, V3<F> origin, V3<F> lightPos,
                       V3<F> lightDir, bool distantLight, int maxMSBounces, S& sg, MSData& d) {  // :165-281
    if (!traceMNEEOccluders(s, o, visOrg, visDir, visTmax, maxMSBounces, sg, d)) return false;
    if (d.hitCount == 0) return true;
    float beta[kMaxMSOccluders];
    for (int i = 0; i < kMaxMSOccluders; ++i) beta[i] = 1.f;
    bool success = false;
    while (!success) {
        success = true;
        for (int idx = 0; idx < d.hitCount; ++idx) {
            MSHit& h = d.hits[idx];
            const DSG<F> dsg = dsgAt(s, h.tri, h.b1, h.b2);
            const V3<F> p0 = idx == 0 ? origin : d.hits[idx - 1].posW;
            const PDV<F> pdv0 = dsg.pdw(p0);
            const bool last = idx + 1 == d.hitCount;
            const V3<F> p2 = last ? lightPos : d.hits[idx + 1].posW;
            // a distant light fixes the last direction (:199-202)
            const PDV<F> pdv2 = last && distantLight ? PDV<F>{lightDir, V3<F>{0, 0, 0}, V3<F>{0, 0, 0}} : dsg.pdw(p2);
            const bool woIsIn = dot(pdv2.v, dsg.n.v) < F(0);
            const PDV<F> wi = woIsIn ? pdv0 : pdv2, wo = woIsIn ? pdv2 : pdv0;
            const SD<F> sd = shadingData(s, h.tri, h.b1, h.b2, wi.v);
            const Instance<F> in = instanceAt(s, sd.materialID);
            const KPDV<F> wos = materialPartialDerivatives(in, wi, dsg, h.lobe, d.wavelength);
            if (wos.v.x == F(0) && wos.v.y == F(0) && wos.v.z == F(0)) return false;
            float lenC;
            V3<F> stepDir;
            if (!manifoldOneStep(wo, wos, dsg.p, lenC, stepDir)) return false;
            if (lenC < o.mneeThreshold) continue;
            success = false;
            for (;;) {  // line search: project the proposal onto the surface
                if (++d.iterations > o.mneeMaxIterations) return false;
                const V3<F> proposal = h.posW - stepDir * beta[idx];
                V3<F> dirTo = proposal - p0;
                const F dist = length(dirTo);
                dirTo = dirTo / dist;
                const OHit ch = shadowRayClosest(s, p0, dirTo, std::fmax(1e-4f, .99f * dist), 1.01f * dist);
                if (ch.tri == 0xFFFFFFFFu || s.triMesh[ch.tri] != s.triMesh[h.tri]) { beta[idx] /= 4.f; continue; }
                h.tri = ch.tri; h.b1 = ch.b1; h.b2 = ch.b2;
                h.posW = p0 + dirTo * ch.t;
                beta[idx] = std::fmin(1.f, 2.f * beta[idx]);
                break;
            }
        }
    }
    const V3<F> p = d.hits[d.hitCount - 1].posW;
    V3<F> dir;
    F dist;
    if (!distantLight) {
        dir = lightPos - p;
        dist = length(dir);
        dir = dir / dist;
    } else {
        dir = lightDir;
        dist = std::numeric_limits<float>::max();
    }
    d.woLastHitToLight = dir;
    return visible(s, p, dir, 1e-4f, dist);
}
void manifoldForwardUnwind(const OScene& s, const Options& o, const MSData& d, V3<F> origin, Beam& beam) {  // :80-93
    for (int idx = d.hitCount - 1; idx >= 0; --idx) {
        const MSHit& h = d.hits[idx];
        const V3<F> p2 = idx == 0 ? origin : d.hits[idx - 1].posW;
        unwindStep(s, o, h.tri, h.b1, h.b2, &h.posW, normalize(p2 - h.posW), h.lobe, beam);
    }
}
// The camera polarizer of measure (pltpt_solve.rt.slang:111-126): the beam rotated to the camera
// frame (t = U, r = -W), then (1 - p) I + p MuellerPolarizer(theta) on each HWSS component.
void applyCameraPolarizer(const OScene& s, Beam& beam) {
    if (!(s.polarizer > .0f)) return;
    beam.rotateFrame(normalize(s.camU), -normalize(s.camW));
    const M4<F> Pm = MuellerPolarizer(s.polarizerTheta);
    M4<F> P;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k) P.m[i][k] = (1.f - s.polarizer) * (i == k ? 1.f : 0.f) + s.polarizer * Pm.m[i][k];
    for (int j = 0; j < beam.hwss_comps; ++j) beam.setSP(j, mul(P, beam.getSP(j)));
}
void measure(const OScene& s, const Options& o, SolveData& sv, int bounceIdx, float pathContribution, Beam& beam) {
    if (!o.directLights && bounceIdx <= 1) return;
    beam.propagate(s.camPos);
    applyCameraPolarizer(s, beam);
    if (o.hwssMIS) {
        float sum = 0;
        for (int j = 0; j < HW; ++j) sum += beam.spectralPDFs[j];
        pathContribution *= sum > 1.192092896e-07f ? beam.spectralPDFs[0] / sum : 0.f;
    } else pathContribution *= 1.f / float(beam.hwss_comps);
    const V3<F> rgb = Spectrum<float>::spectrumToRgb(beam.L, beam.wavelengths, unsigned(beam.hwss_comps)) * pathContribution;
    const V3<F> L = {std::fmax(0.f, rgb.x), std::fmax(0.f, rgb.y), std::fmax(0.f, rgb.z)};
    if (std::isfinite(L.x) && std::isfinite(L.y) && std::isfinite(L.z)) sv.Lrgb = sv.Lrgb + L;
}

template <class S>
void solvePass(const OScene& s, const Options& o, uint32_t px, uint32_t py, const std::vector<Bounce>& rec, float* out) {
    S sg(px, py, o.frame);
    // the per-tile generator for NEE light-type selection (pltpt_solve.rt.slang:412)
    S tilesg(px / o.tileSize, py / o.tileSize, o.frame << 2);
    SolveData sv;
    sv.wavelength = Spectrum<float>::sampleOne(sampleNext1D<float>(sg));
    sv.pathOrigin = s.camPos;
    sv.pathNormal = normalize(s.camW);
    // Scene::useEmissiveLights needs an active (flux > 0) emissive triangle
    const bool useEmissive = o.emissiveLights && !s.activeLights.empty();
    const std::vector<float>& selPdfs = o.uniformEmissive ? s.lightPdfUniform : s.lightPdf;
    const bool useAnalytic = o.analyticLights && !s.analytic.empty();
    const bool useEnv = o.envLight && s.envW > 0;
    // getLightTypeSelectionProbabilities (LightSampling.slangh:37-55)
    float pEnv = useEnv ? 1.f : 0.f, pEmissive = useEmissive ? 1.f : 0.f, pAnalytic = useAnalytic ? 1.f : 0.f;
    {
        const float sum = pEnv + pEmissive + pAnalytic;
        if (sum != 0.f) { const float inv = 1.f / sum; pEnv *= inv; pEmissive *= inv; pAnalytic *= inv; }
    }
    const bool doNEE = o.doNEE != 0, doMIS = doNEE && o.doMIS;
    for (int p = 0;; ++p) {
        const Bounce& bounce = rec[size_t(p)];
        if (bounce.flags & RF_Miss) {
            // onMiss (pltpt_solve.rt.slang:371-396)
            if (useEnv) {
                const bool didNEE = doNEE && p > 0 &&
                                    (!(bounce.flags & RF_LastDelta) || (o.doMNEE && int(bounce.flags >> 16) <= o.mneeMaxOccluders && p > 1));
                if (doMIS || !didNEE) {
                    const bool needMIS = doMIS && didNEE;
                    const V3<F> wo = bounce.wi();
                    const float lightPdf = pEnv * envPdf(s, -wo);
                    const float mis = needMIS ? bounce.lastNonDeltaPdf / (bounce.lastNonDeltaPdf + lightPdf) : 1.f;
                    float lam[HW], Le[HW], rgb[3];
                    hwssLambdas(sv.wavelength, lam);
                    envEval(s, -wo, rgb);
                    for (int j = 0; j < HW; ++j) Le[j] = mis * Spectrum<float>::rgbToSpectrum(V3<F>{rgb[0], rgb[1], rgb[2]}, lam[j]);
                    const int comps = (!o.hwssMIS && (bounce.flags & RF_Dispersive)) ? 1 : HW;
                    Beam beam = sourceBeamDistant(wo, o.maxOmega, Le, lam, o.maxOmega, comps);
                    forwardUnwind(s, o, rec, p, beam);
                    measure(s, o, sv, p, bounce.pathContribution, beam);
                }
            }
            break;
        }
        const V3<F> bwi = bounce.wi();
        const SD<F> sd = shadingData(s, bounce.tri, bounce.b1, bounce.b2, bwi);
        const Instance<F> bsdf = instanceAt(s, sd.materialID);
        const uint32_t props = getPropertyFlags(bsdf);
        if (doNEE) {
            const uint32_t lobes = getLobes(bsdf);
            // a delta reflector can take part in a manifold walk from here (pltpt_solve.rt.slang:254-258)
            const bool allowMS = o.doMNEE && p > 0 && !(bounce.flags & RF_LastDelta) && (props & Prop_MWR);
            const bool sampleRefl = (lobes & Lobe_NonDeltaReflection) != 0 || allowMS;
            const bool sampleTrans = (lobes & Lobe_NonDeltaTransmission) != 0;
            if (sampleRefl || sampleTrans) {
                // selectLightType (LightSampling.slangh:57-85), on the per-tile generator if asked (:181-185)
                float u = o.perTileSG ? sampleNext1D<float>(tilesg) : sampleNext1D<float>(sg);
                int ltype = -1;
                float typePdf = 0.f;
                if (u < pEnv) {
                    ltype = 0; typePdf = pEnv;
                } else {
                    u -= pEnv;
                    if (u < pEmissive) {
                        ltype = 1; typePdf = pEmissive;
                    } else {
                        u -= pEmissive;
                        if (u < pAnalytic) { ltype = 2; typePdf = pAnalytic; }
                    }
                }
                bool valid = false;
                float lsPdf = 0, lsDist = 0;
                V3<F> lsDir{0, 0, 0}, lsOrigin{0, 0, 0}, tPos{0, 0, 0};
                V3<F> bPos{0, 0, 0};
                float bSize = o.sourcingArea;
                bool bDistant = false;
                float Le[HW], lam[HW];
                for (int j = 0; j < HW; ++j) { Le[j] = 0; lam[j] = sv.wavelength; }
                if (ltype == 1) {
                    const float uL = sampleNext1D<float>(sg);
                    uint32_t ti;
                    if (o.uniformEmissive) {
                        // EmissiveUniformSampler.sampleLight (EmissiveUniformSampler.slang:52-72)
                        const uint32_t na = uint32_t(s.activeLights.size());
                        ti = s.activeLights[std::min(uint32_t(uL * float(na)), na - 1)];
                    } else {
                        // EmissivePowerSampler.sampleLight
                        const uint32_t nl = uint32_t(s.lightTri.size());
                        ti = std::min(uint32_t(uL * float(nl)), nl - 1);
                        const uint32_t ax = s.aliasX[ti], ay = s.aliasY[ti];
                        const float threshold = f16_bits_to_f32(uint16_t(ax >> 16u));
                        const uint32_t above = ((ax & 0xFFFFu) << 8u) | ((ay >> 24u) & 0xFFu), below = ay & 0xFFFFFFu;
                        ti = sampleNext1D<float>(sg) >= threshold ? above : below;
                    }
                    const float selPdf = selPdfs[ti];
                    // sampleTriangle
                    const V2<F> u2 = sampleNext2D<float>(sg);
                    const float su = std::sqrt(u2.x);
                    const V3<F> bary = {1.f - (1.f - su) - u2.y * su, 1.f - su, u2.y * su};
                    const uint32_t t = s.lightTri[ti];
                    V3<F> pv[3];
                    for (int k = 0; k < 3; ++k) {
                        const uint32_t v = s.idx[3 * size_t(t) + k];
                        pv[k] = V3<F>{s.pos[3 * size_t(v)], s.pos[3 * size_t(v) + 1], s.pos[3 * size_t(v) + 2]};
                    }
                    tPos = pv[0] * bary.x + pv[1] * bary.y + pv[2] * bary.z;
                    const V3<F> toL = tPos - sd.posW;
                    const float d2 = std::fmax(kFLT_MIN, dot(toL, toL));
                    const float dist = std::sqrt(d2);
                    const V3<F> ldir = toL / dist;
                    const V3<F> ln = s.lightN[ti];
                    const float cosT = dot(ln, -ldir);
                    if (cosT > 0.f) {
                        const MatData& em = s.mats[s.meshMat[s.triMesh[t]]];
                        const SpectralProfile* ep = em.emission < s.nprof ? s.prof + em.emission : nullptr;
                        sampleSpectrum(ep, sv.wavelength, sg, o.isEmitters != 0, Le, lam);
                        for (int j = 0; j < HW; ++j) Le[j] *= Spectrum<float>::rgbToSpectrum(V3<F>{1, 1, 1}, lam[j]);
                        float tpdf = d2 / std::fmax(kFLT_MIN, cosT * s.lightArea[ti]);
                        tpdf *= selPdf;
                        for (int j = 0; j < HW; ++j) Le[j] = tpdf > 0.f ? Le[j] / tpdf : 0.f;
                        lsPdf = tpdf;
                        const V3<F> lightPos = rayOrigin(tPos, ln);
                        const V3<F> rd = lightPos - sd.posW;
                        lsOrigin = rayOrigin(sd.posW, dot(sd.faceN, rd) >= 0.f ? sd.faceN : -sd.faceN);
                        const V3<F> tl = lightPos - lsOrigin;
                        lsDist = length(tl);
                        lsDir = normalize(tl);
                        bPos = tPos;
                        valid = false;
                        for (int j = 0; j < HW; ++j) valid |= Le[j] > 0.f;
                    }
                }
                if (ltype == 0) {
                    // generateEnvMapSample (LightSampling.slangh:87-104)
                    const float u0 = sampleNext1D<float>(sg), u1 = sampleNext1D<float>(sg);
                    float epdf, rgb[3];
                    envSample(s, u0, u1, lsDir, epdf);
                    hwssLambdas(sv.wavelength, lam);
                    envEval(s, lsDir, rgb);
                    bool anyL = false;
                    for (int j = 0; j < HW; ++j) {
                        const float le = Spectrum<float>::rgbToSpectrum(V3<F>{rgb[0], rgb[1], rgb[2]}, lam[j]);
                        Le[j] = epdf > 0.f ? le / epdf : 0.f;
                        anyL |= Le[j] > 0.f;
                    }
                    lsPdf = epdf;
                    lsOrigin = rayOrigin(sd.posW, dot(sd.faceN, lsDir) >= 0.f ? sd.faceN : -sd.faceN);
                    lsDist = std::numeric_limits<float>::max();
                    bDistant = true;
                    bSize = o.maxOmega;
                    valid = anyL;
                }
                if (ltype == 2) {
                    valid = sampleAnalytic(s, o, sd, sv.wavelength, sg, lsDir, lsOrigin, lsDist, lsPdf, bPos, bSize, bDistant, Le, lam);
                    if (valid) {
                        bool anyL = false;
                        for (int j = 0; j < HW; ++j) anyL |= Le[j] > 0.f;
                        valid = anyL;
                    }
                }
                lsPdf *= typePdf;
                const bool isRefl = dot(sd.faceN, lsDir) * dot(sd.faceN, sd.V) > 0.f;
                valid = valid && lsPdf > 1e-6f && !(!sampleRefl && isRefl) && !(!sampleTrans && !isRefl);
                bool anyL = false;
                for (int j = 0; j < HW; ++j) anyL |= Le[j] > 0.f;
                if (valid && anyL) {
                    const int comps = (!o.hwssMIS && (bounce.flags & RF_Dispersive)) ? 1 : HW;
                    Beam beam = bDistant ? sourceBeamDistant(-lsDir, bSize, Le, lam, o.maxOmega, comps)
                                         : sourceBeamArea(bPos, -lsDir, bSize, lsDist, Le, lam, o.maxOmega, comps);
                    beam.scaleAll(1.f / typePdf);
                    const bool misL = doMIS && ltype != 2;  // no MIS for analytic lights (pltpt_solve.rt.slang:275)
                    bool evalCurrent = true, ok = true;
                    const int maxMSBounces = int(o.maxBounces) - p;
                    if (!o.doMNEE || maxMSBounces <= 0) {
                        ok = visible(s, lsOrigin, lsDir, 0.f, lsDist);
                        if (ok && misL) {
                            const float bp = materialPdf(bsdf, sd, lsDir, beam.wavelengths[0]);
                            beam.scaleAll(lsPdf / (lsPdf + bp));
                        }
                    } else {  // manifold sampling on NEE (pltpt_solve.rt.slang:296-366)
                        const V3<F> lightDir = -beam.r;
                        V3<F> dirL = lsDir;
                        float misBsdfPdf = 0.f;
                        MSData ms;
                        ms.wavelength = beam.wavelengths[0];
                        const bool isRefl2 = dot(sd.f.N, lightDir) * dot(sd.f.N, sd.V) > 0.f;
                        const bool currentVertexMS = isRefl2 && allowMS;
                        if (currentVertexMS)
                            ok = manifoldSampleVertex(bounce.tri, bounce.b1, bounce.b2, sd, bsdf, lightDir, sg, ms) &&
                                 (ms.hits[0].lobe & Lobe_Reflection) != 0;
                        const V3<F> origin = !currentVertexMS ? sd.posW : sv.pathOrigin;
                        ok = ok && manifoldNEESample(s, o, lsOrigin, lsDir, lsDist, origin, beam.posW, lightDir, beam.distant,
                                                     maxMSBounces, sg, ms);
                        ok = ok && !(ms.pdf < 1e-6f);
                        if (ok && ms.hitCount > 0) {
                            if (!o.hwssMIS && ms.dispersive) beam.hwss_comps = 1;
                            if (!beam.distant) {
                                dirL = ms.woLastHitToLight;
                                beam.r = -ms.woLastHitToLight;
                                beam.t = perp_stark(beam.r);
                            }
                            beam.scaleAll(1.f / ms.pdf);
                            manifoldForwardUnwind(s, o, ms, origin, beam);
                            if (currentVertexMS) { evalCurrent = false; misBsdfPdf = bounce.lastNonDeltaPdf; }
                        }
                        if (ok && misL) {
                            if (ms.hitCount == 0 || !currentVertexMS) misBsdfPdf = materialPdf(bsdf, sd, dirL, beam.wavelengths[0]);
                            beam.scaleAll(lsPdf / (lsPdf + misBsdfPdf * ms.pdf));
                        }
                    }
                    if (ok) {
                        // evalNEE (pltpt_solve.rt.slang:219-243)
                        if (evalCurrent) {
                            const SD<F> fsd = shadingData(s, bounce.tri, bounce.b1, bounce.b2, -beam.r);
                            beam.propagate(fsd.posW);
                            materialEval<F, HW>(bsdf, fsd, bwi, Lobe_NonDelta, beam);
                            if (o.hwssMIS) {
                                float pd[HW];
                                materialLambdaPdf<F, HW>(bsdf, fsd, bwi, Lobe_NonDelta, beam, pd);
                                for (int j = 0; j < HW; ++j) beam.spectralPDFs[j] *= pd[j];
                            }
                        }
                        forwardUnwind(s, o, rec, p, beam);
                        measure(s, o, sv, p, bounce.pathContribution, beam);
                    }
                }
            }
        }
        // evalEmissive
        // wasNEEAvailable (pltpt_solve.rt.slang:205-210)
        const bool didNEE = doNEE && p > 0 &&
                            (!(bounce.flags & RF_LastDelta) || (o.doMNEE && int(bounce.flags >> 16) <= o.mneeMaxOccluders && p > 1));
        if (useEmissive && (props & Prop_Emissive) && (doMIS || !didNEE)) {
            const bool needMIS = doMIS && didNEE;
            float lightPdf = 0.f;
            const uint32_t lt = s.triToLight[bounce.tri];
            if (lt != 0xFFFFFFFFu) {
                const V3<F> hn = sd.frontFacing ? sd.faceN : -sd.faceN;
                const V3<F> toL = sd.posW - sv.pathOrigin;
                const float d2 = dot(toL, toL);
                if (d2 > kFLT_MIN) {
                    const V3<F> L = toL / std::sqrt(d2);
                    const float ct = dot(hn, -L);
                    if (ct > 0.f) lightPdf = pEmissive * (selPdfs[lt] * (d2 / std::fmax(kFLT_MIN, ct * s.lightArea[lt])));
                }
            }
            const float mis = needMIS ? bounce.lastNonDeltaPdf / (bounce.lastNonDeltaPdf + lightPdf) : 1.f;
            const SpectralProfile* ep = bsdf.emission < s.nprof ? s.prof + bsdf.emission : nullptr;
            float Le[HW], lam[HW];
            sampleSpectrum(ep, sv.wavelength, sg, o.isEmitters != 0, Le, lam);
            bool anyL = false;
            for (int j = 0; j < HW; ++j) { Le[j] *= mis * Spectrum<float>::rgbToSpectrum(V3<F>{1, 1, 1}, lam[j]); anyL |= Le[j] > 0.f; }
            if (anyL) {
                const float dist = length(sv.pathOrigin - sd.posW);
                const int comps = (!o.hwssMIS && (bounce.flags & RF_Dispersive)) ? 1 : HW;
                Beam beam = sourceBeamArea(sd.posW, bwi, o.sourcingArea, dist, Le, lam, o.maxOmega, comps);
                forwardUnwind(s, o, rec, p, beam);
                measure(s, o, sv, p, bounce.pathContribution, beam);
            }
        }
        if ((bounce.flags & RF_Terminated) || p == int(o.maxBounces)) break;
        sv.pathOrigin = sd.posW;
        sv.pathNormal = sd.f.N;
    }
    out[0] = sv.Lrgb.x; out[1] = sv.Lrgb.y; out[2] = sv.Lrgb.z; out[3] = 1.f;
}

void buildLights(OScene& s) {
    s.triToLight.assign(s.ntri, 0xFFFFFFFFu);
    std::vector<float> flux;
    for (uint32_t t = 0; t < s.ntri; ++t) {
        const MatData& m = s.mats[s.meshMat[s.triMesh[t]]];
        if (m.type != MT_DIFFUSE || m.emission == kNoProfile) continue;
        V3<F> p[3];
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = s.idx[3 * size_t(t) + k];
            p[k] = V3<F>{s.pos[3 * size_t(v)], s.pos[3 * size_t(v) + 1], s.pos[3 * size_t(v) + 2]};
        }
        const V3<F> c = cross(p[1] - p[0], p[2] - p[0]);
        const float len = length(c);
        s.triToLight[t] = uint32_t(s.lightTri.size());
        s.lightTri.push_back(t);
        s.lightArea.push_back(0.5f * len);
        s.lightN.push_back(c / len);
        flux.push_back(0.5f * len * float(M_PI));  // luminance x area x pi; one emissive profile -> luminance cancels
    }
    const uint32_t N = uint32_t(flux.size());
    if (!N) return;
    // LightCollection::updateActiveTriangleList (LightCollection.cpp:497-523): flux > 0
    for (uint32_t i = 0; i < N; ++i)
        if (flux[i] > 0.f) s.activeLights.push_back(i);
    const float uni = s.activeLights.empty() ? 0.f : 1.f / float(s.activeLights.size());
    for (uint32_t i = 0; i < N; ++i) s.lightPdfUniform.push_back(flux[i] > 0.f ? uni : 0.f);
    // EmissivePowerSampler::generateAliasTable (EmissivePowerSampler.cpp:86-176)
    double sum = 0;
    for (float f : flux) sum += f;
    const float inv = 1.f / float(sum);
    for (float f : flux) s.lightPdf.push_back(f * inv);
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
    std::uniform_int_distribution<uint32_t> dist;
    for (uint32_t i = 0; i < N; ++i) {
        const uint32_t dst = i + (dist(rng) % (N - i));
        std::swap(thr[i], thr[dst]); std::swap(redirect[i], redirect[dst]); std::swap(perm[i], perm[dst]);
    }
    for (uint32_t i = 0; i < N; ++i) {
        const uint32_t prob = uint32_t(f32_to_f16_bits(thr[i])) << 16u;
        const uint32_t lx = redirect[i] & 0xFFFFFFu, ly = perm[i] & 0xFFFFFFu;
        s.aliasX.push_back(prob | ((lx >> 8u) & 0xFFFFu));
        s.aliasY.push_back(((lx & 0xFFu) << 24u) | ly);
    }
}
}  // namespace

namespace {
// One material query through the material interface in a local frame with N = +z (tests):
// `mat` is a plt_material block, wi = sd.V (towards the light), wo = toEye, 4 HWSS wavelengths,
// 4 incident Stokes vectors (16 values, component-major), coherence area `area` (um^2) for the
// beam (r = -wi, r_m = 1 mm). Writes the 4 outgoing Stokes vectors (16), the pdf of sampling
// `wi` from `wo` (1) and lambdaPdf (4). dbl selects the FP64 instantiation.
template <class T>
static void material_eval_t(const void* mat, const void* prof, uint32_t nprof, const double* wi3, const double* wo3,
                            uint32_t lobe, const double* lambda, const double* stokes, double area, double* out,
                            double* pdf, double* lpdf) {
    const Instance<T> in = setupMaterialInstance<T>(*static_cast<const MatData*>(mat), static_cast<const SpectralProfile*>(prof), nprof);
    SD<T> sd;
    sd.f = Frame<T>{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    sd.posW = {0, 0, 0};
    sd.faceN = {0, 0, 1};
    sd.V = {T(wi3[0]), T(wi3[1]), T(wi3[2])};
    sd.frontFacing = true;
    sd.materialID = 0;
    const V3<T> wo = {T(wo3[0]), T(wo3[1]), T(wo3[2])};
    PLTBeam<T, 4> beam;
    for (int j = 0; j < 4; ++j) {
        beam.L[j] = T(stokes[4 * j]); beam.L1[j] = T(stokes[4 * j + 1]); beam.L2[j] = T(stokes[4 * j + 2]); beam.L3[j] = T(stokes[4 * j + 3]);
        beam.wavelengths[j] = T(lambda[j]);
        beam.spectralPDFs[j] = 1;
    }
    beam.hwss_comps = 4;
    const T d = T(M_PI) * sqr(T(1e-3) * T(lambda[0])) / T(area);
    beam.Theta.D[0] = d; beam.Theta.D[1] = 0; beam.Theta.D[2] = d;
    beam.Theta.r_m = T(1e-3);
    beam.posW = {0, 0, 0};
    beam.r = -sd.V;
    beam.t = perp_stark(beam.r);
    beam.distant = false;
    T lp[4];
    materialLambdaPdf<T, 4>(in, sd, wo, lobe, beam, lp);
    materialEval<T, 4>(in, sd, wo, lobe, beam);
    for (int j = 0; j < 4; ++j) {
        out[4 * j] = double(beam.L[j]); out[4 * j + 1] = double(beam.L1[j]);
        out[4 * j + 2] = double(beam.L2[j]); out[4 * j + 3] = double(beam.L3[j]);
        lpdf[j] = double(lp[j]);
    }
    SD<T> sd2 = sd;
    sd2.V = wo;
    *pdf = double(materialPdf<T>(in, sd2, sd.V, T(lambda[0])));
}

}  // namespace

extern "C" {

// Camera: pos[3], target[3], up[3], fov_y, aspect, focal_distance, jitter[2] (14 floats).
// options: width, height, frame, maxBounces, doNEE, doMIS, doRR, HWSSDoMIS, doImportanceSampleEmitters,
// useDirectLights, useEmissiveLights, doMNEE, MNEEMaxOccluders, MNEEMaxIterations (14 ints); fopts:
// sourcingArea, maxOmega, MNEESolverThreshold.
// Renders pixels [first, first + count) of the row-major frame; color: 4 floats per rendered pixel;
// records (optional): (maxBounces+1) x 9 planes of `count` words, as the product's bounce records.
void* oracle_scene_create(const float* pos, const float* nrm, uint32_t ntri, const uint32_t* idx, const uint32_t* triMesh,
                          const uint32_t* meshMat, const void* mats, const void* prof, uint32_t nprof) {
    OScene* s = new OScene();
    s->pos = pos; s->nrm = nrm; s->idx = idx; s->triMesh = triMesh; s->meshMat = meshMat;
    s->mats = static_cast<const MatData*>(mats);
    s->prof = static_cast<const SpectralProfile*>(prof);
    s->nprof = nprof; s->ntri = ntri;
    s->mesh.pos = pos; s->mesh.idx = idx; s->mesh.ntri = ntri;
    s->bvh.build(s->mesh);
    buildLights(*s);
    return s;
}

void oracle_scene_destroy(void* h) { delete static_cast<OScene*>(h); }

// Analytic lights of a scene made by oracle_scene_create: `count` records of 48 floats in the
// plt_light layout (LightData).
int oracle_scene_set_lights(void* h, const float* lights, uint32_t count) {
    OScene& s = *static_cast<OScene*>(h);
    s.analytic.resize(count);
    static_assert(sizeof(OScene::Light) == 48 * sizeof(float), "light record");
    if (count) std::memcpy(s.analytic.data(), lights, count * sizeof(OScene::Light));
    return 0;
}

// Environment map of a scene made by oracle_scene_create: RGBA32F lat-long image, local-to-world
// rotation (row-major 3x3), radiance scale; the importance map is built here.
int oracle_scene_set_env(void* h, uint32_t w, uint32_t hgt, const float* rgba, const float* M, const float* scale) {
    OScene& s = *static_cast<OScene*>(h);
    s.envW = w; s.envH = hgt;
    s.envRGBA.assign(rgba, rgba + size_t(w) * hgt * 4);
    for (int i = 0; i < 9; ++i) s.envM[i] = M[i];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) s.envMI[3 * i + j] = M[3 * j + i];
    for (int c = 0; c < 3; ++c) s.envScale[c] = scale[c];
    buildEnvImportance(s);
    return 0;
}

// Renders with a scene made by oracle_scene_create (whose arrays the caller keeps alive).
int oracle_scene_render(void* h, const float* cam, const int32_t* iopt, const float* fopt, uint32_t first, uint32_t count,
                        float* color, uint32_t* records, float* aov, int threads) {
    OScene& s = *static_cast<OScene*>(h);
    // Camera.cpp:187-193 in float
    {
        const V3<F> P{cam[0], cam[1], cam[2]}, Tg{cam[3], cam[4], cam[5]}, Up{cam[6], cam[7], cam[8]};
        const float fovY = cam[9], aspect = cam[10], fd = cam[11];
        V3<F> W = normalize(Tg - P) * fd;
        V3<F> U = normalize(cross(W, Up));
        V3<F> V = normalize(cross(U, W));
        const float tn = std::tan(fovY * 0.5f);
        U = U * (fd * tn * aspect);
        V = V * (fd * tn);
        s.camPos = P; s.camU = U; s.camV = V; s.camW = W;
        s.jitterX = cam[12]; s.jitterY = cam[13];
        s.polarizer = cam[14]; s.polarizerTheta = cam[15];
        // glm::lookAt (right-handed) rows: s = normalize(f x up), u = s x f, -f (Camera.cpp:135)
        const V3<F> f = normalize(Tg - P);
        const V3<F> sr = normalize(cross(f, Up));
        s.viewRow[0] = sr;
        s.viewRow[1] = cross(sr, f);
        s.viewRow[2] = -f;
    }
    Options o;
    o.width = uint32_t(iopt[0]); o.height = uint32_t(iopt[1]); o.frame = uint32_t(iopt[2]); o.maxBounces = uint32_t(iopt[3]);
    o.doNEE = iopt[4]; o.doMIS = iopt[5]; o.doRR = iopt[6]; o.hwssMIS = iopt[7]; o.isEmitters = iopt[8];
    o.directLights = iopt[9]; o.emissiveLights = iopt[10];
    o.sourcingArea = fopt[0]; o.maxOmega = fopt[1];
    o.doMNEE = iopt[11];
    o.mneeMaxOccluders = std::min(std::max(iopt[12], 0), kMaxMSOccluders);
    o.mneeMaxIterations = iopt[13];
    o.mneeThreshold = fopt[2];
    o.uniformEmissive = iopt[14];
    o.analyticLights = iopt[15];
    o.envLight = iopt[16];
    o.sampleGenerator = iopt[17];
    o.perTileSG = iopt[18];
    o.tileSize = uint32_t(std::max(iopt[19], 1));
    if (threads > 0) omp_set_num_threads(threads);
    const size_t n = count;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < int64_t(count); ++i) {
        const uint32_t pid = first + uint32_t(i);
        const uint32_t px = pid % o.width, py = pid / o.width;
        std::vector<Bounce> rec(o.maxBounces + 2);
        float* a = aov ? aov + 8 * size_t(i) : nullptr;
        if (o.sampleGenerator == 0) {
            samplePass<TinyUniformSampleGenerator>(s, o, px, py, rec, a);
            solvePass<TinyUniformSampleGenerator>(s, o, px, py, rec, color + 4 * size_t(i));
        } else {
            samplePass<UniformSampleGenerator>(s, o, px, py, rec, a);
            solvePass<UniformSampleGenerator>(s, o, px, py, rec, color + 4 * size_t(i));
        }
        if (records) {
            for (uint32_t p = 0; p <= o.maxBounces; ++p) {
                const Bounce& b = rec[p];
                uint32_t* r = records + size_t(p) * 9 * n + size_t(i);
                r[0 * n] = b.tri;
                std::memcpy(&r[1 * n], &b.b1, 4);
                std::memcpy(&r[2 * n], &b.b2, 4);
                std::memcpy(&r[3 * n], &b.encWi.x, 4);
                std::memcpy(&r[4 * n], &b.encWi.y, 4);
                r[5 * n] = b.lobe;
                r[6 * n] = b.flags;
                std::memcpy(&r[7 * n], &b.pathContribution, 4);
                std::memcpy(&r[8 * n], &b.lastNonDeltaPdf, 4);
            }
        }
    }
    return 0;
}

// One-shot render: builds the scene, renders, frees it.
int oracle_render(const float* pos, const float* nrm, uint32_t ntri, const uint32_t* idx, const uint32_t* triMesh,
                  const uint32_t* meshMat, const void* mats, const void* prof, uint32_t nprof, const float* cam,
                  const int32_t* iopt, const float* fopt, uint32_t first, uint32_t count, float* color, uint32_t* records,
                  int threads) {
    void* h = oracle_scene_create(pos, nrm, ntri, idx, triMesh, meshMat, mats, prof, nprof);
    const int rc = oracle_scene_render(h, cam, iopt, fopt, first, count, color, records, nullptr, threads);
    oracle_scene_destroy(h);
    return rc;
}

// One manifold walk (manifoldNEESample) from `origin` towards the point `lightPos` on a light, with
// the straight segment, stopped 1e-4 short of the light, as the visibility ray. Writes the converged vertices, their sampled lobes, the
// Newton iteration count and the walk's pdf; returns the vertex count, or -1 when the walk fails.
int oracle_mnee_probe(void* h, const float* origin, const float* lightPos, float wavelength, int maxMSBounces, int maxOccluders,
                      int maxIterations, float threshold, uint32_t seed, float* hitPos, uint32_t* hitLobe, int* iterations,
                      float* pdf) {
    const OScene& s = *static_cast<const OScene*>(h);
    Options o;
    o.doMNEE = 1;
    o.mneeMaxOccluders = std::min(std::max(maxOccluders, 0), kMaxMSOccluders);
    o.mneeMaxIterations = maxIterations;
    o.mneeThreshold = threshold;
    const V3<F> org{origin[0], origin[1], origin[2]}, lp{lightPos[0], lightPos[1], lightPos[2]};
    const V3<F> tl = lp - org;
    const float dist = length(tl);
    UniformSampleGenerator sg(seed, 0u, 0u);
    MSData ms;
    ms.wavelength = wavelength;
    const bool ok = manifoldNEESample(s, o, org, tl / dist, dist * (1.f - 1e-4f), org, lp, tl / dist, false, maxMSBounces, sg, ms);
    *iterations = ms.iterations;
    *pdf = ms.pdf;
    for (int i = 0; i < ms.hitCount; ++i) {
        hitPos[3 * i] = ms.hits[i].posW.x; hitPos[3 * i + 1] = ms.hits[i].posW.y; hitPos[3 * i + 2] = ms.hits[i].posW.z;
        hitLobe[i] = ms.hits[i].lobe;
    }
    return ok ? ms.hitCount : -1;
}

// The walk's specular direction at barycentrics (b1, b2) of a triangle (positions P, vertex normals
// N, row-major 3x3) lit from `src`, with its derivatives along the surface parameters
// (materialPartialDerivatives of dsg.pdw(src)), in FP64: out = v, ddu, ddv.
void oracle_walk_derivatives(const void* mat, const void* prof, uint32_t nprof, const double* P, const double* N, double b1,
                             double b2, const double* src, uint32_t lobe, double wavelength, double* out) {
    using D = double;
    const Instance<D> in = setupMaterialInstance<D>(*static_cast<const MatData*>(mat), static_cast<const SpectralProfile*>(prof), nprof);
    V3<D> Pv[3], Nv[3];
    for (int k = 0; k < 3; ++k) {
        Pv[k] = V3<D>{P[3 * k], P[3 * k + 1], P[3 * k + 2]};
        Nv[k] = V3<D>{N[3 * k], N[3 * k + 1], N[3 * k + 2]};
    }
    const DSG<D> dsg = makeDSG<D>(Pv, Nv, b1, b2);
    const PDV<D> wi = dsg.pdw(V3<D>{src[0], src[1], src[2]});
    const KPDV<D> r = materialPartialDerivatives<D>(in, wi, dsg, lobe, wavelength);
    const V3<D> o[3] = {r.v, r.ddu, r.ddv};
    for (int k = 0; k < 3; ++k) { out[3 * k] = o[k].x; out[3 * k + 1] = o[k].y; out[3 * k + 2] = o[k].z; }
}

// The camera polarizer alone: a beam with direction r, polarisation frame tangent t and Stokes
// vector S (one wavelength) through measure's polarizer step for a camera with frame vectors U, W.
void oracle_camera_polarizer(const float* camU, const float* camW, float p, float theta, const float* r, const float* t,
                             const float* S, float* out) {
    OScene s{};
    s.camU = V3<F>{camU[0], camU[1], camU[2]};
    s.camW = V3<F>{camW[0], camW[1], camW[2]};
    s.polarizer = p;
    s.polarizerTheta = theta;
    Beam b;
    b.setZero();
    b.hwss_comps = 1;
    b.r = V3<F>{r[0], r[1], r[2]};
    b.t = V3<F>{t[0], t[1], t[2]};
    b.setSP(0, V4<F>{S[0], S[1], S[2], S[3]});
    applyCameraPolarizer(s, b);
    const V4<F> o = b.getSP(0);
    out[0] = o.x; out[1] = o.y; out[2] = o.z; out[3] = o.w;
}

// The manifold walk's FP64 acos and atan2 (walkAcos, walkAtan2) over arrays.
void oracle_walk_trig(const double* x, const double* y, uint32_t n, double* acosx, double* atan2yx) {
    for (uint32_t i = 0; i < n; ++i) {
        acosx[i] = walkAcos(x[i]);
        atan2yx[i] = walkAtan2(y[i], x[i]);
    }
}

// The power sampler's alias table, decoded: per entry the binary16 threshold (as float), the
// redirect ("above") and permuted ("below") light indices, and the per-light selection pdf.
// Returns the light count (entries beyond `cap` are not written).
int oracle_alias_table(const float* pos, uint32_t ntri, const uint32_t* idx, const uint32_t* triMesh, const uint32_t* meshMat,
                       const void* mats, float* thr, uint32_t* redirect, uint32_t* perm, float* pdf, uint32_t cap) {
    OScene s;
    s.pos = pos; s.idx = idx; s.triMesh = triMesh; s.meshMat = meshMat; s.ntri = ntri;
    s.mats = static_cast<const MatData*>(mats);
    buildLights(s);
    const uint32_t n = uint32_t(s.lightPdf.size());
    for (uint32_t i = 0; i < n && i < cap; ++i) {
        thr[i] = f16_bits_to_f32(uint16_t(s.aliasX[i] >> 16u));
        redirect[i] = ((s.aliasX[i] & 0xFFFFu) << 8u) | (s.aliasY[i] >> 24u);
        perm[i] = s.aliasY[i] & 0xFFFFFFu;
        pdf[i] = s.lightPdf[i];
    }
    return int(n);
}

int oracle_material_eval(const void* mat, const void* prof, uint32_t nprof, const double* wi, const double* wo, uint32_t lobe,
                         const double* lambda, const double* stokes, double area, int dbl, double* out, double* pdf, double* lpdf) {
    if (dbl) material_eval_t<double>(mat, prof, nprof, wi, wo, lobe, lambda, stokes, area, out, pdf, lpdf);
    else material_eval_t<float>(mat, prof, nprof, wi, wo, lobe, lambda, stokes, area, out, pdf, lpdf);
    return 0;
}

// One FP32 sample per query (as the path tracer draws them): query i has its own wi (3 floats,
// local frame) and wavelength and draws from generator (i, 0, frame); writes wo (3n), pdf, weight,
// lobe, valid.
int oracle_material_sample(const void* mat, const void* prof, uint32_t nprof, const float* wi3, const float* lambda, uint32_t n,
                           uint32_t frame, float* wo, float* pdf, float* weight, uint32_t* lobe, uint32_t* valid) {
    const Instance<float> in = setupMaterialInstance<float>(*static_cast<const MatData*>(mat), static_cast<const SpectralProfile*>(prof), nprof);
    SD<float> sd;
    sd.f = Frame<float>{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (uint32_t i = 0; i < n; ++i) {
        sd.V = {wi3[3 * i], wi3[3 * i + 1], wi3[3 * i + 2]};
        UniformSampleGenerator sg(i, 0, frame);
        BSDFSample r;
        valid[i] = materialSample(in, sd, sg, r, lambda[i]) ? 1u : 0u;
        wo[3 * i] = r.wo.x; wo[3 * i + 1] = r.wo.y; wo[3 * i + 2] = r.wo.z;
        pdf[i] = r.pdf; weight[i] = r.weight; lobe[i] = r.lobe;
    }
    return 0;
}

}  // extern "C"

// Scalar entry point for the closed-form test of the grating's Bessel function (test infrastructure).
extern "C" double oracle_bessel_j_f64(int nu, double x) { return BesselJ<double>(nu, x); }

// Lobe intensity of a 1-D grating of the given profile type at normal incidence, for the
// closed-form test of the lobe profiles (test infrastructure).
extern "C" double oracle_grating_lobe_f64(int type, int order, double height_um, double lambda_um) {
    DiffractionGrating<double> g{};
    g.gx = 1; g.gy = 0; g.ipx = 1; g.ipy = 0; g.q = height_um;
    g.lobeCount = 8; g.type = uint32_t(type); g.mult = 1;
    return g.lobeIntensity(order, 0, V3<double>{0, 0, 1}, lambda_um);
}
