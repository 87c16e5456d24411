// Shared device code of the two-pass PLT path tracer (included by the pathtracer*.hip translation
// units, each of which compiles some of its kernels).
// Two-pass PLT path tracer on MI355X (Source/RenderPasses/PLTPT/pltpt_sample.rt.slang and
// pltpt_solve.rt.slang): one lane per path.
//
// Sample pass: camera ray, then per bounce a closest-hit traversal (LDS stack), shading frame,
// Russian roulette, BSDF sampling in the hero wavelength, and a 36 B bounce record written as
// SoA planes. Solve pass: replays the records; per vertex it samples an emissive triangle
// (power alias table, spectral importance sampling), casts an any-hit shadow ray, sources a PLT
// beam and carries it forward through the stored vertices (forwardUnwind: eval + spectral MIS
// per vertex), measures it at the camera, and does the same for emissive hits. With manifold NEE
// on, a shadow ray blocked by specular transmitters becomes a Newton walk of those vertices onto
// the specular connection to the light (Manifolds.slangh), and a delta reflector at the current
// vertex takes part in that walk.
#pragma once
#include <algorithm>
#include "manifold.hpp"
#include "pathtracer.hpp"
#include "plt_internal.hpp"
#include "plt_materials.hpp"
#include "scene_device.hpp"

namespace plt {
namespace {

constexpr int kPtBlock = 64;
constexpr float kFltMax = 3.402823466e+38f;
constexpr float kFltMin = 1.175494351e-38f;

// ---------------------------------------------------------------------------------------
// Helpers: octahedral equal-area direction coding (MathHelpers.slang:199-240), Wachter-Binder
// ray offset (GeometryHelpers.slang:47-61).
// ---------------------------------------------------------------------------------------
PLT_DEV void oct_encode(V3 n, float& u, float& v) {
    const float r = sqrtf(1.f - fabsf(n.z));
    const float phi = atan2f(fabsf(n.y), fabsf(n.x));
    float py = r * phi * 0.636619772367581343076f;
    float px = r - py;
    if (n.z < 0.f) { const float t = px; px = 1.f - py; py = 1.f - t; }
    u = px * sgnf(n.x);
    v = py * sgnf(n.y);
}
PLT_DEV V3 oct_decode(float px, float py) {
    const float d = 1.f - (fabsf(px) + fabsf(py));
    const float r = 1.f - fabsf(d);
    const float phi = r > 0.f ? ((fabsf(py) - fabsf(px)) / r + 1.f) * kPi_4 : 0.f;
    const float f = r * sqrtf(2.f - r * r);
    return v3(f * sgnf(px) * cosf(phi), f * sgnf(py) * sinf(phi), sgnf(d) * (1.f - r * r));
}
PLT_DEV float offset_component(float p, float n) {
#pragma clang fp contract(off)
    const int io = int(n * (3.f * 256.f));
    const float ip = __int_as_float(__float_as_int(p) + (p < 0.f ? -io : io));
    return fabsf(p) < (1.f / 16.f) ? p + n * (3.f / 65536.f) : ip;
}
PLT_DEV V3 ray_origin(V3 p, V3 n) { return v3(offset_component(p.x, n.x), offset_component(p.y, n.y), offset_component(p.z, n.z)); }

// ---------------------------------------------------------------------------------------
// Hit -> shading data (Scene.slang:530-567, MaterialFactory.slang:79-130): barycentric
// interpolation, invented tangent frame (perp_stark; the workload meshes carry no tangents).
// ---------------------------------------------------------------------------------------
struct Surf {
    V3 pos, faceN;
    Frame f;
    uint32_t material;
    bool front;
};
PLT_DEV Surf surface_at(const PtScene& sc, uint32_t tri, float b1, float b2, V3 viewDir) {
    // Uncontracted: the shading point must round as a strict-IEEE evaluation does, because light
    // samples taken from an emitter's own plane decide on the sign of a near-zero cosine.
#pragma clang fp contract(off)
    const uint32_t i0 = sc.indices[3 * tri], i1 = sc.indices[3 * tri + 1], i2 = sc.indices[3 * tri + 2];
    const float w0 = 1.f - b1 - b2;
    const V3 p0 = v3(sc.positions[3 * i0], sc.positions[3 * i0 + 1], sc.positions[3 * i0 + 2]);
    const V3 p1 = v3(sc.positions[3 * i1], sc.positions[3 * i1 + 1], sc.positions[3 * i1 + 2]);
    const V3 p2 = v3(sc.positions[3 * i2], sc.positions[3 * i2 + 1], sc.positions[3 * i2 + 2]);
    const V3 n0 = v3(sc.normals[3 * i0], sc.normals[3 * i0 + 1], sc.normals[3 * i0 + 2]);
    const V3 n1 = v3(sc.normals[3 * i1], sc.normals[3 * i1 + 1], sc.normals[3 * i1 + 2]);
    const V3 n2 = v3(sc.normals[3 * i2], sc.normals[3 * i2 + 1], sc.normals[3 * i2 + 2]);
    Surf s;
    s.pos = v3(p0.x * w0 + p1.x * b1 + p2.x * b2, p0.y * w0 + p1.y * b1 + p2.y * b2, p0.z * w0 + p1.z * b1 + p2.z * b2);
    const V3 N = normalize(n0 * w0 + n1 * b1 + n2 * b2);
    const V3 e1 = p1 - p0, e2 = p2 - p0;
    const V3 c = v3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
    const float cl = sqrtf(c.x * c.x + c.y * c.y + c.z * c.z);
    s.faceN = v3(c.x / cl, c.y / cl, c.z / cl);
    s.f.N = N;
    s.f.T = perp_stark(N);
    s.f.B = cross(N, s.f.T);
    s.front = dot(viewDir, s.faceN) >= 0.f;
    s.material = sc.mesh_material[sc.tri_mesh[tri]];
    return s;
}

// ---------------------------------------------------------------------------------------
// Spectra: CIE fit (Wyman 2013), D65, Rec.709 (SpectrumUtils.slang:43-104).
// ---------------------------------------------------------------------------------------
__constant__ const float kD65[69] = {49.9755f, 52.3118f, 54.6482f, 68.7015f, 82.7549f, 87.1204f, 91.486f, 92.4589f, 93.4318f,
    90.057f, 86.6823f, 95.7736f, 104.865f, 110.936f, 117.008f, 117.41f, 117.812f, 116.336f, 114.861f, 115.392f, 115.923f,
    112.367f, 108.811f, 109.082f, 109.354f, 108.578f, 107.802f, 106.296f, 104.79f, 106.239f, 107.689f, 106.047f, 104.405f,
    104.225f, 104.046f, 102.023f, 100.f, 98.1671f, 96.3342f, 96.0611f, 95.788f, 92.2368f, 88.6856f, 89.3459f, 90.0062f,
    89.8026f, 89.5991f, 88.6489f, 87.6987f, 85.4936f, 83.2886f, 83.4939f, 83.6992f, 81.863f, 80.0268f, 80.1207f, 80.2146f,
    81.2462f, 82.2778f, 80.281f, 78.2842f, 74.0027f, 69.7213f, 70.6652f, 71.6091f, 72.979f, 74.349f, 67.9765f, 61.604f};
PLT_DEV float gl(float x, float l, float a, float b) { const float t = (x - l) * (x < l ? a : b); return expf(-0.5f * t * t); }
PLT_DEV void spectrum_to_rgb_add(float val, float lam, float& r, float& g, float& b) {
    const float X = 0.362f * gl(lam, 442.0f, 0.0624f, 0.0374f) + 1.056f * gl(lam, 599.8f, 0.0264f, 0.0323f) -
                    0.065f * gl(lam, 501.1f, 0.0490f, 0.0382f);
    const float Y = 0.821f * gl(lam, 568.8f, 0.0213f, 0.0247f) + 0.286f * gl(lam, 530.9f, 0.0613f, 0.0322f);
    const float Z = 1.217f * gl(lam, 437.0f, 0.0845f, 0.0278f) + 0.681f * gl(lam, 459.0f, 0.0385f, 0.0725f);
    const float rr = (lam - kMinWavelength) * 0.2f;
    // The table covers 380-720 nm; importance-sampled emitter wavelengths can lie beyond it (where the
    // white basis, and so every spectrum here, is exactly zero), so the indices are clamped to the table.
    const int t = min(68, max(0, int(ceilf(rr) + .5f))), bb = min(68, max(0, int(floorf(rr) + .5f)));
    const float w = rr - float(bb);
    const float d65 = kD65[t] * w + kD65[bb] * (1.f - w);
    const float s = d65 * val;
    const float x = X * s, y = Y * s, z = Z * s;
    r += 3.240969941904522f * x + -1.537383177570094f * y + -0.4986107602930032f * z;
    g += -0.9692436362808803f * x + 1.875967501507721f * y + 0.04155505740717569f * z;
    b += 0.05563007969699373f * x + -0.2039769588889765f * y + 1.056971514242878f * z;
}

// Scene::sampleSpectrum (Scene.slang:296-327).
template <int N>
PLT_DEV void sample_spectrum(const plt_spectral_profile* sp, float hero, Rng& sg, bool importance, float* Le, float* lam) {
    if (!sp) {
#pragma unroll
        for (int j = 0; j < N; ++j) { Le[j] = 0.f; lam[j] = hero; }
        return;
    }
    if (importance) {
        lam[0] = hero;
        Le[0] = eval_spectrum(sp, hero);
        const float bins = float(sp->bins);
        const float x0 = bins * sg.next1D(), x1 = bins * sg.next1D(), x2 = bins * sg.next1D();
        const float xs[3] = {x0, x1, x2};
#pragma unroll
        for (int j = 1; j < N; ++j) {
            const float xv = xs[j - 1];
            const uint32_t idx = uint32_t(xv);
            const float a = idx < sp->bins ? sp->icdf[idx] : 1.f;
            const float b = idx + 1 < sp->bins ? sp->icdf[idx + 1] : 1.f;
            const float fr = xv - floorf(xv);
            const float cdf = a + fr * (b - a);
            lam[j] = sp->min_wavelength + cdf * (sp->max_wavelength - sp->min_wavelength);
            Le[j] = sp->intensity / bins;
        }
    } else {
        const float span = kMaxWavelength - kMinWavelength;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            float h = hero + float(j) * (span / float(N));
            h -= h > kMaxWavelength ? span : 0.f;
            lam[j] = h;
            Le[j] = eval_spectrum(sp, h);
        }
    }
}

// pltSourceBeamArea (PLT.slang:275-290).
template <int N>
PLT_DEV Beam<N> source_beam_area(V3 pos, V3 dir, float A, float dist, const float* Le, const float* lam, float maxOmega, int comps) {
    Beam<N> b;
    const float r2 = (dist * 1e+3f) * (dist * 1e+3f);
    A = fminf(A, maxOmega * r2);
#pragma unroll
    for (int j = 0; j < N; ++j) {
        b.S[j][0] = Le[j]; b.S[j][1] = b.S[j][2] = b.S[j][3] = 0.f;
        b.lambda[j] = lam[j];
        b.spdf[j] = 1.f;
    }
    b.comps = comps;
    b.D00 = A; b.D01 = 0.f; b.D11 = A;
    b.r_m = 0.f;
    b.pos = pos;
    b.r = dir;
    b.t = perp_stark(dir);
    b.distant = false;
    return b;
}

// pltSourceBeamDistant (PLT.slang:253-269): a beam from a source at infinity subtending solidAngle.
template <int N>
PLT_DEV Beam<N> source_beam_distant(V3 dir, float solidAngle, const float* Le, const float* lam, float maxOmega, int comps) {
    Beam<N> b;
    solidAngle = fminf(solidAngle, maxOmega);
#pragma unroll
    for (int j = 0; j < N; ++j) {
        b.S[j][0] = Le[j]; b.S[j][1] = b.S[j][2] = b.S[j][3] = 0.f;
        b.lambda[j] = lam[j];
        b.spdf[j] = 1.f;
    }
    b.comps = comps;
    b.D00 = solidAngle; b.D01 = 0.f; b.D11 = solidAngle;
    b.r_m = 0.f;
    b.pos = v3(0.f, 0.f, 0.f);
    b.r = dir;
    b.t = perp_stark(dir);
    b.distant = true;
    return b;
}

// ---- environment map (Scene/Lights/EnvMap.slang, Rendering/Lights/EnvMapSampler.slang) ----
// Radiance arriving from world direction `dir`: world_to_latlong_map of the local direction, a
// bilinear lookup wrapping in u and clamped in v, times intensity x tint.
PLT_DEV void env_eval(const PtScene& sc, V3 dir, float rgb[3]) {
    const float* M = sc.env_mi;
    const V3 d = v3(M[0] * dir.x + M[1] * dir.y + M[2] * dir.z, M[3] * dir.x + M[4] * dir.y + M[5] * dir.z,
                    M[6] * dir.x + M[7] * dir.y + M[8] * dir.z);
    const float l = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    const V3 p = v3(d.x / l, d.y / l, d.z / l);
    const float u = atan2f(p.x, -p.z) * 0.159154943091895335769f + 0.5f;
    const float v = acosf(fminf(fmaxf(p.y, -1.f), 1.f)) * 0.318309886183790671538f;
    const float s = u * float(sc.env_w) - 0.5f, t = v * float(sc.env_h) - 0.5f;
    const float fs = floorf(s), ft = floorf(t);
    const float a = s - fs, b = t - ft;
    const int W = int(sc.env_w), H = int(sc.env_h);
    int x0 = int(fs) % W;
    if (x0 < 0) x0 += W;
    const int x1 = (x0 + 1) % W;
    const int y0 = min(max(int(ft), 0), H - 1), y1 = min(max(int(ft) + 1, 0), H - 1);
    const float4 c00 = sc.env_rgba[size_t(y0) * W + x0], c10 = sc.env_rgba[size_t(y0) * W + x1];
    const float4 c01 = sc.env_rgba[size_t(y1) * W + x0], c11 = sc.env_rgba[size_t(y1) * W + x1];
    const float t0 = c00.x + (c10.x - c00.x) * a, b0 = c01.x + (c11.x - c01.x) * a;
    const float t1 = c00.y + (c10.y - c00.y) * a, b1 = c01.y + (c11.y - c01.y) * a;
    const float t2 = c00.z + (c10.z - c00.z) * a, b2 = c01.z + (c11.z - c01.z) * a;
    rgb[0] = (t0 + (b0 - t0) * b) * sc.env_scale[0];
    rgb[1] = (t1 + (b1 - t1) * b) * sc.env_scale[1];
    rgb[2] = (t2 + (b2 - t2) * b) * sc.env_scale[2];
}
// EnvMapSampler.evalPdf (:150-163): the point-sampled importance of the direction's octahedral
// texel over the map's mean, per 4 pi steradians.
PLT_DEV float env_pdf(const PtScene& sc, V3 dir) {
    const float* M = sc.env_mi;
    const V3 d = v3(M[0] * dir.x + M[1] * dir.y + M[2] * dir.z, M[3] * dir.x + M[4] * dir.y + M[5] * dir.z,
                    M[6] * dir.x + M[7] * dir.y + M[8] * dir.z);
    float ou, ov;
    oct_encode(d, ou, ov);
    const float u = ou * 0.5f + 0.5f, v = ov * 0.5f + 0.5f;
    const int dim = int(kEnvImportanceDim);
    const int x = min(max(int(floorf(u * float(dim))), 0), dim - 1), y = min(max(int(floorf(v * float(dim))), 0), dim - 1);
    const float avg = sc.env_imp[sc.env_mip[sc.env_base_mip]];
    return sc.env_imp[size_t(y) * dim + x] / avg * (1.f / 12.5663706143591729539f);
}
// EnvMapSampler.sample (:56-148): descend the importance mip chain choosing a child of each 2x2
// block in proportion to its weight, warping the sample; then the octahedral direction of the
// sample's position and its pdf.
PLT_DEV void env_sample(const PtScene& sc, float px, float py, V3& dir, float& pdf) {
    uint32_t x = 0, y = 0;
    for (int mip = int(sc.env_base_mip) - 1; mip >= 0; --mip) {
        x *= 2;
        y *= 2;
        const uint32_t dim = kEnvImportanceDim >> mip;
        const float* L = sc.env_imp + sc.env_mip[mip];
        const float w0 = L[size_t(y) * dim + x], w1 = L[size_t(y) * dim + x + 1];
        const float w2 = L[size_t(y + 1) * dim + x], w3 = L[size_t(y + 1) * dim + x + 1];
        const float q0 = w0 + w2, q1 = w1 + w3;
        const float dd = q0 / (q0 + q1);
        uint32_t ox, oy;
        if (px < dd) {
            ox = 0;
            px = px / dd;
        } else {
            ox = 1;
            px = (px - dd) / (1.f - dd);
        }
        const float e = ox == 0 ? (w0 / q0) : (w1 / q1);
        if (py < e) {
            oy = 0;
            py = py / e;
        } else {
            oy = 1;
            py = (py - e) / (1.f - e);
        }
        x += ox;
        y += oy;
    }
    const float inv = 1.f / float(kEnvImportanceDim);
    const float u = (float(x) + px) * inv, v = (float(y) + py) * inv;
    const V3 dl = oct_decode(u * 2.f - 1.f, v * 2.f - 1.f);
    const float avg = sc.env_imp[sc.env_mip[sc.env_base_mip]];
    pdf = sc.env_imp[size_t(y) * kEnvImportanceDim + x] / avg * 0.0795774715459476678844f;
    const float* M = sc.env_m;
    dir = v3(M[0] * dl.x + M[1] * dl.y + M[2] * dl.z, M[3] * dl.x + M[4] * dl.y + M[5] * dl.z,
             M[6] * dl.x + M[7] * dl.y + M[8] * dl.z);
}
// SpectrumSampler.hwssWavelengths: N equally spaced wavelengths from the hero, wrapping.
template <int N>
PLT_DEV void hwss_lambdas(float hero, float* lam) {
    const float span = kMaxWavelength - kMinWavelength;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        float h = hero + float(j) * (span / float(N));
        h -= h > kMaxWavelength ? span : 0.f;
        lam[j] = h;
    }
}

// An analytic light sample (generateAnalyticLightSample, LightSampling.slangh:122-163, over
// sampleLight, Rendering/Lights/LightHelpers.slang:70-290): uniform choice of a light, its sample,
// radiance scaled by the light count and pdf divided by it. Point and directional lights sample
// with pdf 0 (a Dirac), which the caller's acceptance test then rejects, as the reference does.
// Returns false where the reference's sampler does (an area light seen from behind).
struct AnalyticSample {
    V3 dir, origin, beam_pos;
    float dist, pdf, beam_size;  // beam_size: area (mm^2) or, for distant beams, solid angle
    bool distant;
};
// Kept out of line: inlined, its registers would count against the whole solve kernel, which
// scenes without analytic lights never pay for otherwise.
template <int N>
__device__ __attribute__((noinline)) bool sample_analytic(const PtScene& sc, const Surf& s, float lambda, bool importance,
                                                           Rng& sg, AnalyticSample& a, float* Li, float* lam) {
#pragma clang fp contract(off)
    const uint32_t cnt = sc.n_analytic;
    const uint32_t idx = min(uint32_t(sg.next1D() * float(cnt)), cnt - 1);
    const plt_light& L = sc.analytic[idx];
    const plt_spectral_profile* sp = L.intensity_profile < sc.n_profiles ? sc.profiles + L.intensity_profile : nullptr;
    const float* M = L.trans;
    const float* MI = L.trans_it;
    constexpr float k2Pi = 6.28318530717958647692f;
    float lpdf = 0.f;
    a.distant = false;
    if (L.type == PLT_LIGHT_RECT || L.type == PLT_LIGHT_DISC || L.type == PLT_LIGHT_SPHERE) {
        const float u0 = sg.next1D(), u1 = sg.next1D();
        V3 pl, nl = v3(0.f, 0.f, 1.f);
        if (L.type == PLT_LIGHT_RECT) {
            pl = v3(u0 * 2.f - 1.f, u1 * 2.f - 1.f, 0.f);
        } else if (L.type == PLT_LIGHT_DISC) {  // sample_disk (MathHelpers.slang:255-263)
            const float r = sqrtf(u0), phi = k2Pi * u1;
            pl = v3(r * cosf(phi), r * sinf(phi), 0.f);
        } else {  // sample_sphere (MathHelpers.slang:282-288)
            const float phi = k2Pi * u1, ct = 1.f - 2.f * u0;
            const float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
            pl = v3(st * cosf(phi), st * sinf(phi), ct);
            nl = pl;
        }
        const V3 pw = v3(M[0] * pl.x + M[1] * pl.y + M[2] * pl.z + M[3], M[4] * pl.x + M[5] * pl.y + M[6] * pl.z + M[7],
                         M[8] * pl.x + M[9] * pl.y + M[10] * pl.z + M[11]);
        const V3 nw = normalize(v3(MI[0] * nl.x + MI[1] * nl.y + MI[2] * nl.z, MI[4] * nl.x + MI[5] * nl.y + MI[6] * nl.z,
                                   MI[8] * nl.x + MI[9] * nl.y + MI[10] * nl.z));
        // finalizeAreaLightSample (LightHelpers.slang:70-93)
        const V3 toL = pw - s.pos;
        const float d2 = fmaxf(toL.x * toL.x + toL.y * toL.y + toL.z * toL.z, 1e-9f);
        a.dist = sqrtf(d2);
        a.dir = v3(toL.x / a.dist, toL.y / a.dist, toL.z / a.dist);
        const float ct = nw.x * -a.dir.x + nw.y * -a.dir.y + nw.z * -a.dir.z;
        if (ct <= 0.f) return false;
        sample_spectrum<N>(sp, lambda, sg, importance, Li, lam);
        const float f = L.surface_area * ct / d2;
#pragma unroll
        for (int j = 0; j < N; ++j) Li[j] *= f;
        lpdf = d2 / (ct * L.surface_area);
    } else if (L.type == PLT_LIGHT_DISTANT) {  // sampleDistantLight (:170-198), sample_cone (MathHelpers.slang:270-276)
        const float u0 = sg.next1D(), u1 = sg.next1D();
        const float c = L.cos_subtended_angle;
        const float z = u0 * (1.f - c) + c;
        const float r = sqrtf(1.f - z * z), phi = k2Pi * u1;
        const V3 d = v3(r * cosf(phi), r * sinf(phi), z);
        a.dir = normalize(v3(M[0] * d.x + M[1] * d.y + M[2] * d.z, M[4] * d.x + M[5] * d.y + M[6] * d.z,
                             M[8] * d.x + M[9] * d.y + M[10] * d.z));
        a.dist = kFltMax;
        sample_spectrum<N>(sp, lambda, sg, importance, Li, lam);
        lpdf = 1.f / (k2Pi * (1.f - c));
        a.distant = true;
    } else if (L.type == PLT_LIGHT_DIRECTIONAL) {  // sampleDirectionalLight (:200-225)
        a.dir = v3(-L.dir[0], -L.dir[1], -L.dir[2]);
        a.dist = kFltMax;
        sample_spectrum<N>(sp, lambda, sg, importance, Li, lam);
        lpdf = 0.f;
        a.distant = true;
    } else {  // samplePointLight (:227-269): a Dirac, rejected by the caller (pdf 0)
        const V3 toL = v3(L.pos[0], L.pos[1], L.pos[2]) - s.pos;
        const float d2 = fmaxf(toL.x * toL.x + toL.y * toL.y + toL.z * toL.z, 1e-9f);
        a.dist = sqrtf(d2);
        a.dir = v3(toL.x / a.dist, toL.y / a.dist, toL.z / a.dist);
        sample_spectrum<N>(sp, lambda, sg, importance, Li, lam);
        lpdf = 0.f;
    }
#pragma unroll
    for (int j = 0; j < N; ++j) Li[j] *= float(cnt);
    a.pdf = lpdf / float(cnt);
    // computeNewRayOrigin(sd, dir) (ShadowRays.slangh:50-52)
    const float side = s.faceN.x * a.dir.x + s.faceN.y * a.dir.y + s.faceN.z * a.dir.z;
    a.origin = ray_origin(s.pos, side >= 0.f ? s.faceN : -s.faceN);
    a.beam_pos = v3(L.pos[0], L.pos[1], L.pos[2]);
    a.beam_size = a.distant ? 4.f * 3.14159265358979323846f * (1.f - L.cos_subtended_angle * L.cos_subtended_angle)
                            : L.surface_area;
    return true;
}

struct Rec {
    uint32_t tri;
    float b1, b2;
    V3 wi;
    uint32_t lobe, flags;
    float contrib, lastPdf;
};
// Records are written once and read once, both non-temporally (see ev_get in pathtracer_split.hpp).
// Non-temporal stores, per 1080p frame of the 5.24M-triangle scene (means of three interleaved
// runs): sample kernel 2.484 -> 2.473 ms, light-sampling kernel 2.264 -> 2.245 ms.
PLT_DEV Rec read_rec(const uint32_t* rec, size_t n, size_t i, int p) {
    const uint32_t* b = rec + size_t(p) * kRecWords * n + i;
    Rec r;
    r.tri = __builtin_nontemporal_load(b + kRecTri * n);
    r.b1 = __uint_as_float(__builtin_nontemporal_load(b + kRecB1 * n));
    r.b2 = __uint_as_float(__builtin_nontemporal_load(b + kRecB2 * n));
    r.wi = oct_decode(__uint_as_float(__builtin_nontemporal_load(b + kRecWiX * n)), __uint_as_float(__builtin_nontemporal_load(b + kRecWiY * n)));
    r.lobe = __builtin_nontemporal_load(b + kRecLobe * n);
    r.flags = __builtin_nontemporal_load(b + kRecFlags * n);
    r.contrib = __uint_as_float(__builtin_nontemporal_load(b + kRecContrib * n));
    r.lastPdf = __uint_as_float(__builtin_nontemporal_load(b + kRecLastPdf * n));
    return r;
}
PLT_DEV void write_rec(uint32_t* rec, size_t n, size_t i, int p, uint32_t tri, float b1, float b2, V3 wi, uint32_t lobe,
                       uint32_t flags, float contrib, float lastPdf) {
    uint32_t* b = rec + size_t(p) * kRecWords * n + i;
    float ex, ey;
    oct_encode(wi, ex, ey);
    __builtin_nontemporal_store(tri, b + kRecTri * n);
    __builtin_nontemporal_store(__float_as_uint(b1), b + kRecB1 * n);
    __builtin_nontemporal_store(__float_as_uint(b2), b + kRecB2 * n);
    __builtin_nontemporal_store(__float_as_uint(ex), b + kRecWiX * n);
    __builtin_nontemporal_store(__float_as_uint(ey), b + kRecWiY * n);
    __builtin_nontemporal_store(lobe, b + kRecLobe * n);
    __builtin_nontemporal_store(flags, b + kRecFlags * n);
    __builtin_nontemporal_store(__float_as_uint(contrib), b + kRecContrib * n);
    __builtin_nontemporal_store(__float_as_uint(lastPdf), b + kRecLastPdf * n);
}

PLT_DEV Mat material_at(const PtScene& sc, uint32_t m) { return setup_material(sc.materials[m], sc.profiles, sc.n_profiles); }

// Camera.computeRayPinhole (Camera.slang:46-78). Uncontracted, correctly rounded arithmetic so
// primary rays (and hence primary hits) are bit-identical to a strict-IEEE host computation.
PLT_DEV V3 camera_dir(const PtScene& sc, const PtFrame& fr, uint32_t px, uint32_t py) {
#pragma clang fp contract(off)
    const float sx = (float(px) + 0.5f) / float(fr.width) + -sc.jitter[0];
    const float sy = (float(py) + 0.5f) / float(fr.height) + sc.jitter[1];
    const float nx = 2.f * sx + -1.f, ny = -2.f * sy + 1.f;
    const float x = nx * sc.cam_u[0] + ny * sc.cam_v[0] + sc.cam_w[0];
    const float y = nx * sc.cam_u[1] + ny * sc.cam_v[1] + sc.cam_w[1];
    const float z = nx * sc.cam_u[2] + ny * sc.cam_v[2] + sc.cam_w[2];
    const float l = sqrtf(x * x + y * y + z * z);
    return v3(x / l, y / l, z / l);
}

// ---------------------------------------------------------------------------------------
// Sample pass (pltpt_sample.rt.slang:44-185).
// ---------------------------------------------------------------------------------------
// Camera-path state between vertices (the reference's Path payload, Path.slang:146-196).
struct PathState {
    V3 org, dir;
    float lambda, weight, pdf, thp, lastPdf;
    uint32_t flags, len;
};

// Path start (pltpt_sample.rt.slang:146-185): the hero wavelength is the first draw; albedo and
// normal outputs start at zero; the camera ray.
PLT_DEV PathState start_path(const PtScene& sc, const PtFrame& fr, uint32_t px, uint32_t py, uint32_t pid, Rng& sg) {
    PathState ps;
    ps.lambda = kMinWavelength + sg.next1D() * (kMaxWavelength - kMinWavelength);
    fr.albedo[pid] = make_float4(0.f, 0.f, 0.f, 0.f);
    fr.normal[pid] = make_float4(0.f, 0.f, 0.f, 0.f);
    ps.dir = camera_dir(sc, fr, px, py);
    ps.org = v3(sc.cam_pos[0], sc.cam_pos[1], sc.cam_pos[2]);
    ps.weight = ps.pdf = ps.thp = 1.f;
    ps.lastPdf = 0.f;
    ps.flags = kFlagOnlyDelta;
    ps.len = 0;
    return ps;
}

// One camera-path vertex given the closest hit `h` of the path's current ray
// (scatterMiss / scatterTriangleMeshClosestHit, pltpt_sample.rt.slang:55-141): Russian roulette,
// BSDF sampling in the hero wavelength, the bounce record, the next ray. Returns true when the path
// continues (ps.org / ps.dir then hold the next ray).
PLT_DEV bool sample_vertex(const PtScene& sc, const PtFrame& fr, const Hit& h, uint32_t li, uint32_t pid, PathState& ps, Rng& sg) {
    const size_t n = fr.n_paths;
    const int len = int(ps.len);
    const V3 wi = -ps.dir;
    if (h.tri == kNoHit) {  // scatterMiss
        ps.flags |= kFlagMiss;
        write_rec(fr.records, n, li, len, kNoHit, 0.f, 0.f, wi, 0u, ps.flags, 1.f / (ps.pdf * ps.thp), ps.lastPdf);
        return false;
    }
    const Surf s = surface_at(sc, h.tri, h.b1, h.b2, wi);
    const Mat m = material_at(sc, s.material);
    const uint32_t props = mat_props(m);
    // Russian roulette: RR(1 - weight, 2, u) (pltpt.slangh:170-176)
    float rcp_thp = 1.f;
    if (fr.options & kOptRR) {
        const float u = sg.next1D();
        const float pr = powf(1.f - ps.weight, 2.f);
        rcp_thp = u < pr ? 0.f : 1.f - pr;
    }
    bool end = len == int(fr.max_bounces) || (rcp_thp == 0.f && len > 0);
    BsdfSample bs{v3(0.f, 0.f, 0.f), 0.f, 0.f, 0u};
    if (!end) {
        const bool valid = mat_sample(m, s.f, wi, sg, ps.lambda, bs);
        end = !valid || bs.pdf == 0.f;
        end |= ps.pdf * bs.pdf * ps.thp * rcp_thp < 1e-6f;
    }
    const bool isDelta = (bs.lobe & kLobeDelta) != 0;
    const bool isDisp = (bs.lobe & kLobeDispersiveDelta) != 0;
    const bool isRefl = (bs.lobe & kLobeReflection) != 0;
    const bool isMS = (isRefl && (props & kPropMWReflective)) || (!isRefl && (props & kPropMWTransmissive));
    if (!isDelta && (ps.flags & kFlagOnlyDelta)) {
        float alb[3];
        mat_albedo(m, alb);
        fr.albedo[pid] = make_float4(alb[0], alb[1], alb[2], 1.f);
        // mul(float3x3(viewMat), sd.N): the view-space normal (pltpt_sample.rt.slang:111)
        const float* R = sc.view_rot;
        fr.normal[pid] = make_float4(R[0] * s.f.N.x + R[1] * s.f.N.y + R[2] * s.f.N.z, R[3] * s.f.N.x + R[4] * s.f.N.y + R[5] * s.f.N.z,
                                     R[6] * s.f.N.x + R[7] * s.f.N.y + R[8] * s.f.N.z, 1.f);
    }
    if (end) {
        ps.flags |= kFlagTerminated;
        write_rec(fr.records, n, li, len, h.tri, h.b1, h.b2, wi, 0u, ps.flags, 1.f / (ps.pdf * ps.thp), ps.lastPdf);
        return false;
    }
    ps.weight *= bs.weight;
    ps.thp *= rcp_thp;
    ps.pdf *= bs.pdf;
    if (isDisp) ps.flags |= kFlagDispersive;
    write_rec(fr.records, n, li, len, h.tri, h.b1, h.b2, wi, bs.lobe, ps.flags, 1.f / (ps.pdf * ps.thp), ps.lastPdf);
    ++ps.len;
    ps.lastPdf = !isDelta ? bs.pdf : ps.lastPdf;
    const uint32_t msc = isMS ? (ps.flags >> 16) + 1u : 0u;
    ps.flags = (ps.flags & 0xffffu) | (msc << 16);
    ps.flags = isDelta ? (ps.flags | kFlagLastDelta) : (ps.flags & ~kFlagLastDelta);
    if (!isDelta) ps.flags &= ~kFlagOnlyDelta;
    // ShadingData.computeNewRayOrigin(viewside = isReflection) (ShadingData.slang:77-80)
    ps.org = ray_origin(s.pos, (s.front == isRefl) ? s.faceN : -s.faceN);
    ps.dir = bs.wo;
    return true;
}

// ---------------------------------------------------------------------------------------
// Solve pass (pltpt_solve.rt.slang:37-476).
// ---------------------------------------------------------------------------------------
struct Solve {
    float r, g, b;
    V3 origin, normal;
};

template <int N>
PLT_DEV void measure(const PtScene& sc, const PtFrame& fr, Solve& sv, int bounce, float contrib, Beam<N>& beam) {
    if ((fr.options & kOptDirectLights) == 0 && bounce <= 1) return;
    beam.propagate(v3(sc.cam_pos[0], sc.cam_pos[1], sc.cam_pos[2]));
    // camera polarizer (pltpt_solve.rt.slang:111-126): the beam in the camera frame (t = U, r = -W),
    // then (1 - p) I + p MuellerPolarizer(theta) (Mueller.slang:60-74) on each HWSS component
    if (sc.polarizer > 0.f) {
        beam.rotate_frame(normalize(v3(sc.cam_u[0], sc.cam_u[1], sc.cam_u[2])), -normalize(v3(sc.cam_w[0], sc.cam_w[1], sc.cam_w[2])));
        const float p = sc.polarizer, q = 1.f - sc.polarizer;
        const float c = cosf(sc.polarizer_theta), sn = sinf(sc.polarizer_theta);
        // M = q I + p (P / 2), P = [[1, c, s, 0], [c, c^2, cs, 0], [s, cs, s^2, 0], [0, 0, 0, 0]]
        const float m01 = p * (.5f * c), m02 = p * (.5f * sn), m12 = p * (.5f * (c * sn));
        const float m00 = q + p * .5f, m11 = q + p * (.5f * (c * c)), m22 = q + p * (.5f * (sn * sn));
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (j < beam.comps) {
                const float s0 = beam.S[j][0], s1 = beam.S[j][1], s2 = beam.S[j][2];
                beam.S[j][0] = m00 * s0 + m01 * s1 + m02 * s2;
                beam.S[j][1] = m01 * s0 + m11 * s1 + m12 * s2;
                beam.S[j][2] = m02 * s0 + m12 * s1 + m22 * s2;
                beam.S[j][3] = q * beam.S[j][3];
            }
        }
    }
    if (fr.options & kOptHwssMis) {
        float sum = 0.f;
#pragma unroll
        for (int j = 0; j < N; ++j) sum += beam.spdf[j];
        contrib *= sum > kFltEps ? beam.spdf[0] / sum : 0.f;
    } else {
        contrib *= 1.f / float(beam.comps);
    }
    float r = 0.f, g = 0.f, b = 0.f;
#pragma unroll
    for (int j = 0; j < N; ++j)
        if (j < beam.comps) spectrum_to_rgb_add(beam.S[j][0], beam.lambda[j], r, g, b);
    const float inv = 1.f / 10.5670762f;
    r = fmaxf(0.f, r * inv * contrib);
    g = fmaxf(0.f, g * inv * contrib);
    b = fmaxf(0.f, b * inv * contrib);
    if (isfinite(r) && isfinite(g) && isfinite(b)) { sv.r += r; sv.g += g; sv.b += b; }
}

// One forward step through a stored vertex (forwardUnwindStep, pltpt_solve.rt.slang:48-66).
// posW replaces the interpolated hit position (manifold walk vertices).
template <int N>
PLT_DEV void unwind_step(const PtScene& sc, const PtFrame& fr, uint32_t tri, float b1, float b2, V3 wo, uint32_t lobe, Beam<N>& beam,
                         const V3* posW = nullptr) {
    const V3 V = -beam.r;
    Surf s = surface_at(sc, tri, b1, b2, V);
    if (posW) s.pos = *posW;
    beam.propagate(s.pos);
    const Mat m = material_at(sc, s.material);
    mat_eval<N>(m, s.f, V, wo, lobe, beam);
    if (fr.options & kOptHwssMis) {
        float p[N];
        mat_lambda_pdf<N>(m, s.f, V, wo, lobe, beam, p);
#pragma unroll
        for (int j = 0; j < N; ++j) beam.spdf[j] *= p[j];
    }
}
template <int N>
PLT_DEV void forward_unwind(const PtScene& sc, const PtFrame& fr, size_t n, size_t i, int bounce, Beam<N>& beam) {
    for (int p = bounce - 1; p >= 0; --p) {
        const Rec r = read_rec(fr.records, n, i, p);
        unwind_step<N>(sc, fr, r.tri, r.b1, r.b2, r.wi, r.lobe, beam);
    }
}


// ---------------------------------------------------------------------------------------
// Manifold walks (Manifolds.slangh:57-281).
// ---------------------------------------------------------------------------------------
struct MsHit {
    uint32_t tri;
    float b1, b2;
    V3 pos;
    uint32_t lobe;
};
struct MsData {  // ManifoldSampleData
    MsHit hits[kMaxMsOccluders];
    V3 woLast;
    int count, iterations;
    float wavelength, pdf;
    bool dispersive;
};
PLT_DEV V3 walk_point(V3 o, V3 d, float t) {
#pragma clang fp contract(off)
    return v3(o.x + d.x * t, o.y + d.y * t, o.z + d.z * t);
}
PLT_DEV DSG dsg_at(const PtScene& sc, uint32_t tri, float b1, float b2) {
    V3 P[3], Nv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint32_t v = sc.indices[3 * tri + k];
        P[k] = v3(sc.positions[3 * v], sc.positions[3 * v + 1], sc.positions[3 * v + 2]);
        Nv[k] = v3(sc.normals[3 * v], sc.normals[3 * v + 1], sc.normals[3 * v + 2]);
    }
    return make_dsg(P[0], P[1], P[2], Nv[0], Nv[1], Nv[2], b1, b2);
}
// manifoldSampleVertex (:84-105): a delta lobe towards the light, recorded as the next walk vertex.
PLT_DEV bool ms_sample_vertex(uint32_t tri, float b1, float b2, V3 pos, const Surf& s, const Mat& m, V3 sdV, V3 lightDir, Rng& sg,
                              MsData& d) {
    BsdfSample smp;
    if (!mat_sample_delta_hinted(m, s.f, sdV, sg, d.wavelength, lightDir, smp)) return false;
    d.pdf *= smp.pdf;
    d.hits[d.count] = MsHit{tri, b1, b2, pos, smp.lobe};
    d.dispersive |= (smp.lobe & kLobeDispersiveDelta) != 0;
    ++d.count;
    return true;
}
// traceMNEEOccluders (:107-150): closest hits along the shadow ray; each must be a specular
// transmitter that samples a transmission lobe. Ends after at most maxOcc hits.
PLT_DEV bool ms_trace_occluders(const PtScene& sc, V3 org, V3 dir, float tmax, int maxOcc, int maxMSBounces, Rng& sg, MsData& d,
                                int* stk) {
    const V3 wi = -dir;
    for (;;) {
        const Ray r{org.x, org.y, org.z, dir.x, dir.y, dir.z, 0.f, tmax};
        const Hit h = traverse<false, kTraversalStackDepth>(sc.tsc, r, stk, kPtBlock);
        if (h.tri == kNoHit) return true;
        if (d.count >= maxOcc || d.count >= maxMSBounces) return false;
        const V3 posW = walk_point(org, dir, h.t);
        Surf s = surface_at(sc, h.tri, h.b1, h.b2, wi);
        s.pos = posW;
        const Mat m = material_at(sc, s.material);
        if (!(mat_props(m) & kPropMWTransmissive)) return false;
        if (!ms_sample_vertex(h.tri, h.b1, h.b2, posW, s, m, wi, -wi, sg, d) || (d.hits[d.count - 1].lobe & kLobeTransmission) == 0)
            return false;
        org = ray_origin(posW, wdot(s.faceN, dir) >= 0.f ? s.faceN : -s.faceN);
        tmax -= h.t;
    }
}
// manifoldNEESample (:165-281) for an area light at lightPos: Newton steps per vertex, each
// projected back onto its mesh by a ray from the previous vertex, with a step size that quarters on
// a miss and doubles (up to 1) on a hit; then the visibility of the light from the last vertex.
// Every pass over unconverged vertices costs at least one iteration, so the walk ends after at most
// maxIterations ray casts.
PLT_DEV bool ms_nee_sample(const PtScene& sc, const PtFrame& fr, V3 visOrg, V3 visDir, float visTmax, V3 origin, V3 lightPos,
                           bool distantLight, V3 lightDir, int maxMSBounces, Rng& sg, MsData& d, int* stk) {
    if (!ms_trace_occluders(sc, visOrg, visDir, visTmax, int(fr.mnee_max_occluders), maxMSBounces, sg, d, stk)) return false;
    if (d.count == 0) return true;
    float beta[kMaxMsOccluders];
#pragma unroll
    for (int i = 0; i < int(kMaxMsOccluders); ++i) beta[i] = 1.f;
    bool success = false;
    while (!success) {
        success = true;
        for (int idx = 0; idx < d.count; ++idx) {
            MsHit& h = d.hits[idx];
            const DSG dsg = dsg_at(sc, h.tri, h.b1, h.b2);
            const V3 p0 = idx == 0 ? origin : d.hits[idx - 1].pos;
            const PDV pdv0 = dsg.pdw(p0);
            const bool last = idx + 1 == d.count;
            const V3 p2 = last ? lightPos : d.hits[idx + 1].pos;
            // a distant light fixes the last direction (Manifolds.slangh:199-202)
            const PDV pdv2 = last && distantLight ? PDV{lightDir, v3(0.f, 0.f, 0.f), v3(0.f, 0.f, 0.f)} : dsg.pdw(p2);
            // the scattered direction is taken on the into-material side (avoids total internal reflection)
            const bool woIsIn = wdot(pdv2.v, dsg.n.v) < 0.f;
            const PDV wi = woIsIn ? pdv0 : pdv2, wo = woIsIn ? pdv2 : pdv0;
            const Mat m = material_at(sc, sc.mesh_material[sc.tri_mesh[h.tri]]);
            const PDV wos = mat_partial_derivatives(m, wi, dsg, h.lobe, d.wavelength);
            if (wos.v.x == 0.f && wos.v.y == 0.f && wos.v.z == 0.f) return false;
            float lenC;
            V3 stepDir;
            if (!manifold_one_step(wo, wos, dsg.p, lenC, stepDir)) return false;
            if (lenC < fr.mnee_threshold) continue;
            success = false;
            for (;;) {  // line search
                if (++d.iterations > int(fr.mnee_max_iterations)) return false;
                V3 proposal;
                {
#pragma clang fp contract(off)
                    proposal = h.pos - wscale(stepDir, beta[idx]);
                }
                V3 dirTo = proposal - p0;
                const float dist = wlength(dirTo);
                dirTo = wdiv(dirTo, dist);
                const Ray r{p0.x, p0.y, p0.z, dirTo.x, dirTo.y, dirTo.z, fmaxf(1e-4f, .99f * dist), 1.01f * dist};
                const Hit ch = traverse<false, kTraversalStackDepth>(sc.tsc, r, stk, kPtBlock);
                if (ch.tri == kNoHit || sc.tri_mesh[ch.tri] != sc.tri_mesh[h.tri]) {
                    beta[idx] /= 4.f;
                    continue;
                }
                h.tri = ch.tri;
                h.b1 = ch.b1;
                h.b2 = ch.b2;
                h.pos = walk_point(p0, dirTo, ch.t);
                beta[idx] = fminf(1.f, 2.f * beta[idx]);
                break;
            }
        }
    }
    const V3 p = d.hits[d.count - 1].pos;
    V3 dir;
    float dist;
    if (!distantLight) {
        dir = lightPos - p;
        dist = wlength(dir);
        dir = wdiv(dir, dist);
    } else {
        dir = lightDir;
        dist = kFltMax;
    }
    d.woLast = dir;
    const Ray vr{p.x, p.y, p.z, dir.x, dir.y, dir.z, 1e-4f, dist};
    return traverse<true, kTraversalStackDepth>(sc.tsc, vr, stk, kPtBlock).tri == kNoHit;
}
// manifoldForwardUnwind (pltpt_solve.rt.slang:80-93): the beam through the walk's vertices, last first.
template <int N>
PLT_DEV void ms_forward_unwind(const PtScene& sc, const PtFrame& fr, const MsData& d, V3 origin, Beam<N>& beam) {
    for (int idx = d.count - 1; idx >= 0; --idx) {
        const MsHit& h = d.hits[idx];
        const V3 p2 = idx == 0 ? origin : d.hits[idx - 1].pos;
        unwind_step<N>(sc, fr, h.tri, h.b1, h.b2, wnormalize(p2 - h.pos), h.lobe, beam, &h.pos);
    }
}

// Four waves per SIMD: the solve kernel's live state (beam, light sample, material, walk) needs
// over 256 registers at its peak, which allows one wave per SIMD; capping it at 128 registers
// (spilling the rest to scratch) gives the traversals four waves to hide latency behind. At
// 1080p depth 8 with manifold NEE: 18.0 ms per frame at one wave, 11.3 ms at two, 10.5 ms at
// three, 10.1 ms at four, 13.0 ms at five (96 registers, three times the spills). `Lights` instantiates the analytic-light and environment-map code only
// for scenes that have them, so scenes lit by emissive geometry alone do not carry its registers.
template <int N, bool Lights>
__global__ void __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(4, 4)))
solve_kernel(PtScene sc, PtFrame fr) {
    __shared__ int stack[kTraversalStackDepth * kPtBlock];
    const uint32_t li = blockIdx.x * kPtBlock + threadIdx.x;
    if (li >= fr.n_paths) return;
    if (fr.solve_overflow) {  // the split pass's fallback: only paths whose events did not fit
        const uint32_t* rg = fr.solve_overflow + size_t(li / 64u) * kRegWords;
        const uint32_t lane = li & 63u;
        if (((rg[lane < 32u ? kRegOverflowLo : kRegOverflowHi] >> (lane & 31u)) & 1u) == 0u) return;
    }
    const uint32_t pid = path_pixel(fr.width, fr.height, fr.path_base + li);
    const uint32_t px = pid % fr.width, py = pid / fr.width;
    const size_t n = fr.n_paths;
    int* stk = stack + threadIdx.x;
    const bool doNEE = (fr.options & kOptNEE) != 0, doMIS = doNEE && (fr.options & kOptMIS) != 0;
    const bool useEmissive = (fr.options & kOptEmissiveLights) != 0 && sc.n_lights > 0;
    const bool useAnalytic = Lights && (fr.options & kOptAnalyticLights) != 0 && sc.n_analytic > 0;
    const bool useEnv = Lights && (fr.options & kOptEnvLight) != 0 && sc.env_w > 0;
    const bool importance = (fr.options & kOptISEmitters) != 0;
    const bool doMNEE = doNEE && (fr.options & kOptMNEE) != 0;
    // getLightTypeSelectionProbabilities (LightSampling.slangh:37-55): equal weights over the
    // present light types
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
    // per-tile NEE light-type draws (SampleGenerator(kTile, gFrameCount << 2), pltpt_solve.rt.slang:412)
    const size_t tileBase = fr.tile_draws ? size_t((py / fr.tile_size) * fr.tiles_x + px / fr.tile_size) * (fr.max_bounces + 1) : 0;
    uint32_t neeDraws = 0;
    Solve sv;
    sv.r = sv.g = sv.b = 0.f;
    sv.origin = v3(sc.cam_pos[0], sc.cam_pos[1], sc.cam_pos[2]);
    sv.normal = normalize(v3(sc.cam_w[0], sc.cam_w[1], sc.cam_w[2]));

    for (int p = 0;; ++p) {
        const Rec bounce = read_rec(fr.records, n, li, p);
        if (bounce.flags & kFlagMiss) {
            // onMiss (pltpt_solve.rt.slang:371-396): the environment map, MIS-weighted against its
            // own sampling when NEE could have found it
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
                    const int comps = (!(fr.options & kOptHwssMis) && (bounce.flags & kFlagDispersive)) ? 1 : N;
                    Beam<N> beam = source_beam_distant<N>(wo, fr.max_omega, Le, lam, fr.max_omega, comps);
                    forward_unwind<N>(sc, fr, n, li, p, beam);
                    measure<N>(sc, fr, sv, p, bounce.contrib, beam);
                }
            }
            break;
        }
        const Surf s = surface_at(sc, bounce.tri, bounce.b1, bounce.b2, bounce.wi);
        const Mat m = material_at(sc, s.material);
        const uint32_t props = mat_props(m);

        // ---- NEE (pltpt_solve.rt.slang:245-366) ----
        if (doNEE) {
            const uint32_t lobes = mat_lobes(m);
            // a delta reflector can take part in a manifold walk from here (pltpt_solve.rt.slang:254-258)
            const bool allowMS = doMNEE && p > 0 && !(bounce.flags & kFlagLastDelta) && (props & kPropMWReflective);
            const bool sampleRefl = (lobes & kLobeNonDeltaReflection) != 0 || allowMS;
            const bool sampleTrans = (lobes & kLobeNonDeltaTransmission) != 0;
            if (sampleRefl || sampleTrans) {
                // selectLightType (LightSampling.slangh:57-85), from the per-tile generator when
                // NEEUsePerTileSG is on (LightSampling.slangh:181-185)
                float uType = fr.tile_draws ? fr.tile_draws[tileBase + neeDraws++] : sg.next1D();
                int ltype = -1;  // SampledLightType: 0 environment map, 1 emissive, 2 analytic
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
                V3 lsDir = v3(0.f, 0.f, 0.f), lsOrigin = v3(0.f, 0.f, 0.f), lightPos = v3(0.f, 0.f, 0.f), tPos = v3(0.f, 0.f, 0.f);
                // the sourced beam: position, area (or solid angle) and whether it is distant
                V3 bPos = v3(0.f, 0.f, 0.f);
                float bSize = fr.sourcing_area;
                bool bDistant = false;
                float Le[N], lam[N];
#pragma unroll
                for (int j = 0; j < N; ++j) { Le[j] = 0.f; lam[j] = lambda; }
                if (valid) {
                    // Uncontracted, division-normalised geometry: the cosine test and the shadow ray
                    // round exactly as a strict-IEEE evaluation of the same expressions.
#pragma clang fp contract(off)
                    const float uL = sg.next1D();
                    uint32_t ti;
                    if (fr.options & kOptUniformEmissive) {
                        // EmissiveUniformSampler.sampleLight (EmissiveUniformSampler.slang:52-72); the
                        // pass rejects an empty active list before launch
                        ti = sc.active_lights[min(uint32_t(uL * float(sc.n_active)), sc.n_active - 1)];
                    } else {
                        // EmissivePowerSampler.sampleLight (EmissivePowerSampler.slang:62-83)
                        ti = min(uint32_t(uL * float(sc.n_lights)), sc.n_lights - 1);
                        const uint2 pk = sc.alias[ti];
                        const float thr = float(__builtin_bit_cast(_Float16, uint16_t(pk.x >> 16)));
                        const uint32_t above = ((pk.x & 0xFFFFu) << 8) | ((pk.y >> 24) & 0xFFu);
                        const uint32_t below = pk.y & 0xFFFFFFu;
                        ti = sg.next1D() >= thr ? above : below;
                    }
                    // the sampler's selection pdf (power: flux / total; uniform: 1 / active count)
                    const float selPdf = sc.light_pdf[ti];
                    // sampleTriangle (EmissiveLightSamplerHelpers.slang:76-125)
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
                        lightPos = ray_origin(tPos, ln);
                        // computeNewRayOrigin(sd, dir) (ShadowRays.slangh:50-52)
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
                    // generateEnvMapSample (LightSampling.slangh:87-104): radiance over pdf, sourced
                    // from infinity with the largest beam solid angle
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
                lsPdf *= typePdf;  // account for the light type selection
                // generateLightSample's acceptance (LightSampling.slangh:196-205)
                const bool isRefl = dot(s.faceN, lsDir) * dot(s.faceN, bounce.wi) > 0.f;
                valid = valid && lsPdf > 1e-6f && !(!sampleRefl && isRefl) && !(!sampleTrans && !isRefl);
                bool anyLe = false;
#pragma unroll
                for (int j = 0; j < N; ++j) anyLe |= Le[j] > 0.f;
                if (valid && anyLe) {
                    const int comps = (!(fr.options & kOptHwssMis) && (bounce.flags & kFlagDispersive)) ? 1 : N;
                    Beam<N> beam = bDistant ? source_beam_distant<N>(-lsDir, bSize, Le, lam, fr.max_omega, comps)
                                            : source_beam_area<N>(bPos, -lsDir, bSize, lsDist, Le, lam, fr.max_omega, comps);
                    beam.scale(1.f / typePdf);
                    // analytic lights are not hit by BSDF samples: no MIS (pltpt_solve.rt.slang:275)
                    const bool misL = doMIS && ltype != 2;
                    bool evalCurrent = true, ok = true;
                    const int maxMSBounces = int(fr.max_bounces) - p;
                    if (!doMNEE || maxMSBounces <= 0) {
                        const Ray vr{lsOrigin.x, lsOrigin.y, lsOrigin.z, lsDir.x, lsDir.y, lsDir.z, 0.f, lsDist};
                        ok = traverse<true, kTraversalStackDepth>(sc.tsc, vr, stk, kPtBlock).tri == kNoHit;
                        if (ok && misL) {
                            const float bp = mat_pdf(m, s.f, bounce.wi, lsDir, lam[0]);
                            beam.scale(lsPdf / (lsPdf + bp));
                        }
                    } else {  // manifold sampling on NEE (pltpt_solve.rt.slang:296-366)
                        const V3 lightDir = -beam.r;
                        V3 dirL = lsDir;
                        float misBsdfPdf = 0.f;
                        MsData ms;
                        ms.count = 0;
                        ms.iterations = 0;
                        ms.wavelength = beam.lambda[0];
                        ms.pdf = 1.f;
                        ms.dispersive = false;
                        ms.woLast = v3(0.f, 0.f, 0.f);
                        // MS on the current vertex too (never on the first bounce)?
                        const bool isRefl2 = dot(s.f.N, lightDir) * dot(s.f.N, bounce.wi) > 0.f;
                        const bool currentVertexMS = isRefl2 && allowMS;
                        if (currentVertexMS)
                            ok = ms_sample_vertex(bounce.tri, bounce.b1, bounce.b2, s.pos, s, m, bounce.wi, lightDir, sg, ms) &&
                                 (ms.hits[0].lobe & kLobeReflection) != 0;
                        const V3 origin = !currentVertexMS ? s.pos : sv.origin;
                        ok = ok && ms_nee_sample(sc, fr, lsOrigin, lsDir, lsDist, origin, beam.pos, beam.distant, lightDir,
                                                 maxMSBounces, sg, ms, stk);
                        ok = ok && !(ms.pdf < 1e-6f);
                        if (ok && ms.count > 0) {
                            if (!(fr.options & kOptHwssMis) && ms.dispersive) beam.comps = 1;
                            if (!beam.distant) {
                                dirL = ms.woLast;
                                beam.r = -ms.woLast;
                                beam.t = perp_stark(beam.r);
                            }
                            beam.scale(1.f / ms.pdf);
                            ms_forward_unwind<N>(sc, fr, ms, origin, beam);
                            if (currentVertexMS) {
                                evalCurrent = false;
                                misBsdfPdf = bounce.lastPdf;
                            }
                        }
                        if (ok && misL) {
                            if (ms.count == 0 || !currentVertexMS) misBsdfPdf = mat_pdf(m, s.f, bounce.wi, dirL, beam.lambda[0]);
                            beam.scale(lsPdf / (lsPdf + misBsdfPdf * ms.pdf));
                        }
                    }
                    if (ok) {
                        // evalNEE (pltpt_solve.rt.slang:219-243): the current vertex's non-delta lobes
                        if (evalCurrent) unwind_step<N>(sc, fr, bounce.tri, bounce.b1, bounce.b2, bounce.wi, kLobeNonDelta, beam);
                        forward_unwind<N>(sc, fr, n, li, p, beam);
                        measure<N>(sc, fr, sv, p, bounce.contrib, beam);
                    }
                }
            }
        }

        // ---- emissive hit (evalEmissive, pltpt_solve.rt.slang:152-200) ----
        // wasNEEAvailable (pltpt_solve.rt.slang:205-210)
        const bool didNEE = doNEE && p > 0 &&
                            (!(bounce.flags & kFlagLastDelta) ||
                             (doMNEE && (bounce.flags >> 16) <= fr.mnee_max_occluders && p > 1));
        if (useEmissive && (props & kPropEmissive) && (doMIS || !didNEE)) {
            const bool needMIS = doMIS && didNEE;
            float lightPdf = 0.f;
            const uint32_t lt = sc.tri_light[bounce.tri];
            if (lt != kNoHit) {
                // evalTrianglePdf (EmissiveLightSamplerHelpers.slang:134-150)
                const V3 hn = s.front ? s.faceN : -s.faceN;
                const V3 toL = s.pos - sv.origin;
                const float d2 = dot(toL, toL);
                if (d2 > kFltMin) {
                    const V3 L = toL * (1.f / sqrtf(d2));
                    const float ct = dot(hn, -L);
                    // getEmissiveSelectionProbability() * emissiveSampler.evalPdf (pltpt_solve.rt.slang:172)
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
                const V3 dd = sv.origin - s.pos;
                const float dist = sqrtf(dot(dd, dd));
                const int comps = (!(fr.options & kOptHwssMis) && (bounce.flags & kFlagDispersive)) ? 1 : N;
                Beam<N> beam = source_beam_area<N>(s.pos, bounce.wi, fr.sourcing_area, dist, Le, lam, fr.max_omega, comps);
                forward_unwind<N>(sc, fr, n, li, p, beam);
                measure<N>(sc, fr, sv, p, bounce.contrib, beam);
            }
        }

        if ((bounce.flags & kFlagTerminated) || p == int(fr.max_bounces)) break;
        sv.origin = s.pos;
        sv.normal = s.f.N;
    }
    fr.color[pid] = make_float4(sv.r, sv.g, sv.b, 1.f);
}

}  // namespace
}  // namespace plt
