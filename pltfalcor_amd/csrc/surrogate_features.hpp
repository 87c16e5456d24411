// Inputs of the learned BSDF surrogate (no counterpart in the reference), shared by the exact-sample
// generator that trains it and by the path tracer that queries it, so both feed the network the
// same 22 numbers for the same lobe evaluation.
//
// The surrogate stands for the Mueller matrix of one wavelength's first-order scattering (SPM) lobe
// of three material families: the coated opaque dielectric and the coated conductor
// (PLTCoatedOpaqueDielectricBSDF.slang:69-124, PLTCoatedConductorBSDF.slang:72-128, scatter lobe
// only) and the grated conductor's base (zeroth-order, non-delta) lobe
// (PLTDiffractionGratedConductorBSDF.slang:68-107). The matrix maps the incident Stokes vector, in
// the frame whose tangent is perp_stark(-wi) in the shading frame, to the exitant one in the frame
// of construct_local_frame(wo) -- the frames in which the generator's beams arrive and leave.
#pragma once
#include "plt_materials.hpp"

namespace plt {

enum SurFamily : int { kSurCoatedOpaque = 0, kSurCoatedConductor = 1, kSurGrated = 2, kSurFamilies = 3 };
constexpr int kSurIn = PLT_SURROGATE_INPUTS, kSurOut = PLT_SURROGATE_OUTPUTS;
// Coherence areas (um^2) the network is trained on, log-uniform over [1, 1e8]: from beams coherent
// over a micrometre to beams coherent over centimetres. The render clamps larger areas to
// kSurAreaClamp, inside the trained range (whose sampled maximum falls just short of 1e8), where the
// coherence factor of a coat of at most 5 um is already 1 to within 1e-6. The input is ln(area)
// scaled to [-1, 1]; kSurAreaNeutral is the value for lobes that do not depend on it.
constexpr float kSurLogAreaMid = 9.21034037f, kSurAreaClamp = 5e7f, kSurAreaNeutral = 1e4f;

// The family whose evaluation of `lobe` the surrogate stands for, or -1.
PLT_DEV int sur_family(const Mat& m, uint32_t lobe) {
    const bool scatterOnly = (lobe & kLobeScatterReflection) != 0 && (lobe & kLobeDelta) == 0;
    switch (m.type) {
        case PLT_MAT_COATED_OPAQUE_DIELECTRIC: return scatterOnly ? int(kSurCoatedOpaque) : -1;
        case PLT_MAT_COATED_CONDUCTOR: return scatterOnly ? int(kSurCoatedConductor) : -1;
        case PLT_MAT_GRATED_CONDUCTOR: {
            int lx, ly;
            decode_order(lobe, lx, ly);
            return scatterOnly && lx == 0 && ly == 0 ? int(kSurGrated) : -1;
        }
        default: return -1;
    }
}

// The 22 inputs for directions wi (towards the light) and wo (towards the eye) in the shading
// frame, one wavelength (nm) and the beam's coherence area (um^2): wi, wo, wavelength, roughness q
// and correlation length sqrt(T) of the K-correlation surface, log coherence area, the family one-hot,
// four features of the K-correlation lobe (log of its PSD argument k l |wi_t + wo_t|, of k q, of
// k l and of |wi_t + wo_t|) and the family's decoded (binary16-rounded) parameters, each scaled to
// about [-1, 1] over the training ranges.
PLT_DEV void sur_features(const Mat& m, int fam, V3 wi, V3 wo, float lambda, float area, float* x) {
    const float k = k2Pi / (1e-3f * lambda);
    const float q = m.fs.q, l = sqrtf(m.fs.T);
    float p[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const float ht = sqrtf((wi.x + wo.x) * (wi.x + wo.x) + (wi.y + wo.y) * (wi.y + wo.y));
    p[0] = (logf(k * l * ht + 1e-3f) - 1.f) / 3.f;
    p[1] = (logf(k * q) - 1.f) / 1.2f;
    p[2] = (logf(k * l) - 2.f) / 1.5f;
    p[3] = (logf(ht + 1e-4f) + 1.5f) / 2.5f;
    if (fam != kSurGrated) {
        p[4] = fam == kSurCoatedOpaque ? 2.f * rgb_to_spectrum(m.col[0], m.col[1], m.col[2], lambda) - 1.f : 0.f;
        p[5] = 2.f * (m.coatIor - 1.3f) / .5f - 1.f;
        p[6] = 2.f * (m.tau - .2f) / 4.8f - 1.f;
    } else {
        p[4] = 2.f * (1.f / m.gr.ipx - .5f) / 1.5f - 1.f;
        p[5] = m.gr.gx;
        p[6] = m.gr.gy;
        p[7] = 2.f * (m.gr.q - .05f) / .45f - 1.f;
    }
    x[0] = wi.x; x[1] = wi.y; x[2] = wi.z; x[3] = wo.x; x[4] = wo.y; x[5] = wo.z;
    x[6] = (lambda - 550.f) / 170.f;
    x[7] = (q - .325f) / .275f;
    x[8] = (l - .46f) / .41f;
    x[9] = (logf(area) - kSurLogAreaMid) / kSurLogAreaMid;
#pragma unroll
    for (int f = 0; f < kSurFamilies; ++f) x[10 + f] = f == fam ? 1.f : 0.f;
#pragma unroll
    for (int f = 0; f < 9; ++f) x[13 + f] = p[f];
}

// The beams the surrogate is trained on: one wavelength arriving along -wi with coherence area
// `area` (um^2) at 1 mm from its source, frame tangent perp_stark(-wi), as the batched material
// interface sources them (material_batch.hip).
PLT_DEV Beam<1> unit_beam(V3 wi, float lambda, float area, const float* S) {
    Beam<1> b;
    b.S[0][0] = S[0]; b.S[0][1] = S[1]; b.S[0][2] = S[2]; b.S[0][3] = S[3];
    b.lambda[0] = lambda;
    b.spdf[0] = 1.f;
    b.comps = 1;
    const float l0 = 1e-3f * lambda;
    const float d = kPi * (l0 * l0) / area;
    b.D00 = d; b.D01 = 0.f; b.D11 = d;
    b.r_m = 1e-3f;
    b.pos = v3(0.f, 0.f, 0.f);
    b.r = -wi;
    b.t = perp_stark(b.r);
    b.distant = false;
    return b;
}
PLT_DEV Frame identity_frame() {
    Frame f;
    f.T = v3(1.f, 0.f, 0.f);
    f.B = v3(0.f, 1.f, 0.f);
    f.N = v3(0.f, 0.f, 1.f);
    return f;
}
// Exact outgoing Stokes vector of `lobe` for incident S (the surrogate's training target, one
// column of the Mueller matrix per unit S).
PLT_DEV void exact_stokes(const Mat& m, V3 wi, V3 wo, uint32_t lobe, float lambda, float area, const float* S, float* out) {
    Beam<1> b = unit_beam(wi, lambda, area, S);
    mat_eval<1>(m, identity_frame(), wi, wo, lobe, b);
#pragma unroll
    for (int c = 0; c < 4; ++c) out[c] = b.S[0][c];
}

}  // namespace plt
