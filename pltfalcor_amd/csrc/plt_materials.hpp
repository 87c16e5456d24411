// Device material layer of the path tracer: decodes a plt_material (the reference's binary16
// BasicMaterialData block) into a register-resident instance at each hit, and implements the PLT
// material interface (Rendering/Materials/IMaterialInstance.slang:122-184) for the five materials
// of the path workloads: eval on a PLT beam with N HWSS wavelengths, sample, pdf, lambdaPdf,
// getLobes and the properties the transport reads.
//
// Dispatch is a switch on the type tag in registers; the wavefront scheduler sorts hits by
// material so that a wave mostly executes one case.
#pragma once
#include "plt_device.hpp"

namespace plt {

enum : uint32_t {
    kLobeDiffuseReflection = 0x01, kLobeDispersiveDeltaReflection = 0x08, kLobeNonDeltaReflection = 0x03,
    kLobeNonDeltaTransmission = 0x30, kLobeDispersiveDelta = 0x88, kLobeAll = 0xffff, kLobeMaterialMask = 0xff00
};
enum : uint32_t { kPropTransmissive = 0x1, kPropEmissive = 0x2, kPropMWReflective = 0x4, kPropMWTransmissive = 0x8 };

constexpr float kMeanK = k2Pi / (550.f * 1e-3f);   // FractalSurface / opaque dielectric mean wavenumber
constexpr float kQcutoffForMS = .15f;               // IMaterialInstance.slang:44

struct Frame {
    V3 T, B, N;
    PLT_DEV V3 to_local(V3 v) const { return v3(dot(v, T), dot(v, B), dot(v, N)); }
    PLT_DEV V3 from_local(V3 v) const { return T * v.x + B * v.y + N * v.z; }
};

struct BsdfSample {
    V3 wo;
    float pdf, weight;
    uint32_t lobe;
};

// K-correlation surface (FractalSurface.slang:38-135).
struct Fractal {
    float T, gamma, q;
    PLT_DEV static Fractal from_roughness(float rough, float gamma) {
        Fractal f;
        f.gamma = fmaxf(1.1f, gamma);
        const float a = fminf(fmaxf(rough, 0.f), .75f);
        const float a2 = a * a;
        f.T = fminf(5000.f, (1.f - a2) / (4.f * (kMeanK * kMeanK) * a2));
        f.q = 1.f / sqrtf(1.f + 4.f * (kMeanK * kMeanK) * f.T);
        return f;
    }
    PLT_DEV float pgf(float zx, float zy) const { return fractal_pgf(T, gamma, zx, zy); }
    PLT_DEV float alpha(float ci, float co, float k) const { return surface_alpha(ci, co, q, k); }
    PLT_DEV float one_minus_alpha(float ci, float co, float k) const { return surface_one_minus_alpha(ci, co, q, k); }
    // Sampling after Holzschuch & Pacanowski 2017 (FractalSurface.slang:97-124).
    PLT_DEV V3 sample(V3 wi, Rng& sg, float k, float& pdf) const {
        const float s = sqrtf(fmaxf(0.f, 1.f - wi.z * wi.z));
        const float phi_i = s > 0.f ? atan2f(wi.y, wi.x) : 0.f;
        const float sqrtT = sqrtf(T);
        const float u0 = sg.next1D(), u1 = sg.next1D();
        (void)sg.next1D();
        (void)sg.next1D();
        const float M = 1.f - powf(1.f + k * k * T * ((1.f + s) * (1.f + s)), -(gamma - 1.f) * 0.5f);
        const float f = sqrtf(powf(1.f - M * u0, -2.f / (gamma - 1.f)) - 1.f) / sqrtT;
        const float fk = f / k;
        const float phi_max = (f == 0.f || s == 0.f) ? kPi
                              : acosf(fminf(fmaxf((fk * fk + s * s - 1.f) / (2.f * fk * s), -1.f), 1.f));
        const float phi_f = phi_i + (2.f * u1 - 1.f) * phi_max;
        float sp, cp;
        sincosf(phi_f, &sp, &cp);
        const float zx = f * cp, zy = f * sp;
        const float wx = zx / k - wi.x, wy = zy / k - wi.y;
        const float z = sqrtf(fmaxf(0.f, 1.f - (wx * wx + wy * wy)));
        pdf = kPi / phi_max * z * pgf(zx, zy);
        return v3(wx, wy, sgnf(wi.z) * z);
    }
    PLT_DEV float pdf(V3 wi, V3 wo, float k) const { return pgf(k * (wi.x + wo.x), k * (wi.y + wo.y)); }
};

// Gaussian surface (GaussianSurface.slang:37-110).
struct Gauss {
    float sigma2, q;
    PLT_DEV static float rough_to_sigma2(float r) { return powf(fminf(fmaxf(1.f - .5f * r, 0.f), 1.f), 28.f) * 2500.f; }
    PLT_DEV static Gauss from_roughness(float r) {
        Gauss g;
        const float s2 = rough_to_sigma2(r), s2max = rough_to_sigma2(.1f), s2r0 = rough_to_sigma2(0.f);
        const float t = (s2r0 - s2) / (s2r0 - s2max);
        g.q = 0.f + t * (1.f - 0.f);
        g.sigma2 = fminf(s2max, s2);
        return g;
    }
    PLT_DEV bool delta_only() const { return q == 0.f; }
    // Truncated Box-Muller sample around -wi.xy (MathHelpers.slang:416-434).
    PLT_DEV V3 sample(V3 wi, Rng& sg, float k, float& pdf) const {
        const float s2 = (1.f / sigma2) / (k * k);
        const float mx = -wi.x, my = -wi.y;
        const float u0 = sg.next1D(), u1 = sg.next1D();
        const float mm = mx * mx + my * my;
        const float l = sqrtf(fminf(1.f, mm));
        const float coso = sqrtf(fmaxf(0.f, 1.f - mm));
        const float phi_i = (mx != 0.f || my != 0.f) ? atan2f(my, mx) : 0.f;
        const float s = expf(-.5f * (1.f + l) * (1.f + l) / s2);
        const float x = (1.f - s) * fmaxf(kFltEps, u0) + s;
        const float r = sqrtf(-2.f * s2 * logf(x));
        const float max_phi = (r < kFltEps || l < kFltEps) ? kPi
                              : acosf(fminf(fmaxf((r * r + l * l - 1.f) / (2.f * r * l), -1.f), 1.f));
        const float phi = phi_i + kPi + max_phi * (2.f * u1 - 1.f);
        float sp, cp;
        sincosf(phi, &sp, &cp);
        pdf = .5f * x / (max_phi * s2) * coso;
        const float wx = r * cp + mx, wy = r * sp + my;
        const float z = sqrtf(fmaxf(0.f, 1.f - (wx * wx + wy * wy)));
        return v3(wx, wy, sgnf(wi.z) * z);
    }
    PLT_DEV float pdf(V3 wi, V3 wo, float k) const {
        return boxmueller_trunc_pdf(wo.x, wo.y, -wi.x, -wi.y, (1.f / sigma2) / (k * k));
    }
};

// Bessel J_nu for |nu| < 7 by blending a rational small-argument form and the asymptotic
// form (MathHelpers.slang:851-876).
PLT_DEV float bessel_j(int nu, float x) {
    const float sgn = (nu > 0 || ((-nu) % 2) == 0) ? 1.f : -1.f;
    nu = abs(nu);
    const float cutoff = float(nu) + 1.f, interp = cutoff / 10.f;
    const float laa = sqrtf(2.f / kPi / x) * cosf(x - float(nu) * kPi / 2.f - kPi / 4.f);
    float saa = __builtin_nanf("");
    if (nu < 7) {
        const uint32_t fact[13] = {1u, 1u, 2u, 6u, 24u, 120u, 720u, 5040u, 40320u, 362880u, 3628800u, 39916800u, 479001600u};
        const float x2 = x * x, x4 = x2 * x2, x6 = x2 * x4, x8 = x4 * x4;
        const float s = 6.f * powf(x / 2.f, float(nu));
        const float a = 1.f / (6.f * float(fact[nu]));
        const float b = -x2 / (24.f * float(fact[nu + 1]));
        const float c1 = 11.f * x8 - 864.f * float(6 + nu) *
                         (x6 - 12.f * (5.f + float(nu)) * (3.f * x4 + 32.f * float(4 + nu) * (-2.f * x2 + 27.f * float(3 + nu))));
        const float c = c1 / float(fact[6 + nu]) * x4 * 5.652695401144601e-10f;
        saa = s * (a + b + c);
    }
    if (nu >= 7 || fabsf(x) >= cutoff + interp) return laa * sgn;
    const float w = fmaxf(0.f, fminf(1.f, .5f * (x - cutoff - interp) / interp));
    return (saa + w * (laa - saa)) * sgn;
}

// One-dimensional or 2-D sinusoidal phase grating (DiffractionGrating.slang:41-186).
struct Grating {
    float gx, gy;      // grating direction in the shading tangent plane
    float ipx, ipy;    // 1/pitch (1/um), 0 = no grating along that axis
    float q;           // height (um)
    uint32_t lobes, type;
    float mult;
    PLT_DEV bool is1d() const { return ipy == 0.f; }
    PLT_DEV float alpha(float wiz, float k) const { const float a = wiz * q * k; return expf(-(a * a)); }
    PLT_DEV float lobe_intensity(int lx, int ly, float wiz, float lambda_um) const {
        const uint32_t t = type & 0xFu;
        const float a = 4.f * kPi * q / (lambda_um * fabsf(wiz));
        float ix, iy;
        if (t == 0) {
            ix = iy = lx == 0 ? 1.f : bessel_j(lx, a);
            if (!is1d()) iy = ly == 0 ? 1.f : bessel_j(ly, a);
        } else if (t == 1) {
            const float s = sinf(a / 2.f);
            const float ax = kPi * float(lx) / 2.f;
            ix = iy = lx == 0 ? 1.f : s * (sinf(ax) / ax);
            if (!is1d()) { const float ay = kPi * float(ly) / 2.f; iy = ly == 0 ? 1.f : s * (sinf(ay) / ay); }
        } else {
            ix = iy = lx == 0 ? 1.f : 1.f / sqrtf(fabsf(float(lx)));
            if (!is1d()) iy = ly == 0 ? 1.f : 1.f / sqrtf(fabsf(float(ly)));
        }
        return mult * ix * iy;
    }
    // Grating equation for order (lx, ly) (DiffractionGrating.slang:142-156).
    PLT_DEV V3 diffract(V3 wi, int lx, int ly, float lambda_um) const {
        const float px = sqrtf(wi.x * wi.x + wi.z * wi.z), py = sqrtf(wi.y * wi.y + wi.z * wi.z);
        const float sx = px > kFltEps ? wi.x / px : 0.f, sy = py > kFltEps ? wi.y / py : 0.f;
        const float a = lambda_um * float(lx) * ipx - sx;
        const float b = lambda_um * float(ly) * ipy - sy;
        const float m = (a * a - 1.f) / ((a * b) * (a * b) - 1.f);
        const float qq = 1.f - b * b * m;
        if (!(fabsf(a) <= 1.f && fabsf(b) <= 1.f)) return v3(0.f, 0.f, 0.f);
        return v3(a * sqrtf(fmaxf(0.f, qq)), b * sqrtf(m), sqrtf(fmaxf(0.f, 1.f - a * a * qq - b * b * m)));
    }
    // Samples a diffraction order proportionally to its efficiency (DiffractionGrating.slang:109-140).
    PLT_DEV void sample_lobe(V3 wi, float lambda_um, Rng& sg, int& lx, int& ly, float& px, float& py) const {
        float inten[8];
        float total = 0.f;
        for (uint32_t l = 0; l < lobes; ++l) {
            const float li = lobe_intensity(int(l) + 1, 0, wi.z, lambda_um);
            total += li;
            inten[l] = li;
        }
        const float rx = (sg.next1D() - .5f) * 2.f, ry = (sg.next1D() - .5f) * 2.f;
        float cdf = 0.f;
        lx = ly = 0;
        px = py = 0.f;
        for (uint32_t l = 0; l < lobes; ++l) {
            const float p = inten[l] / total;
            cdf += p;
            const bool sx = fabsf(rx) < cdf && lx == 0, sy = fabsf(ry) < cdf && ly == 0;
            if (sx) { px = p; lx = int(l) + 1; }
            if (sy) { py = p; ly = int(l) + 1; }
        }
        px /= 2.f;
        py /= 2.f;
        lx = rx < 0.f ? -lx : lx;
        ly = ry < 0.f ? -ly : ly;
        if (is1d()) { ly = 0; py = 1.f; }
    }
};
// Diffraction order packed in the material-specific lobe bits (PLTDiffractionGratedConductorBSDF.slang:54-66).
PLT_DEV uint32_t encode_order(int lx, int ly, uint32_t lobe) {
    const int sx = lx < 0 ? 0 : 1, sy = ly < 0 ? 0 : 1;
    const int ax = min(abs(lx), 7), ay = min(abs(ly), 7);
    lobe &= ~kLobeMaterialMask;
    lobe += uint32_t((ay << 3) + ax + (sy << 7) + (sx << 6)) << 8;
    return lobe;
}
PLT_DEV void decode_order(uint32_t lobe, int& lx, int& ly) {
    const int x = int(lobe) >> 8;
    lx = x & 0x7;
    ly = (x >> 3) & 0x7;
    lx *= 2 * ((x >> 6) & 0x1) - 1;
    ly *= 2 * ((x >> 7) & 0x1) - 1;
}

// Register-resident material instance.
struct Mat {
    uint32_t type;
    float col[3];      // base colour / specular / albedo
    float trans[3];    // dielectric transmission colour
    // dielectric
    float etaInt, etaExt, dispB;
    Gauss gs;
    // K-correlation surfaces and conductors
    Fractal fs;
    const plt_spectral_profile* pn;
    const plt_spectral_profile* pk;
    float extEta;
    // coated (tau in um) / thin dielectric (tau in mm)
    float coatIor, tau;
    // thin dielectric: birefringence and optic axis xy
    float bir, oax, oay;
    // multilayer stack
    float mlIor[2], mlThick[2];
    uint32_t layers;
    // grating
    Grating gr;
    // emission
    uint32_t emission;
};

PLT_DEV float h16(uint16_t h) { return float(__builtin_bit_cast(_Float16, h)); }

// A parameter block this layer evaluates: a PLT MaterialType (2-11) and, for the gratings, not a
// radial grating (type bit 0x10, DiffractionGrating.slangh:37), whose direction comes from texture
// coordinates that the meshes here do not carry.
__host__ __device__ inline bool material_supported(const plt_material& m) {
    if (m.type < PLT_MAT_DIFFUSE || m.type > PLT_MAT_MULTILAYER_STACK) return false;
    const bool grated = m.type == PLT_MAT_GRATED_CONDUCTOR || m.type == PLT_MAT_GRATED_OPAQUE_DIELECTRIC;
    return !(grated && ((m.data1[3] >> 8) & 0x10u));
}

// setupMaterialInstance of each *Material.slang.
PLT_DEV Mat setup_material(const plt_material& m, const plt_spectral_profile* prof, uint32_t nprof) {
    Mat r;
    r.type = m.type;
    r.col[0] = h16(m.base_color[0]); r.col[1] = h16(m.base_color[1]); r.col[2] = h16(m.base_color[2]);
    r.trans[0] = h16(m.transmission[0]); r.trans[1] = h16(m.transmission[1]); r.trans[2] = h16(m.transmission[2]);
    r.pn = m.ior_n_profile < nprof ? prof + m.ior_n_profile : nullptr;
    r.pk = m.ior_k_profile < nprof ? prof + m.ior_k_profile : nullptr;
    r.emission = m.emission_profile;
    r.extEta = h16(m.ior);
    const float sx = h16(m.specular[0]), sy = h16(m.specular[1]), sz = h16(m.specular[2]), sw = h16(m.specular[3]);
    switch (m.type) {
        case PLT_MAT_DIELECTRIC: {  // PLTDielectricMaterial.slang:40-56
            r.etaExt = sw;
            r.dispB = (r.trans[0] > 0.f || r.trans[1] > 0.f || r.trans[2] > 0.f) ? sx : 0.f;
            r.etaInt = fmaxf(1.f, h16(m.ior) - 2.897f * r.dispB);
            r.gs = Gauss::from_roughness(sy);
            break;
        }
        case PLT_MAT_CONDUCTOR:  // PLTConductorMaterial.slang:40-57
            r.fs = Fractal::from_roughness(sy, sx);
            break;
        case PLT_MAT_OPAQUE_DIELECTRIC:  // PLTOpaqueDielectricMaterial.slang:40-54
            r.fs = Fractal::from_roughness(sy, 3.f);
            break;
        case PLT_MAT_COATED_OPAQUE_DIELECTRIC:  // PLTCoatedOpaqueDielectricMaterial.slang:40-60
            r.fs = Fractal::from_roughness(sy, 3.f);
            r.tau = h16(m.data1[0]) * sz;
            r.coatIor = h16(m.data1[3]);
            break;
        case PLT_MAT_GRATED_CONDUCTOR:           // PLTDiffractionGratedConductorMaterial.slang:40-70
        case PLT_MAT_GRATED_OPAQUE_DIELECTRIC: {  // PLTDiffractionGratedOpaqueDielectricMaterial.slang:40-70
            r.fs = Fractal::from_roughness(h16(m.data2[1]), 3.f);
            const float ptx = h16(m.data1[0]), pty = h16(m.data1[1]);
            r.gr.ipx = ptx > 0.f ? 1.f / ptx : 0.f;
            r.gr.ipy = pty > 0.f ? 1.f / pty : 0.f;
            r.gr.q = h16(m.data2[0]) * sx;
            r.gr.mult = sy;
            const uint32_t tl = m.data1[3];
            r.gr.lobes = min(max(tl & 0xFFu, 1u), 8u);
            r.gr.type = tl >> 8;
            const float dir = h16(m.data1[2]);
            r.gr.gx = cosf(dir);
            r.gr.gy = sinf(dir);
            break;
        }
        case PLT_MAT_COATED_CONDUCTOR:  // PLTCoatedConductorMaterial.slang:40-60
            r.fs = Fractal::from_roughness(sy, 3.f);
            r.tau = h16(m.data1[0]) * sz;
            r.coatIor = h16(m.data1[3]);
            break;
        case PLT_MAT_THIN_DIELECTRIC:  // PLTThinDielectricMaterial.slang:40-56
            r.tau = sw;
            r.bir = r.col[0] * sx;
            r.oax = r.col[1];
            r.oay = r.col[2];
            break;
        case PLT_MAT_MULTILAYER_STACK:  // PLTMultiLayeredStackMaterial.slang:40-57
            r.mlIor[0] = h16(m.data2[0]);
            r.mlIor[1] = h16(m.data2[1]);
            r.mlThick[0] = h16(m.data1[0]);
            r.mlThick[1] = h16(m.data1[1]);
            r.layers = m.data1[3];
            break;
        default:
            break;
    }
    return r;
}

PLT_DEV uint32_t mat_lobes(const Mat& m) {
    switch (m.type) {
        case PLT_MAT_DIELECTRIC: {
            const uint32_t dt = m.dispB != 0.f ? kLobeDispersiveDeltaTransmission : kLobeDeltaTransmission;
            return m.gs.delta_only() ? (kLobeDeltaReflection | dt)
                                     : (kLobeDeltaReflection | dt | kLobeScatterReflection | kLobeScatterTransmission);
        }
        case PLT_MAT_DIFFUSE: return kLobeDiffuseReflection;
        case PLT_MAT_THIN_DIELECTRIC: return kLobeDeltaReflection | kLobeDeltaTransmission;
        case PLT_MAT_MULTILAYER_STACK: return kLobeDeltaReflection;
        case PLT_MAT_GRATED_CONDUCTOR:
        case PLT_MAT_GRATED_OPAQUE_DIELECTRIC:
            return (m.fs.q == 0.f ? kLobeDeltaReflection : (kLobeDeltaReflection | kLobeScatterReflection)) |
                   kLobeDispersiveDeltaReflection;
        default:
            return m.fs.q == 0.f ? kLobeDeltaReflection : (kLobeDeltaReflection | kLobeScatterReflection);
    }
}
PLT_DEV uint32_t mat_props(const Mat& m) {
    switch (m.type) {
        case PLT_MAT_DIELECTRIC: return kPropTransmissive | (m.gs.q < kQcutoffForMS ? kPropMWTransmissive : 0u);
        case PLT_MAT_DIFFUSE: return m.emission != PLT_NO_PROFILE ? kPropEmissive : 0u;
        case PLT_MAT_GRATED_CONDUCTOR:
        case PLT_MAT_GRATED_OPAQUE_DIELECTRIC: return kPropMWReflective;
        case PLT_MAT_MULTILAYER_STACK: return kPropMWReflective;
        case PLT_MAT_THIN_DIELECTRIC: return kPropTransmissive | kPropMWReflective | kPropMWTransmissive;
        default: return m.fs.q < kQcutoffForMS ? kPropMWReflective : 0u;
    }
}

// BSDFProperties.albedo of getProperties (the albedo output of the sample pass): the specular
// colour times the normal-incidence Fresnel reflectance at 550 nm for the dielectric
// (PLTDielectricBSDF.slang:342-346) and the conductors (PLTConductorBSDF.slang:226-231,
// PLTDiffractionGratedConductorBSDF.slang:273-280, n and k not divided by the exterior IoR), 1 for
// the thin dielectric (PLTThinDielectricBSDF.slang:259), the base colour for the others.
PLT_DEV void mat_albedo(const Mat& m, float out[3]) {
    float f = 1.f;
    switch (m.type) {
        case PLT_MAT_DIELECTRIC: {
            const float uk = kMeanK * k1_2Pi;  // indexOfRefraction(2 pi / 0.55 um): Cauchy in 1 / lambda^2
            float cT;
            f = fresnel_dielectric_F((m.etaInt + m.dispB * (uk * uk)) / m.etaExt, 1.f, cT);
            break;
        }
        case PLT_MAT_CONDUCTOR:
        case PLT_MAT_GRATED_CONDUCTOR: {
            const float n = m.pn ? eval_spectrum(m.pn, 550.f) : 0.f, k = m.pk ? eval_spectrum(m.pk, 550.f) : 0.f;
            float rsr, rsi, rpr, rpi;
            f = fresnel_conductor_amp(1.f, n, k, rsr, rsi, rpr, rpi) ? .5f * ((rsr * rsr + rsi * rsi) + (rpr * rpr + rpi * rpi)) : 0.f;
            break;
        }
        case PLT_MAT_THIN_DIELECTRIC:
            out[0] = out[1] = out[2] = 1.f;
            return;
        default:
            break;
    }
    out[0] = m.col[0] * f;
    out[1] = m.col[1] * f;
    out[2] = m.col[2] * f;
}

PLT_DEV bool is_grated(const Mat& m) { return m.type == PLT_MAT_GRATED_CONDUCTOR || m.type == PLT_MAT_GRATED_OPAQUE_DIELECTRIC; }

// Shading frame used for eval / sample / pdf: the grated materials evaluate in a frame whose
// tangent follows the grating direction (DiffractionGrating.slang:72-79).
PLT_DEV Frame mat_frame(const Mat& m, const Frame& sd) {
    if (!is_grated(m)) return sd;
    Frame g;
    g.T = sd.from_local(v3(m.gr.gx, m.gr.gy, 0.f));
    g.N = sd.N;
    g.B = cross(g.N, g.T);
    return g;
}

// ---------------------------------------------------------------------------------------
// eval on the local frame (beam already rotated to the exitant frame).
// ---------------------------------------------------------------------------------------
template <int N>
PLT_DEV void conductor_like_eval(const Mat& m, V3 wi, V3 wo, uint32_t lobe, Beam<N>& beam) {
    Conductor c;
    c.etaN = m.pn; c.etaK = m.pk; c.extEta = m.extEta; c.T = m.fs.T; c.gamma = m.fs.gamma; c.q = m.fs.q;
    c.spec[0] = m.col[0]; c.spec[1] = m.col[1]; c.spec[2] = m.col[2];
    conductor_eval<N>(c, wi, wo, lobe, beam);
}

// interfere (Interference.slang:45-57): two Stokes vectors combined per transverse component with
// phasors ps, pp and mutual coherence coh; the result replaces a.
PLT_DEV void interfere4(float a[4], const float b[4], float ps, float pp, float coh) {
    const float S1 = (a[0] + a[1]) / 2.f, P1 = (a[0] - a[1]) / 2.f, S2 = (b[0] + b[1]) / 2.f, P2 = (b[0] - b[1]) / 2.f;
    const float S = S1 + S2 + 2.f * sqrtf(fmaxf(0.f, S1 * S2)) * ps * coh;
    const float P = P1 + P2 + 2.f * sqrtf(fmaxf(0.f, P1 * P2)) * pp * coh;
    const float nrm = a[0] + b[0] > 1e-6f ? (S + P) / (a[0] + b[0]) : 0.f;
    const float a2 = (a[2] + b[2]) * nrm, a3 = (a[3] + b[3]) * nrm;
    a[0] = S + P;
    a[1] = S - P;
    a[2] = a2;
    a[3] = a3;
}

// PLTCoatedConductorBxDF::eval (PLTCoatedConductorBSDF.slang:72-128): coat transmission in and
// out around the conductor's Fresnel at the half vector; the direct lobe interferes with the
// coat's own reflection (phase 2 tau / cos k eta + arg r per polarisation, mutual coherence
// over the coat thickness).
template <int N>
PLT_DEV void coated_conductor_eval(const Mat& m, V3 wi, V3 wo, uint32_t lobe, Beam<N>& beam) {
    const bool isRefl = wi.z * wo.z > 0.f;
    const bool isScatter = (lobe & kLobeScatterReflection) != 0, isDirect = (lobe & kLobeDeltaReflection) != 0;
    const float coatior = m.extEta + (m.coatIor - m.extEta) * (fminf(fmaxf(m.tau, 0.f), .05f) / .05f);
    const float eta1 = coatior / m.extEta;
    float cT;
    (void)fresnel_dielectric_F(eta1, wi.z, cT);
    const V3 wiIn = -refract_dir(wi, cT, eta1);
    V3 woIn = v3(-wiIn.x, -wiIn.y, wiIn.z);
    if (!isDirect) {
        float cT2;
        (void)fresnel_dielectric_F(eta1, wo.z, cT2);
        woIn = -refract_dir(wo, cT2, eta1);
    }
    if (!isRefl || (!isScatter && !isDirect) || wiIn.z <= 0.f || woIn.z <= 0.f) { beam.set_zero(); return; }
    float k[N], coh[N];
#pragma unroll
    for (int j = 0; j < N; ++j) k[j] = k2Pi / (1e-3f * beam.lambda[j]);
    const V3 h = wiIn + woIn;
    const V3 mh = normalize(h);
    const float LdotH = fmaxf(0.f, dot(wiIn, mh));
    const V3 Z = beam.to_local(v3(0.f, 0.f, 1.f));
    beam.mutual_coherence(k, Z * m.tau, coh);
    const Fres MiR = fresnel_dielectric_mueller(wi.z, eta1, true);
    const Fres MiT = fresnel_dielectric_mueller(wi.z, eta1, false);
    const Fres MoT = fresnel_dielectric_mueller(-woIn.z, eta1, false);
    const float icoat = 1.f / coatior;
    const float dist = 2.f * m.tau / wiIn.z;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (j < beam.comps) {
            const float n = (m.pn ? eval_spectrum(m.pn, beam.lambda[j]) : 0.f) * icoat;
            const float kk = (m.pk ? eval_spectrum(m.pk, beam.lambda[j]) : 0.f) * icoat;
            float rsr, rsi, rpr, rpi;
            const bool ok = fresnel_conductor_amp(LdotH, n, kk, rsr, rsi, rpr, rpi);
            const Fres R2 = ok ? fres_from_amp(rsr, rsi, rpr, rpi) : Fres{0.f, 0.f, 0.f, 0.f};
            const Fres M = fres_mul(fres_mul(MiT, R2), MoT);
            const float m00 = rgb_to_spectrum(m.col[0], m.col[1], m.col[2], beam.lambda[j]);
            float a[4] = {beam.S[j][0], beam.S[j][1], beam.S[j][2], beam.S[j][3]};
            float b[4] = {a[0], a[1], a[2], a[3]};
            apply(M, a[0], a[1], a[2], a[3]);
            if (isScatter) {
                const float w = m00 * m.fs.one_minus_alpha(wiIn.z, woIn.z, k[j]) * fabsf(wi.z) * m.fs.pgf(k[j] * h.x, k[j] * h.y);
                a[0] *= w; a[1] *= w; a[2] *= w; a[3] *= w;
            } else {
                const float w = m00 * m.fs.alpha(wiIn.z, woIn.z, k[j]);
                a[0] *= w; a[1] *= w; a[2] *= w; a[3] *= w;
                apply(MiR, b[0], b[1], b[2], b[3]);
                // phaseShift(eta, dist, k, r) = sin(dist k eta + arg r) (Interference.slang:36-39)
                const float pS = sinf(dist * k[j] * coatior + atan2f(rsi, rsr));
                const float pP = sinf(dist * k[j] * coatior + atan2f(rpi, rpr));
                interfere4(a, b, pS, pP, coh[j]);
            }
            beam.S[j][0] = a[0]; beam.S[j][1] = a[1]; beam.S[j][2] = a[2]; beam.S[j][3] = a[3];
        }
    }
}

// Uniaxial slab optics (Rendering/Materials/PLT/Birefringence.slang:33-197): y is the surface
// normal, incidence in the yz plane, unit wavenumber.
struct Biref {
    PLT_DEV static float qo(float eo, float K) { return sqrtf(eo * eo - K * K); }
    PLT_DEV static void qe(float eo, float ee, V3 A, float K, float& qp, float& qm) {
        const float e = ee * ee - eo * eo;
        const float G = A.y * A.y * e + eo * eo;
        const float r = sqrtf(fmaxf(0.f, K * K * (A.x * A.x * e - ee * ee) + ee * ee * G));
        const float c = A.z * A.y * K * e;
        qp = (eo * r + c) / G;
        qm = (-eo * r + c) / G;
    }
    PLT_DEV static V3 E_e(float eo, V3 A, float K, float q) {
        return normalize(v3(A.x * eo * eo, K * A.z * q + A.y * (eo * eo - q * q), K * A.y * q + A.z * (eo * eo - K * K)));
    }
    PLT_DEV static V3 E_o(V3 A, float K, float q) { return normalize(v3(-A.y * K - A.z * q, A.x * K, A.x * q)); }
    PLT_DEV static V3 poynting(float q, V3 E, float K) {
        return normalize(v3(E.x * (K * E.z - q * E.y), K * E.z * E.y + q * (E.x * E.x + E.z * E.z),
                            -q * E.z * E.y - K * (E.x * E.x + E.y * E.y)));
    }
};

// PLTThinDielectricBxDF::handle_birefringence (PLTThinDielectricBSDF.slang:60-118): ordinary and
// extraordinary transmission through the slab and their interference (mutual coherence over the
// lateral o/e offset), applied to the (x, y) intensities of the beam.
template <int N>
PLT_DEV void thin_birefringence(const Mat& m, Beam<N>& beam, V3 wi, const float* k) {
    const float oaz = sqrtf(fmaxf(0.f, 1.f - (m.oax * m.oax + m.oay * m.oay)));
    const float phii = atan2f(wi.y, wi.x);
    const float phiA = atan2f(m.oay, m.oax) - phii;
    const V3 I = v3(0.f, fabsf(wi.z), sqrtf(1.f - wi.z * wi.z));
    const float sA = sqrtf(1.f - oaz * oaz);
    const V3 A = v3(cosf(phiA) * sA, sgnf(wi.z) * oaz, sinf(phiA) * sA);
    const V3 A2 = v3(A.x, -A.y, A.z);
    const V3 Z = beam.to_local(v3(cosf(phii), sinf(phii), 0.f));
    const float ei = 1.f, eo = m.extEta - .5f * m.bir, ee = m.extEta + .5f * m.bir;
    const float tauum = m.tau * 1e+3f;
    const float ct = I.y, st = I.z;
    // vectors_in_slab (:168-195)
    float qepA, qemA;
    const float Kin = ei * I.z;
    Biref::qe(eo, ee, A, Kin, qepA, qemA);
    const float qomA = -Biref::qo(eo, Kin);
    const V3 Io = Biref::poynting(qomA, Biref::E_o(A, Kin, qomA), Kin);
    const V3 Ie = Biref::poynting(qemA, Biref::E_e(eo, A, Kin, qemA), Kin);
    const V3 We = normalize(v3(0.f, qemA, -Kin));
    const float e_eff = Kin / sqrtf(fmaxf(0.f, 1.f - We.y * We.y));
    // fresnel_iso_aniso (:98-125), transmission coefficients
    float tso, tse, tpo, tpe;
    {
        const float K = ei * st;
        const float qi = -Biref::qo(ei, K);
        const float q2om = -Biref::qo(eo, K);
        float qp, q2em;
        Biref::qe(eo, ee, A, K, qp, q2em);
        const V3 E2om = Biref::E_o(A, K, q2om);
        const V3 E2em = Biref::E_e(eo, A, K, q2em);
        const float A_1 = -E2em.x * E2om.y * K * ct;
        const float A_2 = -E2em.y * E2om.x * K * ct;
        const float C_1 = qi * ct - K * st;
        const float C_2 = E2em.x * E2om.z * (C_1 + q2om * ct) - A_1;
        const float C_3 = E2em.z * E2om.x * (C_1 + q2em * ct) - A_2;
        const float Ns = C_2 * (qi + q2em) - C_3 * (qi + q2om);
        tpo = -(2.f * E2em.x * ct * (q2em + qi) * C_1) / Ns;
        tpe = (2.f * E2om.x * ct * (q2om + qi) * C_1) / Ns;
        tso = -(2.f * qi * (E2em.y * K * ct + E2em.z * C_1 + E2em.z * q2em * ct)) / Ns;
        tse = (2.f * qi * (E2om.y * K * ct + E2om.z * C_1 + E2om.z * q2om * ct)) / Ns;
    }
    // fresnel_aniso_iso (:126-166) with the mirrored axis, transmission coefficients
    float tos, top, tes, tep;
    {
        const float K = ei * st;
        const float q2om = -Biref::qo(ei, K);
        const float qop = Biref::qo(eo, K);
        float qep, qem;
        Biref::qe(eo, ee, A2, K, qep, qem);
        const V3 Eop = Biref::E_o(A2, K, qop), Eom = Biref::E_o(A2, K, -qop);
        const V3 Eep = Biref::E_e(eo, A2, K, qep), Eem = Biref::E_e(eo, A2, K, qem);
        const float B_1 = -Eep.x * K * ct * (q2om - qep);
        const float B_2 = -Eom.x * K * ct * (q2om + qop);
        const float B_3 = -Eop.x * K * ct * (qop - q2om);
        const float D_1 = K * st - q2om * ct;
        const float D_2 = Eep.z * (D_1 + qep * ct);
        const float D_3 = Eem.z * (D_1 + qem * ct);
        const float D_4 = Eop.z * (D_1 + qop * ct);
        const float N2 = Eep.x * (qep - q2om) * D_4 + Eop.x * (q2om - qop) * D_2 + Eop.y * B_1 + Eep.y * B_3;
        const float roo = -(Eep.x * Eom.z * (q2om - qep) * (qop * ct - D_1) + Eom.x * (q2om + qop) * D_2 + Eom.y * B_1 - Eep.y * B_2) / N2;
        const float roe = -(-Eom.x * Eop.z * (qop * (qop * ct + K * st) + q2om * D_1) +
                            Eom.z * Eop.x * (qop * (qop * ct - K * st) + q2om * D_1) + Eop.y * B_2 + Eom.y * B_3) / N2;
        tos = (Eep.x * qep * roe - Eom.x * qop + Eop.x * qop * roo) / q2om;
        top = -(Eom.z * qop - Eom.y * K - (Eep.y * K + Eep.z * qep) * roe - (Eop.y * K + Eop.z * qop) * roo) / D_1;
        const float reo = (Eem.x * (qem - q2om) * (D_2 + Eep.y * K * ct) + Eep.x * (q2om - qep) * (D_3 + Eem.y * K * ct)) / N2;
        const float ree = -(Eem.x * (qem - q2om) * D_4 - Eem.x * Eop.y * K * ct * (q2om - qem) + Eem.y * B_3 + Eop.x * (q2om - qop) * D_3) / N2;
        tes = (Eem.x * qem + Eep.x * qep * ree + Eop.x * qop * reo) / q2om;
        tep = -(Eem.z + Eep.z * ree + Eop.z * reo) / ct;
    }
    const float ss = tso * tos + tse * tes, sp = tso * top + tse * tep;
    const float ps = tpo * tos + tpe * tes, pp = tpo * top + tpe * tep;
    const float Ioz = fabsf(tauum / Io.y) * Io.z, Iez = fabsf(tauum / Ie.y) * Ie.z;
    const float OPLo = fabsf(tauum / Io.y) * eo, OPLe = fabsf(tauum / Ie.y) * e_eff;
    const float oez = Ioz - Iez;
    float coh[N], nLx[N], nLy[N], newL[N];
    beam.mutual_coherence(k, Z * oez, coh);
    bool allPos = true, allFinite = true;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const float t = coh[j] * 2.f * sinf(-k[j] * (ei * oez * I.z + OPLo - OPLe));
        const float Lx = .5f * (beam.S[j][0] + beam.S[j][1]), Ly = .5f * (beam.S[j][0] - beam.S[j][1]);
        const float sq = sqrtf(Lx * Ly);
        nLx[j] = fmaxf(0.f, ss * ss * Lx + ps * ps * Ly + ss * ps * t * sq);
        nLy[j] = fmaxf(0.f, sp * sp * Lx + pp * pp * Ly + sp * pp * t * sq);
        newL[j] = nLx[j] + nLy[j];
        allFinite = allFinite && isfinite(newL[j]);
        allPos = allPos && beam.S[j][0] > 1e-6f && isfinite(newL[j]);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const float nrm = allPos ? newL[j] / beam.S[j][0] : 0.f;
        beam.S[j][0] = allFinite ? newL[j] : 0.f;
        beam.S[j][1] = allFinite ? nLx[j] - nLy[j] : 0.f;
        beam.S[j][2] *= nrm;
        beam.S[j][3] *= nrm;
    }
}

// PLTThinDielectricBxDF::eval (PLTThinDielectricBSDF.slang:120-149): reflection R + T R (1-R)^-1 T,
// transmission T (1-R)^-1 T (Mueller blocks), birefringence in transmission.
template <int N>
PLT_DEV void thin_dielectric_eval(const Mat& m, V3 wi, V3 wo, uint32_t lobe, Beam<N>& beam) {
    const float eta = m.extEta;
    const bool isRefl = wi.z * wo.z > 0.f;
    const bool isDirect = (lobe & kLobeDelta) != 0;
    const bool hasT = (lobe & kLobeTransmission) != 0, hasR = (lobe & kLobeReflection) != 0;
    if ((isRefl && eta == 1.f) || wi.z * wo.z == 0.f || !isDirect || (!hasT && !isRefl) || (!hasR && isRefl)) {
        beam.set_zero();
        return;
    }
    const float c = fabsf(wi.z);
    const Fres R = fresnel_dielectric_mueller(c, eta, true);
    const Fres T = fresnel_dielectric_mueller(c, eta, false);
    const Fres iR = fres_inv_one_minus(R);
    const Fres M = isRefl ? (iR.a > 0.f ? fres_add(R, fres_mul(fres_mul(T, R), fres_mul(iR, T))) : R)
                          : (iR.a > 0.f ? fres_mul(T, fres_mul(iR, T)) : fres_mul(T, T));
    if (m.bir > 0.f && !isRefl) {
        float k[N];
#pragma unroll
        for (int j = 0; j < N; ++j) k[j] = k2Pi / (1e-3f * beam.lambda[j]);
        thin_birefringence<N>(m, beam, wi, k);
    }
#pragma unroll
    for (int j = 0; j < N; ++j)
        if (j < beam.comps) apply(M, beam.S[j][0], beam.S[j][1], beam.S[j][2], beam.S[j][3]);
}

// PLTMultiLayeredStackBxDF::eval (PLTMultiLayeredStackBSDF.slang:58-95): Airy recursion of the
// amplitude G <- (R + G p) / (1 + R G p) over the layers from the base, output Re(G)^2. The
// recursion runs in FP64 (20 complex steps per wavelength at the workload's stack).
template <int N>
PLT_DEV void multilayer_eval(const Mat& m, V3 wi, V3 wo, uint32_t lobe, Beam<N>& beam) {
    const bool isRefl = wi.z * wo.z > 0.f;
    if ((lobe & kLobeDelta) == 0 || !isRefl) { beam.set_zero(); return; }
    const float eta1 = m.mlIor[0] / m.extEta, eta2 = m.mlIor[1] / m.mlIor[0];
    float cT0, cT1, cT2;
    const float Rout = sqrtf(fresnel_dielectric_F(eta1, wi.z, cT0));
    const float r0 = sqrtf(fresnel_dielectric_F(eta2, -cT0, cT1));
    const float r1 = sqrtf(fresnel_dielectric_F(eta2, cT1, cT2));
    const double c0 = fabs(double(cT0)), c1 = fabs(double(cT1));
    const double ph0 = 2.0 * double(m.mlThick[0]) / c0 * double(m.mlIor[0]);
    const double ph1 = 2.0 * double(m.mlThick[1]) / c1 * double(m.mlIor[1]);
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (j < beam.comps) {
            const double k = 6.283185307179586 / (1e-3 * double(beam.lambda[j]));
            double gr = sqrt(double(rgb_to_spectrum(m.col[0], m.col[1], m.col[2], beam.lambda[j]))), gi = 0.0;
            for (uint32_t l = 0; l < m.layers; ++l) {
                const bool odd = (l & 1u) != 0;
                double ps, pc;
                sincos((odd ? ph1 : ph0) * k, &ps, &pc);
                const double R = l + 1 == m.layers ? double(Rout) : double(odd ? r1 : r0);
                const double tr = gr * pc - gi * ps, ti = gr * ps + gi * pc;
                // (R + G) / (1 + R G)
                const double nr = R + tr, ni = ti, dr = 1.0 + R * tr, di = R * ti;
                const double id = 1.0 / (dr * dr + di * di);
                gr = (nr * dr + ni * di) * id;
                gi = (ni * dr - nr * di) * id;
            }
            const float w = float(gr * gr);
            beam.S[j][0] *= w; beam.S[j][1] *= w; beam.S[j][2] *= w; beam.S[j][3] *= w;
        }
    }
}

// PLTOpaqueDielectricBxDF::eval (PLTOpaqueDielectricBSDF.slang:57-82): a depolarising rough
// dielectric evaluated at the mean wavenumber, scaled by the albedo spectrum.
template <int N>
PLT_DEV void opaque_dielectric_eval(const Mat& m, V3 wi, V3 wo, uint32_t lobe, Beam<N>& beam) {
    const bool isDirect = (lobe & kLobeDelta) != 0, isScatter = (lobe & kLobeNonDelta) != 0;
    if ((!isScatter && !isDirect) || (lobe & kLobeReflection) == 0) { beam.set_zero(); return; }
    float M = m.fs.alpha(wi.z, wo.z, kMeanK);
    if (isScatter) {
        const V3 h = wi + wo;
        M = m.fs.one_minus_alpha(wi.z, wo.z, kMeanK) * fabsf(wi.z) * m.fs.pgf(kMeanK * h.x, kMeanK * h.y);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (j < beam.comps) {
            const float w = M * rgb_to_spectrum(m.col[0], m.col[1], m.col[2], beam.lambda[j]);
            beam.S[j][0] *= w; beam.S[j][1] *= w; beam.S[j][2] *= w; beam.S[j][3] *= w;
        }
    }
}

template <int N>
PLT_DEV void mat_eval_local(const Mat& m, V3 wi, V3 wo, uint32_t lobe, Beam<N>& beam) {
    switch (m.type) {
        case PLT_MAT_DIELECTRIC: {
            Dielectric d;
            d.etaInt = m.etaInt; d.etaExt = m.etaExt; d.B = m.dispB; d.sigma2 = m.gs.sigma2; d.q = m.gs.q;
            for (int c = 0; c < 3; ++c) { d.spec[c] = m.col[c]; d.trans[c] = m.trans[c]; }
            dielectric_eval<N>(d, wi, wo, lobe, beam);
            break;
        }
        case PLT_MAT_CONDUCTOR:
            conductor_like_eval<N>(m, wi, wo, lobe, beam);
            break;
        case PLT_MAT_DIFFUSE: {  // PLTDiffuseBSDF.slang:43-55
            const bool valid = (lobe & kLobeDiffuseReflection) != 0;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const float w = valid ? fabsf(wi.z) * k1_Pi * rgb_to_spectrum(m.col[0], m.col[1], m.col[2], beam.lambda[j]) : 0.f;
                beam.S[j][0] *= w;
                beam.S[j][1] = beam.S[j][2] = beam.S[j][3] = 0.f;
            }
            break;
        }
        case PLT_MAT_OPAQUE_DIELECTRIC:
            opaque_dielectric_eval<N>(m, wi, wo, lobe, beam);
            break;
        case PLT_MAT_COATED_OPAQUE_DIELECTRIC: {  // PLTCoatedOpaqueDielectricBSDF.slang:69-124
            const bool isRefl = wi.z * wo.z > 0.f;
            const bool isScatter = (lobe & kLobeScatterReflection) != 0, isDirect = (lobe & kLobeDeltaReflection) != 0;
            const float coatior = m.extEta + (m.coatIor - m.extEta) * (fminf(fmaxf(m.tau, 0.f), .05f) / .05f);
            const float eta1 = coatior / m.extEta;
            float cT;
            (void)fresnel_dielectric_F(eta1, wi.z, cT);
            const V3 wiIn = -refract_dir(wi, cT, eta1);
            V3 woIn = v3(-wiIn.x, -wiIn.y, wiIn.z);
            if (!isDirect) {
                float cT2;
                (void)fresnel_dielectric_F(eta1, wo.z, cT2);
                woIn = -refract_dir(wo, cT2, eta1);
            }
            if (!isRefl || (!isScatter && !isDirect) || wiIn.z <= 0.f || woIn.z <= 0.f) { beam.set_zero(); break; }
            float k[N], coh[N];
#pragma unroll
            for (int j = 0; j < N; ++j) k[j] = k2Pi / (1e-3f * beam.lambda[j]);
            const V3 h = wiIn + woIn;
            const V3 mh = normalize(h);
            const float LdotH = fmaxf(0.f, dot(wiIn, mh));
            const V3 Z = beam.to_local(v3(0.f, 0.f, 1.f));
            beam.mutual_coherence(k, Z * m.tau, coh);
            const float dist = 2.f * m.tau / wiIn.z;
            // M = MiT * MoT: products of two transmission Mueller blocks
            const Fres Ti = fresnel_dielectric_mueller(wi.z, eta1, false);
            const Fres To = fresnel_dielectric_mueller(-woIn.z, eta1, false);
            const Fres M{Ti.a * To.a + Ti.b * To.b, Ti.a * To.b + Ti.b * To.a, Ti.c * To.c - Ti.d * To.d, Ti.c * To.d + Ti.d * To.c};
            const Fres Rl = fresnel_dielectric_mueller(isScatter ? LdotH : wi.z, eta1, true);
#pragma unroll
            for (int j = 0; j < N; ++j) {
                if (j < beam.comps) {
                    const float m00 = rgb_to_spectrum(m.col[0], m.col[1], m.col[2], beam.lambda[j]);
                    float w;
                    if (isScatter) {
                        w = m.fs.one_minus_alpha(wiIn.z, woIn.z, k[j]) * fabsf(wi.z) * m.fs.pgf(k[j] * h.x, k[j] * h.y);
                    } else {
                        w = m.fs.alpha(wiIn.z, woIn.z, k[j]);
                    }
                    float a0 = beam.S[j][0], a1 = beam.S[j][1], a2 = beam.S[j][2], a3 = beam.S[j][3];
                    float b0 = a0, b1 = a1, b2 = a2, b3 = a3;
                    apply(M, a0, a1, a2, a3);
                    const float s1 = m00 * w;
                    a0 *= s1; a1 *= s1; a2 *= s1; a3 *= s1;
                    apply(Rl, b0, b1, b2, b3);
                    b0 *= w; b1 *= w; b2 *= w; b3 *= w;
                    // interfere (Interference.slang:45-57) with phasor sin(dist k eta)
                    const float ph = sinf(dist * k[j] * coatior);
                    const float S1 = (a0 + a1) / 2.f, P1 = (a0 - a1) / 2.f, S2 = (b0 + b1) / 2.f, P2 = (b0 - b1) / 2.f;
                    const float S = S1 + S2 + 2.f * sqrtf(fmaxf(0.f, S1 * S2)) * ph * coh[j];
                    const float P = P1 + P2 + 2.f * sqrtf(fmaxf(0.f, P1 * P2)) * ph * coh[j];
                    const float nrm = a0 + b0 > 1e-6f ? (S + P) / (a0 + b0) : 0.f;
                    beam.S[j][0] = S + P;
                    beam.S[j][1] = S - P;
                    beam.S[j][2] = (a2 + b2) * nrm;
                    beam.S[j][3] = (a3 + b3) * nrm;
                }
            }
            break;
        }
        case PLT_MAT_COATED_CONDUCTOR:
            coated_conductor_eval<N>(m, wi, wo, lobe, beam);
            break;
        case PLT_MAT_THIN_DIELECTRIC:
            thin_dielectric_eval<N>(m, wi, wo, lobe, beam);
            break;
        case PLT_MAT_MULTILAYER_STACK:
            multilayer_eval<N>(m, wi, wo, lobe, beam);
            break;
        case PLT_MAT_GRATED_CONDUCTOR:           // PLTDiffractionGratedConductorBSDF.slang:68-107
        case PLT_MAT_GRATED_OPAQUE_DIELECTRIC: {  // PLTDiffractionGratedOpaqueDielectricBSDF.slang:68-101
            const bool opaque = m.type == PLT_MAT_GRATED_OPAQUE_DIELECTRIC;
            const bool isDirect = (lobe & kLobeDelta) != 0;
            float a[N], lam[N];
            bool anyPos = false, anyBelow1 = false;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                lam[j] = 1e-3f * beam.lambda[j];
                a[j] = m.gr.alpha(wi.z, k2Pi / lam[j]);
                anyPos |= a[j] > 0.f;
                anyBelow1 |= a[j] < 1.f;
            }
            int lx, ly;
            decode_order(lobe, lx, ly);
            if (lx == 0 && ly == 0 && anyPos) {
                if (opaque) opaque_dielectric_eval<N>(m, wi, wo, lobe, beam);
                else conductor_like_eval<N>(m, wi, wo, lobe, beam);
                beam.scale_each(a);
            } else if ((lx != 0 || ly != 0) && anyBelow1 && isDirect) {
                const V3 wiBeam = beam.to_local(wi);
                const float iext = 1.f / m.extEta;
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    if (j < beam.comps) {
                        float coh = 1.f;
                        if (j != 0) coh = beam.mutual_coherence_angular(beam.to_local(m.gr.diffract(wo, lx, ly, lam[j])), wiBeam);
                        const float inten = m.gr.lobe_intensity(lx, ly, wi.z, lam[j]);
                        const float w = rgb_to_spectrum(m.col[0], m.col[1], m.col[2], beam.lambda[j]) * (1.f - a[j]) * coh * inten;
                        if (!opaque) {  // the conductor's Fresnel Mueller matrix; the opaque dielectric has none
                            const float n = (m.pn ? eval_spectrum(m.pn, beam.lambda[j]) : 0.f) * iext;
                            const float kk = (m.pk ? eval_spectrum(m.pk, beam.lambda[j]) : 0.f) * iext;
                            const Fres F = fresnel_conductor_mueller(wi.z, n, kk);
                            apply(F, beam.S[j][0], beam.S[j][1], beam.S[j][2], beam.S[j][3]);
                        }
                        beam.S[j][0] *= w; beam.S[j][1] *= w; beam.S[j][2] *= w; beam.S[j][3] *= w;
                    }
                }
            } else {
                beam.set_zero();
            }
            break;
        }
        default:
            beam.set_zero();
    }
}

// Coherence transform after an interaction (coherenceTransform of each BxDF).
template <int N>
PLT_DEV void mat_coherence_transform(const Mat& m, V3 wo, uint32_t lobe, Beam<N>& beam) {
    bool scatter;
    switch (m.type) {
        case PLT_MAT_DIFFUSE: scatter = true; break;
        case PLT_MAT_GRATED_CONDUCTOR:
        case PLT_MAT_GRATED_OPAQUE_DIELECTRIC: {
            int lx, ly;
            decode_order(lobe, lx, ly);
            scatter = (lx == 0 && ly == 0) && (lobe & kLobeDelta) == 0;
            break;
        }
        default: scatter = (lobe & kLobeDelta) == 0; break;
    }
    if (scatter) beam.scatter_coherence(wo.z);
}

// IMaterialInstance::eval: world <-> local frame change around the local eval
// (e.g. PLTConductorBSDF.slang:177-193).
template <int N>
PLT_DEV void mat_eval(const Mat& m, const Frame& sd, V3 sdV, V3 toEye, uint32_t lobe, Beam<N>& beam) {
    const Frame f = mat_frame(m, sd);
    beam.r = f.to_local(beam.r);
    beam.t = f.to_local(beam.t);
    const V3 wi = f.to_local(sdV), wo = f.to_local(toEye);
    beam.construct_local_frame(wo);
    mat_eval_local<N>(m, wi, wo, lobe, beam);
    mat_coherence_transform<N>(m, wo, lobe, beam);
    beam.r = f.from_local(beam.r);
    beam.t = f.from_local(beam.t);
}

// ---------------------------------------------------------------------------------------
// lambdaPdf (spectral MIS); frame per material as in the reference (the grated conductor uses
// the plain shading frame here).
// ---------------------------------------------------------------------------------------
template <int N>
PLT_DEV void mat_lambda_pdf(const Mat& m, const Frame& sd, V3 sdV, V3 woW, uint32_t lobe, const Beam<N>& beamW, float* ret) {
#pragma unroll
    for (int j = 0; j < N; ++j) ret[j] = 1.f;
    Beam<N> beam = beamW;
    beam.r = sd.to_local(beam.r);
    beam.t = sd.to_local(beam.t);
    const V3 toLight = sd.to_local(woW);  // bxdf.lambdaPdf(toLocal(sd.V), toLocal(wo)) names them (toLight, toEye)
    const V3 toEye = sd.to_local(sdV);
    (void)toLight;
    const V3 a0 = sd.to_local(sdV), a1 = sd.to_local(woW);  // argument order of the call: (sd.V, wo)
    const bool isScatter = (lobe & kLobeNonDelta) != 0, isDirect = (lobe & kLobeDelta) != 0;
    switch (m.type) {
        case PLT_MAT_DIELECTRIC: {  // PLTDielectricBSDF.slang:238-270 (toLight = a0, toEye = a1)
            const bool isRefl = a0.z * a1.z > 0.f;
            const V3 lb = beam.to_local(a0);
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const float k = k2Pi / (1e-3f * beam.lambda[j]);
                const float u = k * k1_2Pi;
                const float eta = (m.etaInt + m.dispB * (u * u)) / m.etaExt;
                if (isDirect && !isRefl) {
                    float cT;
                    (void)fresnel_dielectric_F(eta, a1.z, cT);
                    ret[j] = j == 0 ? 1.f : beam.mutual_coherence_angular(beam.to_local(refract_dir(a1, cT, eta)), lb);
                } else if (isScatter) {
                    const V3 w = isRefl ? a0 : scattered_wo(a0, eta);
                    ret[j] = m.gs.pdf(a1, w, k);
                }
            }
            break;
        }
        case PLT_MAT_CONDUCTOR:  // PLTConductorBSDF.slang:149-163
            if (isScatter) {
#pragma unroll
                for (int j = 0; j < N; ++j) ret[j] = m.fs.pdf(a1, a0, k2Pi / (1e-3f * beam.lambda[j]));
            }
            break;
        case PLT_MAT_COATED_CONDUCTOR:          // PLTCoatedConductorBSDF.slang:190-213 (same as below)
        case PLT_MAT_COATED_OPAQUE_DIELECTRIC:  // PLTCoatedOpaqueDielectricBSDF.slang:190-213
            if (isScatter) {
                const float coatior = m.extEta + (m.coatIor - m.extEta) * (fminf(fmaxf(m.tau, 0.f), .05f) / .05f);
                const float eta1 = coatior / m.extEta;
                float cT, cT1;
                (void)fresnel_dielectric_F(eta1, a1.z, cT);
                (void)fresnel_dielectric_F(eta1, a0.z, cT1);
                const V3 wiIn = -refract_dir(a1, cT, eta1);
                const V3 woIn = -refract_dir(a0, cT1, eta1);
#pragma unroll
                for (int j = 0; j < N; ++j) ret[j] = m.fs.pdf(wiIn, woIn, k2Pi / (1e-3f * beam.lambda[j]));
            }
            break;
        case PLT_MAT_GRATED_CONDUCTOR:           // PLTDiffractionGratedConductorBSDF.slang:167-190
        case PLT_MAT_GRATED_OPAQUE_DIELECTRIC: {  // PLTDiffractionGratedOpaqueDielectricBSDF.slang:164-188
            int lx, ly;
            decode_order(lobe, lx, ly);
            bool anyBelow1 = false;
            float lam[N];
#pragma unroll
            for (int j = 0; j < N; ++j) {
                lam[j] = 1e-3f * beam.lambda[j];
                anyBelow1 |= m.gr.alpha(a0.z, k2Pi / lam[j]) < 1.f;
            }
            if ((lx != 0 || ly != 0) && anyBelow1 && isDirect) {
                const V3 lb = beam.to_local(a0);
#pragma unroll
                for (int j = 0; j < N; ++j)
                    ret[j] = j == 0 ? 1.f : beam.mutual_coherence_angular(beam.to_local(m.gr.diffract(a1, lx, ly, lam[j])), lb);
            }
            break;
        }
        default:
            break;
    }
    (void)toEye;
}

// ---------------------------------------------------------------------------------------
// pdf(sd, toLight): bxdf.pdf(toLocal(toLight), toLocal(sd.V)).
// ---------------------------------------------------------------------------------------
PLT_DEV float mat_pdf(const Mat& m, const Frame& sd, V3 sdV, V3 toLightW, float lambda) {
    const Frame f = mat_frame(m, sd);
    const V3 wi = f.to_local(toLightW), wo = f.to_local(sdV);
    const float k = k2Pi / (1e-3f * lambda);
    switch (m.type) {
        case PLT_MAT_DIELECTRIC: {
            Dielectric d;
            d.etaInt = m.etaInt; d.etaExt = m.etaExt; d.B = m.dispB; d.sigma2 = m.gs.sigma2; d.q = m.gs.q;
            return dielectric_pdf(d, wi, wo, lambda);
        }
        case PLT_MAT_CONDUCTOR:
            return m.fs.one_minus_alpha(wi.z, wi.z, k) * m.fs.pdf(wi, wo, k);
        case PLT_MAT_DIFFUSE:
            return wo.z * k1_Pi;  // cosine_hemisphere_concentric_pdf(woLocal), woLocal = toLocal(sd.V)
        case PLT_MAT_OPAQUE_DIELECTRIC:
            return m.fs.one_minus_alpha(wi.z, wi.z, kMeanK) * m.fs.pdf(wi, wo, kMeanK);
        case PLT_MAT_COATED_CONDUCTOR:
        case PLT_MAT_COATED_OPAQUE_DIELECTRIC: {
            if (wi.z <= 0.f || wo.z <= 0.f) return 0.f;
            const float coatior = m.extEta + (m.coatIor - m.extEta) * (fminf(fmaxf(m.tau, 0.f), .05f) / .05f);
            const float eta1 = coatior / m.extEta;
            float cT, cT1;
            const float Fd = fresnel_dielectric_F(eta1, wi.z, cT);
            (void)fresnel_dielectric_F(eta1, wo.z, cT1);
            const V3 wiIn = -refract_dir(wi, cT, eta1), woIn = -refract_dir(wo, cT1, eta1);
            return wiIn.z * woIn.z == 0.f ? 0.f : m.fs.one_minus_alpha(wiIn.z, wiIn.z, k) * (1.f - Fd) * m.fs.pdf(wiIn, woIn, k);
        }
        case PLT_MAT_GRATED_CONDUCTOR:
            return m.fs.one_minus_alpha(wi.z, wi.z, k) * m.fs.pdf(wi, wo, k) * m.gr.alpha(wi.z, k);
        case PLT_MAT_GRATED_OPAQUE_DIELECTRIC:  // PLTDiffractionGratedOpaqueDielectricBSDF.slang:148-156
            return m.fs.one_minus_alpha(wi.z, wi.z, kMeanK) * m.fs.pdf(wi, wo, kMeanK) * m.gr.alpha(wi.z, k);
        default:
            return 0.f;
    }
}

// ---------------------------------------------------------------------------------------
// sample(sd, sg, lambda): BSDF importance sampling in the hero wavelength.
// ---------------------------------------------------------------------------------------
PLT_DEV bool mat_sample(const Mat& m, const Frame& sd, V3 sdV, Rng& sg, float lambda, BsdfSample& s) {
    const Frame f = mat_frame(m, sd);
    const V3 wi = f.to_local(sdV);
    const float k = k2Pi / (1e-3f * lambda);
    V3 wo = v3(0.f, 0.f, 0.f);
    float pdf = 0.f, weight = 0.f;
    uint32_t lobe = 0;
    bool valid = false;
    switch (m.type) {
        case PLT_MAT_DIELECTRIC: {  // PLTDielectricBSDF.slang:159-211 (lobe mask All)
            const float a = m.gs.delta_only() ? 1.f : surface_alpha(wi.z, wi.z, m.gs.q, k);
            const float u = k * k1_2Pi;
            const float eta = (m.etaInt + m.dispB * (u * u)) / m.etaExt;
            const float pdfDirect = a;
            bool isRefl;
            const float uD = sg.next1D_if(pdfDirect != 1.f);
            const bool isDirect = pdfDirect == 1.f || uD < pdfDirect;
            pdf = isDirect ? pdfDirect : 1.f - pdfDirect;
            if (isDirect) {
                float cT;
                const float Fd = fresnel_dielectric_F(eta, wi.z, cT);
                const float uR = sg.next1D_if(Fd > 0.f);
                isRefl = Fd > 0.f && uR <= Fd;
                wo = isRefl ? v3(-wi.x, -wi.y, wi.z) : refract_dir(wi, cT, eta);
                pdf *= isRefl ? Fd : 1.f - Fd;
                lobe = isRefl ? kLobeDeltaReflection : (m.dispB != 0.f ? kLobeDispersiveDeltaTransmission : kLobeDeltaTransmission);
            } else {
                float sp;
                wo = m.gs.sample(wi, sg, k, sp);
                pdf *= sp;
                const V3 mh = normalize(wi + wo);
                float cT;
                const float F = fresnel_dielectric_F(eta, sgnf(wi.z) * dot(wi, mh), cT);
                const float uR = sg.next1D_if(F > 0.f);
                isRefl = F > 0.f && uR <= F;
                if (!isRefl) wo = scattered_wo(wo, eta);
                pdf *= isRefl ? F : 1.f - F;
                lobe = isRefl ? kLobeScatterReflection : kLobeScatterTransmission;
            }
            weight = rgb_to_spectrum(isRefl ? m.col[0] : m.trans[0], isRefl ? m.col[1] : m.trans[1],
                                     isRefl ? m.col[2] : m.trans[2], lambda);
            valid = !((isRefl && wi.z * wo.z <= 0.f) || (!isRefl && wi.z * wo.z >= 0.f));
            break;
        }
        case PLT_MAT_CONDUCTOR:
        case PLT_MAT_OPAQUE_DIELECTRIC: {  // PLTConductorBSDF.slang:101-133, PLTOpaqueDielectricBSDF.slang:89-123
            const float kk = m.type == PLT_MAT_CONDUCTOR ? k : kMeanK;
            const float a = m.fs.alpha(wi.z, wi.z, kk);
            weight = rgb_to_spectrum(m.col[0], m.col[1], m.col[2], lambda);
            const float uD = sg.next1D_if(a != 1.f);
            const bool isDirect = a == 1.f || uD < a;
            pdf = isDirect ? a : 1.f - a;
            if (isDirect) { wo = v3(-wi.x, -wi.y, wi.z); lobe = kLobeDeltaReflection; }
            else { float sp; wo = m.fs.sample(wi, sg, kk, sp); pdf *= sp; lobe = kLobeScatterReflection; }
            valid = true;
            break;
        }
        case PLT_MAT_DIFFUSE: {  // PLTDiffuseBSDF.slang:57-65
            weight = rgb_to_spectrum(m.col[0], m.col[1], m.col[2], lambda);
            const float u0 = sg.next1D(), u1 = sg.next1D();
            wo = cosine_hemisphere(u0, u1);
            pdf = wo.z * k1_Pi;
            lobe = kLobeDiffuseReflection;
            valid = true;
            break;
        }
        case PLT_MAT_THIN_DIELECTRIC: {  // PLTThinDielectricBSDF.slang:164-183 (both lobes allowed)
            float cTd;
            const float Fd = fresnel_dielectric_F(m.extEta, fabsf(wi.z), cTd);
            const float uR = sg.next1D_if(Fd > 0.f);
            const bool isRefl = Fd > 0.f && uR <= Fd;
            wo = isRefl ? v3(-wi.x, -wi.y, wi.z) : -wi;
            weight = 1.f;
            pdf = isRefl ? Fd : 1.f - Fd;
            lobe = isRefl ? kLobeDeltaReflection : kLobeDeltaTransmission;
            valid = true;
            break;
        }
        case PLT_MAT_MULTILAYER_STACK: {  // PLTMultiLayeredStackBSDF.slang:102-115
            if (wi.z <= 0.f) break;
            pdf = 1.f;
            wo = v3(-wi.x, -wi.y, wi.z);
            lobe = kLobeDeltaReflection;
            weight = rgb_to_spectrum(m.col[0], m.col[1], m.col[2], lambda);
            valid = wi.z * wo.z > 0.f;
            break;
        }
        case PLT_MAT_COATED_CONDUCTOR:
        case PLT_MAT_COATED_OPAQUE_DIELECTRIC: {  // PLTCoatedOpaqueDielectricBSDF.slang:131-180 (= coated conductor's)
            if (wi.z <= 0.f) break;
            const float coatior = m.extEta + (m.coatIor - m.extEta) * (fminf(fmaxf(m.tau, 0.f), .05f) / .05f);
            const float eta1 = coatior / m.extEta;
            float cT;
            const float Fd = fresnel_dielectric_F(eta1, wi.z, cT);
            const float uC = sg.next1D_if(Fd != .1f);
            const bool isReflCoat = Fd == .1f || uC < Fd;
            weight = 1.f;
            pdf = isReflCoat ? Fd : 1.f - Fd;
            if (isReflCoat) {
                wo = v3(-wi.x, -wi.y, wi.z);
                lobe = kLobeDeltaReflection;
                valid = true;
                break;
            }
            const V3 wiIn = -refract_dir(wi, cT, eta1);
            if (wiIn.z == 0.f) break;
            weight = rgb_to_spectrum(m.col[0], m.col[1], m.col[2], lambda);
            const float a = m.fs.alpha(wiIn.z, wiIn.z, k);
            const float uD = sg.next1D_if(a != 1.f);
            const bool isDirect = a == 1.f || uD < a;
            pdf *= isDirect ? a : 1.f - a;
            if (isDirect) {
                wo = v3(-wi.x, -wi.y, wi.z);
                lobe = kLobeDeltaReflection;
            } else {
                float sp;
                const V3 woIn = m.fs.sample(wiIn, sg, k, sp);
                pdf *= sp;
                float cT1;
                (void)fresnel_dielectric_F(eta1, -woIn.z, cT1);
                wo = refract_dir(-woIn, cT1, eta1);
                lobe = kLobeScatterReflection;
            }
            valid = wi.z * wo.z > 0.f;
            break;
        }
        case PLT_MAT_GRATED_CONDUCTOR:           // PLTDiffractionGratedConductorBSDF.slang:115-153 (no hint)
        case PLT_MAT_GRATED_OPAQUE_DIELECTRIC: {  // PLTDiffractionGratedOpaqueDielectricBSDF.slang:110-146
            const float lam = 1e-3f * lambda;
            const float a = m.gr.alpha(wi.z, k2Pi / lam);
            const float uD = sg.next1D_if(a != 1.f);
            const bool isDirect = a == 1.f || uD < a;
            pdf = isDirect ? a : 1.f - a;
            // the base lobe's wavenumber: the conductor's own, the opaque dielectric's mean
            const float kb = m.type == PLT_MAT_GRATED_CONDUCTOR ? k : kMeanK;
            if (isDirect) {
                const float ab = m.fs.alpha(wi.z, wi.z, kb);
                weight = rgb_to_spectrum(m.col[0], m.col[1], m.col[2], lambda);
                float bp = 1.f;
                const float uB = sg.next1D_if(ab != 1.f);
                const bool bd = ab == 1.f || uB < ab;
                bp = bd ? ab : 1.f - ab;
                if (bd) { wo = v3(-wi.x, -wi.y, wi.z); lobe = kLobeDeltaReflection; }
                else { float sp; wo = m.fs.sample(wi, sg, kb, sp); bp *= sp; lobe = kLobeScatterReflection; }
                pdf *= bp;
                valid = true;
                break;
            }
            weight = rgb_to_spectrum(m.col[0], m.col[1], m.col[2], lambda);
            int lx, ly;
            float px, py;
            m.gr.sample_lobe(wi, lam, sg, lx, ly, px, py);
            wo = m.gr.diffract(wi, lx, ly, lam);
            lobe = encode_order(lx, ly, kLobeDispersiveDeltaReflection);
            pdf *= px * py;
            valid = true;
            break;
        }
        default:
            break;
    }
    s.wo = f.from_local(wo);
    s.pdf = pdf;
    s.weight = weight;
    s.lobe = lobe;
    return valid;
}

}  // namespace plt
