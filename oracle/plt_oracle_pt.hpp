// PLT oracle, path-tracing part: the remaining PLT materials of the path workloads, material
// setup from the binary16 parameter block, the emissive-light sampler, and the two PLTPT passes
// (sample pass writing bounce records, solve pass replaying them), restated on the CPU.
//
// TEST INFRASTRUCTURE ONLY (see plt_oracle.hpp). FP32 instantiation: per-pixel parity reference.
#pragma once
#include <limits>
#include <random>
#include "plt_oracle.hpp"
#include "plt_oracle_geom.hpp"

namespace pltoracle {

// Material parameter block (Scene/Material/BasicMaterialData.slang:63-107), binary16 as bits.
struct MatData {
    uint32_t type;
    uint16_t base_color[4], specular[4], data1[4], data2[4], transmission[3], ior;
    uint32_t emission, ior_n, ior_k;
};
enum : uint32_t { MT_DIELECTRIC = 0, MT_CONDUCTOR = 1, MT_DIFFUSE = 2, MT_OPAQUE = 3, MT_COATED_OPAQUE = 4, MT_GRATED_CONDUCTOR = 5 };
constexpr uint32_t kNoProfile = 0xFFFFFFFFu;
inline float hf(uint16_t h) { return f16_bits_to_f32(h); }

// PLTDiffuseBxDF (PLTDiffuseBSDF.slang:39-88)
template <class T> struct DiffuseBxDF {
    V3<T> albedo;
    template <int N> void eval(V3<T> toLight, V3<T> /*toEye*/, uint32_t lobe, PLTBeam<T, N>& beam) const {
        const bool valid = (lobe & Lobe_DiffuseReflection) != 0;
        for (int j = 0; j < N; ++j) {
            const T colour = Spectrum<T>::rgbToSpectrum(albedo, beam.wavelengths[j]);
            const T w = valid ? std::fabs(toLight.z) * K<T>::_1_PI * colour : T(0);
            beam.L[j] *= w;
            beam.L1[j] = beam.L2[j] = beam.L3[j] = 0;
        }
    }
    template <class S> bool sample(V3<T> /*wi*/, V3<T>& wo, T& pdf, T& weight, uint32_t& lobe, S& sg, T wavelength, uint32_t lobeMask) const {
        weight = Spectrum<T>::rgbToSpectrum(albedo, wavelength);
        wo = sample_cosine_hemisphere_concentric(sampleNext2D<T>(sg), pdf);
        lobe = Lobe_DiffuseReflection;
        return (lobeMask & Lobe_DiffuseReflection) != 0;
    }
    T pdf(V3<T> /*wi*/, V3<T> wo, T /*wavelength*/) const { return wo.z * K<T>::_1_PI; }
    M2<T> coherenceTransform(V3<T>, V3<T> wo, uint32_t) const {
        const T c = std::sqrt(hmax(T(1e-2), std::fabs(wo.z)));
        return m2<T>(c, 0, 0, T(1) / c);
    }
};

// PLTOpaqueDielectricBxDF (PLTOpaqueDielectricBSDF.slang:43-145)
template <class T> struct OpaqueDielectricBxDF {
    V3<T> albedo{1, 1, 1};
    FractalSurface<T> gfs;
    static T meank() { return K<T>::_2PI / (T(550) * T(1e-3)); }
    template <int N> void eval(V3<T> toLight, V3<T> toEye, uint32_t lobe, PLTBeam<T, N>& beam) const {
        const bool hasReflection = (lobe & Lobe_Reflection) != 0, isDirect = (lobe & Lobe_Delta) != 0, isScatter = (lobe & Lobe_NonDelta) != 0;
        if ((!isScatter && !isDirect) || !hasReflection) { beam.setZero(); return; }
        const T a = gfs.alpha(toLight, toEye, meank());
        T M = a;
        if (isScatter) {
            const V3<T> h = toLight + toEye;
            M = (T(1) - a) * std::fabs(toLight.z) * gfs.eval(meank(), h);
        }
        for (int j = 0; j < beam.hwss_comps; ++j) beam.setSP(j, (M * Spectrum<T>::rgbToSpectrum(albedo, beam.wavelengths[j])) * beam.getSP(j));
    }
    template <class S> bool sample(V3<T> wi, V3<T>& wo, T& pdf, T& weight, uint32_t& lobe, S& sg, T wavelength, uint32_t lobeMask) const {
        const bool hasReflection = (lobeMask & Lobe_Reflection) != 0, hasDirect = (lobeMask & Lobe_Delta) != 0, hasScatter = (lobeMask & Lobe_NonDelta) != 0;
        if (!hasReflection || (!hasDirect && !hasScatter)) return false;
        const T a = gfs.alpha(wi, wi, meank());
        pdf = 1;
        bool isDirect = hasDirect;
        if (hasDirect && hasScatter) { isDirect = a == T(1) || sampleNext1D<T>(sg) < a; pdf = isDirect ? a : T(1) - a; }
        weight = Spectrum<T>::rgbToSpectrum(albedo, wavelength);
        if (isDirect) { wo = V3<T>{-wi.x, -wi.y, wi.z}; lobe = Lobe_DeltaReflection; }
        else { T sp; wo = gfs.sample(wi, sg, meank(), sp); pdf *= sp; lobe = Lobe_ScatterReflection; }
        return true;
    }
    T pdf(V3<T> wi, V3<T> wo, T) const { return (T(1) - gfs.alpha(wi, wi, meank())) * gfs.pdf(wi, wo, meank()); }
    M2<T> coherenceTransform(V3<T>, V3<T> wo, uint32_t lobe) const {
        const T c = std::sqrt(hmax(T(1e-2), std::fabs(wo.z)));
        return (lobe & Lobe_Delta) ? ident2<T>() : m2<T>(c, 0, 0, T(1) / c);
    }
};

// Interference (Interference.slang:36-57)
template <class T> inline V4<T> interfere(V4<T> SP1, V4<T> SP2, T phs, T php, T coh) {
    const T S1 = (SP1.x + SP1.y) / T(2), P1 = (SP1.x - SP1.y) / T(2);
    const T S2 = (SP2.x + SP2.y) / T(2), P2 = (SP2.x - SP2.y) / T(2);
    const T S = S1 + S2 + T(2) * std::sqrt(hmax(T(0), S1 * S2)) * phs * coh;
    const T P = P1 + P2 + T(2) * std::sqrt(hmax(T(0), P1 * P2)) * php * coh;
    const T n = SP1.x + SP2.x > T(1e-6) ? (S + P) / (SP1.x + SP2.x) : T(0);
    return {S + P, S - P, (SP1.z + SP2.z) * n, (SP1.w + SP2.w) * n};
}
template <class T> inline M4<T> mul4(const M4<T>& a, const M4<T>& b) {
    M4<T> r = zero4<T>();
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k) r.m[i][j] += a.m[i][k] * b.m[k][j];
    return r;
}

// PLTCoatedOpaqueDielectricBxDF (PLTCoatedOpaqueDielectricBSDF.slang:45-246)
template <class T> struct CoatedOpaqueBxDF {
    V3<T> albedo;
    FractalSurface<T> gfs;
    T extEta, coatIOR, tau;
    static V3<T> refractCoat(V3<T> wi, T eta, T cosThetaT) {
        const T s = cosThetaT < T(0) ? T(1) / eta : eta;
        return {-s * wi.x, -s * wi.y, cosThetaT};
    }
    T coatEta() const { return hlerp(extEta, coatIOR, clampT(tau, T(0), T(.05)) / T(.05)); }
    template <int N> void eval(V3<T> toLight, V3<T> toEye, uint32_t lobe, PLTBeam<T, N>& beam) const {
        const bool isReflection = toLight.z * toEye.z > T(0);
        const bool isScatter = (lobe & Lobe_ScatterReflection) != 0, isDirect = (lobe & Lobe_DeltaReflection) != 0;
        const T coatior = coatEta();
        const T eta1 = coatior / extEta;
        T cTin;
        (void)evalFresnelDielectric(eta1, toLight.z, cTin);
        const V3<T> wiIn = -refractCoat(toLight, eta1, cTin);
        V3<T> woIn = {-wiIn.x, -wiIn.y, wiIn.z};
        if (!isDirect) { T cTout; (void)evalFresnelDielectric(eta1, toEye.z, cTout); woIn = -refractCoat(toEye, eta1, cTout); }
        if (!isReflection || (!isScatter && !isDirect) || wiIn.z <= T(0) || woIn.z <= T(0)) { beam.setZero(); return; }
        T k[N], a[N], coh[N];
        for (int j = 0; j < N; ++j) { k[j] = K<T>::_2PI / (T(1e-3) * beam.wavelengths[j]); a[j] = gfs.alpha(wiIn, woIn, k[j]); }
        const V3<T> h = wiIn + woIn;
        const V3<T> m = normalize(h);
        const T LdotH = hmax(T(0), dot(wiIn, m));
        const V3<T> Z = beam.toLocal(V3<T>{0, 0, 1});
        beam.mutualCoherence(k, Z * tau, coh);
        const M4<T> M = mul4(MuellerFresnelTDielectric(toLight.z, eta1), MuellerFresnelTDielectric(-woIn.z, eta1));
        for (int j = 0; j < beam.hwss_comps; ++j) {
            const V4<T> SP = beam.getSP(j);
            const T m00 = Spectrum<T>::rgbToSpectrum(albedo, beam.wavelengths[j]);
            const T phase = std::sin(T(2) * tau / wiIn.z * k[j] * coatior);
            V4<T> newSP = m00 * mul(M, SP), refl;
            if (isScatter) {
                const T D = gfs.eval(k[j], h);
                const T w = (T(1) - a[j]) * std::fabs(toLight.z) * D;
                newSP = w * newSP;
                refl = w * mul(MuellerFresnelRDielectric(LdotH, eta1), SP);
            } else {
                newSP = a[j] * newSP;
                refl = a[j] * mul(MuellerFresnelRDielectric(toLight.z, eta1), SP);
            }
            beam.setSP(j, interfere(newSP, refl, phase, phase, coh[j]));
        }
    }
    template <class S> bool sample(V3<T> wi, V3<T>& wo, T& pdf, T& weight, uint32_t& lobe, S& sg, T wavelength, uint32_t lobeMask) const {
        const bool hasReflection = (lobeMask & Lobe_Reflection) != 0, hasDirect = (lobeMask & Lobe_Delta) != 0, hasScatter = (lobeMask & Lobe_NonDelta) != 0;
        if (!hasReflection || (!hasDirect && !hasScatter) || wi.z <= T(0)) return false;
        const T k = K<T>::_2PI / (T(1e-3) * wavelength);
        const T eta1 = coatEta() / extEta;
        T cT;
        const T Fd = evalFresnelDielectric(eta1, wi.z, cT);
        const bool isReflCoat = Fd == T(.1f) || sampleNext1D<T>(sg) < Fd;
        weight = 1;
        pdf = isReflCoat ? Fd : T(1) - Fd;
        if (isReflCoat) { wo = V3<T>{-wi.x, -wi.y, wi.z}; lobe = Lobe_DeltaReflection; return true; }
        const V3<T> wiIn = -refractCoat(wi, eta1, cT);
        if (wiIn.z == T(0)) return false;
        weight = Spectrum<T>::rgbToSpectrum(albedo, wavelength);
        const T a = gfs.alpha(wiIn, wiIn, k);
        const T pdfDirect = hasDirect && hasScatter ? a : hasDirect ? T(1) : T(0);
        const bool isDirect = pdfDirect == T(1) || sampleNext1D<T>(sg) < pdfDirect;
        if (hasDirect && hasScatter) pdf *= isDirect ? pdfDirect : T(1) - pdfDirect;
        if (isDirect) { wo = V3<T>{-wi.x, -wi.y, wi.z}; lobe = Lobe_DeltaReflection; }
        else {
            T sp;
            const V3<T> woIn = gfs.sample(wiIn, sg, k, sp);
            pdf *= sp;
            T cT1;
            (void)evalFresnelDielectric(eta1, -woIn.z, cT1);
            wo = refractCoat(-woIn, eta1, cT1);
            lobe = Lobe_ScatterReflection;
        }
        return wi.z * wo.z > T(0);
    }
    T pdf(V3<T> wi, V3<T> wo, T wavelength) const {
        if (wi.z <= T(0) || wo.z <= T(0)) return 0;
        const T k = K<T>::_2PI / (T(1e-3) * wavelength);
        const T eta1 = coatEta() / extEta;
        T cT, cT1;
        const T Fd = evalFresnelDielectric(eta1, wi.z, cT);
        (void)evalFresnelDielectric(eta1, wo.z, cT1);
        const V3<T> wiIn = -refractCoat(wi, eta1, cT), woIn = -refractCoat(wo, eta1, cT1);
        const T a = gfs.alpha(wiIn, wiIn, k);
        return wiIn.z * woIn.z == T(0) ? T(0) : (T(1) - a) * (T(1) - Fd) * gfs.pdf(wiIn, woIn, k);
    }
    template <int N> void lambdaPdf(V3<T> toLight, V3<T> toEye, uint32_t lobe, const PLTBeam<T, N>& beam, T* ret) const {
        for (int j = 0; j < N; ++j) ret[j] = 1;
        if (!(lobe & Lobe_NonDelta)) return;
        const T eta1 = coatEta() / extEta;
        T cT, cT1;
        (void)evalFresnelDielectric(eta1, toEye.z, cT);
        (void)evalFresnelDielectric(eta1, toLight.z, cT1);
        const V3<T> wiIn = -refractCoat(toEye, eta1, cT), woIn = -refractCoat(toLight, eta1, cT1);
        for (int j = 0; j < N; ++j) ret[j] = gfs.pdf(wiIn, woIn, K<T>::_2PI / (T(1e-3) * beam.wavelengths[j]));
    }
    M2<T> coherenceTransform(V3<T>, V3<T> wo, uint32_t lobe) const {
        const T c = std::sqrt(hmax(T(1e-2), std::fabs(wo.z)));
        return (lobe & Lobe_Delta) ? ident2<T>() : m2<T>(c, 0, 0, T(1) / c);
    }
};

// BesselJ (Utils/Math/MathHelpers.slang:813-876)
template <class T> inline T BesselJ(int nu, T x) {
    static const uint32_t table[13] = {1, 1, 2, 6, 24, 120, 720, 5040, 40320, 362880, 3628800, 39916800, 479001600};
    auto fact = [&](int n) { return n < 13 ? table[n] : 0u; };
    const T t = nu > 0 || (-nu) % 2 == 0 ? T(1) : T(-1);
    nu = std::abs(nu);
    const T cutoff = T(nu) + T(1), interp = cutoff / T(10);
    const T laa = std::sqrt(T(2) / K<T>::PI / x) * std::cos(x - T(nu) * K<T>::PI / T(2) - K<T>::PI / T(4));
    T saa = std::numeric_limits<T>::quiet_NaN();
    if (nu < 7) {
        const T x2 = sqr(x), x4 = sqr(x2), x6 = x2 * x4, x8 = sqr(x4);
        const T s = T(6) * std::pow(x / T(2), T(nu));
        const T a = T(1) / (T(6) * T(fact(nu)));
        const T b = -x2 / (T(24) * T(fact(nu + 1)));
        const T c1 = T(11) * x8 - T(864) * T(6 + nu) * (x6 - T(12) * (T(5) + T(nu)) * (T(3) * x4 + T(32) * T(4 + nu) * (T(-2) * x2 + T(27) * T(3 + nu))));
        const T c = c1 / T(fact(6 + nu)) * x4 * T(5.652695401144601e-10);
        saa = s * (a + b + c);
    }
    return nu >= 7 || std::fabs(x) >= cutoff + interp ? laa * t
                                                        : hlerp(saa, laa, hmax(T(0), hmin(T(1), T(.5) * (x - cutoff - interp) / interp))) * t;
}

// DiffractionGrating (Rendering/Materials/PLT/DiffractionGrating.slang:41-186), 2-D vectors as pairs.
template <class T> struct DiffractionGrating {
    T gx, gy, ipx, ipy, q;
    uint32_t lobeCount, type;
    T mult;
    bool is1D() const { return ipy == T(0); }
    T alpha(V3<T> wi, T k) const { return std::exp(-sqr(wi.z * q * k)); }
    T lobeIntensity(int lx, int ly, V3<T> wi, T lambda) const {
        const uint32_t gt = type & 0xFu;
        const T a = K<T>::_4PI * q / (lambda * std::fabs(wi.z));
        T ix, iy;
        if (gt == 0) { iy = ix = lx == 0 ? T(1) : BesselJ<T>(lx, a); if (!is1D()) iy = ly == 0 ? T(1) : BesselJ<T>(ly, a); }
        else if (gt == 1) {
            auto sinc = [](T v) { return std::sin(v) / v; };
            iy = ix = lx == 0 ? T(1) : std::sin(a / T(2)) * sinc(K<T>::PI * T(lx) / T(2));
            if (!is1D()) iy = ly == 0 ? T(1) : std::sin(a / T(2)) * sinc(K<T>::PI * T(ly) / T(2));
        } else { iy = ix = lx == 0 ? T(1) : T(1) / std::sqrt(std::fabs(T(lx))); if (!is1D()) iy = ly == 0 ? T(1) : T(1) / std::sqrt(std::fabs(T(ly))); }
        return mult * ix * iy;
    }
    V3<T> diffract(V3<T> wi, int lx, int ly, T lambda) const {
        const T px = std::sqrt(sqr(wi.x) + sqr(wi.z)), py = std::sqrt(sqr(wi.y) + sqr(wi.z));
        const T sx = px > K<T>::FLT_EPS ? wi.x / px : T(0), sy = py > K<T>::FLT_EPS ? wi.y / py : T(0);
        const T a = lambda * T(lx) * ipx - sx, b = lambda * T(ly) * ipy - sy;
        const T m = (sqr(a) - T(1)) / (sqr(a * b) - T(1));
        const T qq = T(1) - sqr(b) * m;
        return (std::fabs(a) <= T(1) && std::fabs(b) <= T(1))
                   ? V3<T>{a * std::sqrt(hmax(T(0), qq)), b * std::sqrt(m), std::sqrt(hmax(T(0), T(1) - sqr(a) * qq - sqr(b) * m))}
                   : V3<T>{0, 0, 0};
    }
    template <class S> void sampleLobe(V3<T> wi, T& pdfx, T& pdfy, int& lx, int& ly, T lambda, S& sg) const {
        T intensity[16];
        T total = 0;
        for (uint32_t l = 0; l < lobeCount; ++l) { const T li = lobeIntensity(int(l) + 1, 0, wi, lambda); total += li; intensity[l] = li; }
        const V2<T> u = sampleNext2D<T>(sg);
        const T rx = (u.x - T(.5)) * T(2), ry = (u.y - T(.5)) * T(2);
        T cdf = 0;
        lx = ly = 0;
        pdfx = pdfy = 0;
        for (uint32_t l = 0; l < lobeCount; ++l) {
            const T p = intensity[l] / total;
            cdf += p;
            const bool sx = std::fabs(rx) < cdf && lx == 0, sy = std::fabs(ry) < cdf && ly == 0;
            if (sx) { pdfx = p; lx = int(l) + 1; }
            if (sy) { pdfy = p; ly = int(l) + 1; }
        }
        pdfx /= T(2); pdfy /= T(2);
        lx = rx < T(0) ? -lx : lx;
        ly = ry < T(0) ? -ly : ly;
        if (is1D()) { ly = 0; pdfy = 1; }
    }
};

// PLTDiffractionGratedConductorBxDF (PLTDiffractionGratedConductorBSDF.slang:45-213)
template <class T> struct GratedConductorBxDF {
    ConductorBxDF<T> base;
    DiffractionGrating<T> grating;
    static void decode(uint32_t bsdfLobe, int& lx, int& ly) {
        const int x = int(bsdfLobe) >> 8;
        lx = x & 0x7; ly = (x >> 3) & 0x7;
        lx *= 2 * ((x >> 6) & 0x1) - 1;
        ly *= 2 * ((x >> 7) & 0x1) - 1;
    }
    static uint32_t encode(int lx, int ly, uint32_t bsdfLobe) {
        const int sx = lx < 0 ? 0 : 1, sy = ly < 0 ? 0 : 1;
        lx = std::min(std::abs(lx), 7); ly = std::min(std::abs(ly), 7);
        bsdfLobe &= ~0xff00u;
        bsdfLobe += uint32_t((ly << 3) + lx + (sy << 7) + (sx << 6)) << 8;
        return bsdfLobe;
    }
    template <int N> void eval(V3<T> toLight, V3<T> toEye, uint32_t bsdfLobe, PLTBeam<T, N>& beam) const {
        const bool isDirect = (bsdfLobe & Lobe_Delta) != 0;
        T lam[N], a[N];
        bool anyPos = false, anyBelow = false;
        for (int j = 0; j < N; ++j) {
            lam[j] = T(1e-3) * beam.wavelengths[j];
            a[j] = grating.alpha(toLight, K<T>::_2PI / lam[j]);
            anyPos |= a[j] > T(0); anyBelow |= a[j] < T(1);
        }
        int lx, ly;
        decode(bsdfLobe, lx, ly);
        const V3<T> toLightBeam = beam.toLocal(toLight);
        if (lx == 0 && ly == 0 && anyPos) {
            base.eval(toLight, toEye, bsdfLobe, beam);
            beam.scaleEach(a);
        } else if ((lx != 0 || ly != 0) && anyBelow && isDirect) {
            for (int j = 0; j < beam.hwss_comps; ++j) {
                const T m00 = Spectrum<T>::rgbToSpectrum(base.specular, beam.wavelengths[j]);
                const V2<T> ior = {evalSpectrum<T>(base.etaN, beam.wavelengths[j]) / base.extEta, evalSpectrum<T>(base.etaK, beam.wavelengths[j]) / base.extEta};
                const M4<T> M0 = MuellerFresnelRConductor(toLight.z, ior);
                const V3<T> lobeWi = beam.toLocal(grating.diffract(toEye, lx, ly, lam[j]));
                const T coh = j == 0 ? T(1) : beam.mutualCoherenceAngular(lobeWi, toLightBeam);
                const T inten = grating.lobeIntensity(lx, ly, toLight, lam[j]);
                beam.setSP(j, (m00 * (T(1) - a[j]) * coh * inten) * mul(M0, beam.getSP(j)));
            }
        } else beam.setZero();
    }
    template <class S> bool sample(V3<T> wi, V3<T>& wo, T& pdf, T& weight, uint32_t& bsdfLobe, S& sg, T wavelength, uint32_t lobeMask) const {
        const bool hasReflection = (lobeMask & Lobe_Reflection) != 0, hasDirect = (lobeMask & Lobe_Delta) != 0, hasScatter = (lobeMask & Lobe_NonDelta) != 0;
        if (!hasReflection || (!hasDirect && !hasScatter)) return false;
        const T lam = T(1e-3) * wavelength;
        const T a = grating.alpha(wi, K<T>::_2PI / lam);
        bool isDirect = hasDirect;
        if (hasDirect && hasScatter) { isDirect = a == T(1) || sampleNext1D<T>(sg) < a; pdf = isDirect ? a : T(1) - a; }
        if (isDirect) {
            T basePdf;
            base.sample(wi, wo, basePdf, weight, bsdfLobe, sg, wavelength, lobeMask);
            pdf *= basePdf;
            return true;
        }
        weight = Spectrum<T>::rgbToSpectrum(base.specular, wavelength);
        T px, py;
        int lx, ly;
        grating.sampleLobe(wi, px, py, lx, ly, lam, sg);
        wo = grating.diffract(wi, lx, ly, lam);
        bsdfLobe = encode(lx, ly, Lobe_DispersiveDeltaReflection);
        pdf *= px * py;
        return true;
    }
    T pdf(V3<T> wi, V3<T> wo, T wavelength) const {
        const T k = K<T>::_2PI / (T(1e-3) * wavelength);
        return base.pdf(wi, wo, wavelength) * grating.alpha(wi, k);
    }
    template <int N> void lambdaPdf(V3<T> toLight, V3<T> toEye, uint32_t bsdfLobe, const PLTBeam<T, N>& beam, T* ret) const {
        for (int j = 0; j < N; ++j) ret[j] = 1;
        const bool isDirect = (bsdfLobe & Lobe_Delta) != 0;
        T lam[N];
        bool anyBelow = false;
        for (int j = 0; j < N; ++j) { lam[j] = T(1e-3) * beam.wavelengths[j]; anyBelow |= grating.alpha(toLight, K<T>::_2PI / lam[j]) < T(1); }
        int lx, ly;
        decode(bsdfLobe, lx, ly);
        const V3<T> toLightBeam = beam.toLocal(toLight);
        if ((lx != 0 || ly != 0) && anyBelow && isDirect)
            for (int j = 0; j < N; ++j) ret[j] = j == 0 ? T(1) : beam.mutualCoherenceAngular(beam.toLocal(grating.diffract(toEye, lx, ly, lam[j])), toLightBeam);
    }
    M2<T> coherenceTransform(V3<T> wi, V3<T> wo, uint32_t bsdfLobe) const {
        int lx, ly;
        decode(bsdfLobe, lx, ly);
        return (lx == 0 && ly == 0) ? base.coherenceTransform(wi, wo, bsdfLobe) : ident2<T>();
    }
};

// Material instance (the IMaterialInstance of each type), built as setupMaterialInstance does.
template <class T> struct Instance {
    uint32_t type = 0;
    DielectricBxDF<T> die;
    ConductorBxDF<T> con;
    DiffuseBxDF<T> dif;
    OpaqueDielectricBxDF<T> opq;
    CoatedOpaqueBxDF<T> cot;
    GratedConductorBxDF<T> grt;
    uint32_t emission = kNoProfile;
};
template <class T> inline Instance<T> setupMaterialInstance(const MatData& md, const SpectralProfile* prof, uint32_t nprof) {
    Instance<T> in;
    in.type = md.type;
    const V3<T> base = {T(hf(md.base_color[0])), T(hf(md.base_color[1])), T(hf(md.base_color[2]))};
    const float sx = hf(md.specular[0]), sy = hf(md.specular[1]), sz = hf(md.specular[2]), sw = hf(md.specular[3]);
    const SpectralProfile* pn = md.ior_n < nprof ? prof + md.ior_n : nullptr;
    const SpectralProfile* pk = md.ior_k < nprof ? prof + md.ior_k : nullptr;
    switch (md.type) {
        case MT_DIELECTRIC: {  // PLTDielectricMaterial.slang:40-56
            const V3<T> tran = {T(hf(md.transmission[0])), T(hf(md.transmission[1])), T(hf(md.transmission[2]))};
            const float B = (tran.x > 0 || tran.y > 0 || tran.z > 0) ? sx : 0.f;
            in.die.specular = base;
            in.die.transmission = tran;
            in.die.gfs = GaussianSurface<T>::makeRef(T(sy));
            in.die.etaInt = T(std::fmax(1.f, hf(md.ior) - 2.897f * B));
            in.die.etaExt = T(sw);
            in.die.dispersionB = T(B);
            break;
        }
        case MT_CONDUCTOR:  // PLTConductorMaterial.slang:40-57
            in.con.specular = base;
            in.con.gfs = FractalSurface<T>::makeRef(T(sy), T(sx));
            in.con.etaN = pn; in.con.etaK = pk;
            in.con.extEta = T(hf(md.ior));
            break;
        case MT_DIFFUSE:
            in.dif.albedo = base;
            in.emission = md.emission;
            break;
        case MT_OPAQUE:
            in.opq.albedo = base;
            in.opq.gfs = FractalSurface<T>::makeRef(T(sy));
            break;
        case MT_COATED_OPAQUE:  // PLTCoatedOpaqueDielectricMaterial.slang:40-60
            in.cot.albedo = base;
            in.cot.gfs = FractalSurface<T>::makeRef(T(sy));
            in.cot.extEta = T(hf(md.ior));
            in.cot.tau = T(hf(md.data1[0]) * sz);
            in.cot.coatIOR = T(hf(md.data1[3]));
            break;
        case MT_GRATED_CONDUCTOR: {  // PLTDiffractionGratedConductorMaterial.slang:40-70
            in.grt.base.specular = base;
            in.grt.base.gfs = FractalSurface<T>::makeRef(T(hf(md.data2[1])));
            in.grt.base.etaN = pn; in.grt.base.etaK = pk;
            in.grt.base.extEta = T(hf(md.ior));
            const float px = hf(md.data1[0]), py = hf(md.data1[1]);
            in.grt.grating.ipx = T(px > 0.f ? 1.f / px : 0.f);
            in.grt.grating.ipy = T(py > 0.f ? 1.f / py : 0.f);
            in.grt.grating.q = T(hf(md.data2[0]) * sx);
            in.grt.grating.mult = T(sy);
            const uint32_t tl = md.data1[3];
            in.grt.grating.lobeCount = std::min(std::max(tl & 0xFFu, 1u), 8u);
            in.grt.grating.type = tl >> 8;
            const T dir = T(hf(md.data1[2]));
            in.grt.grating.gx = std::cos(dir);
            in.grt.grating.gy = std::sin(dir);
            break;
        }
    }
    return in;
}

template <class T> inline bool isEmissive(const Instance<T>& in) { return in.type == MT_DIFFUSE && in.emission != kNoProfile; }
template <class T> inline uint32_t getLobes(const Instance<T>& in) {
    switch (in.type) {
        case MT_DIELECTRIC: {
            const uint32_t dt = in.die.isDispersive() ? Lobe_DispersiveDeltaTransmission : Lobe_DeltaTransmission;
            return in.die.gfs.isDeltaOnly() ? (Lobe_DeltaReflection | dt) : (Lobe_DeltaReflection | dt | Lobe_ScatterReflection | Lobe_ScatterTransmission);
        }
        case MT_CONDUCTOR: return in.con.gfs.isDeltaOnly() ? Lobe_DeltaReflection : (Lobe_DeltaReflection | Lobe_ScatterReflection);
        case MT_DIFFUSE: return Lobe_DiffuseReflection;
        case MT_OPAQUE: return in.opq.gfs.isDeltaOnly() ? Lobe_DeltaReflection : (Lobe_DeltaReflection | Lobe_ScatterReflection);
        case MT_COATED_OPAQUE: return in.cot.gfs.isDeltaOnly() ? Lobe_DeltaReflection : (Lobe_DeltaReflection | Lobe_ScatterReflection);
        default:
            return (in.grt.base.gfs.isDeltaOnly() ? Lobe_DeltaReflection : (Lobe_DeltaReflection | Lobe_ScatterReflection)) | Lobe_DispersiveDeltaReflection;
    }
}
enum : uint32_t { Prop_Transmissive = 1, Prop_Emissive = 2, Prop_MWR = 4, Prop_MWT = 8 };
template <class T> inline uint32_t getPropertyFlags(const Instance<T>& in) {
    const T cut = T(.15f);
    switch (in.type) {
        case MT_DIELECTRIC: return Prop_Transmissive | (in.die.gfs.q < cut ? Prop_MWT : 0u);
        case MT_CONDUCTOR: return in.con.gfs.q < cut ? Prop_MWR : 0u;
        case MT_DIFFUSE: return isEmissive(in) ? Prop_Emissive : 0u;
        case MT_OPAQUE: return in.opq.gfs.q < cut ? Prop_MWR : 0u;
        case MT_COATED_OPAQUE: return in.cot.gfs.q < cut ? Prop_MWR : 0u;
        default: return Prop_MWR;
    }
}

// Shading data (Scene/ShadingData.slang): frame + position + face normal.
template <class T> struct SD {
    Frame<T> f;
    V3<T> posW, faceN, V;
    bool frontFacing;
    uint32_t materialID;
};
// The grating evaluates in a frame aligned with its direction (DiffractionGrating.slang:72-79).
template <class T> inline Frame<T> bsdfFrame(const Instance<T>& in, const Frame<T>& sd) {
    if (in.type != MT_GRATED_CONDUCTOR) return sd;
    Frame<T> g;
    g.T_ = sd.fromLocal(V3<T>{in.grt.grating.gx, in.grt.grating.gy, 0});
    g.N = sd.N;
    g.B = cross(g.N, g.T_);
    return g;
}

template <class T, int N> inline void materialEval(const Instance<T>& in, const SD<T>& sd, V3<T> toEye, uint32_t lobe, PLTBeam<T, N>& beam) {
    const Frame<T> f = bsdfFrame(in, sd.f);
    switch (in.type) {
        case MT_DIELECTRIC: bsdfEval<T, N>(in.die, f, sd.V, toEye, lobe, beam); break;
        case MT_CONDUCTOR: bsdfEval<T, N>(in.con, f, sd.V, toEye, lobe, beam); break;
        case MT_DIFFUSE: bsdfEval<T, N>(in.dif, f, sd.V, toEye, lobe, beam); break;
        case MT_OPAQUE: bsdfEval<T, N>(in.opq, f, sd.V, toEye, lobe, beam); break;
        case MT_COATED_OPAQUE: bsdfEval<T, N>(in.cot, f, sd.V, toEye, lobe, beam); break;
        default: bsdfEval<T, N>(in.grt, f, sd.V, toEye, lobe, beam); break;
    }
}
template <class T, int N> inline void materialLambdaPdf(const Instance<T>& in, const SD<T>& sd, V3<T> wo, uint32_t lobe, PLTBeam<T, N> beam, T* ret) {
    for (int j = 0; j < N; ++j) ret[j] = 1;
    beam.r = sd.f.toLocal(beam.r);
    beam.t = sd.f.toLocal(beam.t);
    const V3<T> a = sd.f.toLocal(sd.V), b = sd.f.toLocal(wo);
    switch (in.type) {
        case MT_DIELECTRIC: in.die.lambdaPdf(a, b, lobe, beam, ret); break;
        case MT_CONDUCTOR: in.con.lambdaPdf(a, b, lobe, beam, ret); break;
        case MT_COATED_OPAQUE: in.cot.lambdaPdf(a, b, lobe, beam, ret); break;
        case MT_GRATED_CONDUCTOR: in.grt.lambdaPdf(a, b, lobe, beam, ret); break;
        default: break;
    }
}
template <class T> inline T materialPdf(const Instance<T>& in, const SD<T>& sd, V3<T> toLight, T wavelength) {
    const Frame<T> f = bsdfFrame(in, sd.f);
    const V3<T> a = f.toLocal(toLight), b = f.toLocal(sd.V);
    switch (in.type) {
        case MT_DIELECTRIC: return in.die.pdf(a, b, wavelength);
        case MT_CONDUCTOR: return in.con.pdf(a, b, wavelength);
        case MT_DIFFUSE: return in.dif.pdf(a, b, wavelength);
        case MT_OPAQUE: return in.opq.pdf(a, b, wavelength);
        case MT_COATED_OPAQUE: return in.cot.pdf(a, b, wavelength);
        default: return in.grt.pdf(a, b, wavelength);
    }
}
struct BSDFSample { V3<float> wo{0, 0, 0}; float pdf = 0, weight = 0; uint32_t lobe = 0; };
template <class S> inline bool materialSample(const Instance<float>& in, const SD<float>& sd, S& sg, BSDFSample& r, float wavelength) {
    const Frame<float> f = bsdfFrame(in, sd.f);
    const V3<float> wi = f.toLocal(sd.V);
    bool valid = false;
    switch (in.type) {
        case MT_DIELECTRIC: valid = in.die.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, Lobe_All); r.wo = f.fromLocal(r.wo); return valid;
        case MT_CONDUCTOR: valid = in.con.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, Lobe_All); break;
        case MT_DIFFUSE: valid = in.dif.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, Lobe_All); r.wo = f.fromLocal(r.wo); return valid;
        case MT_OPAQUE: valid = in.opq.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, Lobe_All); break;
        case MT_COATED_OPAQUE: valid = in.cot.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, Lobe_All); break;
        default: valid = in.grt.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, Lobe_All); break;
    }
    if (valid) r.wo = f.fromLocal(r.wo);
    return valid;
}

}  // namespace pltoracle
