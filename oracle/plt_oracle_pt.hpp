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
// MaterialType (Scene/Material/MaterialTypes.slang:36-53)
enum : uint32_t { MT_DIFFUSE = 2, MT_CONDUCTOR = 3, MT_DIELECTRIC = 4, MT_OPAQUE = 5, MT_THIN_DIELECTRIC = 6, MT_COATED_CONDUCTOR = 7,
                  MT_COATED_OPAQUE = 8, MT_GRATED_CONDUCTOR = 9, MT_GRATED_OPAQUE = 10, MT_MULTILAYER = 11 };
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

// PLTCoatedConductorBxDF (PLTCoatedConductorBSDF.slang:45-217): a thin dielectric coat over a
// K-correlation conductor with tabulated n,k. sample / pdf / lambdaPdf / coherenceTransform are
// those of the coated opaque dielectric (the `albedo` member is the specular colour here).
template <class T> struct CoatedConductorBxDF : CoatedOpaqueBxDF<T> {
    const SpectralProfile* etaN = nullptr;
    const SpectralProfile* etaK = nullptr;
    template <int N> void eval(V3<T> toLight, V3<T> toEye, uint32_t lobe, PLTBeam<T, N>& beam) const {  // :72-128
        using B = CoatedOpaqueBxDF<T>;
        const bool isReflection = toLight.z * toEye.z > T(0);
        const bool isScatter = (lobe & Lobe_ScatterReflection) != 0, isDirect = (lobe & Lobe_DeltaReflection) != 0;
        const T coatior = this->coatEta();
        const T eta1 = coatior / this->extEta;
        T cTin;
        (void)evalFresnelDielectric(eta1, toLight.z, cTin);
        const V3<T> wiIn = -B::refractCoat(toLight, eta1, cTin);
        V3<T> woIn = {-wiIn.x, -wiIn.y, wiIn.z};
        if (!isDirect) { T cTout; (void)evalFresnelDielectric(eta1, toEye.z, cTout); woIn = -B::refractCoat(toEye, eta1, cTout); }
        if (!isReflection || (!isScatter && !isDirect) || wiIn.z <= T(0) || woIn.z <= T(0)) { beam.setZero(); return; }
        T k[N], a[N], coh[N], iorN[N], iorK[N];
        for (int j = 0; j < N; ++j) {
            k[j] = K<T>::_2PI / (T(1e-3) * beam.wavelengths[j]);
            a[j] = this->gfs.alpha(wiIn, woIn, k[j]);
            iorN[j] = evalSpectrum<T>(etaN, beam.wavelengths[j]);
            iorK[j] = evalSpectrum<T>(etaK, beam.wavelengths[j]);
        }
        const V3<T> h = wiIn + woIn;
        const V3<T> m = normalize(h);
        const T LdotH = hmax(T(0), dot(wiIn, m));
        const V3<T> Z = beam.toLocal(V3<T>{0, 0, 1});
        beam.mutualCoherence(k, Z * this->tau, coh);
        const M4<T> MiR = MuellerFresnelRDielectric(toLight.z, eta1);
        const M4<T> MiT = MuellerFresnelTDielectric(toLight.z, eta1);
        const M4<T> MoT = MuellerFresnelTDielectric(-woIn.z, eta1);
        for (int j = 0; j < beam.hwss_comps; ++j) {
            V2<T> rs, rp;
            const V2<T> ior = {iorN[j] / coatior, iorK[j] / coatior};
            fresnel_conductor(LdotH, ior, rs, rp);
            const M4<T> M = mul4(mul4(MiT, MuellerFresnel(rs, rp)), MoT);
            const V4<T> SP = beam.getSP(j);
            const T m00 = Spectrum<T>::rgbToSpectrum(this->albedo, beam.wavelengths[j]);
            V4<T> newSP = m00 * mul(M, SP);
            if (isScatter) {
                const T D = this->gfs.eval(k[j], h);
                newSP = ((T(1) - a[j]) * std::fabs(toLight.z) * D) * newSP;
            } else {
                // phaseShift(eta, dist, k, r) = sin(dist k eta + arg r) (Interference.slang:36-39)
                const T dist = T(2) * this->tau / wiIn.z;
                const T pS = std::sin(dist * k[j] * coatior + std::atan2(rs.y, rs.x));
                const T pP = std::sin(dist * k[j] * coatior + std::atan2(rp.y, rp.x));
                newSP = interfere(a[j] * newSP, mul(MiR, SP), pS, pP, coh[j]);
            }
            beam.setSP(j, newSP);
        }
    }
};

// Birefringence (Rendering/Materials/PLT/Birefringence.slang:33-197): normal modes, fields, Poynting
// vectors and Fresnel coefficients of an isotropic / uniaxial interface (y up, incidence in yz).
template <class T> struct Biref {
    static T qo(T eo, T Kt) { return std::sqrt(sqr(eo) - sqr(Kt)); }
    static V2<T> qe(T eo, T ee, V3<T> A, T Kt) {
        const T e = sqr(ee) - sqr(eo);
        const T a = A.x, b = A.y, c = A.z;
        const T G = sqr(b) * e + sqr(eo);
        const T r = std::sqrt(hmax(T(0), sqr(Kt) * (sqr(a) * e - sqr(ee)) + sqr(ee) * G));
        return {(eo * r + c * b * Kt * e) / G, (-eo * r + c * b * Kt * e) / G};
    }
    static V3<T> E_e(T eo, V3<T> A, T Kt, T q) {
        return normalize(V3<T>{A.x * sqr(eo), Kt * A.z * q + A.y * (sqr(eo) - sqr(q)), Kt * A.y * q + A.z * (sqr(eo) - sqr(Kt))});
    }
    static V3<T> E_o(V3<T> A, T Kt, T q) { return normalize(V3<T>{-A.y * Kt - A.z * q, A.x * Kt, A.x * q}); }
    static V3<T> Poynting(T q, V3<T> E, T Kt) {
        return normalize(V3<T>{E.x * (Kt * E.z - q * E.y), Kt * E.z * E.y + q * (sqr(E.x) + sqr(E.z)),
                               -q * E.z * E.y - Kt * (sqr(E.x) + sqr(E.y))});
    }
    static void fresnel_iso_aniso(T cos_theta, T sin_theta, T ei, T eo, T ee, V3<T> A, T& rss, T& rsp, T& tso, T& tse,
                                  T& rps, T& rpp, T& tpo, T& tpe) {
        const T Kt = ei * sin_theta;
        const T qi = -qo(ei, Kt);
        const T q2om = -qo(eo, Kt);
        const T q2em = qe(eo, ee, A, Kt).y;
        const V3<T> E2om = E_o(A, Kt, q2om);
        const V3<T> E2em = E_e(eo, A, Kt, q2em);
        const T A_1 = -E2em.x * E2om.y * Kt * cos_theta;
        const T A_2 = -E2em.y * E2om.x * Kt * cos_theta;
        const T C_1 = qi * cos_theta - Kt * sin_theta;
        const T C_2 = E2em.x * E2om.z * (C_1 + q2om * cos_theta) - A_1;
        const T C_3 = E2em.z * E2om.x * (C_1 + q2em * cos_theta) - A_2;
        const T Ns = C_2 * (qi + q2em) - C_3 * (qi + q2om);
        rps = (T(2) * E2em.x * E2om.x * cos_theta * C_1 * (q2om - q2em)) / Ns;
        rpp = (-E2em.x * E2om.z * (C_1 - q2om * cos_theta) * (qi + q2em) + E2em.z * E2om.x * (C_1 - q2em * cos_theta) * (qi + q2om) -
               A_1 * (q2em + qi) + A_2 * (q2om + qi)) / Ns;
        tpo = -(T(2) * E2em.x * cos_theta * (q2em + qi) * C_1) / Ns;
        tpe = (T(2) * E2om.x * cos_theta * (q2om + qi) * C_1) / Ns;
        rss = (C_2 * (qi - q2em) + C_3 * (q2om - qi)) / Ns;
        rsp = (T(2) * qi * (E2em.z * E2om.y * Kt - E2em.y * E2om.z * Kt + E2em.z * E2om.z * (q2om - q2em))) / Ns;
        tso = -(T(2) * qi * (E2em.y * Kt * cos_theta + E2em.z * C_1 + E2em.z * q2em * cos_theta)) / Ns;
        tse = (T(2) * qi * (E2om.y * Kt * cos_theta + E2om.z * C_1 + E2om.z * q2om * cos_theta)) / Ns;
    }
    static void fresnel_aniso_iso(T cos_phi, T sin_phi, T ei, T eo, T ee, V3<T> A, T& roo, T& roe, T& tos, T& top, T& reo,
                                  T& ree, T& tes, T& tep) {
        const T Kt = ei * sin_phi;
        const T q2om = -qo(ei, Kt);
        const T qop = qo(eo, Kt);
        const V2<T> q2 = qe(eo, ee, A, Kt);
        const T qep = q2.x, qem = q2.y;
        const V3<T> Eop = E_o(A, Kt, qop), Eom = E_o(A, Kt, -qop);
        const V3<T> Eep = E_e(eo, A, Kt, qep), Eem = E_e(eo, A, Kt, qem);
        const T B_1 = -Eep.x * Kt * cos_phi * (q2om - qep);
        const T B_2 = -Eom.x * Kt * cos_phi * (q2om + qop);
        const T B_3 = -Eop.x * Kt * cos_phi * (qop - q2om);
        const T D_1 = Kt * sin_phi - q2om * cos_phi;
        const T D_2 = Eep.z * (D_1 + qep * cos_phi);
        const T D_3 = Eem.z * (D_1 + qem * cos_phi);
        const T D_4 = Eop.z * (D_1 + qop * cos_phi);
        const T N2 = Eep.x * (qep - q2om) * D_4 + Eop.x * (q2om - qop) * D_2 + Eop.y * B_1 + Eep.y * B_3;
        roo = -(Eep.x * Eom.z * (q2om - qep) * (qop * cos_phi - D_1) + Eom.x * (q2om + qop) * D_2 + Eom.y * B_1 - Eep.y * B_2) / N2;
        roe = -(-Eom.x * Eop.z * (qop * (qop * cos_phi + Kt * sin_phi) + q2om * D_1) +
                Eom.z * Eop.x * (qop * (qop * cos_phi - Kt * sin_phi) + q2om * D_1) + Eop.y * B_2 + Eom.y * B_3) / N2;
        tos = (Eep.x * qep * roe - Eom.x * qop + Eop.x * qop * roo) / q2om;
        top = -(Eom.z * qop - Eom.y * Kt - (Eep.y * Kt + Eep.z * qep) * roe - (Eop.y * Kt + Eop.z * qop) * roo) / D_1;
        reo = (Eem.x * (qem - q2om) * (D_2 + Eep.y * Kt * cos_phi) + Eep.x * (q2om - qep) * (D_3 + Eem.y * Kt * cos_phi)) / N2;
        ree = -(Eem.x * (qem - q2om) * D_4 - Eem.x * Eop.y * Kt * cos_phi * (q2om - qem) + Eem.y * B_3 + Eop.x * (q2om - qop) * D_3) / N2;
        tes = (Eem.x * qem + Eep.x * qep * ree + Eop.x * qop * reo) / q2om;
        tep = -(Eem.z + Eep.z * ree + Eop.z * reo) / cos_phi;
    }
    // The ordinary and extraordinary ray directions in the slab and the effective indices (:168-195).
    static void vectors_in_slab(V3<T> I, T ei, T eo, T ee, V3<T> A, V3<T>& Io, V3<T>& Ie, T& e_eff) {
        const T Kt = ei * I.z;
        const T qom = -qo(eo, Kt);
        const T qem = qe(eo, ee, A, Kt).y;
        const V3<T> Eom = E_o(A, Kt, qom);
        const V3<T> Eem = E_e(eo, A, Kt, qem);
        Io = Poynting(qom, Eom, Kt);
        Ie = Poynting(qem, Eem, Kt);
        const V3<T> We = normalize(V3<T>{0, qem, -Kt});
        e_eff = Kt / std::sqrt(hmax(T(0), T(1) - sqr(We.y)));
    }
};

// PLTThinDielectricBxDF (PLTThinDielectricBSDF.slang:43-200): a slab in air with the incoherent
// sum of internal reflections T(1-R)^-1 T, and optional uniaxial birefringence in transmission.
template <class T> struct ThinDielectricBxDF {
    V2<T> opticAxisXY{T(.15), T(.04)};
    T etaAvg, tau, birefringence;
    V3<T> opticAxis() const {
        return {opticAxisXY.x, opticAxisXY.y, std::sqrt(hmax(T(0), T(1) - dot(opticAxisXY, opticAxisXY)))};
    }
    template <int N> void handle_birefringence(PLTBeam<T, N>& beam, V3<T> wi, const T* k) const {  // :60-118
        const V3<T> OA = opticAxis();
        const T phii = std::atan2(wi.y, wi.x);
        const T phiA = std::atan2(OA.y, OA.x) - phii;
        const V3<T> I = {0, std::fabs(wi.z), std::sqrt(T(1) - sqr(wi.z))};
        const T s = std::sqrt(T(1) - sqr(OA.z));
        V3<T> A = {std::cos(phiA) * s, 0, std::sin(phiA) * s};
        A.y = sgn(wi.z) * OA.z;
        const V3<T> A2 = {A.x, -A.y, A.z};
        const V3<T> Z = beam.toLocal(V3<T>{std::cos(phii), std::sin(phii), 0});
        const T ei = 1, eo = etaAvg - T(.5) * birefringence, ee = etaAvg + T(.5) * birefringence;
        const T tauum = tau * T(1e+3);
        V3<T> Io, Ie;
        T e_eff;
        Biref<T>::vectors_in_slab(I, ei, eo, ee, A, Io, Ie, e_eff);
        T rss, rsp, tso, tse, rps, rpp, tpo, tpe, roo, roe, tos, top, reo, ree, tes, tep;
        Biref<T>::fresnel_iso_aniso(I.y, I.z, ei, eo, ee, A, rss, rsp, tso, tse, rps, rpp, tpo, tpe);
        Biref<T>::fresnel_aniso_iso(I.y, I.z, ei, eo, ee, A2, roo, roe, tos, top, reo, ree, tes, tep);
        const T ss = tso * tos + tse * tes, sp = tso * top + tse * tep;
        const T ps = tpo * tos + tpe * tes, pp = tpo * top + tpe * tep;
        const T Ioz = std::fabs(tauum / Io.y) * Io.z, Iez = std::fabs(tauum / Ie.y) * Ie.z;
        const T OPLo = std::fabs(tauum / Io.y) * eo, OPLe = std::fabs(tauum / Ie.y) * e_eff;
        const T oez = Ioz - Iez;
        T coh[N], nLx[N], nLy[N], newL[N];
        beam.mutualCoherence(k, Z * oez, coh);
        bool allPos = true, allFinite = true;
        for (int j = 0; j < N; ++j) {
            const T t = coh[j] * T(2) * std::sin(-k[j] * (ei * oez * I.z + OPLo - OPLe));
            const T Lx = T(.5) * (beam.L[j] + beam.L1[j]), Ly = T(.5) * (beam.L[j] - beam.L1[j]);
            const T sq = std::sqrt(Lx * Ly);
            nLx[j] = std::fmax(T(0), sqr(ss) * Lx + sqr(ps) * Ly + ss * ps * t * sq);
            nLy[j] = std::fmax(T(0), sqr(sp) * Lx + sqr(pp) * Ly + sp * pp * t * sq);
            newL[j] = nLx[j] + nLy[j];
            allFinite = allFinite && std::isfinite(newL[j]);
            allPos = allPos && beam.L[j] > T(1e-6) && std::isfinite(newL[j]);
        }
        for (int j = 0; j < N; ++j) {
            const T norm = allPos ? newL[j] / beam.L[j] : T(0);
            beam.L[j] = allFinite ? newL[j] : T(0);
            beam.L1[j] = allFinite ? nLx[j] - nLy[j] : T(0);
            beam.L2[j] *= norm;
            beam.L3[j] *= norm;
        }
    }
    template <int N> void eval(V3<T> toLight, V3<T> toEye, uint32_t lobe, PLTBeam<T, N>& beam) const {  // :120-149
        const T eta = etaAvg;
        const bool isReflection = toLight.z * toEye.z > T(0);
        const bool isDirect = (lobe & Lobe_Delta) != 0;
        const bool hasTransmission = (lobe & Lobe_Transmission) != 0, hasReflection = (lobe & Lobe_Reflection) != 0;
        if ((isReflection && eta == T(1)) || toLight.z * toEye.z == T(0) || !isDirect || (!hasTransmission && !isReflection) ||
            (!hasReflection && isReflection)) {
            beam.setZero();
            return;
        }
        T k[N];
        for (int j = 0; j < N; ++j) k[j] = K<T>::_2PI / (T(1e-3) * beam.wavelengths[j]);
        const T c = std::fabs(toLight.z);
        const M4<T> R = MuellerFresnelRDielectric(c, eta), iR = invOneMinusMuellerFresnelRDielectric(c, eta),
                    Tm = MuellerFresnelTDielectric(c, eta);
        const M4<T> M = isReflection ? (iR.m[0][0] > T(0) ? add4(R, mul4(mul4(Tm, R), mul4(iR, Tm))) : R)
                                     : (iR.m[0][0] > T(0) ? mul4(Tm, mul4(iR, Tm)) : mul4(Tm, Tm));
        if (birefringence > T(0) && !isReflection) handle_birefringence(beam, toLight, k);
        for (int j = 0; j < beam.hwss_comps; ++j) beam.setSP(j, mul(M, beam.getSP(j)));
    }
    template <class S> bool sample(V3<T> wi, V3<T>& wo, T& pdf, T& weight, uint32_t& lobe, S& sg, T /*wavelength*/, uint32_t lobeMask) const {
        const bool hasReflection = (lobeMask & Lobe_Reflection) != 0, hasTransmission = (lobeMask & Lobe_Transmission) != 0;
        if (!hasReflection && !hasTransmission) return false;
        const T Fd = evalFresnelDielectric(etaAvg, std::fabs(wi.z));
        const bool isReflection = hasReflection ? (Fd > T(0) && sampleNext1D<T>(sg) <= Fd) : false;
        wo = isReflection ? V3<T>{-wi.x, -wi.y, wi.z} : -wi;
        weight = 1;
        pdf = hasReflection && hasTransmission ? (isReflection ? Fd : T(1) - Fd) : T(1);
        lobe = isReflection ? Lobe_DeltaReflection : Lobe_DeltaTransmission;
        return true;
    }
    T pdf(V3<T>, V3<T>, T) const { return 0; }
    M2<T> coherenceTransform(V3<T>, V3<T>, uint32_t) const { return ident2<T>(); }
};

// PLTMultiLayeredStackBxDF (PLTMultiLayeredStackBSDF.slang:42-139): alternating two-index stack
// over a base of reflectance albedo; scalar Airy recursion of the amplitude, output Re(G)^2.
template <class T> struct MultiLayerBxDF {
    V3<T> albedo{1, 1, 1};
    T extEta;
    T ior[2], thickness[2];
    uint32_t layers;
    template <int N> void eval(V3<T> toLight, V3<T> toEye, uint32_t lobe, PLTBeam<T, N>& beam) const {  // :58-95
        const bool isReflection = toLight.z * toEye.z > T(0);
        const bool isDirect = (lobe & Lobe_Delta) != 0;
        if (!isDirect || !isReflection) { beam.setZero(); return; }
        const T eta1 = ior[0] / extEta, eta2 = ior[1] / ior[0];
        T cT[2], r[2];
        const T Rout = std::sqrt(evalFresnelDielectric(eta1, toLight.z, cT[0]));
        r[0] = std::sqrt(evalFresnelDielectric(eta2, -cT[0], cT[1]));
        r[1] = std::sqrt(evalFresnelDielectric(eta2, cT[1]));
        cT[0] = std::fabs(cT[0]);
        cT[1] = std::fabs(cT[1]);
        for (int j = 0; j < beam.hwss_comps; ++j) {
            const T k = K<T>::_2PI / (T(1e-3) * beam.wavelengths[j]);
            V2<T> G = {std::sqrt(Spectrum<T>::rgbToSpectrum(albedo, beam.wavelengths[j])), 0};
            for (uint32_t l = 0; l < layers; ++l) {
                const uint32_t idx = l % 2;
                const T d = T(2) * thickness[idx] / cT[idx] * k * ior[idx];
                const V2<T> p = {std::cos(d), std::sin(d)};
                const T R = l == layers - 1 ? Rout : r[idx];
                G = cmul(G, p);
                G = cdiv(V2<T>{R + G.x, G.y}, V2<T>{T(1) + R * G.x, R * G.y});
            }
            beam.setSP(j, sqr(G.x) * beam.getSP(j));
        }
    }
    template <class S> bool sample(V3<T> wi, V3<T>& wo, T& pdf, T& weight, uint32_t& lobe, S&, T wavelength, uint32_t lobeMask) const {
        const bool hasReflection = (lobeMask & Lobe_Reflection) != 0, hasDirect = (lobeMask & Lobe_Delta) != 0;
        if (!hasReflection || !hasDirect || wi.z <= T(0)) return false;
        pdf = 1;
        wo = {-wi.x, -wi.y, wi.z};
        lobe = Lobe_DeltaReflection;
        weight = Spectrum<T>::rgbToSpectrum(albedo, wavelength);
        return wi.z * wo.z > T(0);
    }
    T pdf(V3<T>, V3<T>, T) const { return 0; }
    M2<T> coherenceTransform(V3<T>, V3<T>, uint32_t) const { return ident2<T>(); }
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
    // closestLobeHint (DiffractionGrating.slang:92-107): the diffraction order whose direction is
    // nearest the hint (manifold walks).
    void closestLobeHint(V3<T> wi, V3<T> woHint, T alpha, T lambda, int& lx, int& ly) const {
        lx = ly = 0;
        if (alpha == T(1)) return;
        const T lix = sqr(wi.x) + sqr(wi.z), liy = sqr(wi.y) + sqr(wi.z);
        const T lox = sqr(woHint.x) + sqr(woHint.z), loy = sqr(woHint.y) + sqr(woHint.z);
        const T six = lix > K<T>::FLT_EPS ? wi.x / std::sqrt(lix) : T(0), siy = liy > K<T>::FLT_EPS ? wi.y / std::sqrt(liy) : T(0);
        const T sox = lox > K<T>::FLT_EPS ? -woHint.x / std::sqrt(lox) : T(0), soy = loy > K<T>::FLT_EPS ? -woHint.y / std::sqrt(loy) : T(0);
        const int n = int(lobeCount);
        auto toLobe = [&](T l) { return std::min(std::max(int(l + sgn(l) * T(.5)), -n), n); };
        // (a 1-D grating has no y period: its y order is 0 whatever the hint)
        const T l_x = (six - sox) / (ipx * lambda);
        const T l_y = is1D() ? T(0) : (siy - soy) / (ipy * lambda);
        lx = toLobe(l_x);
        ly = toLobe(l_y);
        if (alpha == T(0) && lx == 0) lx = int(sgn(l_x));
        if (alpha == T(0) && ly == 0) ly = int(sgn(l_y));
        if (is1D()) ly = 0;
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

// The base lobe's share in the diffraction orders: the conductor's colour and Fresnel Mueller
// matrix (PLTDiffractionGratedConductorBSDF.slang:84-98), the opaque dielectric's albedo and no
// Mueller matrix (PLTDiffractionGratedOpaqueDielectricBSDF.slang:84-93).
template <class T> inline T gratingBaseColor(const ConductorBxDF<T>& b, T wavelength) { return Spectrum<T>::rgbToSpectrum(b.specular, wavelength); }
template <class T> inline T gratingBaseColor(const OpaqueDielectricBxDF<T>& b, T wavelength) { return Spectrum<T>::rgbToSpectrum(b.albedo, wavelength); }
template <class T> inline V4<T> gratingBaseMueller(const ConductorBxDF<T>& b, V3<T> toLight, T wavelength, V4<T> sp) {
    const V2<T> ior = {evalSpectrum<T>(b.etaN, wavelength) / b.extEta, evalSpectrum<T>(b.etaK, wavelength) / b.extEta};
    return mul(MuellerFresnelRConductor(toLight.z, ior), sp);
}
template <class T> inline V4<T> gratingBaseMueller(const OpaqueDielectricBxDF<T>&, V3<T>, T, V4<T> sp) { return sp; }

// PLTDiffractionGratedConductorBxDF (PLTDiffractionGratedConductorBSDF.slang:45-213) and
// PLTDiffractionGratedOpaqueDielectricBxDF (PLTDiffractionGratedOpaqueDielectricBSDF.slang:45-209):
// the same grating over a conductor or an opaque dielectric base.
template <class T, class Base> struct GratedBxDF {
    Base base;
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
                const T m00 = gratingBaseColor(base, beam.wavelengths[j]);
                const V3<T> lobeWi = beam.toLocal(grating.diffract(toEye, lx, ly, lam[j]));
                const T coh = j == 0 ? T(1) : beam.mutualCoherenceAngular(lobeWi, toLightBeam);
                const T inten = grating.lobeIntensity(lx, ly, toLight, lam[j]);
                beam.setSP(j, (m00 * (T(1) - a[j]) * coh * inten) * gratingBaseMueller(base, toLight, beam.wavelengths[j], beam.getSP(j)));
            }
        } else beam.setZero();
    }
    template <class S> bool sample(V3<T> wi, V3<T>& wo, T& pdf, T& weight, uint32_t& bsdfLobe, S& sg, T wavelength, uint32_t lobeMask,
                                   V3<T> woHint = V3<T>{0, 0, 0}) const {  // :115-166
        const bool hasReflection = (lobeMask & Lobe_Reflection) != 0, hasDirect = (lobeMask & Lobe_Delta) != 0, hasScatter = (lobeMask & Lobe_NonDelta) != 0;
        if (!hasReflection || (!hasDirect && !hasScatter)) return false;
        const T lam = T(1e-3) * wavelength;
        const T a = grating.alpha(wi, K<T>::_2PI / lam);
        const bool hasHint = hasDirect && (woHint.x != T(0) || woHint.y != T(0) || woHint.z != T(0));
        int lx = 0, ly = 0;
        if (hasHint && a < T(1)) grating.closestLobeHint(wi, woHint, a, lam, lx, ly);
        bool isDirect = hasDirect;
        pdf = 1;
        if (hasHint) { isDirect = lx == 0 && ly == 0; pdf = 1; }
        else if (hasDirect && hasScatter) { isDirect = a == T(1) || sampleNext1D<T>(sg) < a; pdf = isDirect ? a : T(1) - a; }
        if (isDirect) {
            T basePdf;
            base.sample(wi, wo, basePdf, weight, bsdfLobe, sg, wavelength, lobeMask);
            pdf *= basePdf;
            return true;
        }
        weight = gratingBaseColor(base, wavelength);  // base.weight: the base's colour spectrum
        T px = 1, py = 1;
        if (!hasHint) grating.sampleLobe(wi, px, py, lx, ly, lam, sg);
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
    GratedBxDF<T, ConductorBxDF<T>> grt;
    GratedBxDF<T, OpaqueDielectricBxDF<T>> gro;
    CoatedConductorBxDF<T> ccn;
    ThinDielectricBxDF<T> thn;
    MultiLayerBxDF<T> mls;
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
        case MT_GRATED_OPAQUE: {  // PLTDiffractionGratedOpaqueDielectricMaterial.slang:40-70
            in.gro.base.albedo = base;
            in.gro.base.gfs = FractalSurface<T>::makeRef(T(hf(md.data2[1])));
            const float px = hf(md.data1[0]), py = hf(md.data1[1]);
            in.gro.grating.ipx = T(px > 0.f ? 1.f / px : 0.f);
            in.gro.grating.ipy = T(py > 0.f ? 1.f / py : 0.f);
            in.gro.grating.q = T(hf(md.data2[0]) * sx);
            in.gro.grating.mult = T(sy);
            const uint32_t tl = md.data1[3];
            in.gro.grating.lobeCount = std::min(std::max(tl & 0xFFu, 1u), 8u);
            in.gro.grating.type = tl >> 8;
            const T dir = T(hf(md.data1[2]));
            in.gro.grating.gx = std::cos(dir);
            in.gro.grating.gy = std::sin(dir);
            break;
        }
        case MT_COATED_CONDUCTOR:  // PLTCoatedConductorMaterial.slang:40-60
            in.ccn.albedo = base;
            in.ccn.gfs = FractalSurface<T>::makeRef(T(sy));
            in.ccn.etaN = pn; in.ccn.etaK = pk;
            in.ccn.extEta = T(hf(md.ior));
            in.ccn.tau = T(hf(md.data1[0]) * sz);
            in.ccn.coatIOR = T(hf(md.data1[3]));
            break;
        case MT_THIN_DIELECTRIC:  // PLTThinDielectricMaterial.slang:40-56
            in.thn.tau = T(sw);
            in.thn.birefringence = T(hf(md.base_color[0]) * sx);
            in.thn.etaAvg = T(hf(md.ior));
            in.thn.opticAxisXY = V2<T>{T(hf(md.base_color[1])), T(hf(md.base_color[2]))};
            break;
        case MT_MULTILAYER:  // PLTMultiLayeredStackMaterial.slang:40-57
            in.mls.albedo = base;
            in.mls.extEta = T(hf(md.ior));
            in.mls.ior[0] = T(hf(md.data2[0])); in.mls.ior[1] = T(hf(md.data2[1]));
            in.mls.thickness[0] = T(hf(md.data1[0])); in.mls.thickness[1] = T(hf(md.data1[1]));
            in.mls.layers = md.data1[3];
            break;
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
        case MT_COATED_CONDUCTOR: return in.ccn.gfs.isDeltaOnly() ? Lobe_DeltaReflection : (Lobe_DeltaReflection | Lobe_ScatterReflection);
        case MT_THIN_DIELECTRIC: return Lobe_DeltaReflection | Lobe_DeltaTransmission;
        case MT_MULTILAYER: return Lobe_DeltaReflection;
        case MT_GRATED_OPAQUE:
            return (in.gro.base.gfs.isDeltaOnly() ? Lobe_DeltaReflection : (Lobe_DeltaReflection | Lobe_ScatterReflection)) | Lobe_DispersiveDeltaReflection;
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
        case MT_COATED_CONDUCTOR: return in.ccn.gfs.q < cut ? Prop_MWR : 0u;
        case MT_THIN_DIELECTRIC: return Prop_Transmissive | Prop_MWR | Prop_MWT;
        default: return Prop_MWR;  // grated conductor and opaque dielectric, multilayer stack
    }
}

// BSDFProperties.albedo of each material's getProperties (the sample pass's albedo output):
// PLTDielectricBSDF.slang:342-346, PLTConductorBSDF.slang:226-231, PLTDiffractionGratedConductorBSDF.slang:273-280,
// PLTThinDielectricBSDF.slang:259, the base colour elsewhere.
template <class T> inline V3<T> materialAlbedo(const Instance<T>& in) {
    const T meanK = K<T>::_2PI / (T(550) * T(1e-3));
    switch (in.type) {
        case MT_DIELECTRIC: return in.die.specular * evalFresnelDielectric(in.die.indexOfRefraction(meanK), T(1));
        case MT_CONDUCTOR:
            return in.con.specular * evalFresnelConductor(V2<T>{evalSpectrum<T>(in.con.etaN, T(550)), evalSpectrum<T>(in.con.etaK, T(550))}, T(1));
        case MT_GRATED_CONDUCTOR:
            return in.grt.base.specular *
                   evalFresnelConductor(V2<T>{evalSpectrum<T>(in.grt.base.etaN, T(550)), evalSpectrum<T>(in.grt.base.etaK, T(550))}, T(1));
        case MT_DIFFUSE: return in.dif.albedo;
        case MT_OPAQUE: return in.opq.albedo;
        case MT_COATED_OPAQUE: return in.cot.albedo;
        case MT_COATED_CONDUCTOR: return in.ccn.albedo;
        case MT_GRATED_OPAQUE: return in.gro.base.albedo;
        case MT_MULTILAYER: return in.mls.albedo;
        default: return V3<T>{1, 1, 1};  // thin dielectric
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
    if (in.type != MT_GRATED_CONDUCTOR && in.type != MT_GRATED_OPAQUE) return sd;
    const DiffractionGrating<T>& gr = in.type == MT_GRATED_CONDUCTOR ? in.grt.grating : in.gro.grating;
    Frame<T> g;
    g.T_ = sd.fromLocal(V3<T>{gr.gx, gr.gy, 0});
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
        case MT_COATED_CONDUCTOR: bsdfEval<T, N>(in.ccn, f, sd.V, toEye, lobe, beam); break;
        case MT_THIN_DIELECTRIC: bsdfEval<T, N>(in.thn, f, sd.V, toEye, lobe, beam); break;
        case MT_MULTILAYER: bsdfEval<T, N>(in.mls, f, sd.V, toEye, lobe, beam); break;
        case MT_GRATED_OPAQUE: bsdfEval<T, N>(in.gro, f, sd.V, toEye, lobe, beam); break;
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
        case MT_COATED_CONDUCTOR: in.ccn.lambdaPdf(a, b, lobe, beam, ret); break;
        case MT_GRATED_CONDUCTOR: in.grt.lambdaPdf(a, b, lobe, beam, ret); break;
        case MT_GRATED_OPAQUE: in.gro.lambdaPdf(a, b, lobe, beam, ret); break;
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
        case MT_COATED_CONDUCTOR: return in.ccn.pdf(a, b, wavelength);
        case MT_THIN_DIELECTRIC: return in.thn.pdf(a, b, wavelength);
        case MT_MULTILAYER: return in.mls.pdf(a, b, wavelength);
        case MT_GRATED_OPAQUE: return in.gro.pdf(a, b, wavelength);
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
        case MT_COATED_CONDUCTOR: valid = in.ccn.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, Lobe_All); break;
        case MT_THIN_DIELECTRIC: valid = in.thn.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, Lobe_All); break;
        case MT_MULTILAYER: valid = in.mls.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, Lobe_All); break;
        case MT_GRATED_OPAQUE: valid = in.gro.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, Lobe_All); break;
        default: valid = in.grt.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, Lobe_All); break;
    }
    if (valid) r.wo = f.fromLocal(r.wo);
    return valid;
}

// ------------------------------------------------------------------------------------------
// Differential geometry for manifold walks (Rendering/Utils/DifferentialShadingGeometry.slang).
// ------------------------------------------------------------------------------------------
// Spherical angles of the manifold walk's Newton step, in double precision: the Cephes library's
// rational approximation of atan and its range reduction (S. L. Moshier, atan.c), atan2 by quadrant
// and acos(x) = 2 atan(sqrt((1 - x) / (1 + x))), written with +, -, *, / and sqrt only, so that a
// strict-IEEE host and the device compute the same bits.
inline double walkAtan(double xx) {
    double x = std::fabs(xx), y;
    int flag;
    if (x > 2.41421356237309504880) { y = 1.57079632679489661923; flag = 1; x = -(1.0 / x); }
    else if (x <= 0.66) { y = 0.0; flag = 0; }
    else { y = 0.785398163397448309616; flag = 2; x = (x - 1.0) / (x + 1.0); }
    const double z = x * x;
    const double P = (((-8.750608600031904122785e-1 * z - 1.615753718733365076637e1) * z - 7.500855792314704667340e1) * z -
                      1.228866684490136173410e2) * z - 6.485021904942025371773e1;
    const double Q = ((((z + 2.485846490142306297962e1) * z + 1.650270098316988542046e2) * z + 4.328810604912902668951e2) * z +
                      4.853903996359136964868e2) * z + 1.945506571482613964425e2;
    double r = x * (z * P / Q) + x;
    if (flag == 2) r += 0.5 * 6.123233995736765886130e-17;
    else if (flag == 1) r += 6.123233995736765886130e-17;
    y += r;
    return xx < 0.0 ? -y : y;
}
inline double walkAtan2(double y, double x) {
    const int code = (x < 0.0 ? 2 : 0) | (y < 0.0 ? 1 : 0);
    if (x == 0.0) return (code & 1) ? -1.57079632679489661923 : (y == 0.0 ? 0.0 : 1.57079632679489661923);
    if (y == 0.0) return (code & 2) ? 3.14159265358979323846 : 0.0;
    const double w = code == 2 ? 3.14159265358979323846 : code == 3 ? -3.14159265358979323846 : 0.0;
    return w + walkAtan(y / x);
}
inline double walkAcos(double x) { return 2.0 * walkAtan(std::sqrt((1.0 - x) / (1.0 + x))); }

template <class T> struct PDV {
    V3<T> v, ddu, ddv;
    void flip() { v = -v; ddu = -ddu; ddv = -ddv; }
};
template <class T> struct KPDV { V3<T> v{0, 0, 0}, ddu{0, 0, 0}, ddv{0, 0, 0}, ddl{0, 0, 0}; };
template <class T> struct DSG {
    PDV<T> p, n, s, t;
    V3<T> ngeo;
    PDV<T> pdw(V3<T> dst) const {  // :50-58
        PDV<T> w;
        w.v = dst - p.v;
        const T ilo = T(1) / length(w.v);
        w.v = w.v * ilo;
        w.ddu = (p.ddu - w.v * dot(w.v, p.ddu)) * -ilo;
        w.ddv = (p.ddv - w.v * dot(w.v, p.ddv)) * -ilo;
        return w;
    }
    V3<T> fromLocal(V3<T> v) const { return s.v * v.x + t.v * v.y + n.v * v.z; }
};
// differentialShadingGeometry (:120-190) for a triangle without texture coordinates: the basic
// parameterisation uv0 = 0, uv1 = (1, 0), uv2 = (0, 1).
template <class T> inline DSG<T> makeDSG(const V3<T> P[3], const V3<T> Nv[3], T b1, T b2) {
    const T b0 = T(1) - b1 - b2;
    DSG<T> g;
    g.ngeo = normalize(cross(P[1] - P[0], P[2] - P[0]));
    g.p.v = P[0] * b0 + P[1] * b1 + P[2] * b2;
    g.n.v = Nv[0] * b0 + Nv[1] * b1 + Nv[2] * b2;
    const T il = T(1) / length(g.n.v);
    g.n.v = g.n.v * il;
    g.n.ddu = (Nv[1] - Nv[0]) * il;
    g.n.ddv = (Nv[2] - Nv[0]) * il;
    g.n.ddu = g.n.ddu - g.n.v * dot(g.n.v, g.n.ddu);
    g.n.ddv = g.n.ddv - g.n.v * dot(g.n.v, g.n.ddv);
    // det of (duv0, duv1) = 1: the position derivatives are the edges
    g.p.ddu = P[1] - P[0];
    g.p.ddv = P[2] - P[0];
    g.s.v = g.p.ddu - g.n.v * dot(g.n.v, g.p.ddu);
    T ils = T(1) / length(g.s.v);
    g.s.v = g.s.v * ils;
    if (g.p.ddu.x == T(0) && g.p.ddu.y == T(0) && g.p.ddu.z == T(0)) { g.s.v = perp_stark(g.n.v); ils = 1; }
    g.t.v = cross(g.n.v, g.s.v);
    g.s.ddu = (-g.n.ddu * dot(g.n.v, g.p.ddu) - g.n.v * dot(g.n.ddu, g.p.ddu)) * ils;
    g.s.ddv = (-g.n.ddv * dot(g.n.v, g.p.ddu) - g.n.v * dot(g.n.ddv, g.p.ddu)) * ils;
    g.s.ddu = g.s.ddu - g.s.v * dot(g.s.ddu, g.s.v);
    g.s.ddv = g.s.ddv - g.s.v * dot(g.s.ddv, g.s.v);
    g.t.ddu = cross(g.n.ddu, g.s.v) + cross(g.n.v, g.s.ddu);
    g.t.ddv = cross(g.n.ddv, g.s.v) + cross(g.n.v, g.s.ddv);
    return g;
}
template <class T> inline KPDV<T> reflectDerivatives(const PDV<T>& wi, const PDV<T>& n) {  // :192-205
    const T wn = dot(wi.v, n.v);
    const T dwdun = dot(wi.ddu, n.v), dwdvn = dot(wi.ddv, n.v);
    const T dnduw = dot(n.ddu, wi.v), dndvw = dot(n.ddv, wi.v);
    KPDV<T> r;
    r.v = n.v * (T(2) * wn) - wi.v;
    r.ddu = (n.v * (dwdun + dnduw) + n.ddu * wn) * T(2) - wi.ddu;
    r.ddv = (n.v * (dwdvn + dndvw) + n.ddv * wn) * T(2) - wi.ddv;
    return r;
}
template <class T> inline KPDV<T> refractDerivatives(const PDV<T>& wi, PDV<T> n, T eta) {  // :206-222 (no dispersion)
    const T wn = dot(wi.v, n.v);
    eta = wn < T(0) ? eta : T(1) / eta;
    if (wn < T(0)) n.flip();
    const T eta2 = sqr(eta);
    const T dwdun = dot(wi.ddu, n.v), dwdvn = dot(wi.ddv, n.v);
    const T dnduw = dot(n.ddu, wi.v), dndvw = dot(n.ddv, wi.v);
    const T root = std::sqrt(hmax(T(0), T(1) - eta2 * (T(1) - sqr(wn))));
    if (root == T(0)) return KPDV<T>{};
    const T a = eta - eta2 * wn / root;
    KPDV<T> r;
    r.v = (wi.v - n.v * wn) * -eta - n.v * root;
    r.ddu = n.v * (a * (dnduw + dwdun)) + n.ddu * (eta * wn - root) - wi.ddu * eta;
    r.ddv = n.v * (a * (dndvw + dwdvn)) + n.ddv * (eta * wn - root) - wi.ddv * eta;
    return r;
}
// DiffractionGrating::partialDerivatives (DiffractionGrating.slang:188-265) in world space.
template <class T> inline KPDV<T> gratingDerivatives(const DiffractionGrating<T>& g, const DSG<T>& dsg, const PDV<T>& wi, int lx, int ly, T wavelength) {
    const PDV<T>& n = dsg.n;
    PDV<T> X;
    X.v = dsg.fromLocal(V3<T>{g.gx, g.gy, 0});
    X.ddu = dsg.s.ddu * g.gx + dsg.t.ddu * g.gy;
    X.ddv = dsg.s.ddv * g.gx + dsg.t.ddv * g.gy;
    if (lx == 0 && ly == 0) return reflectDerivatives(wi, n);
    const T lambda = T(1e-3) * wavelength;
    const T lpx = T(lx) * g.ipx, lpy = T(ly) * g.ipy;
    const V3<T> Y = cross(n.v, X.v);
    const T x = dot(wi.v, X.v), y = dot(wi.v, Y), z = dot(wi.v, n.v);
    const T px = std::sqrt(sqr(x) + sqr(z)), py = std::sqrt(sqr(y) + sqr(z));
    const T ipx_ = px > K<T>::FLT_EPS ? T(1) / px : T(0), ipy_ = py > K<T>::FLT_EPS ? T(1) / py : T(0);
    const T a = lambda * lpx - x * ipx_, b = lambda * lpy - y * ipy_;
    const T oneMinusA2 = hmax(T(0), T(1) - sqr(a));
    const T m = -oneMinusA2 / (sqr(a * b) - T(1));
    const T q = T(1) - sqr(b) * m;
    if (a > T(1) || a < T(-1) || b > T(1) || b < T(-1) || q < T(0) || !(m >= T(0))) return KPDV<T>{};
    const T sqrtOneM = std::sqrt(q), sqrtM = std::sqrt(m);
    const T c = T(.5) / (sqrtOneM * std::sqrt(oneMinusA2));
    const T dxdu = dot(wi.v, X.ddu) + dot(wi.ddu, X.v), dxdv = dot(wi.v, X.ddv) + dot(wi.ddv, X.v);
    const T dzdu = dot(wi.v, n.ddu) + dot(wi.ddu, n.v), dzdv = dot(wi.v, n.ddv) + dot(wi.ddv, n.v);
    const T ipx3 = sqr(ipx_) * ipx_, ipy3 = sqr(ipy_) * ipy_;
    const T dadu = z * (x * dzdu - z * dxdu) * ipx3, dbdu = z * (y * dzdu - z * dxdu) * ipy3;
    const T dadv = z * (x * dzdv - z * dxdv) * ipx3, dbdv = z * (y * dzdv - z * dxdv) * ipy3;
    const T dadl = lpx, dbdl = lpy;
    const T den = sqr(sqr(a * b) - T(1));
    const T dmdu = T(2) * a * (a * b * oneMinusA2 * dbdu - dadu * (T(1) - sqr(b))) / den;
    const T dmdv = T(2) * a * (a * b * oneMinusA2 * dbdv - dadv * (T(1) - sqr(b))) / den;
    const T dmdl = T(2) * a * (a * b * oneMinusA2 * dbdl - dadl * (T(1) - sqr(b))) / den;
    KPDV<T> r;
    r.v = X.v * (a * sqrtOneM) + Y * (b * sqrtM) + n.v * std::sqrt(hmax(T(0), T(1) - sqr(a) * q - sqr(b) * m));
    auto d = [&](T dad, T dbd, T dmd) {
        return X.v * ((q * dad - a * b * (b * dmd / T(2) + m * dbd)) / sqrtOneM) + Y * ((b * dmd / T(2) + m * dbd) / sqrtM) +
               n.v * ((T(-2) * a * dad * q - b * (b * dmd + T(2) * m * dbd) + sqr(a) * b * (b * dmd + T(2) * m * dbd)) * c);
    };
    r.ddu = d(dadu, dbdu, dmdu);
    r.ddv = d(dadv, dbdv, dmdv);
    r.ddl = d(dadl, dbdl, dmdl);
    return r;
}

// IMaterialInstance::partialDerivatives of each material (world space; lobe = the sampled lobe).
template <class T> inline KPDV<T> materialPartialDerivatives(const Instance<T>& in, const PDV<T>& wi, const DSG<T>& dsg, uint32_t lobe, T wavelength) {
    const bool isTransmission = (lobe & Lobe_Transmission) != 0;
    switch (in.type) {
        case MT_DIELECTRIC:  // PLTDielectricBSDF.slang:139-157
            if (isTransmission) return refractDerivatives(wi, dsg.n, in.die.indexOfRefraction(K<T>::_2PI / (T(1e-3) * wavelength)));
            return reflectDerivatives(wi, dsg.n);
        case MT_DIFFUSE: return KPDV<T>{};  // PLTDiffuseBSDF.slang:113-117 (unsupported)
        case MT_THIN_DIELECTRIC:  // PLTThinDielectricBSDF.slang:151-162
            if (isTransmission) { PDV<T> w = wi; w.flip(); KPDV<T> r; r.v = w.v; r.ddu = w.ddu; r.ddv = w.ddv; return r; }
            return reflectDerivatives(wi, dsg.n);
        case MT_GRATED_CONDUCTOR:     // PLTDiffractionGratedConductorBSDF.slang:104-112
        case MT_GRATED_OPAQUE: {  // PLTDiffractionGratedOpaqueDielectricBSDF.slang:101-107
            int lx, ly;
            GratedBxDF<T, ConductorBxDF<T>>::decode(lobe, lx, ly);
            if (lx == 0 && ly == 0) return reflectDerivatives(wi, dsg.n);
            return gratingDerivatives(in.type == MT_GRATED_CONDUCTOR ? in.grt.grating : in.gro.grating, dsg, wi, lx, ly, wavelength);
        }
        default: return reflectDerivatives(wi, dsg.n);  // conductor, opaque, coated, multilayer stack
    }
}

// IMaterialInstance::sample with a lobe mask and a direction hint (manifold walks): the dielectric
// and thin dielectric restrict themselves to reflection or transmission by the side of the hint,
// the grating picks the diffraction order nearest the hint; the others ignore it.
template <class S> inline bool materialSampleHinted(const Instance<float>& in, const SD<float>& sd, S& sg, BSDFSample& r, float wavelength,
                                                    uint32_t lobeMask, V3<float> woHint) {
    const Frame<float> f = bsdfFrame(in, sd.f);
    const V3<float> wi = f.toLocal(sd.V);
    const float dh = dot(sd.V, woHint);
    const uint32_t forced = (dh > 0.f ? uint32_t(Lobe_Reflection) : dh < 0.f ? uint32_t(Lobe_Transmission) : uint32_t(Lobe_All)) & lobeMask;
    bool valid = false;
    switch (in.type) {
        case MT_DIELECTRIC:  // PLTDielectricBSDF.slang:309-320
            valid = in.die.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, forced);
            r.wo = f.fromLocal(r.wo);
            return valid;
        case MT_THIN_DIELECTRIC: valid = in.thn.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, forced); break;
        case MT_GRATED_CONDUCTOR: valid = in.grt.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, lobeMask, f.toLocal(woHint)); break;
        case MT_GRATED_OPAQUE: valid = in.gro.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, lobeMask, f.toLocal(woHint)); break;
        case MT_CONDUCTOR: valid = in.con.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, lobeMask); break;
        case MT_DIFFUSE: valid = in.dif.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, lobeMask); r.wo = f.fromLocal(r.wo); return valid;
        case MT_OPAQUE: valid = in.opq.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, lobeMask); break;
        case MT_COATED_OPAQUE: valid = in.cot.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, lobeMask); break;
        case MT_COATED_CONDUCTOR: valid = in.ccn.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, lobeMask); break;
        case MT_MULTILAYER: valid = in.mls.sample(wi, r.wo, r.pdf, r.weight, r.lobe, sg, wavelength, lobeMask); break;
        default: break;
    }
    if (valid) r.wo = f.fromLocal(r.wo);
    return valid;
}

}  // namespace pltoracle
