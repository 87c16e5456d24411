// Device pieces of manifold next-event estimation (Source/RenderPasses/PLTPT/Manifolds.slangh):
// the differential shading geometry of a triangle (Rendering/Utils/DifferentialShadingGeometry.slang),
// the partial derivatives of each material's specular direction (IMaterialInstance::partialDerivatives),
// the delta-lobe sample with a direction hint that seeds a walk vertex, and one Newton step on the
// angle-difference constraint.
//
// Every expression is evaluated uncontracted, with correctly rounded division and square root, in
// the operation order of a plain strict-IEEE evaluation, and the Newton step's spherical angles use
// the fixed rational atan below (FP64): the walk's vertices then round identically on the host and
// the device. That matters because the walk's last visibility ray ends on the light's own surface,
// where an ulp decides whether the light's triangle occludes it.
#pragma once
#include "plt_materials.hpp"

namespace plt {

// ---------------------------------------------------------------------------------------
// atan / atan2 / acos in double precision for the Newton step: the Cephes library's rational
// approximation of atan with its range reduction (S. L. Moshier, atan.c; within 1 ulp), atan2 by
// quadrant, acos(x) = 2 atan(sqrt((1 - x) / (1 + x))). Written with +, -, *, / and sqrt only
// (correctly rounded in double on both sides), so host and device compute the same bits.
// ---------------------------------------------------------------------------------------
PLT_DEV double walk_atan(double xx) {
#pragma clang fp contract(off)
    double x = fabs(xx), y;
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
PLT_DEV double walk_atan2(double y, double x) {
#pragma clang fp contract(off)
    const int code = (x < 0.0 ? 2 : 0) | (y < 0.0 ? 1 : 0);
    if (x == 0.0) return (code & 1) ? -1.57079632679489661923 : (y == 0.0 ? 0.0 : 1.57079632679489661923);
    if (y == 0.0) return (code & 2) ? 3.14159265358979323846 : 0.0;
    const double w = code == 2 ? 3.14159265358979323846 : code == 3 ? -3.14159265358979323846 : 0.0;
    return w + walk_atan(y / x);
}
PLT_DEV double walk_acos(double x) {
#pragma clang fp contract(off)
    return 2.0 * walk_atan(sqrt((1.0 - x) / (1.0 + x)));
}

// Strict-order vector helpers of the walk (division by the length, as a plain evaluation does).
PLT_DEV float wdot(V3 a, V3 b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
PLT_DEV V3 wcross(V3 a, V3 b) {
#pragma clang fp contract(off)
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
PLT_DEV V3 wscale(V3 a, float s) {
#pragma clang fp contract(off)
    return v3(a.x * s, a.y * s, a.z * s);
}
PLT_DEV V3 wdiv(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
PLT_DEV float wlength(V3 a) { return sqrtf(wdot(a, a)); }
PLT_DEV V3 wnormalize(V3 a) { return wdiv(a, sqrtf(wdot(a, a))); }
// a * s + b * t, left to right, rounded after each operation
PLT_DEV V3 wmad2(V3 a, float s, V3 b, float t) {
#pragma clang fp contract(off)
    return v3(a.x * s + b.x * t, a.y * s + b.y * t, a.z * s + b.z * t);
}

// A direction with its derivatives along the surface parameters u, v (float3pdv).
struct PDV {
    V3 v, ddu, ddv;
};
PLT_DEV PDV pdv_flip(PDV w) { return PDV{-w.v, -w.ddu, -w.ddv}; }

// Differential shading geometry (DifferentialShadingGeometry.slang:120-190) of a triangle without
// texture coordinates (the basic parameterisation uv0 = 0, uv1 = (1, 0), uv2 = (0, 1)).
struct DSG {
    PDV p, n, s, t;
    // pdw (:50-58): the unit direction from the point to dst and its derivatives.
    PLT_DEV PDV pdw(V3 dst) const {
#pragma clang fp contract(off)
        PDV w;
        w.v = dst - p.v;
        const float ilo = 1.f / wlength(w.v);
        w.v = wscale(w.v, ilo);
        w.ddu = wscale(p.ddu - wscale(w.v, wdot(w.v, p.ddu)), -ilo);
        w.ddv = wscale(p.ddv - wscale(w.v, wdot(w.v, p.ddv)), -ilo);
        return w;
    }
    PLT_DEV V3 from_local(V3 l) const {
#pragma clang fp contract(off)
        return v3(s.v.x * l.x + t.v.x * l.y + n.v.x * l.z, s.v.y * l.x + t.v.y * l.y + n.v.y * l.z,
                  s.v.z * l.x + t.v.z * l.y + n.v.z * l.z);
    }
};
PLT_DEV DSG make_dsg(V3 P0, V3 P1, V3 P2, V3 N0, V3 N1, V3 N2, float b1, float b2) {
#pragma clang fp contract(off)
    const float b0 = 1.f - b1 - b2;
    DSG g;
    g.p.v = v3(P0.x * b0 + P1.x * b1 + P2.x * b2, P0.y * b0 + P1.y * b1 + P2.y * b2, P0.z * b0 + P1.z * b1 + P2.z * b2);
    g.n.v = v3(N0.x * b0 + N1.x * b1 + N2.x * b2, N0.y * b0 + N1.y * b1 + N2.y * b2, N0.z * b0 + N1.z * b1 + N2.z * b2);
    const float il = 1.f / wlength(g.n.v);
    g.n.v = wscale(g.n.v, il);
    g.n.ddu = wscale(N1 - N0, il);
    g.n.ddv = wscale(N2 - N0, il);
    g.n.ddu = g.n.ddu - wscale(g.n.v, wdot(g.n.v, g.n.ddu));
    g.n.ddv = g.n.ddv - wscale(g.n.v, wdot(g.n.v, g.n.ddv));
    // the parameterisation's Jacobian has determinant 1: the position derivatives are the edges
    g.p.ddu = P1 - P0;
    g.p.ddv = P2 - P0;
    g.s.v = g.p.ddu - wscale(g.n.v, wdot(g.n.v, g.p.ddu));
    float ils = 1.f / wlength(g.s.v);
    g.s.v = wscale(g.s.v, ils);
    if (g.p.ddu.x == 0.f && g.p.ddu.y == 0.f && g.p.ddu.z == 0.f) {
        // perp_stark with division normalisation
        const V3 u = g.n.v;
        const V3 a = v3(fabsf(u.x), fabsf(u.y), fabsf(u.z));
        const uint32_t uyx = (a.x - a.y) < 0.f ? 1u : 0u, uzx = (a.x - a.z) < 0.f ? 1u : 0u, uzy = (a.y - a.z) < 0.f ? 1u : 0u;
        const uint32_t xm = uyx & uzx, ym = (1u ^ xm) & uzy, zm = 1u ^ (xm | ym);
        g.s.v = wnormalize(wcross(u, v3(float(xm), float(ym), float(zm))));
        ils = 1.f;
    }
    g.t.v = wcross(g.n.v, g.s.v);
    const float npu = wdot(g.n.v, g.p.ddu);
    g.s.ddu = wscale(wscale(-g.n.ddu, npu) - wscale(g.n.v, wdot(g.n.ddu, g.p.ddu)), ils);
    g.s.ddv = wscale(wscale(-g.n.ddv, npu) - wscale(g.n.v, wdot(g.n.ddv, g.p.ddu)), ils);
    g.s.ddu = g.s.ddu - wscale(g.s.v, wdot(g.s.ddu, g.s.v));
    g.s.ddv = g.s.ddv - wscale(g.s.v, wdot(g.s.ddv, g.s.v));
    g.t.ddu = wcross(g.n.ddu, g.s.v) + wcross(g.n.v, g.s.ddu);
    g.t.ddv = wcross(g.n.ddv, g.s.v) + wcross(g.n.v, g.s.ddv);
    return g;
}

// Specular reflection / refraction directions with their derivatives
// (DifferentialShadingGeometry.slang:192-222; no wavelength derivative).
PLT_DEV PDV reflect_derivatives(const PDV& wi, const PDV& n) {
#pragma clang fp contract(off)
    const float wn = wdot(wi.v, n.v);
    const float dwdun = wdot(wi.ddu, n.v), dwdvn = wdot(wi.ddv, n.v);
    const float dnduw = wdot(n.ddu, wi.v), dndvw = wdot(n.ddv, wi.v);
    PDV r;
    r.v = wscale(n.v, 2.f * wn) - wi.v;
    r.ddu = wscale(wscale(n.v, dwdun + dnduw) + wscale(n.ddu, wn), 2.f) - wi.ddu;
    r.ddv = wscale(wscale(n.v, dwdvn + dndvw) + wscale(n.ddv, wn), 2.f) - wi.ddv;
    return r;
}
PLT_DEV PDV refract_derivatives(const PDV& wi, PDV n, float eta) {
#pragma clang fp contract(off)
    const float wn = wdot(wi.v, n.v);
    eta = wn < 0.f ? eta : 1.f / eta;
    if (wn < 0.f) n = pdv_flip(n);
    const float eta2 = eta * eta;
    const float dwdun = wdot(wi.ddu, n.v), dwdvn = wdot(wi.ddv, n.v);
    const float dnduw = wdot(n.ddu, wi.v), dndvw = wdot(n.ddv, wi.v);
    const float root = sqrtf(fmaxf(0.f, 1.f - eta2 * (1.f - wn * wn)));
    if (root == 0.f) return PDV{v3(0.f, 0.f, 0.f), v3(0.f, 0.f, 0.f), v3(0.f, 0.f, 0.f)};
    const float a = eta - eta2 * wn / root;
    PDV r;
    r.v = wscale(wi.v - wscale(n.v, wn), -eta) - wscale(n.v, root);
    r.ddu = wscale(n.v, a * (dnduw + dwdun)) + wscale(n.ddu, eta * wn - root) - wscale(wi.ddu, eta);
    r.ddv = wscale(n.v, a * (dndvw + dwdvn)) + wscale(n.ddv, eta * wn - root) - wscale(wi.ddv, eta);
    return r;
}
// DiffractionGrating::partialDerivatives (DiffractionGrating.slang:188-265) in world space.
PLT_DEV PDV grating_derivatives(const Grating& g, const DSG& dsg, const PDV& wi, int lx, int ly, float wavelength) {
#pragma clang fp contract(off)
    const PDV& n = dsg.n;
    PDV X;
    X.v = dsg.from_local(v3(g.gx, g.gy, 0.f));
    X.ddu = wscale(dsg.s.ddu, g.gx) + wscale(dsg.t.ddu, g.gy);
    X.ddv = wscale(dsg.s.ddv, g.gx) + wscale(dsg.t.ddv, g.gy);
    if (lx == 0 && ly == 0) return reflect_derivatives(wi, n);
    const PDV zero{v3(0.f, 0.f, 0.f), v3(0.f, 0.f, 0.f), v3(0.f, 0.f, 0.f)};
    const float lambda = 1e-3f * wavelength;
    const float lpx = float(lx) * g.ipx, lpy = float(ly) * g.ipy;
    const V3 Y = wcross(n.v, X.v);
    const float x = wdot(wi.v, X.v), y = wdot(wi.v, Y), z = wdot(wi.v, n.v);
    const float px = sqrtf(x * x + z * z), py = sqrtf(y * y + z * z);
    const float ipx_ = px > kFltEps ? 1.f / px : 0.f, ipy_ = py > kFltEps ? 1.f / py : 0.f;
    const float a = lambda * lpx - x * ipx_, b = lambda * lpy - y * ipy_;
    const float oneMinusA2 = fmaxf(0.f, 1.f - a * a);
    const float ab = a * b;
    const float m = -oneMinusA2 / (ab * ab - 1.f);
    const float q = 1.f - b * b * m;
    if (a > 1.f || a < -1.f || b > 1.f || b < -1.f || q < 0.f || !(m >= 0.f)) return zero;
    const float sqrtOneM = sqrtf(q), sqrtM = sqrtf(m);
    const float c = .5f / (sqrtOneM * sqrtf(oneMinusA2));
    const float dxdu = wdot(wi.v, X.ddu) + wdot(wi.ddu, X.v), dxdv = wdot(wi.v, X.ddv) + wdot(wi.ddv, X.v);
    const float dzdu = wdot(wi.v, n.ddu) + wdot(wi.ddu, n.v), dzdv = wdot(wi.v, n.ddv) + wdot(wi.ddv, n.v);
    const float ipx3 = ipx_ * ipx_ * ipx_, ipy3 = ipy_ * ipy_ * ipy_;
    const float dadu = z * (x * dzdu - z * dxdu) * ipx3, dbdu = z * (y * dzdu - z * dxdu) * ipy3;
    const float dadv = z * (x * dzdv - z * dxdv) * ipx3, dbdv = z * (y * dzdv - z * dxdv) * ipy3;
    const float abm1 = ab * ab - 1.f;
    const float den = abm1 * abm1;
    const float dmdu = 2.f * a * (a * b * oneMinusA2 * dbdu - dadu * (1.f - b * b)) / den;
    const float dmdv = 2.f * a * (a * b * oneMinusA2 * dbdv - dadv * (1.f - b * b)) / den;
    PDV r;
    r.v = wscale(X.v, a * sqrtOneM) + wscale(Y, b * sqrtM) + wscale(n.v, sqrtf(fmaxf(0.f, 1.f - a * a * q - b * b * m)));
    auto d = [&](float dad, float dbd, float dmd) {
        return wscale(X.v, (q * dad - a * b * (b * dmd / 2.f + m * dbd)) / sqrtOneM) +
               wscale(Y, (b * dmd / 2.f + m * dbd) / sqrtM) +
               wscale(n.v, (-2.f * a * dad * q - b * (b * dmd + 2.f * m * dbd) + a * a * b * (b * dmd + 2.f * m * dbd)) * c);
    };
    r.ddu = d(dadu, dbdu, dmdu);
    r.ddv = d(dadv, dbdv, dmdv);
    return r;
}

// IMaterialInstance::partialDerivatives of each material (world space; lobe = the walk vertex's
// sampled lobe). A zero direction means the material takes no part in walks.
PLT_DEV PDV mat_partial_derivatives(const Mat& m, const PDV& wi, const DSG& dsg, uint32_t lobe, float wavelength) {
#pragma clang fp contract(off)
    const bool isTransmission = (lobe & kLobeTransmission) != 0;
    switch (m.type) {
        case PLT_MAT_DIELECTRIC: {  // PLTDielectricBSDF.slang:139-157
            if (!isTransmission) return reflect_derivatives(wi, dsg.n);
            const float k = k2Pi / (1e-3f * wavelength);
            const float u = k * k1_2Pi;
            return refract_derivatives(wi, dsg.n, (m.etaInt + m.dispB * (u * u)) / m.etaExt);
        }
        case PLT_MAT_DIFFUSE:  // PLTDiffuseBSDF.slang:113-117
            return PDV{v3(0.f, 0.f, 0.f), v3(0.f, 0.f, 0.f), v3(0.f, 0.f, 0.f)};
        case PLT_MAT_THIN_DIELECTRIC:  // PLTThinDielectricBSDF.slang:151-162
            return isTransmission ? pdv_flip(wi) : reflect_derivatives(wi, dsg.n);
        case PLT_MAT_GRATED_CONDUCTOR:           // PLTDiffractionGratedConductorBSDF.slang:104-112
        case PLT_MAT_GRATED_OPAQUE_DIELECTRIC: {  // PLTDiffractionGratedOpaqueDielectricBSDF.slang:101-107
            int lx, ly;
            decode_order(lobe, lx, ly);
            if (lx == 0 && ly == 0) return reflect_derivatives(wi, dsg.n);
            return grating_derivatives(m.gr, dsg, wi, lx, ly, wavelength);
        }
        default:  // conductor, opaque and coated dielectrics, coated conductor, multilayer stack
            return reflect_derivatives(wi, dsg.n);
    }
}

// DiffractionGrating::closestLobeHint (DiffractionGrating.slang:92-107): the diffraction order
// whose direction lies nearest the hint.
PLT_DEV void grating_closest_lobe(const Grating& g, V3 wi, V3 hint, float alpha, float lambda, int& lx, int& ly) {
#pragma clang fp contract(off)
    lx = ly = 0;
    if (alpha == 1.f) return;
    const float lix = wi.x * wi.x + wi.z * wi.z, liy = wi.y * wi.y + wi.z * wi.z;
    const float lox = hint.x * hint.x + hint.z * hint.z, loy = hint.y * hint.y + hint.z * hint.z;
    const float six = lix > kFltEps ? wi.x / sqrtf(lix) : 0.f, siy = liy > kFltEps ? wi.y / sqrtf(liy) : 0.f;
    const float sox = lox > kFltEps ? -hint.x / sqrtf(lox) : 0.f, soy = loy > kFltEps ? -hint.y / sqrtf(loy) : 0.f;
    const int n = int(g.lobes);
    auto to_lobe = [&](float l) { return min(max(int(l + sgnf(l) * .5f), -n), n); };
    // a 1-D grating has no y period: its y order is 0 whatever the hint
    const float l_x = (six - sox) / (g.ipx * lambda);
    const float l_y = g.is1d() ? 0.f : (siy - soy) / (g.ipy * lambda);
    lx = to_lobe(l_x);
    ly = to_lobe(l_y);
    if (alpha == 0.f && lx == 0) lx = int(sgnf(l_x));
    if (alpha == 0.f && ly == 0) ly = int(sgnf(l_y));
    if (g.is1d()) ly = 0;
}

// IMaterialInstance::sample restricted to the delta lobes with a direction hint, as a walk vertex is
// seeded (Manifolds.slangh:84-105): the dielectrics keep to reflection or transmission by the side
// of the hint, the grating takes the order nearest the hint, the others reflect specularly.
// Random numbers are drawn exactly where the full sample would draw them under that lobe mask.
PLT_DEV bool mat_sample_delta_hinted(const Mat& m, const Frame& sd, V3 sdV, Rng& sg, float lambda, V3 hintW, BsdfSample& s) {
    const Frame f = mat_frame(m, sd);
    const V3 wi = f.to_local(sdV);
    const float dh = dot(sdV, hintW);
    const uint32_t forced = (dh > 0.f ? kLobeReflection : dh < 0.f ? kLobeTransmission : kLobeAll) & kLobeDelta;
    const bool fR = (forced & kLobeReflection) != 0, fT = (forced & kLobeTransmission) != 0;
    const float k = k2Pi / (1e-3f * lambda);
    V3 wo = v3(0.f, 0.f, 0.f);
    float pdf = 1.f, weight = 0.f;
    uint32_t lobe = 0;
    bool valid = false;
    switch (m.type) {
        case PLT_MAT_DIELECTRIC: {  // PLTDielectricBSDF.slang:159-211, delta lobes only
            const float u = k * k1_2Pi;
            const float eta = (m.etaInt + m.dispB * (u * u)) / m.etaExt;
            float cT;
            const float Fd = fresnel_dielectric_F(eta, wi.z, cT);
            const float uR = sg.next1D_if(fR && Fd > 0.f);
            const bool isRefl = fR && Fd > 0.f && uR <= Fd;
            wo = isRefl ? v3(-wi.x, -wi.y, wi.z) : refract_dir(wi, cT, eta);
            pdf = (fR && fT) ? (isRefl ? Fd : 1.f - Fd) : 1.f;
            lobe = isRefl ? kLobeDeltaReflection : (m.dispB != 0.f ? kLobeDispersiveDeltaTransmission : kLobeDeltaTransmission);
            weight = rgb_to_spectrum(isRefl ? m.col[0] : m.trans[0], isRefl ? m.col[1] : m.trans[1],
                                     isRefl ? m.col[2] : m.trans[2], lambda);
            valid = !((isRefl && wi.z * wo.z <= 0.f) || (!isRefl && wi.z * wo.z >= 0.f));
            break;
        }
        case PLT_MAT_CONDUCTOR:
        case PLT_MAT_OPAQUE_DIELECTRIC:  // PLTConductorBSDF.slang:101-133, PLTOpaqueDielectricBSDF.slang:89-123
            if (!fR) break;
            wo = v3(-wi.x, -wi.y, wi.z);
            lobe = kLobeDeltaReflection;
            weight = rgb_to_spectrum(m.col[0], m.col[1], m.col[2], lambda);
            valid = true;
            break;
        case PLT_MAT_DIFFUSE: {  // PLTDiffuseBSDF.slang:57-65: draws its direction, then refuses the mask
            (void)sg.next1D();
            (void)sg.next1D();
            break;
        }
        case PLT_MAT_THIN_DIELECTRIC: {  // PLTThinDielectricBSDF.slang:164-183
            float cTd;
            const float Fd = fresnel_dielectric_F(m.extEta, fabsf(wi.z), cTd);
            const float uR = sg.next1D_if(fR && Fd > 0.f);
            const bool isRefl = fR && Fd > 0.f && uR <= Fd;
            wo = isRefl ? v3(-wi.x, -wi.y, wi.z) : -wi;
            weight = 1.f;
            pdf = (fR && fT) ? (isRefl ? Fd : 1.f - Fd) : 1.f;
            lobe = isRefl ? kLobeDeltaReflection : kLobeDeltaTransmission;
            valid = true;
            break;
        }
        case PLT_MAT_MULTILAYER_STACK:  // PLTMultiLayeredStackBSDF.slang:102-115
            if (!fR || wi.z <= 0.f) break;
            wo = v3(-wi.x, -wi.y, wi.z);
            lobe = kLobeDeltaReflection;
            weight = rgb_to_spectrum(m.col[0], m.col[1], m.col[2], lambda);
            valid = wi.z * wo.z > 0.f;
            break;
        case PLT_MAT_COATED_CONDUCTOR:
        case PLT_MAT_COATED_OPAQUE_DIELECTRIC: {  // PLTCoatedOpaqueDielectricBSDF.slang:131-180
            if (!fR || wi.z <= 0.f) break;
            const float coatior = m.extEta + (m.coatIor - m.extEta) * (fminf(fmaxf(m.tau, 0.f), .05f) / .05f);
            const float eta1 = coatior / m.extEta;
            float cT;
            const float Fd = fresnel_dielectric_F(eta1, wi.z, cT);
            const float uC = sg.next1D_if(Fd != .1f);
            const bool isReflCoat = Fd == .1f || uC < Fd;
            weight = 1.f;
            pdf = isReflCoat ? Fd : 1.f - Fd;
            wo = v3(-wi.x, -wi.y, wi.z);
            lobe = kLobeDeltaReflection;
            if (isReflCoat) { valid = true; break; }
            const V3 wiIn = -refract_dir(wi, cT, eta1);
            if (wiIn.z == 0.f) break;
            weight = rgb_to_spectrum(m.col[0], m.col[1], m.col[2], lambda);
            valid = wi.z * wo.z > 0.f;
            break;
        }
        case PLT_MAT_GRATED_CONDUCTOR:           // PLTDiffractionGratedConductorBSDF.slang:115-166 with a hint
        case PLT_MAT_GRATED_OPAQUE_DIELECTRIC: {  // PLTDiffractionGratedOpaqueDielectricBSDF.slang:110-146
            const float lam = 1e-3f * lambda;
            const float a = m.gr.alpha(wi.z, k2Pi / lam);
            const V3 hint = f.to_local(hintW);
            int lx = 0, ly = 0;
            if (a < 1.f) grating_closest_lobe(m.gr, wi, hint, a, lam, lx, ly);
            weight = rgb_to_spectrum(m.col[0], m.col[1], m.col[2], lambda);
            if (lx == 0 && ly == 0) {
                wo = v3(-wi.x, -wi.y, wi.z);
                lobe = kLobeDeltaReflection;
            } else {
                wo = m.gr.diffract(wi, lx, ly, lam);
                lobe = encode_order(lx, ly, kLobeDispersiveDeltaReflection);
            }
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

// manifoldOneStep (Manifolds.slangh:283-333): the angle-difference constraint between the current
// direction w and the specular direction ws in spherical coordinates, and the Newton step of the
// vertex position along the surface. The step runs in FP64 (the directions and their derivatives
// widened from FP32): the constraint, its 2 x 2 Jacobian, the inverse and the step; the result is
// rounded back to FP32 where the walk moves the vertex.
PLT_DEV void walk_sph(V3 w, double& theta, double& phi) {
#pragma clang fp contract(off)
    theta = walk_acos(fmin(fmax(double(w.z), -1.0), 1.0));
    phi = walk_atan2(double(w.y), double(w.x));
    phi += phi < 0.0 ? 6.28318530717958647692 : 0.0;
}
PLT_DEV void walk_sph_derivs(V3 v, V3 ddu, V3 ddv, double out[4]) {
#pragma clang fp contract(off)
    const double vx = v.x, vy = v.y, vz = v.z;
    const double dAcos = -1.0 / sqrt(fmax(0.0, 1.0 - vz * vz));
    const double invNorm = 1.0 / (vx * vx + vy * vy);
    out[0] = dAcos * double(ddu.z);
    out[1] = dAcos * double(ddv.z);
    out[2] = invNorm * (vx * double(ddu.y) - vy * double(ddu.x));
    out[3] = invNorm * (vx * double(ddv.y) - vy * double(ddv.x));
}
PLT_DEV bool manifold_one_step(const PDV& w, const PDV& ws, const PDV& p, float& lenC, V3& stepDir) {
#pragma clang fp contract(off)
    double to, po, ti, pi;
    walk_sph(w.v, to, po);
    walk_sph(ws.v, ti, pi);
    const double dt = to - ti;
    double dp = po - pi;
    dp = fabs(dp) > 3.14159265358979323846 ? dp - (dp > 0.0 ? 6.28318530717958647692 : -6.28318530717958647692) : dp;
    lenC = float(sqrt(dt * dt + dp * dp));
    double a[4], b[4];
    walk_sph_derivs(w.v, w.ddu, w.ddv, a);
    walk_sph_derivs(ws.v, ws.ddu, ws.ddv, b);
    const double j00 = a[0] - b[0], j01 = a[1] - b[1], j10 = a[2] - b[2], j11 = a[3] - b[3];
    // tryInvert (MathHelpers.slang:484-496)
    const double det = j00 * j11 - j01 * j10;
    if (fabs(det) < 4e-6) return false;
    const double id = 1.0 / det;
    const double i00 = j11 * id, i11 = j00 * id, i01 = -j01 * id, i10 = -j10 * id;
    const double dx = i00 * dt + i01 * dp, dy = i10 * dt + i11 * dp;
    stepDir = v3(float(double(p.ddu.x) * dx + double(p.ddv.x) * dy), float(double(p.ddu.y) * dx + double(p.ddv.y) * dy),
                 float(double(p.ddu.z) * dx + double(p.ddv.z) * dy));
    return true;
}

}  // namespace plt
