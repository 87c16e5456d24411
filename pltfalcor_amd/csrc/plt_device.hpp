// Device-side PLT math for CDNA4 (gfx950): sample generator, spectral lookups, Fresnel /
// Mueller algebra, surface PSD lobes and the PLT beam, written for one lane per query with
// everything in VGPRs. Mueller matrices are never materialised as 4x4: a Fresnel Mueller
// matrix is its four distinct entries (m00, m01, m22, m23) and a frame rotation is the pair
// (sin 2theta, cos 2theta) obtained from the frame vectors without atan2/sin/cos.
//
// Reference semantics follow Source/Falcor/Rendering/PLT.slang, Rendering/Materials/Fresnel.slang,
// Rendering/Materials/PLT/{Mueller,GaussianSurface,FractalSurface}.slang (cited per function).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/plt_abi.h"

#define PLT_DEV __device__ __forceinline__

namespace plt {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kPi_2 = 1.57079632679489661923f;
constexpr float kPi_4 = 0.785398163397448309616f;
constexpr float k1_Pi = 0.318309886183790671538f;
constexpr float k2Pi = 6.28318530717958647693f;
constexpr float k1_2Pi = 0.159154943091895335769f;
constexpr float kFltEps = 1.192092896e-07f;
constexpr float kSqrt4Pi = 3.54490770181103205460f;  // sqrt(4 pi), MathConstants M_4PI under sqrt
constexpr float kMinWavelength = 380.f, kMaxWavelength = 720.f;

// Lobe bits (Rendering/Materials/LobeType.slang:36-69)
enum : uint32_t {
    kLobeScatterReflection = 0x02, kLobeDeltaReflection = 0x04, kLobeScatterTransmission = 0x20,
    kLobeDeltaTransmission = 0x40, kLobeDispersiveDeltaTransmission = 0x80, kLobeDelta = 0xCC,
    kLobeNonDelta = 0x33, kLobeReflection = 0x0f, kLobeTransmission = 0xf0
};

struct V3 { float x, y, z; };
PLT_DEV V3 v3(float x, float y, float z) { return V3{x, y, z}; }
PLT_DEV V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PLT_DEV V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PLT_DEV V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
PLT_DEV V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
PLT_DEV float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PLT_DEV V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
PLT_DEV float sq(float v) { return v * v; }
PLT_DEV float sgnf(float v) { return v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f); }
PLT_DEV V3 normalize(V3 a) { return a * (1.f / sqrtf(dot(a, a))); }

// ---------------------------------------------------------------------------------------
// Sample generator: SplitMix64 -> xoshiro128** (UniformSampleGenerator.slang:51-69,
// SplitMix64.slang:54-60, Xoshiro.slang:52-66, BitTricks.slang:45-62). 16 B of state.
// ---------------------------------------------------------------------------------------
PLT_DEV uint32_t spread16(uint32_t x) {
    x &= 0xffffu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}
PLT_DEV uint64_t splitmix_next(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// blockCipherTEA (Utils/Math/HashUtils.slang:59-71): TEA encryption of (v0, v1) with Falcor's
// fixed 128-bit key and 16 cycles; the Tiny generator takes .x as its seed.
PLT_DEV uint32_t tea_seed(uint32_t v0, uint32_t v1) {
    uint32_t sum = 0u;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        sum += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

// Either of the reference's sample generators (SampleGeneratorType.slangh:32-36), chosen per
// launch: `lcg` false = UniformSampleGenerator (xoshiro128** in s0..s3), true =
// TinyUniformSampleGenerator (TinyUniformSampleGenerator.slang:55-66: the Numerical Recipes LCG,
// Pseudorandom/LCG.slang:45-61, with its 32-bit state in s0).
struct Rng {
    uint32_t s0, s1, s2, s3;
    bool lcg;
    PLT_DEV static Rng make(uint32_t px, uint32_t py, uint32_t sampleNumber, bool tiny = false) {
        const uint32_t pix = spread16(px) | (spread16(py) << 1);  // interleave_32bit(pixel)
        if (tiny) return Rng{tea_seed(pix, sampleNumber), 0u, 0u, 0u, true};
        uint64_t st = (uint64_t(sampleNumber) << 32) | uint64_t(pix);
        const uint64_t a = splitmix_next(st), b = splitmix_next(st);
        return Rng{uint32_t(a), uint32_t(a >> 32), uint32_t(b), uint32_t(b >> 32), false};
    }
    PLT_DEV static Rng from_state(uint4 st, bool tiny) { return Rng{st.x, st.y, st.z, st.w, tiny}; }
    PLT_DEV uint32_t next() {
        if (lcg) {
            s0 = 1664525u * s0 + 1013904223u;
            return s0;
        }
        const uint32_t m = s0 * 5u;
        const uint32_t res = ((m << 7) | (m >> 25)) * 9u;
        const uint32_t t = s1 << 9;
        s2 ^= s0; s3 ^= s1; s1 ^= s2; s0 ^= s3;
        s2 ^= t;
        s3 = (s3 << 11) | (s3 >> 21);
        return res;
    }
    // (bits >> 8) * 2^-24: exact in float (SampleGeneratorInterface.slang:46-52)
    PLT_DEV float next1D() { return float(next() >> 8) * 0x1p-24f; }
    // A draw taken only when `take` (the state is unchanged otherwise; the value is then 0), as
    // straight-line selects: sampling decisions of the form `p == 1 || u < p` stay branch-free,
    // a shape that hipcc -O2/-O3 for gfx950 miscompiles when `u` is drawn on a divergent branch
    // (a value hoisted above the branch is clobbered inside it for lanes that rejoin later).
    PLT_DEV float next1D_if(bool take) {
        Rng n = *this;
        const float u = n.next1D();
        s0 = take ? n.s0 : s0; s1 = take ? n.s1 : s1; s2 = take ? n.s2 : s2; s3 = take ? n.s3 : s3;
        return take ? u : 0.f;
    }
};

// ---------------------------------------------------------------------------------------
// binary16 parameters (BasicMaterialData.slang:70-83): hardware RNE conversions.
// ---------------------------------------------------------------------------------------
PLT_DEV float half_lo(uint32_t w) { return float(__builtin_bit_cast(_Float16, uint16_t(w & 0xffffu))); }
PLT_DEV float half_hi(uint32_t w) { return float(__builtin_bit_cast(_Float16, uint16_t(w >> 16))); }
PLT_DEV uint32_t to_half_bits(float f) { return uint32_t(__builtin_bit_cast(uint16_t, _Float16(f))); }

// ---------------------------------------------------------------------------------------
// Spectral tables.
// ---------------------------------------------------------------------------------------
// rgbToSpectrum's 11-bin basis (Utils/Color/SpectrumUtils.slang:106-173), rows:
// white, cyan, magenta, yellow, red, green, blue.
__constant__ const float kRgbBasis[7][11] = {
    {1.0000f, 1.0000f, 0.9999f, 0.9993f, 0.9992f, 0.9998f, 1.0000f, 1.0000f, 1.0000f, 1.0000f, 0.0f},
    {0.9710f, 0.9426f, 1.0007f, 1.0007f, 1.0007f, 1.0007f, 0.1564f, 0.0000f, 0.0000f, 0.0000f, 0.0f},
    {1.0000f, 1.0000f, 0.968f, 0.22295f, 0.0000f, 0.0458f, 0.8369f, 1.0000f, 1.0000f, 0.9959f, 0.0f},
    {0.0001f, 0.0000f, 0.1088f, 0.6651f, 1.0000f, 1.0000f, 0.9996f, 0.9586f, 0.9685f, 0.9840f, 0.0f},
    {0.1012f, 0.0515f, 0.0000f, 0.0000f, 0.0000f, 0.0000f, 0.8325f, 1.0149f, 1.0149f, 1.014f, 0.0f},
    {0.0000f, 0.0000f, 0.0273f, 0.7937f, 1.0000f, 0.9418f, 0.1719f, 0.0000f, 0.0000f, 0.0025f, 0.0f},
    {1.0000f, 1.0000f, 0.8916f, 0.3323f, 0.0000f, 0.0000f, 0.0003f, 0.0369f, 0.0483f, 0.0496f, 0.0f}};

PLT_DEV int rgb_bin(float wavelength) {
    const float t = fminf(fmaxf((wavelength - kMinWavelength) / (kMaxWavelength - kMinWavelength), 0.f), 1.f);
    return int(t * 10.f);
}
// Smits-style RGB -> spectral value at one wavelength (SpectrumUtils.slang:106-173).
PLT_DEV float rgb_to_spectrum(float r, float g, float b, float wavelength) {
    const int i = rgb_bin(wavelength);
    const float w = kRgbBasis[0][i];
    float sd = 0.f;
    if (r <= g && r <= b) {
        sd += w * r;
        if (g <= b) { sd += kRgbBasis[1][i] * (g - r); sd += kRgbBasis[6][i] * (b - g); }
        else { sd += kRgbBasis[1][i] * (b - r); sd += kRgbBasis[5][i] * (g - b); }
    } else if (g <= r && g <= b) {
        sd += w * g;
        if (r <= b) { sd += kRgbBasis[2][i] * (r - g); sd += kRgbBasis[6][i] * (b - r); }
        else { sd += kRgbBasis[2][i] * (b - g); sd += kRgbBasis[4][i] * (r - b); }
    } else {
        sd += w * b;
        if (r <= g) { sd += kRgbBasis[3][i] * (r - b); sd += kRgbBasis[5][i] * (g - r); }
        else { sd += kRgbBasis[3][i] * (g - b); sd += kRgbBasis[4][i] * (r - g); }
    }
    return sd;
}
// White (1,1,1) reflectance: every colour term multiplies zero, leaving the white basis.
PLT_DEV float white_spectrum(float wavelength) { return kRgbBasis[0][rgb_bin(wavelength)]; }

// Scene::evalSpectrum (Scene/Scene.slang:251-269): linear interpolation, zero outside range.
PLT_DEV float eval_spectrum(const plt_spectral_profile* __restrict__ sp, float lambda) {
    const float minl = sp->min_wavelength, maxl = sp->max_wavelength;
    const int bins = int(sp->bins);
    const float x = (lambda - minl) / (maxl - minl) * (float(bins) - 1.0f);
    const float fi = floorf(x);
    const float w = x - fi;
    const int i0 = max(0, min(bins - 1, int(fi)));
    const int i1 = max(0, min(bins - 1, int(fi + 1.f)));
    const float a = sp->pdf[i0], b = sp->pdf[i1];
    return (lambda < minl || lambda > maxl) ? 0.f : sp->intensity * (a + w * (b - a));
}

// ---------------------------------------------------------------------------------------
// Directions.
// ---------------------------------------------------------------------------------------
// perp_stark (Utils/Math/MathHelpers.slang:609-621): cross with the axis of the smallest |component|.
PLT_DEV V3 perp_stark(V3 u) {
    const float ax = fabsf(u.x), ay = fabsf(u.y), az = fabsf(u.z);
    const uint32_t uyx = (ax - ay) < 0.f, uzx = (ax - az) < 0.f, uzy = (ay - az) < 0.f;
    const uint32_t xm = uyx & uzx, ym = (1u ^ xm) & uzy, zm = 1u ^ (xm | ym);
    return normalize(cross(u, v3(float(xm), float(ym), float(zm))));
}
// Shirley concentric disk -> cosine hemisphere (MathHelpers.slang:305-334).
PLT_DEV V3 cosine_hemisphere(float u0, float u1) {
    float x = 2.f * u0 - 1.f, y = 2.f * u1 - 1.f;
    float dx = x, dy = y;
    if (!(x == 0.f && y == 0.f)) {
        float r, phi;
        if (fabsf(x) > fabsf(y)) { r = x; phi = (y / x) * kPi_4; }
        else { r = y; phi = kPi_2 - (x / y) * kPi_4; }
        dx = r * cosf(phi);
        dy = r * sinf(phi);
    }
    return v3(dx, dy, sqrtf(fmaxf(0.f, 1.f - (dx * dx + dy * dy))));
}

// ---------------------------------------------------------------------------------------
// Fresnel (Rendering/Materials/Fresnel.slang:30-73) as the distinct Mueller entries
// (Rendering/Materials/PLT/Mueller.slang:77-88): M = [[a,b,0,0],[b,a,0,0],[0,0,c,d],[0,0,-d,c]].
// ---------------------------------------------------------------------------------------
struct Fres { float a, b, c, d; };
PLT_DEV void apply(const Fres& m, float& s0, float& s1, float& s2, float& s3) {
    const float n0 = m.a * s0 + m.b * s1, n1 = m.b * s0 + m.a * s1;
    const float n2 = m.c * s2 + m.d * s3, n3 = m.c * s3 - m.d * s2;
    s0 = n0; s1 = n1; s2 = n2; s3 = n3;
}
PLT_DEV Fres fres_real(float fs, float fp) {
    const float Rp = fp * fp, Rs = fs * fs;
    return Fres{(Rp + Rs) * 0.5f, (Rp - Rs) * 0.5f, fs * fp, 0.f};
}
// Dielectric amplitude coefficients; returns reflection (refl) or transmission Mueller.
PLT_DEV Fres fresnel_dielectric_mueller(float cosI_, float eta, bool refl) {
    if (eta == 1.f) return refl ? fres_real(0.f, 0.f) : fres_real(1.f, 1.f);
    eta = cosI_ > 0.f ? eta : 1.f / eta;
    const float cosI = fabsf(cosI_);
    const float cosT = sqrtf(fmaxf(0.f, 1.f - (1.f - cosI * cosI) / (eta * eta)));
    const float rs = (cosI - eta * cosT) / (cosI + eta * cosT);
    const float rp = (eta * cosI - cosT) / (eta * cosI + cosT);
    if (refl) return fres_real(rs, rp);
    const float imp = sqrtf(eta * cosT / cosI);
    return fres_real(imp * (rs + 1.f), imp * (rp + 1.f) / eta);
}
// Unpolarised dielectric reflectance .5(rs^2+rp^2) and signed cos of the refracted ray
// (evalFresnelDielectric, Fresnel.slang:75-80).
PLT_DEV float fresnel_dielectric_F(float eta, float cosI_, float& cosT_signed) {
    if (eta == 1.f) { cosT_signed = -cosI_; return 0.f; }
    eta = cosI_ > 0.f ? eta : 1.f / eta;
    const float cosI = fabsf(cosI_);
    const float cosT = sqrtf(fmaxf(0.f, 1.f - (1.f - cosI * cosI) / (eta * eta)));
    const float rs = (cosI - eta * cosT) / (cosI + eta * cosT);
    const float rp = (eta * cosI - cosT) / (eta * cosI + cosT);
    cosT_signed = cosT * -sgnf(cosI_);
    return 0.5f * (rs * rs + rp * rp);
}
// Mueller block products and sums: [[a,b],[b,a]] and [[c,d],[-d,c]] blocks compose like the
// Mueller matrices of Mueller.slang:77-88 (X*Y, X+Y).
PLT_DEV Fres fres_mul(const Fres& x, const Fres& y) {
    return Fres{x.a * y.a + x.b * y.b, x.a * y.b + x.b * y.a, x.c * y.c - x.d * y.d, x.c * y.d + x.d * y.c};
}
PLT_DEV Fres fres_add(const Fres& x, const Fres& y) { return Fres{x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d}; }
// (1 - M)^-1 block by block, zero if a block's determinant is below 4e-6 (Mueller.slang:89-110,
// tryInvert MathHelpers.slang:484-496).
PLT_DEV Fres fres_inv_one_minus(const Fres& m) {
    const float a = 1.f - m.a, c = 1.f - m.c;
    const float det1 = a * a - m.b * m.b, det2 = c * c + m.d * m.d;
    if (fabsf(det1) < 4e-6f || fabsf(det2) < 4e-6f) return Fres{0.f, 0.f, 0.f, 0.f};
    const float i1 = 1.f / det1, i2 = 1.f / det2;
    return Fres{a * i1, m.b * i1, c * i2, m.d * i2};
}
// Conductor amplitude coefficients rs, rp (complex) for a complex relative IoR (n, k)
// (fresnel_conductor, Fresnel.slang:60-73); false (and zeros) for n = k = 0 or cosI < 0.
PLT_DEV bool fresnel_conductor_amp(float cosI, float n, float k, float& rs_r, float& rs_i, float& rp_r, float& rp_i) {
    rs_r = rs_i = rp_r = rp_i = 0.f;
    if ((n == 0.f && k == 0.f) || cosI < 0.f) return false;
    // 1/eta = conj(eta)/|eta|^2 ; t2 = 1 - sin^2 * (1/eta)^2 ; t = sqrt(t2)
    const float inv = 1.f / (n * n + k * k);
    const float ir = n * inv, ii = -k * inv;
    const float s2 = 1.f - cosI * cosI;
    const float t2r = 1.f - s2 * (ir * ir - ii * ii);
    const float t2i = -s2 * (2.f * ir * ii);
    // Principal sqrt without the |z| - Re z cancellation of the textbook form: the smaller
    // component is |Im z| / (2 * larger). Same branch and sign convention (sign(0) = 0).
    const float mag = sqrtf(t2r * t2r + t2i * t2i);
    float tr, ti;
    if (t2r >= 0.f) {
        tr = sqrtf((mag + t2r) * 0.5f);
        ti = tr > 0.f ? sgnf(t2i) * (fabsf(t2i) / (2.f * tr)) : 0.f;
    } else {
        const float b = sqrtf((mag - t2r) * 0.5f);
        tr = b > 0.f ? fabsf(t2i) / (2.f * b) : 0.f;
        ti = sgnf(t2i) * b;
    }
    // eta*t
    const float etr = n * tr - k * ti, eti = n * ti + k * tr;
    // rs = (c - eta t)/(c + eta t)
    const float rsn_r = cosI - etr, rsn_i = -eti, rsd_r = cosI + etr, rsd_i = eti;
    const float rsd = 1.f / (rsd_r * rsd_r + rsd_i * rsd_i);
    rs_r = (rsn_r * rsd_r + rsn_i * rsd_i) * rsd;
    rs_i = (rsn_i * rsd_r - rsn_r * rsd_i) * rsd;
    // rp = (eta c - t)/(eta c + t)
    const float ecr = n * cosI, eci = k * cosI;
    const float rpn_r = ecr - tr, rpn_i = eci - ti, rpd_r = ecr + tr, rpd_i = eci + ti;
    const float rpd = 1.f / (rpd_r * rpd_r + rpd_i * rpd_i);
    rp_r = (rpn_r * rpd_r + rpn_i * rpd_i) * rpd;
    rp_i = (rpn_i * rpd_r - rpn_r * rpd_i) * rpd;
    return true;
}
PLT_DEV Fres fres_from_amp(float rs_r, float rs_i, float rp_r, float rp_i) {
    const float Rp = rp_r * rp_r + rp_i * rp_i, Rs = rs_r * rs_r + rs_i * rs_i;
    // rs * conj(rp)
    const float cr = rs_r * rp_r + rs_i * rp_i, ci = rs_i * rp_r - rs_r * rp_i;
    return Fres{(Rp + Rs) * 0.5f, (Rp - Rs) * 0.5f, cr, ci};
}
// Conductor reflection Mueller block (MuellerFresnelRConductor, Mueller.slang:111-116).
PLT_DEV Fres fresnel_conductor_mueller(float cosI, float n, float k) {
    float a, b, c, d;
    if (!fresnel_conductor_amp(cosI, n, k, a, b, c, d)) return Fres{0.f, 0.f, 0.f, 0.f};
    return fres_from_amp(a, b, c, d);
}

// ---------------------------------------------------------------------------------------
// PLT beam (Rendering/PLT.slang:83-245) with N HWSS components.
// ---------------------------------------------------------------------------------------
template <int N> struct Beam {
    float S[N][4];
    float lambda[N];
    float spdf[N];
    int comps;
    float D00, D01, D11;  // diffusivity (PLTCoherence, PLT.slang:41-78)
    float r_m;            // distance travelled from the source (m)
    V3 pos, r, t;
    bool distant;

    PLT_DEV float r_mm() const { return r_m * 1e+3f; }
    PLT_DEV void set_zero() {
#pragma unroll
        for (int j = 0; j < N; ++j) S[j][0] = S[j][1] = S[j][2] = S[j][3] = 0.f;
    }
    // Rotates the polarisation frame to (newt, newr): Stokes by MuellerT, coherence by
    // R^T D R (PLT.slang:203-215, Mueller.slang:34-53). theta = atan2(y, x) is never formed:
    // sin/cos come from normalising (x, y); rx = sin 2theta, ry = cos 2theta.
    PLT_DEV void rotate_frame(V3 newt, V3 newr) {
        const float y = dot(r, cross(t, newt)), x = dot(t, newt);
        const float rr = x * x + y * y;
        float s, c;
        if (rr > 0.f && rr <= 3.4e38f) {
            const float inv = 1.f / sqrtf(rr);
            s = y * inv;
            c = x * inv;
        } else {
            float th = atan2f(y, x);  // zero / non-finite frames: reference fallback
            th = isfinite(th) ? th : 0.f;
            s = sinf(th);
            c = cosf(th);
        }
        const float rx = 2.f * s * c, ry = 1.f - 2.f * s * s;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (j < comps) {
                const float s1 = S[j][1], s2 = S[j][2];
                S[j][1] = rx * s1 - ry * s2;
                S[j][2] = ry * s1 + rx * s2;
            }
        }
        // D' = R^T D R, R = [[c,-s],[s,c]]
        const float a = D00, b = D01, d = D11;
        const float a0 = a * c + b * s, a1 = -a * s + b * c;  // row 0 of D R
        const float b0 = b * c + d * s, b1 = -b * s + d * c;  // row 1 of D R
        D00 = c * a0 + s * b0;
        D01 = c * a1 + s * b1;
        D11 = -s * a1 + c * b1;
        r = newr;
        t = newt;
    }
    // constructLocalFrame (PLT.slang:217-223): tangent = z x wo, or perp_stark at the pole.
    PLT_DEV void construct_local_frame(V3 wo) {
        V3 nt = v3(-wo.y, wo.x, 0.f);
        const float l2 = nt.x * nt.x + nt.y * nt.y;
        nt = l2 > 1e-8f ? nt * (1.f / sqrtf(l2)) : perp_stark(wo);
        rotate_frame(nt, wo);
    }
    // Transverse rotation Q (PLT.slang:151-155): columns t.xy and (t x r).xy.
    PLT_DEV void Q(float& q00, float& q01, float& q10, float& q11) const {
        const V3 b = cross(t, r);
        q00 = t.x; q01 = b.x; q10 = t.y; q11 = b.y;
    }
    PLT_DEV V3 to_local(V3 v) const { return v3(dot(v, t), dot(v, cross(t, r)), dot(v, r)); }
    // mutualCoherenceAngular (PLT.slang:172-177): invTheta()*r^2 = D.
    PLT_DEV float mutual_coherence_angular(V3 d1, V3 d2) const {
        const float vx = 1.f / fmaxf(kSqrt4Pi * fabsf(d1.x - d2.x), kFltEps);
        const float vy = 1.f / fmaxf(kSqrt4Pi * fabsf(d1.y - d2.y), kFltEps);
        const float rr = r_mm();
        const float s = (1.f / rr) * (1.f / rr) * (rr * rr);
        const float q = vx * (s * D00 * vx + s * D01 * vy) + vy * (s * D01 * vx + s * D11 * vy);
        return expf(-0.5f * (1.f / q));
    }
    // Coherence transform on a non-delta interaction: U = diag(c, 1/c), D' = U^T D U.
    PLT_DEV void scatter_coherence(float woz) {
        const float c = sqrtf(fmaxf(1e-2f, fabsf(woz)));
        const float ic = 1.f / c;
        D00 = c * (D00 * c);
        D01 = c * (D01 * ic);
        D11 = ic * (D11 * ic);
    }
    // Free-space propagation to p (PLT.slang:224-230): distance accumulates unless distant.
    PLT_DEV void propagate(V3 p) {
        if (!distant) {
            const V3 d = p - pos;
            r_m += sqrtf(dot(d, d));
        }
        pos = p;
    }
    // Spatial mutual coherence per HWSS component for a transverse offset v (um)
    // (PLT.slang:165-170): exp(-(k/2pi)^2 v^T (D/r^2) v / 2). A beam that has not propagated
    // (r = 0: distant sources) takes the expression's limit, 0 at any nonzero offset and 1 at none;
    // evaluated as written it is 0 * inf or inf - inf there, NaN or 0 depending on rounding.
    PLT_DEV void mutual_coherence(const float* k, V3 v, float* out) const {
        if (r_m == 0.f) {
            const float c = (v.x == 0.f && v.y == 0.f) ? 1.f : 0.f;
#pragma unroll
            for (int j = 0; j < N; ++j) out[j] = c;
            return;
        }
        const float ir = 1.f / r_mm();
        const float s = ir * ir;
        const float q = v.x * (s * D00 * v.x + s * D01 * v.y) + v.y * (s * D01 * v.x + s * D11 * v.y);
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const float rl = k[j] * k1_2Pi;
            out[j] = expf(-0.5f * (rl * rl * q));
        }
    }
    PLT_DEV void scale(float s) {
#pragma unroll
        for (int j = 0; j < N; ++j) { S[j][0] *= s; S[j][1] *= s; S[j][2] *= s; S[j][3] *= s; }
    }
    PLT_DEV void scale_each(const float* s) {
#pragma unroll
        for (int j = 0; j < N; ++j) { S[j][0] *= s[j]; S[j][1] *= s[j]; S[j][2] *= s[j]; S[j][3] *= s[j]; }
    }
};

// ---------------------------------------------------------------------------------------
// Surface PSD lobes.
// ---------------------------------------------------------------------------------------
// Coherent fraction alpha (GaussianSurface.slang:61-72, FractalSurface.slang:77-88) and the
// scattered fraction 1 - alpha, formed with expm1 so it keeps full precision when alpha -> 1.
PLT_DEV float surface_alpha(float ciz, float coz, float q, float k) {
    const float e = (fabsf(ciz) + fabsf(coz)) * q * k;
    return expf(-(e * e));
}
PLT_DEV float surface_one_minus_alpha(float ciz, float coz, float q, float k) {
    const float e = (fabsf(ciz) + fabsf(coz)) * q * k;
    return -expm1f(-(e * e));
}
// Gaussian PSD convolved with the beam's coherence (GaussianSurface.slang:97-109):
// S = I/sigma2 + Q invTheta Q^T, D = exp(-k^2/2 h^T S^-1 h) k^2 / (2 pi sqrt(det S)).
PLT_DEV float gaussian_eval(float k, float sigma2, float i00, float i01, float i11, float q00, float q01,
                            float q10, float q11, float hx, float hy) {
    // invTheta' = Q invTheta Q^T (invTheta symmetric)
    const float p00 = i00 * q00 + i01 * q01, p01 = i01 * q00 + i11 * q01;  // row 0 of Q I
    const float p10 = i00 * q10 + i01 * q11, p11 = i01 * q10 + i11 * q11;  // row 1 of Q I
    const float is2 = 1.f / sigma2;
    const float s00 = is2 + (p00 * q00 + p01 * q01);
    const float s01 = p00 * q10 + p01 * q11;
    const float s10 = p10 * q00 + p11 * q01;
    const float s11 = is2 + (p10 * q10 + p11 * q11);
    const float det = s00 * s11 - s01 * s10;
    const float idet = 1.f / det;
    const float quad = (hx * (s11 * hx - s01 * hy) + hy * (-s10 * hx + s00 * hy)) * idet;
    return expf(-0.5f * k * k * quad) * (k * k * k1_2Pi / sqrtf(det));
}
// Truncated Box-Muller pdf (MathHelpers.slang:435-450).
PLT_DEV float boxmueller_trunc_pdf(float wx, float wy, float mx, float my, float sigma2) {
    const float mm = mx * mx + my * my;
    const float l = sqrtf(fminf(1.f, mm));
    const float coso = sqrtf(fmaxf(0.f, 1.f - mm));
    const float dx = wx - mx, dy = wy - my;
    const float r2 = dx * dx + dy * dy;
    const float x = expf(-0.5f * r2 / sigma2);
    const float r = sqrtf(r2);
    const float max_phi = (r < kFltEps || l < kFltEps) ? kPi
                          : acosf(fminf(fmaxf((r * r + l * l - 1.f) / (2.f * r * l), -1.f), 1.f));
    return 0.5f * x / (max_phi * sigma2) * coso;
}
// K-correlation PSD pGF (FractalSurface.slang:92-95); gamma = 3 (exponent 2) avoids pow.
PLT_DEV float fractal_pgf(float T, float gamma, float zx, float zy) {
    const float base = 1.f + T * (zx * zx + zy * zy);
    const float e = (gamma + 1.f) * 0.5f;
    const float p = e == 2.f ? base * base : powf(base, e);
    return k1_2Pi * (gamma - 1.f) * T * (1.f / p);
}

// ---------------------------------------------------------------------------------------
// Dielectric (PLTDielectricBSDF.slang:43-279), Gaussian PSD.
// ---------------------------------------------------------------------------------------
struct Dielectric {
    float etaInt, etaExt, B;  // Cauchy: eta(k) = (etaInt + B (k/2pi)^2) / etaExt
    float sigma2, q;
    float spec[3], trans[3];
    PLT_DEV float eta_at(float k) const { const float u = k * k1_2Pi; return (etaInt + B * (u * u)) / etaExt; }
};
// scattered_wo (PLTDielectricBSDF.slang:71-77)
PLT_DEV V3 scattered_wo(V3 wo, float eta) {
    const float s = wo.z > 0.f ? 1.f / eta : eta;
    const float x = wo.x * s, y = wo.y * s;
    const float l2 = x * x + y * y;
    return l2 > 1.f ? v3(0.f, 0.f, 0.f) : v3(x, y, -sgnf(wo.z) * sqrtf(fmaxf(0.f, 1.f - l2)));
}
// refract (PLTDielectricBSDF.slang:66-69)
PLT_DEV V3 refract_dir(V3 wi, float cosT, float eta) {
    const float s = cosT < 0.f ? 1.f / eta : eta;
    return v3(-s * wi.x, -s * wi.y, cosT);
}

// eval in the local frame, after the beam has been rotated to the exitant frame
// (PLTDielectricBSDF.slang:79-137).
template <int N>
PLT_DEV void dielectric_eval(const Dielectric& m, V3 wi, V3 wo, uint32_t lobe, Beam<N>& beam) {
    const bool isRefl = wi.z * wo.z > 0.f;
    const bool isScatter = (lobe & kLobeNonDelta) != 0, isDirect = (lobe & kLobeDelta) != 0;
    const bool hasT = (lobe & kLobeTransmission) != 0, hasR = (lobe & kLobeReflection) != 0;
    if (wi.z * wo.z == 0.f || (!isScatter && !isDirect) || (!hasT && !isRefl) || (!hasR && isRefl)) {
        beam.set_zero();
        return;
    }
    float k[N], eta[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        k[j] = k2Pi / (1e-3f * beam.lambda[j]);
        eta[j] = m.eta_at(k[j]);
    }
    const float* col = isRefl ? m.spec : m.trans;
    const V3 up = isRefl ? wo : scattered_wo(wo, eta[0]);
    if (isScatter) {
        const V3 h = wi + up;
        const V3 mh = normalize(h);
        const float LdotH = fmaxf(0.f, dot(wi, mh));
        float q00, q01, q10, q11;
        beam.Q(q00, q01, q10, q11);
        const float ir = 1.f / beam.r_mm();
        const float sc = (k1_2Pi * ir) * (k1_2Pi * ir);  // invTheta(k=1) = (1/(2 pi r))^2 D
        const float i00 = sc * beam.D00, i01 = sc * beam.D01, i11 = sc * beam.D11;
        const float cosL = sgnf(wi.z) * LdotH;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (j < beam.comps) {
                const float oma = m.q == 0.f ? 0.f : surface_one_minus_alpha(wi.z, up.z, m.q, k[j]);
                const float J = isRefl ? 1.f : sq(wo.z > 0.f ? eta[j] : 1.f / eta[j]);
                const float D = gaussian_eval(k[j], m.sigma2, i00, i01, i11, q00, q01, q10, q11, h.x, h.y);
                const Fres F = fresnel_dielectric_mueller(cosL, eta[j], isRefl);
                const float m00 = rgb_to_spectrum(col[0], col[1], col[2], beam.lambda[j]);
                const float w = m00 * oma * J * fabsf(wi.z) * D;
                apply(F, beam.S[j][0], beam.S[j][1], beam.S[j][2], beam.S[j][3]);
                beam.S[j][0] *= w; beam.S[j][1] *= w; beam.S[j][2] *= w; beam.S[j][3] *= w;
            }
        }
    } else {
        const V3 wiBeam = beam.to_local(wi);
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (j < beam.comps) {
                const float a = m.q == 0.f ? 1.f : surface_alpha(wi.z, up.z, m.q, k[j]);
                const float J = isRefl ? 1.f : sq(wo.z > 0.f ? eta[j] : 1.f / eta[j]);
                const Fres F = fresnel_dielectric_mueller(wi.z, eta[j], isRefl);
                float coh = 1.f;
                if (!isRefl && j != 0) {
                    float cosT;
                    (void)fresnel_dielectric_F(eta[j], wo.z, cosT);
                    coh = beam.mutual_coherence_angular(beam.to_local(refract_dir(wo, cosT, eta[j])), wiBeam);
                }
                const float m00 = rgb_to_spectrum(col[0], col[1], col[2], beam.lambda[j]);
                const float w = m00 * a * coh * J;
                apply(F, beam.S[j][0], beam.S[j][1], beam.S[j][2], beam.S[j][3]);
                beam.S[j][0] *= w; beam.S[j][1] *= w; beam.S[j][2] *= w; beam.S[j][3] *= w;
            }
        }
    }
}
// pdf(wi -> wo) of the non-delta lobes (PLTDielectricBSDF.slang:213-229).
PLT_DEV float dielectric_pdf(const Dielectric& m, V3 wi, V3 wo, float lambda) {
    const float k = k2Pi / (1e-3f * lambda);
    const float oma = m.q == 0.f ? 0.f : surface_one_minus_alpha(wi.z, wi.z, m.q, k);
    const float eta = m.eta_at(k);
    const bool isRefl = wi.z * wo.z > 0.f;
    const V3 w = isRefl ? wo : scattered_wo(wo, eta);
    const V3 mh = normalize(wi + w);
    float cosT;
    const float F = fresnel_dielectric_F(eta, sgnf(wi.z) * dot(wi, mh), cosT);
    const float pF = isRefl ? F : 1.f - F;
    const float k2 = k * k;
    return oma * pF * boxmueller_trunc_pdf(w.x, w.y, -wi.x, -wi.y, (1.f / m.sigma2) / k2);
}

// ---------------------------------------------------------------------------------------
// Conductor (PLTConductorBSDF.slang:44-173), K-correlation PSD, tabulated n,k.
// ---------------------------------------------------------------------------------------
struct Conductor {
    const plt_spectral_profile* etaN;
    const plt_spectral_profile* etaK;
    float extEta, T, gamma, q;
    float spec[3];
};
template <int N>
PLT_DEV void conductor_eval(const Conductor& m, V3 wi, V3 wo, uint32_t lobe, Beam<N>& beam) {
    const bool isScatter = (lobe & kLobeScatterReflection) != 0, isDirect = (lobe & kLobeDeltaReflection) != 0;
    if ((!isScatter && !isDirect) || (lobe & kLobeReflection) == 0) {
        beam.set_zero();
        return;
    }
    const V3 h = wi + wo;
    const V3 mh = normalize(h);
    const float LdotH = fmaxf(0.f, dot(wi, mh));
    const float iext = 1.f / m.extEta;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (j < beam.comps) {
            const float lam = beam.lambda[j];
            const float k = k2Pi / (1e-3f * lam);
            const float n = (m.etaN ? eval_spectrum(m.etaN, lam) : 0.f) * iext;
            const float kk = (m.etaK ? eval_spectrum(m.etaK, lam) : 0.f) * iext;
            float w;
            Fres F;
            if (isScatter) {
                F = fresnel_conductor_mueller(LdotH, n, kk);
                w = surface_one_minus_alpha(wi.z, wo.z, m.q, k) * fabsf(wi.z) * fractal_pgf(m.T, m.gamma, k * h.x, k * h.y);
            } else {
                F = fresnel_conductor_mueller(wi.z, n, kk);
                w = surface_alpha(wi.z, wo.z, m.q, k);
            }
            w *= rgb_to_spectrum(m.spec[0], m.spec[1], m.spec[2], lam);
            apply(F, beam.S[j][0], beam.S[j][1], beam.S[j][2], beam.S[j][3]);
            beam.S[j][0] *= w; beam.S[j][1] *= w; beam.S[j][2] *= w; beam.S[j][3] *= w;
        }
    }
}
// (PLTConductorBSDF.slang:135-142)
PLT_DEV float conductor_pdf(const Conductor& m, V3 wi, V3 wo, float lambda) {
    const float k = k2Pi / (1e-3f * lambda);
    return surface_one_minus_alpha(wi.z, wi.z, m.q, k) * fractal_pgf(m.T, m.gamma, k * (wi.x + wo.x), k * (wi.y + wo.y));
}

}  // namespace plt
