// C entry points of the PLT oracle (liboracle.so), called from tests/ and bench.py's
// cpu_baseline leg through ctypes. TEST INFRASTRUCTURE ONLY: never linked by the product.
//
// The BSDF batch layout mirrors the product's documented SoA contract: 16 planes of n
// 32-bit words per query batch, 5 planes of n floats per result batch.
#include "plt_oracle.hpp"
#include "plt_oracle_geom.hpp"
#include <omp.h>

using namespace pltoracle;

namespace {
enum Plane {
    Q_WI_X = 0, Q_WI_Y, Q_WI_Z, Q_WO_X, Q_WO_Y, Q_WO_Z, Q_LAMBDA, Q_MAT, Q_PARAM,
    Q_ROUGH_Q, Q_ROUGH_L, Q_COH_AREA, Q_S0, Q_S1, Q_S2, Q_S3, Q_PLANES
};
enum MatType { MAT_CONDUCTOR = 3, MAT_DIELECTRIC = 4 };  // MaterialTypes.slang:36-53

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// Synthetic BSDF batch generator: query i draws from UniformSampleGenerator(pixel=(i&0xFFFF, i>>16),
// sampleNumber=seed) in a fixed order (14 draws). Arithmetic is plain IEEE float.
void c0_generate_one(uint32_t i, uint32_t seed, float* q, size_t n) {
    UniformSampleGenerator sg(i & 0xFFFFu, i >> 16, seed);
    const float u_type = sampleNext1D<float>(sg);
    const float u_ior = sampleNext1D<float>(sg);
    const float u_q = sampleNext1D<float>(sg);
    const float u_l = sampleNext1D<float>(sg);
    const float u_ac = sampleNext1D<float>(sg);
    const V2<float> u_wi = sampleNext2D<float>(sg);
    const V2<float> u_wo = sampleNext2D<float>(sg);
    const float u_side = sampleNext1D<float>(sg);
    const float u_lam = sampleNext1D<float>(sg);
    const V2<float> u_st = sampleNext2D<float>(sg);
    const float u_dop = sampleNext1D<float>(sg);

    const bool dielectric = u_type < 0.5f;
    float pdfc;
    V3<float> wi = sample_cosine_hemisphere_concentric(u_wi, pdfc);
    V3<float> wo = sample_cosine_hemisphere_concentric(u_wo, pdfc);
    if (dielectric && u_side < 0.5f) wo.z = -wo.z;
    const float lambda = 380.f + 400.f * u_lam;
    const float z = 1.f - 2.f * u_st.x;
    const float phi = 6.28318530717958647693f * u_st.y;
    const float s = std::sqrt(std::fmax(0.f, 1.f - z * z));
    const float dop = u_dop;

    uint32_t mat, param;
    if (dielectric) {
        const uint16_t ior16 = f32_to_f16_bits(1.3f + 0.5f * u_ior);
        const float B = 0.52f * (f16_bits_to_f32(ior16) - 1.f) / 100.f;  // Abbe number 100
        mat = uint32_t(MAT_DIELECTRIC) << 28;
        param = uint32_t(ior16) | (uint32_t(f32_to_f16_bits(B)) << 16);
    } else {
        mat = (uint32_t(MAT_CONDUCTOR) << 28) | 0u /*n profile*/ | (1u << 14) /*k profile*/;
        param = uint32_t(f32_to_f16_bits(1.f)) | (uint32_t(f32_to_f16_bits(3.f)) << 16);
    }
    q[Q_WI_X * n + i] = wi.x; q[Q_WI_Y * n + i] = wi.y; q[Q_WI_Z * n + i] = wi.z;
    q[Q_WO_X * n + i] = wo.x; q[Q_WO_Y * n + i] = wo.y; q[Q_WO_Z * n + i] = wo.z;
    q[Q_LAMBDA * n + i] = lambda;
    q[Q_MAT * n + i] = bitsf(mat);
    q[Q_PARAM * n + i] = bitsf(param);
    q[Q_ROUGH_Q * n + i] = 0.01f + 0.29f * u_q;
    q[Q_ROUGH_L * n + i] = 1.f + 19.f * u_l;
    q[Q_COH_AREA * n + i] = 1.f + 99.f * u_ac;
    q[Q_S0 * n + i] = 1.f;
    q[Q_S1 * n + i] = dop * s * std::cos(phi);
    q[Q_S2 * n + i] = dop * s * std::sin(phi);
    q[Q_S3 * n + i] = dop * z;
}

// One query: beam sourced along -wi with isotropic diffusivity chosen so that
// coherenceArea(k) equals the query's coherence area (r = 1 mm), identity shading frame,
// eval on the NonDelta lobe (as next-event estimation does), pdf(wi -> wo).
template <class T>
void c0_eval_one(const float* q, size_t n, size_t i, const SpectralProfile* prof, uint32_t nprof, T* out) {
    auto F = [&](int p) { return T(q[size_t(p) * n + i]); };
    const uint32_t mat = fbits(q[size_t(Q_MAT) * n + i]);
    const uint32_t param = fbits(q[size_t(Q_PARAM) * n + i]);
    const uint32_t type = mat >> 28;
    const V3<T> wi = {F(Q_WI_X), F(Q_WI_Y), F(Q_WI_Z)};
    const V3<T> wo = {F(Q_WO_X), F(Q_WO_Y), F(Q_WO_Z)};
    const T lambda = F(Q_LAMBDA);
    const T rq = F(Q_ROUGH_Q), rl = F(Q_ROUGH_L), Ac = F(Q_COH_AREA);

    PLTBeam<T, 1> beam;
    beam.L[0] = F(Q_S0); beam.L1[0] = F(Q_S1); beam.L2[0] = F(Q_S2); beam.L3[0] = F(Q_S3);
    beam.wavelengths[0] = lambda;
    beam.spectralPDFs[0] = 1;
    beam.hwss_comps = 1;
    const T d = K<T>::PI * sqr(T(1e-3) * lambda) / Ac;
    beam.Theta.D[0] = d; beam.Theta.D[1] = 0; beam.Theta.D[2] = d;
    beam.Theta.r_m = T(1e-3);
    beam.posW = {0, 0, 0};
    beam.r = -wi;
    beam.t = perp_stark(beam.r);
    beam.distant = false;

    Frame<T> sd{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    T pdf = 0;
    if (type == MAT_DIELECTRIC) {
        DielectricBxDF<T> b;
        const float ior = f16_bits_to_f32(uint16_t(param & 0xffffu));
        const float B = f16_bits_to_f32(uint16_t(param >> 16));
        b.gfs.sigma2 = sqr(rl);
        b.gfs.q = rq;
        // PLTDielectricMaterial.slang:49-51: eta = max(1, IoR - 2.897*B)
        b.etaInt = T(std::fmax(1.f, ior - 2.897f * B));
        b.etaExt = T(1);
        b.dispersionB = T(B);
        bsdfEval<T, 1>(b, sd, wi, wo, Lobe_NonDelta, beam);
        pdf = b.pdf(wi, wo, lambda);
    } else if (type == MAT_CONDUCTOR) {
        ConductorBxDF<T> b;
        const uint32_t nid = mat & 0x3fffu, kid = (mat >> 14) & 0x3fffu;
        b.etaN = nid < nprof ? &prof[nid] : nullptr;
        b.etaK = kid < nprof ? &prof[kid] : nullptr;
        b.extEta = T(f16_bits_to_f32(uint16_t(param & 0xffffu)));
        b.gfs.gamma = hmax(T(1.1), T(f16_bits_to_f32(uint16_t(param >> 16))));
        b.gfs.Tc = hmin(T(5000), sqr(rl));
        b.gfs.q = rq;
        bsdfEval<T, 1>(b, sd, wi, wo, Lobe_NonDelta, beam);
        pdf = b.pdf(wi, wo, lambda);
    } else {
        beam.setZero();
    }
    out[0 * n + i] = beam.L[0];
    out[1 * n + i] = beam.L1[0];
    out[2 * n + i] = beam.L2[0];
    out[3 * n + i] = beam.L3[0];
    out[4 * n + i] = pdf;
}
}  // namespace

extern "C" {

uint32_t oracle_interleave_32bit(uint32_t x, uint32_t y) { return interleave_32bit(x, y); }

void oracle_splitmix64_seq(uint64_t seed, uint32_t n, uint64_t* out) {
    SplitMix64 s{seed};
    for (uint32_t i = 0; i < n; ++i) out[i] = s.next64();
}
void oracle_xoshiro128ss_seq(const uint32_t* state, uint32_t n, uint32_t* out) {
    Xoshiro128SS x;
    for (int j = 0; j < 4; ++j) x.s[j] = state[j];
    for (uint32_t i = 0; i < n; ++i) out[i] = x.next();
}
// Raw uint32 stream of UniformSampleGenerator(pixel, sampleNumber).
void oracle_usg_seq(uint32_t px, uint32_t py, uint32_t sampleNumber, uint32_t n, uint32_t* out) {
    UniformSampleGenerator sg(px, py, sampleNumber);
    for (uint32_t i = 0; i < n; ++i) out[i] = sg.next();
}
// Raw uint32 streams of either generator for a block of pixels: stream s = pixel (x0 + s % w,
// y0 + s / w), `draws` values each, out[s * draws + d]. type 0 = TinyUniformSampleGenerator,
// 1 = UniformSampleGenerator.
void oracle_sg_streams(int type, uint32_t x0, uint32_t y0, uint32_t w, uint32_t nStreams, uint32_t sampleNumber,
                       uint32_t draws, uint32_t* out) {
    for (uint32_t s = 0; s < nStreams; ++s) {
        const uint32_t px = x0 + s % w, py = y0 + s / w;
        uint32_t* o = out + size_t(s) * draws;
        if (type == 0) {
            TinyUniformSampleGenerator sg(px, py, sampleNumber);
            for (uint32_t d = 0; d < draws; ++d) o[d] = sg.next();
        } else {
            UniformSampleGenerator sg(px, py, sampleNumber);
            for (uint32_t d = 0; d < draws; ++d) o[d] = sg.next();
        }
    }
}
// TEA encryption of (v0, v1) with a 128-bit key and `iterations` cycles (blockCipherTEA's loop).
void oracle_tea(uint32_t v0, uint32_t v1, const uint32_t* key, uint32_t iterations, uint32_t* out) {
    teaEncrypt(v0, v1, key, iterations);
    out[0] = v0;
    out[1] = v1;
}
uint16_t oracle_f32_to_f16(float f) { return f32_to_f16_bits(f); }
float oracle_f16_to_f32(uint16_t h) { return f16_bits_to_f32(h); }

uint32_t oracle_spectral_profile_size(void) { return uint32_t(sizeof(SpectralProfile)); }

// PiecewiseLinearSpectrum(wl,val) -> SampledSpectrum(5 nm) -> genSpectralProfile.
int oracle_build_profile(uint32_t npts, const float* wl, const float* val, int computeICdf, void* out) {
    if (npts < 2) return -1;
    PiecewiseLinearSpectrum p;
    p.wl.assign(wl, wl + npts);
    p.val.assign(val, val + npts);
    SampledSpectrumF ss(p);
    SpectralProfile sp = genSpectralProfile(ss, computeICdf != 0);
    std::memcpy(out, &sp, sizeof(sp));
    return 0;
}
float oracle_eval_spectrum_f32(const void* profile, float lambda) {
    return evalSpectrum<float>(static_cast<const SpectralProfile*>(profile), lambda);
}

void oracle_c0_generate(uint32_t n, uint32_t seed, float* q, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < int64_t(n); ++i) c0_generate_one(uint32_t(i), seed, q, n);
}
void oracle_c0_eval_f32(uint32_t n, const float* q, const void* profiles, uint32_t nprof, float* out, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    const SpectralProfile* p = static_cast<const SpectralProfile*>(profiles);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < int64_t(n); ++i) c0_eval_one<float>(q, n, size_t(i), p, nprof, out);
}
void oracle_c0_eval_f64(uint32_t n, const float* q, const void* profiles, uint32_t nprof, double* out, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    const SpectralProfile* p = static_cast<const SpectralProfile*>(profiles);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < int64_t(n); ++i) c0_eval_one<double>(q, n, size_t(i), p, nprof, out);
}

// --- scalar entry points for closed-form / invariant tests (FP64 and FP32) ---
void oracle_fresnel_dielectric_f64(double cosI, double eta, double* out5) {
    double rs, rp, ts, tp, ct;
    fresnel_dielectric(cosI, eta, rs, rp, ts, tp, ct);
    out5[0] = rs; out5[1] = rp; out5[2] = ts; out5[3] = tp; out5[4] = ct;
}
void oracle_fresnel_conductor_f64(double cosI, double n, double k, double* out4) {
    V2<double> rs, rp;
    fresnel_conductor(cosI, V2<double>{n, k}, rs, rp);
    out4[0] = rs.x; out4[1] = rs.y; out4[2] = rp.x; out4[3] = rp.y;
}
// Mueller rotation applied to a Stokes vector: frame (n1,t1) -> t2.
void oracle_mueller_rotate_f64(const double* n1, const double* t1, const double* t2, const double* s_in, double* s_out) {
    double si, co;
    M4<double> Tm = MuellerT(V3<double>{n1[0], n1[1], n1[2]}, V3<double>{t1[0], t1[1], t1[2]}, V3<double>{0, 0, 0},
                             V3<double>{t2[0], t2[1], t2[2]}, si, co);
    V4<double> r = mul(Tm, V4<double>{s_in[0], s_in[1], s_in[2], s_in[3]});
    s_out[0] = r.x; s_out[1] = r.y; s_out[2] = r.z; s_out[3] = r.w;
}
double oracle_fractal_pgf_f64(double T, double gamma, double zx, double zy) {
    FractalSurface<double> f; f.Tc = T; f.gamma = gamma; f.q = 0.1;
    return f.pGF(V2<double>{zx, zy});
}
double oracle_gaussian_eval_f64(double sigma2, double k, const double* invTheta4, const double* Q4, double hx, double hy) {
    GaussianSurface<double> g; g.sigma2 = sigma2; g.q = 0.1;
    return g.eval(k, m2<double>(invTheta4[0], invTheta4[1], invTheta4[2], invTheta4[3]),
                  m2<double>(Q4[0], Q4[1], Q4[2], Q4[3]), V3<double>{hx, hy, 0});
}
double oracle_boxmueller_trunc_pdf_f64(double wx, double wy, double mx, double my, double sigma2) {
    return boxmueller_truncated_pdf(V2<double>{wx, wy}, V2<double>{mx, my}, sigma2);
}
double oracle_rgb_to_spectrum_f64(double r, double g, double b, double lambda) {
    return Spectrum<double>::rgbToSpectrum(V3<double>{r, g, b}, lambda);
}
void oracle_wavelength_to_xyz_f64(double lambda, double* xyz) {
    V3<double> v = Spectrum<double>::wavelengthToXYZ(lambda);
    xyz[0] = v.x; xyz[1] = v.y; xyz[2] = v.z;
}
void oracle_hwss_wavelengths_f32(float hero, int n, float* out) { Spectrum<float>::hwssWavelengths(hero, n, out); }

// Surface sampling helpers for sample/pdf consistency tests: draws `count` samples of the
// Fractal (type 1) or Gaussian (type 0) lobe around wi; writes wo (3 floats) and pdf.
void oracle_surface_sample_f64(int type, double p0, double p1, double q, const double* wi3, double k, uint32_t seed,
                               uint32_t count, double* wo3, double* pdf) {
    UniformSampleGenerator sg(seed & 0xffffu, seed >> 16, 7u);
    const V3<double> wi = {wi3[0], wi3[1], wi3[2]};
    for (uint32_t i = 0; i < count; ++i) {
        V3<double> w; double p;
        if (type == 1) { FractalSurface<double> f; f.Tc = p0; f.gamma = p1; f.q = q; w = f.sample(wi, sg, k, p); }
        else { GaussianSurface<double> g; g.sigma2 = p0; g.q = q; w = g.sample(wi, sg, k, p); }
        wo3[3 * i] = w.x; wo3[3 * i + 1] = w.y; wo3[3 * i + 2] = w.z; pdf[i] = p;
    }
}
double oracle_surface_pdf_f64(int type, double p0, double p1, double q, const double* wi3, const double* wo3, double k) {
    const V3<double> wi = {wi3[0], wi3[1], wi3[2]}, wo = {wo3[0], wo3[1], wo3[2]};
    if (type == 1) { FractalSurface<double> f; f.Tc = p0; f.gamma = p1; f.q = q; return f.pdf(wi, wo, k); }
    GaussianSurface<double> g; g.sigma2 = p0; g.q = q; return g.pdf(wi, wo, k);
}

// Ray casting: rays are 8 planes of n floats (o, d, tmin, tmax). mode 0 writes 4 planes (tri id,
// b1, b2, t as raw 32-bit words), mode 1 writes one plane (any hit 1/0). brute != 0 tests every
// triangle; otherwise the oracle's own median-split BVH is used.
int oracle_trace(const float* positions, uint32_t ntri, const uint32_t* indices, uint32_t n, const float* rays,
                 uint32_t* hits, int mode, int brute, int threads) {
    OMesh m;
    m.pos = positions; m.idx = indices; m.ntri = ntri;
    OBvh bvh;
    if (!brute) bvh.build(m);
    if (threads > 0) omp_set_num_threads(threads);
    const size_t N = n;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        ORay r;
        for (int a = 0; a < 3; ++a) { r.o[a] = rays[a * N + size_t(i)]; r.d[a] = rays[(3 + a) * N + size_t(i)]; }
        r.tmin = rays[6 * N + size_t(i)];
        r.tmax = rays[7 * N + size_t(i)];
        OHit h;
        if (brute) {
            if (mode == 1) {
                const WoopPrep p = woop_prep(r);
                for (uint32_t t = 0; t < ntri; ++t) {
                    float tt, b1, b2;
                    if (woop_intersect(r, p, m.vtx(t, 0), m.vtx(t, 1), m.vtx(t, 2), r.tmin, r.tmax, tt, b1, b2)) { h.tri = t; break; }
                }
            } else h = closest_brute(m, r);
        } else h = bvh.closest(m, r, mode == 1);
        if (mode == 1) hits[size_t(i)] = h.tri != 0xFFFFFFFFu ? 1u : 0u;
        else {
            hits[0 * N + size_t(i)] = h.tri;
            std::memcpy(&hits[1 * N + size_t(i)], &h.b1, 4);
            std::memcpy(&hits[2 * N + size_t(i)], &h.b2, 4);
            std::memcpy(&hits[3 * N + size_t(i)], &h.t, 4);
        }
    }
    return 0;
}

void oracle_compute_ray_origin(const float* pos, const float* n, float* out) { compute_ray_origin(pos, n, out); }

}  // extern "C"
