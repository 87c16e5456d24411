// Exact samples for the learned BSDF surrogate, and its evaluation with an exact fallback.
//
// The generator draws a material of one of the three surrogate families (the scattering lobes of
// the coated opaque dielectric, the coated conductor and the grated conductor; the multilayer
// stack is a delta lobe with no scattering lobe and stays exact), a geometry and a coherence
// area per sample, writes the material as a plt_material block (binary16 fields, exactly what the
// path tracer would read) and evaluates its exact Mueller matrix with the path tracer's material
// layer: four evaluations of a one-wavelength beam whose Stokes vector is each unit vector in turn
// give the matrix's columns (the evaluation is linear in the Stokes vector).
#include "plt_internal.hpp"
#include "plt_materials.hpp"
#include "surrogate.hpp"
#include "surrogate_features.hpp"

namespace plt {
namespace {

constexpr int kSBlock = 128;
constexpr int kIn = PLT_SURROGATE_INPUTS, kOut = PLT_SURROGATE_OUTPUTS;
constexpr int kFamCoatedOpaque = kSurCoatedOpaque, kFamCoatedConductor = kSurCoatedConductor, kFamGrated = kSurGrated,
              kFamilies = kSurFamilies;

__device__ __forceinline__ uint16_t h16_bits(float x) { return __builtin_bit_cast(uint16_t, static_cast<_Float16>(x)); }

__global__ void __launch_bounds__(kSBlock) generate_kernel(uint32_t n, uint32_t seed, plt_material* __restrict__ mats,
                                                         float* __restrict__ geo, float* __restrict__ X, float* __restrict__ M,
                                                         const plt_spectral_profile* prof, uint32_t nprof) {
    const uint32_t i = blockIdx.x * kSBlock + threadIdx.x;
    if (i >= n) return;
    Rng sg = Rng::make(i & 0xFFFFu, i >> 16, seed);
    const float u_fam = sg.next1D();
    float u[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) u[k] = sg.next1D();
    const float u_wi0 = sg.next1D(), u_wi1 = sg.next1D(), u_wo0 = sg.next1D(), u_wo1 = sg.next1D();
    const float u_lam = sg.next1D(), u_area = sg.next1D();
    const int fam = min(kFamilies - 1, int(float(kFamilies) * u_fam));

    plt_material pm;
    for (int k = 0; k < 4; ++k) pm.base_color[k] = h16_bits(1.f);
    for (int k = 0; k < 3; ++k) pm.transmission[k] = h16_bits(1.f);
    for (int k = 0; k < 4; ++k) { pm.specular[k] = h16_bits(0.f); pm.data1[k] = 0; pm.data2[k] = 0; }
    pm.ior = h16_bits(1.f);
    pm.emission_profile = pm.ior_n_profile = pm.ior_k_profile = PLT_NO_PROFILE;
    const uint32_t lobe = kLobeScatterReflection;
    switch (fam) {
        case kFamCoatedOpaque:  // albedo U[0.1,0.9]^3, roughness U[0.05,0.6], coat IoR U[1.3,1.8], tau U[0.2,5] um
        case kFamCoatedConductor: {
            pm.type = fam == kFamCoatedOpaque ? PLT_MAT_COATED_OPAQUE_DIELECTRIC : PLT_MAT_COATED_CONDUCTOR;
            if (fam == kFamCoatedOpaque) {
                for (int k = 0; k < 3; ++k) pm.base_color[k] = h16_bits(.1f + .8f * u[k]);
            } else {
                pm.ior_n_profile = 0; pm.ior_k_profile = 1;
            }
            pm.specular[1] = h16_bits(.05f + .55f * u[3]);
            pm.specular[2] = h16_bits(1.f);
            pm.data1[3] = h16_bits(1.3f + .5f * u[4]);
            pm.data1[0] = h16_bits(.2f + 4.8f * u[5]);
            break;
        }
        default: {  // grated conductor: pitch U[0.5,2] um, direction U[0,pi), height U[0.1,1] x 0.5 um, roughness U[0.05,0.6], 3 lobes
            pm.type = PLT_MAT_GRATED_CONDUCTOR;
            pm.ior_n_profile = 0; pm.ior_k_profile = 1;
            pm.data1[0] = h16_bits(.5f + 1.5f * u[0]);
            pm.data1[2] = h16_bits(kPi * u[1]);
            pm.data2[0] = h16_bits(.1f + .9f * u[2]);
            pm.data2[1] = h16_bits(.05f + .55f * u[3]);
            pm.specular[0] = h16_bits(.5f);
            pm.specular[1] = h16_bits(1.f);
            pm.data1[3] = 3u;
            break;
        }
    }
    const Mat m = setup_material(pm, prof, nprof);
    const V3 wi = cosine_hemisphere(u_wi0, u_wi1);
    const V3 wo = cosine_hemisphere(u_wo0, u_wo1);
    const float lambda = 380.f + 340.f * u_lam;
    const float area = expf(2.f * kSurLogAreaMid * u_area);  // log-uniform over [1, 1e8] um^2
    mats[i] = pm;
    const size_t N = n;
    geo[PLT_SG_WI_X * N + i] = wi.x; geo[PLT_SG_WI_Y * N + i] = wi.y; geo[PLT_SG_WI_Z * N + i] = wi.z;
    geo[PLT_SG_WO_X * N + i] = wo.x; geo[PLT_SG_WO_Y * N + i] = wo.y; geo[PLT_SG_WO_Z * N + i] = wo.z;
    geo[PLT_SG_LAMBDA * N + i] = lambda;
    geo[PLT_SG_LOBE * N + i] = __uint_as_float(lobe);
    geo[PLT_SG_AREA * N + i] = area;
    sur_features(m, fam, wi, wo, lambda, area, X + size_t(i) * kIn);
    float* Mi = M + size_t(i) * kOut;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float e[4] = {c == 0 ? 1.f : 0.f, c == 1 ? 1.f : 0.f, c == 2 ? 1.f : 0.f, c == 3 ? 1.f : 0.f};
        float col[4];
        exact_stokes(m, wi, wo, lobe, lambda, area, e, col);
#pragma unroll
        for (int r = 0; r < 4; ++r) Mi[4 * r + c] = col[r];
    }
}

// Exact evaluation: of every sample (mode 1), or (mode 0) of the samples outside the training box
// (stats[32..54) / [54..76)), counting them; the surrogate has written the others.
__global__ void __launch_bounds__(kSBlock) exact_kernel(uint32_t n, const plt_material* __restrict__ mats, const float* __restrict__ geo,
                                                      const float* __restrict__ X, const float* __restrict__ stats,
                                                      const float* __restrict__ Sin, float* __restrict__ Sout, int mode,
                                                      uint32_t* __restrict__ fallbacks, const plt_spectral_profile* prof, uint32_t nprof) {
    const uint32_t i = blockIdx.x * kSBlock + threadIdx.x;
    if (i >= n) return;
    if (mode == 0) {
        bool inside = true;
        for (int k = 0; k < kIn; ++k) {
            const float v = X[size_t(i) * kIn + k];
            inside &= v >= stats[2 * kOut + k] && v <= stats[2 * kOut + kIn + k];
        }
        if (inside) return;
        atomicAdd(fallbacks, 1u);
    }
    const size_t N = n;
    const V3 wi = v3(geo[PLT_SG_WI_X * N + i], geo[PLT_SG_WI_Y * N + i], geo[PLT_SG_WI_Z * N + i]);
    const V3 wo = v3(geo[PLT_SG_WO_X * N + i], geo[PLT_SG_WO_Y * N + i], geo[PLT_SG_WO_Z * N + i]);
    const Mat m = setup_material(mats[i], prof, nprof);
    exact_stokes(m, wi, wo, __float_as_uint(geo[PLT_SG_LOBE * N + i]), geo[PLT_SG_LAMBDA * N + i], geo[PLT_SG_AREA * N + i],
                 Sin + size_t(i) * 4, Sout + size_t(i) * 4);
}

}  // namespace
}  // namespace plt

using namespace plt;

extern "C" {

int plt_surrogate_generate(plt_context* ctx, uint32_t n, uint32_t seed, plt_material* materials, float* geometry, float* X,
                           float* mueller, plt_stream stream) {
    if (!ctx || (n && (!materials || !geometry || !X || !mueller))) return set_error(PLT_EINVAL, "plt_surrogate_generate: bad arguments");
    if (ctx->n_profiles < 2) return set_error(PLT_ESTATE, "plt_surrogate_generate: the context needs conductor profiles 0 (n) and 1 (k)");
    if (n == 0) return PLT_OK;
    hipLaunchKernelGGL(generate_kernel, dim3((n + kSBlock - 1) / kSBlock), dim3(kSBlock), 0, static_cast<hipStream_t>(stream), n, seed,
                       materials, geometry, X, mueller, ctx->d_profiles, ctx->n_profiles);
    return check_launch("surrogate generate_kernel");
}

int plt_surrogate_set_stats(plt_surrogate* s, const float* y_mean, const float* y_std, const float* x_lo, const float* x_hi) {
    if (!s || !y_mean || !y_std || !x_lo || !x_hi) return set_error(PLT_EINVAL, "plt_surrogate_set_stats: bad arguments");
    float h[2 * kOut + 2 * kIn];
    for (int k = 0; k < kOut; ++k) { h[k] = y_mean[k]; h[kOut + k] = y_std[k]; }
    for (int k = 0; k < kIn; ++k) { h[2 * kOut + k] = x_lo[k]; h[2 * kOut + kIn + k] = x_hi[k]; }
    if (hipMemcpy(s->stats, h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) return set_error(PLT_EHIP, "plt_surrogate_set_stats: upload failed");
    s->has_stats = true;
    return PLT_OK;
}

int plt_surrogate_eval(plt_surrogate* s, uint32_t n, const plt_material* materials, const float* geometry, const float* X,
                       const float* stokes_in, float* stokes_out, int mode, uint32_t* fallbacks, plt_stream stream) {
    if (!s || (n && (!materials || !geometry || !X || !stokes_in || !stokes_out)) || (mode != 0 && mode != 1))
        return set_error(PLT_EINVAL, "plt_surrogate_eval: bad arguments");
    if (mode == 0 && !s->has_stats) return set_error(PLT_ESTATE, "plt_surrogate_eval: plt_surrogate_set_stats has not been called");
    if (n == 0) return PLT_OK;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    uint32_t* dcount = s->counter;
    int rc = PLT_OK;
    if (hipMemsetAsync(dcount, 0, sizeof(uint32_t), st) != hipSuccess) return set_error(PLT_EHIP, "plt_surrogate_eval: clear failed");
    if (n > (1u << 30)) return set_error(PLT_EINVAL, "plt_surrogate_eval: n too large");
    if (mode == 0 && (rc = surrogate_infer(s, int(n), X, nullptr, stokes_in, stokes_out, st)) < 0) return rc;
    hipLaunchKernelGGL(exact_kernel, dim3((n + kSBlock - 1) / kSBlock), dim3(kSBlock), 0, st, n, materials, geometry, X, s->stats, stokes_in,
                       stokes_out, mode, dcount, s->ctx->d_profiles, s->ctx->n_profiles);
    if ((rc = check_launch("surrogate exact_kernel")) < 0) return rc;
    if (fallbacks) {
        uint32_t c = 0;
        if (hipMemcpyAsync(&c, dcount, sizeof(c), hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
            return set_error(PLT_EHIP, "plt_surrogate_eval: read-back failed");
        *fallbacks += c;
    }
    return PLT_OK;
}

}  // extern "C"
