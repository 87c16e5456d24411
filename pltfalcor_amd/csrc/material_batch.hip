// Batched PLT material interface (Rendering/Materials/IMaterialInstance.slang:122-184): eval +
// pdf + lambdaPdf, and sample, of any plt_material for SoA query planes, one lane per query.
// Queries are in a local frame with N = +z (the shading frame of a surface whose tangent frame is
// the identity); the incident beam is sourced along -wi with the given coherence area, as the
// BSDF batch does. The device code is the path tracer's own material layer.
#include "plt_internal.hpp"
#include "plt_materials.hpp"

namespace plt {
namespace {

constexpr int kMatBlock = 256;

__global__ void __launch_bounds__(kMatBlock)
material_eval_kernel(uint32_t n, const float* __restrict__ q, float* __restrict__ r, const plt_material* __restrict__ mats,
                     uint32_t nmats, const plt_spectral_profile* prof, uint32_t nprof) {
    const uint32_t i = blockIdx.x * kMatBlock + threadIdx.x;
    if (i >= n) return;
    auto Q = [&](int p) { return q[size_t(p) * n + i]; };
    const uint32_t mi = __float_as_uint(Q(PLT_MQ_MAT));
    const uint32_t lobe = __float_as_uint(Q(PLT_MQ_LOBE));
    float* out = r + i;
    if (mi >= nmats || !material_supported(mats[mi])) {  // unknown material or radial grating: zeros
        for (int p = 0; p < PLT_MR_PLANES; ++p) out[size_t(p) * n] = 0.f;
        return;
    }
    const Mat m = setup_material(mats[mi], prof, nprof);
    Frame sd;
    sd.T = v3(1.f, 0.f, 0.f);
    sd.B = v3(0.f, 1.f, 0.f);
    sd.N = v3(0.f, 0.f, 1.f);
    const V3 wi = v3(Q(PLT_MQ_WI_X), Q(PLT_MQ_WI_Y), Q(PLT_MQ_WI_Z));
    const V3 wo = v3(Q(PLT_MQ_WO_X), Q(PLT_MQ_WO_Y), Q(PLT_MQ_WO_Z));
    Beam<4> beam;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int c = 0; c < 4; ++c) beam.S[j][c] = Q(PLT_MQ_S + 4 * j + c);
        beam.lambda[j] = Q(PLT_MQ_LAMBDA + j);
        beam.spdf[j] = 1.f;
    }
    beam.comps = 4;
    const float l0 = 1e-3f * beam.lambda[0];
    const float d = kPi * (l0 * l0) / Q(PLT_MQ_AREA);
    beam.D00 = d; beam.D01 = 0.f; beam.D11 = d;
    beam.r_m = 1e-3f;
    beam.pos = v3(0.f, 0.f, 0.f);
    beam.r = -wi;
    beam.t = perp_stark(beam.r);
    beam.distant = false;
    float lp[4];
    mat_lambda_pdf<4>(m, sd, wi, wo, lobe, beam, lp);
    mat_eval<4>(m, sd, wi, wo, lobe, beam);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int c = 0; c < 4; ++c) out[size_t(PLT_MR_S + 4 * j + c) * n] = beam.S[j][c];
        out[size_t(PLT_MR_LAMBDA_PDF + j) * n] = lp[j];
    }
    out[size_t(PLT_MR_PDF) * n] = mat_pdf(m, sd, wo, wi, beam.lambda[0]);
}

__global__ void __launch_bounds__(kMatBlock)
material_sample_kernel(uint32_t n, uint32_t frame, const float* __restrict__ q, float* __restrict__ r,
                       const plt_material* __restrict__ mats, uint32_t nmats, const plt_spectral_profile* prof,
                       uint32_t nprof) {
    const uint32_t i = blockIdx.x * kMatBlock + threadIdx.x;
    if (i >= n) return;
    auto Q = [&](int p) { return q[size_t(p) * n + i]; };
    const uint32_t mi = __float_as_uint(Q(PLT_MSQ_MAT));
    float* out = r + i;
    BsdfSample bs{v3(0.f, 0.f, 0.f), 0.f, 0.f, 0u};
    bool valid = false;
    if (mi < nmats && material_supported(mats[mi])) {
        const Mat m = setup_material(mats[mi], prof, nprof);
        Frame sd;
        sd.T = v3(1.f, 0.f, 0.f);
        sd.B = v3(0.f, 1.f, 0.f);
        sd.N = v3(0.f, 0.f, 1.f);
        Rng sg = Rng::make(i, 0u, frame);
        valid = mat_sample(m, sd, v3(Q(PLT_MSQ_WI_X), Q(PLT_MSQ_WI_Y), Q(PLT_MSQ_WI_Z)), sg, Q(PLT_MSQ_LAMBDA), bs);
    }
    out[size_t(PLT_MSR_WO_X) * n] = bs.wo.x;
    out[size_t(PLT_MSR_WO_Y) * n] = bs.wo.y;
    out[size_t(PLT_MSR_WO_Z) * n] = bs.wo.z;
    out[size_t(PLT_MSR_PDF) * n] = bs.pdf;
    out[size_t(PLT_MSR_WEIGHT) * n] = bs.weight;
    out[size_t(PLT_MSR_LOBE) * n] = __uint_as_float(bs.lobe);
    out[size_t(PLT_MSR_VALID) * n] = __uint_as_float(valid ? 1u : 0u);
}

}  // namespace
}  // namespace plt

extern "C" {

int plt_material_eval_batch(plt_context* ctx, const plt_material* materials, uint32_t n_materials, uint32_t n,
                            const float* queries, float* results, plt_stream stream) {
    if (!ctx || (n && (!queries || !results)) || (n_materials && !materials))
        return plt::set_error(PLT_EINVAL, "plt_material_eval_batch: bad arguments");
    if (n == 0) return PLT_OK;
    hipLaunchKernelGGL(plt::material_eval_kernel, dim3((n + plt::kMatBlock - 1) / plt::kMatBlock), dim3(plt::kMatBlock), 0,
                       static_cast<hipStream_t>(stream), n, queries, results, materials, n_materials, ctx->d_profiles,
                       ctx->n_profiles);
    return plt::check_launch("material_eval_kernel");
}

int plt_material_sample_batch(plt_context* ctx, const plt_material* materials, uint32_t n_materials, uint32_t n,
                              uint32_t frame, const float* queries, float* results, plt_stream stream) {
    if (!ctx || (n && (!queries || !results)) || (n_materials && !materials))
        return plt::set_error(PLT_EINVAL, "plt_material_sample_batch: bad arguments");
    if (n == 0) return PLT_OK;
    hipLaunchKernelGGL(plt::material_sample_kernel, dim3((n + plt::kMatBlock - 1) / plt::kMatBlock), dim3(plt::kMatBlock),
                       0, static_cast<hipStream_t>(stream), n, frame, queries, results, materials, n_materials,
                       ctx->d_profiles, ctx->n_profiles);
    return plt::check_launch("material_sample_kernel");
}

}  // extern "C"
