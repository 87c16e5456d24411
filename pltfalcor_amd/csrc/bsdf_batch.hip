// Batched BSDF queries on MI355X: synthetic batch generator and the eval + pdf kernel.
//
// Layout: a batch is 16 SoA planes of n words (plt_abi.h), so every plane load of a wave is one
// contiguous 256 B segment. One lane per query; spectral profiles stay in global memory
// (2 x 544 B per conductor, L1/L2 resident).
#include "plt_device.hpp"
#include "plt_internal.hpp"

namespace plt {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ float as_f(uint32_t u) { return __builtin_bit_cast(float, u); }
__device__ __forceinline__ uint32_t as_u(float f) { return __builtin_bit_cast(uint32_t, f); }

// Synthetic query generator. Linear arithmetic is kept uncontracted so that every field that
// does not pass through cos/sin is bit-identical to a strict-IEEE host computation.
__global__ void __launch_bounds__(kBlock) generate_queries_kernel(uint32_t n, uint32_t seed, uint32_t* __restrict__ q) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    Rng sg = Rng::make(i & 0xFFFFu, i >> 16, seed);
    const float u_type = sg.next1D();
    const float u_ior = sg.next1D();
    const float u_q = sg.next1D();
    const float u_l = sg.next1D();
    const float u_ac = sg.next1D();
    const float u_wi0 = sg.next1D(), u_wi1 = sg.next1D();
    const float u_wo0 = sg.next1D(), u_wo1 = sg.next1D();
    const float u_side = sg.next1D();
    const float u_lam = sg.next1D();
    const float u_st0 = sg.next1D(), u_st1 = sg.next1D();
    const float u_dop = sg.next1D();

    const bool dielectric = u_type < 0.5f;
    const V3 wi = cosine_hemisphere(u_wi0, u_wi1);
    V3 wo = cosine_hemisphere(u_wo0, u_wo1);
    if (dielectric && u_side < 0.5f) wo.z = -wo.z;
    const float z = 1.f - 2.f * u_st0;
    const float phi = k2Pi * u_st1;
    const float s = sqrtf(fmaxf(0.f, 1.f - z * z));

    uint32_t mat, param;
    if (dielectric) {
        const uint32_t ior16 = to_half_bits(1.3f + 0.5f * u_ior);
        const float B = 0.52f * (half_lo(ior16) - 1.f) / 100.f;  // Abbe number 100
        mat = uint32_t(PLT_MAT_DIELECTRIC) << 28;
        param = ior16 | (to_half_bits(B) << 16);
    } else {
        mat = (uint32_t(PLT_MAT_CONDUCTOR) << 28) | 0u | (1u << 14);
        param = to_half_bits(1.f) | (to_half_bits(3.f) << 16);
    }
    const size_t N = n;
    q[PLT_Q_WI_X * N + i] = as_u(wi.x);
    q[PLT_Q_WI_Y * N + i] = as_u(wi.y);
    q[PLT_Q_WI_Z * N + i] = as_u(wi.z);
    q[PLT_Q_WO_X * N + i] = as_u(wo.x);
    q[PLT_Q_WO_Y * N + i] = as_u(wo.y);
    q[PLT_Q_WO_Z * N + i] = as_u(wo.z);
    q[PLT_Q_LAMBDA * N + i] = as_u(380.f + 400.f * u_lam);
    q[PLT_Q_MAT * N + i] = mat;
    q[PLT_Q_PARAM * N + i] = param;
    q[PLT_Q_ROUGH_Q * N + i] = as_u(0.01f + 0.29f * u_q);
    q[PLT_Q_ROUGH_L * N + i] = as_u(1.f + 19.f * u_l);
    q[PLT_Q_COH_AREA * N + i] = as_u(1.f + 99.f * u_ac);
    q[PLT_Q_S0 * N + i] = as_u(1.f);
    q[PLT_Q_S1 * N + i] = as_u(u_dop * s * cosf(phi));
    q[PLT_Q_S2 * N + i] = as_u(u_dop * s * sinf(phi));
    q[PLT_Q_S3 * N + i] = as_u(u_dop * z);
}

// One query: identity shading frame, beam arriving along -wi with isotropic diffusivity
// d = pi (lambda_um)^2 / A_c at r = 1 mm (coherence area A_c), rotated into the exitant frame,
// evaluated on the NonDelta lobe; pdf(wi -> wo).
__device__ __forceinline__ void eval_one(const uint32_t* __restrict__ q, float* __restrict__ res, size_t N, size_t i,
                                         const plt_spectral_profile* __restrict__ prof, uint32_t nprof) {
    const V3 wi = v3(as_f(q[PLT_Q_WI_X * N + i]), as_f(q[PLT_Q_WI_Y * N + i]), as_f(q[PLT_Q_WI_Z * N + i]));
    const V3 wo = v3(as_f(q[PLT_Q_WO_X * N + i]), as_f(q[PLT_Q_WO_Y * N + i]), as_f(q[PLT_Q_WO_Z * N + i]));
    const float lambda = as_f(q[PLT_Q_LAMBDA * N + i]);
    const uint32_t mat = q[PLT_Q_MAT * N + i];
    const uint32_t param = q[PLT_Q_PARAM * N + i];
    const float rq = as_f(q[PLT_Q_ROUGH_Q * N + i]);
    const float rl = as_f(q[PLT_Q_ROUGH_L * N + i]);
    const float Ac = as_f(q[PLT_Q_COH_AREA * N + i]);

    Beam<1> beam;
    beam.S[0][0] = as_f(q[PLT_Q_S0 * N + i]);
    beam.S[0][1] = as_f(q[PLT_Q_S1 * N + i]);
    beam.S[0][2] = as_f(q[PLT_Q_S2 * N + i]);
    beam.S[0][3] = as_f(q[PLT_Q_S3 * N + i]);
    beam.lambda[0] = lambda;
    beam.spdf[0] = 1.f;
    beam.comps = 1;
    const float lum = 1e-3f * lambda;
    const float d = kPi * (lum * lum) / Ac;
    beam.D00 = d; beam.D01 = 0.f; beam.D11 = d;
    beam.r_m = 1e-3f;
    beam.pos = v3(0.f, 0.f, 0.f);
    beam.r = -wi;
    beam.t = perp_stark(beam.r);
    beam.distant = false;
    beam.construct_local_frame(wo);

    const uint32_t type = mat >> 28;
    float pdf = 0.f;
    if (type == PLT_MAT_DIELECTRIC) {
        Dielectric m;
        const float ior = half_lo(param), B = half_hi(param);
        m.etaInt = fmaxf(1.f, ior - 2.897f * B);  // PLTDielectricMaterial.slang:49-51
        m.etaExt = 1.f;
        m.B = B;
        m.sigma2 = rl * rl;
        m.q = rq;
        m.spec[0] = m.spec[1] = m.spec[2] = 1.f;
        m.trans[0] = m.trans[1] = m.trans[2] = 1.f;
        dielectric_eval<1>(m, wi, wo, kLobeNonDelta, beam);
        pdf = dielectric_pdf(m, wi, wo, lambda);
    } else if (type == PLT_MAT_CONDUCTOR) {
        Conductor m;
        const uint32_t nid = mat & 0x3fffu, kid = (mat >> 14) & 0x3fffu;
        m.etaN = nid < nprof ? prof + nid : nullptr;
        m.etaK = kid < nprof ? prof + kid : nullptr;
        m.extEta = half_lo(param);
        m.gamma = fmaxf(1.1f, half_hi(param));
        m.T = fminf(5000.f, rl * rl);
        m.q = rq;
        m.spec[0] = m.spec[1] = m.spec[2] = 1.f;
        conductor_eval<1>(m, wi, wo, kLobeNonDelta, beam);
        pdf = conductor_pdf(m, wi, wo, lambda);
    } else {
        beam.set_zero();
    }
    res[PLT_R_S0 * N + i] = beam.S[0][0];
    res[PLT_R_S1 * N + i] = beam.S[0][1];
    res[PLT_R_S2 * N + i] = beam.S[0][2];
    res[PLT_R_S3 * N + i] = beam.S[0][3];
    res[PLT_R_PDF * N + i] = pdf;
}

__global__ void __launch_bounds__(kBlock) eval_queries_kernel(uint32_t n, const uint32_t* __restrict__ q,
                                                                float* __restrict__ res,
                                                                const plt_spectral_profile* __restrict__ prof,
                                                                uint32_t nprof) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    eval_one(q, res, n, i, prof, nprof);
}

__global__ void __launch_bounds__(kBlock) sample_generator_kernel(uint32_t n, uint32_t width, uint32_t sn, uint32_t draws,
                                                                    bool tiny, uint32_t* __restrict__ out) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    Rng g = Rng::make(s % width, s / width, sn, tiny);
    for (uint32_t d = 0; d < draws; ++d) out[size_t(d) * n + s] = g.next();
}

}  // namespace

int launch_generate_queries(uint32_t n, uint32_t seed, uint32_t* q, hipStream_t st) {
    if (n == 0) return PLT_OK;
    hipLaunchKernelGGL(generate_queries_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, n, seed, q);
    return check_launch("generate_queries_kernel");
}
int launch_eval_queries(uint32_t n, const uint32_t* q, float* res, const plt_spectral_profile* prof, uint32_t nprof,
                        hipStream_t st) {
    if (n == 0) return PLT_OK;
    hipLaunchKernelGGL(eval_queries_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, n, q, res, prof, nprof);
    return check_launch("eval_queries_kernel");
}
int launch_sample_generator(uint32_t n, uint32_t width, uint32_t sn, uint32_t draws, bool tiny, uint32_t* out, hipStream_t st) {
    if (n == 0 || draws == 0) return PLT_OK;
    hipLaunchKernelGGL(sample_generator_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, n, width, sn, draws, tiny, out);
    return check_launch("sample_generator_kernel");
}

}  // namespace plt
