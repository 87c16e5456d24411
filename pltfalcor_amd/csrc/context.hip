// C ABI implementation: context lifetime, error reporting, spectral-profile construction and
// the batched BSDF entry points (include/plt_abi.h).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include "plt_internal.hpp"


namespace plt {
namespace {
thread_local std::string g_last_error;

// Host spectral pipeline: piecewise-linear table -> uniform samples (>= 5 nm bins) -> profile
// clamped to 380-720 nm (Utils/Color/Spectrum.h:88-123, SampledSpectrum.h:64-74,308-317,
// Scene/SceneBuilder.cpp:911-960). Interpolation weights use (1-t)a + t b as the host code does.
inline float mix(float a, float b, float t) { return (1.f - t) * a + t * b; }

struct Tabulated {
    const float* wl;
    const float* v;
    uint32_t n;
    float at(float w) const {
        if (w < wl[0] || w > wl[n - 1]) return 0.f;
        const float* it = std::lower_bound(wl, wl + n, w);
        if (it == wl) return v[0];
        const uint32_t i = uint32_t(it - wl) - 1;
        return mix(v[i], v[i + 1], (w - wl[i]) / (wl[i + 1] - wl[i]));
    }
};
}  // namespace

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
int check_launch(const char* kernel) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(PLT_EHIP, std::string(kernel) + ": " + hipGetErrorString(e));
    return PLT_OK;
}
}  // namespace plt

using plt::set_error;

namespace {

// genSpectralProfile (SceneBuilder.cpp:911-960) of a SampledSpectrum: `ns` samples spaced
// uniformly over [lo, hi], evaluated by linear interpolation between them.
int gen_profile(const float* smp, size_t ns, float lo, float hi, int compute_icdf, plt_spectral_profile* out) {
    auto sampled = [&](float l) {
        if (l < lo || l > hi) return 0.f;
        const float x = ((l - lo) / (hi - lo)) * (float(ns) - 1.0f);
        const size_t i = size_t(std::floor(x));
        if (i + 1 >= ns) return smp[ns - 1];
        return plt::mix(smp[i], smp[i + 1], x - float(i));
    };
    plt_spectral_profile sp;
    std::memset(&sp, 0, sizeof(sp));
    sp.intensity = 1.f;
    sp.min_wavelength = std::max(lo, 380.f);
    sp.max_wavelength = std::min(hi, 720.f);
    sp.bins = std::min<uint32_t>(PLT_SPECTRAL_PROFILE_MAX_BINS,
                                 uint32_t(float(ns) * (sp.max_wavelength - sp.min_wavelength) / (hi - lo) + .5f));
    if (sp.bins > 0) {
        const float r = sp.bins > 1 ? 1.f / float(sp.bins - 1) : 1.f;
        float sum = 0.f;
        for (uint32_t i = 0; i < sp.bins; ++i) {
            sp.pdf[i] = sampled(plt::mix(sp.min_wavelength, sp.max_wavelength, float(i) * r));
            sum += sp.pdf[i];
        }
        if (sum >= 1.192092896e-07f && compute_icdf) {
            for (uint32_t i = 0; i < sp.bins; ++i) sp.pdf[i] /= sum;
            sp.intensity = sum;
            float cdfa = 0.f, cdfb = sp.pdf[0];
            uint32_t pdfi = 0;
            for (uint32_t ic = 0; ic < sp.bins; ++ic) {
                const float cdf = float(ic) * r;
                while (cdf >= cdfb && pdfi + 1 < sp.bins) {
                    cdfa = cdfb;
                    ++pdfi;
                    cdfb += sp.pdf[pdfi];
                }
                const float l = std::min(1.f, (cdf - cdfa) / (cdfb - cdfa));
                sp.icdf[ic] = plt::mix(float(std::max<uint32_t>(1, pdfi)) - 1.f, float(pdfi), l) * r;
            }
        }
    }
    *out = sp;
    return PLT_OK;
}

}  // namespace

extern "C" {

const char* plt_last_error(void) { return plt::g_last_error.c_str(); }
int plt_abi_version(void) { return 103; }

int plt_context_create(int device, plt_context** out) {
    if (!out) return set_error(PLT_EINVAL, "plt_context_create: out is NULL");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return set_error(PLT_EHIP, "plt_context_create: no HIP device");
    if (device < 0 || device >= count) return set_error(PLT_EINVAL, "plt_context_create: device index out of range");
    if (hipSetDevice(device) != hipSuccess) return set_error(PLT_EHIP, "plt_context_create: hipSetDevice failed");
    plt_context* c = new (std::nothrow) plt_context();
    if (!c) return set_error(PLT_ENOMEM, "plt_context_create: out of host memory");
    c->device = device;
    *out = c;
    return PLT_OK;
}

int plt_context_destroy(plt_context* ctx) {
    if (!ctx) return PLT_OK;
    if (ctx->d_profiles) (void)hipFree(ctx->d_profiles);
    delete ctx;
    return PLT_OK;
}

int plt_spectral_profile_build(uint32_t npts, const float* wavelengths, const float* values, int compute_icdf,
                               plt_spectral_profile* out) {
    if (npts < 2 || !wavelengths || !values || !out) return set_error(PLT_EINVAL, "plt_spectral_profile_build: bad arguments");
    for (uint32_t i = 1; i < npts; ++i)
        if (!(wavelengths[i] >= wavelengths[i - 1]))
            return set_error(PLT_EINVAL, "plt_spectral_profile_build: wavelengths must be non-decreasing");
    const plt::Tabulated tab{wavelengths, values, npts};
    const float lo = wavelengths[0], hi = wavelengths[npts - 1];
    if (!(hi > lo)) return set_error(PLT_EINVAL, "plt_spectral_profile_build: empty wavelength range");
    float least = hi - lo;
    for (uint32_t i = 1; i < npts; ++i) least = std::min(least, wavelengths[i] - wavelengths[i - 1]);
    const float bin = std::max(5.f, least / 2.f);
    const size_t ns = std::max<size_t>(1, size_t((hi - lo) / bin + .5f));
    std::vector<float> smp(ns);
    for (size_t i = 0; i < ns; ++i) smp[i] = tab.at(lo + (hi - lo) / float(ns - 1) * float(i));
    return gen_profile(smp.data(), ns, lo, hi, compute_icdf, out);
}

int plt_spectral_profile_build_sampled(uint32_t n, float lambda_start, float lambda_end, const float* values,
                                       int compute_icdf, plt_spectral_profile* out) {
    if (n < 1 || !values || !out || !(lambda_end > lambda_start))
        return set_error(PLT_EINVAL, "plt_spectral_profile_build_sampled: bad arguments");
    return gen_profile(values, n, lambda_start, lambda_end, compute_icdf, out);
}

int plt_context_set_spectral_profiles(plt_context* ctx, const plt_spectral_profile* profiles, uint32_t count) {
    if (!ctx || (count && !profiles)) return set_error(PLT_EINVAL, "plt_context_set_spectral_profiles: bad arguments");
    if (ctx->d_profiles) {
        (void)hipFree(ctx->d_profiles);
        ctx->d_profiles = nullptr;
        ctx->n_profiles = 0;
    }
    if (count == 0) return PLT_OK;
    for (uint32_t i = 0; i < count; ++i)
        if (profiles[i].bins > PLT_SPECTRAL_PROFILE_MAX_BINS)
            return set_error(PLT_EINVAL, "plt_context_set_spectral_profiles: bins > 128");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_error(PLT_EHIP, "hipSetDevice failed");
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(plt_spectral_profile) * count) != hipSuccess)
        return set_error(PLT_ENOMEM, "plt_context_set_spectral_profiles: hipMalloc failed");
    if (hipMemcpy(p, profiles, sizeof(plt_spectral_profile) * count, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p);
        return set_error(PLT_EHIP, "plt_context_set_spectral_profiles: hipMemcpy failed");
    }
    ctx->d_profiles = static_cast<plt_spectral_profile*>(p);
    ctx->n_profiles = count;
    return PLT_OK;
}

int plt_sample_generator_stream_typed(int type, uint32_t n_streams, uint32_t width, uint32_t sample_number, uint32_t draws,
                                      uint32_t* out, plt_stream stream) {
    if (type != 0 && type != 1) return set_error(PLT_EINVAL, "plt_sample_generator_stream: type must be 0 (tiny) or 1 (uniform)");
    if (width == 0 || (n_streams && draws && !out)) return set_error(PLT_EINVAL, "plt_sample_generator_stream: bad arguments");
    if (uint64_t(n_streams) * draws > (uint64_t(1) << 34)) return set_error(PLT_EINVAL, "plt_sample_generator_stream: too large");
    return plt::launch_sample_generator(n_streams, width, sample_number, draws, type == 0, out, static_cast<hipStream_t>(stream));
}

int plt_sample_generator_stream(uint32_t n_streams, uint32_t width, uint32_t sample_number, uint32_t draws, uint32_t* out,
                                plt_stream stream) {
    return plt_sample_generator_stream_typed(1, n_streams, width, sample_number, draws, out, stream);
}

int plt_bsdf_generate_queries(plt_context* ctx, uint32_t n, uint32_t seed, uint32_t* queries, plt_stream stream) {
    if (!ctx || (n && !queries)) return set_error(PLT_EINVAL, "plt_bsdf_generate_queries: bad arguments");
    return plt::launch_generate_queries(n, seed, queries, static_cast<hipStream_t>(stream));
}

int plt_bsdf_eval_queries(plt_context* ctx, uint32_t n, const uint32_t* queries, float* results, plt_stream stream) {
    if (!ctx || (n && (!queries || !results))) return set_error(PLT_EINVAL, "plt_bsdf_eval_queries: bad arguments");
    return plt::launch_eval_queries(n, queries, results, ctx->d_profiles, ctx->n_profiles, static_cast<hipStream_t>(stream));
}

}  // extern "C"
