// Temporal accumulation of a frame sequence (Source/RenderPasses/AccumulatePass): the pass that
// averages PLTPT.color (Double) and PLTPT.albedo / PLTPT.normal (Single) over frames in the
// reference's render graph (scripts/PLTPathTracer.py:19-24,36-38).
//
// Host logic: frame counting, reset triggers and the overflow modes of AccumulatePass::execute /
// accumulate (AccumulatePass.cpp:144-301). Device: one kernel per (precision, moving-average,
// input format), each thread one pixel, 16-byte loads and stores, HBM-bound
// (Double: 16 B in + 32 B sum in + 32 B sum out + 16 B out = 96 B per pixel).
//
// Arithmetic follows Accumulate.cs.slang operation by operation with FP contraction off, so the
// output is bit-identical to a scalar IEEE restatement:
//  - weights: `1.0 / (gAccumCount + 1)` has an unsuffixed (float) literal and a uint operand, so
//    it is a float division; in Double mode the float weight is then widened to double
//    (Accumulate.cs.slang:84,157);
//  - Single: sum = last + cur, output = sum * w; moving average: lerp(last, cur, w) =
//    last + w * (cur - last) (:84-102);
//  - SingleCompensated (Kahan): y = cur - c, s' = s + y, output = s' / (count + 1),
//    c' = (s' - s) - y (:123-135);
//  - Double: sum += (double)cur, output = (float)(sum * w); moving average:
//    sum = sum * (1 - w) + (double)cur * w (:157-188).
// Cleared state (prepareAccumulation clears the buffers when the frame count is 0,
// AccumulatePass.cpp:403-407) is folded into the kernel: a `first` launch reads zeros instead of
// the state buffer, so no clear pass runs.
#include <hip/hip_runtime.h>
#include <cstring>
#include <string>
#include "plt_internal.hpp"

using plt::set_error;

namespace {

#pragma clang fp contract(off)

enum : int { kDouble = 0, kSingle = 1, kSingleCompensated = 2 };

template <int Fmt>
__device__ __forceinline__ float4 load_input(const void* src, uint32_t i) {
    if constexpr (Fmt == PLT_ACCUM_INPUT_FLOAT) {
        return reinterpret_cast<const float4*>(src)[i];
    } else if constexpr (Fmt == PLT_ACCUM_INPUT_UINT) {
        const uint4 u = reinterpret_cast<const uint4*>(src)[i];
        return make_float4(float(u.x), float(u.y), float(u.z), float(u.w));
    } else {
        const int4 s = reinterpret_cast<const int4*>(src)[i];
        return make_float4(float(s.x), float(s.y), float(s.z), float(s.w));
    }
}

struct AccumArgs {
    const void* src;
    float* dst;
    void* state;        // Double: double4 per pixel; Single: float4 sum; Compensated: float4 sum, float4 corr
    uint32_t n;         // pixels
    uint32_t count;     // frames accumulated before this one (gAccumCount)
    float wf;           // 1.0f / float(count + 1)
    float denom;        // float(count + 1)
    int first;          // state is cleared: read zeros
};

__device__ __forceinline__ float ema(float last, float cur, float w) { return last + w * (cur - last); }

template <int Mode, bool Ema, int Fmt>
__global__ __launch_bounds__(256) void accumulate_kernel(AccumArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    const float4 cur = load_input<Fmt>(a.src, i);
    float4 out;
    if constexpr (Mode == kDouble) {
        double4* st = reinterpret_cast<double4*>(a.state);
        double4 s = a.first ? make_double4(0.0, 0.0, 0.0, 0.0) : st[i];
        const double w = double(a.wf);
        if constexpr (Ema) {
            const double om = 1.0 - w;
            s.x = s.x * om + double(cur.x) * w;
            s.y = s.y * om + double(cur.y) * w;
            s.z = s.z * om + double(cur.z) * w;
            s.w = s.w * om + double(cur.w) * w;
            out = make_float4(float(s.x), float(s.y), float(s.z), float(s.w));
        } else {
            s.x += double(cur.x);
            s.y += double(cur.y);
            s.z += double(cur.z);
            s.w += double(cur.w);
            out = make_float4(float(s.x * w), float(s.y * w), float(s.z * w), float(s.w * w));
        }
        st[i] = s;
    } else if constexpr (Mode == kSingle) {
        float4* st = reinterpret_cast<float4*>(a.state);
        const float4 last = a.first ? make_float4(0.f, 0.f, 0.f, 0.f) : st[i];
        if constexpr (Ema) {
            out = make_float4(ema(last.x, cur.x, a.wf), ema(last.y, cur.y, a.wf), ema(last.z, cur.z, a.wf),
                              ema(last.w, cur.w, a.wf));
            st[i] = out;
        } else {
            const float4 s = make_float4(last.x + cur.x, last.y + cur.y, last.z + cur.z, last.w + cur.w);
            out = make_float4(s.x * a.wf, s.y * a.wf, s.z * a.wf, s.w * a.wf);
            st[i] = s;
        }
    } else {
        float4* st = reinterpret_cast<float4*>(a.state);
        const float4 s = a.first ? make_float4(0.f, 0.f, 0.f, 0.f) : st[2 * i];
        const float4 c = a.first ? make_float4(0.f, 0.f, 0.f, 0.f) : st[2 * i + 1];
        const float4 y = make_float4(cur.x - c.x, cur.y - c.y, cur.z - c.z, cur.w - c.w);
        const float4 sn = make_float4(s.x + y.x, s.y + y.y, s.z + y.z, s.w + y.w);
        out = make_float4(__fdiv_rn(sn.x, a.denom), __fdiv_rn(sn.y, a.denom), __fdiv_rn(sn.z, a.denom),
                          __fdiv_rn(sn.w, a.denom));
        st[2 * i] = sn;
        st[2 * i + 1] = make_float4((sn.x - s.x) - y.x, (sn.y - s.y) - y.y, (sn.z - s.z) - y.z, (sn.w - s.w) - y.w);
    }
    reinterpret_cast<float4*>(a.dst)[i] = out;
}

// Output = input converted to float (the disabled pass: a blit for float input,
// AccumulatePass.cpp:222-226; the gAccumulate = false branch for integer input).
template <int Fmt>
__global__ __launch_bounds__(256) void passthrough_kernel(const void* src, float* dst, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    reinterpret_cast<float4*>(dst)[i] = load_input<Fmt>(src, i);
}

template <int Mode, bool Ema>
void launch_mode(int fmt, const AccumArgs& a, uint32_t blocks, hipStream_t st) {
    switch (fmt) {
        case PLT_ACCUM_INPUT_FLOAT: accumulate_kernel<Mode, Ema, PLT_ACCUM_INPUT_FLOAT><<<blocks, 256, 0, st>>>(a); break;
        case PLT_ACCUM_INPUT_UINT: accumulate_kernel<Mode, Ema, PLT_ACCUM_INPUT_UINT><<<blocks, 256, 0, st>>>(a); break;
        default: accumulate_kernel<Mode, Ema, PLT_ACCUM_INPUT_SINT><<<blocks, 256, 0, st>>>(a); break;
    }
}

struct OptionSpec {
    const char* key;
    double def;
};
// Keys and defaults of AccumulatePass (AccumulatePass.h:106-114, AccumulatePass.cpp:63-71).
// outputSize 0 = IOSize::Default (RenderPassHelpers.h:45-53); outputFormat 0 = Unknown (the pass
// then writes RGBA32Float, AccumulatePass.cpp:137).
const OptionSpec kOptions[] = {
    {"enabled", 1}, {"outputSize", 0}, {"autoReset", 1}, {"precisionMode", kSingle},
    {"maxFrameCount", 0}, {"overflowMode", PLT_ACCUM_OVERFLOW_STOP}, {"outputFormat", 0},
};
constexpr int kNumOptions = int(sizeof(kOptions) / sizeof(kOptions[0]));

int option_index(const char* key) {
    if (!key) return -1;
    for (int i = 0; i < kNumOptions; ++i)
        if (std::strcmp(key, kOptions[i].key) == 0) return i;
    return -1;
}

}  // namespace

struct plt_accumulate {
    plt_context* ctx = nullptr;
    bool enabled = true;
    bool auto_reset = true;
    int precision = kSingle;
    uint32_t max_frames = 0;
    int overflow = PLT_ACCUM_OVERFLOW_STOP;
    uint32_t frame_count = 0;
    uint32_t width = 0, height = 0;  // mFrameDim
    void* d_state = nullptr;         // 32 B per pixel, enough for every mode
    size_t state_pixels = 0;
};

extern "C" {

int plt_accumulate_create(plt_context* ctx, plt_accumulate** out) {
    if (!out) return set_error(PLT_EINVAL, "plt_accumulate_create: out is NULL");
    auto* a = new (std::nothrow) plt_accumulate();
    if (!a) return set_error(PLT_ENOMEM, "plt_accumulate_create: out of host memory");
    a->ctx = ctx;
    *out = a;
    return PLT_OK;
}

int plt_accumulate_destroy(plt_accumulate* a) {
    if (!a) return PLT_OK;
    if (a->d_state) {
        (void)hipSetDevice(a->ctx->device);
        (void)hipFree(a->d_state);
    }
    delete a;
    return PLT_OK;
}

const char* plt_accumulate_option_name(uint32_t i) { return i < uint32_t(kNumOptions) ? kOptions[i].key : nullptr; }

int plt_accumulate_set_option(plt_accumulate* a, const char* key, double v) {
    if (!a || !key) return set_error(PLT_EINVAL, "plt_accumulate_set_option: bad arguments");
    const int k = option_index(key);
    if (k < 0) return set_error(PLT_WARN_UNKNOWN_KEY, std::string("Unknown field '") + key + "' in AccumulatePass dictionary.");
    const std::string s = key;
    if (s == "enabled") return plt_accumulate_set_enabled(a, v != 0.0);
    if (s == "autoReset") {
        a->auto_reset = v != 0.0;
    } else if (s == "precisionMode") {
        if (v != kDouble && v != kSingle && v != kSingleCompensated)
            return set_error(PLT_EINVAL, "AccumulatePass: precisionMode must be 0 (Double), 1 (Single) or 2 (SingleCompensated)");
        a->precision = int(v);
        a->frame_count = 0;  // mode change resets (AccumulatePass.cpp:321-326)
    } else if (s == "maxFrameCount") {
        if (!(v >= 0.0 && v <= 4294967295.0) || v != double(uint32_t(v)))
            return set_error(PLT_EINVAL, "AccumulatePass: maxFrameCount must be an unsigned integer");
        a->max_frames = uint32_t(v);
        a->frame_count = 0;  // setMaxAccumulatedFrames resets (AccumulatePass.h:67)
    } else if (s == "overflowMode") {
        if (v != PLT_ACCUM_OVERFLOW_STOP && v != PLT_ACCUM_OVERFLOW_RESET && v != PLT_ACCUM_OVERFLOW_EMA)
            return set_error(PLT_EINVAL, "AccumulatePass: overflowMode must be 0 (Stop), 1 (Reset) or 2 (EMA)");
        a->overflow = int(v);
        a->frame_count = 0;
    } else if (s == "outputSize") {
        if (v != 0.0) return set_error(PLT_ENOTIMPL, "AccumulatePass: only outputSize Default (0) is built (output = input size)");
    } else if (s == "outputFormat") {
        if (v != 0.0) return set_error(PLT_ENOTIMPL, "AccumulatePass: only the default output format (RGBA32Float) is built");
    }
    return PLT_OK;
}

int plt_accumulate_get_option(const plt_accumulate* a, const char* key, double* v) {
    if (!a || !key || !v) return set_error(PLT_EINVAL, "plt_accumulate_get_option: bad arguments");
    const std::string s = key;
    if (s == "enabled") *v = a->enabled;
    else if (s == "autoReset") *v = a->auto_reset;
    else if (s == "precisionMode") *v = a->precision;
    else if (s == "maxFrameCount") *v = a->max_frames;
    else if (s == "overflowMode") *v = a->overflow;
    else if (s == "outputSize" || s == "outputFormat") *v = 0.0;
    else return set_error(PLT_WARN_UNKNOWN_KEY, std::string("Unknown field '") + key + "' in AccumulatePass dictionary.");
    return PLT_OK;
}

int plt_accumulate_set_enabled(plt_accumulate* a, int enabled) {
    if (!a) return set_error(PLT_EINVAL, "plt_accumulate_set_enabled: NULL pass");
    if ((enabled != 0) != a->enabled) {
        a->enabled = enabled != 0;
        a->frame_count = 0;
    }
    return PLT_OK;
}

int plt_accumulate_reset(plt_accumulate* a) {
    if (!a) return set_error(PLT_EINVAL, "plt_accumulate_reset: NULL pass");
    a->frame_count = 0;
    return PLT_OK;
}

int plt_accumulate_get_frame_count(const plt_accumulate* a, uint32_t* count) {
    if (!a || !count) return set_error(PLT_EINVAL, "plt_accumulate_get_frame_count: bad arguments");
    *count = a->frame_count;
    return PLT_OK;
}

int plt_accumulate_execute(plt_accumulate* a, const void* src, int src_format, uint32_t width, uint32_t height,
                           float* dst, uint32_t refresh_flags, plt_stream stream) {
    if (!a) return set_error(PLT_EINVAL, "plt_accumulate_execute: NULL pass");
    if (!a->ctx) return set_error(PLT_ESTATE, "plt_accumulate_execute: pass created without a context");
    if (!src || !dst || width == 0 || height == 0)
        return set_error(PLT_EINVAL, "plt_accumulate_execute: bad arguments");
    if (src_format != PLT_ACCUM_INPUT_FLOAT && src_format != PLT_ACCUM_INPUT_UINT && src_format != PLT_ACCUM_INPUT_SINT)
        return set_error(PLT_EINVAL, "plt_accumulate_execute: src_format must be PLT_ACCUM_INPUT_FLOAT, _UINT or _SINT");
    if ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15u)
        return set_error(PLT_EINVAL, "plt_accumulate_execute: src and dst must be 16-byte aligned");
    const uint64_t n64 = uint64_t(width) * height;
    if (n64 > (1ull << 31)) return set_error(PLT_EINVAL, "plt_accumulate_execute: frame larger than 2^31 pixels");
    const uint32_t n = uint32_t(n64);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (hipSetDevice(a->ctx->device) != hipSuccess) return set_error(PLT_EHIP, "plt_accumulate_execute: hipSetDevice failed");

    // Refresh flags / scene changes (AccumulatePass.cpp:146-174): the caller reports them.
    if (a->auto_reset && refresh_flags != 0) a->frame_count = 0;
    // Overflow (AccumulatePass.cpp:176-189).
    if (a->max_frames > 0 && a->frame_count == a->max_frames) {
        if (a->overflow == PLT_ACCUM_OVERFLOW_STOP) return PLT_OK;
        if (a->overflow == PLT_ACCUM_OVERFLOW_RESET) a->frame_count = 0;
    }
    // Resolution change resets (AccumulatePass.cpp:200-205).
    if (width != a->width || height != a->height) {
        a->width = width;
        a->height = height;
        a->frame_count = 0;
    }
    const uint32_t blocks = (n + 255u) / 256u;
    const bool integer_input = src_format != PLT_ACCUM_INPUT_FLOAT;
    if (!a->enabled && !integer_input) {
        // Blit (AccumulatePass.cpp:222-226): no frame counted.
        passthrough_kernel<PLT_ACCUM_INPUT_FLOAT><<<blocks, 256, 0, st>>>(src, dst, n);
        return plt::check_launch("accumulate passthrough");
    }
    // accumulate() (AccumulatePass.cpp:238-301).
    if (a->enabled && a->state_pixels != n) {
        if (a->d_state) (void)hipFree(a->d_state);
        a->d_state = nullptr;
        a->state_pixels = 0;
        if (hipMalloc(&a->d_state, size_t(n) * 32u) != hipSuccess)
            return set_error(PLT_ENOMEM, "plt_accumulate_execute: out of device memory for the accumulation buffers");
        a->state_pixels = n;
        a->frame_count = 0;  // prepareBuffer re-creates and resets (AccumulatePass.cpp:396-401)
    }
    AccumArgs args;
    args.src = src;
    args.dst = dst;
    args.state = a->d_state;
    args.n = n;
    args.count = a->frame_count;
    args.denom = float(a->frame_count + 1u);
    args.wf = 1.0f / args.denom;
    args.first = a->frame_count == 0;
    const bool ema = a->max_frames > 0;
    if (a->max_frames == 0 || a->precision == kSingleCompensated || a->frame_count < a->max_frames) a->frame_count++;
    if (!a->enabled) {
        switch (src_format) {
            case PLT_ACCUM_INPUT_UINT: passthrough_kernel<PLT_ACCUM_INPUT_UINT><<<blocks, 256, 0, st>>>(src, dst, n); break;
            default: passthrough_kernel<PLT_ACCUM_INPUT_SINT><<<blocks, 256, 0, st>>>(src, dst, n); break;
        }
        return plt::check_launch("accumulate passthrough");
    }
    if (a->precision == kDouble) {
        if (ema) launch_mode<kDouble, true>(src_format, args, blocks, st);
        else launch_mode<kDouble, false>(src_format, args, blocks, st);
    } else if (a->precision == kSingle) {
        if (ema) launch_mode<kSingle, true>(src_format, args, blocks, st);
        else launch_mode<kSingle, false>(src_format, args, blocks, st);
    } else {
        launch_mode<kSingleCompensated, false>(src_format, args, blocks, st);
    }
    return plt::check_launch("accumulate_kernel");
}

}  // extern "C"
