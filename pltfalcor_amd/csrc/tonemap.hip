// Tone mapping (Source/RenderPasses/ToneMapper): the display pass at the end of the reference's
// render graph (scripts/PLTPathTracer.py:26-27,51-53: autoExposure on, exposureCompensation 0).
//
// Host: the option setters with the reference's clamps and coupled exposure values
// (ToneMapper.cpp:157-178,395-515) and the colour transform = white balance (von Kries in CAT02
// LMS, Utils/Color/ColorUtils.h:190-218) x exposure compensation x manual exposure
// (ToneMapper.cpp:535-547), computed in float as the reference's float matrices are.
//
// Device, auto exposure (ToneMapper.cpp:230-241, Luminance.ps.slang:28-42): the luminance image is
// bit_floor(W) x bit_floor(H) texels, each a bilinear sample of the source at the texel centre,
// holding log2(max(1e-4, dot(rgb, (.299,.587,.114)))); its mip chain (2x2 box per level,
// Texture::generateMips) ends in one texel, the mean log luminance, which the tone-mapping pass
// reads (kLuminanceLod = 16, ToneMapping.ps.slang:38,150-153). Here the chain is computed without
// storing it: each 32x32 tile of a level is reduced through five levels in LDS (a tile aligned
// with the power-of-two grid covers exactly the texels its level-5 texel averages), and the
// per-tile means form the next image until one value is left (two launches at 1080p).
//
// Device, tone mapping (ToneMapping.ps.slang:145-168): exposure by 0.042 / exp2(mean log
// luminance), the 3x3 colour transform, one of six operators, optional saturate; alpha passes
// through. One thread per pixel, 16-byte loads and stores (32 B per pixel).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include "plt_internal.hpp"

using plt::set_error;

namespace {

constexpr int kTile = 32;  // texels per tile side; 256 threads reduce 32x32 texels

struct LumArgs {
    const float4* src;   // level-0 pass: the source image
    const float* img;    // later passes: the previous level's per-tile means
    float* out;          // per-tile means of this pass
    uint32_t sw, sh;     // source size (level-0 pass)
    uint32_t w, h;       // size of the image this pass reduces (powers of two)
    float sx, sy;        // source texels per luminance texel (level-0 pass)
};

__device__ __forceinline__ float log_luminance(float4 c) {
    const float l = c.x * 0.299f + c.y * 0.587f + c.z * 0.114f;
    return log2f(fmaxf(1e-4f, l));
}

// Bilinear sample of the source at the centre of luminance texel (x, y). The luminance image is
// never larger than the source, so the four taps stay inside it.
__device__ __forceinline__ float4 bilinear(const LumArgs& a, uint32_t x, uint32_t y) {
    const float u = (float(x) + 0.5f) * a.sx - 0.5f;
    const float v = (float(y) + 0.5f) * a.sy - 0.5f;
    const float fu0 = floorf(u), fv0 = floorf(v);
    const uint32_t x0 = uint32_t(fu0), y0 = uint32_t(fv0);
    const uint32_t x1 = min(x0 + 1u, a.sw - 1u), y1 = min(y0 + 1u, a.sh - 1u);
    const float fu = u - fu0, fv = v - fv0;
    const float4 c00 = a.src[size_t(y0) * a.sw + x0], c10 = a.src[size_t(y0) * a.sw + x1];
    const float4 c01 = a.src[size_t(y1) * a.sw + x0], c11 = a.src[size_t(y1) * a.sw + x1];
    float4 r;
    r.x = (c00.x + (c10.x - c00.x) * fu) + ((c01.x + (c11.x - c01.x) * fu) - (c00.x + (c10.x - c00.x) * fu)) * fv;
    r.y = (c00.y + (c10.y - c00.y) * fu) + ((c01.y + (c11.y - c01.y) * fu) - (c00.y + (c10.y - c00.y) * fu)) * fv;
    r.z = (c00.z + (c10.z - c00.z) * fu) + ((c01.z + (c11.z - c01.z) * fu) - (c00.z + (c10.z - c00.z) * fu)) * fv;
    r.w = 0.f;
    return r;
}

// One pass of the mip-chain reduction: block (bx, by) reduces the tile of `w x h` texels at
// (bx*32, by*32) (clipped to the image) to its mean and writes it to out[by * tiles_x + bx].
template <bool FromSource>
__global__ __launch_bounds__(256) void luminance_reduce_kernel(LumArgs a) {
    __shared__ float buf[2][kTile * kTile];
    const uint32_t tw = min(a.w, uint32_t(kTile)), th = min(a.h, uint32_t(kTile));
    const uint32_t ox = blockIdx.x * kTile, oy = blockIdx.y * kTile;
    for (uint32_t t = threadIdx.x; t < tw * th; t += 256) {
        const uint32_t x = ox + t % tw, y = oy + t / tw;
        buf[0][t] = FromSource ? log_luminance(bilinear(a, x, y)) : a.img[size_t(y) * a.w + x];
    }
    __syncthreads();
    uint32_t w = tw, h = th, cur = 0;
    while (w > 1 || h > 1) {
        const uint32_t nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
        const uint32_t dx = w > 1 ? 1 : 0, dy = h > 1 ? 1 : 0;
        const float scale = (dx && dy) ? 0.25f : 0.5f;
        for (uint32_t t = threadIdx.x; t < nw * nh; t += 256) {
            const uint32_t x = (t % nw) << dx, y = (t / nw) << dy;
            const float* s = buf[cur];
            const float a0 = s[y * w + x], a1 = s[y * w + x + dx];
            const float b0 = s[(y + dy) * w + x], b1 = s[(y + dy) * w + x + dx];
            // With one axis collapsed, a0 and b1 are the two texels along the other axis.
            buf[cur ^ 1][t] = (dx && dy) ? ((a0 + a1) + (b0 + b1)) * scale : (a0 + b1) * scale;
        }
        __syncthreads();
        cur ^= 1;
        w = nw;
        h = nh;
    }
    if (threadIdx.x == 0) a.out[blockIdx.y * gridDim.x + blockIdx.x] = buf[cur][0];
}

struct TonemapArgs {
    const float4* src;
    float4* dst;
    const float* mean_log_lum;  // device scalar, read when auto exposure is on
    uint32_t n;
    float m[9];                 // colour transform, row-major
    float white_scale;          // 1 / uc2(whiteScale), precomputed (HableUc2)
    float white_max_lum2;       // whiteMaxLuminance^2 (ReinhardModified)
};

__device__ __forceinline__ float lum_weights(float r, float g, float b) { return r * 0.299f + g * 0.587f + b * 0.114f; }
__device__ __forceinline__ float saturate(float x) { return fminf(fmaxf(x, 0.f), 1.f); }

// The Uncharted 2 curve (ToneMapping.ps.slang:85-105).
__host__ __device__ __forceinline__ float uc2(float x) {
    const float A = 0.22f, B = 0.3f, C = 0.1f, D = 0.2f, E = 0.01f, F = 0.3f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}

template <int Op, bool Auto, bool Clamp>
__global__ __launch_bounds__(256) void tonemap_kernel(TonemapArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    const float4 c = a.src[i];
    float r = c.x, g = c.y, b = c.z;
    if constexpr (Auto) {
        const float k = 0.042f / exp2f(*a.mean_log_lum);
        r *= k;
        g *= k;
        b *= k;
    }
    {
        const float tr = a.m[0] * r + a.m[1] * g + a.m[2] * b;
        const float tg = a.m[3] * r + a.m[4] * g + a.m[5] * b;
        const float tb = a.m[6] * r + a.m[7] * g + a.m[8] * b;
        r = tr;
        g = tg;
        b = tb;
    }
    if constexpr (Op == PLT_TONEMAP_REINHARD) {
        const float l = lum_weights(r, g, b);
        const float s = (l / (l + 1.f)) / l;
        r *= s;
        g *= s;
        b *= s;
    } else if constexpr (Op == PLT_TONEMAP_REINHARD_MODIFIED) {
        // The reference multiplies by (1 + L) where Reinhard's operator divides
        // (ToneMapping.ps.slang:65-70); kept as it is.
        const float l = lum_weights(r, g, b);
        const float s = (l * (1.f + l / a.white_max_lum2) * (1.f + l)) / l;
        r *= s;
        g *= s;
        b *= s;
    } else if constexpr (Op == PLT_TONEMAP_HEJI_HABLE_ALU) {
        auto f = [](float x) {
            x = fmaxf(0.f, x - 0.004f);
            x = (x * (6.2f * x + 0.5f)) / (x * (6.2f * x + 1.7f) + 0.06f);
            return powf(x, 2.2f);
        };
        r = f(r);
        g = f(g);
        b = f(b);
    } else if constexpr (Op == PLT_TONEMAP_HABLE_UC2) {
        r = uc2(2.f * r) * a.white_scale;
        g = uc2(2.f * g) * a.white_scale;
        b = uc2(2.f * b) * a.white_scale;
    } else if constexpr (Op == PLT_TONEMAP_ACES) {
        auto f = [](float x) {
            x *= 0.6f;
            return saturate((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f));
        };
        r = f(r);
        g = f(g);
        b = f(b);
    }
    if constexpr (Clamp) {
        r = saturate(r);
        g = saturate(g);
        b = saturate(b);
    }
    a.dst[i] = make_float4(r, g, b, c.w);
}

template <int Op, bool Auto>
void launch_tonemap_clamp(bool clamp, const TonemapArgs& a, uint32_t blocks, hipStream_t st) {
    if (clamp) tonemap_kernel<Op, Auto, true><<<blocks, 256, 0, st>>>(a);
    else tonemap_kernel<Op, Auto, false><<<blocks, 256, 0, st>>>(a);
}

template <int Op>
void launch_tonemap_op(bool autoexp, bool clamp, const TonemapArgs& a, uint32_t blocks, hipStream_t st) {
    if (autoexp) launch_tonemap_clamp<Op, true>(clamp, a, blocks, st);
    else launch_tonemap_clamp<Op, false>(clamp, a, blocks, st);
}

// ---------------------------------------------------------------------------- colour transform

using Mat3 = float[3][3];

void mat_mul(const Mat3 a, const Mat3 b, Mat3 out) {
    Mat3 t;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
    std::memcpy(out, t, sizeof(t));
}

void mat_vec(const Mat3 a, const float v[3], float out[3]) {
    float t[3];
    for (int i = 0; i < 3; ++i) t[i] = a[i][0] * v[0] + a[i][1] * v[1] + a[i][2] * v[2];
    std::memcpy(out, t, sizeof(t));
}

// Published colour-science constants, stored as float as the reference's float matrices hold
// them: linear Rec.709 <-> CIE XYZ (D65) and the CIECAM02 chromatic adaptation matrix CAT02 with
// its inverse (ColorUtils.h:65-95).
const Mat3 kRgbToXyz = {{0.4123907992659595f, 0.3575843393838780f, 0.1804807884018343f},
                        {0.2126390058715104f, 0.7151686787677559f, 0.0721923153607337f},
                        {0.0193308187155918f, 0.1191947797946259f, 0.9505321522496608f}};
const Mat3 kXyzToRgb = {{3.2409699419045213f, -1.5373831775700935f, -0.4986107602930033f},
                        {-0.9692436362808798f, 1.8759675015077206f, 0.0415550574071756f},
                        {0.0556300796969936f, -0.2039769588889765f, 1.0569715142428784f}};
const Mat3 kXyzToLmsCat02 = {{0.7328f, 0.4296f, -0.1624f}, {-0.7036f, 1.6975f, 0.0061f}, {0.0030f, 0.0136f, 0.9834f}};
const Mat3 kLmsToXyzCat02 = {{1.096123820835514f, -0.278869000218287f, 0.182745179382773f},
                             {0.454369041975359f, 0.473533154307412f, 0.072097803717229f},
                             {-0.009627608738429f, -0.005698031216113f, 1.015325639954543f}};

// CIE XYZ (Y = 1) of a blackbody at T kelvin, from the piecewise rational fit of the Planckian
// locus in Kang et al., "Design of Advanced Color Temperature Control System for HDTV
// Applications" (2002), evaluated in double, chromaticities rounded to float
// (ColorUtils.h:134-187).
void color_temperature_xyz(float T, float out[3]) {
    const double t = T, t2 = t * t, t3 = t * t * t;
    const double xc = T < 4000.f ? -0.2661239e9 / t3 - 0.2343580e6 / t2 + 0.8776956e3 / t + 0.179910
                                 : -3.0258469e9 / t3 + 2.1070379e6 / t2 + 0.2226347e3 / t + 0.240390;
    const double x = xc, x2 = x * x, x3 = x * x * x;
    double yc;
    if (T < 2222.f) yc = -1.1063814 * x3 - 1.34811020 * x2 + 2.18555832 * x - 0.20219683;
    else if (T < 4000.f) yc = -0.9549476 * x3 - 1.37418593 * x2 + 2.09137015 * x - 0.16748867;
    else yc = 3.0817580 * x3 - 5.87338670 * x2 + 3.75112997 * x - 0.37001483;
    const float fx = float(xc), fy = float(yc), Y = 1.f;
    out[0] = fx * Y / fy;
    out[1] = Y;
    out[2] = (1.f - fx - fy) * Y / fy;
}

// Von Kries white balance from colour temperature T to D65-like 6500 K in CAT02 LMS
// (ColorUtils.h:202-218).
void white_balance_rec709(float T, Mat3 out) {
    Mat3 ma, inv_ma;
    mat_mul(kXyzToLmsCat02, kRgbToXyz, ma);
    mat_mul(kXyzToRgb, kLmsToXyzCat02, inv_ma);
    float xyz_d[3], xyz_s[3], wd[3], ws[3];
    color_temperature_xyz(6500.f, xyz_d);
    color_temperature_xyz(T, xyz_s);
    mat_vec(kXyzToLmsCat02, xyz_d, wd);
    mat_vec(kXyzToLmsCat02, xyz_s, ws);
    const Mat3 d = {{wd[0] / ws[0], 0.f, 0.f}, {0.f, wd[1] / ws[1], 0.f}, {0.f, 0.f, wd[2] / ws[2]}};
    Mat3 t;
    mat_mul(inv_ma, d, t);
    mat_mul(t, ma, out);
}

struct OptionSpec {
    const char* key;
    double def;
};
// Keys of ToneMapper::getScriptingDictionary (ToneMapper.cpp:181-200) with the defaults of
// ToneMapper.h:109-139 (operator 4 = HableUc2, exposureMode 0 = AperturePriority).
const OptionSpec kOptions[] = {
    {"outputSize", 0}, {"useSceneMetadata", 1}, {"exposureCompensation", 0}, {"autoExposure", 0},
    {"filmSpeed", 100}, {"whiteBalance", 0}, {"whitePoint", 6500}, {"operator", PLT_TONEMAP_HABLE_UC2},
    {"clamp", 1}, {"whiteMaxLuminance", 1}, {"whiteScale", 11.2}, {"fNumber", 1}, {"shutter", 1},
    {"exposureMode", 0},
};
constexpr int kNumOptions = int(sizeof(kOptions) / sizeof(kOptions[0]));

// Ranges of the setters (ToneMapper.cpp:78-100).
constexpr float kEcMin = -12.f, kEcMax = 12.f;
constexpr float kIsoMin = 1.f, kIsoMax = 6400.f;
constexpr float kFMin = 0.1f, kFMax = 100.f;
constexpr float kShutterMin = 0.1f, kShutterMax = 10000.f;
constexpr float kWpMin = 1905.f, kWpMax = 25000.f;

}  // namespace

struct plt_tonemapper {
    plt_context* ctx = nullptr;
    bool use_scene_metadata = true;
    float exposure_compensation = 0.f;
    bool auto_exposure = false;
    float exposure_value = 0.f;
    float film_speed = 100.f;
    float f_number = 1.f;
    float shutter = 1.f;
    int exposure_mode = 0;
    bool white_balance = false;
    float white_point = 6500.f;
    int op = PLT_TONEMAP_HABLE_UC2;
    bool clamp = true;
    float white_max_luminance = 1.f;
    float white_scale = 11.2f;
    float* d_lum = nullptr;  // per-tile means of the reduction passes, ping-pong halves
    size_t lum_cap = 0;      // floats per half
};

namespace {

void update_exposure_value(plt_tonemapper* t) {
    t->exposure_value = std::log2(t->shutter * t->f_number * t->f_number);
}

void color_transform(const plt_tonemapper* t, float out[9]) {
    Mat3 wb = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
    if (t->white_balance) white_balance_rec709(t->white_point, wb);
    const float exposure_scale = std::pow(2.f, t->exposure_compensation);
    float manual = 1.f;
    if (!t->auto_exposure) manual = (1.f / 100.f * t->film_speed) / (t->shutter * t->f_number * t->f_number);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[3 * i + j] = wb[i][j] * exposure_scale * manual;
}

}  // namespace

extern "C" {

int plt_tonemapper_create(plt_context* ctx, plt_tonemapper** out) {
    if (!out) return set_error(PLT_EINVAL, "plt_tonemapper_create: out is NULL");
    auto* t = new (std::nothrow) plt_tonemapper();
    if (!t) return set_error(PLT_ENOMEM, "plt_tonemapper_create: out of host memory");
    t->ctx = ctx;
    *out = t;
    return PLT_OK;
}

int plt_tonemapper_destroy(plt_tonemapper* t) {
    if (!t) return PLT_OK;
    if (t->d_lum) {
        (void)hipSetDevice(t->ctx->device);
        (void)hipFree(t->d_lum);
    }
    delete t;
    return PLT_OK;
}

const char* plt_tonemapper_option_name(uint32_t i) { return i < uint32_t(kNumOptions) ? kOptions[i].key : nullptr; }

int plt_tonemapper_set_option(plt_tonemapper* t, const char* key, double value) {
    if (!t || !key) return set_error(PLT_EINVAL, "plt_tonemapper_set_option: bad arguments");
    const std::string s = key;
    const float v = float(value);
    if (s == "outputSize") {
        if (value != 0.0) return set_error(PLT_ENOTIMPL, "ToneMapper: only outputSize Default (0) is built (output = input size)");
    } else if (s == "outputFormat") {
        if (value != 0.0) return set_error(PLT_ENOTIMPL, "ToneMapper: only the RGBA32Float output is built");
    } else if (s == "useSceneMetadata") {
        t->use_scene_metadata = value != 0.0;
    } else if (s == "exposureCompensation") {
        t->exposure_compensation = std::clamp(v, kEcMin, kEcMax);
    } else if (s == "autoExposure") {
        t->auto_exposure = value != 0.0;
    } else if (s == "exposureValue") {
        // setExposureValue (ToneMapper.cpp:419-441); a property, not a dictionary key.
        const float ev_min = std::log2(kShutterMin * kFMin * kFMin), ev_max = std::log2(kShutterMax * kFMax * kFMax);
        t->exposure_value = std::clamp(v, ev_min, ev_max);
        if (t->exposure_mode == 0) {
            t->shutter = std::clamp(std::pow(2.f, t->exposure_value) / (t->f_number * t->f_number), kShutterMin, kShutterMax);
        } else {
            t->f_number = std::clamp(std::sqrt(std::pow(2.f, t->exposure_value) / t->shutter), kFMin, kFMax);
        }
        update_exposure_value(t);
    } else if (s == "filmSpeed") {
        t->film_speed = std::clamp(v, kIsoMin, kIsoMax);
    } else if (s == "whiteBalance") {
        t->white_balance = value != 0.0;
    } else if (s == "whitePoint") {
        t->white_point = std::clamp(v, kWpMin, kWpMax);
    } else if (s == "operator") {
        if (!(value >= 0.0 && value <= 5.0) || value != double(int(value)))
            return set_error(PLT_EINVAL, "ToneMapper: operator must be 0..5 (Linear, Reinhard, ReinhardModified, HejiHableAlu, HableUc2, Aces)");
        t->op = int(value);
    } else if (s == "clamp") {
        t->clamp = value != 0.0;
    } else if (s == "whiteMaxLuminance") {
        t->white_max_luminance = v;
    } else if (s == "whiteScale") {
        t->white_scale = std::max(0.001f, v);
    } else if (s == "fNumber") {
        t->f_number = std::clamp(v, kFMin, kFMax);
        update_exposure_value(t);
    } else if (s == "shutter") {
        t->shutter = std::clamp(v, kShutterMin, kShutterMax);
        update_exposure_value(t);
    } else if (s == "exposureMode") {
        if (value != 0.0 && value != 1.0)
            return set_error(PLT_EINVAL, "ToneMapper: exposureMode must be 0 (AperturePriority) or 1 (ShutterPriority)");
        t->exposure_mode = int(value);
    } else {
        return set_error(PLT_WARN_UNKNOWN_KEY, std::string("Unknown field '") + key + "' in a ToneMapping dictionary.");
    }
    return PLT_OK;
}

int plt_tonemapper_get_option(const plt_tonemapper* t, const char* key, double* v) {
    if (!t || !key || !v) return set_error(PLT_EINVAL, "plt_tonemapper_get_option: bad arguments");
    const std::string s = key;
    if (s == "outputSize" || s == "outputFormat") *v = 0.0;
    else if (s == "useSceneMetadata") *v = t->use_scene_metadata;
    else if (s == "exposureCompensation") *v = t->exposure_compensation;
    else if (s == "autoExposure") *v = t->auto_exposure;
    else if (s == "exposureValue") *v = t->exposure_value;
    else if (s == "filmSpeed") *v = t->film_speed;
    else if (s == "whiteBalance") *v = t->white_balance;
    else if (s == "whitePoint") *v = t->white_point;
    else if (s == "operator") *v = t->op;
    else if (s == "clamp") *v = t->clamp;
    else if (s == "whiteMaxLuminance") *v = t->white_max_luminance;
    else if (s == "whiteScale") *v = t->white_scale;
    else if (s == "fNumber") *v = t->f_number;
    else if (s == "shutter") *v = t->shutter;
    else if (s == "exposureMode") *v = t->exposure_mode;
    else return set_error(PLT_WARN_UNKNOWN_KEY, std::string("Unknown field '") + key + "' in a ToneMapping dictionary.");
    return PLT_OK;
}

int plt_tonemapper_set_scene_metadata(plt_tonemapper* t, float film_iso, float f_number, float shutter_speed) {
    if (!t) return set_error(PLT_EINVAL, "plt_tonemapper_set_scene_metadata: NULL pass");
    if (!t->use_scene_metadata) return PLT_OK;
    if (!std::isnan(film_iso)) plt_tonemapper_set_option(t, "filmSpeed", film_iso);
    if (!std::isnan(f_number)) plt_tonemapper_set_option(t, "fNumber", f_number);
    if (!std::isnan(shutter_speed)) plt_tonemapper_set_option(t, "shutter", shutter_speed);
    return PLT_OK;
}

int plt_tonemapper_get_color_transform(const plt_tonemapper* t, float* out9) {
    if (!t || !out9) return set_error(PLT_EINVAL, "plt_tonemapper_get_color_transform: bad arguments");
    color_transform(t, out9);
    return PLT_OK;
}

}  // extern "C"

namespace {

// The fixed luminance-reduction workspace of plt_tonemapper_execute_noalloc: two halves of
// kNoAllocTiles floats (the per-tile means ping-pong between them), 512 KiB on each device, held
// until the process ends. Host code never passes this name to a kernel: it is a host-side stand-in,
// and the device address of each GPU's copy comes from hipGetSymbolAddress (noalloc_scratch).
constexpr size_t kNoAllocTiles = 65536;  // 256 x 256 tiles: bit_floor(width), bit_floor(height) <= 8192
__device__ __align__(256) float scratch[2 * kNoAllocTiles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch`, looked up once per device with that device
// current (each GPU holds its own copy at its own address).
int noalloc_scratch(int device, float** out) {
    if (device < 0 || device >= kMaxDevices) return set_error(PLT_EINVAL, "plt_tonemapper_execute_noalloc: device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
            return set_error(PLT_EHIP, "plt_tonemapper_execute_noalloc: hipGetSymbolAddress of the scratch array failed");
        g_scratch_ptr[device] = static_cast<float*>(ptr);
    }
    *out = g_scratch_ptr[device];
    return PLT_OK;
}

// Size of the auto-exposure reduction: bit_floor of each side, and the level-0 tile count.
void lum_extent(uint32_t width, uint32_t height, uint32_t& w, uint32_t& h, size_t& tiles0) {
    auto bit_floor = [](uint32_t v) { uint32_t b = 1; while (b <= v / 2) b <<= 1; return b; };
    w = bit_floor(width);
    h = bit_floor(height);
    tiles0 = size_t((w + kTile - 1) / kTile) * ((h + kTile - 1) / kTile);
}

int check_execute_args(const plt_tonemapper* t, const float* src, uint32_t width, uint32_t height, const float* dst,
                       const char* fn) {
    if (!t) return set_error(PLT_EINVAL, std::string(fn) + ": NULL pass");
    if (!t->ctx) return set_error(PLT_ESTATE, std::string(fn) + ": pass created without a context");
    if (!src || !dst || width == 0 || height == 0 || width > 65536 || height > 65536)
        return set_error(PLT_EINVAL, std::string(fn) + ": bad arguments (sizes 1..65536)");
    if ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15u)
        return set_error(PLT_EINVAL, std::string(fn) + ": src and dst must be 16-byte aligned");
    if (uint64_t(width) * height > (1ull << 31)) return set_error(PLT_EINVAL, std::string(fn) + ": image larger than 2^31 pixels");
    return PLT_OK;
}

// The launches of one execute: the luminance reduction into `lum` (two halves of `half_cap` floats,
// at least the level-0 tile count each) when auto exposure is on, then the tone-mapping pass.
// Nothing is allocated here.
int execute_launches(const plt_tonemapper* t, const float* src, uint32_t width, uint32_t height, float* dst,
                     float* mean_log_luminance, float* lum, size_t half_cap, hipStream_t st) {
    const uint64_t n64 = uint64_t(width) * height;
    const float* mean = nullptr;
    if (t->auto_exposure) {
        uint32_t w, h;
        size_t need;
        lum_extent(width, height, w, h, need);
        auto tiles = [](uint32_t v) { return (v + kTile - 1) / kTile; };
        LumArgs a{};
        a.src = reinterpret_cast<const float4*>(src);
        a.sw = width;
        a.sh = height;
        a.w = w;
        a.h = h;
        a.sx = float(width) / float(w);
        a.sy = float(height) / float(h);
        int half = 0;
        a.out = lum;
        luminance_reduce_kernel<true><<<dim3(tiles(w), tiles(h)), 256, 0, st>>>(a);
        if (int rc = plt::check_launch("luminance_reduce_kernel")) return rc;
        w = tiles(w);
        h = tiles(h);
        while (w > 1 || h > 1) {
            a.img = lum + size_t(half) * half_cap;
            half ^= 1;
            a.out = lum + size_t(half) * half_cap;
            a.w = w;
            a.h = h;
            luminance_reduce_kernel<false><<<dim3(tiles(w), tiles(h)), 256, 0, st>>>(a);
            if (int rc = plt::check_launch("luminance_reduce_kernel")) return rc;
            w = tiles(w);
            h = tiles(h);
        }
        mean = lum + size_t(half) * half_cap;
        if (mean_log_luminance &&
            hipMemcpyAsync(mean_log_luminance, mean, sizeof(float), hipMemcpyDeviceToDevice, st) != hipSuccess)
            return set_error(PLT_EHIP, "plt_tonemapper_execute: copy of the mean log luminance failed");
    }
    TonemapArgs a{};
    a.src = reinterpret_cast<const float4*>(src);
    a.dst = reinterpret_cast<float4*>(dst);
    a.mean_log_lum = mean;
    a.n = uint32_t(n64);
    color_transform(t, a.m);
    a.white_scale = 1.f / uc2(t->white_scale);
    a.white_max_lum2 = t->white_max_luminance * t->white_max_luminance;
    const uint32_t blocks = uint32_t((n64 + 255) / 256);
    const bool ae = t->auto_exposure, cl = t->clamp;
    switch (t->op) {
        case PLT_TONEMAP_LINEAR: launch_tonemap_op<PLT_TONEMAP_LINEAR>(ae, cl, a, blocks, st); break;
        case PLT_TONEMAP_REINHARD: launch_tonemap_op<PLT_TONEMAP_REINHARD>(ae, cl, a, blocks, st); break;
        case PLT_TONEMAP_REINHARD_MODIFIED: launch_tonemap_op<PLT_TONEMAP_REINHARD_MODIFIED>(ae, cl, a, blocks, st); break;
        case PLT_TONEMAP_HEJI_HABLE_ALU: launch_tonemap_op<PLT_TONEMAP_HEJI_HABLE_ALU>(ae, cl, a, blocks, st); break;
        case PLT_TONEMAP_HABLE_UC2: launch_tonemap_op<PLT_TONEMAP_HABLE_UC2>(ae, cl, a, blocks, st); break;
        default: launch_tonemap_op<PLT_TONEMAP_ACES>(ae, cl, a, blocks, st); break;
    }
    return plt::check_launch("tonemap_kernel");
}

}  // namespace

extern "C" {

int plt_tonemapper_execute(plt_tonemapper* t, const float* src, uint32_t width, uint32_t height, float* dst,
                           float* mean_log_luminance, plt_stream stream) {
    if (int rc = check_execute_args(t, src, width, height, dst, "plt_tonemapper_execute")) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (hipSetDevice(t->ctx->device) != hipSuccess) return set_error(PLT_EHIP, "plt_tonemapper_execute: hipSetDevice failed");
    if (t->auto_exposure) {
        uint32_t w, h;
        size_t need;
        lum_extent(width, height, w, h, need);
        if (need > t->lum_cap) {
            if (t->d_lum) (void)hipFree(t->d_lum);
            t->d_lum = nullptr;
            t->lum_cap = 0;
            if (hipMalloc(&t->d_lum, 2 * need * sizeof(float)) != hipSuccess)
                return set_error(PLT_ENOMEM, "plt_tonemapper_execute: out of device memory for the luminance reduction");
            t->lum_cap = need;
        }
    }
    return execute_launches(t, src, width, height, dst, mean_log_luminance, t->d_lum, t->lum_cap, st);
}

int plt_tonemapper_execute_noalloc(plt_tonemapper* t, const float* src, uint32_t width, uint32_t height, float* dst,
                                   float* mean_log_luminance, plt_stream stream) {
    if (int rc = check_execute_args(t, src, width, height, dst, "plt_tonemapper_execute_noalloc")) return rc;
    uint32_t w, h;
    size_t need;
    lum_extent(width, height, w, h, need);
    if (t->auto_exposure && need > kNoAllocTiles)
        return set_error(PLT_EINVAL, "plt_tonemapper_execute_noalloc: the auto-exposure reduction of this image needs " +
                                         std::to_string(need) + " tiles; the fixed workspace holds " + std::to_string(kNoAllocTiles) +
                                         " (bit_floor(width) and bit_floor(height) up to 8192)");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (hipSetDevice(t->ctx->device) != hipSuccess) return set_error(PLT_EHIP, "plt_tonemapper_execute_noalloc: hipSetDevice failed");
    float* lum = nullptr;
    if (int rc = noalloc_scratch(t->ctx->device, &lum)) return rc;
    // every texel the launches read is written earlier in the same sequence, so nothing needs clearing
    return execute_launches(t, src, width, height, dst, mean_log_luminance, lum, kNoAllocTiles, st);
}

}  // extern "C"
