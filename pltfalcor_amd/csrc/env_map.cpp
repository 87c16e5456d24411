// Environment map light, host side: the lat-long radiance image with its transform and intensity
// (Scene/Lights/EnvMap.{h,cpp}, EnvMapData.slang) and the hierarchical importance map that
// EnvMapSampler builds for it (Rendering/Lights/EnvMapSampler.cpp:86-120,
// EnvMapSamplerSetup.cs.slang): 512 x 512 texels over the equal-area octahedral map, each the mean
// luminance of 8 x 8 bilinear samples of the image, and its 2 x 2 box mip chain down to one texel
// (Texture::generateMips). Built once per scene on the host, like the BVH.
#include <cmath>
#include <cstring>
#include <string>
#include <algorithm>
#include "env_map.hpp"
#include "plt_internal.hpp"
#include "scene_device.hpp"

namespace plt {

namespace {

float sgnf(float x) { return x >= 0.f ? 1.f : -1.f; }

}  // namespace

// oct_to_ndir_equal_area_snorm (MathHelpers.slang:223-239).
void env_oct_to_ndir(float px, float py, float out[3]) {
    const float d = 1.f - (std::fabs(px) + std::fabs(py));
    const float r = 1.f - std::fabs(d);
    const float phi = r > 0.f ? ((std::fabs(py) - std::fabs(px)) / r + 1.f) * 0.785398163397448309616f : 0.f;
    const float f = r * std::sqrt(2.f - r * r);
    out[0] = f * sgnf(px) * std::cos(phi);
    out[1] = f * sgnf(py) * std::sin(phi);
    out[2] = sgnf(d) * (1.f - r * r);
}

// world_to_latlong_map (MathHelpers.slang:106-113).
void env_latlong(const float dir[3], float& u, float& v) {
    const float l = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    const float p[3] = {dir[0] / l, dir[1] / l, dir[2] / l};
    u = std::atan2(p[0], -p[2]) * 0.159154943091895335769f + 0.5f;
    v = std::acos(std::fmin(std::fmax(p[1], -1.f), 1.f)) * 0.318309886183790671538f;
}

// Bilinear lookup at uv, U wrapping and V clamped (the env map sampler, EnvMap.cpp:123-126).
void env_bilinear(const HostEnvMap& e, float u, float v, float rgb[3]) {
    const float s = u * float(e.width) - 0.5f, t = v * float(e.height) - 0.5f;
    const float fs = std::floor(s), ft = std::floor(t);
    const float a = s - fs, b = t - ft;
    const int W = int(e.width), H = int(e.height);
    int x0 = int(fs) % W;
    if (x0 < 0) x0 += W;
    const int x1 = (x0 + 1) % W;
    const int y0 = std::min(std::max(int(ft), 0), H - 1), y1 = std::min(std::max(int(ft) + 1, 0), H - 1);
    for (int c = 0; c < 3; ++c) {
        const float c00 = e.rgba[4 * (size_t(y0) * W + x0) + c], c10 = e.rgba[4 * (size_t(y0) * W + x1) + c];
        const float c01 = e.rgba[4 * (size_t(y1) * W + x0) + c], c11 = e.rgba[4 * (size_t(y1) * W + x1) + c];
        const float top = c00 + (c10 - c00) * a, bot = c01 + (c11 - c01) * a;
        rgb[c] = top + (bot - top) * b;
    }
}

void build_env_importance(HostEnvMap& e) {
    const uint32_t dim = kEnvImportanceDim, spp = 8;  // 64 samples per texel (8 x 8)
    e.importance.clear();
    e.mip_offset.clear();
    std::vector<float> level(size_t(dim) * dim);
    const float inv = 1.f / float(spp * spp);
    const float dimS = float(dim * spp);
    for (uint32_t y = 0; y < dim; ++y)
        for (uint32_t x = 0; x < dim; ++x) {
            float L = 0.f;
            for (uint32_t sy = 0; sy < spp; ++sy)
                for (uint32_t sx = 0; sx < spp; ++sx) {
                    const float px = (float(x * spp + sx) + 0.5f) / dimS, py = (float(y * spp + sy) + 0.5f) / dimS;
                    float dir[3], u, v, rgb[3];
                    env_oct_to_ndir(px * 2.f - 1.f, py * 2.f - 1.f, dir);
                    env_latlong(dir, u, v);
                    env_bilinear(e, u, v, rgb);
                    L += rgb[0] * 0.2126f + rgb[1] * 0.7152f + rgb[2] * 0.0722f;  // luminance (ColorHelpers.slang)
                }
            level[size_t(y) * dim + x] = L * inv;
        }
    uint32_t d = dim;
    for (;;) {
        e.mip_offset.push_back(uint32_t(e.importance.size()));
        e.importance.insert(e.importance.end(), level.begin(), level.end());
        if (d == 1) break;
        const uint32_t h = d / 2;
        std::vector<float> next(size_t(h) * h);
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < h; ++x) {
                const float a = level[size_t(2 * y) * d + 2 * x], b = level[size_t(2 * y) * d + 2 * x + 1];
                const float c = level[size_t(2 * y + 1) * d + 2 * x], dd = level[size_t(2 * y + 1) * d + 2 * x + 1];
                next[size_t(y) * h + x] = ((a + b) + (c + dd)) * 0.25f;
            }
        level.swap(next);
        d = h;
    }
    e.base_mip = uint32_t(e.mip_offset.size() - 1);
}

}  // namespace plt

using plt::set_error;

extern "C" int plt_scene_set_env_map(plt_scene* scene, uint32_t width, uint32_t height, const float* rgba,
                                     const float* transform, float intensity, const float* tint) {
    if (!scene) return set_error(PLT_EINVAL, "plt_scene_set_env_map: NULL scene");
    plt::HostEnvMap& e = scene->host.env;
    if (width == 0 || height == 0) {  // remove
        e = plt::HostEnvMap();
        return PLT_OK;
    }
    if (!rgba || width > 65536 || height > 65536)
        return set_error(PLT_EINVAL, "plt_scene_set_env_map: bad image (1..65536 texels per side)");
    plt::HostEnvMap n;
    n.width = width;
    n.height = height;
    n.rgba.assign(rgba, rgba + size_t(width) * height * 4);
    // local-to-world transform (row-major 3x3) and its inverse (a rotation: the transpose)
    const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const float* M = transform ? transform : I;
    for (int i = 0; i < 9; ++i) n.transform[i] = M[i];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) n.inv_transform[3 * i + j] = M[3 * j + i];
    for (int c = 0; c < 3; ++c) n.scale[c] = intensity * (tint ? tint[c] : 1.f);  // EnvMap::getIntensity: intensity * tint
    plt::build_env_importance(n);
    e = std::move(n);
    return PLT_OK;
}
