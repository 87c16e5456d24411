// Camera jitter patterns of the G-buffer pass that precedes PLTPT in the reference's render graph
// (scripts/PLTPathTracer.py:16: GBufferRT with samplePattern Halton, sampleCount 16). The G-buffer
// pass attaches a CPU sample generator to the camera (GBufferBase.cpp:202-234); every frame the
// camera draws the next 2D offset in [-0.5, 0.5)^2, scales it by 1 / frame size and stores it as
// its jitter (Camera.cpp:63-70), which the primary rays then apply (Camera.slang:68-78).
//
// Patterns (GBufferBase.h:39-45):
//  - Center: no generator; the jitter is the camera's own (zero unless set).
//  - DirectX: the 8 standard Direct3D 8x MSAA positions (DxSamplePattern.cpp:32-39), cyclic.
//  - Halton: radical inverses in bases 2 and 3, accumulated in float, shifted so that 0 maps to
//    the pixel centre: fract(h + 0.5) - 0.5 (HaltonSamplePattern.cpp:36-70).
//  - Stratified: an M x N grid (M, N the closest factor pair), bins visited in a fresh random
//    permutation each round, uniform offset within the bin; std::mt19937 default-seeded,
//    std::shuffle and std::uniform_real_distribution<float> (StratifiedSamplePattern.cpp:34-88).
//    These two library algorithms are implementation-defined, so the sequence is the one of the
//    C++ standard library this file is built with.
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>
#include "plt_internal.hpp"
#include "sample_pattern.hpp"

namespace plt {

namespace {

float halton(uint32_t index, uint32_t base) {
    float result = 0.0f, factor = 1.0f;
    for (; index > 0; index /= base) {
        factor /= float(base);
        result += factor * float(index % base);
    }
    return result;
}

float fract(float x) { return x - std::floor(x); }

}  // namespace

int SamplePattern::configure(int pattern_type, uint32_t sample_count) {
    if (pattern_type < PLT_SAMPLE_PATTERN_CENTER || pattern_type > PLT_SAMPLE_PATTERN_STRATIFIED)
        return set_error(PLT_EINVAL, "sample pattern must be 0 (Center), 1 (DirectX), 2 (Halton) or 3 (Stratified)");
    type = pattern_type;
    count = sample_count;
    if (type == PLT_SAMPLE_PATTERN_DIRECTX) count = 8;  // DxSamplePattern supports 8 only (DxSamplePattern.cpp:43-47)
    if (type == PLT_SAMPLE_PATTERN_STRATIFIED) {
        count = std::clamp(count, 1u, 1024u);  // StratifiedSamplePattern.cpp:41-44
        bins_x = uint32_t(std::sqrt(double(count)));
        bins_y = count / bins_x;
        while (bins_x * bins_y != count) {
            ++bins_x;
            bins_y = count / bins_x;
        }
        perm.resize(count);
        for (uint32_t i = 0; i < count; ++i) perm[i] = i;
    }
    cur = 0;
    rng = std::mt19937();
    return PLT_OK;
}

void SamplePattern::next(float out[2]) {
    switch (type) {
        case PLT_SAMPLE_PATTERN_DIRECTX: {
            static const float kDx[8][2] = {{1.f / 16, -3.f / 16}, {-1.f / 16, 3.f / 16}, {5.f / 16, 1.f / 16},
                                            {-3.f / 16, -5.f / 16}, {-5.f / 16, 5.f / 16}, {-7.f / 16, -1.f / 16},
                                            {3.f / 16, 7.f / 16}, {7.f / 16, -7.f / 16}};
            out[0] = kDx[cur % 8][0];
            out[1] = kDx[cur % 8][1];
            ++cur;
            return;
        }
        case PLT_SAMPLE_PATTERN_HALTON: {
            const float hx = halton(cur, 2), hy = halton(cur, 3);
            ++cur;
            if (count != 0) cur %= count;
            out[0] = fract(hx + 0.5f) - 0.5f;
            out[1] = fract(hy + 0.5f) - 0.5f;
            return;
        }
        case PLT_SAMPLE_PATTERN_STRATIFIED: {
            std::uniform_real_distribution<float> dist;
            if (cur == 0) std::shuffle(perm.begin(), perm.end(), rng);
            const uint32_t bin = perm[cur];
            const uint32_t i = bin % bins_x, j = bin / bins_x;
            cur = (cur + 1) % count;
            const float x = (float(i) + dist(rng)) / float(bins_x);
            const float y = (float(j) + dist(rng)) / float(bins_y);
            out[0] = x - 0.5f;
            out[1] = y - 0.5f;
            return;
        }
        default:
            out[0] = out[1] = 0.f;
    }
}

}  // namespace plt

extern "C" int plt_sample_pattern_generate(int pattern, uint32_t sample_count, uint32_t n, float* xy) {
    if (!xy && n) return plt::set_error(PLT_EINVAL, "plt_sample_pattern_generate: xy is NULL");
    plt::SamplePattern sp;
    if (int rc = sp.configure(pattern, sample_count)) return rc;
    for (uint32_t k = 0; k < n; ++k) sp.next(xy + 2 * k);
    return PLT_OK;
}
