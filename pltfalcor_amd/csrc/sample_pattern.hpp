// Host-side camera jitter pattern generator (the G-buffer pass's CPU sample generators).
#pragma once
#include <cstdint>
#include <random>
#include <vector>

namespace plt {

struct SamplePattern {
    int type = 0;  // PLT_SAMPLE_PATTERN_*
    uint32_t count = 0;
    uint32_t cur = 0;
    uint32_t bins_x = 1, bins_y = 1;
    std::vector<uint32_t> perm;
    std::mt19937 rng;
    // Selects the pattern and restarts it (the G-buffer pass creates a new generator).
    int configure(int pattern_type, uint32_t sample_count);
    // Next offset in [-0.5, 0.5)^2 pixels.
    void next(float out[2]);
};

}  // namespace plt
