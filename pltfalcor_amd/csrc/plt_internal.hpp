// Internal (non-ABI) declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include "../../include/plt_abi.h"

// A device context: the GPU ordinal and the spectral profiles (device copy) that materials and
// BSDF batches reference by id.
struct plt_context {
    int device = 0;
    plt_spectral_profile* d_profiles = nullptr;
    uint32_t n_profiles = 0;
};

namespace plt {

// Records a message as the calling thread's last error and returns `code`.
int set_error(int code, const std::string& msg);
// Converts a pending launch error into PLT_EHIP with the kernel name.
int check_launch(const char* kernel);

int launch_generate_queries(uint32_t n, uint32_t seed, uint32_t* q, hipStream_t st);
int launch_eval_queries(uint32_t n, const uint32_t* q, float* res, const plt_spectral_profile* prof, uint32_t nprof,
                        hipStream_t st);
int launch_sample_generator(uint32_t n, uint32_t width, uint32_t sn, uint32_t draws, bool tiny, uint32_t* out, hipStream_t st);

}  // namespace plt
