// Solve pass of the PLT path tracer with four HWSS wavelengths per beam, with the analytic-light
// and environment-map code (one translation unit per large kernel, so they compile in parallel).
#include "pathtracer_kernels.hpp"

namespace plt {

int launch_solve_4_lights(const PtScene& sc, const PtFrame& fr, hipStream_t st) {
    const dim3 g((fr.n_paths + kPtBlock - 1) / kPtBlock), b(kPtBlock);
    hipLaunchKernelGGL((solve_kernel<4, true>), g, b, 0, st, sc, fr);
    return check_launch("solve_kernel<4>");
}

}  // namespace plt
