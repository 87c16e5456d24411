// Solve pass of the PLT path tracer with four HWSS wavelengths per beam, with the analytic-light
// and environment-map code: the megakernel and the split pass's light-sampling kernel (one
// translation unit per large kernel group, so they compile in parallel).
#include "pathtracer_split.hpp"

namespace plt {

int launch_solve_4_lights(const PtScene& sc, const PtFrame& fr, hipStream_t st) {
    const dim3 g((fr.n_paths + kPtBlock - 1) / kPtBlock), b(kPtBlock);
    hipLaunchKernelGGL((solve_kernel<4, true>), g, b, 0, st, sc, fr);
    return check_launch("solve_kernel<4>");
}

int launch_solve_sample_4_lights(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, hipStream_t st) {
    const dim3 g((fr.n_paths + kPtBlock - 1) / kPtBlock), b(kPtBlock);
    hipLaunchKernelGGL((solve_sample_kernel<4, true>), g, b, 0, st, sc, fr, sb);
    return check_launch("solve_sample_kernel<4>");
}

}  // namespace plt
