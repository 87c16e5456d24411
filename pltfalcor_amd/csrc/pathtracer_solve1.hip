// Solve pass of the PLT path tracer with one wavelength per beam (HWSS 1), with and without the
// analytic-light and environment-map code: the megakernels and the split pass's light-sampling
// kernels. One translation unit per instantiation group, so the large solve kernels compile in
// parallel.
#include "pathtracer_split.hpp"

namespace plt {

int launch_solve_1(const PtScene& sc, const PtFrame& fr, hipStream_t st, bool lights) {
    const dim3 g((fr.n_paths + kPtBlock - 1) / kPtBlock), b(kPtBlock);
    if (lights) hipLaunchKernelGGL((solve_kernel<1, true>), g, b, 0, st, sc, fr);
    else hipLaunchKernelGGL((solve_kernel<1, false>), g, b, 0, st, sc, fr);
    return check_launch("solve_kernel<1>");
}

int launch_solve_sample_1(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, hipStream_t st, bool lights) {
    const dim3 g((fr.n_paths + kPtBlock - 1) / kPtBlock), b(kPtBlock);
    if (lights) hipLaunchKernelGGL((solve_sample_kernel<1, true>), g, b, 0, st, sc, fr, sb);
    else hipLaunchKernelGGL((solve_sample_kernel<1, false>), g, b, 0, st, sc, fr, sb);
    return check_launch("solve_sample_kernel<1>");
}

}  // namespace plt
