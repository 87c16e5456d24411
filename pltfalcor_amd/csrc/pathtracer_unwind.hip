// The split solve pass's exact unwind kernels (four and one HWSS wavelengths per beam): beam
// transport through the stored vertices and the per-path sums.
#include "pathtracer_split.hpp"

namespace plt {

int launch_solve_unwind(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, hipStream_t st) {
    const dim3 g((fr.n_paths + kPtBlock - 1) / kPtBlock), b(kPtBlock);
    if (fr.hwss == 4) hipLaunchKernelGGL((solve_unwind_kernel<4, kUnwindExact>), g, b, 0, st, sc, fr, sb);
    else hipLaunchKernelGGL((solve_unwind_kernel<1, kUnwindExact>), g, b, 0, st, sc, fr, sb);
    return check_launch("solve_unwind_kernel");
}

int launch_solve_walk_unwind(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, hipStream_t st) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        return set_error(PLT_EHIP, "solve_walk_unwind_kernel: device query failed");
    const dim3 g(uint32_t(cus) * 8u), b(kPtBlock);  // a grid-stride loop over the queued walks
    if (fr.hwss == 4) hipLaunchKernelGGL((solve_walk_unwind_kernel<4>), g, b, 0, st, sc, fr, sb);
    else hipLaunchKernelGGL((solve_walk_unwind_kernel<1>), g, b, 0, st, sc, fr, sb);
    return check_launch("solve_walk_unwind_kernel");
}

int launch_solve_sum(const PtFrame& fr, const SolveBuffers& sb, hipStream_t st) {
    hipLaunchKernelGGL(solve_sum_kernel, dim3((fr.n_paths + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, st, fr, sb);
    return check_launch("solve_sum_kernel");
}

}  // namespace plt
