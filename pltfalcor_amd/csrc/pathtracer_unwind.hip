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

}  // namespace plt
