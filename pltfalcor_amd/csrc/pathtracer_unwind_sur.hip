// The split solve pass's unwind with the learned BSDF surrogate: clear the query counters, the
// query pass (the unwind appending the network's inputs of every eligible step), one batched
// inference on the matrix cores, and the apply pass (the unwind using the network's matrices).
#include "pathtracer_split.hpp"
#include "surrogate.hpp"

namespace plt {

int launch_solve_unwind_surrogate(const PtScene& sc, const PtFrame& fr, const SolveBuffers& sb, plt_surrogate* s, hipStream_t st) {
    const dim3 g((fr.n_paths + kPtBlock - 1) / kPtBlock), b(kPtBlock);
    hipLaunchKernelGGL(sur_clear_kernel, dim3(1), dim3(64), 0, st, sb.sur.counter);
    if (fr.hwss == 4) hipLaunchKernelGGL((solve_unwind_kernel<4, kUnwindQuery>), g, b, 0, st, sc, fr, sb);
    else hipLaunchKernelGGL((solve_unwind_kernel<1, kUnwindQuery>), g, b, 0, st, sc, fr, sb);
    int rc = check_launch("solve_unwind_kernel (surrogate queries)");
    if (rc < 0) return rc;
    if (!sb.sur.teacher && (rc = surrogate_infer_counted(s, sb.sur.counter, sb.sur.qcap, sb.sur.X, const_cast<float*>(sb.sur.Y), st)) < 0)
        return rc;
    if (fr.hwss == 4) hipLaunchKernelGGL((solve_unwind_kernel<4, kUnwindApply>), g, b, 0, st, sc, fr, sb);
    else hipLaunchKernelGGL((solve_unwind_kernel<1, kUnwindApply>), g, b, 0, st, sc, fr, sb);
    return check_launch("solve_unwind_kernel (surrogate applied)");
}

}  // namespace plt
