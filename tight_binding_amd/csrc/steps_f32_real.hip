// Step-kernel instantiations for fp32 vectors, real matrix values.
#include "step_launch.h"

namespace tbk {
bool launch_f32_real(int width, int kch, int mode, dim3 grid, hipStream_t st, const StepArgs& a) {
    return launch_width<float, false>(width, kch, mode, grid, st, a);
}
}  // namespace tbk
