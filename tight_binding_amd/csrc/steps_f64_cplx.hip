// Step-kernel instantiations for fp64 vectors, complex matrix values.
#include "step_launch.h"

namespace tbk {
bool launch_f64_cplx(int width, int kch, int cpl, int mode, dim3 grid, hipStream_t st, const StepArgs& a) {
    return launch_width<double, true>(width, kch, cpl, mode, grid, st, a);
}
bool launch_slots_f64_cplx(int Rp, int cpl, dim3 grid, hipStream_t st, const SlotArgs& a) {
    return launch_slots<double, true>(Rp, cpl, grid, st, a);
}
bool launch_persist_f64_cplx(int Rp, int cpl, dim3 grid, hipStream_t st, const PersistSched& s) {
    return launch_persist<double, true>(Rp, cpl, grid, st, s);
}
}  // namespace tbk
