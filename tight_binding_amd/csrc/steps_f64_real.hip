// Step-kernel instantiations for fp64 vectors, real matrix values.
#include "step_launch.h"

namespace tbk {
bool launch_f64_real(int width, int kch, int cpl, int mode, dim3 grid, hipStream_t st, const StepArgs& a) {
    return launch_width<double, false>(width, kch, cpl, mode, grid, st, a);
}
bool launch_slots_f64_real(int Rp, int cpl, dim3 grid, hipStream_t st, const SlotArgs& a) {
    return launch_slots<double, false>(Rp, cpl, grid, st, a);
}
bool launch_persist_f64_real(int Rp, int cpl, dim3 grid, hipStream_t st, const PersistSched& s) {
    return launch_persist<double, false>(Rp, cpl, grid, st, s);
}
int64_t launch_sync_f64_real(int Rp, int cpl, int bpc, int num_cu, hipStream_t st, const SyncSched& s) {
    return launch_sync<double, false>(Rp, cpl, bpc, num_cu, st, s);
}
}  // namespace tbk
