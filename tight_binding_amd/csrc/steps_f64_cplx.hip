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
int64_t launch_sync_f64_cplx(int Rp, int cpl, int bpc, int num_cu, hipStream_t st, const SyncSched& s) {
    return launch_sync<double, true>(Rp, cpl, bpc, num_cu, st, s);
}
}  // namespace tbk
