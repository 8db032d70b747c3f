// Step-kernel launchers: StepArgs and the template switch from runtime
// (width, kernel family, mode) to a kernel instantiation.  The instantiations
// are split over four translation units (steps_f64_real.hip, ...) so that
// they compile in parallel; tbkpm.hip calls launch_steps().
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "cheb_kernels.h"

namespace tbk {

struct StepArgs {
    const int32_t* rp;
    const int32_t* ci;
    const void* hv;
    const void* x;
    const void* z;
    void* y;
    const void* bra;
    int32_t rows, Rp;
    double alpha, beta;
    int32_t opts;  // kOptShift | kOptStream
    int64_t ntiles;
    TileOrder order;
    int32_t xcd_split;
    double2* partial;
    const void* hv2 = nullptr;  // kDual: values of the second matrix (same pattern)
    void* y2 = nullptr;         // kDual: its product with x
    void* y3 = nullptr;         // kPairFirst, kPairRec: the right half's split words
    int64_t wbs = 0;            // split words: block stride of a k-blocked stack (word_offset), 0 contiguous
};

// Arguments of cheb_slots_kernel (one step of every slot of a class; the
// grid is the total of the slots' workgroups).
struct SlotArgs {
    const SlotDesc* slots;
    const int32_t* wg_slot;
    int32_t s;
    int32_t Rp;
    int32_t opts;  // kOptStream
    double2* partial;
};

bool launch_slots_f64_real(int Rp, int cpl, dim3 grid, hipStream_t st, const SlotArgs& a);
bool launch_slots_f64_cplx(int Rp, int cpl, dim3 grid, hipStream_t st, const SlotArgs& a);
bool launch_slots_f32_real(int Rp, int cpl, dim3 grid, hipStream_t st, const SlotArgs& a);
bool launch_slots_f32_cplx(int Rp, int cpl, dim3 grid, hipStream_t st, const SlotArgs& a);

template <typename VT, bool CH>
bool launch_slots(int Rp, int cpl, dim3 grid, hipStream_t st, const SlotArgs& a) {
#define TBK_SLOT_CASE(RP, C)                                                                                 \
    if (Rp == RP && cpl == C) {                                                                              \
        hipLaunchKernelGGL((cheb_slots_kernel<VT, CH, RP / C, C>), grid, dim3(kBlock), 0, st, a.slots,       \
                           a.wg_slot, a.s, a.Rp, a.opts, a.partial);                                         \
        return true;                                                                                         \
    }
    TBK_SLOT_CASE(1, 1)
    TBK_SLOT_CASE(2, 1)
    TBK_SLOT_CASE(2, 2)
    TBK_SLOT_CASE(4, 1)
    TBK_SLOT_CASE(4, 2)
    TBK_SLOT_CASE(8, 1)
    TBK_SLOT_CASE(8, 2)
    TBK_SLOT_CASE(8, 4)
    TBK_SLOT_CASE(16, 1)
    TBK_SLOT_CASE(16, 2)
    TBK_SLOT_CASE(16, 4)
    TBK_SLOT_CASE(32, 1)
    TBK_SLOT_CASE(32, 2)
    TBK_SLOT_CASE(32, 4)
    TBK_SLOT_CASE(64, 1)
    TBK_SLOT_CASE(64, 2)
    TBK_SLOT_CASE(128, 2)
    TBK_SLOT_CASE(128, 4)
    TBK_SLOT_CASE(256, 4)
#undef TBK_SLOT_CASE
    return false;
}

bool launch_persist_f64_real(int Rp, int cpl, dim3 grid, hipStream_t st, const PersistSched& s);
bool launch_persist_f64_cplx(int Rp, int cpl, dim3 grid, hipStream_t st, const PersistSched& s);
bool launch_persist_f32_real(int Rp, int cpl, dim3 grid, hipStream_t st, const PersistSched& s);
bool launch_persist_f32_cplx(int Rp, int cpl, dim3 grid, hipStream_t st, const PersistSched& s);

// the persistent batched recursion for a slot class (the widths of cheb_slots_kernel)
template <typename VT, bool CH>
bool launch_persist(int Rp, int cpl, dim3 grid, hipStream_t st, const PersistSched& s) {
#define TBK_PERSIST_CASE(RP, C)                                                                               \
    if (Rp == RP && cpl == C) {                                                                               \
        hipLaunchKernelGGL((cheb_slots_persistent_kernel<VT, CH, RP / C, C>), grid, dim3(kBlock), 0, st, s); \
        return true;                                                                                          \
    }
    TBK_PERSIST_CASE(1, 1)
    TBK_PERSIST_CASE(2, 1)
    TBK_PERSIST_CASE(2, 2)
    TBK_PERSIST_CASE(4, 1)
    TBK_PERSIST_CASE(4, 2)
    TBK_PERSIST_CASE(8, 1)
    TBK_PERSIST_CASE(8, 2)
    TBK_PERSIST_CASE(8, 4)
    TBK_PERSIST_CASE(16, 1)
    TBK_PERSIST_CASE(16, 2)
    TBK_PERSIST_CASE(16, 4)
    TBK_PERSIST_CASE(32, 1)
    TBK_PERSIST_CASE(32, 2)
    TBK_PERSIST_CASE(32, 4)
    TBK_PERSIST_CASE(64, 1)
    TBK_PERSIST_CASE(64, 2)
    TBK_PERSIST_CASE(128, 2)
    TBK_PERSIST_CASE(128, 4)
    TBK_PERSIST_CASE(256, 4)
#undef TBK_PERSIST_CASE
    return false;
}

// The grid-synchronised batched recursion (cheb_slots_sync_kernel) for a
// slot class: grid = min(bpc, resident workgroups per CU of the kernel)
// workgroups per CU times num_cu, rounded down to a multiple of 8 (all of
// them resident, so that every grid barrier can complete).  Returns the
// grid launched, 0 if no kernel exists for the width.
int64_t launch_sync_f64_real(int Rp, int cpl, int bpc, int num_cu, hipStream_t st, const SyncSched& s);
int64_t launch_sync_f64_cplx(int Rp, int cpl, int bpc, int num_cu, hipStream_t st, const SyncSched& s);
int64_t launch_sync_f32_real(int Rp, int cpl, int bpc, int num_cu, hipStream_t st, const SyncSched& s);
int64_t launch_sync_f32_cplx(int Rp, int cpl, int bpc, int num_cu, hipStream_t st, const SyncSched& s);

template <typename VT, bool CH, int LPR, int CPL>
int64_t launch_sync_one(int bpc, int num_cu, hipStream_t st, const SyncSched& s) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, cheb_slots_sync_kernel<VT, CH, LPR, CPL>, kBlock, 0) !=
            hipSuccess ||
        per_cu < 1)
        return 0;
    const int64_t grid = (static_cast<int64_t>(std::min(bpc, per_cu)) * num_cu / 8) * 8;
    if (grid < 8) return 0;
    hipLaunchKernelGGL((cheb_slots_sync_kernel<VT, CH, LPR, CPL>), dim3(static_cast<unsigned>(grid)), dim3(kBlock), 0,
                       st, s);
    return grid;
}

template <typename VT, bool CH>
int64_t launch_sync(int Rp, int cpl, int bpc, int num_cu, hipStream_t st, const SyncSched& s) {
#define TBK_SYNC_CASE(RP, C) \
    if (Rp == RP && cpl == C) return launch_sync_one<VT, CH, RP / C, C>(bpc, num_cu, st, s);
    TBK_SYNC_CASE(1, 1)
    TBK_SYNC_CASE(2, 1)
    TBK_SYNC_CASE(2, 2)
    TBK_SYNC_CASE(4, 1)
    TBK_SYNC_CASE(4, 2)
    TBK_SYNC_CASE(8, 1)
    TBK_SYNC_CASE(8, 2)
    TBK_SYNC_CASE(8, 4)
    TBK_SYNC_CASE(16, 1)
    TBK_SYNC_CASE(16, 2)
    TBK_SYNC_CASE(16, 4)
    TBK_SYNC_CASE(32, 1)
    TBK_SYNC_CASE(32, 2)
    TBK_SYNC_CASE(32, 4)
    TBK_SYNC_CASE(64, 1)
    TBK_SYNC_CASE(64, 2)
    TBK_SYNC_CASE(128, 2)
    TBK_SYNC_CASE(128, 4)
    TBK_SYNC_CASE(256, 4)
#undef TBK_SYNC_CASE
    return 0;
}

// Launch one step; returns false if no kernel exists for the combination.
bool launch_f64_real(int width, int kch, int cpl, int mode, dim3 grid, hipStream_t st, const StepArgs& a);
bool launch_f64_cplx(int width, int kch, int cpl, int mode, dim3 grid, hipStream_t st, const StepArgs& a);
bool launch_f32_real(int width, int kch, int cpl, int mode, dim3 grid, hipStream_t st, const StepArgs& a);
bool launch_f32_cplx(int width, int kch, int cpl, int mode, dim3 grid, hipStream_t st, const StepArgs& a);

// KCH = 0: cheb_step_kernel (rows of any length); KCH in {1, 2, 4}:
// cheb_tile_kernel for rows of at most 4*KCH entries.
template <typename VT, bool CH, int LPR, int CPL, int MODE, int KCH>
bool launch_step(dim3 grid, hipStream_t st, const StepArgs& a) {
    using V2 = typename Vec2<VT>::type;
    using H = typename HVal<VT, CH>::type;
    if constexpr (KCH == 0) {
        hipLaunchKernelGGL((cheb_step_kernel<VT, CH, LPR, CPL, MODE>), grid, dim3(kBlock), 0, st, a.rp, a.ci,
                           static_cast<const H*>(a.hv), static_cast<const V2*>(a.x), static_cast<const V2*>(a.z),
                           static_cast<V2*>(a.y), static_cast<const V2*>(a.bra), a.rows, a.Rp,
                           static_cast<VT>(a.alpha), static_cast<VT>(a.beta), a.opts, a.ntiles, a.order, a.xcd_split,
                           a.partial, static_cast<const H*>(a.hv2), static_cast<V2*>(a.y2), static_cast<V2*>(a.y3),
                           a.wbs);
        return true;
    } else if constexpr (MODE >= kDual) {
        return false;  // the dual product and split outputs exist for the row kernel only
    } else if constexpr (tile_fits(LPR, CPL, KCH, sizeof(H))) {
        hipLaunchKernelGGL((cheb_tile_kernel<VT, CH, LPR, CPL, MODE, KCH>), grid, dim3(kBlock), 0, st, a.rp, a.ci,
                           static_cast<const H*>(a.hv), static_cast<const V2*>(a.x), static_cast<const V2*>(a.z),
                           static_cast<V2*>(a.y), static_cast<const V2*>(a.bra), a.rows, a.Rp,
                           static_cast<VT>(a.alpha), static_cast<VT>(a.beta), a.opts, a.ntiles, a.order,
                           a.xcd_split, a.partial);
        return true;
    } else {
        return false;
    }
}

template <typename VT, bool CH, int LPR, int CPL, int KCH>
bool launch_mode(int mode, dim3 grid, hipStream_t st, const StepArgs& a) {
    switch (mode) {
        case kFirst: return launch_step<VT, CH, LPR, CPL, kFirst, KCH>(grid, st, a);
        case kRec: return launch_step<VT, CH, LPR, CPL, kRec, KCH>(grid, st, a);
        // the dual product and the pre-split stack outputs: fp64 complex row
        // kernel (the Kubo stacks) only
        case kDual:
            if constexpr (KCH == 0 && CH && sizeof(VT) == 8) return launch_step<VT, CH, LPR, CPL, kDual, 0>(grid, st, a);
            return false;
        case kFirstSplit:
            if constexpr (KCH == 0 && CH && sizeof(VT) == 8)
                return launch_step<VT, CH, LPR, CPL, kFirstSplit, 0>(grid, st, a);
            return false;
        case kRecDblSplit:
            if constexpr (KCH == 0 && CH && sizeof(VT) == 8)
                return launch_step<VT, CH, LPR, CPL, kRecDblSplit, 0>(grid, st, a);
            return false;
        case kDualSplit:
            if constexpr (KCH == 0 && CH && sizeof(VT) == 8)
                return launch_step<VT, CH, LPR, CPL, kDualSplit, 0>(grid, st, a);
            return false;
        // the paired chains: complex row kernel, fp64 or fp32 vectors
        case kPairFirst:
            if constexpr (KCH == 0 && CH) return launch_step<VT, CH, LPR, CPL, kPairFirst, 0>(grid, st, a);
            return false;
        case kPairRec:
            if constexpr (KCH == 0 && CH) return launch_step<VT, CH, LPR, CPL, kPairRec, 0>(grid, st, a);
            return false;
        default: return launch_step<VT, CH, LPR, CPL, kRecDbl, KCH>(grid, st, a);
    }
}

template <typename VT, bool CH, int LPR, int CPL>
bool launch_kch(int kch, int mode, dim3 grid, hipStream_t st, const StepArgs& a) {
    switch (kch) {
        case 0: return launch_mode<VT, CH, LPR, CPL, 0>(mode, grid, st, a);
        case 1: return launch_mode<VT, CH, LPR, CPL, 1>(mode, grid, st, a);
        case 2: return launch_mode<VT, CH, LPR, CPL, 2>(mode, grid, st, a);
        case 4: return launch_mode<VT, CH, LPR, CPL, 4>(mode, grid, st, a);
        default: return false;
    }
}

// width: padded column count for the row kernel, slab width for the tiled one.
// cpl: complex values per lane of the row kernel; the default (1, or width/64
// above 64 columns) uses the width switch below, larger values put more rows
// in one wave (LPR = width / cpl lanes per row).
template <typename VT, bool CH>
bool launch_width(int width, int kch, int cpl, int mode, dim3 grid, hipStream_t st, const StepArgs& a) {
    if (kch == 0 && cpl > (width > kWave ? width / kWave : 1)) {
        switch (width * 8 + cpl) {
            case 2 * 8 + 2: return launch_mode<VT, CH, 1, 2, 0>(mode, grid, st, a);
            case 4 * 8 + 2: return launch_mode<VT, CH, 2, 2, 0>(mode, grid, st, a);
            case 4 * 8 + 4: return launch_mode<VT, CH, 1, 4, 0>(mode, grid, st, a);
            case 8 * 8 + 2: return launch_mode<VT, CH, 4, 2, 0>(mode, grid, st, a);
            case 8 * 8 + 4: return launch_mode<VT, CH, 2, 4, 0>(mode, grid, st, a);
            case 16 * 8 + 2: return launch_mode<VT, CH, 8, 2, 0>(mode, grid, st, a);
            case 16 * 8 + 4: return launch_mode<VT, CH, 4, 4, 0>(mode, grid, st, a);
            case 32 * 8 + 2: return launch_mode<VT, CH, 16, 2, 0>(mode, grid, st, a);
            case 32 * 8 + 4: return launch_mode<VT, CH, 8, 4, 0>(mode, grid, st, a);
            case 64 * 8 + 2: return launch_mode<VT, CH, 32, 2, 0>(mode, grid, st, a);
            case 64 * 8 + 4: return launch_mode<VT, CH, 16, 4, 0>(mode, grid, st, a);
            case 128 * 8 + 4: return launch_mode<VT, CH, 32, 4, 0>(mode, grid, st, a);
            default: return false;
        }
    }
    switch (width) {
        case 1: return launch_kch<VT, CH, 1, 1>(kch, mode, grid, st, a);
        case 2: return launch_kch<VT, CH, 2, 1>(kch, mode, grid, st, a);
        case 4: return launch_kch<VT, CH, 4, 1>(kch, mode, grid, st, a);
        case 8: return launch_kch<VT, CH, 8, 1>(kch, mode, grid, st, a);
        case 16: return launch_kch<VT, CH, 16, 1>(kch, mode, grid, st, a);
        case 32: return launch_kch<VT, CH, 32, 1>(kch, mode, grid, st, a);
        case 64: return launch_kch<VT, CH, 64, 1>(kch, mode, grid, st, a);
        case 128: return launch_kch<VT, CH, 64, 2>(kch, mode, grid, st, a);
        case 256: return launch_kch<VT, CH, 64, 4>(kch, mode, grid, st, a);
        default: return false;
    }
}

}  // namespace tbk
