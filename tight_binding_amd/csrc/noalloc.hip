// tbk_moments_noalloc: the moment recursion of tbk_moments for callers that
// cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).
//
// tbk_moments keeps two rolling vector blocks and the per-step moment
// partials in buffers the context grows on demand.  Here the caller passes
// the two rolling blocks (they are outputs: on return they hold the last two
// Chebyshev blocks) and a device array for the n moments; the partials and
// per-step sums live in `scratch`, a fixed array in device memory with one
// copy per GPU, whose address is looked up once per device.
//
// Partials are held for kScratchRing steps and then reduced, chunk by
// chunk, by the same reduce_partials_kernel that tbk_moments launches over
// all steps at once (each step's sum depends only on that step's partials,
// summed in index order), and the moments are formed on the device by the
// same arithmetic the host applies in tbk_moments, so the moments are the
// same bits.  Nothing is allocated, nothing is copied to the host and the
// stream is not synchronised.
#include <hip/hip_runtime.h>

#include <mutex>

#include "aux_kernels.h"
#include "engine.h"
#include "tbkpm.h"

using namespace engine;

namespace {

constexpr int64_t kScratchRing = 32;             // steps of partials held before one reduce launch
constexpr int64_t kScratchMaxWorkgroups = 4096;  // workgroups of one step launch
constexpr int64_t kScratchMaxSteps = 32768;      // per-step sums: n <= 65536 doubling, 32769 plain
constexpr int64_t kScratchElems = kScratchRing * kScratchMaxWorkgroups + kScratchMaxSteps;  // 2.5 MiB

// One copy per GPU, held until the process ends.  Host code never uses this
// name as an address: scratch_on() looks the device address up.
static __device__ __align__(256) double2 scratch[kScratchElems];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double2* g_scratch_ptr[kMaxDevices] = {};

// The address of `scratch` on the current device (set by the caller), looked
// up on the first call on that device.
double2* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) raise(TBK_E_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) raise(TBK_E_DEVICE, "hipGetSymbolAddress returned a null scratch address");
        g_scratch_ptr[device] = static_cast<double2*>(p);
    }
    return g_scratch_ptr[device];
}

// mu[i] from the per-step sums, as tbk_moments forms them on the host:
// mu_0, mu_1 from step 0; doubling: mu_2k = 2 s_k.x - mu_0,
// mu_2k+1 = 2 s_k.y - mu_1; one moment per step: mu_k+1 = s_k.x.
__global__ __launch_bounds__(256) void moments_from_sums_kernel(const double2* __restrict__ sums, int32_t n,
                                                                int32_t doubling, double* __restrict__ mu) {
    const double mu0 = sums[0].x, mu1 = sums[0].y;
    for (int32_t i = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x); i < n;
         i += static_cast<int32_t>(gridDim.x * blockDim.x)) {
        double v;
        if (i == 0)
            v = mu0;
        else if (i == 1)
            v = mu1;
        else if (doubling)
            v = (i & 1) ? 2.0 * sums[i / 2].y - mu1 : 2.0 * sums[i / 2].x - mu0;
        else
            v = sums[i - 1].x;
        mu[i] = v;
    }
}

}  // namespace

extern "C" int tbk_moments_noalloc(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* bras,
                                   const tbk_block* kets, int32_t n, int32_t flags, void* work0, void* work1,
                                   double* mu_dev, void* stream) {
    return guarded([&] {
        if (!ctx || !M || !kets || !work0 || !work1 || !mu_dev) raise(TBK_E_INVALID, "null argument");
        if (n < 2) raise(TBK_E_INVALID, "n must be greater than or equal to 2");
        if (!(a != 0.0) || !std::isfinite(a) || !std::isfinite(b)) raise(TBK_E_INVALID, "invalid rescaling a, b");
        if (M->ctx != ctx || kets->ctx != ctx || (bras && bras->ctx != ctx))
            raise(TBK_E_INVALID, "objects belong to another context");
        if (kets->rows != M->rows) raise(TBK_E_INVALID, "kets rows do not match the matrix");
        if (bras && (bras->rows != kets->rows || bras->R != kets->R || bras->precision != kets->precision))
            raise(TBK_E_INVALID, "bras shape or precision does not match kets");
        if (work0 == work1 || work0 == kets->data.p || work1 == kets->data.p ||
            (bras && (work0 == bras->data.p || work1 == bras->data.p)))
            raise(TBK_E_INVALID, "work blocks must be distinct from each other and from the kets and bras");
        if (flags & TBK_FLAG_TIME_LAUNCHES) raise(TBK_E_INVALID, "per-launch timing needs events: use tbk_moments");
        if (kets->precision == TBK_FP32 && !M->have32)
            raise(TBK_E_INVALID, "the matrix's fp32 values are formed on its first fp32 tbk_moments call");

        const bool doubling = (bras == nullptr || bras == kets) && !(flags & TBK_FLAG_PLAIN);
        const int64_t steps = doubling ? (static_cast<int64_t>(n) + 1) / 2 : (static_cast<int64_t>(n) - 1);
        if (steps > kScratchMaxSteps)
            raise(TBK_E_UNSUPPORTED, "more than 32768 recursion steps (n > 65536 with doubling, n > 32769 plain)");
        set_device(ctx);
        const StepPlan plan = make_plan(ctx, M, kets->precision, kets->Rp, kets->rows, b,
                                        (flags & TBK_FLAG_ROW_KERNEL) != 0, ctx->tuning, true);
        const int64_t nb = plan.nb;
        if (nb > kScratchMaxWorkgroups) raise(TBK_E_UNSUPPORTED, "step launch wider than 4096 workgroups");
        double2* ring = scratch_on(ctx->device);
        double2* sums = ring + kScratchRing * kScratchMaxWorkgroups;
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;

        const tbk_block* bra = bras ? bras : kets;
        auto buf = [&](int64_t k) -> void* {  // storage of T_k, as in tbk_moments
            if (k == 0) return kets->data.p;
            return (k & 1) ? work1 : work0;
        };
        for (int64_t k = 0; k < steps; ++k) {
            const int64_t slot = k % kScratchRing;
            double2* partial = ring + slot * nb;
            if (k == 0)
                launch_plan(plan, st, tbk::kFirst, 1.0 / a, -b / a, buf(0), nullptr, buf(1), bra->data.p, partial);
            else
                launch_plan(plan, st, doubling ? tbk::kRecDbl : tbk::kRec, 2.0 / a, -2.0 * b / a, buf(k),
                            buf(k - 1), buf(k + 1), bra->data.p, partial);
            if (slot == kScratchRing - 1 || k == steps - 1) {
                hipLaunchKernelGGL(tbk::reduce_partials_kernel, dim3(static_cast<unsigned>(slot + 1)),
                                   dim3(tbk::kBlock), 0, st, static_cast<const double2*>(ring),
                                   static_cast<int32_t>(nb), sums + (k - slot));
                HIPC(hipGetLastError());
            }
        }
        const unsigned grid = static_cast<unsigned>(std::min<int64_t>(64, (static_cast<int64_t>(n) + 255) / 256));
        hipLaunchKernelGGL(moments_from_sums_kernel, dim3(grid), dim3(256), 0, st,
                           static_cast<const double2*>(sums), n, static_cast<int32_t>(doubling), mu_dev);
        HIPC(hipGetLastError());
    });
}

extern "C" int tbk_block_device_ptr(const tbk_block* blk, void** out) {
    return guarded([&] {
        if (!blk || !out) raise(TBK_E_INVALID, "null argument");
        *out = blk->data.p;
    });
}
