// Chebyshev reconstruction of a spectral function from its moments, on the
// device: the evaluation half of GreenFunctionImag (lattice.h:616-640).
//
// With e = cos(theta), Im alpha_0(e) = -1/sin(theta) and
// Im alpha_n(e) = -2 cos(n theta)/sin(theta) (numerical.h:304-316), so the
// reference's direct sum (operator(), lattice.h:617-623) is
//   f(e) = sum_n g_n Im alpha_n(e) mu_n = -r(e),
//   r(e) = [g_0 mu_0 + 2 sum_{n>=1} g_n mu_n cos(n theta)] / sin(theta),
// and its fft() (lattice.h:625-640) is r at the nodes
// e_k = cos((k + 1/2) pi / N): FFTW REDFT01 of g_n mu_n,
// y_k = x_0 + 2 sum_{n>=1} x_n cos(pi n (k + 1/2) / N), times 1/sin(theta_k).
// DoS(e) = -f(e)/(R pi) = r(e)/(R pi) (functions.h:27-30).
//
// One workgroup per energy: each thread sums the moments n = t, t + 256, ...
// with cos(n theta) from the library cosine (no recurrence, so no error
// growth with n), then a fixed-order tree; deterministic.  Out of (-1, 1) the
// value is not finite, as the reference's is at e = -1.
//
// Node mode (no energies given): theta_k = (k + 1/2) pi / n_e directly, with
// sin(theta_k) in place of sqrt(1 - e_k^2).  Near the band edges an energy
// carries its rounding into theta = acos(e) amplified by 1/sin(theta), and n
// theta multiplies that by up to N: at N = 2048 the edge nodes would move by
// ~3e-10 rad.  fft() is defined on the nodes, so it never forms e.  There
// n theta_k = m pi / (2 n_e) with m = n (2k + 1) mod 4 n_e reduced exactly in
// integers, so each cosine's argument is rounded once, below 2 pi: the phase
// error stays ~1e-16 for every n instead of growing as n theta_k eps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "engine.h"
#include "tbkpm.h"

namespace tbk {
namespace {

__global__ __launch_bounds__(kBlock) void cheb_reconstruct_kernel(const double* __restrict__ mu,
                                                                  const double* __restrict__ g, int32_t N,
                                                                  const double* __restrict__ e, int32_t n_e,
                                                                  double scale, double* __restrict__ out) {
    __shared__ double sh[kBlock];
    for (int32_t j = blockIdx.x; j < n_e; j += gridDim.x) {
        const double th = e ? acos(e[j]) : (static_cast<double>(j) + 0.5) * (3.141592653589793 / n_e);
        const double sn = e ? sqrt((1.0 - e[j]) * (1.0 + e[j])) : sin(th);
        const double quarter = 3.141592653589793 / (2.0 * n_e);  // node mode: phase unit pi / (2 n_e)
        const int64_t period = 4 * static_cast<int64_t>(n_e), odd = 2 * static_cast<int64_t>(j) + 1;
        double acc = 0.0;
        for (int32_t n = threadIdx.x; n < N; n += kBlock) {
            const double w = (g ? g[n] : 1.0) * mu[n];
            const double c = e ? cos(static_cast<double>(n) * th)
                               : cos(static_cast<double>((static_cast<int64_t>(n) * odd) % period) * quarter);
            acc = acc + (n == 0 ? w : 2.0 * w * c);
        }
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int s = kBlock / 2; s >= 1; s >>= 1) {
            if (static_cast<int>(threadIdx.x) < s) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[j] = scale * sh[0] / sn;
        __syncthreads();
    }
}

}  // namespace
}  // namespace tbk

using namespace engine;

extern "C" {

int tbk_cheb_reconstruct(tbk_context* ctx, const double* mu, const double* g, int32_t N, const double* e,
                         int32_t n_e, double scale, double* out, double* kernel_ms) {
    return guarded([&] {
        if (!mu || !out) raise(TBK_E_INVALID, "null argument");
        if (N < 1 || n_e < 0) raise(TBK_E_INVALID, "N must be >= 1 and n_e >= 0");
        if (!ctx) ctx = default_context();
        set_device(ctx);
        if (n_e == 0) {
            if (kernel_ms) *kernel_ms = 0.0;
            return;
        }
        DevBuf d;
        const size_t nb = static_cast<size_t>(N) * 8, eb = static_cast<size_t>(n_e) * 8;
        d.reserve(2 * nb + 2 * eb);
        double* dmu = d.as<double>();
        double* dg = dmu + N;
        double* de = dg + N;
        double* dout = de + n_e;
        hipStream_t st = ctx->stream;
        HIPC(hipMemcpyAsync(dmu, mu, nb, hipMemcpyHostToDevice, st));
        if (g) HIPC(hipMemcpyAsync(dg, g, nb, hipMemcpyHostToDevice, st));
        if (e) HIPC(hipMemcpyAsync(de, e, eb, hipMemcpyHostToDevice, st));
        ensure_events(ctx, 2);
        HIPC(hipEventRecord(ctx->events[0], st));
        const unsigned grid = static_cast<unsigned>(std::min<int32_t>(n_e, 65535));
        hipLaunchKernelGGL(tbk::cheb_reconstruct_kernel, dim3(grid), dim3(tbk::kBlock), 0, st, dmu,
                           g ? dg : nullptr, N, e ? de : nullptr, n_e, scale, dout);
        HIPC(hipGetLastError());
        HIPC(hipEventRecord(ctx->events[1], st));
        HIPC(hipMemcpyAsync(out, dout, eb, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        if (kernel_ms) {
            float ms = 0.0f;
            HIPC(hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
            *kernel_ms = ms;
        }
    });
}

}  // extern "C"
