// Spectral range of a Hermitian matrix: the bounds E_min, E_max behind the
// rescaling H~ = (H - b)/a of every Chebyshev expansion.
//
// The reference gets them from numerical::linalg::matrix_spectral_range_sparse
// (numerical.h:23-46): Spectra's SymEigsSolver for the two extreme
// eigenvalues, which refuses matrices above 5000 rows and then returns
// {0, 0}, so GreenFunction(N, HE, ref_HE, epsilon) (lattice.h:560-569) divides
// by a = 0.  Complex matrices are first embedded as 2n x 2n real ones
// (numerical.h:42-44, 51-72), which doubles every eigenvalue's multiplicity
// and leaves the extremes unchanged.
//
// Here: FP64 Lanczos on the device, any size, real or complex values.
//   w = H v_j - beta_j v_{j-1};  alpha_j = Re<v_j, w>;  w -= alpha_j v_j;
//   beta_{j+1} = |w|;  v_{j+1} = w / beta_{j+1}
// from a seeded random-phase unit vector, without reorthogonalisation (lost
// orthogonality duplicates converged Ritz values but does not move the
// extreme ones outside the spectrum).  Each step is three launches whose
// reductions are fixed-order block trees followed by a fixed-order sum over
// the workgroups, so results are deterministic.  Every `check_every` steps
// the host reads alpha, beta and finds the extreme eigenvalues theta of the
// tridiagonal T_k by Sturm bisection, and the residual bound
// r = beta_{k+1} |y_k| of each (y = the unit eigenvector of T_k, from
// inverse iteration with a shift just outside the spectrum, where T - sigma
// is definite and the elimination needs no pivoting).  An eigenvalue of H
// lies within r of theta.  The Gershgorin discs, computed once at upload,
// hold the whole spectrum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "aux_kernels.h"
#include "engine.h"
#include "tbkpm.h"

namespace tbk {
namespace {

constexpr int kLanczosGroups = 1024;  // workgroups per Lanczos launch (at most)

// sum over the workgroup of one double per thread, fixed tree; every thread
// gets the result
__device__ __forceinline__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kBlock / 2; w >= 1; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// fixed-order sum of np workgroup partials, the same in every workgroup
__device__ __forceinline__ double sum_partials(const double* __restrict__ partial, int np, double* sh) {
    double t = 0.0;
    for (int b = threadIdx.x; b < np; b += kBlock) t = t + partial[b];
    return block_sum(t, sh);
}

// w = H v - beta[j] u; partial[g] = sum over the group's rows of Re(conj(v) w)
template <bool CPLX>
__global__ __launch_bounds__(kBlock) void lanczos_apply_kernel(const int32_t* __restrict__ rp,
                                                               const int32_t* __restrict__ ci,
                                                               const double* __restrict__ hv,
                                                               const double2* __restrict__ v,
                                                               const double2* __restrict__ u,
                                                               double2* __restrict__ w,
                                                               const double* __restrict__ beta, int j,
                                                               int32_t n, double* __restrict__ partial) {
    __shared__ double sh[kBlock];
    const double bj = beta[j];
    double acc = 0.0;
    for (int32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        double sx = 0.0, sy = 0.0;
        const int32_t p1 = rp[i + 1];
        for (int32_t p = rp[i]; p < p1; ++p) {
            const double2 x = v[ci[p]];
            if (CPLX) {
                const double hr = hv[2 * static_cast<int64_t>(p)], hi = hv[2 * static_cast<int64_t>(p) + 1];
                sx = sx + (hr * x.x - hi * x.y);
                sy = sy + (hr * x.y + hi * x.x);
            } else {
                const double h = hv[p];
                sx = sx + h * x.x;
                sy = sy + h * x.y;
            }
        }
        const double2 ui = u[i];
        sx = sx - bj * ui.x;
        sy = sy - bj * ui.y;
        w[i] = make_double2(sx, sy);
        const double2 vi = v[i];
        acc = acc + (vi.x * sx + vi.y * sy);
    }
    const double r = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// alpha_j = sum of the apply partials; w -= alpha_j v; partial2 = |w|^2 per group
__global__ __launch_bounds__(kBlock) void lanczos_orth_kernel(const double2* __restrict__ v, double2* __restrict__ w,
                                                              const double* __restrict__ partial, int np,
                                                              double* __restrict__ alpha, int j, int32_t n,
                                                              double* __restrict__ partial2) {
    __shared__ double sh[kBlock];
    const double a = sum_partials(partial, np, sh);
    if (blockIdx.x == 0 && threadIdx.x == 0) alpha[j] = a;
    double acc = 0.0;
    for (int32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double2 vi = v[i];
        double2 wi = w[i];
        wi.x = wi.x - a * vi.x;
        wi.y = wi.y - a * vi.y;
        w[i] = wi;
        acc = acc + (wi.x * wi.x + wi.y * wi.y);
    }
    const double r = block_sum(acc, sh);
    if (threadIdx.x == 0) partial2[blockIdx.x] = r;
}

// beta_{j+1} = |w|; vnext = w / beta_{j+1} (zero when beta_{j+1} = 0)
__global__ __launch_bounds__(kBlock) void lanczos_next_kernel(const double2* __restrict__ w,
                                                              double2* __restrict__ vnext,
                                                              const double* __restrict__ partial2, int np,
                                                              double* __restrict__ beta, int j, int32_t n) {
    __shared__ double sh[kBlock];
    const double b = sqrt(sum_partials(partial2, np, sh));
    if (blockIdx.x == 0 && threadIdx.x == 0) beta[j + 1] = b;
    const double inv = b > 0.0 ? 1.0 / b : 0.0;
    for (int32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double2 wi = w[i];
        vnext[i] = make_double2(wi.x * inv, wi.y * inv);
    }
}

}  // namespace
}  // namespace tbk

namespace engine {

// the one-shot range (tbk_spectral_range): Lanczos steps, residual tolerance
// relative to the spectral scale, start-vector seed
constexpr int kRangeIters = 1000;
constexpr double kRangeTol = 1e-8;
constexpr uint64_t kRangeSeed = 0x5EEDB0D5ull;

// ------------------------------------------------- tridiagonal eigen-work
// T_k: diagonal alpha[0..k), off-diagonal beta[1..k) (beta[i] couples i-1, i)

// number of eigenvalues of T_k below x (Sturm sequence of the LDL^T pivots)
int sturm_below(const double* alpha, const double* beta, int k, double x) {
    int c = 0;
    double d = 1.0;
    for (int i = 0; i < k; ++i) {
        const double off = i > 0 ? beta[i] * beta[i] / d : 0.0;
        d = alpha[i] - x - off;
        if (d == 0.0) d = -1e-300;  // x is an eigenvalue of the leading block: count it as below
        if (d < 0.0) ++c;
    }
    return c;
}

// the m-th smallest eigenvalue (m = 0 .. k-1) of T_k by bisection on [lo, hi]
double tridiag_eigenvalue(const double* alpha, const double* beta, int k, int m, double lo, double hi) {
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;  // adjacent doubles
        if (sturm_below(alpha, beta, k, mid) > m)
            hi = mid;
        else
            lo = mid;
    }
    return 0.5 * (lo + hi);
}

// |last component| of the unit eigenvector of T_k for its extreme eigenvalue
// theta: inverse iteration with sigma = theta +- delta just outside the
// spectrum of T_k, where T - sigma is definite and LDL^T without pivoting is
// stable.
double tridiag_last_component(const double* alpha, const double* beta, int k, double theta, double delta) {
    if (k == 1) return 1.0;
    const double sigma = theta + delta;
    std::vector<double> y(static_cast<size_t>(k), 1.0), d(static_cast<size_t>(k)), l(static_cast<size_t>(k));
    // factor T - sigma = L D L^T
    d[0] = alpha[0] - sigma;
    for (int i = 1; i < k; ++i) {
        l[i] = beta[i] / d[i - 1];
        d[i] = alpha[i] - sigma - l[i] * beta[i];
    }
    for (int rep = 0; rep < 3; ++rep) {
        for (int i = 1; i < k; ++i) y[i] -= l[i] * y[i - 1];
        for (int i = 0; i < k; ++i) y[i] /= d[i];
        for (int i = k - 2; i >= 0; --i) y[i] -= l[i + 1] * y[i + 1];
        double nrm = 0.0;
        for (double t : y) nrm += t * t;
        nrm = std::sqrt(nrm);
        if (!(nrm > 0.0) || !std::isfinite(nrm)) return 1.0;  // no estimate: report the worst case
        for (double& t : y) t /= nrm;
    }
    return std::fabs(y[static_cast<size_t>(k - 1)]);
}

struct RitzEnds {
    double tmin, tmax, rmin, rmax;
};

RitzEnds ritz_ends(const std::vector<double>& alpha, const std::vector<double>& beta, int k) {
    // Gershgorin interval of T_k for the bisection
    double lo = alpha[0], hi = alpha[0];
    for (int i = 0; i < k; ++i) {
        const double r = (i > 0 ? std::fabs(beta[i]) : 0.0) + (i + 1 < k ? std::fabs(beta[i + 1]) : 0.0);
        lo = std::min(lo, alpha[i] - r);
        hi = std::max(hi, alpha[i] + r);
    }
    const double pad = 1e-12 * std::max(1.0, std::max(std::fabs(lo), std::fabs(hi)));
    lo -= pad;
    hi += pad;
    RitzEnds e{};
    e.tmin = tridiag_eigenvalue(alpha.data(), beta.data(), k, 0, lo, hi);
    e.tmax = tridiag_eigenvalue(alpha.data(), beta.data(), k, k - 1, lo, hi);
    const double scale = std::max({std::fabs(e.tmin), std::fabs(e.tmax), e.tmax - e.tmin, 1e-300});
    const double delta = 1e-9 * scale;
    e.rmin = std::fabs(beta[k]) * tridiag_last_component(alpha.data(), beta.data(), k, e.tmin, -delta);
    e.rmax = std::fabs(beta[k]) * tridiag_last_component(alpha.data(), beta.data(), k, e.tmax, delta);
    return e;
}

void spectral_bounds(tbk_context* ctx, const tbk_matrix* M, int max_iters, double tol, uint64_t seed,
                     hipStream_t st, tbk_bounds* out) {
    if (max_iters < 1 || max_iters > 100000) raise(TBK_E_INVALID, "max_iters must be in [1, 100000]");
    if (!(tol > 0.0) || !(tol < 1.0)) raise(TBK_E_INVALID, "tol must be in (0, 1)");
    const int32_t n = M->rows;
    const int kmax = static_cast<int>(std::min<int64_t>(max_iters, n));  // T_n is exact at k = n
    const int np = static_cast<int>(std::min<int64_t>(tbk::kLanczosGroups, (static_cast<int64_t>(n) + tbk::kBlock - 1) / tbk::kBlock));
    DevBuf vec[3], partial, coef;
    for (auto& b : vec) b.reserve(static_cast<size_t>(n) * 16);
    partial.reserve(2 * static_cast<size_t>(np) * 8);
    coef.reserve(2 * (static_cast<size_t>(kmax) + 2) * 8);
    double* dalpha = coef.as<double>();
    double* dbeta = dalpha + kmax + 2;
    double* p1 = partial.as<double>();
    double* p2 = p1 + np;
    HIPC(hipMemsetAsync(dbeta, 0, 8, st));  // beta_0 = 0
    // v_0: unit random-phase vector
    {
        const int grid = static_cast<int>(std::min<int64_t>((static_cast<int64_t>(n) + tbk::kBlock - 1) / tbk::kBlock, 65536));
        hipLaunchKernelGGL(tbk::random_phase_kernel<double2>, dim3(grid), dim3(tbk::kBlock), 0, st,
                           vec[0].as<double2>(), n, 1, 1, seed, 1.0 / std::sqrt(static_cast<double>(n)), 1, 0);
        HIPC(hipGetLastError());
    }
    const int32_t* rp = M->rp.as<const int32_t>();
    const int32_t* ci = M->ci.as<const int32_t>();
    const double* hv = M->val64.as<const double>();
    int cur = 0, prev = 0;  // step 0: u = v_0, multiplied by beta_0 = 0
    double2* w = vec[2].as<double2>();
    std::vector<double> ha(static_cast<size_t>(kmax) + 2), hb(static_cast<size_t>(kmax) + 2);
    const int check_every = 16;
    RitzEnds e{};
    int k = 0;
    bool converged = false;
    for (int j = 0; j < kmax; ++j) {
        const double2* v = vec[cur].as<const double2>();
        const double2* u = vec[prev].as<const double2>();
        if (M->is_complex)
            hipLaunchKernelGGL(tbk::lanczos_apply_kernel<true>, dim3(np), dim3(tbk::kBlock), 0, st, rp, ci, hv, v, u,
                               w, dbeta, j, n, p1);
        else
            hipLaunchKernelGGL(tbk::lanczos_apply_kernel<false>, dim3(np), dim3(tbk::kBlock), 0, st, rp, ci, hv, v,
                               u, w, dbeta, j, n, p1);
        hipLaunchKernelGGL(tbk::lanczos_orth_kernel, dim3(np), dim3(tbk::kBlock), 0, st, v, w, p1, np, dalpha, j, n,
                           p2);
        // v_{j+1} overwrites v_{j-1} (read by this step's apply, already done)
        const int nxt = (j == 0) ? 1 : prev;
        hipLaunchKernelGGL(tbk::lanczos_next_kernel, dim3(np), dim3(tbk::kBlock), 0, st, w, vec[nxt].as<double2>(),
                           p2, np, dbeta, j, n);
        HIPC(hipGetLastError());
        prev = cur;
        cur = nxt;
        k = j + 1;
        if (k % check_every == 0 || k == kmax) {
            HIPC(hipMemcpyAsync(ha.data(), dalpha, static_cast<size_t>(k) * 8, hipMemcpyDeviceToHost, st));
            HIPC(hipMemcpyAsync(hb.data(), dbeta, static_cast<size_t>(k + 1) * 8, hipMemcpyDeviceToHost, st));
            HIPC(hipStreamSynchronize(st));
            // an invariant subspace: beta_{i+1} at rounding level against the
            // entries so far makes T_{i+1}'s Ritz values eigenvalues of M;
            // the steps after it only continued from rounding noise
            double sc = 0.0;
            for (int i = 0; i < k; ++i) {
                sc = std::max({sc, std::fabs(ha[static_cast<size_t>(i)]), hb[static_cast<size_t>(i)]});
                if (hb[static_cast<size_t>(i) + 1] <= 1e-12 * sc) {
                    k = i + 1;
                    converged = true;
                    break;
                }
            }
            e = ritz_ends(ha, hb, k);
            const double scale = std::max({std::fabs(e.tmin), std::fabs(e.tmax), e.tmax - e.tmin, 1e-300});
            if (converged || (e.rmin <= tol * scale && e.rmax <= tol * scale)) {
                converged = true;
                break;
            }
        }
    }
    out->ritz_min = e.tmin;
    out->ritz_max = e.tmax;
    out->resid_min = e.rmin;
    out->resid_max = e.rmax;
    matrix_gershgorin(M, st);
    out->gersh_min = M->gersh_min;
    out->gersh_max = M->gersh_max;
    out->lower = std::max(M->gersh_min, e.tmin - e.rmin);
    out->upper = std::min(M->gersh_max, e.tmax + e.rmax);
    out->iters = k;
    out->converged = converged ? 1 : 0;
}

}  // namespace engine

using namespace engine;

extern "C" {

int tbk_spectral_bounds(tbk_context* ctx, const tbk_matrix* M, int32_t max_iters, double tol, uint64_t seed,
                        tbk_bounds* out) {
    return guarded([&] {
        if (!ctx || !M || !out) raise(TBK_E_INVALID, "null argument");
        if (M->ctx != ctx) raise(TBK_E_INVALID, "matrix belongs to another context");
        set_device(ctx);
        spectral_bounds(ctx, M, max_iters, tol, seed, ctx->stream, out);
    });
}

int tbk_spectral_range(const tbk_sparse* M, double* e_min, double* e_max) {
    return guarded([&] {
        if (!M || !e_min || !e_max) raise(TBK_E_INVALID, "null argument");
        tbk_context* ctx = default_context();
        tbk_matrix* m = nullptr;
        int rc = tbk_matrix_create(ctx, M, &m);
        if (rc != TBK_OK) raise(rc, g_last_error);
        struct Drop {
            tbk_matrix* m;
            ~Drop() { tbk_matrix_destroy(m); }
        } drop{m};
        tbk_bounds b{};
        spectral_bounds(ctx, m, kRangeIters, kRangeTol, kRangeSeed, ctx->stream, &b);
        *e_min = b.ritz_min;
        *e_max = b.ritz_max;
    });
}

}  // extern "C"
