// Operators formed on the device from resident matrices, as new resident
// matrices:
//   * H~ = (M - b I)/a over the union pattern of M and I (rescale /
//     GreenFunction::scale, lattice.h:550-559): every diagonal entry
//     explicit, (m_ii - b)/a on it, (m_ij - 0)/a elsewhere, (0 - b)/a for a
//     diagonal M lacks, the host rescale's operations in its order;
//   * the velocity operator v = X M - M X without the factor i
//     (commutator_with_diagonal, lattice.h:338-345; snippets.h:316-317) of the
//     lattice's position operator X along one axis (op_position,
//     lattice.h:314-333), x_i m_ij - m_ij x_j on M's pattern.
// Both are one pass over M's entries (HBM-bound, a few bytes per entry), so
// a sigma_xy problem's three operators never exist on the host.  Compiled
// without contraction: every value has the host's bits.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "cheb_kernels.h"
#include "engine.h"
#include "tbkpm.h"

namespace tbk {
namespace {

__device__ __forceinline__ bool row_has_diagonal(const int32_t* __restrict__ ci, int32_t p0, int32_t p1, int32_t i) {
    for (int32_t p = p0; p < p1; ++p)
        if (ci[p] == i) return true;
    return false;
}

// entries of row i of the union pattern; the longest row
__global__ __launch_bounds__(kBlock) void rescale_count_kernel(const int32_t* __restrict__ rp,
                                                               const int32_t* __restrict__ ci, int32_t n,
                                                               int32_t* __restrict__ cnt,
                                                               int32_t* __restrict__ max_row) {
    int32_t mx = 0;
    for (int64_t ii = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; ii < n;
         ii += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int32_t i = static_cast<int32_t>(ii);
        const int32_t p0 = rp[i], p1 = rp[i + 1];
        const int32_t k = (p1 - p0) + (row_has_diagonal(ci, p0, p1, i) ? 0 : 1);
        cnt[i] = k;
        mx = max(mx, k);
    }
    __shared__ int32_t sh[kBlock];
    sh[threadIdx.x] = mx;
    __syncthreads();
    for (int s = kBlock / 2; s >= 1; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(max_row, sh[0]);
}

// row i of H~ at rp2[i]: M's entries rescaled, the diagonal inserted before
// the first column past i when M lacks it
template <bool CPLX>
__global__ __launch_bounds__(kBlock) void rescale_fill_kernel(const int32_t* __restrict__ rp,
                                                              const int32_t* __restrict__ ci,
                                                              const double* __restrict__ val, int32_t n, double a,
                                                              double b, const int32_t* __restrict__ rp2,
                                                              int32_t* __restrict__ ci2, double* __restrict__ val2) {
    constexpr int W = CPLX ? 2 : 1;
    for (int64_t ii = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; ii < n;
         ii += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int32_t i = static_cast<int32_t>(ii);
        int64_t q = rp2[i];
        bool done = false;
        auto put = [&](int32_t col, double re, double im) {
            ci2[q] = col;
            val2[W * q] = re;
            if (CPLX) val2[W * q + 1] = im;
            ++q;
        };
        for (int32_t p = rp[i]; p < rp[i + 1]; ++p) {
            const int32_t j = ci[p];
            if (!done && j > i) {
                put(i, (0.0 - b) / a, (0.0 - 0.0) / a);
                done = true;
            }
            const double re = val[W * static_cast<int64_t>(p)];
            const double im = CPLX ? val[W * static_cast<int64_t>(p) + 1] : 0.0;
            if (j == i) {
                put(j, (re - b) / a, (im - 0.0) / a);
                done = true;
            } else {
                put(j, (re - 0.0) / a, (im - 0.0) / a);
            }
        }
        if (!done) put(i, (0.0 - b) / a, (0.0 - 0.0) / a);
    }
}

struct Geometry {
    int32_t dim, n_atoms, n;
    int32_t span[3];
    double base[3];  // component `axis` of each lattice vector
};

// x_s = atom_pos[a][axis] + (sum_d base_d i_d), s = (cell i, atom a): the
// host's displacement_of_lattice sum in its order, then the atom's offset
__global__ __launch_bounds__(kBlock) void position_kernel(Geometry G, const double* __restrict__ atom_axis,
                                                          double* __restrict__ x) {
    for (int64_t ss = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; ss < G.n;
         ss += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int32_t s = static_cast<int32_t>(ss);
        const int32_t a = s % G.n_atoms;
        int32_t c = s / G.n_atoms;
        int32_t iv[3] = {0, 0, 0};
        for (int d = G.dim - 1; d >= 0; --d) {
            iv[d] = c % G.span[d];
            c /= G.span[d];
        }
        double L = G.base[0] * static_cast<double>(iv[0]);
        for (int d = 1; d < G.dim; ++d) L = L + G.base[d] * static_cast<double>(iv[d]);
        x[s] = atom_axis[a] + L;
    }
}

template <bool CPLX>
__global__ __launch_bounds__(kBlock) void velocity_fill_kernel(const int32_t* __restrict__ rp,
                                                               const int32_t* __restrict__ ci,
                                                               const double* __restrict__ val, int32_t n,
                                                               const double* __restrict__ x,
                                                               double* __restrict__ out) {
    constexpr int W = CPLX ? 2 : 1;
    for (int64_t ii = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; ii < n;
         ii += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int32_t i = static_cast<int32_t>(ii);
        const double xi = x[i];
        for (int32_t p = rp[i]; p < rp[i + 1]; ++p) {
            const double xj = x[ci[p]];
            const int64_t o = W * static_cast<int64_t>(p);
            out[o] = xi * val[o] - val[o] * xj;
            if (CPLX) out[o + 1] = xi * val[o + 1] - val[o + 1] * xj;
        }
    }
}

unsigned row_grid(int64_t n) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 8192)));
}

}  // namespace
}  // namespace tbk

namespace engine {

void rescale_matrix(tbk_context* ctx, const tbk_matrix* M, double a, double b, hipStream_t st, tbk_matrix** out) {
    if (!(a != 0.0) || !std::isfinite(a) || !std::isfinite(b)) raise(TBK_E_INVALID, "invalid rescaling a, b");
    const int32_t n = M->rows;
    if (static_cast<int64_t>(M->nnz) + n > INT32_MAX) raise(TBK_E_UNSUPPORTED, "too many entries for int32 row pointers");
    auto m = std::make_unique<tbk_matrix>();
    m->ctx = ctx;
    m->rows = n;
    m->is_complex = M->is_complex;
    m->rp.reserve_placed((static_cast<size_t>(n) + 1) * 4);
    DevBuf dcnt, dmax, dtemp;
    dcnt.reserve((static_cast<size_t>(n) + 1) * 4);
    dmax.reserve(4);
    HIPC(hipMemsetAsync(dmax.p, 0, 4, st));
    HIPC(hipMemsetAsync(dcnt.as<int32_t>() + n, 0, 4, st));
    const unsigned grid = tbk::row_grid(n);
    hipLaunchKernelGGL(tbk::rescale_count_kernel, dim3(grid), dim3(tbk::kBlock), 0, st, M->rp.as<const int32_t>(),
                       M->ci.as<const int32_t>(), n, dcnt.as<int32_t>(), dmax.as<int32_t>());
    HIPC(hipGetLastError());
    size_t temp = 0;
    HIPC(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, dcnt.as<int32_t>(), m->rp.as<int32_t>(), n + 1, st));
    dtemp.reserve(std::max<size_t>(temp, 16));
    HIPC(hipcub::DeviceScan::ExclusiveSum(dtemp.p, temp, dcnt.as<int32_t>(), m->rp.as<int32_t>(), n + 1, st));
    int32_t nnz = 0, max_row = 0;
    HIPC(hipMemcpyAsync(&nnz, m->rp.as<int32_t>() + n, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(&max_row, dmax.p, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (nnz < M->nnz || nnz > M->nnz + n) raise(TBK_E_DEVICE, "rescale counted a wrong number of entries");
    m->nnz = nnz;
    m->max_row_nnz = max_row;
    const int vw = m->is_complex ? 2 : 1;
    m->ci.reserve_placed(std::max<size_t>(static_cast<size_t>(nnz) * 4, 4));
    m->val64.reserve_placed(std::max<size_t>(static_cast<size_t>(nnz) * vw * 8, 8));
    if (m->is_complex)
        hipLaunchKernelGGL(tbk::rescale_fill_kernel<true>, dim3(grid), dim3(tbk::kBlock), 0, st,
                           M->rp.as<const int32_t>(), M->ci.as<const int32_t>(), M->val64.as<const double>(), n, a, b,
                           m->rp.as<const int32_t>(), m->ci.as<int32_t>(), m->val64.as<double>());
    else
        hipLaunchKernelGGL(tbk::rescale_fill_kernel<false>, dim3(grid), dim3(tbk::kBlock), 0, st,
                           M->rp.as<const int32_t>(), M->ci.as<const int32_t>(), M->val64.as<const double>(), n, a, b,
                           m->rp.as<const int32_t>(), m->ci.as<int32_t>(), m->val64.as<double>());
    HIPC(hipGetLastError());
    m->bandwidth95 = distance_percentile95(m->rp.as<const int32_t>(), m->ci.as<const int32_t>(), n, nnz, st);
    HIPC(hipStreamSynchronize(st));
    ctx->held_bytes += static_cast<int64_t>(m->rp.bytes + m->ci.bytes + m->val64.bytes);
    *out = m.release();
}

void velocity_matrix(tbk_context* ctx, const tbk_matrix* M, const tbk_lattice_desc* d, int axis, hipStream_t st,
                     tbk_matrix** out) {
    if (d->dim < 1 || d->dim > 3) raise(TBK_E_INVALID, "dim must be 1, 2 or 3");
    if (d->n_atoms < 1) raise(TBK_E_INVALID, "need at least one atom");
    if (!d->atom_pos) raise(TBK_E_INVALID, "the description carries no atom positions");
    if (axis < 0 || axis >= d->dim) raise(TBK_E_INVALID, "axis out of range");
    for (int k = 0; k < d->dim; ++k)
        if (d->span[k] < 1) raise(TBK_E_INVALID, "span must be positive");
    if (lattice_rows(d) != M->rows) raise(TBK_E_INVALID, "the matrix rows are not the lattice's sites");
    const int32_t n = M->rows;
    tbk::Geometry G{};
    G.dim = d->dim;
    G.n_atoms = d->n_atoms;
    G.n = n;
    for (int k = 0; k < 3; ++k) {
        G.span[k] = k < d->dim ? d->span[k] : 1;
        G.base[k] = k < d->dim ? d->bases[3 * k + axis] : 0.0;
    }
    std::vector<double> atom_axis(static_cast<size_t>(d->n_atoms));
    for (int a = 0; a < d->n_atoms; ++a) atom_axis[static_cast<size_t>(a)] = d->atom_pos[3 * a + axis];
    DevBuf dapos, dx;
    dapos.reserve(atom_axis.size() * 8);
    dx.reserve(std::max<size_t>(static_cast<size_t>(n) * 8, 8));
    HIPC(hipMemcpyAsync(dapos.p, atom_axis.data(), atom_axis.size() * 8, hipMemcpyHostToDevice, st));
    const unsigned grid = tbk::row_grid(n);
    hipLaunchKernelGGL(tbk::position_kernel, dim3(grid), dim3(tbk::kBlock), 0, st, G, dapos.as<const double>(),
                       dx.as<double>());
    HIPC(hipGetLastError());

    auto m = std::make_unique<tbk_matrix>();
    m->ctx = ctx;
    m->rows = n;
    m->nnz = M->nnz;
    m->is_complex = M->is_complex;
    m->max_row_nnz = M->max_row_nnz;
    m->bandwidth95 = M->bandwidth95;
    m->pattern = M->pattern;  // the same rows, row pointers and columns
    const int vw = m->is_complex ? 2 : 1;
    m->rp.reserve_placed((static_cast<size_t>(n) + 1) * 4);
    m->ci.reserve_placed(std::max<size_t>(static_cast<size_t>(M->nnz) * 4, 4));
    m->val64.reserve_placed(std::max<size_t>(static_cast<size_t>(M->nnz) * vw * 8, 8));
    HIPC(hipMemcpyAsync(m->rp.p, M->rp.p, (static_cast<size_t>(n) + 1) * 4, hipMemcpyDeviceToDevice, st));
    if (M->nnz > 0)
        HIPC(hipMemcpyAsync(m->ci.p, M->ci.p, static_cast<size_t>(M->nnz) * 4, hipMemcpyDeviceToDevice, st));
    if (m->is_complex)
        hipLaunchKernelGGL(tbk::velocity_fill_kernel<true>, dim3(grid), dim3(tbk::kBlock), 0, st,
                           M->rp.as<const int32_t>(), M->ci.as<const int32_t>(), M->val64.as<const double>(), n,
                           dx.as<const double>(), m->val64.as<double>());
    else
        hipLaunchKernelGGL(tbk::velocity_fill_kernel<false>, dim3(grid), dim3(tbk::kBlock), 0, st,
                           M->rp.as<const int32_t>(), M->ci.as<const int32_t>(), M->val64.as<const double>(), n,
                           dx.as<const double>(), m->val64.as<double>());
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(st));  // dapos and dx are freed on return
    ctx->held_bytes += static_cast<int64_t>(m->rp.bytes + m->ci.bytes + m->val64.bytes);
    *out = m.release();
}

}  // namespace engine

using namespace engine;

extern "C" {

int tbk_matrix_rescale(tbk_context* ctx, const tbk_matrix* M, double a, double b, tbk_matrix** out) {
    return guarded([&] {
        if (!ctx || !M || !out) raise(TBK_E_INVALID, "null argument");
        if (M->ctx != ctx) raise(TBK_E_INVALID, "objects belong to another context");
        set_device(ctx);
        rescale_matrix(ctx, M, a, b, ctx->stream, out);
    });
}

int tbk_matrix_velocity(tbk_context* ctx, const tbk_matrix* M, const tbk_lattice_desc* d, int32_t axis,
                        tbk_matrix** out) {
    return guarded([&] {
        if (!ctx || !M || !d || !out) raise(TBK_E_INVALID, "null argument");
        if (M->ctx != ctx) raise(TBK_E_INVALID, "objects belong to another context");
        set_device(ctx);
        velocity_matrix(ctx, M, d, axis, ctx->stream, out);
    });
}

}  // extern "C"
