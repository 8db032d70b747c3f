// Hamiltonian assembly on the device, straight into the engine's CSR: the
// builders get_Hamiltonian_sparse (open boundaries, lattice.h:262-305),
// get_Hamiltonian_periodic_extend_bare_sparse (lattice.h:216-258) and
// get_Hamiltonian_bare_sparse (lattice.h:347-377) for lattices without
// removed sites.
//
// The host builders push, for every cell c (index order) and hopping h
// (table order) whose target lies inside (or wraps), the triplet
// (row, col, t) and, off the diagonal, (col, row, conj t); setFromTriplets
// sums duplicates in insertion order.  Row i of the result therefore holds
//   * the forward pushes of its own cell's hoppings out of its atom, and
//   * the transposed pushes of the hoppings into its atom from the one
//     source cell that maps onto it,
// each with the insertion key (c * n_hops + h) * 2 + {0, 1}.  One thread per
// row forms these candidates, sorts them by (column, key) and merges equal
// columns in key order: the same entries, in the same order, with the same
// sums as the host, without a global sort.  Two passes (count, scan, fill).
//
// Values: t from the table; times the host's own phase table
// exp(+i A(r_from).d) per (cell, hopping) when a gauge field is set (the
// complex product written out as the host's: re = a c - b d, im = a d + b c);
// plus the site-keyed disorder W (u(seed, site) - 1/2) on the diagonal pushes
// (the same splitmix64 counter, so the same bits).  Compiled without
// contraction, so every value is bit-identical to the host builder's.
//
// The matrix statistics the engine keeps (longest row, the 95th percentile
// of |row - col| by an exact three-pass radix select, the Gershgorin
// interval) are formed on the device as well.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "engine.h"
#include "tbkpm.h"

namespace tbk {
namespace {

constexpr int kMaxRowCand = 32;  // candidate pushes per row (forward + transposed)
constexpr int kSelBits = 11;     // radix-select digit
constexpr int kSelBins = 1 << kSelBits;

struct LatticeArgs {
    int32_t dim, n_atoms, n_hops, builder, is_complex, has_disorder;
    int32_t span[3], ext[3], span_s[3];
    int32_t n_rows;
    double W;
    uint64_t key;  // splitmix64(seed) of the disorder generator
    const int32_t* fwd_off;
    const int32_t* fwd;  // per atom: hoppings with from == atom, table order
    const int32_t* bwd_off;
    const int32_t* bwd;  // per atom: hoppings with to == atom, table order
    const int32_t* hop_from;
    const int32_t* hop_to;
    const int32_t* hop_lv;  // 3 per hopping
    const double* hop_t;    // (re, im) per hopping
    const double2* phase;   // (cell, hopping) phase factors, or null
};

struct Cand {
    int32_t col;
    int64_t key;
    double re, im;
};

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// t of hopping h pushed from cell c, with the gauge phase when fields apply
__device__ __forceinline__ void hop_value(const LatticeArgs& A, int64_t cell, int h, bool fields, double& re,
                                          double& im) {
    re = A.hop_t[2 * h];
    im = A.hop_t[2 * h + 1];
    if (fields && A.phase) {
        const double2 p = A.phase[cell * A.n_hops + h];
        const double r2 = re * p.x - im * p.y;
        const double i2 = re * p.y + im * p.x;
        re = r2;
        im = i2;
    }
}

// the pushes that land in row i, unsorted; returns their number
__device__ int row_candidates(const LatticeArgs& A, int32_t i, Cand* c) {
    const int32_t a = i % A.n_atoms;
    const int32_t cell = i / A.n_atoms;
    int32_t iv[3] = {0, 0, 0};
    {
        int32_t t = cell;
        for (int d = A.dim - 1; d >= 0; --d) {
            iv[d] = t % A.span[d];
            t /= A.span[d];
        }
    }
    const bool fields = A.builder != 2;
    int n = 0;
    // forward pushes (i, col, t) of this cell's hoppings out of atom a
    for (int32_t q = A.fwd_off[a]; q < A.fwd_off[a + 1]; ++q) {
        const int h = A.fwd[q];
        bool ok = true;
        int32_t col = A.hop_to[h];
        for (int d = 0; d < A.dim; ++d) {
            int32_t x = iv[d] + A.hop_lv[3 * h + d];
            const int32_t L = A.span[d];
            if (A.builder == 1) {
                if (!(A.ext[d] != 0 || (x >= 0 && x < L))) ok = false;
                int32_t y = x - L * (x / L);
                y = y + L;
                x = y - L * (y / L);
            } else if (x < 0 || x >= L) {
                ok = false;
            }
            col += x * A.span_s[d];
        }
        if (!ok) continue;
        double re, im;
        hop_value(A, cell, h, fields, re, im);
        if (fields && A.has_disorder && col == i) {
            const double u = static_cast<double>(mix64(A.key ^ static_cast<uint64_t>(i)) >> 11) * 0x1.0p-53;
            re = re + A.W * (u - 0.5);
        }
        c[n++] = Cand{col, (static_cast<int64_t>(cell) * A.n_hops + h) * 2, re, im};
    }
    // transposed pushes (i, r, conj t) of the hoppings into atom a, from the
    // one source cell whose push lands on this cell
    for (int32_t q = A.bwd_off[a]; q < A.bwd_off[a + 1]; ++q) {
        const int h = A.bwd[q];
        bool ok = true;
        int32_t scell = 0;
        for (int d = 0; d < A.dim; ++d) {
            const int32_t L = A.span[d];
            int32_t x = iv[d] - A.hop_lv[3 * h + d];
            if (A.builder == 1 && A.ext[d] != 0) {
                x = x % L;
                if (x < 0) x += L;
            } else if (x < 0 || x >= L) {
                ok = false;
            }
            scell = scell * L + x;
        }
        if (!ok) continue;
        const int32_t r = scell * A.n_atoms + A.hop_from[h];
        if (r == i) continue;  // that push was diagonal: no transposed copy
        double re, im;
        hop_value(A, scell, h, fields, re, im);
        c[n++] = Cand{r, (static_cast<int64_t>(scell) * A.n_hops + h) * 2 + 1, re, -im};
    }
    // insertion sort by (column, insertion key)
    for (int x = 1; x < n; ++x) {
        const Cand v = c[x];
        int y = x - 1;
        while (y >= 0 && (c[y].col > v.col || (c[y].col == v.col && c[y].key > v.key))) {
            c[y + 1] = c[y];
            --y;
        }
        c[y + 1] = v;
    }
    return n;
}

__global__ __launch_bounds__(kBlock) void assemble_count_kernel(LatticeArgs A, int32_t* __restrict__ cnt,
                                                                int32_t* __restrict__ max_row) {
    Cand c[kMaxRowCand];
    int32_t mx = 0;
    for (int64_t ii = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; ii < A.n_rows;
         ii += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int32_t i = static_cast<int32_t>(ii);
        const int n = row_candidates(A, i, c);
        int32_t k = 0;
        for (int x = 0; x < n; ++x)
            if (x == 0 || c[x].col != c[x - 1].col) ++k;
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

// writes row i's merged entries at rp[i]; per-workgroup Gershgorin partials
__global__ __launch_bounds__(kBlock) void assemble_fill_kernel(LatticeArgs A, const int32_t* __restrict__ rp,
                                                               int32_t* __restrict__ ci, double* __restrict__ val,
                                                               double2* __restrict__ gersh) {
    Cand c[kMaxRowCand];
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t ii = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; ii < A.n_rows;
         ii += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int32_t i = static_cast<int32_t>(ii);
        const int n = row_candidates(A, i, c);
        int32_t p = rp[i];
        double center = 0.0, radius = 0.0;
        int x = 0;
        while (x < n) {
            double re = c[x].re, im = c[x].im;
            int y = x + 1;
            for (; y < n && c[y].col == c[x].col; ++y) {  // duplicates: sum in insertion order
                re = re + c[y].re;
                im = im + c[y].im;
            }
            ci[p] = c[x].col;
            if (A.is_complex) {
                val[2 * static_cast<int64_t>(p)] = re;
                val[2 * static_cast<int64_t>(p) + 1] = im;
            } else {
                val[p] = re;
            }
            if (c[x].col == i)
                center += re;
            else
                radius += A.is_complex ? sqrt(re * re + im * im) : fabs(re);
            ++p;
            x = y;
        }
        lo = fmin(lo, center - radius);
        hi = fmax(hi, center + radius);
    }
    __shared__ double slo[kBlock], shi[kBlock];
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    for (int s = kBlock / 2; s >= 1; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) {
            slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + s]);
            shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) gersh[blockIdx.x] = make_double2(slo[0], shi[0]);
}

// histogram of the digit (d >> shift) & (bins - 1) over the entries whose
// d >> (shift + kSelBits) equals prefix, d = |row - col|
__global__ __launch_bounds__(kBlock) void distance_digit_kernel(const int32_t* __restrict__ rp,
                                                                const int32_t* __restrict__ ci, int32_t n,
                                                                int shift, uint64_t prefix,
                                                                uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kSelBins];
    for (int b = threadIdx.x; b < kSelBins; b += kBlock) h[b] = 0;
    __syncthreads();
    for (int64_t ii = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; ii < n;
         ii += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int32_t i = static_cast<int32_t>(ii);
        const int32_t p1 = rp[i + 1];
        for (int32_t p = rp[i]; p < p1; ++p) {
            const int64_t dd = static_cast<int64_t>(ci[p]) - i;
            const uint64_t d = static_cast<uint64_t>(dd < 0 ? -dd : dd);
            if ((d >> (shift + kSelBits)) == prefix) atomicAdd(&h[(d >> shift) & (kSelBins - 1)], 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kSelBins; b += kBlock)
        if (h[b]) atomicAdd(&hist[b], h[b]);
}

}  // namespace
}  // namespace tbk

namespace engine {

namespace {

uint64_t host_mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

}  // namespace

int64_t lattice_rows(const tbk_lattice_desc* d) {
    int64_t n = d->n_atoms;
    for (int k = 0; k < d->dim; ++k) n *= d->span[k];
    return n;
}

// upper bound of the stored entries: two pushes per (cell, hopping)
int64_t lattice_nnz_bound(const tbk_lattice_desc* d) {
    return (lattice_rows(d) / std::max(d->n_atoms, 1)) * 2 * static_cast<int64_t>(std::max(d->n_hops, 0));
}

// The exact 95th percentile of |row - col| over the entries of a resident
// CSR pattern (the entry nth_element would pick at rank 0.95 nnz): a
// three-pass radix select over 11-bit digits.
int32_t distance_percentile95(const int32_t* rp, const int32_t* ci, int32_t n, int64_t nnz, hipStream_t st) {
    if (nnz <= 0) return 0;
    const unsigned grid = static_cast<unsigned>(
        std::min<int64_t>((static_cast<int64_t>(n) + tbk::kBlock - 1) / tbk::kBlock, 8192));
    uint64_t rank = std::min<uint64_t>(static_cast<uint64_t>(nnz) - 1,
                                       static_cast<uint64_t>(0.95 * static_cast<double>(nnz)));
    uint64_t prefix = 0;
    DevBuf dhist;
    dhist.reserve(tbk::kSelBins * 4);
    std::vector<uint32_t> hh(tbk::kSelBins);
    for (int shift = 2 * tbk::kSelBits; shift >= 0; shift -= tbk::kSelBits) {
        HIPC(hipMemsetAsync(dhist.p, 0, tbk::kSelBins * 4, st));
        hipLaunchKernelGGL(tbk::distance_digit_kernel, dim3(grid), dim3(tbk::kBlock), 0, st, rp, ci, n, shift, prefix,
                           dhist.as<uint32_t>());
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(hh.data(), dhist.p, tbk::kSelBins * 4, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        int b = 0;
        for (; b < tbk::kSelBins; ++b) {
            if (rank < hh[b]) break;
            rank -= hh[b];
        }
        if (b == tbk::kSelBins) raise(TBK_E_DEVICE, "distance select lost its rank");
        prefix = (prefix << tbk::kSelBits) | static_cast<uint64_t>(b);
    }
    return static_cast<int32_t>(prefix);
}

void assemble_matrix(tbk_context* ctx, const tbk_lattice_desc* d, hipStream_t st, tbk_matrix** out, double* ms) {
    const auto t0 = std::chrono::steady_clock::now();
    if (d->dim < 1 || d->dim > 3) raise(TBK_E_INVALID, "dim must be 1, 2 or 3");
    if (d->n_atoms < 1 || d->n_hops < 0) raise(TBK_E_INVALID, "need at least one atom");
    if (d->builder < 0 || d->builder > 2) raise(TBK_E_INVALID, "unknown builder");
    if (d->n_hops > 0 && (!d->hop_from || !d->hop_to || !d->hop_lv || !d->hop_t))
        raise(TBK_E_INVALID, "null hopping table");
    for (int k = 0; k < d->dim; ++k)
        if (d->span[k] < 1) raise(TBK_E_INVALID, "span must be positive");
    const int64_t rows64 = lattice_rows(d);
    if (rows64 > INT32_MAX) raise(TBK_E_UNSUPPORTED, "more than 2^31 - 1 sites");
    const int32_t n = static_cast<int32_t>(rows64);
    const int64_t cells = rows64 / d->n_atoms;
    // per-atom candidate lists
    std::vector<int32_t> fwd_off(d->n_atoms + 1, 0), bwd_off(d->n_atoms + 1, 0), fwd, bwd;
    for (int h = 0; h < d->n_hops; ++h) {
        if (d->hop_from[h] < 0 || d->hop_from[h] >= d->n_atoms || d->hop_to[h] < 0 || d->hop_to[h] >= d->n_atoms)
            raise(TBK_E_INVALID, "hopping atom index out of range");
        fwd_off[d->hop_from[h] + 1]++;
        bwd_off[d->hop_to[h] + 1]++;
    }
    int most = 0;
    for (int a = 0; a < d->n_atoms; ++a) {
        most = std::max(most, fwd_off[a + 1] + bwd_off[a + 1]);
        fwd_off[a + 1] += fwd_off[a];
        bwd_off[a + 1] += bwd_off[a];
    }
    if (most > tbk::kMaxRowCand) raise(TBK_E_UNSUPPORTED, "more than 32 hoppings touch one atom");
    if (rows64 * std::max(most, 1) > INT32_MAX) raise(TBK_E_UNSUPPORTED, "too many entries for int32 row pointers");
    fwd.resize(fwd_off[d->n_atoms]);
    bwd.resize(bwd_off[d->n_atoms]);
    {
        std::vector<int32_t> fp(fwd_off.begin(), fwd_off.end() - 1), bp(bwd_off.begin(), bwd_off.end() - 1);
        for (int h = 0; h < d->n_hops; ++h) {
            fwd[fp[d->hop_from[h]]++] = h;
            bwd[bp[d->hop_to[h]]++] = h;
        }
    }
    std::vector<int32_t> lv3(static_cast<size_t>(3) * std::max(d->n_hops, 1), 0);
    for (int h = 0; h < d->n_hops; ++h)
        for (int k = 0; k < d->dim; ++k) lv3[3 * h + k] = d->hop_lv[h * d->dim + k];
    // small tables in one upload
    std::vector<int32_t> itab;
    auto put = [&](const std::vector<int32_t>& v) {
        const size_t at = itab.size();
        itab.insert(itab.end(), v.begin(), v.end());
        itab.push_back(0);
        return at;
    };
    const size_t o_fo = put(fwd_off), o_f = put(fwd), o_bo = put(bwd_off), o_b = put(bwd), o_lv = put(lv3);
    const size_t o_from = put(std::vector<int32_t>(d->hop_from, d->hop_from + d->n_hops));
    const size_t o_to = put(std::vector<int32_t>(d->hop_to, d->hop_to + d->n_hops));
    DevBuf dtab, dhop, dphase, dcnt, dmax, dtemp, dgersh;
    dtab.reserve(itab.size() * 4);
    HIPC(hipMemcpyAsync(dtab.p, itab.data(), itab.size() * 4, hipMemcpyHostToDevice, st));
    dhop.reserve(static_cast<size_t>(std::max(d->n_hops, 1)) * 16);
    if (d->n_hops > 0)
        HIPC(hipMemcpyAsync(dhop.p, d->hop_t, static_cast<size_t>(d->n_hops) * 16, hipMemcpyHostToDevice, st));
    const bool use_phase = d->phase && d->builder != 2;
    if (use_phase) {
        const size_t pb = static_cast<size_t>(cells) * d->n_hops * 16;
        dphase.reserve(std::max<size_t>(pb, 16));
        if (pb) HIPC(hipMemcpyAsync(dphase.p, d->phase, pb, hipMemcpyHostToDevice, st));
    }
    tbk::LatticeArgs A{};
    A.dim = d->dim;
    A.n_atoms = d->n_atoms;
    A.n_hops = d->n_hops;
    A.builder = d->builder;
    A.is_complex = d->is_complex ? 1 : 0;
    A.has_disorder = d->has_disorder ? 1 : 0;
    for (int k = 0; k < 3; ++k) {
        A.span[k] = k < d->dim ? d->span[k] : 1;
        A.ext[k] = k < d->dim ? d->extend[k] : 0;
    }
    A.span_s[d->dim - 1] = d->n_atoms;
    for (int k = d->dim - 2; k >= 0; --k) A.span_s[k] = A.span_s[k + 1] * A.span[k + 1];
    A.n_rows = n;
    A.W = d->disorder_W;
    A.key = host_mix64(d->disorder_seed);
    const int32_t* tb = dtab.as<const int32_t>();
    A.fwd_off = tb + o_fo;
    A.fwd = tb + o_f;
    A.bwd_off = tb + o_bo;
    A.bwd = tb + o_b;
    A.hop_lv = tb + o_lv;
    A.hop_from = tb + o_from;
    A.hop_to = tb + o_to;
    A.hop_t = dhop.as<const double>();
    A.phase = use_phase ? dphase.as<const double2>() : nullptr;

    auto m = std::make_unique<tbk_matrix>();
    m->ctx = ctx;
    m->rows = n;
    m->is_complex = d->is_complex != 0;
    m->rp.reserve_placed((static_cast<size_t>(n) + 1) * 4);
    dcnt.reserve((static_cast<size_t>(n) + 1) * 4);
    dmax.reserve(4);
    HIPC(hipMemsetAsync(dmax.p, 0, 4, st));
    HIPC(hipMemsetAsync(dcnt.as<int32_t>() + n, 0, 4, st));
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>((rows64 + tbk::kBlock - 1) / tbk::kBlock, 8192));
    hipLaunchKernelGGL(tbk::assemble_count_kernel, dim3(grid), dim3(tbk::kBlock), 0, st, A, dcnt.as<int32_t>(),
                       dmax.as<int32_t>());
    HIPC(hipGetLastError());
    size_t temp = 0;
    HIPC(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, dcnt.as<int32_t>(), m->rp.as<int32_t>(), n + 1, st));
    dtemp.reserve(std::max<size_t>(temp, 16));
    HIPC(hipcub::DeviceScan::ExclusiveSum(dtemp.p, temp, dcnt.as<int32_t>(), m->rp.as<int32_t>(), n + 1, st));
    int32_t nnz = 0, max_row = 0;
    HIPC(hipMemcpyAsync(&nnz, m->rp.as<int32_t>() + n, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(&max_row, dmax.p, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    m->nnz = nnz;
    m->max_row_nnz = max_row;
    const int vw = m->is_complex ? 2 : 1;
    m->ci.reserve_placed(std::max<size_t>(static_cast<size_t>(nnz) * 4, 4));
    m->val64.reserve_placed(std::max<size_t>(static_cast<size_t>(nnz) * vw * 8, 8));
    dgersh.reserve(static_cast<size_t>(grid) * 16);
    hipLaunchKernelGGL(tbk::assemble_fill_kernel, dim3(grid), dim3(tbk::kBlock), 0, st, A, m->rp.as<const int32_t>(),
                       m->ci.as<int32_t>(), m->val64.as<double>(), dgersh.as<double2>());
    HIPC(hipGetLastError());
    std::vector<double2> gp(grid);
    HIPC(hipMemcpyAsync(gp.data(), dgersh.p, static_cast<size_t>(grid) * 16, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    m->bandwidth95 = distance_percentile95(m->rp.as<const int32_t>(), m->ci.as<const int32_t>(), n, nnz, st);
    double lo = INFINITY, hi = -INFINITY;
    for (const double2& g : gp) {
        lo = std::min(lo, g.x);
        hi = std::max(hi, g.y);
    }
    m->gersh_min = std::isfinite(lo) ? lo : 0.0;
    m->gersh_max = std::isfinite(hi) ? hi : 0.0;
    m->have_gersh = true;
    ctx->held_bytes += static_cast<int64_t>(m->rp.bytes + m->ci.bytes + m->val64.bytes);
    if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = m.release();
}

}  // namespace engine

using namespace engine;

extern "C" {

int tbk_matrix_assemble(tbk_context* ctx, const tbk_lattice_desc* d, tbk_matrix** out, double* wall_ms) {
    return guarded([&] {
        if (!ctx || !d || !out) raise(TBK_E_INVALID, "null argument");
        set_device(ctx);
        assemble_matrix(ctx, d, ctx->stream, out, wall_ms);
    });
}

int tbk_matrix_download(const tbk_matrix* m, int32_t* rowptr, int32_t* colidx, double* values) {
    return guarded([&] {
        if (!m || !rowptr || (m->nnz > 0 && (!colidx || !values))) raise(TBK_E_INVALID, "null argument");
        set_device(m->ctx);
        HIPC(hipStreamSynchronize(m->ctx->stream));
        HIPC(hipMemcpy(rowptr, m->rp.p, (static_cast<size_t>(m->rows) + 1) * 4, hipMemcpyDeviceToHost));
        if (m->nnz > 0) {
            HIPC(hipMemcpy(colidx, m->ci.p, static_cast<size_t>(m->nnz) * 4, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(values, m->val64.p, static_cast<size_t>(m->nnz) * (m->is_complex ? 16 : 8),
                           hipMemcpyDeviceToHost));
        }
    });
}

}  // extern "C"
