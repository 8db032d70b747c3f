// Small device kernels of the engine other than the recursion step:
// moment-partial reductions, random-phase block generation, block copy and
// fp64 -> fp32 value conversion.  The kernels have internal linkage: each
// translation unit that includes this header launches its own copy.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cheb_kernels.h"

namespace tbk {
namespace {

// out[s] = sum_b partial[s*nb + b], fixed order; one workgroup per step.
__global__ __launch_bounds__(kBlock) void reduce_partials_kernel(const double2* __restrict__ partial,
                                                                 int32_t nb,
                                                                 double2* __restrict__ out) {
    const int s = blockIdx.x;
    const double2* ps = partial + static_cast<size_t>(s) * nb;
    double a0 = 0.0, a1 = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) {
        a0 = a0 + ps[b].x;
        a1 = a1 + ps[b].y;
    }
    __shared__ double r0[kBlock], r1[kBlock];
    r0[threadIdx.x] = a0;
    r1[threadIdx.x] = a1;
    __syncthreads();
    for (int w = kBlock / 2; w >= 1; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            r0[threadIdx.x] = r0[threadIdx.x] + r0[threadIdx.x + w];
            r1[threadIdx.x] = r1[threadIdx.x] + r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = make_double2(r0[0], r1[0]);
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Columns [col0, col0 + R) of the random-phase block of width Rtot:
// e^{i 2 pi u(seed, site*Rtot + col0 + col)} / sqrt(rows), stored R wide in a
// block padded to Rp columns (padding zero).  col0 = 0, Rtot = R is the
// whole block; a narrower call gives exactly its columns of that block.
template <typename V2>
__global__ __launch_bounds__(kBlock) void random_phase_kernel(V2* __restrict__ out, int32_t rows,
                                                              int32_t R, int32_t Rp, uint64_t seed,
                                                              double inv_sqrt_n, int32_t Rtot, int32_t col0) {
    const int64_t total = static_cast<int64_t>(rows) * Rp;
    const uint64_t key = splitmix64(seed);
    for (int64_t idx = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; idx < total;
         idx += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int64_t i = idx / Rp;
        const int col = static_cast<int>(idx - i * Rp);
        V2 v;
        if (col < R) {
            const uint64_t h = splitmix64(key ^ static_cast<uint64_t>(i * Rtot + col0 + col));
            const double u = static_cast<double>(h >> 11) * 0x1.0p-53;
            double sn, cs;
            sincos(2.0 * 3.141592653589793 * u, &sn, &cs);
            v.x = static_cast<decltype(v.x)>(cs * inv_sqrt_n);
            v.y = static_cast<decltype(v.y)>(sn * inv_sqrt_n);
        } else {
            v.x = 0;
            v.y = 0;
        }
        out[idx] = v;
    }
}

// Columns [col0, col0 + cols) of a row-major rows x Rp complex128 block into a
// rows x W block (padding columns zero).
__global__ __launch_bounds__(kBlock) void column_group_kernel(const double2* __restrict__ in, int32_t rows,
                                                              int32_t Rp, int32_t col0, int32_t cols, int32_t W,
                                                              double2* __restrict__ out) {
    const int64_t total = static_cast<int64_t>(rows) * W;
    for (int64_t idx = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; idx < total;
         idx += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int64_t i = idx / W;
        const int c = static_cast<int>(idx - i * W);
        out[idx] = c < cols ? in[i * Rp + col0 + c] : make_double2(0.0, 0.0);
    }
}

// dst = src, 16-byte elements (a vector block copy), non-temporal: a block
// is far larger than the caches.  Each thread moves kCopyUnroll elements, all
// loads issued before the stores.
constexpr int kCopyUnroll = 8;
__global__ __launch_bounds__(kBlock) void copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                        int64_t n) {
    typedef uint32_t u4v __attribute__((ext_vector_type(4)));
    const u4v* s = reinterpret_cast<const u4v*>(src);
    u4v* d = reinterpret_cast<u4v*>(dst);
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kBlock * kCopyUnroll + threadIdx.x;
    u4v v[kCopyUnroll];
#pragma unroll
    for (int k = 0; k < kCopyUnroll; ++k) {
        const int64_t i = base + static_cast<int64_t>(k) * kBlock;
        if (i < n) v[k] = __builtin_nontemporal_load(s + i);
    }
#pragma unroll
    for (int k = 0; k < kCopyUnroll; ++k) {
        const int64_t i = base + static_cast<int64_t>(k) * kBlock;
        if (i < n) __builtin_nontemporal_store(v[k], d + i);
    }
}

// Bandwidth references with non-temporal accesses (grid-stride, 4 elements
// per thread in flight): a copy, the two-reads-one-write mix of a recursion
// step without its gathers (y = a + b), and a read-only reduction.
typedef double nt_d2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(kBlock) void stream_copy_kernel(const nt_d2* __restrict__ a, nt_d2* __restrict__ y,
                                                             int64_t n) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock * 4 + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kBlock * 4) {
        nt_d2 v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + k * kBlock < n) v[k] = __builtin_nontemporal_load(a + i + k * kBlock);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + k * kBlock < n) __builtin_nontemporal_store(v[k], y + i + k * kBlock);
    }
}
__global__ __launch_bounds__(kBlock) void stream_2r1w_kernel(const nt_d2* __restrict__ a, const nt_d2* __restrict__ b,
                                                             nt_d2* __restrict__ y, int64_t n) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock * 4 + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kBlock * 4) {
        nt_d2 u[4], w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + k * kBlock < n) {
                u[k] = __builtin_nontemporal_load(a + i + k * kBlock);
                w[k] = __builtin_nontemporal_load(b + i + k * kBlock);
            }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + k * kBlock < n) __builtin_nontemporal_store(u[k] + w[k], y + i + k * kBlock);
    }
}
__global__ __launch_bounds__(kBlock) void stream_read_kernel(const nt_d2* __restrict__ a, double* __restrict__ out,
                                                             int64_t n) {
    double s = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kBlock) {
        const nt_d2 v = __builtin_nontemporal_load(a + i);
        s = s + (v.x + v.y);
    }
    __shared__ double sh[kBlock];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w >= 1; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// Gershgorin discs of a Hermitian CSR matrix: per workgroup the min of
// Re h_ii - r_i and the max of Re h_ii + r_i, r_i = sum_{j != i} |h_ij| (in
// column order, |h| = sqrt(re^2 + im^2))
__global__ __launch_bounds__(kBlock) void gershgorin_kernel(const int32_t* __restrict__ rp,
                                                            const int32_t* __restrict__ ci,
                                                            const double* __restrict__ val, int cplx, int32_t n,
                                                            double2* __restrict__ out) {
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t ii = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; ii < n;
         ii += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int32_t i = static_cast<int32_t>(ii);
        double c = 0.0, r = 0.0;
        for (int32_t p = rp[i]; p < rp[i + 1]; ++p) {
            const double re = cplx ? val[2 * static_cast<int64_t>(p)] : val[p];
            const double im = cplx ? val[2 * static_cast<int64_t>(p) + 1] : 0.0;
            if (ci[p] == i)
                c = c + re;
            else
                r = r + (cplx ? sqrt(re * re + im * im) : fabs(re));
        }
        lo = fmin(lo, c - r);
        hi = fmax(hi, c + r);
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
    if (threadIdx.x == 0) out[blockIdx.x] = make_double2(slo[0], shi[0]);
}

// Position-keyed fingerprint of a CSR pattern: per workgroup the sum mod 2^64
// of mix(position, value) over the row pointers and the column indices
// (order-free, so the result does not depend on the launch shape)
__device__ __forceinline__ uint64_t fp_mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__global__ __launch_bounds__(kBlock) void pattern_hash_kernel(const int32_t* __restrict__ rp,
                                                              const int32_t* __restrict__ ci, int32_t n,
                                                              int64_t nnz, unsigned long long* __restrict__ out) {
    unsigned long long h = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i <= n;
         i += static_cast<int64_t>(gridDim.x) * kBlock)
        h += fp_mix((static_cast<uint64_t>(i) << 32) | static_cast<uint32_t>(rp[i]));
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; p < nnz;
         p += static_cast<int64_t>(gridDim.x) * kBlock)
        h += fp_mix(((static_cast<uint64_t>(p) << 32) | static_cast<uint32_t>(ci[p])) ^ 0xC01C01C01C01C01Cull);
    __shared__ unsigned long long sh[kBlock];
    sh[threadIdx.x] = h;
    __syncthreads();
    for (int s = kBlock / 2; s >= 1; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// fp64 -> fp32 copy of matrix values
__global__ __launch_bounds__(kBlock) void to_fp32_kernel(const double* __restrict__ in,
                                                         float* __restrict__ out, int64_t n) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kBlock)
        out[i] = static_cast<float>(in[i]);
}

// Moments of step k0 + s of every slot (one workgroup per slot): the
// slot's workgroup partials summed in a fixed order, then mu_0, mu_1 (step
// 0) or the doubling identities mu_2k = 2 s.x - mu_0, mu_2k+1 = 2 s.y - mu_1.
__global__ __launch_bounds__(kBlock) void slots_moments_kernel(const SlotDesc* __restrict__ slots,
                                                                    int32_t s, const double2* __restrict__ partial) {
    slot_step_moments(slots, static_cast<int32_t>(blockIdx.x), s, partial);
}

// Moments of the persistent batched recursion (one workgroup per slot):
// for each of the slot's steps k, its nshare item partials at
// part[pbase + k * nshare .. ] summed in share order, then mu_0, mu_1
// (k = 0) or the doubling identities mu_2k = 2 s.x - mu_0,
// mu_2k+1 = 2 s.y - mu_1.
__global__ __launch_bounds__(kBlock) void slots_history_moments_kernel(const SlotDesc* __restrict__ slots,
                                                                       const int32_t* __restrict__ nshare,
                                                                       const int64_t* __restrict__ pbase,
                                                                       const double2* __restrict__ part) {
    const SlotDesc& d = slots[blockIdx.x];
    const int32_t ns = nshare[blockIdx.x];
    const double2* p = part + pbase[blockIdx.x];
    double m0 = 0.0, m1 = 0.0;
    for (int32_t j = 0; j < ns; ++j) {
        m0 = m0 + p[j].x;
        m1 = m1 + p[j].y;
    }
    double* m = d.mu;
    for (int64_t k = threadIdx.x; k < d.steps; k += kBlock) {
        if (k == 0) {
            m[0] = m0;
            m[1] = m1;
            continue;
        }
        const double2* pk = p + k * ns;
        double a0 = 0.0, a1 = 0.0;
        for (int32_t j = 0; j < ns; ++j) {
            a0 = a0 + pk[j].x;
            a1 = a1 + pk[j].y;
        }
        if (2 * k < d.n) m[2 * k] = 2.0 * a0 - m0;
        if (2 * k + 1 < d.n) m[2 * k + 1] = 2.0 * a1 - m1;
    }
}

// Resets before a persistent launch (a kernel, not a memset, so that it can
// be captured): the done counts of n slots, the nq ticket counters, and the
// error flag when given.
__global__ __launch_bounds__(kBlock) void persist_reset_kernel(int32_t* __restrict__ done, int32_t n,
                                                               int32_t* __restrict__ counters, int32_t nq,
                                                               int32_t* __restrict__ error) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    for (int64_t i = gid; i < n; i += static_cast<int64_t>(gridDim.x) * kBlock) done[i] = 0;
    if (gid < nq) counters[gid] = 0;
    if (gid == 0 && error) *error = 0;  // the first launch of a run
}

// After a persistent launch: *error if a queue was not served to its end (a
// workgroup that draws a ticket works on it, so a served queue's counter
// reaches its item count).
__global__ void persist_check_kernel(const int32_t* __restrict__ counters, const int64_t* __restrict__ totals,
                                     int32_t nq, int32_t* __restrict__ error) {
    const int q = static_cast<int>(threadIdx.x);
    if (q < nq && counters[q] < totals[q]) *error = 1;
}

// out = c * t (init) or out + c * t, elementwise over n complex values, c
// complex; the product is formed as (c.re t.re - c.im t.im, c.re t.im +
// c.im t.re) in the block's precision.  With c.im = 0 each part is c.re times
// the part, the scalar-times-matrix product of a real coefficient.
template <typename V2>
__global__ __launch_bounds__(kBlock) void block_axpy_kernel(V2* __restrict__ out, const V2* __restrict__ t,
                                                            double cr, double ci, int64_t n, int32_t init) {
    using E = decltype(t->x);
    const E r = static_cast<E>(cr), im = static_cast<E>(ci);
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kBlock) {
        const V2 v = t[i];
        V2 p;
        if (ci == 0.0) {
            p.x = r * v.x;
            p.y = r * v.y;
        } else {
            p.x = r * v.x - im * v.y;
            p.y = r * v.y + im * v.x;
        }
        if (!init) {
            const V2 o = out[i];
            p.x = o.x + p.x;
            p.y = o.y + p.y;
        }
        out[i] = p;
    }
}

// out = [a | b]: two blocks of half columns per row side by side in one of
// 2 half columns (complex128 or rounded to complex64)
template <typename V2>
__global__ __launch_bounds__(kBlock) void pair_blocks_kernel(const double2* __restrict__ a,
                                                             const double2* __restrict__ b,
                                                             V2* __restrict__ out, int64_t rows, int32_t half) {
    const int64_t w = 2 * static_cast<int64_t>(half), n = rows * w;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int64_t r = i / w;
        const int64_t c = i - r * w;
        const double2 v = c < half ? a[r * half + c] : b[r * half + (c - half)];
        V2 o;
        o.x = static_cast<decltype(o.x)>(v.x);
        o.y = static_cast<decltype(o.y)>(v.y);
        out[i] = o;
    }
}

}  // namespace
}  // namespace tbk
