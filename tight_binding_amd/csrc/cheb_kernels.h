// Device kernels of the Chebyshev moment recursion (gfx950, wave64).
//
// One step applies H~ = (H - bI)/a to a whole block of R vectors stored
// row-major [site][Rp] and fuses, per site row r:
//     acc  = sum_{p in row r, ascending column} H[p] * x[col[p]]
//     y_r  = alpha*acc + beta*x_r - z_r          (three-term update)
//     partial moments from y_r, x_r and (optionally) a bra row
// A row is handled by LPR lanes (each lane owns CPL consecutive columns), so
// the gather of a neighbour row is one coalesced LPR*CPL*16-byte access.
// Workgroups take tiles of consecutive rows; with xcd_split the tiles are
// dealt so that each XCD (blocks b with equal b % 8) walks one contiguous
// eighth of the rows, keeping a lattice neighbourhood inside one L2.
//
// Rounding: this file is compiled without floating-point contraction, so
// each product is rounded and then added -- the operation order of the
// CSC scatter-axpy in numerical.h:368-374 (Eigen sparse x dense, columns
// ascending).  With alpha = 2, beta = 0 (a = 1, b = 0) the fp64 vectors are
// bit-identical to that CPU recursion.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tbk {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

enum StepMode : int {
    kFirst = 0,   // T1 = H~ T0;           partials (Re<bra|T0>, Re<bra|T1>)
    kRec = 1,     // T_{k+1} = 2H~T_k - T_{k-1}; partial Re<bra|T_{k+1}>
    kRecDbl = 2,  // same update;          partials (Re<T_k|T_k>, Re<T_{k+1}|T_k>)
};

template <typename VT>
struct Vec2;
template <>
struct Vec2<double> {
    using type = double2;
};
template <>
struct Vec2<float> {
    using type = float2;
};

template <typename VT, bool CH>
struct HVal {
    using type = VT;
};
template <typename VT>
struct HVal<VT, true> {
    using type = typename Vec2<VT>::type;
};

__device__ __forceinline__ void madd(double2& acc, double h, double2 x) {
    acc.x = acc.x + h * x.x;
    acc.y = acc.y + h * x.y;
}
__device__ __forceinline__ void madd(float2& acc, float h, float2 x) {
    acc.x = acc.x + h * x.x;
    acc.y = acc.y + h * x.y;
}
// (h.re x.re - h.im x.im, h.re x.im + h.im x.re), then added
__device__ __forceinline__ void madd(double2& acc, double2 h, double2 x) {
    double re = h.x * x.x - h.y * x.y;
    double im = h.x * x.y + h.y * x.x;
    acc.x = acc.x + re;
    acc.y = acc.y + im;
}
__device__ __forceinline__ void madd(float2& acc, float2 h, float2 x) {
    float re = h.x * x.x - h.y * x.y;
    float im = h.x * x.y + h.y * x.x;
    acc.x = acc.x + re;
    acc.y = acc.y + im;
}
// Re(conj(a) b) accumulated in fp64
__device__ __forceinline__ double redot(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ double redot(float2 a, float2 b) {
    return static_cast<double>(a.x) * static_cast<double>(b.x) +
           static_cast<double>(a.y) * static_cast<double>(b.y);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, kWave);
    return v;
}

// Tile range of this workgroup (see file comment).
__device__ __forceinline__ void tile_range(int64_t ntiles, int xcd_split, int64_t& first,
                                           int64_t& end, int64_t& stride) {
    const int64_t nb = gridDim.x;
    const int64_t b = blockIdx.x;
    if (xcd_split) {
        const int64_t xcd = b & 7, j = b >> 3, per = nb >> 3;
        first = (ntiles * xcd) / 8 + j;
        end = (ntiles * (xcd + 1)) / 8;
        stride = per;
    } else {
        first = b;
        end = ntiles;
        stride = nb;
    }
}

template <typename VT, bool CH, int LPR, int CPL, int MODE>
__global__ __launch_bounds__(kBlock) void cheb_step_kernel(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
    const typename HVal<VT, CH>::type* __restrict__ hv, const typename Vec2<VT>::type* __restrict__ x,
    const typename Vec2<VT>::type* z, typename Vec2<VT>::type* y,
    const typename Vec2<VT>::type* __restrict__ bra, int32_t rows, int32_t Rp, VT alpha, VT beta,
    int32_t shift, int64_t ntiles, int32_t xcd_split, double2* __restrict__ partial) {
    using V2 = typename Vec2<VT>::type;
    using H = typename HVal<VT, CH>::type;
    constexpr int RPW = kWave / LPR;
    constexpr int TR = kWavesPerBlock * RPW;
    static_assert(LPR >= 1 && LPR <= kWave && (kWave % LPR) == 0, "bad lanes per row");

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int sub = lane / LPR;
    const int q = lane % LPR;

    int64_t first, end, stride;
    tile_range(ntiles, xcd_split, first, end, stride);

    double p0 = 0.0, p1 = 0.0;
    for (int64_t t = first; t < end; t += stride) {
        int r = static_cast<int>(t * TR) + wave * RPW + sub;
        if (LPR == kWave) r = __builtin_amdgcn_readfirstlane(r);
        if (r >= rows) continue;
        const size_t own = static_cast<size_t>(r) * Rp + static_cast<size_t>(q) * CPL;

        V2 zv[CPL], bv[CPL], acc[CPL], xs[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            acc[c].x = 0;
            acc[c].y = 0;
            xs[c].x = 0;
            xs[c].y = 0;
        }
        if (MODE != kFirst) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) zv[c] = z[own + c];
        }
        if (MODE != kRecDbl) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) bv[c] = bra[own + c];
        }
        const int s = rp[r], e = rp[r + 1];
        bool have_self = false;
        int p = s;
        for (; p + 4 <= e; p += 4) {
            int cc[4];
            H hh[4];
            V2 xv[4][CPL];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                cc[k] = ci[p + k];
                hh[k] = hv[p + k];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const V2* xr = x + static_cast<size_t>(cc[k]) * Rp + static_cast<size_t>(q) * CPL;
#pragma unroll
                for (int c = 0; c < CPL; ++c) xv[k][c] = xr[c];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
#pragma unroll
                for (int c = 0; c < CPL; ++c) madd(acc[c], hh[k], xv[k][c]);
                if (cc[k] == r) {
                    have_self = true;
#pragma unroll
                    for (int c = 0; c < CPL; ++c) xs[c] = xv[k][c];
                }
            }
        }
        for (; p < e; ++p) {
            const int cc = ci[p];
            const H hh = hv[p];
            const V2* xr = x + static_cast<size_t>(cc) * Rp + static_cast<size_t>(q) * CPL;
            V2 xv[CPL];
#pragma unroll
            for (int c = 0; c < CPL; ++c) xv[c] = xr[c];
#pragma unroll
            for (int c = 0; c < CPL; ++c) madd(acc[c], hh, xv[c]);
            if (cc == r) {
                have_self = true;
#pragma unroll
                for (int c = 0; c < CPL; ++c) xs[c] = xv[c];
            }
        }
        if (!have_self) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) xs[c] = x[own + c];
        }

#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            V2 v;
            v.x = alpha * acc[c].x;
            v.y = alpha * acc[c].y;
            if (shift) {
                v.x = v.x + beta * xs[c].x;
                v.y = v.y + beta * xs[c].y;
            }
            if (MODE != kFirst) {
                v.x = v.x - zv[c].x;
                v.y = v.y - zv[c].y;
            }
            y[own + c] = v;
            if (MODE == kFirst) {
                p0 = p0 + redot(bv[c], xs[c]);
                p1 = p1 + redot(bv[c], v);
            } else if (MODE == kRec) {
                p0 = p0 + redot(bv[c], v);
            } else {
                p0 = p0 + redot(xs[c], xs[c]);
                p1 = p1 + redot(v, xs[c]);
            }
        }
    }

    // deterministic block reduction: butterfly in the wave, then waves in order
    __shared__ double red[2][kWavesPerBlock];
    p0 = wave_sum(p0);
    p1 = wave_sum(p1);
    if (lane == 0) {
        red[0][wave] = p0;
        red[1][wave] = p1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = red[0][0], s1 = red[1][0];
        for (int w = 1; w < kWavesPerBlock; ++w) {
            s0 = s0 + red[0][w];
            s1 = s1 + red[1][w];
        }
        partial[blockIdx.x] = make_double2(s0, s1);
    }
}

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

// e^{i 2 pi u(seed, site*R + col)} / sqrt(rows); padding columns zero
template <typename V2>
__global__ __launch_bounds__(kBlock) void random_phase_kernel(V2* __restrict__ out, int32_t rows,
                                                              int32_t R, int32_t Rp, uint64_t seed,
                                                              double inv_sqrt_n) {
    const int64_t total = static_cast<int64_t>(rows) * Rp;
    const uint64_t key = splitmix64(seed);
    for (int64_t idx = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; idx < total;
         idx += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int64_t i = idx / Rp;
        const int col = static_cast<int>(idx - i * Rp);
        V2 v;
        if (col < R) {
            const uint64_t h = splitmix64(key ^ static_cast<uint64_t>(i * R + col));
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

// fp64 -> fp32 copy of matrix values
__global__ __launch_bounds__(kBlock) void to_fp32_kernel(const double* __restrict__ in,
                                                         float* __restrict__ out, int64_t n) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kBlock)
        out[i] = static_cast<float>(in[i]);
}

}  // namespace tbk
