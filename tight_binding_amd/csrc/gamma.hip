// Gamma[n][m] = sum_k conj(A[n][k]) B[m][k] on gfx950 matrix cores.
//
// Each workgroup (4 waves) owns a 64 x 64 tile of Gamma for one chunk of k;
// each wave a 32 x 32 sub-tile = 2 x 2 blocks of 16 x 16.  Complex products
// are four real ones: Re = Ar.Br + Ai.Bi, Im = Ar.Bi - Ai.Br.
//
// fp64 path, v_mfma_f64_16x16x4f64 (A 16x4, B 4x16, D 16x16 in fp64): lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; D[4*t+l/16][l%16] is item t of
// lane l (rows interleaved across lane groups, unlike the fp32-accumulating
// forms).  One k-step covers 4 complex columns per 16-row block.
//
// split-BF16 path, v_mfma_f32_16x16x32_bf16 (A 16x32, B 32x16 in bf16, D in
// fp32): lane l supplies A[l%16][8*(l/16) .. +7] and B[8*(l/16) .. +7][l%16];
// D[4*(l/16)+t][l%16] is item t of lane l.
// The 32 real k of one instruction are 16 complex columns interleaved
// (re, im); against B the same pairs give Re, against (Bi, -Br) they give
// Im.  x = hi + lo with hi = bf16(x), lo = bf16(x - hi); the products
// hi*hi + hi*lo + lo*hi carry ~16 significant bits of each operand.
//
// Partial sums of each k chunk go to P[chunk]; launch_gamma_reduce adds the
// chunks in order, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include "gamma_launch.h"

namespace tbk {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

namespace {

__device__ __forceinline__ double2 ld_or_zero(const double2* p, bool ok) {
    return ok ? *p : make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(256) void gamma_f64_kernel(const double2* __restrict__ A, const double2* __restrict__ B,
                                                        int N, int64_t K, int64_t kc, double2* __restrict__ P) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int n0 = blockIdx.x * kGammaTile + (wave >> 1) * 32;
    const int m0 = blockIdx.y * kGammaTile + (wave & 1) * 32;
    const int64_t kb = static_cast<int64_t>(blockIdx.z) * kc;
    const int64_t ke = (kb + kc < K) ? kb + kc : K;

    const double2* a[2];
    const double2* b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int rn = n0 + 16 * i + li, rm = m0 + 16 * i + li;
        a[i] = A + static_cast<int64_t>(rn < N ? rn : N - 1) * K;
        b[i] = B + static_cast<int64_t>(rm < N ? rm : N - 1) * K;
    }
    v4d re[2][2], im[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            re[i][j] = v4d{0, 0, 0, 0};
            im[i][j] = v4d{0, 0, 0, 0};
        }

    double2 av[2], bv[2];
    {
        const int64_t k = kb + lk;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            av[i] = ld_or_zero(a[i] + k, k < ke);
            bv[i] = ld_or_zero(b[i] + k, k < ke);
        }
    }
    for (int64_t k0 = kb; k0 < ke; k0 += 4) {
        double2 an[2], bn[2];
        const int64_t k = k0 + 4 + lk;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            an[i] = ld_or_zero(a[i] + k, k < ke);
            bn[i] = ld_or_zero(b[i] + k, k < ke);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].x, bv[j].x, re[i][j], 0, 0, 0);
                im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].x, bv[j].y, im[i][j], 0, 0, 0);
            }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].y, bv[j].y, re[i][j], 0, 0, 0);
                im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[i].y, bv[j].x, im[i][j], 0, 0, 0);
            }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            av[i] = an[i];
            bv[i] = bn[i];
        }
    }

    double2* Pc = P + static_cast<int64_t>(blockIdx.z) * N * N;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = n0 + 16 * i + 4 * t + lk, col = m0 + 16 * j + li;
                if (row < N && col < N) Pc[static_cast<int64_t>(row) * N + col] = make_double2(re[i][j][t], im[i][j][t]);
            }
}

struct Split {
    v8bf hi, lo;
};

// 4 complex values (8 reals, interleaved re/im) -> hi/lo bf16 vectors.
// neg_swap: (re, im) -> (im, -re), the operand that yields the imaginary part.
__device__ __forceinline__ Split split4(const double2* v, bool neg_swap) {
    Split s;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double x0 = v[c].x, x1 = v[c].y;
        if (neg_swap) {
            const double t = x0;
            x0 = x1;
            x1 = -t;
        }
        const __bf16 h0 = static_cast<__bf16>(static_cast<float>(x0));
        const __bf16 h1 = static_cast<__bf16>(static_cast<float>(x1));
        s.hi[2 * c] = h0;
        s.hi[2 * c + 1] = h1;
        s.lo[2 * c] = static_cast<__bf16>(static_cast<float>(x0 - static_cast<double>(static_cast<float>(h0))));
        s.lo[2 * c + 1] = static_cast<__bf16>(static_cast<float>(x1 - static_cast<double>(static_cast<float>(h1))));
    }
    return s;
}

__global__ __launch_bounds__(256) void gamma_bf16x3_kernel(const double2* __restrict__ A,
                                                           const double2* __restrict__ B, int N, int64_t K,
                                                           int64_t kc, double2* __restrict__ P) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int n0 = blockIdx.x * kGammaTile + (wave >> 1) * 32;
    const int m0 = blockIdx.y * kGammaTile + (wave & 1) * 32;
    const int64_t kb = static_cast<int64_t>(blockIdx.z) * kc;
    const int64_t ke = (kb + kc < K) ? kb + kc : K;

    const double2* a[2];
    const double2* b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int rn = n0 + 16 * i + li, rm = m0 + 16 * i + li;
        a[i] = A + static_cast<int64_t>(rn < N ? rn : N - 1) * K;
        b[i] = B + static_cast<int64_t>(rm < N ? rm : N - 1) * K;
    }
    v4f re[2][2], im[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            re[i][j] = v4f{0, 0, 0, 0};
            im[i][j] = v4f{0, 0, 0, 0};
        }

    // one instruction covers 16 complex k; lane group lk owns k0 + 4*lk .. +3
    for (int64_t k0 = kb; k0 < ke; k0 += 16) {
        double2 av[2][4], bv[2][4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int64_t k = k0 + 4 * lk + c;
                av[i][c] = ld_or_zero(a[i] + k, k < ke);
                bv[i][c] = ld_or_zero(b[i] + k, k < ke);
            }
        Split sa[2], sb[2], sbi[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            sa[i] = split4(av[i], false);
            sb[i] = split4(bv[i], false);
            sbi[i] = split4(bv[i], true);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                re[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sa[i].hi, sb[j].hi, re[i][j], 0, 0, 0);
                im[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sa[i].hi, sbi[j].hi, im[i][j], 0, 0, 0);
                re[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sa[i].hi, sb[j].lo, re[i][j], 0, 0, 0);
                im[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sa[i].hi, sbi[j].lo, im[i][j], 0, 0, 0);
                re[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sa[i].lo, sb[j].hi, re[i][j], 0, 0, 0);
                im[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sa[i].lo, sbi[j].hi, im[i][j], 0, 0, 0);
            }
    }

    double2* Pc = P + static_cast<int64_t>(blockIdx.z) * N * N;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = n0 + 16 * i + 4 * lk + t, col = m0 + 16 * j + li;
                if (row < N && col < N)
                    Pc[static_cast<int64_t>(row) * N + col] =
                        make_double2(static_cast<double>(re[i][j][t]), static_cast<double>(im[i][j][t]));
            }
}

__global__ __launch_bounds__(256) void gamma_reduce_kernel(const double2* __restrict__ P, int nchunks, int64_t NN,
                                                           double2* __restrict__ out) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < NN;
         i += static_cast<int64_t>(gridDim.x) * 256) {
        double2 s = P[i];
        for (int c = 1; c < nchunks; ++c) {
            const double2 v = P[static_cast<int64_t>(c) * NN + i];
            s.x = s.x + v.x;
            s.y = s.y + v.y;
        }
        out[i] = s;
    }
}

dim3 gamma_grid(int N, int nchunks) {
    const unsigned t = static_cast<unsigned>((N + kGammaTile - 1) / kGammaTile);
    return dim3(t, t, static_cast<unsigned>(nchunks));
}

}  // namespace

void launch_gamma_f64(const double2* A, const double2* B, int N, int64_t K, int64_t kc, int nchunks, double2* P,
                      hipStream_t st) {
    hipLaunchKernelGGL(gamma_f64_kernel, gamma_grid(N, nchunks), dim3(256), 0, st, A, B, N, K, kc, P);
}

void launch_gamma_bf16x3(const double2* A, const double2* B, int N, int64_t K, int64_t kc, int nchunks, double2* P,
                         hipStream_t st) {
    hipLaunchKernelGGL(gamma_bf16x3_kernel, gamma_grid(N, nchunks), dim3(256), 0, st, A, B, N, K, kc, P);
}

void launch_gamma_reduce(const double2* P, int nchunks, int64_t NN, double2* out, hipStream_t st) {
    const int64_t blocks = (NN + 255) / 256;
    hipLaunchKernelGGL(gamma_reduce_kernel, dim3(static_cast<unsigned>(blocks < 65536 ? blocks : 65536)), dim3(256), 0,
                       st, P, nchunks, NN, out);
}

}  // namespace tbk
