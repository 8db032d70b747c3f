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
#include "split_bf16.h"

namespace tbk {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

namespace {

constexpr int kWavesPerGroup = 4;

__device__ __forceinline__ double2 ld_or_zero(const double2* p, bool ok) {
    return ok ? *p : make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(256) void gamma_f64_kernel(const double2* __restrict__ A, const double2* __restrict__ B,
                                                        int N, int64_t K, int64_t ldk, int64_t kc,
                                                        double2* __restrict__ P) {
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
        a[i] = A + static_cast<int64_t>(rn < N ? rn : N - 1) * ldk;
        b[i] = B + static_cast<int64_t>(rm < N ? rm : N - 1) * ldk;
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

// ---------------------------------------------------------------------------
// Large-N fp64 contraction: a 128 x 128 Gamma tile per workgroup.  Tiles of A
// and B (128 rows x 8 complex columns, two MFMA k-steps) are staged through
// LDS, double-buffered, rows padded to 9 double2 to spread the fragment reads
// over the banks; the next tile's global loads are in flight while the matrix
// cores work on the current one.  A wave's staging load covers 8 rows x 8
// columns: every stack row is its own page, and the address translations a
// load needs (one per row it touches) bound the kernel at C3 sizes, so wider
// row pieces per load mean fewer translations per byte.
//
// Work item w = tile + tiles * chunk.  Workgroup b runs on XCD b % 8; the
// items are dealt so that XCD x takes w in [x * per_xcd, (x + 1) * per_xcd):
// all tiles of one k chunk then share that XCD's L2, and every A and B row of
// the chunk is fetched from HBM about once instead of once per tile.
__device__ __forceinline__ bool xcd_work_item(const GammaMap& m, int& tn, int& tm, int64_t& c) {
    const int x = static_cast<int>(blockIdx.x & 7u);
    const int64_t q = static_cast<int64_t>(blockIdx.x >> 3);
    if (q >= m.per_xcd) return false;
    const int64_t w = static_cast<int64_t>(x) * m.per_xcd + q;
    const int64_t tiles = static_cast<int64_t>(m.tiles_n) * m.tiles_n;
    if (w >= tiles * m.chunks) return false;
    c = w / tiles;
    const int t = static_cast<int>(w - c * tiles);
    tn = t / m.tiles_n;
    tm = t - tn * m.tiles_n;
    return true;
}

constexpr int kBigKt = 8;               // complex k per LDS stage (two MFMA k-steps)
constexpr int kBigStride = kBigKt + 1;  // padded LDS row stride (double2)

__global__ __launch_bounds__(512) void gamma_f64_big_kernel(const double2* __restrict__ A,
                                                            const double2* __restrict__ B, int N, int64_t K,
                                                            int64_t ldk, int64_t kc, GammaMap map,
                                                            double2* __restrict__ P) {
    __shared__ double2 As[2][kBigTile * kBigStride];
    __shared__ double2 Bs[2][kBigTile * kBigStride];
    int tn, tm;
    int64_t c;
    if (!xcd_work_item(map, tn, tm, c)) return;  // uniform over the workgroup
    // 8 waves: wave w owns the 64 x 32 block (rows (w >> 2) * 64, columns
    // (w & 3) * 32) of the 128 x 128 tile, as 4 x 2 fragments of 16 x 16
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int wn = wave >> 2, wm = wave & 3;
    const int n0 = tn * kBigTile, m0 = tm * kBigTile;
    const int64_t kb = c * kc;
    const int64_t ke = (kb + kc < K) ? kb + kc : K;

    // staging: thread t loads column t % 8 of rows t / 8 and t / 8 + 64 of
    // each operand
    const int cc = tid & 7, r = tid >> 3;
    const double2* ga[2];
    const double2* gb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int rn = (n0 + r + 64 * i < N) ? n0 + r + 64 * i : N - 1;
        const int rm = (m0 + r + 64 * i < N) ? m0 + r + 64 * i : N - 1;
        ga[i] = A + static_cast<int64_t>(rn) * ldk + cc;
        gb[i] = B + static_cast<int64_t>(rm) * ldk + cc;
    }
    double2 ra[2], rb[2];
    auto gload = [&](int64_t k0) {
        if (k0 + kBigKt <= ke) {  // a full stage: plain loads
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                ra[i] = ga[i][k0];
                rb[i] = gb[i][k0];
            }
        } else {
            const bool ok = k0 + cc < ke;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                ra[i] = ld_or_zero(ga[i] + k0, ok);
                rb[i] = ld_or_zero(gb[i] + k0, ok);
            }
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            As[buf][(r + 64 * i) * kBigStride + cc] = ra[i];
            Bs[buf][(r + 64 * i) * kBigStride + cc] = rb[i];
        }
    };

    // Three real products per complex one (Gauss):
    //   t1 = Ar.Br, t2 = Ai.Bi, t3 = (Ar - Ai).(Br + Bi),
    //   Re = t1 + t2, Im = t3 - t1 + t2 = Ar.Bi - Ai.Br,
    // 24 MFMAs per stage and wave where the four-product form needs 32.
    v4d t1[4][2], t2[4][2], t3[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            t1[i][j] = v4d{0, 0, 0, 0};
            t2[i][j] = v4d{0, 0, 0, 0};
            t3[i][j] = v4d{0, 0, 0, 0};
        }

    gload(kb);
    sstore(0);
    __syncthreads();
    int buf = 0;
    for (int64_t k0 = kb; k0 < ke; k0 += kBigKt) {
        const bool more = k0 + kBigKt < ke;
        if (more) gload(k0 + kBigKt);
#pragma unroll
        for (int ks = 0; ks < kBigKt / 4; ++ks) {  // the stage's MFMA k-steps of 4
            double2 a[4], b[2];
            double bs[2];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[buf][(wn * 64 + 16 * i + li) * kBigStride + 4 * ks + lk];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                b[j] = Bs[buf][(wm * 32 + 16 * j + li) * kBigStride + 4 * ks + lk];
                bs[j] = b[j].x + b[j].y;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double ad = a[i].x - a[i].y;
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    t1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].x, t1[i][j], 0, 0, 0);
                    t2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, b[j].y, t2[i][j], 0, 0, 0);
                    t3[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ad, bs[j], t3[i][j], 0, 0, 0);
                }
            }
        }
        if (more) sstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }

    double2* Pc = P + c * static_cast<int64_t>(N) * N;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = n0 + wn * 64 + 16 * i + 4 * t + lk;
                const int col = m0 + wm * 32 + 16 * j + li;
                const double p1 = t1[i][j][t], p2 = t2[i][j][t];
                if (row < N && col < N)
                    Pc[static_cast<int64_t>(row) * N + col] = make_double2(p1 + p2, (t3[i][j][t] - p1) + p2);
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
                                                           int64_t ldk, int64_t kc, double2* __restrict__ P) {
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
        a[i] = A + static_cast<int64_t>(rn < N ? rn : N - 1) * ldk;
        b[i] = B + static_cast<int64_t>(rm < N ? rm : N - 1) * ldk;
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

// ---------------------------------------------------------------------------
// Large-N split-bf16 x3 contraction, same 128 x 128 tiling and XCD placement
// as gamma_f64_big_kernel.  Each double is split once, when its tile is
// staged (hi = bf16(x), lo = bf16(x - hi)); pre-split stacks arrive as those
// words already.  In LDS the real and imaginary parts are kept apart: an LDS
// word holds one part of two consecutive complex columns, and a row of a
// 16-column stage is, per hi / lo part, [re 8 words | im 8 words] (virtual
// k: 16 real parts, then 16 imaginary parts), plus 4 words of padding (rows
// of 36 words: the 16 rows of a fragment read fall on distinct banks).  Lane
// group g of a fragment reads words 4g .. 4g + 3.  B read that way gives
// (re, im): A.B = sum ar br + ai bi = Re conj(a) b.  Groups 0, 1 reading the
// im words and groups 2, 3 the re words with the sign bits flipped give
// (im, -re): A.B' = sum ar bi - ai br = Im conj(a) b.  Chunks are short
// (about 32K complex columns) so fp32 accumulation never runs long; chunk
// partials are summed in fp64.
constexpr int kSplitKt = 16;        // complex k per stage = one bf16 MFMA k-step
constexpr int kSplitRowWords = 36;  // hi re, hi im, lo re, lo im (8 words each), 4 pad

__device__ __forceinline__ v8bf as_v8bf(uint4 u) { return __builtin_bit_cast(v8bf, u); }

// the (hi, lo) words of an operand element: fp64 values are split here,
// pre-split stack entries are used as stored
__device__ __forceinline__ uint2 operand_words(double2 v) { return split_word(v); }
__device__ __forceinline__ uint2 operand_words(uint2 w) { return w; }
__device__ __forceinline__ uint2 ld_or_zero(const uint2* p, bool ok) { return ok ? *p : make_uint2(0u, 0u); }

// two consecutive operand elements e, e + 1 of a row.  rows16: K even, so
// every pair (2m, 2m + 1) is one 16-byte aligned load of pre-split words.
__device__ __forceinline__ void ld_pair(const uint2* p, bool rows16, uint2& x0, uint2& x1) {
    if (rows16) {
        const uint4 u = *reinterpret_cast<const uint4*>(p);
        x0 = make_uint2(u.x, u.y);
        x1 = make_uint2(u.z, u.w);
    } else {
        x0 = p[0];
        x1 = p[1];
    }
}
__device__ __forceinline__ void ld_pair(const double2* p, bool, double2& x0, double2& x1) {
    x0 = p[0];
    x1 = p[1];
}

// real parts of two (re, im) bf16 words, and their imaginary parts
__device__ __forceinline__ uint32_t re_pair(uint32_t w0, uint32_t w1) { return (w0 & 0xffffu) | (w1 << 16); }
__device__ __forceinline__ uint32_t im_pair(uint32_t w0, uint32_t w1) { return (w0 >> 16) | (w1 & 0xffff0000u); }

// T = double2: fp64 stacks, split at staging, row n at n * ldk; T = uint2:
// stacks written pre-split by the recursion (the same words), half the bytes
// to stage, k-blocked with block stride wbs (word_offset; wbs = 0: rows at
// n * ldk)
template <typename T>
__global__ __launch_bounds__(512) void gamma_bf16x3_big_kernel(const T* __restrict__ A, const T* __restrict__ B,
                                                               int N, int64_t K, int64_t ldk, int64_t wbs, int64_t kc,
                                                               GammaMap map, double2* __restrict__ P) {
    __shared__ __align__(16) uint32_t As[2][kBigTile * kSplitRowWords];
    __shared__ __align__(16) uint32_t Bs[2][kBigTile * kSplitRowWords];
    int tn, tm;
    int64_t c;
    if (!xcd_work_item(map, tn, tm, c)) return;  // uniform over the workgroup
    // 8 waves: wave w owns the 64 x 32 block (rows (w >> 2) * 64, columns
    // (w & 3) * 32) of the 128 x 128 tile; two waves per SIMD, so one wave's
    // staging overlaps another's matrix-core work
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int wn = wave >> 2, wm = wave & 3;
    const int n0 = tn * kBigTile, m0 = tm * kBigTile;
    const int64_t kb = c * kc;
    const int64_t ke = (kb + kc < K) ? kb + kc : K;

    // staging: thread t takes complex columns 2p, 2p + 1 (p = t % 8) of rows
    // q and q + 64, q = 8 (t / 64) + 2 ((t / 8) % 4) + (t / 32) % 2: the 32
    // lanes of one LDS write pass cover rows r, r + 2, r + 4, r + 6, whose
    // 8-word groups fall on distinct banks
    const int p = tid & 7;
    const int q8 = (tid >> 3) & 7;
    const int r0 = 8 * wave + 2 * (q8 & 3) + (q8 >> 2);
    const T* ga[2];
    const T* gb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int r = r0 + 64 * i;
        const int rn = (n0 + r < N) ? n0 + r : N - 1;
        const int rm = (m0 + r < N) ? m0 + r : N - 1;
        const int64_t row_step = wbs ? kWordBlock : ldk;
        ga[i] = A + static_cast<int64_t>(rn) * row_step;
        gb[i] = B + static_cast<int64_t>(rm) * row_step;
    }
    T ra[2][2], rb[2][2];
    // row starts, and so column pairs, 16-byte aligned (blocked: always)
    const bool rows16 = wbs != 0 || (ldk & 1) == 0;
    auto gload = [&](int64_t k0) {
        const int64_t e = k0 + 2 * p;
        // a stage of 16 columns lies within one word block
        const int64_t o = word_offset(k0, wbs) + 2 * p;
        if (k0 + kSplitKt <= ke) {  // a full stage (all but a row's last, short one): plain loads
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                ld_pair(ga[i] + o, rows16, ra[i][0], ra[i][1]);
                ld_pair(gb[i] + o, rows16, rb[i][0], rb[i][1]);
            }
        } else {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                ra[i][0] = ld_or_zero(ga[i] + o, e < ke);
                ra[i][1] = ld_or_zero(ga[i] + o + 1, e + 1 < ke);
                rb[i][0] = ld_or_zero(gb[i] + o, e < ke);
                rb[i][1] = ld_or_zero(gb[i] + o + 1, e + 1 < ke);
            }
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int r = r0 + 64 * i;
            const uint2 a0 = operand_words(ra[i][0]), a1 = operand_words(ra[i][1]);
            const uint2 b0 = operand_words(rb[i][0]), b1 = operand_words(rb[i][1]);
            uint32_t* ar = &As[buf][r * kSplitRowWords + p];
            ar[0] = re_pair(a0.x, a1.x);
            ar[8] = im_pair(a0.x, a1.x);
            ar[16] = re_pair(a0.y, a1.y);
            ar[24] = im_pair(a0.y, a1.y);
            uint32_t* br = &Bs[buf][r * kSplitRowWords + p];
            br[0] = re_pair(b0.x, b1.x);
            br[8] = im_pair(b0.x, b1.x);
            br[16] = re_pair(b0.y, b1.y);
            br[24] = im_pair(b0.y, b1.y);
        }
    };

    v4f re[4][2], im[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            re[i][j] = v4f{0, 0, 0, 0};
            im[i][j] = v4f{0, 0, 0, 0};
        }

    // Fragments of one part (hi or lo) of a stage: A of the wave's 4 row
    // blocks; B and B' = (im, -re) of its 2 column blocks.  B' lanes: groups
    // 0, 1 read the im words, 2, 3 the re words with the sign bits flipped.
    struct FragB {
        v8bf b[2], bp[2];
    };
    const int prime_off = lk < 2 ? 8 + 4 * lk : 4 * lk - 8;
    const uint32_t prime_neg = lk < 2 ? 0u : 0x80008000u;
    auto read_a = [&](v8bf (&a)[4], int buf, int part) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            a[i] = as_v8bf(*reinterpret_cast<const uint4*>(
                &As[buf][(wn * 64 + 16 * i + li) * kSplitRowWords + 16 * part + 4 * lk]));
    };
    auto read_b = [&](FragB& f, int buf, int part) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const uint32_t* row = &Bs[buf][(wm * 32 + 16 * j + li) * kSplitRowWords + 16 * part];
            f.b[j] = as_v8bf(*reinterpret_cast<const uint4*>(row + 4 * lk));
            const uint4 u = *reinterpret_cast<const uint4*>(row + prime_off);
            f.bp[j] = as_v8bf(make_uint4(u.x ^ prime_neg, u.y ^ prime_neg, u.z ^ prime_neg, u.w ^ prime_neg));
        }
    };
    // 16 products with independent accumulators
    auto phase = [&](const v8bf (&a)[4], const FragB& f) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                re[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], f.b[j], re[i][j], 0, 0, 0);
                im[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], f.bp[j], im[i][j], 0, 0, 0);
            }
    };

    // Stage s sits in LDS buffer s % 2 and is multiplied in three phases,
    // hi.hi, hi.lo, lo.hi, each reading the next phase's fragments while its
    // products run; phase 3 reads stage s + 1's hi fragments.  The one
    // barrier per stage comes after the last read of stage s's buffer, which
    // then takes stage s + 2 (its global loads were issued at the start of
    // the stage).
    const int64_t stages = (ke - kb + kSplitKt - 1) / kSplitKt;
    v8bf ah[4], al[4];
    FragB bhi, blo;
    gload(kb);
    sstore(0);
    __syncthreads();
    read_a(ah, 0, 0);
    read_b(bhi, 0, 0);
    if (stages > 1) {
        gload(kb + kSplitKt);
        sstore(1);
    }
    for (int64_t st = 0; st < stages; ++st) {
        const int a = static_cast<int>(st & 1), b = a ^ 1;
        if (st + 2 < stages) gload(kb + (st + 2) * kSplitKt);
        read_b(blo, a, 1);
        phase(ah, bhi);
        read_a(al, a, 1);
        phase(ah, blo);
        __syncthreads();
        if (st + 2 < stages) sstore(a);
        v8bf ah_n[4];
        FragB bhi_n;
        if (st + 1 < stages) {
            read_a(ah_n, b, 0);
            read_b(bhi_n, b, 0);
        }
        phase(al, bhi);
#pragma unroll
        for (int i = 0; i < 4; ++i) ah[i] = ah_n[i];
        bhi = bhi_n;
    }

    double2* Pc = P + c * static_cast<int64_t>(N) * N;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = n0 + wn * 64 + 16 * i + 4 * lk + t;
                const int cl = m0 + wm * 32 + 16 * j + li;
                if (row < N && cl < N)
                    Pc[static_cast<int64_t>(row) * N + cl] =
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

// zeta(e_j) = sum_{n,m} Gamma[n][m] (u_n v_m + conj(v_n) u_m) / (1 - e_j^2)^2 / volume
// with e = cos(theta), s = sqrt(1 - e^2), a_0 = 1, a_{n>0} = 2,
//   u_n = a_n cos(n theta),  v_n = a_n (e - i n s) e^{i n theta}:
// the double sum of Gamma_nm g_mn(e) (g_mn of the Kubo-Bastin integrand) in
// separable form.  One workgroup takes EB energies; their u, v tables sit
// in LDS; wave w takes rows n = w, w+4, ..., lane l columns m = l, l+64, ...
// (a coalesced read of each Gamma row).  Partial sums are reduced in a
// fixed order.
template <int EB>
__global__ __launch_bounds__(256) void kubo_zeta_kernel(const double2* __restrict__ G, int N,
                                                        const double* __restrict__ e, int ne, double inv_volume,
                                                        double2* __restrict__ out) {
    extern __shared__ double tab[];  // [EB][N] u, then [EB][N] v.re, then [EB][N] v.im
    double* tu = tab;
    double* tvr = tab + EB * N;
    double* tvi = tab + 2 * EB * N;
    const int j0 = blockIdx.x * EB;
    for (int idx = threadIdx.x; idx < EB * N; idx += 256) {
        const int j = idx / N, n = idx - j * N;
        const int jj = (j0 + j < ne) ? j0 + j : ne - 1;
        const double ee = e[jj];
        const double th = acos(ee), es = sqrt(1.0 - ee * ee);
        double sn, cn;
        sincos(static_cast<double>(n) * th, &sn, &cn);
        const double an = (n == 0) ? 1.0 : 2.0;
        const double ns = static_cast<double>(n) * es;
        tu[idx] = an * cn;
        tvr[idx] = an * (ee * cn + ns * sn);
        tvi[idx] = an * (ee * sn - ns * cn);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double2 acc[EB];
#pragma unroll
    for (int j = 0; j < EB; ++j) acc[j] = make_double2(0.0, 0.0);
    for (int n = wave; n < N; n += kWavesPerGroup) {
        const double2* row = G + static_cast<int64_t>(n) * N;
        for (int m = lane; m < N; m += 64) {
            const double2 g = row[m];
#pragma unroll
            for (int j = 0; j < EB; ++j) {
                const double un = tu[j * N + n], vnr = tvr[j * N + n], vni = tvi[j * N + n];
                const double um = tu[j * N + m], vmr = tvr[j * N + m], vmi = tvi[j * N + m];
                const double cr = un * vmr + vnr * um;
                const double ci = un * vmi - vni * um;
                acc[j].x = acc[j].x + (g.x * cr - g.y * ci);
                acc[j].y = acc[j].y + (g.x * ci + g.y * cr);
            }
        }
    }
    __shared__ double2 red[kWavesPerGroup][EB];
#pragma unroll
    for (int j = 0; j < EB; ++j) {
        double x = acc[j].x, y = acc[j].y;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            x += __shfl_xor(x, off);
            y += __shfl_xor(y, off);
        }
        if (lane == 0) red[wave][j] = make_double2(x, y);
    }
    __syncthreads();
    if (threadIdx.x < EB && j0 + static_cast<int>(threadIdx.x) < ne) {
        const int j = threadIdx.x;
        double2 z = red[0][j];
        for (int w = 1; w < kWavesPerGroup; ++w) {
            z.x = z.x + red[w][j].x;
            z.y = z.y + red[w][j].y;
        }
        const double ee = e[j0 + j];
        const double d = (1.0 - ee * ee) * (1.0 - ee * ee);
        out[j0 + j] = make_double2(z.x / d * inv_volume, z.y / d * inv_volume);
    }
}

// Gamma[n][m] *= g_n g_m (kernel damping; the product is formed first, as
// the host-side weighting does).
__global__ __launch_bounds__(256) void gamma_weight_kernel(double2* __restrict__ G, const double* __restrict__ g,
                                                           int N) {
    const int64_t NN = static_cast<int64_t>(N) * N;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < NN;
         i += static_cast<int64_t>(gridDim.x) * 256) {
        const int64_t n = i / N, m = i - n * N;
        const double w = g[n] * g[m];
        double2 v = G[i];
        v.x = v.x * w;
        v.y = v.y * w;
        G[i] = v;
    }
}

dim3 gamma_grid(int N, int nchunks) {
    const unsigned t = static_cast<unsigned>((N + kGammaTile - 1) / kGammaTile);
    return dim3(t, t, static_cast<unsigned>(nchunks));
}

}  // namespace

void launch_gamma_f64(const double2* A, const double2* B, int N, int64_t K, int64_t ldk, int64_t kc, int nchunks,
                      double2* P, hipStream_t st) {
    hipLaunchKernelGGL(gamma_f64_kernel, gamma_grid(N, nchunks), dim3(256), 0, st, A, B, N, K, ldk, kc, P);
}

void launch_gamma_bf16x3(const double2* A, const double2* B, int N, int64_t K, int64_t ldk, int64_t kc, int nchunks,
                         double2* P, hipStream_t st) {
    hipLaunchKernelGGL(gamma_bf16x3_kernel, gamma_grid(N, nchunks), dim3(256), 0, st, A, B, N, K, ldk, kc, P);
}

void launch_gamma_reduce(const double2* P, int nchunks, int64_t NN, double2* out, hipStream_t st) {
    const int64_t blocks = (NN + 255) / 256;
    hipLaunchKernelGGL(gamma_reduce_kernel, dim3(static_cast<unsigned>(blocks < 65536 ? blocks : 65536)), dim3(256), 0,
                       st, P, nchunks, NN, out);
}


void launch_gamma_f64_big(const double2* A, const double2* B, int N, int64_t K, int64_t ldk, int64_t kc,
                          const GammaMap& map, double2* P, hipStream_t st) {
    hipLaunchKernelGGL(gamma_f64_big_kernel, dim3(static_cast<unsigned>(8 * map.per_xcd)), dim3(512), 0, st, A, B, N,
                       K, ldk, kc, map, P);
}

void launch_gamma_bf16x3_big(const double2* A, const double2* B, int N, int64_t K, int64_t ldk, int64_t kc,
                             const GammaMap& map, double2* P, hipStream_t st) {
    hipLaunchKernelGGL(gamma_bf16x3_big_kernel<double2>, dim3(static_cast<unsigned>(8 * map.per_xcd)), dim3(512), 0,
                       st, A, B, N, K, ldk, int64_t{0}, kc, map, P);
}

void launch_gamma_bf16x3_big_presplit(const uint2* A, const uint2* B, int N, int64_t K, int64_t wbs, int64_t kc,
                                      const GammaMap& map, double2* P, hipStream_t st) {
    hipLaunchKernelGGL(gamma_bf16x3_big_kernel<uint2>, dim3(static_cast<unsigned>(8 * map.per_xcd)), dim3(512), 0,
                       st, A, B, N, K, K, wbs, kc, map, P);
}

namespace {
__global__ __launch_bounds__(256) void split_block_kernel(const double2* __restrict__ in, uint2* __restrict__ out,
                                                          int64_t n, int64_t bs) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * 256)
        out[word_offset(i, bs)] = split_word(in[i]);
}
}  // namespace

void launch_split_block(const double2* in, uint2* out, int64_t n, int64_t bs, hipStream_t st) {
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>((n + 255) / 256, 65536));
    hipLaunchKernelGGL(split_block_kernel, dim3(grid), dim3(256), 0, st, in, out, n, bs);
}

void launch_gamma_weight(double2* G, const double* g, int N, hipStream_t st) {
    const int64_t NN = static_cast<int64_t>(N) * N;
    const int64_t blocks = (NN + 255) / 256;
    hipLaunchKernelGGL(gamma_weight_kernel, dim3(static_cast<unsigned>(blocks < 65536 ? blocks : 65536)), dim3(256), 0,
                       st, G, g, N);
}

int kubo_zeta_energies_per_group(int N) {
    const int64_t per = static_cast<int64_t>(N) * 3 * sizeof(double);
    if (per * 4 <= kZetaLdsBytes) return 4;
    if (per * 2 <= kZetaLdsBytes) return 2;
    if (per <= kZetaLdsBytes) return 1;
    return 0;
}

void launch_kubo_zeta(const double2* G, int N, const double* e, int ne, double inv_volume, double2* out,
                      hipStream_t st) {
    const int eb = kubo_zeta_energies_per_group(N);
    const size_t lds = static_cast<size_t>(eb) * N * 3 * sizeof(double);
    const dim3 grid(static_cast<unsigned>((ne + eb - 1) / eb));
    if (eb == 4)
        hipLaunchKernelGGL(kubo_zeta_kernel<4>, grid, dim3(256), lds, st, G, N, e, ne, inv_volume, out);
    else if (eb == 2)
        hipLaunchKernelGGL(kubo_zeta_kernel<2>, grid, dim3(256), lds, st, G, N, e, ne, inv_volume, out);
    else
        hipLaunchKernelGGL(kubo_zeta_kernel<1>, grid, dim3(256), lds, st, G, N, e, ne, inv_volume, out);
}

}  // namespace tbk
