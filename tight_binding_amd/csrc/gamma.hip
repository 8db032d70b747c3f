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

#include <type_traits>

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
    const int64_t tiles = static_cast<int64_t>(m.tiles_n) * m.tiles_m;
    if (w >= tiles * m.chunks) return false;
    c = w / tiles;
    const int t = static_cast<int>(w - c * tiles);
    tn = t / m.tiles_m;
    tm = t - tn * m.tiles_m;
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
// as gamma_f64_big_kernel.  An operand value x is split as hi = bf16(x),
// lo = bf16(x - hi) per real part, and a column's (hi, lo) words are
// (re, im) bf16 pairs (split_bf16.h); pre-split stacks arrive as those words
// already, fp64 stacks are split when staged.  Chunks are short (about 32K
// complex columns) so fp32 accumulation never runs long; chunk partials are
// summed in fp64.
//
// A stage row (16 complex columns of one entry) is kept as stored: (hi, lo)
// per column, 32 words, 128 bytes.  The k order of the matrix-core products
// is a free choice as long as both operands use it, so lane group g of a
// fragment takes columns 4g .. 4g + 3: its hi fragment is their hi words and
// its lo fragment their lo words, both out of one 32-byte read.  A word is
// (re, im), so A.B = sum ar br + ai bi = Re conj(a) b; B' = (im, -re) per
// column, formed from B in registers (a 16-bit rotate and a sign flip),
// gives A.B' = sum ar bi - ai br = Im conj(a) b.  Products hi.hi, hi.lo and
// lo.hi carry ~16 significant bits of each operand.
//
// LDS: two stage buffers of a 128-row A tile and a 128-row B tile, rows of
// 128 bytes = 8 chunks of 16 bytes, chunk c of row r at position
// c ^ ((r / 2) % 8): the 16 rows of a fragment read fall on distinct banks.
// Pre-split stacks are copied global -> LDS by the load unit
// (global_load_lds_dwordx4: lane L of a wave writes 16 bytes at the wave's
// base + 16 L, so lane L loads the chunk that belongs there), with no
// registers, VALU or LDS write instructions; fp64 stacks are loaded, split
// and written by the lanes in the same layout.  Stage s + 1 is fetched
// while stage s is multiplied, one barrier per stage.
//
// Measured at 512 x 512 x 2^23 (C3): staging through registers with the
// real and imaginary parts repacked (16 LDS writes and ~60 VALU per lane per
// stage) ran 48.4 ms in the solve; this form 43.7 ms.  Without any staging
// (LDS contents left as they are) the same loop runs at 99% of the matrix
// cores' peak, so staging is what remains.  Fetching two or three stages
// ahead (three or four buffers) was no faster (44.7, 45.6 ms).
constexpr int kSplitKt = 16;  // complex k per stage = one bf16 MFMA k-step

__device__ __forceinline__ v8bf as_v8bf(uint4 u) { return __builtin_bit_cast(v8bf, u); }
__device__ __forceinline__ uint2 ld_or_zero(const uint2* p, bool ok) { return ok ? *p : make_uint2(0u, 0u); }

constexpr int kRawRowBytes = 128;
constexpr int kRawTileBytes = kBigTile * kRawRowBytes;  // 16 KB

__device__ __forceinline__ uint32_t prime_word(uint32_t w) {
    return __builtin_amdgcn_alignbit(w, w, 16) ^ 0x80000000u;  // (re, im) -> (im, -re)
}

// Register staging of one 16-byte chunk (columns e, e + 1 of a row's stage;
// zero past ke): fp64 stacks always, pre-split stacks for a short last stage.
struct RawChunk {
    uint2 w0, w1;
    double2 v0, v1;
};
__device__ __forceinline__ void raw_fetch(RawChunk& ch, const uint2* src, bool full, int64_t e, int64_t ke) {
    ch.w0 = ld_or_zero(src, full || e < ke);
    ch.w1 = ld_or_zero(src + 1, full || e + 1 < ke);
}
__device__ __forceinline__ void raw_fetch(RawChunk& ch, const double2* src, bool full, int64_t e, int64_t ke) {
    ch.v0 = ld_or_zero(src, full || e < ke);
    ch.v1 = ld_or_zero(src + 1, full || e + 1 < ke);
}
__device__ __forceinline__ uint4 raw_words(const RawChunk& ch, const uint2*) {
    return make_uint4(ch.w0.x, ch.w0.y, ch.w1.x, ch.w1.y);
}
__device__ __forceinline__ uint4 raw_words(const RawChunk& ch, const double2*) {
    const uint2 w0 = split_word(ch.v0), w1 = split_word(ch.v1);
    return make_uint4(w0.x, w0.y, w1.x, w1.y);
}

template <typename T>
__global__ __launch_bounds__(512) void gamma_bf16x3_big_kernel(const T* __restrict__ A, const T* __restrict__ B,
                                                               int N, int64_t K, int64_t ldk, int64_t wbs, int64_t kc,
                                                               GammaMap map, double2* __restrict__ P) {
    // two stage buffers (A tile, then B tile), as separate objects, each use
    // site naming its own, so the compiler can tell the copies into one from
    // the reads of the other
    constexpr int kStageBytes = 2 * kRawTileBytes;
    __shared__ __align__(1024) uint8_t S0[kStageBytes];
    __shared__ __align__(1024) uint8_t S1[kStageBytes];
    int tn, tm;
    int64_t c;
    if (!xcd_work_item(map, tn, tm, c)) return;  // uniform over the workgroup
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int wn = wave >> 2, wm = wave & 3;
    const int n0 = tn * kBigTile, m0 = tm * kBigTile;
    const int64_t kb = c * kc;
    const int64_t ke = (kb + kc < K) ? kb + kc : K;

    // copies: wave w moves 8-row groups 2w, 2w + 1 of A (i = 0, 1) and of B
    // (i = 2, 3); lane L takes row r = 8 g + L / 8 and the logical chunk that
    // belongs at byte 16 L of the group, (L % 8) ^ ((r / 2) % 8)
    constexpr bool kPre = std::is_same<T, uint2>::value;  // pre-split, k-blocked (wbs != 0)
    const int lr = lane >> 3;
    int64_t roff[4];  // elements from a sub-stage's start to the lane's chunk
    int lcol[4];      // the chunk's first column within the sub-stage
    int dst_off[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int g = 2 * wave + (i & 1);
        const int r = 8 * g + lr;
        const int lci = (lane & 7) ^ ((r >> 1) & 7);
        const bool isb = i >= 2;
        const int row = isb ? ((m0 + r < N) ? m0 + r : N - 1) : ((n0 + r < N) ? n0 + r : N - 1);
        const int64_t row_step = kPre ? kWordBlock : ldk;
        roff[i] = static_cast<int64_t>(row) * row_step + 2 * lci;
        lcol[i] = 2 * lci;
        dst_off[i] = (isb ? kRawTileBytes : 0) + 8 * g * kRawRowBytes;
    }
    RawChunk ch[4];
    bool dma = false;  // the newest fetched stage went by the load unit
    // stage s into buffer buf: a full stage of pre-split words by the load
    // unit (a uniform base plus the lane's 32-bit offset), anything else
    // through registers (put writes them)
    auto fetch = [&](int64_t s, uint8_t* buf) {
        const int64_t k0 = kb + s * kSplitKt;
        const bool full = k0 + kSplitKt <= ke;
        const int64_t o = word_offset(k0, wbs);
        dma = kPre && full;
        if (dma) {
            const T* pa = A + o;
            const T* pb = B + o;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                // The host pass of this kernel never runs it and cannot form
                // an LDS address: the call exists in the device pass.
#if defined(__HIP_DEVICE_COMPILE__)
                __builtin_amdgcn_global_load_lds((i >= 2 ? pb : pa) + static_cast<uint32_t>(roff[i]),
                                                 (__attribute__((address_space(3))) void*)(buf + dst_off[i]), 16, 0,
                                                 0);
#endif
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) raw_fetch(ch[i], (i >= 2 ? B : A) + o + roff[i], full, k0 + lcol[i], ke);
        }
    };
    auto put = [&](uint8_t* buf) {
        if (dma) return;
#pragma unroll
        for (int i = 0; i < 4; ++i) *reinterpret_cast<uint4*>(buf + dst_off[i] + 16 * lane) = raw_words(ch[i], A);
    };

    v4f re[4][2], im[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            re[i][j] = v4f{0, 0, 0, 0};
            im[i][j] = v4f{0, 0, 0, 0};
        }
    // the 32 bytes (chunks 2 lk, 2 lk + 1) of tile row r: hi words of columns
    // 4 lk .. 4 lk + 3 in x, z of both, lo words in y, w
    auto rd = [&](const uint8_t* tile, int r, uint4& x0, uint4& x1) {
        const uint8_t* row = tile + r * kRawRowBytes;
        const int sw = (r >> 1) & 7;
        x0 = *reinterpret_cast<const uint4*>(row + (((2 * lk) ^ sw) << 4));
        x1 = *reinterpret_cast<const uint4*>(row + (((2 * lk + 1) ^ sw) << 4));
    };
    // fragments of a sub-stage: B, B' = (im, -re) and A, hi and lo parts
    struct Frags {
        v8bf bh[2], bl[2], bph[2], bpl[2], ah[4], al[4];
    };
    auto read_sub = [&](const uint8_t* buf, Frags& f) {
        const uint8_t* At = buf;
        const uint8_t* Bt = buf + kRawTileBytes;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            uint4 x0, x1;
            rd(Bt, wm * 32 + 16 * j + li, x0, x1);
            f.bh[j] = as_v8bf(make_uint4(x0.x, x0.z, x1.x, x1.z));
            f.bl[j] = as_v8bf(make_uint4(x0.y, x0.w, x1.y, x1.w));
            f.bph[j] = as_v8bf(make_uint4(prime_word(x0.x), prime_word(x0.z), prime_word(x1.x), prime_word(x1.z)));
            f.bpl[j] = as_v8bf(make_uint4(prime_word(x0.y), prime_word(x0.w), prime_word(x1.y), prime_word(x1.w)));
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint4 x0, x1;
            rd(At, wn * 64 + 16 * i + li, x0, x1);
            f.ah[i] = as_v8bf(make_uint4(x0.x, x0.z, x1.x, x1.z));
            f.al[i] = as_v8bf(make_uint4(x0.y, x0.w, x1.y, x1.w));
        }
    };
    // products of a sub-stage in three phases, hi.hi, hi.lo, lo.hi, each over
    // the 16 accumulators, so consecutive products into one accumulator are
    // 16 instructions apart
    auto mul_sub = [&](const Frags& f) {
#pragma unroll
        for (int ph = 0; ph < 3; ++ph)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const v8bf& a = ph == 2 ? f.al[i] : f.ah[i];
                    const v8bf& b = ph == 1 ? f.bl[j] : f.bh[j];
                    const v8bf& bp = ph == 1 ? f.bpl[j] : f.bph[j];
                    re[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, re[i][j], 0, 0, 0);
                    im[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bp, im[i][j], 0, 0, 0);
                }
    };

    const int64_t stages = (ke - kb + kSplitKt - 1) / kSplitKt;
    fetch(0, S0);
    put(S0);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    // Stage st: its fragments are read from cur, stage st + 1 is fetched
    // into nxt (after the reads: the compiler makes LDS reads that follow a
    // copy into LDS wait for it), then the products.  Before the barrier this
    // wave's copies have landed and its LDS writes are done; the barrier
    // publishes stage st + 1 and frees cur.
    auto step = [&](int64_t st, const uint8_t* cur, uint8_t* nxt) {
        Frags f;
        read_sub(cur, f);
        const bool ahead = st + 1 < stages;
        if (ahead) fetch(st + 1, nxt);
        mul_sub(f);
        if (ahead) put(nxt);
        __syncthreads();
    };
    for (int64_t st = 0; st < stages; st += 2) {
        step(st, S0, S1);
        if (st + 1 < stages) step(st + 1, S1, S0);
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

// ---------------------------------------------------------------------------
// Wide form of the pre-split contraction: each workgroup owns a 256 x 128
// tile (8 waves, 64 x 64 each = 4 x 4 blocks of 16 x 16), so every staged
// word feeds twice the products of the 128 x 128 form.  Per stage a wave
// reads the hi and lo fragments of its 4 A and 4 B blocks once (256 bytes a
// lane); for each half of its B blocks it makes the 24 Re products (hi.hi,
// hi.lo, lo.hi), turns those B fragments into B' = (im, -re) in place and
// makes the 24 Im products: the same sums as the 128 x 128 kernel (the same
// fragments in the same k order, the same phases per accumulator), so Gamma
// is the same bits.  LDS: three stage
// buffers of a 256-row A tile and a 128-row B tile (rows of 128 bytes,
// chunk c of row r at c ^ wide_swizzle(r)), 144 KB; a full stage goes global
// -> LDS by the load unit two stages ahead of its products, a short last
// stage through registers, written at once.
constexpr int kWideStageBytes = (kWideTile + kBigTile) * kRawRowBytes;  // 48 KB

// Position of a row's 16-byte chunks: chunk c of row r sits at c ^ wide_swizzle(r).
// Rows are 128 bytes, so a chunk's banks depend on its position (and, with
// 64 banks, on the row's parity): any 8 consecutive rows of a fragment read
// get distinct positions, and rows r, r + 8 of one 16-row read differ too.
__device__ __forceinline__ int wide_swizzle(int r) { return (r & 7) ^ ((r >> 3) & 1); }

// End of a stage of the three-buffer pipeline: this wave's copies of the
// next stage have landed, its LDS accesses are done, then the workgroup
// barrier.  Copies complete in issue order, so with later_in_flight the 6
// copies of the stage after the next one (issued last) may stay in flight.
// The empty asm statements keep the compiler from moving memory accesses
// across the barrier.  The waits are s_waitcnt builtins, not asm text, so
// the compiler's own wait placement knows which copies have landed and adds
// no wait before reading a landed buffer.  gfx9 encoding: vmcnt in bits 3:0
// (and 15:14), expcnt in 6:4 (7: no wait), lgkmcnt in 11:8.
__device__ __forceinline__ void wide_stage_barrier(bool later_in_flight) {
    asm volatile("" ::: "memory");
    if (later_in_flight)
        __builtin_amdgcn_s_waitcnt(0x0076);  // vmcnt(6) lgkmcnt(0)
    else
        __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

__global__ __launch_bounds__(512) void gamma_bf16x3_wide_kernel(const uint2* __restrict__ A,
                                                                const uint2* __restrict__ B, int N, int64_t K,
                                                                int64_t wbs, int64_t kc, GammaMap map,
                                                                double2* __restrict__ P) {
    __shared__ __align__(1024) uint8_t S0[kWideStageBytes];
    __shared__ __align__(1024) uint8_t S1[kWideStageBytes];
    __shared__ __align__(1024) uint8_t S2[kWideStageBytes];
    int tn, tm;
    int64_t c;
    if (!xcd_work_item(map, tn, tm, c)) return;  // uniform over the workgroup
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int li = lane & 15, lk = lane >> 4;
    const int wn = wave >> 1, wm = wave & 1;
    const int n0 = tn * kWideTile, m0 = tm * kBigTile;
    const int64_t kb = c * kc;
    const int64_t ke = (kb + kc < K) ? kb + kc : K;

    // copies: wave w moves 8-row groups 4w .. 4w + 3 of A (i = 0 .. 3) and
    // 2w, 2w + 1 of B (i = 4, 5); lane L takes row r = 8 g + L / 8 and the
    // logical chunk that belongs at byte 16 L of the group, (L % 8) ^ ((r / 2) % 8)
    const int lr = lane >> 3;
    uint32_t roff[6];  // words from a sub-stage's start to the lane's chunk
    int lcol[6];       // the chunk's first column within the sub-stage
    int dst_off[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const bool isb = i >= 4;
        const int g = isb ? 2 * wave + (i - 4) : 4 * wave + i;
        const int r = 8 * g + lr;
        const int lci = (lane & 7) ^ wide_swizzle(r);
        const int row = isb ? ((m0 + r < N) ? m0 + r : N - 1) : ((n0 + r < N) ? n0 + r : N - 1);
        roff[i] = static_cast<uint32_t>(row) * kWordBlock + 2 * lci;
        lcol[i] = 2 * lci;
        dst_off[i] = (isb ? kWideTile * kRawRowBytes : 0) + 8 * g * kRawRowBytes;
    }
    // a short last stage goes through registers and is written at once (no
    // registers held across the products)
    // stage counts are small (a chunk is at most a few thousand stages), so
    // the stage tests are 32-bit scalar compares
    const int nfull = static_cast<int>((ke - kb) / kSplitKt);
    auto fetch_full = [&](int s, uint8_t* buf) {
        const int64_t o = word_offset(kb + static_cast<int64_t>(s) * kSplitKt, wbs);
        const uint2* pa = A + o;
        const uint2* pb = B + o;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
#if defined(__HIP_DEVICE_COMPILE__)
            __builtin_amdgcn_global_load_lds((i >= 4 ? pb : pa) + roff[i],
                                             (__attribute__((address_space(3))) void*)(buf + dst_off[i]), 16, 0, 0);
#endif
        }
    };
    auto fetch = [&](int s, uint8_t* buf) {
        if (s < nfull) {
            fetch_full(s, buf);
        } else {
            const int64_t k0 = kb + static_cast<int64_t>(s) * kSplitKt;
            const int64_t o = word_offset(k0, wbs);
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                const uint2* src = (i >= 4 ? B : A) + o + roff[i];
                const int64_t e = k0 + lcol[i];
                const uint2 w0 = ld_or_zero(src, e < ke), w1 = ld_or_zero(src + 1, e + 1 < ke);
                *reinterpret_cast<uint4*>(buf + dst_off[i] + 16 * lane) = make_uint4(w0.x, w0.y, w1.x, w1.y);
            }
        }
    };

    v4f re[4][4], im[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            re[i][j] = v4f{0, 0, 0, 0};
            im[i][j] = v4f{0, 0, 0, 0};
        }
    auto rd = [&](const uint8_t* tile, int r, uint4& x0, uint4& x1) {
        const uint8_t* row = tile + r * kRawRowBytes;
        const int sw = wide_swizzle(r);
        x0 = *reinterpret_cast<const uint4*>(row + (((2 * lk) ^ sw) << 4));
        x1 = *reinterpret_cast<const uint4*>(row + (((2 * lk + 1) ^ sw) << 4));
    };
    // B fragments are read, used and turned into B' two 16-column blocks at a
    // time (j = 2h, 2h + 1), so the A fragments, two B blocks and the 128
    // accumulators fit the register budget of two waves per SIMD
    auto half = [&](const uint8_t* Bt, int h, const v8bf (&ah)[4], const v8bf (&al)[4]) {
        v8bf bh[2], bl[2];
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
            uint4 x0, x1;
            rd(Bt, wm * 64 + 16 * (2 * h + jj) + li, x0, x1);
            bh[jj] = as_v8bf(make_uint4(x0.x, x0.z, x1.x, x1.z));
            bl[jj] = as_v8bf(make_uint4(x0.y, x0.w, x1.y, x1.w));
        }
        // Re: hi.hi, hi.lo, lo.hi
#pragma unroll
        for (int ph = 0; ph < 3; ++ph)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) {
                    const int j = 2 * h + jj;
                    re[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ph == 2 ? al[i] : ah[i],
                                                                       ph == 1 ? bl[jj] : bh[jj], re[i][j], 0, 0, 0);
                }
        // B' = (im, -re) per word, in place
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
            uint4 hw = __builtin_bit_cast(uint4, bh[jj]), lw = __builtin_bit_cast(uint4, bl[jj]);
            hw = make_uint4(prime_word(hw.x), prime_word(hw.y), prime_word(hw.z), prime_word(hw.w));
            lw = make_uint4(prime_word(lw.x), prime_word(lw.y), prime_word(lw.z), prime_word(lw.w));
            bh[jj] = as_v8bf(hw);
            bl[jj] = as_v8bf(lw);
        }
#pragma unroll
        for (int ph = 0; ph < 3; ++ph)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) {
                    const int j = 2 * h + jj;
                    im[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ph == 2 ? al[i] : ah[i],
                                                                       ph == 1 ? bl[jj] : bh[jj], im[i][j], 0, 0, 0);
                }
    };
    const int stages = static_cast<int>((ke - kb + kSplitKt - 1) / kSplitKt);
    {
        // Copies two stages ahead: stage st + 2 goes into the buffer that
        // stage st - 1 was read from (freed by the barrier that ended stage
        // st - 1), so a stage's copies have a whole stage of products to land
        // in; with two buffers they had only the products of the stage that
        // issued them (38.6 against 37.7 ms at N = 512, K = 2^23).  A full
        // stage is 6 copies a lane; a short one is written through registers
        // before its fetch returns.
        // Steps whose stage st + 2 is full (main loop) end with vmcnt(6):
        // every path through them issues the 6 copies, so the compiler's wait
        // placement can tell which buffer is still being filled.  The last
        // few steps (a short or missing stage ahead) wait for everything.
        auto step3 = [&](int st, const uint8_t* cur, uint8_t* nxt2, bool tail) {
            v8bf ah[4], al[4];
            const uint8_t* At = cur;
            const uint8_t* Bt = cur + kWideTile * kRawRowBytes;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                uint4 x0, x1;
                rd(At, wn * 64 + 16 * i + li, x0, x1);
                ah[i] = as_v8bf(make_uint4(x0.x, x0.z, x1.x, x1.z));
                al[i] = as_v8bf(make_uint4(x0.y, x0.w, x1.y, x1.w));
            }
            if (!tail)
                fetch_full(st + 2, nxt2);
            else if (st + 2 < stages)
                fetch(st + 2, nxt2);
            half(Bt, 0, ah, al);
            half(Bt, 1, ah, al);
            wide_stage_barrier(!tail);
        };
        fetch(0, S0);
        if (stages > 1) fetch(1, S1);
        wide_stage_barrier(false);
        int st = 0;
        for (; st + 4 < nfull; st += 3) {
            step3(st, S0, S2, false);
            step3(st + 1, S1, S0, false);
            step3(st + 2, S2, S1, false);
        }
        for (; st < stages; st += 3) {
            step3(st, S0, S2, true);
            if (st + 1 < stages) step3(st + 1, S1, S0, true);
            if (st + 2 < stages) step3(st + 2, S2, S1, true);
        }
    }

    double2* Pc = P + c * static_cast<int64_t>(N) * N;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = n0 + wn * 64 + 16 * i + 4 * lk + t;
                const int cl = m0 + wm * 64 + 16 * j + li;
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
    gamma_bf16x3_big_kernel<double2><<<dim3(static_cast<unsigned>(8 * map.per_xcd)), dim3(512), 0, st>>>(
        A, B, N, K, ldk, int64_t{0}, kc, map, P);
}

void launch_gamma_bf16x3_big_presplit(const uint2* A, const uint2* B, int N, int64_t K, int64_t wbs, int64_t kc,
                                      const GammaMap& map, double2* P, hipStream_t st) {
    gamma_bf16x3_big_kernel<uint2><<<dim3(static_cast<unsigned>(8 * map.per_xcd)), dim3(512), 0, st>>>(
        A, B, N, K, K, wbs, kc, map, P);
}

void launch_gamma_bf16x3_wide_presplit(const uint2* A, const uint2* B, int N, int64_t K, int64_t wbs, int64_t kc,
                                       const GammaMap& map, double2* P, hipStream_t st) {
    gamma_bf16x3_wide_kernel<<<dim3(static_cast<unsigned>(8 * map.per_xcd)), dim3(512), 0, st>>>(A, B, N, K, wbs, kc,
                                                                                               map, P);
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
