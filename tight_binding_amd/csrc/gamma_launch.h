// Kubo-Bastin moment contraction Gamma[n][m] = sum_k conj(A[n][k]) B[m][k]
// over two stacks of N vector blocks (A = T_n(H~)|r>, B = v_x T_m(H~) v_y|r>,
// k running over sites x vector columns).  Launchers implemented in
// gamma.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tbk {

constexpr int kGammaTile = 64;  // Gamma rows/cols per workgroup
constexpr int kBigTile = 128;   // Gamma rows/cols per workgroup of the large-N kernels

// Work placement of the large-N kernels: tiles_n x tiles_m tiles x `chunks`
// k chunks, per_xcd work items per XCD (the grid is 8 * per_xcd workgroups).
struct GammaMap {
    int32_t tiles_n;
    int32_t chunks;
    int64_t per_xcd;
    int32_t tiles_m;
    int32_t pad_;
};

// Gamma rows (A entries) per workgroup of the wide pre-split kernel; its
// columns (B entries) are kBigTile.
constexpr int kWideTile = 256;

// Split-K partial sums P[c][n][m] for k in chunk c = [c*kc, min(K, (c+1)*kc)),
// fp64 matrix cores.  A, B: N rows of K complex128 each, row n at n * ldk
// (ldk >= K; every launcher below takes the same row stride).
void launch_gamma_f64(const double2* A, const double2* B, int N, int64_t K, int64_t ldk, int64_t kc, int nchunks,
                      double2* P, hipStream_t st);

// Same partial sums with 128 x 128 tiles staged through LDS and the chunks
// placed per XCD (for N >= 256).
void launch_gamma_f64_big(const double2* A, const double2* B, int N, int64_t K, int64_t ldk, int64_t kc,
                          const GammaMap& map, double2* P, hipStream_t st);

// Split-BF16 three-term variant: each fp64 operand value x is split as
// hi = bf16(x), lo = bf16(x - hi); products hi*hi + hi*lo + lo*hi are
// accumulated in fp32 on matrix cores within a chunk and summed in fp64
// across chunks.
void launch_gamma_bf16x3(const double2* A, const double2* B, int N, int64_t K, int64_t ldk, int64_t kc, int nchunks,
                         double2* P, hipStream_t st);

// Split-bf16 x3 partial sums with 128 x 128 tiles (operands split once, at
// LDS staging) and the chunks placed per XCD (for N >= 256).
void launch_gamma_bf16x3_big(const double2* A, const double2* B, int N, int64_t K, int64_t ldk, int64_t kc,
                             const GammaMap& map, double2* P, hipStream_t st);

// The same contraction over stacks written pre-split by the recursion: each
// complex element is its (hi, lo) words (split_bf16.h), 8 bytes instead of 16,
// so the partial sums equal launch_gamma_bf16x3_big's on the fp64 stacks.  The
// stacks are k-blocked with block stride wbs = N * kWordBlock (word_offset).
void launch_gamma_bf16x3_big_presplit(const uint2* A, const uint2* B, int N, int64_t K, int64_t wbs, int64_t kc,
                                      const GammaMap& map, double2* P, hipStream_t st);

// The same partial sums over pre-split stacks with 256 x 128 tiles, 64 x 64
// per wave (wider reuse of every staged word; map.tiles_n counts 256-row tiles).
void launch_gamma_bf16x3_wide_presplit(const uint2* A, const uint2* B, int N, int64_t K, int64_t wbs, int64_t kc,
                                       const GammaMap& map, double2* P, hipStream_t st);

// out[word_offset(i, bs)] = split words of in[i] (n complex values)
void launch_split_block(const double2* in, uint2* out, int64_t n, int64_t bs, hipStream_t st);

// out[i] = sum_{c < nchunks} P[c*NN + i], chunks in order.
void launch_gamma_reduce(const double2* P, int nchunks, int64_t NN, double2* out, hipStream_t st);

// Gamma[n][m] *= g[n] * g[m] in place.
void launch_gamma_weight(double2* G, const double* g, int N, hipStream_t st);

// Kubo-Bastin integrand at rescaled energies e[0..ne) (|e| < 1) from the
// N x N complex Gamma (row-major, kernel weights already applied):
//   out[j] = sum_{n,m} Gamma[n][m] g_mn(e_j) / (1 - e_j^2)^2 * inv_volume.
// Tables of 3 N doubles per energy live in LDS; N must satisfy
// kubo_zeta_energies_per_group(N) > 0.
constexpr int64_t kZetaLdsBytes = 64 * 1024;
int kubo_zeta_energies_per_group(int N);
void launch_kubo_zeta(const double2* G, int N, const double* e, int ne, double inv_volume, double2* out,
                      hipStream_t st);

}  // namespace tbk
