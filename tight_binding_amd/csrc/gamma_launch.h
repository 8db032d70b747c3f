// Kubo-Bastin moment contraction Gamma[n][m] = sum_k conj(A[n][k]) B[m][k]
// over two stacks of N vector blocks (A = T_n(H~)|r>, B = v_x T_m(H~) v_y|r>,
// k running over sites x vector columns).  Launchers implemented in
// gamma.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tbk {

constexpr int kGammaTile = 64;  // Gamma rows/cols per workgroup

// Split-K partial sums P[c][n][m] for k in chunk c = [c*kc, min(K, (c+1)*kc)),
// fp64 matrix cores.  A, B: N rows of K complex128 each.
void launch_gamma_f64(const double2* A, const double2* B, int N, int64_t K, int64_t kc, int nchunks, double2* P,
                      hipStream_t st);

// Split-BF16 three-term variant: each fp64 operand value x is split as
// hi = bf16(x), lo = bf16(x - hi); products hi*hi + hi*lo + lo*hi are
// accumulated in fp32 on matrix cores within a chunk and summed in fp64
// across chunks.
void launch_gamma_bf16x3(const double2* A, const double2* B, int N, int64_t K, int64_t kc, int nchunks, double2* P,
                         hipStream_t st);

// out[i] = sum_{c < nchunks} P[c*NN + i], chunks in order.
void launch_gamma_reduce(const double2* P, int nchunks, int64_t NN, double2* out, hipStream_t st);

}  // namespace tbk
