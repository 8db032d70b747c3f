// Drop-in for the reference's numerics_cuda.cuh (numerics_cuda.cuh:8-14):
// the two overloads of cuda::Cheb_first_kind_seq_braket, here backed by the
// HIP engine through the C ABI of include/tbkpm.h.
//
//   mu[i] = sum over rows and columns of Re(conj(bras_T) * T_i(M) kets),
//   T_0 = kets, T_1 = M kets, T_i = 2 M T_{i-1} - T_{i-2}, i < n,
//
// M is the already rescaled matrix (GreenFunction::scaled_HE), stored
// column-major compressed (an uncompressed Eigen matrix is rejected, as the
// reference's comments at numerics_cuda.cu:189,266 require); bras_T and kets
// are rows x R complex blocks of the same shape.  The two-moments-per-step
// path runs only for bras_T == kets (same storage) and an exactly Hermitian
// M; anything else takes one moment per step.  n < 2 prints the
// reference's message and returns an empty vector (numerics_cuda.cu:194-197);
// a device failure throws std::runtime_error (the reference ignores CUDA
// status codes).  Nothing is printed per step.
//
// Accepted types: tb::SparseMatrix / tb::MatrixXcd of this package, and any
// type with Eigen's accessors (rows(), cols(), nonZeros(), outerIndexPtr(),
// innerIndexPtr(), valuePtr(), data()) -- including Eigen::SparseMatrix and
// Eigen::MatrixXcd when the caller has Eigen.
#pragma once

#include <complex>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "tbkpm.h"

namespace cuda {

using Complex = std::complex<double>;

namespace detail {

// false only for a sparse type that reports itself uncompressed (Eigen's
// isCompressed()); types without the accessor are compressed by construction
template <typename Sparse>
auto is_compressed(const Sparse& M, int) -> decltype(M.isCompressed(), bool()) {
    return M.isCompressed();
}
template <typename Sparse>
bool is_compressed(const Sparse&, long) {
    return true;
}

template <typename Sparse, typename Dense>
std::vector<double> cheb_braket(const Sparse& M, int n, const Dense& bras_T, const Dense& kets) {
    using Scalar = std::remove_cv_t<std::remove_pointer_t<decltype(M.valuePtr())>>;
    static_assert(std::is_same_v<Scalar, double> || std::is_same_v<Scalar, Complex>,
                  "matrix scalar must be double or std::complex<double>");
    if (n < 2) {
        std::cerr << "n must be greater than or equal to 2" << std::endl;
        return {};
    }
    if (bras_T.rows() != kets.rows() || bras_T.cols() != kets.cols())
        throw std::invalid_argument("Cheb_first_kind_seq_braket: bras_T and kets differ in shape");
    if (!is_compressed(M, 0))
        throw std::invalid_argument("Cheb_first_kind_seq_braket: M must be compressed (makeCompressed())");
    tbk_sparse s{};
    s.rows = static_cast<int32_t>(M.rows());
    s.cols = static_cast<int32_t>(M.cols());
    s.nnz = static_cast<int64_t>(M.nonZeros());
    s.outer = reinterpret_cast<const int32_t*>(M.outerIndexPtr());
    s.inner = reinterpret_cast<const int32_t*>(M.innerIndexPtr());
    s.values = reinterpret_cast<const double*>(M.valuePtr());
    s.is_complex = std::is_same_v<Scalar, Complex> ? 1 : 0;
    s.layout = TBK_CSC;
    std::vector<double> mu(static_cast<size_t>(n));
    const double* b = reinterpret_cast<const double*>(bras_T.data());
    const double* k = reinterpret_cast<const double*>(kets.data());
    const int rc = tbk_cheb_first_kind_seq_braket(&s, n, b, k, static_cast<int32_t>(kets.rows()),
                                                  static_cast<int32_t>(kets.cols()), mu.data());
    if (rc != TBK_OK) throw std::runtime_error(std::string("Cheb_first_kind_seq_braket: ") + tbk_last_error());
    return mu;
}

}  // namespace detail

// real-M overload (numerics_cuda.cu:190-255)
template <typename Sparse, typename Dense>
auto Cheb_first_kind_seq_braket(const Sparse& M, int n, const Dense& bras_T, const Dense& kets)
    -> std::enable_if_t<std::is_same_v<std::remove_cv_t<std::remove_pointer_t<decltype(M.valuePtr())>>, double>,
                        std::vector<double>> {
    return detail::cheb_braket(M, n, bras_T, kets);
}

// complex-M overload (numerics_cuda.cu:267-323)
template <typename Sparse, typename Dense>
auto Cheb_first_kind_seq_braket(const Sparse& M, int n, const Dense& bras_T, const Dense& kets)
    -> std::enable_if_t<std::is_same_v<std::remove_cv_t<std::remove_pointer_t<decltype(M.valuePtr())>>, Complex>,
                        std::vector<double>> {
    return detail::cheb_braket(M, n, bras_T, kets);
}

}  // namespace cuda
