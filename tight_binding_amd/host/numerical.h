// Host numerics of the KPM front end (numerical.h:120-144, 304-316,
// 383-447 of the reference): damping kernels, Chebyshev expansion
// coefficients of the resolvent, random-phase kets and block traces.
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <numbers>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "lattice.h"
#include "linalg.h"
#include "tbkpm.h"

namespace tb::numerical {

using std::numbers::pi;

namespace linalg {

// Generator state behind bases_of_uniform_ampl_with_random_phases: every call
// draws the next seed, so a program is reproducible run to run (the
// reference draws from std::rand through Eigen's Random).
inline uint64_t& ket_seed_state() {
    static uint64_t s = 0x5EEDu;
    return s;
}
inline void seed_random_kets(uint64_t seed) { ket_seed_state() = seed; }

// kets(i, r) = e^{i 2 pi u(seed, i*m + r)} / sqrt(n)  (numerical.h:120-122)
inline MatrixXcd bases_of_uniform_ampl_with_random_phases(int n, int m) {
    const uint64_t seed = ket_seed_state()++;
    MatrixXcd k(n, m);
    const double inv = 1.0 / std::sqrt(static_cast<double>(n));
    for (int r = 0; r < m; ++r)
        for (int i = 0; i < n; ++i) {
            const double u = counter_uniform(seed, static_cast<uint64_t>(i) * m + r);
            const double th = 2.0 * pi * u;
            k(i, r) = Complex(std::cos(th) * inv, std::sin(th) * inv);
        }
    return k;
}

// {E_min, E_max} of a Hermitian sparse matrix (numerical.h:23-46, both
// overloads): FP64 Lanczos on the device (tbk_spectral_range), any size.
// The reference's Spectra solver refuses more than 5000 rows and returns
// {0, 0}; this throws std::runtime_error on a device failure instead.  A
// template so that code which never calls it does not link libtbkpm.
template <typename Sparse>
std::pair<double, double> matrix_spectral_range_sparse(const Sparse& A) {
    using Scalar = std::remove_cv_t<std::remove_pointer_t<decltype(A.valuePtr())>>;
    tbk_sparse s{};
    s.rows = static_cast<int32_t>(A.rows());
    s.cols = static_cast<int32_t>(A.cols());
    s.nnz = static_cast<int64_t>(A.nonZeros());
    s.outer = reinterpret_cast<const int32_t*>(A.outerIndexPtr());
    s.inner = reinterpret_cast<const int32_t*>(A.innerIndexPtr());
    s.values = reinterpret_cast<const double*>(A.valuePtr());
    s.is_complex = std::is_same_v<Scalar, Complex> ? 1 : 0;
    s.layout = TBK_CSC;
    double lo = 0.0, hi = 0.0;
    if (tbk_spectral_range(&s, &lo, &hi) != TBK_OK)
        throw std::runtime_error(std::string("matrix_spectral_range_sparse: ") + tbk_last_error());
    return {lo, hi};
}

// Re tr(bras^H kets) (numerical.h:127-135)
inline double trace_braket(const MatrixXcd& bras_T, const MatrixXcd& kets) {
    double t = 0.0;
    for (int64_t i = 0; i < bras_T.rows(); ++i)
        for (int64_t j = 0; j < bras_T.cols(); ++j)
            t += bras_T(i, j).real() * kets(i, j).real() + bras_T(i, j).imag() * kets(i, j).imag();
    return t;
}

// tr(bras^H kets) = sum conj(bras_T(i, j)) kets(i, j)  (numerical.h:137-144)
inline Complex trace_braket_complex(const MatrixXcd& bras_T, const MatrixXcd& kets) {
    Complex t = 0.0;
    for (int64_t i = 0; i < bras_T.rows(); ++i)
        for (int64_t j = 0; j < bras_T.cols(); ++j) t += std::conj(bras_T(i, j)) * kets(i, j);
    return t;
}

}  // namespace linalg

namespace Chebyshev {

// 1/(z - x) = sum alpha_n(z) T_n(x)  (numerical.h:304-316)
inline std::vector<Complex> expn_coef_alpha_seq(Complex z, int n) {
    using namespace std::complex_literals;
    std::vector<Complex> g(static_cast<size_t>(n));
    const Complex s1 = std::sqrt(1.0 - z * z);
    const Complex s2 = z - 1.0i * s1;
    Complex sn = -2.0i / s1;
    g[0] = -1.0i / s1;
    for (int i = 1; i < n; ++i) {
        sn *= s2;
        g[i] = sn;
    }
    return g;
}

// [T_0(x) v, ..., T_{n-1}(x) v] for a sparse x (already rescaled) and a
// rows x R complex block v (numerical.h:334-344), formed in FP64 on the
// device (tbk_cheb_first_kind_seq_on_bases).  n = 1 gives {v}; n < 1 and
// device failures throw (the reference's loop writes past its vector for
// n < 2).  A template so that code which never calls it does not link
// libtbkpm.
template <typename Sparse>
std::vector<MatrixXcd> Cheb_first_kind_seq_on_bases(const Sparse& x, int n, const MatrixXcd& v) {
    using Scalar = std::remove_cv_t<std::remove_pointer_t<decltype(x.valuePtr())>>;
    tbk_sparse s{};
    s.rows = static_cast<int32_t>(x.rows());
    s.cols = static_cast<int32_t>(x.cols());
    s.nnz = static_cast<int64_t>(x.nonZeros());
    s.outer = reinterpret_cast<const int32_t*>(x.outerIndexPtr());
    s.inner = reinterpret_cast<const int32_t*>(x.innerIndexPtr());
    s.values = reinterpret_cast<const double*>(x.valuePtr());
    s.is_complex = std::is_same_v<Scalar, Complex> ? 1 : 0;
    s.layout = TBK_CSC;
    if (n < 1) throw std::invalid_argument("Cheb_first_kind_seq_on_bases: n must be at least 1");
    const size_t blk = static_cast<size_t>(v.rows() * v.cols());
    std::vector<Complex> all(static_cast<size_t>(n) * blk);
    if (tbk_cheb_first_kind_seq_on_bases(&s, n, reinterpret_cast<const double*>(v.data()),
                                         static_cast<int32_t>(v.rows()), static_cast<int32_t>(v.cols()),
                                         reinterpret_cast<double*>(all.data())) != TBK_OK)
        throw std::runtime_error(std::string("Cheb_first_kind_seq_on_bases: ") + tbk_last_error());
    std::vector<MatrixXcd> Tn;
    Tn.reserve(static_cast<size_t>(n));
    for (int k = 0; k < n; ++k) {
        MatrixXcd t(v.rows(), v.cols());
        std::copy(all.begin() + static_cast<std::ptrdiff_t>(k * blk),
                  all.begin() + static_cast<std::ptrdiff_t>((k + 1) * blk), t.data());
        Tn.push_back(std::move(t));
    }
    return Tn;
}

}  // namespace Chebyshev

namespace kernels {

struct Dirichlet {
    double operator()(int) const { return 1.0; }
};
struct Fejer {
    int N;
    explicit Fejer(int N) : N(N) {}
    double operator()(int n) const { return 1.0 - static_cast<double>(n) / N; }
};
struct Lorentz {
    int N;
    double lambda;
    Lorentz(int N, double lambda) : N(N), lambda(lambda) {}
    double operator()(int n) const { return std::sinh(lambda * (1.0 - static_cast<double>(n) / N)) / std::sinh(lambda); }
};
struct Jackson {
    int N1;
    explicit Jackson(int N) : N1(N + 1) {}
    double operator()(int n) const {
        const double s1 = pi / N1;
        const double s2 = n * s1;
        return ((N1 - n) * std::cos(s2) + std::sin(s2) / std::tan(s1)) / N1;
    }
};
// as numerical.h:428-435: (sin(n pi/N) / (pi/N))^M (no 1/n factor)
struct Lanczos {
    double pi_N;
    int M;
    Lanczos(int N, int M) : pi_N(pi / N), M(M) {}
    double operator()(int n) const { return std::pow(std::sin(n * pi_N) / pi_N, M); }
};
struct Wang_and_Zunger {
    double alpha_N, beta;
    Wang_and_Zunger(int N, double alpha, double beta) : alpha_N(alpha / N), beta(beta) {}
    double operator()(int n) const { return std::exp(-std::pow(n * alpha_N, beta)); }
};

template <typename K>
std::vector<double> kernel_seq(int N, const K& kernel) {
    std::vector<double> g(static_cast<size_t>(N));
    for (int n = 0; n < N; ++n) g[n] = kernel(n);
    return g;
}

}  // namespace kernels

}  // namespace tb::numerical
