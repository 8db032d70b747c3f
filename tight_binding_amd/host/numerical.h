// Host numerics of the KPM front end (numerical.h:120-144, 304-316,
// 334-361, 383-447 of the reference): damping kernels, Chebyshev expansion
// coefficients of the resolvent, random-phase kets, block traces, and the
// stored / streamed Chebyshev sequences (formed on the device).
#pragma once

#include <cmath>
#include <complex>
#include <algorithm>
#include <cstdint>
#include <memory>
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

namespace detail {

// The C-ABI description of a compressed sparse matrix (CSC, values fp64 or
// complex128) of the host layer.
template <typename Sparse>
tbk_sparse describe(const Sparse& A) {
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
    return s;
}

[[noreturn]] inline void fail(const char* where) {
    throw std::runtime_error(std::string(where) + ": " + tbk_last_error());
}

// Owner of a device context, a matrix and blocks of the C ABI.
struct DeviceScope {
    tbk_context* ctx = nullptr;
    tbk_matrix* m = nullptr;
    std::vector<tbk_block*> blocks;
    DeviceScope() = default;
    DeviceScope(const DeviceScope&) = delete;
    DeviceScope& operator=(const DeviceScope&) = delete;
    ~DeviceScope() {
        for (tbk_block* b : blocks) tbk_block_destroy(b);
        if (m) tbk_matrix_destroy(m);
        if (ctx) tbk_context_destroy(ctx);
    }
    template <typename Sparse>
    void open(const Sparse& x, const char* where) {
        if (tbk_context_create(0, &ctx) != TBK_OK) fail(where);
        const tbk_sparse s = describe(x);
        if (tbk_matrix_create(ctx, &s, &m) != TBK_OK) fail(where);
    }
    // a rows x R complex128 block from column-major data (null: zeros)
    tbk_block* block(const Complex* colmajor, int64_t rows, int64_t R, const char* where) {
        tbk_block* b = nullptr;
        if (tbk_block_create(ctx, reinterpret_cast<const double*>(colmajor), static_cast<int32_t>(rows),
                             static_cast<int32_t>(R), TBK_FP64, &b) != TBK_OK)
            fail(where);
        blocks.push_back(b);
        return b;
    }
    void release(tbk_block* b) {
        blocks.erase(std::remove(blocks.begin(), blocks.end(), b), blocks.end());
        tbk_block_destroy(b);
    }
};

}  // namespace detail

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
    const tbk_sparse s = detail::describe(A);
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
    const tbk_sparse s = detail::describe(x);
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

// Chebyshev::Cheb_first_kind_seq_view (numerical.h:346-361): next()
// returns {i, T_i(x) v} for i = 0, 1, 2, ..., the recursion running two
// terms ahead as the reference's v_0 / v_1 rotation does.  x is a sparse
// matrix (already rescaled), v a rows x R block.  The recursion state lives
// on device 0 (a context of its own, two blocks): each next() downloads T_i
// and forms T_{i+2} in its place with one step (tbk_cheb_step), so memory
// stays at two blocks however far the sequence runs.  x is copied to the
// device at construction; the view is movable, not copyable.
template <typename Mat, typename Vec = MatrixXcd>
struct Cheb_first_kind_seq_view {
    static_assert(std::is_same_v<Vec, MatrixXcd>, "the device view runs on complex blocks (MatrixXcd)");
    int i;
    const Mat& x;

    Cheb_first_kind_seq_view(const Mat& x, const Vec& v)
        : i(0), x(x), dev_(std::make_unique<detail::DeviceScope>()), rows_(v.rows()), cols_(v.cols()) {
        if (x.rows() != x.cols() || v.rows() != x.rows())
            throw std::invalid_argument("Cheb_first_kind_seq_view: x must be square with v's row count");
        dev_->open(x, "Cheb_first_kind_seq_view");
        v0_ = dev_->block(v.data(), rows_, cols_, "Cheb_first_kind_seq_view");
        v1_ = dev_->block(nullptr, rows_, cols_, "Cheb_first_kind_seq_view");
        if (tbk_cheb_step(dev_->ctx, dev_->m, 1.0, 0.0, v0_, nullptr, v1_) != TBK_OK)
            detail::fail("Cheb_first_kind_seq_view");
    }

    auto next() -> std::pair<int, Vec> {
        Vec t(rows_, cols_);
        if (tbk_block_download(v0_, reinterpret_cast<double*>(t.data())) != TBK_OK)
            detail::fail("Cheb_first_kind_seq_view::next");
        // T_{i+2} = 2 x T_{i+1} - T_i over T_i, then the roles swap
        if (tbk_cheb_step(dev_->ctx, dev_->m, 1.0, 0.0, v1_, v0_, v0_) != TBK_OK)
            detail::fail("Cheb_first_kind_seq_view::next");
        std::swap(v0_, v1_);
        return {i++, std::move(t)};
    }

private:
    std::unique_ptr<detail::DeviceScope> dev_;
    tbk_block* v0_ = nullptr;  // T_i
    tbk_block* v1_ = nullptr;  // T_{i+1}
    int64_t rows_, cols_;
};

// sum_{k < c.size()} c_k T_k(x) for a sparse x (already rescaled) as an
// n x n column-major matrix: the expansion applied to the identity on
// device 0, 256 columns at a time (tbk_cheb_expand), each T_k formed and
// added in order of k and never stored.
template <typename Sparse>
MatrixXcd Cheb_expand_on_identity(const Sparse& x, const std::vector<Complex>& c) {
    const char* where = "Cheb_expand_on_identity";
    if (x.rows() != x.cols()) throw std::invalid_argument("Cheb_expand_on_identity: x must be square");
    if (c.empty()) throw std::invalid_argument("Cheb_expand_on_identity: no coefficients");
    const int64_t n = x.rows();
    MatrixXcd out(n, n);
    out.setZero();
    detail::DeviceScope dev;
    dev.open(x, where);
    constexpr int64_t kCols = 256;
    for (int64_t c0 = 0; c0 < n; c0 += kCols) {
        const int64_t w = std::min(kCols, n - c0);
        std::vector<Complex> e(static_cast<size_t>(n * w), Complex(0.0, 0.0));
        for (int64_t j = 0; j < w; ++j) e[static_cast<size_t>(j * n + c0 + j)] = 1.0;
        tbk_block* v = dev.block(e.data(), n, w, where);
        tbk_block* o = dev.block(nullptr, n, w, where);
        if (tbk_cheb_expand(dev.ctx, dev.m, 1.0, 0.0, v, static_cast<int32_t>(c.size()),
                            reinterpret_cast<const double*>(c.data()), o, nullptr) != TBK_OK ||
            tbk_block_download(o, reinterpret_cast<double*>(out.data() + c0 * n)) != TBK_OK)
            detail::fail(where);
        dev.release(v);
        dev.release(o);
    }
    return out;
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
