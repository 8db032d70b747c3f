// conductivity_Kubo_Bastin_test (snippets.h:306-456 of the reference) on the
// HIP engine, returning what the reference plots.
//
// The reference builds GreenFunction<Complex>(N, H_og, E_min - 1e-8,
// E_max + 1e-8), v_x = X H~ - H~ X and v_y = Y H~ - H~ Y from the position
// operators, R = 4 random-phase kets and the Lorentz kernel (lambda = 6),
// forms both Chebyshev stacks, zeta at the N Chebyshev nodes
// e_k = cos((k + 1/2) pi / N) through DCT-III / FFT sums divided by
// (1 - e^2)^2 and Volume, drops N/24 nodes at each end, and integrates Re
// and Im zeta outward from the node pair around e = 0 (the two nodes next to
// zero seeded by a 3/4-1/4 trapezoid for the real part only), then plots
// both integrals.
//
// Here Gamma_nm = <aL_n|aR_m> comes from tbk_kubo_gamma (FP64) and zeta from
// tbk_kubo_zeta, the same sums in the double-sum form the reference states
// beside its FFT route (snippets.h:337-349); the kept nodes and the
// integrals follow the reference line for line.  Kets are drawn from the
// seeded generator of numerical.h (bases_of_uniform_ampl_with_random_phases).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

#include "green.h"
#include "numerical.h"
#include "tbkpm.h"

namespace tb {

struct KuboBastinResult {
    std::vector<double> E;            // rescaled Fermi energies of the kept nodes, ascending
    std::vector<double> zeta_real;    // Re zeta on E
    std::vector<double> zeta_imag;    // Im zeta on E
    std::vector<double> zeta_re_int;  // integral of Re zeta from e = 0 (sigma_xy up to the unit factor)
    std::vector<double> zeta_im_int;  // integral of Im zeta, as the reference forms it
};

namespace detail {

template <typename Scalar>
tbk_sparse csc_view(const SparseMatrix<Scalar>& A) {
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

inline void check(int rc, const char* what) {
    if (rc != TBK_OK) throw std::runtime_error(std::string("conductivity_Kubo_Bastin_test: ") + what + ": " +
                                               tbk_last_error());
}

// device objects of one call, released on every exit
struct KuboDevice {
    tbk_context* ctx = nullptr;
    tbk_matrix *H = nullptr, *vx = nullptr, *vy = nullptr;
    tbk_block* kets = nullptr;
    ~KuboDevice() {
        if (kets) tbk_block_destroy(kets);
        for (tbk_matrix* m : {H, vx, vy})
            if (m) tbk_matrix_destroy(m);
        if (ctx) tbk_context_destroy(ctx);
    }
};

}  // namespace detail

inline KuboBastinResult conductivity_Kubo_Bastin_test(const SparseMatrix<Complex>& H_og,
                                                      const SparseMatrix<double>& op_x,
                                                      const SparseMatrix<double>& op_y, int N, double E_min,
                                                      double E_max, double Volume = 1.0) {
    if (N < 2) throw std::invalid_argument("conductivity_Kubo_Bastin_test: N must be at least 2");
    typename PartialLattice<2>::template GreenFunction<Complex> G(N, H_og, E_min - 1e-8, E_max + 1e-8);
    const auto& H = G.scaled_HE;
    const auto v_x = PartialLattice<2>::commutator_with_diagonal(op_x, H);
    const auto v_y = PartialLattice<2>::commutator_with_diagonal(op_y, H);
    const int R = 4;
    const auto g = numerical::kernels::kernel_seq(N, numerical::kernels::Lorentz(N, 6.));
    const MatrixXcd kets = numerical::linalg::bases_of_uniform_ampl_with_random_phases(H_og.rows(), R);

    detail::KuboDevice d;
    detail::check(tbk_context_create(0, &d.ctx), "context");
    const tbk_sparse sH = detail::csc_view(H), sx = detail::csc_view(v_x), sy = detail::csc_view(v_y);
    detail::check(tbk_matrix_create(d.ctx, &sH, &d.H), "H~");
    detail::check(tbk_matrix_create(d.ctx, &sx, &d.vx), "v_x");
    detail::check(tbk_matrix_create(d.ctx, &sy, &d.vy), "v_y");
    detail::check(tbk_block_create(d.ctx, reinterpret_cast<const double*>(kets.data()), kets.rows(), R, TBK_FP64,
                                   &d.kets),
                  "kets");
    std::vector<double> gamma(static_cast<size_t>(2) * N * N);
    detail::check(tbk_kubo_gamma(d.ctx, d.H, 1.0, 0.0, d.vx, d.vy, d.kets, N, TBK_GAMMA_FP64, g.data(),
                                 gamma.data(), nullptr),
                  "Gamma");
    std::vector<double> epsilon(static_cast<size_t>(N));
    for (int k = 0; k < N; ++k) epsilon[k] = std::cos((k + 0.5) / N * numerical::pi);
    std::vector<double> zeta(static_cast<size_t>(2) * N);
    detail::check(tbk_kubo_zeta(d.ctx, gamma.data(), N, epsilon.data(), N, Volume, zeta.data(), nullptr), "zeta");

    // the kept nodes, ascending (snippets.h:420-425)
    const int cut = N / 24;
    KuboBastinResult r;
    for (int k = N - cut - 1; k >= cut; --k) {
        r.E.push_back(epsilon[k]);
        r.zeta_real.push_back(zeta[2 * k]);
        r.zeta_imag.push_back(zeta[2 * k + 1]);
    }
    // integrals outward from e = 0 (snippets.h:436-449)
    const auto& E = r.E;
    const size_t n = E.size();
    const size_t E0_i = static_cast<size_t>(std::upper_bound(E.begin(), E.end(), 0.0) - E.begin());
    if (E0_i == 0 || E0_i >= n)
        throw std::invalid_argument("conductivity_Kubo_Bastin_test: the kept nodes do not straddle e = 0");
    r.zeta_re_int.assign(n, 0.0);
    r.zeta_im_int.assign(n, 0.0);
    const auto& zr = r.zeta_real;
    const auto& zi = r.zeta_imag;
    r.zeta_re_int[E0_i] = (zr[E0_i] * 3 / 4 + zr[E0_i - 1] / 4) * (E[E0_i] - E[E0_i - 1]) / 2;
    r.zeta_re_int[E0_i - 1] = -(zr[E0_i] / 4 + zr[E0_i - 1] * 3 / 4) * (E[E0_i] - E[E0_i - 1]) / 2;
    for (size_t i = E0_i + 1; i < n; ++i) {
        r.zeta_re_int[i] = r.zeta_re_int[i - 1] + (zr[i] + zr[i - 1]) / 2 * (E[i] - E[i - 1]);
        r.zeta_im_int[i] = r.zeta_im_int[i - 1] + (zi[i] + zi[i - 1]) / 2 * (E[i] - E[i - 1]);
    }
    for (size_t j = E0_i - 1; j-- > 0;) {
        r.zeta_re_int[j] = r.zeta_re_int[j + 1] - (zr[j] + zr[j + 1]) / 2 * (E[j + 1] - E[j]);
        r.zeta_im_int[j] = r.zeta_im_int[j + 1] - (zi[j] + zi[j + 1]) / 2 * (E[j + 1] - E[j]);
    }
    return r;
}

}  // namespace tb
