// KPM front end of PartialLattice: GreenFunction<Scalar> and its
// GreenFunctionImag functor (lattice.h:532-652 of the reference), and
// density_of_states_seq (functions.h:15-33).  The small-system constructor
// GreenFunction(N, H) and direct_eval (lattice.h:570-592) run the Chebyshev
// sum on the device.
//
// GreenFunction(N, H, E_min, E_max) rescales H~ = (H - b I)/a with
// a = (E_max - E_min)/2, b = (E_max + E_min)/2 over the union pattern of H
// and I (lattice.h:553-559).  GreenFunctionImag computes the N Chebyshev
// moments with the HIP engine (cuda::Cheb_first_kind_seq_braket, the path
// the reference takes under USE_CUDA, lattice.h:608-612) and the Lorentz
// kernel with lambda = 4 (lattice.h:613); operator()(e) evaluates
// sum g_n Im(alpha_n(e)) mu_n (lattice.h:618-624).
#pragma once

#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "lattice.h"
#include "numerical.h"
#include "numerics_cuda.cuh"

namespace tb {

template <int Dim>
template <typename Scalar>
struct PartialLattice<Dim>::GreenFunction {
    using Mat = SparseMatrix<Scalar>;
    using Kets = MatrixXcd;

    int N;
    double a, b;
    double E_min, E_max;
    Mat scaled_HE;
    bool direct_ok = false;  // direct_eval available (small-system constructor)

    GreenFunction(int N, const Mat& HE, double E_min, double E_max) : N(N), E_min(E_min), E_max(E_max) {
        a = (E_max - E_min) / 2;
        b = (E_max + E_min) / 2;
        scaled_HE = rescale(HE, a, b);
    }
    // bounds from the spectrum of ref_HE, contracted by epsilon
    // (lattice.h:560-569): a = (E_max - E_min)/(2 - epsilon)
    GreenFunction(int N, const Mat& HE, const Mat& ref_HE, double epsilon = 0.01) : N(N) {
        std::tie(E_min, E_max) = numerical::linalg::matrix_spectral_range_sparse(ref_HE);
        a = (E_max - E_min) / (2 - epsilon);
        b = (E_max + E_min) / 2;
        scaled_HE = rescale(HE, a, b);
    }
    // for small systems (lattice.h:570-577): the range of HE's own spectrum,
    // and direct_eval below 5000 rows.  The reference stores T_n(H~) for
    // n < N here; direct_eval below forms its sum on the device instead, so
    // nothing is stored.
    GreenFunction(int N, const Mat& HE) : GreenFunction(N, HE, HE) {
        direct_ok = HE.rows() < 5000;
        if (!direct_ok && PartialLattice<Dim>::auto_info)
            std::cout << "GreenFunctionKP refuses to generate Chebyt seq of H (H is too big)." << std::endl;
    }
    // sum_{n<N} Im(alpha_n(z~)) T_n(H~) / a at z~ = (z - b)/a (lattice.h:579-587):
    // the expansion applied to the identity on the device (tbk_cheb_expand),
    // the terms added in order of n as the reference's loop adds them.  The
    // result holds the entries that are not zero.  Outside the small-system
    // constructor (or at 5000 rows and above) it warns and returns zero, as
    // the reference's loop over an empty sequence does.  sgn is unused, as
    // in the reference.
    Mat direct_eval(Complex z, int sgn = 1) const {
        (void)sgn;
        Mat Gf(scaled_HE.rows(), scaled_HE.cols());
        if (!direct_ok) {
            std::cerr << "Warning: Too big to implement direct eval." << std::endl;
            return Gf;
        }
        z = scale(z, Complex(1.0, 0.0));
        const auto alpha = numerical::Chebyshev::expn_coef_alpha_seq(z, N);
        std::vector<Complex> c(alpha.size());
        for (size_t n = 0; n < alpha.size(); ++n) c[n] = Complex(alpha[n].imag(), 0.0);
        const MatrixXcd sum = numerical::Chebyshev::Cheb_expand_on_identity(scaled_HE, c);
        std::vector<Triplet<Scalar>> t;
        for (int32_t j = 0; j < Gf.cols(); ++j)
            for (int32_t i = 0; i < Gf.rows(); ++i) {
                Scalar v;
                if constexpr (is_complex_v<Scalar>) v = sum(i, j) / a;
                else v = sum(i, j).real() / a;
                if (v != Scalar(0)) t.emplace_back(i, j, v);
            }
        Gf.setFromTriplets(t.begin(), t.end());
        return Gf;
    }
    // z = omega + i eta (lattice.h:589-592)
    Mat direct_eval(double omega, double eta = 0.01) const {
        const int sgn = eta > 0 ? 1 : -1;
        return direct_eval(Complex(omega, eta), sgn);
    }
    double scale(double x) const { return (x - b) / a; }
    // (x - b id) / a, the reference's two-argument form (lattice.h:550)
    Complex scale(Complex x, Complex id) const { return (x - b * id) / a; }
    std::pair<double, double> spectral_range() const { return {E_min, E_max}; }

    struct GreenFunctionImag {
        GreenFunction* G;
        std::vector<double> moments;
        std::vector<double> g;  // kernel

        GreenFunctionImag(GreenFunction* G, const Kets& bras_T, const Kets& kets) : G(G) {
            moments = cuda::Cheb_first_kind_seq_braket(G->scaled_HE, G->N, bras_T, kets);
            g = numerical::kernels::kernel_seq(G->N, numerical::kernels::Lorentz(G->N, 4));
        }
        static double weight(double x) { return 1. / std::sqrt(1. - x * x); }
        // sum_n g_n Im(alpha_n(e)) mu_n at a rescaled energy e
        double operator()(Complex e) const {
            const auto alpha = numerical::Chebyshev::expn_coef_alpha_seq(e, G->N);
            double sum = 0;
            for (int n = 0; n < G->N && n < static_cast<int>(moments.size()); ++n)
                sum += g[n] * alpha[n].imag() * moments[n];
            return sum;
        }
        // the expansion at the n N Chebyshev nodes E_k = cos((k + 1/2) pi / (n N)):
        // REDFT01 of g_n mu_n times weight(E_k) (lattice.h:625-640), evaluated on
        // the device at theta_k directly (tbk_cheb_reconstruct, node mode).  For
        // n > 1 the moments are zero-padded (the reference reads past its N
        // moments there).  Nothing is printed per node.
        std::pair<std::vector<double>, std::vector<double>> fft(int n = 1) const {
            if (n < 1) throw std::invalid_argument("fft: n must be >= 1");
            const int N0 = static_cast<int>(moments.size()), Nn = n * N0;
            std::vector<double> E(static_cast<size_t>(Nn)), r(static_cast<size_t>(Nn));
            for (int k = 0; k < Nn; ++k) E[k] = std::cos((k + 0.5) * numerical::pi / Nn);
            if (tbk_cheb_reconstruct(nullptr, moments.data(), g.data(), N0, nullptr, Nn, 1.0, r.data(), nullptr) !=
                TBK_OK)
                throw std::runtime_error(std::string("fft: ") + tbk_last_error());
            return {E, r};
        }
    };
    GreenFunctionImag imag(const Kets& bras_T, const Kets& kets) { return GreenFunctionImag(this, bras_T, kets); }
};

// A functor E -> DoS(E) of a GreenFunction (functions.h:5-13): 4 random-phase
// vectors, f = G.imag(kets, kets) (moments on the device), and
// DoS(E) = -f(omega)/(4 pi) at omega = G.scale(E, i) = (E - i b)/a, the
// reference's argument (its imaginary part is -b/a; zero for symmetric
// bounds).  The functor refers to G, which must outlive it.
template <typename GF>
auto density_of_states_function(GF& G) {
    MatrixXcd kets = numerical::linalg::bases_of_uniform_ampl_with_random_phases(G.scaled_HE.rows(), 4);
    auto f = G.imag(kets, kets);
    return [f, &G](double E) -> double {
        const Complex omega_t = G.scale(Complex(E, 0.0), Complex(0.0, 1.0));
        return -1.0 / 4 / numerical::pi * f(omega_t);
    };
}

// DoS(e) = -f(e)/(R pi) at rescaled energies e (functions.h:16-33): periodic
// H (extend 1 along every axis), bounds padded by 1e-8, R = 4 random-phase
// vectors.
template <typename Scalar, int Dim>
std::vector<double> density_of_states_seq(PartialLattice<Dim>& X, int N, const std::vector<double>& E_list,
                                          double E_min, double E_max) {
    LatticeVector<Dim> ext;
    for (int d = 0; d < Dim; ++d) ext(d) = 1;
    auto H = X.template get_Hamiltonian_periodic_extend_bare_sparse<Scalar>(ext);
    typename PartialLattice<Dim>::template GreenFunction<Scalar> G(N, H, E_min - 1e-8, E_max + 1e-8);
    const int R = 4;
    MatrixXcd kets = numerical::linalg::bases_of_uniform_ampl_with_random_phases(X.get_total_number(), R);
    auto f = G.imag(kets, kets);
    std::vector<double> dos;
    dos.reserve(E_list.size());
    for (double e : E_list) dos.push_back(-1.0 / R / numerical::pi * f(Complex(e, 0.0)));
    return dos;
}

}  // namespace tb
