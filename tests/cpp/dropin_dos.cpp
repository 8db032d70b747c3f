// A caller written against the reference's API (models, UnitCell,
// PartialLattice, GreenFunction, density_of_states_seq -- the flow of
// snippets.h:743-771 / functions.h:16-33), built on this package's headers.
// Prints the N moments of GreenFunctionImag, the expansion at the Chebyshev
// nodes, stored blocks T_k(H~) kets, DoS at a few energies and the spectral
// range.
//
// usage: dropin_dos <L> <N>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "green.h"
#include "models.h"

using namespace tb;

int main(int argc, char** argv) {
    const int L = argc > 1 ? std::atoi(argv[1]) : 16;
    const int N = argc > 2 ? std::atoi(argv[2]) : 64;
    auto d = models::graphene();  // Honeycomb + 1_NN t = 1
    PartialLattice<2> X{d, LatticeVector<2>{L, L}};
    X.set_scalar_field_indexed(uniform_disorder(0.2, 99));
    X.gauge_field = [](Vector3d r) { return 0.01 * Vector3d{-r(1), r(0), 0} / 2; };  // symmetric gauge

    // moments exactly as GreenFunctionImag computes them
    auto H = X.get_Hamiltonian_periodic_extend_bare_sparse<Complex>(LatticeVector<2>{1, 1});
    PartialLattice<2>::GreenFunction<Complex> G(N, H, -3.5 - 1e-8, 3.5 + 1e-8);
    numerical::linalg::seed_random_kets(0x5EED);
    MatrixXcd kets = numerical::linalg::bases_of_uniform_ampl_with_random_phases(X.get_total_number(), 4);
    auto f = G.imag(kets, kets);
    for (double m : f.moments) std::printf("mu %.17g\n", m);
    auto [En, rn] = f.fft();  // GreenFunctionImag::fft(): the expansion at the N Chebyshev nodes
    for (size_t k = 0; k < En.size(); ++k) std::printf("fft %.17g %.17g\n", En[k], rn[k]);

    // the DoS driver (draws the next ket seed)
    std::vector<double> E = {-0.5, -0.1, 0.0, 0.3, 0.7};
    auto dos = density_of_states_seq<Complex>(X, N, E, -3.5, 3.5);
    for (size_t i = 0; i < E.size(); ++i) std::printf("dos %.17g %.17g\n", E[i], dos[i]);

    // bounds from the spectrum of H itself, contracted by epsilon = 0.01
    // (GreenFunction(N, HE, ref_HE, epsilon), lattice.h:560-569)
    PartialLattice<2>::GreenFunction<Complex> G3(N, H, H);
    std::printf("range %.17g %.17g %.17g %.17g\n", G3.E_min, G3.E_max, G3.a, G3.b);

    // stored blocks T_k(scaled_HE) kets, k < 6 (Cheb_first_kind_seq_on_bases):
    // column 1 of each
    auto Tn = numerical::Chebyshev::Cheb_first_kind_seq_on_bases(G.scaled_HE, 6, kets);
    for (size_t k = 0; k < Tn.size(); ++k)
        for (int64_t i = 0; i < Tn[k].rows(); ++i)
            std::printf("tn %zu %lld %.17g %.17g\n", k, static_cast<long long>(i), Tn[k](i, 1).real(),
                        Tn[k](i, 1).imag());

    // n < 2 keeps the reference's behaviour: message and an empty result
    auto none = cuda::Cheb_first_kind_seq_braket(G.scaled_HE, 1, kets, kets);
    std::printf("empty %zu\n", none.size());
    // bras_T must have the kets' shape
    try {
        MatrixXcd narrow = numerical::linalg::bases_of_uniform_ampl_with_random_phases(X.get_total_number(), 2);
        cuda::Cheb_first_kind_seq_braket(G.scaled_HE, N, narrow, kets);
        std::printf("shape_error none\n");
    } catch (const std::invalid_argument&) {
        std::printf("shape_error thrown\n");
    }
    return 0;
}
