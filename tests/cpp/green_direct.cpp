// A caller of the small-system Green function and the streamed Chebyshev
// sequence (lattice.h:570-592, numerical.h:346-361), built on this package's
// headers: GreenFunction(N, H) on a small square lattice with a Landau gauge
// (complex H) and on an ungauged one (real H), direct_eval at two energies,
// and Cheb_first_kind_seq_view over a seeded block.  Prints the scaling, the
// entries of direct_eval and the view's blocks for the test to compare.
//
// usage: green_direct <L> <N>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "green.h"
#include "models.h"

using namespace tb;

template <typename Scalar>
void dump(const char* tag, const SparseMatrix<Scalar>& G) {
    for (int32_t j = 0; j < G.cols(); ++j)
        for (int32_t p = G.outerIndexPtr()[j]; p < G.outerIndexPtr()[j + 1]; ++p) {
            const Complex v = G.valuePtr()[p];
            std::printf("%s %d %d %.17g %.17g\n", tag, G.innerIndexPtr()[p], j, v.real(), v.imag());
        }
}

int main(int argc, char** argv) {
    const int L = argc > 1 ? std::atoi(argv[1]) : 6;
    const int N = argc > 2 ? std::atoi(argv[2]) : 96;
    UnitCell<2> d = models::Square();
    d.add_hopping("A", 1_NN, 1.0);
    PartialLattice<2> X{d, LatticeVector<2>{L, L}};
    X.set_scalar_field_indexed(uniform_disorder(1.0, 5));
    auto Hr = X.get_Hamiltonian_periodic_extend_bare_sparse<double>(LatticeVector<2>{1, 1});
    X.gauge_field = [](Vector3d r) { return Vector3d{0, 2 * numerical::pi / 6 * r(0), 0}; };  // Landau gauge
    auto Hc = X.get_Hamiltonian_sparse<Complex>();

    // complex H: bounds from its own spectrum (epsilon 0.01), direct_eval
    PartialLattice<2>::GreenFunction<Complex> G(N, Hc);
    std::printf("scale_c %.17g %.17g\n", G.a, G.b);
    dump("gc", G.direct_eval(0.3, 0.3));
    // real H
    PartialLattice<2>::GreenFunction<double> Gr(N, Hr);
    std::printf("scale_r %.17g %.17g\n", Gr.a, Gr.b);
    dump("gr", Gr.direct_eval(Complex(-1.1, 0.4)));
    // outside the small-system constructor: a warning and zero
    PartialLattice<2>::GreenFunction<double> Gb(N, Hr, -5.0, 5.0);
    std::printf("big_nnz %lld\n", static_cast<long long>(Gb.direct_eval(0.0).nonZeros()));

    // the streamed sequence over a seeded block: T_0 .. T_5, column 1
    numerical::linalg::seed_random_kets(0x77);
    MatrixXcd v = numerical::linalg::bases_of_uniform_ampl_with_random_phases(X.get_total_number(), 3);
    numerical::Chebyshev::Cheb_first_kind_seq_view view(G.scaled_HE, v);
    for (int k = 0; k < 6; ++k) {
        auto [i, t] = view.next();
        for (int64_t r = 0; r < t.rows(); ++r)
            std::printf("view %d %lld %.17g %.17g\n", i, static_cast<long long>(r), t(r, 1).real(), t(r, 1).imag());
    }
    return 0;
}
