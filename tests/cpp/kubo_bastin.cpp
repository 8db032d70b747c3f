// A caller of conductivity_Kubo_Bastin_test (snippets.h:306-456), the
// reference's Kubo-Bastin driver, built on this package's headers: a square
// lattice with a Landau gauge and site disorder, open boundaries, its
// position operators, then the driver's kept nodes, zeta and both integrals
// (the reference plots them; here they are printed).  Also prints
// trace_braket_complex of two seeded blocks and density_of_states_function
// at a few energies.
//
// usage: kubo_bastin <L> <N>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "kubo.h"
#include "models.h"

using namespace tb;

int main(int argc, char** argv) {
    const int L = argc > 1 ? std::atoi(argv[1]) : 12;
    const int N = argc > 2 ? std::atoi(argv[2]) : 64;
    auto d = models::Square();
    d.add_hopping("A", 1_NN, 1.0);
    PartialLattice<2> X{d, LatticeVector<2>{L, L}};
    X.set_scalar_field_indexed(uniform_disorder(1.0, 31));
    const double B = 2 * numerical::pi / 8;
    X.gauge_field = [B](Vector3d r) { return Vector3d{0.0, B * r(0), 0.0}; };  // Landau gauge
    auto H = X.get_Hamiltonian_sparse<Complex>();
    auto op_x = X.op_position(0);
    auto op_y = X.op_position(1);

    numerical::linalg::seed_random_kets(0x5EED);
    auto r = conductivity_Kubo_Bastin_test(H, op_x, op_y, N, -5.0, 5.0, 2.0);
    for (size_t i = 0; i < r.E.size(); ++i)
        std::printf("kubo %.17g %.17g %.17g %.17g %.17g\n", r.E[i], r.zeta_real[i], r.zeta_imag[i], r.zeta_re_int[i],
                    r.zeta_im_int[i]);

    numerical::linalg::seed_random_kets(77);
    MatrixXcd a = numerical::linalg::bases_of_uniform_ampl_with_random_phases(H.rows(), 3);
    MatrixXcd b = numerical::linalg::bases_of_uniform_ampl_with_random_phases(H.rows(), 3);
    const Complex t = numerical::linalg::trace_braket_complex(a, b);
    std::printf("trace %.17g %.17g\n", t.real(), t.imag());

    // density_of_states_function on the periodic H (draws the next ket seed, 79)
    auto Hp = X.get_Hamiltonian_periodic_extend_bare_sparse<Complex>(LatticeVector<2>{1, 1});
    PartialLattice<2>::GreenFunction<Complex> G(N, Hp, -5.0, 5.0);
    auto dos = density_of_states_function(G);
    for (double E : {-3.0, -1.0, 0.5, 2.0}) std::printf("dosf %.17g %.17g\n", E, dos(E));
    return 0;
}
