// The reference's named entry test_conductivity_Kubo_xy (snippets.h:743-778,
// called from main.cpp:43), built on this package's headers:
//   graphene (Honeycomb + 1_NN, t = 1) with zero-valued complex second
//   neighbours (t = 0: +-i t on A-A and B-B bonds), 128 x 128 cells,
//   symmetric gauge A = Bz (-y, x, 0)/2 with Bz = 300/205102 * 2 pi,
//   DoS from density_of_states_seq<Complex>(X, 1024, E, -6, 6) at
//   utils::linspace(-1, 1, 2000) (endpoint excluded, utils.h:15-21),
//   then the open Hamiltonian get_Hamiltonian_sparse<Complex>().
// The plot becomes printed lines.  The reference's disorder line is
// commented out there, so there is none here; kets come from the seeded
// generator (the reference draws std::rand).  Its Kubo call is commented out
// too; sigma_xy runs through tbk_kubo_gamma instead.
//
// usage: kubo_xy_entry [L] [N] [n_E]
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "green.h"
#include "models.h"

using namespace tb;
using namespace std::complex_literals;

int main(int argc, char** argv) {
    const int L = argc > 1 ? std::atoi(argv[1]) : 128;
    const int N = argc > 2 ? std::atoi(argv[2]) : 1024;
    const int n_E = argc > 3 ? std::atoi(argv[3]) : 2000;
    auto d = models::graphene();
    const double t = 0.0;
    d.add_hopping("A", "A", LatticeVector<2>{0, 1}, 1.0i * t);
    d.add_hopping("A", "A", LatticeVector<2>{1, -1}, 1.0i * t);
    d.add_hopping("A", "A", LatticeVector<2>{-1, 0}, 1.0i * t);
    d.add_hopping("B", "B", LatticeVector<2>{0, 1}, -1.0i * t);
    d.add_hopping("B", "B", LatticeVector<2>{1, -1}, -1.0i * t);
    d.add_hopping("B", "B", LatticeVector<2>{-1, 0}, -1.0i * t);
    PartialLattice<2> X{d, LatticeVector<2>{L, L}};
    const double Bz = 300. / 205102. * 2 * numerical::pi;
    X.gauge_field = [Bz](Vector3d r) { return Bz * Vector3d{-r(1), r(0), 0} / 2; };

    std::vector<double> E(static_cast<size_t>(n_E));
    const double step = (1.0 - -1.0) / n_E;
    for (int i = 0; i < n_E; ++i) E[static_cast<size_t>(i)] = -1.0 + i * step;
    numerical::linalg::seed_random_kets(0x5EED);
    auto DoS = density_of_states_seq<Complex>(X, N, E, -6, 6);
    for (size_t i = 0; i < E.size(); ++i) std::printf("dos %.17g %.17g\n", E[i], DoS[i]);

    auto H = X.get_Hamiltonian_sparse<Complex>();
    std::printf("open %d %lld\n", H.rows(), static_cast<long long>(H.nonZeros()));
    return 0;
}
