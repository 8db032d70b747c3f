// PartialLattice<Dim>: a finite block of unit cells and the sparse
// Hamiltonian builders on the KPM path, restated from the reference
// (lattice.h:22-117, 216-305, 314-345, 389-505).
//
// Site numbering (lattice.h:28, 106-117): bare index = i_v . span_s + atom,
// last lattice index fastest, atoms innermost.  Cells are visited in that
// order and, inside a cell, hoppings in table order (lattice.h:466-493).
//
// Value pipeline per (cell, hopping) (lattice.h:274-298):
//   t  <- t * exp(+i A(r_from).d)          if a gauge field is set
//        (A and d are real vectors: `1.0i * A.dot(d)` in lattice.h:284 is
//        i times the real scalar A.d)
//   t  <- t + V(r_from)                     if row == col and a scalar field
//                                           (or a site-keyed field) is set
//   push (row, col, t) and, if row != col, (col, row, conj t); a real Scalar
//   keeps Re t.  setFromTriplets then sums duplicates in insertion order.
// The matrix is total_number x total_number even when sites are removed.
//
// Addition on top of the reference: set_scalar_field_indexed(f) takes a
// function of the bare site index, so that disorder can be generated by a
// counter-based generator keyed by site (identical in every build and
// independent of iteration order); uniform_disorder(W, seed) is that field.
#pragma once

#include <cmath>
#include <cstdint>
#include <functional>
#include <iostream>
#include <numeric>
#include <random>
#include <string>
#include <unordered_set>
#include <vector>

#include "linalg.h"
#include "unit_cell.h"

namespace tb {

// splitmix64 finaliser: the counter-based generator behind uniform_disorder
// and the seeded random-phase kets.
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
// uniform double in [0, 1) keyed by (seed, counter); exact in binary64
inline double counter_uniform(uint64_t seed, uint64_t counter) {
    uint64_t h = splitmix64(splitmix64(seed) ^ counter);
    return static_cast<double>(h >> 11) * 0x1.0p-53;
}
// on-site energies uniform in [-W/2, W/2), keyed by bare site index
inline std::function<double(int)> uniform_disorder(double W, uint64_t seed) {
    return [W, seed](int site) { return W * (counter_uniform(seed, static_cast<uint64_t>(site)) - 0.5); };
}

// exp(+i theta) as (cos theta, sin theta) from two separate libm calls.
// std::exp of an imaginary argument goes through sincos, whose rounding
// differs from cos/sin in near-tie cases (theta = -2.4183991523122903 is
// one); the separate, non-inlined calls keep the compiler from fusing them
// into sincos, so every phase has the bits of libm's cos and sin.
__attribute__((noinline)) inline double phase_cos(double theta) { return std::cos(theta); }
__attribute__((noinline)) inline double phase_sin(double theta) { return std::sin(theta); }
inline Complex peierls_phase(double theta) { return Complex(phase_cos(theta), phase_sin(theta)); }

template <int Dim>
class PartialLattice {
public:
    static constexpr int dim = Dim;
    PartialLattice(const UnitCell<Dim>& cell, LatticeVector<Dim> span_size)
        : cell_(cell), span_v(span_size), cell_number(1) {
        for (int i = 0; i < Dim; ++i) {
            if (span_v(i) <= 0) throw std::invalid_argument("PartialLattice: span must be positive");
            cell_number *= span_v(i);
        }
        total_number = cell_number * cell.get_atom_number();
        span_s(Dim - 1) = cell_.get_atom_number();
        for (int i = Dim - 2; i >= 0; --i) span_s(i) = span_s(i + 1) * span_v(i + 1);
        if (auto_info)
            std::cout << "PartialLattice generated with total number: " << total_number << std::endl;
    }

    int index_bare(const LatticeVector<Dim>& i_v, int i_a) const { return i_v.dot(span_s) + i_a; }
    int index_comp(int i_b) const { return index_c[i_b]; }
    int index_comp(const LatticeVector<Dim>& i_v, int i_a) const { return index_c[index_bare(i_v, i_a)]; }
    bool index_bare_is_removed(int i_b) const { return index_c[i_b] == -1; }
    bool index_inside_lattice(const LatticeVector<Dim>& i) const {
        for (int d = 0; d < Dim; ++d)
            if (i(d) < 0 || i(d) >= span_v(d)) return false;
        return true;
    }
    bool index_inside_lattice_periodic_extend(const LatticeVector<Dim>& i,
                                              const LatticeVector<Dim>& extend) const {
        for (int d = 0; d < Dim; ++d)
            if (!(extend(d) != 0 || (i(d) >= 0 && i(d) < span_v(d)))) return false;
        return true;
    }

    int get_total_number() const { return total_number; }
    int get_cell_number() const { return cell_number; }
    const UnitCell<Dim>& cell() const { return cell_; }
    LatticeVector<Dim> span() const { return span_v; }

    // Open boundaries, compressed indices, gauge and scalar fields
    // (lattice.h:262-305).
    template <typename Scalar>
    SparseMatrix<Scalar> get_Hamiltonian_sparse() {
        if (index_c.empty()) index_comp_init();
        return assemble<Scalar>(Mode::Open, LatticeVector<Dim>::Zero());
    }
    // Periodic wrap along every axis with extend(d) != 0, bare indices,
    // removed sites ignored (lattice.h:216-258).
    template <typename Scalar>
    SparseMatrix<Scalar> get_Hamiltonian_periodic_extend_bare_sparse(LatticeVector<Dim> extend) {
        return assemble<Scalar>(Mode::Periodic, extend);
    }
    // Open boundaries, bare indices, no fields (lattice.h:347-377).
    template <typename Scalar>
    SparseMatrix<Scalar> get_Hamiltonian_bare_sparse() {
        return assemble<Scalar>(Mode::Bare, LatticeVector<Dim>::Zero());
    }

    // Diagonal position operator over compressed indices (lattice.h:314-336).
    SparseMatrix<double> op_position(int axis) {
        if (axis >= Dim) {
            std::cerr << "Error: axis out of range." << std::endl;
            axis = 0;
        }
        if (index_c.empty()) index_comp_init();
        std::vector<Triplet<double>> trip;
        trip.reserve(total_number);
        const int na = cell_.get_atom_number();
        for (int c = 0; c < cell_number; ++c) {
            LatticeVector<Dim> i = cell_vector(c);
            Vector3d L = cell_.displacement_of_lattice(i);
            for (int a = 0; a < na; ++a) {
                int row = index_comp(i, a);
                if (row == -1) continue;
                double pos = cell_.get_atom_pos(a)(axis) + L(axis);
                trip.emplace_back(row, row, pos);
            }
        }
        SparseMatrix<double> op(total_number, total_number);
        op.setFromTriplets(trip.begin(), trip.end());
        return op;
    }

    // [X, H] without the factor i (lattice.h:338-345): entries
    // x_i*H_ij - H_ij*x_j over the pattern of H (explicit zeros kept).
    template <typename Scalar = double>
    SparseMatrix<Scalar> op_velocity(int axis) {
        auto X = op_position(axis);
        auto H = get_Hamiltonian_sparse<Scalar>();
        return commutator_with_diagonal(X, H);
    }

    // C = X*M - M*X for diagonal X with an explicit diagonal
    template <typename Scalar>
    static SparseMatrix<Scalar> commutator_with_diagonal(const SparseMatrix<double>& X,
                                                         const SparseMatrix<Scalar>& M) {
        SparseMatrix<Scalar> C = M;
        auto xdiag = [&](int i) { return X.coeff(i, i); };
        const int32_t* op = M.outerIndexPtr();
        const int32_t* ip = M.innerIndexPtr();
        const Scalar* vp = M.valuePtr();
        Scalar* cp = C.valuePtr();
        for (int32_t j = 0; j < M.cols(); ++j) {
            double xj = xdiag(j);
            for (int32_t p = op[j]; p < op[j + 1]; ++p) {
                double xi = xdiag(ip[p]);
                cp[p] = mul_real_left(xi, vp[p]) - mul_real_right(vp[p], xj);
            }
        }
        return C;
    }

    void remove_atoms(const std::string& atom_name, const std::vector<LatticeVector<Dim>>& i) {
        if (!cell_.atom_exists(atom_name)) {
            std::cerr << "Atom " << atom_name << " does not exist in the unit cell." << std::endl;
            return;
        }
        int ai = cell_.get_atom_index(atom_name);
        for (const auto& v : i) removed_atoms.push_back(index_bare(v, ai));
        index_comp_construct();
    }
    // n distinct random cells lose atom `atom_name` (lattice.h:404-444);
    // seeded so that runs are reproducible.
    void remove_atoms_random(const std::string& atom_name, int n, uint64_t seed = 0) {
        if (!cell_.atom_exists(atom_name)) {
            std::cerr << "Atom " << atom_name << " does not exist in the unit cell." << std::endl;
            return;
        }
        if (n > cell_number) {
            std::cerr << "The random selection is larger than the total number." << std::endl;
            return;
        }
        int ai = cell_.get_atom_index(atom_name);
        std::vector<int> seq(cell_number);
        std::iota(seq.begin(), seq.end(), 0);
        std::mt19937_64 gen(seed);
        std::shuffle(seq.begin(), seq.end(), gen);
        for (int k = 0; k < n; ++k) removed_atoms.push_back(seq[k] * cell_.get_atom_number() + ai);
        index_comp_construct();
    }

    // scalar field V(r) of position (lattice.h:496-505): ensures every atom
    // has an on-site entry so the diagonal exists
    void set_scalar_field(std::function<double(Vector3d)> f) {
        scalar_field = std::move(f);
        ensure_onsite();
    }
    // scalar field V(site) of the bare site index
    void set_scalar_field_indexed(std::function<double(int)> f) {
        site_field = std::move(f);
        ensure_onsite();
    }

    std::function<Vector3d(Vector3d)> gauge_field;
    inline static bool auto_info = false;

    // exp(+i A(r_from) . d_h) for every (cell, hopping), cells in index order
    // and hoppings in table order: the factor the open and periodic builders
    // apply (lattice.h:282-285), computed by the same expressions.
    std::vector<Complex> phase_table() const {
        const auto& hops = cell_.hopping_view();
        std::vector<Complex> out;
        if (!gauge_field) return out;
        out.reserve(static_cast<size_t>(cell_number) * hops.size());
        std::vector<Vector3d> disp(hops.size());
        for (size_t h = 0; h < hops.size(); ++h) disp[h] = cell_.displacement_of_hopping(hops[h].first);
        for (int c = 0; c < cell_number; ++c) {
            const LatticeVector<Dim> i_v = cell_vector(c);
            for (size_t h = 0; h < hops.size(); ++h) {
                Vector3d A = gauge_field(cell_.atom_position(std::get<0>(hops[h].first), i_v));
                double theta = A.dot(disp[h]);
                out.push_back(peierls_phase(theta));
            }
        }
        return out;
    }
    bool has_position_field() const { return static_cast<bool>(scalar_field); }
    int removed_count() const { return static_cast<int>(removed_atoms.size()); }

    template <typename Scalar>
    struct GreenFunction;

private:
    enum class Mode { Open, Periodic, Bare };

    static Complex mul_real_left(double x, Complex v) { return Complex(x * v.real(), x * v.imag()); }
    static double mul_real_left(double x, double v) { return x * v; }
    static Complex mul_real_right(Complex v, double x) { return Complex(v.real() * x, v.imag() * x); }
    static double mul_real_right(double v, double x) { return v * x; }

    LatticeVector<Dim> cell_vector(int c) const {
        LatticeVector<Dim> i;
        for (int d = Dim - 1; d >= 0; --d) {
            i(d) = c % span_v(d);
            c /= span_v(d);
        }
        return i;
    }

    void ensure_onsite() {
        for (int i = 0; i < cell_.get_atom_number(); ++i) {
            Hopping<Dim> h{i, i, LatticeVector<Dim>::Zero()};
            if (cell_.hopping_exists(h)) continue;
            cell_.add_hopping_direct(h, 0.0);
        }
    }

    void index_comp_construct() {
        index_c.assign(total_number, 0);
        for (int i : removed_atoms) index_c[i] = -1;
        int count = 0;
        for (int i = 0; i < total_number; ++i)
            if (index_c[i] != -1) index_c[i] = count++;
    }
    void index_comp_init() {
        index_c.resize(total_number);
        std::iota(index_c.begin(), index_c.end(), 0);
    }

    template <typename Scalar>
    SparseMatrix<Scalar> assemble(Mode mode, LatticeVector<Dim> extend) {
        const auto& hops = cell_.hopping_view();
        if (hops.empty()) std::cerr << "Warning: no hopping is added!" << std::endl;
        const bool fields = (mode != Mode::Bare);
        // per-hopping constant displacement d (independent of the cell)
        std::vector<Vector3d> disp(hops.size());
        for (size_t h = 0; h < hops.size(); ++h) disp[h] = cell_.displacement_of_hopping(hops[h].first);

        std::vector<Triplet<Scalar>> trip;
        trip.reserve(static_cast<size_t>(cell_number) * hops.size() * 2);
        bool warned = false;
        for (int c = 0; c < cell_number; ++c) {
            const LatticeVector<Dim> i_v = cell_vector(c);
            for (size_t h = 0; h < hops.size(); ++h) {
                const auto& [hp, t0] = hops[h];
                const auto& [from, to, lat_v] = hp;
                LatticeVector<Dim> to_v = i_v + lat_v;
                int row, col;
                if (mode == Mode::Periodic) {
                    if (!(index_inside_lattice_periodic_extend(to_v, extend) || index_inside_lattice(to_v)))
                        continue;
                    LatticeVector<Dim> m;
                    for (int d = 0; d < Dim; ++d) {
                        int L = span_v(d);
                        int x = to_v(d) - L * (to_v(d) / L);
                        x = x + L;
                        m(d) = x - L * (x / L);
                    }
                    row = index_bare(i_v, from);
                    col = index_bare(m, to);
                } else if (mode == Mode::Open) {
                    if (!index_inside_lattice(to_v)) continue;
                    int rb = index_bare(i_v, from), cb = index_bare(to_v, to);
                    if (index_bare_is_removed(rb) || index_bare_is_removed(cb)) continue;
                    row = index_c[rb];
                    col = index_c[cb];
                } else {
                    if (!index_inside_lattice(to_v)) continue;
                    row = index_bare(i_v, from);
                    col = index_bare(to_v, to);
                }
                Complex t = t0;
                if (fields) {
                    if (gauge_field) {
                        Vector3d A = gauge_field(cell_.atom_position(from, i_v));
                        double theta = A.dot(disp[h]);
                        t *= peierls_phase(theta);
                    }
                    if (row == col) {
                        if (scalar_field) t += scalar_field(cell_.atom_position(from, i_v));
                        if (site_field) t += site_field(index_bare(i_v, from));
                    }
                }
                if constexpr (is_complex_v<Scalar>) {
                    trip.emplace_back(row, col, t);
                    if (row == col) continue;
                    trip.emplace_back(col, row, std::conj(t));
                } else {
                    if (std::abs(t.imag()) > 1e-10 && !warned) {
                        std::cerr << "Warning: Imaginary part of hopping is discarded." << std::endl;
                        warned = true;
                    }
                    trip.emplace_back(row, col, t.real());
                    if (row == col) continue;
                    trip.emplace_back(col, row, t.real());
                }
            }
        }
        SparseMatrix<Scalar> H(total_number, total_number);
        H.setFromTriplets(trip.begin(), trip.end());
        if (auto_info)
            std::cout << "Sparse Hamiltonian generated with number of nonzero: " << H.nonZeros()
                      << std::endl;
        return H;
    }

    UnitCell<Dim> cell_;
    LatticeVector<Dim> span_v, span_s;
    int cell_number, total_number;
    std::vector<int> removed_atoms;
    std::vector<int> index_c;
    std::function<double(Vector3d)> scalar_field;
    std::function<double(int)> site_field;
};

// Spectral rescaling H~ = (H - b I)/a over the union of the pattern of H and
// the identity (lattice.h:550-559): off-diagonal entries become t/a, diagonal
// entries (h_ii - b)/a, and a missing diagonal entry becomes (0 - b)/a.
template <typename Scalar>
SparseMatrix<Scalar> rescale(const SparseMatrix<Scalar>& H, double a, double b) {
    SparseMatrix<Scalar> out(H.rows(), H.cols());
    auto& op = out.outer();
    auto& ip = out.inner();
    auto& vp = out.values();
    ip.reserve(H.nonZeros() + H.cols());
    vp.reserve(H.nonZeros() + H.cols());
    const Scalar bI = Scalar(b);
    for (int32_t j = 0; j < H.cols(); ++j) {
        bool diag_done = (j >= H.rows());
        for (int32_t p = H.outerIndexPtr()[j]; p < H.outerIndexPtr()[j + 1]; ++p) {
            int32_t i = H.innerIndexPtr()[p];
            if (!diag_done && i > j) {
                ip.push_back(j);
                vp.push_back((Scalar(0) - bI) / a);
                diag_done = true;
            }
            if (i == j) {
                ip.push_back(i);
                vp.push_back((H.valuePtr()[p] - bI) / a);
                diag_done = true;
            } else {
                ip.push_back(i);
                vp.push_back((H.valuePtr()[p] - Scalar(0)) / a);
            }
        }
        if (!diag_done) {
            ip.push_back(j);
            vp.push_back((Scalar(0) - bI) / a);
        }
        op[j + 1] = static_cast<int32_t>(ip.size());
    }
    return out;
}

}  // namespace tb
