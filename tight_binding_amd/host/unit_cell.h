// UnitCell<Dim> / UnitCellNN<Dim>: lattice vectors, atoms and the canonical
// hopping table.  Same public surface and semantics as the reference
// (unit_cell.h:40-303, 527-547, 665-741):
//
//  * hoppings are stored with from >= to; a hopping with from < to is stored
//    as its Hermitian conjugate (to, from, -dR) with conj(t)
//    (unit_cell.h:545-546);
//  * a hopping whose (from,to,dR) or conjugate is already present is rejected
//    (unit_cell.h:527-539); on-site terms must be real (unit_cell.h:221-224);
//  * add_hopping(from, k_NN, t) adds the k-th shell of neighbours of atom
//    `from` found by a distance min-heap over cells reached through the
//    3^Dim-1 neighbouring lattice vectors, with shells grouped at 1e-7
//    (unit_cell.h:680-707).  Shell 0 is distance zero.
//
// Deliberate difference: the hopping table iterates in insertion order (the
// reference iterates an unordered_map, whose order is unspecified).  The
// order only decides the summation order of duplicate matrix entries.
#pragma once

#include <cmath>
#include <cstdint>
#include <functional>
#include <iostream>
#include <map>
#include <numbers>
#include <queue>
#include <set>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <utility>
#include <vector>

#include "linalg.h"

namespace tb {

using std::numbers::pi;

template <int Dim>
using Hopping = std::tuple<int, int, LatticeVector<Dim>>;

using HoppingType = unsigned long long int;
constexpr HoppingType On_Site = 0;

inline namespace literals {
inline HoppingType operator""_NN(unsigned long long int n) { return static_cast<HoppingType>(n); }
}  // namespace literals

template <int Dim>
struct HoppingLess {
    bool operator()(const Hopping<Dim>& a, const Hopping<Dim>& b) const {
        if (std::get<0>(a) != std::get<0>(b)) return std::get<0>(a) < std::get<0>(b);
        if (std::get<1>(a) != std::get<1>(b)) return std::get<1>(a) < std::get<1>(b);
        for (int i = 0; i < Dim; ++i)
            if (std::get<2>(a)(i) != std::get<2>(b)(i)) return std::get<2>(a)(i) < std::get<2>(b)(i);
        return false;
    }
};

template <int Dim>
class UnitCell;

// Nearest-neighbour shell generator for one atom (unit_cell.h:162-186).
template <int Dim>
class UnitCellNN {
public:
    explicit UnitCellNN(int atom_no) : A(atom_no) {}

    using NN_info = std::tuple<double, Hopping<Dim>>;
    struct Greater {
        bool operator()(const NN_info& n, const NN_info& m) const {
            return std::get<0>(n) > std::get<0>(m);
        }
    };

    // generate shells 0..NN (unit_cell.h:680-695)
    void generate_hoppings_to(const UnitCell<Dim>& cell, HoppingType NN) {
        if (registered.empty()) add_cell(cell, LatticeVector<Dim>::Zero());
        while (NN_solved.size() <= NN) {
            auto [d, hp] = heap.top();
            double d0 = d;
            std::vector<Hopping<Dim>> shell;
            while (std::abs(d - d0) < 1e-7) {
                for (const auto& dx : UnitCell<Dim>::NN_vectors())
                    add_cell(cell, std::get<2>(hp) + dx);
                shell.emplace_back(hp);
                heap.pop();
                std::tie(d, hp) = heap.top();
            }
            NN_solved.emplace_back(std::move(shell));
        }
    }
    const std::vector<Hopping<Dim>>& shell(HoppingType k) const { return NN_solved.at(k); }

private:
    // queue every atom of cell `a` as a candidate (unit_cell.h:699-707)
    void add_cell(const UnitCell<Dim>& cell, LatticeVector<Dim> a) {
        if (!registered.insert(key(a)).second) return;
        for (int B = 0; B < cell.get_atom_number(); ++B) {
            Hopping<Dim> hp{A, B, a};
            heap.emplace(cell.distance_of_hopping(hp), hp);
        }
    }
    static std::vector<int> key(const LatticeVector<Dim>& a) {
        return std::vector<int>(a.v, a.v + Dim);
    }

    int A;
    std::priority_queue<NN_info, std::vector<NN_info>, Greater> heap;
    std::set<std::vector<int>> registered;
    std::vector<std::vector<Hopping<Dim>>> NN_solved;
};

template <int Dim>
class UnitCell {
public:
    static constexpr int dim = Dim;
    UnitCell(std::vector<Vector3d>&& lattice_vectors,
             std::vector<std::pair<std::string, Vector3d>>&& atoms) {
        if (static_cast<int>(lattice_vectors.size()) != Dim)
            throw std::runtime_error(
                "Number of lattice vectors must be equal to the dimension of the unit cell");
        for (auto&& v : lattice_vectors)
            if (v.isZero()) throw std::runtime_error("Lattice vectors must not be zero");
        if (atoms.empty()) throw std::runtime_error("Unit cell must contain at least one atom");
        for (int i = 0; i < Dim; ++i) bases_[i] = lattice_vectors[i];
        for (size_t i = 0; i < atoms.size(); ++i) {
            if (atom_no.count(atoms[i].first))
                throw std::runtime_error("Atom " + atoms[i].first + " is already defined");
            atom_no[atoms[i].first] = static_cast<int>(i);
            atom_name.push_back(atoms[i].first);
            atom_pos.push_back(atoms[i].second);
        }
    }

    // explicit hopping (unit_cell.h:212-228)
    UnitCell& add_hopping(const std::string& from, const std::string& to,
                          LatticeVector<Dim> lattice_vector, Complex hopping_strength) {
        auto f = atom_no.find(from);
        if (f == atom_no.end()) throw std::runtime_error("Atom " + from + " not found in the unit cell");
        auto t = atom_no.find(to);
        if (t == atom_no.end()) throw std::runtime_error("Atom " + to + " not found in the unit cell");
        if (from == to && lattice_vector.isZero() && std::imag(hopping_strength) != 0) {
            std::cerr << "On-site hopping must be real" << std::endl;
            return *this;
        }
        add_hopping_direct(Hopping<Dim>{f->second, t->second, lattice_vector}, hopping_strength);
        return *this;
    }
    // k-th nearest-neighbour shell (unit_cell.h:230-256)
    UnitCell& add_hopping(const std::string& from, HoppingType NN, Complex hopping_strength) {
        auto f = atom_no.find(from);
        if (f == atom_no.end()) {
            std::cerr << "Atom " << from << " not found in the unit cell" << std::endl;
            return *this;
        }
        int from_no = f->second;
        if (NN == On_Site) {
            if (std::imag(hopping_strength) != 0) {
                std::cerr << "On-site hopping must be real" << std::endl;
                return *this;
            }
            add_hopping_direct(Hopping<Dim>{from_no, from_no, LatticeVector<Dim>::Zero()},
                               hopping_strength);
            return *this;
        }
        auto it = hopping_gen.find(from_no);
        if (it == hopping_gen.end()) it = hopping_gen.emplace(from_no, UnitCellNN<Dim>(from_no)).first;
        it->second.generate_hoppings_to(*this, NN);
        for (auto hp : it->second.shell(NN)) add_hopping_direct(std::move(hp), hopping_strength);
        return *this;
    }

    // bases * a, accumulated left to right (Eigen lazy product order)
    Vector3d displacement_of_lattice(const LatticeVector<Dim>& a) const {
        Vector3d r;
        for (int i = 0; i < 3; ++i) {
            double s = bases_[0](i) * static_cast<double>(a(0));
            for (int j = 1; j < Dim; ++j) s = s + bases_[j](i) * static_cast<double>(a(j));
            r(i) = s;
        }
        return r;
    }
    // -pos(from) + bases*dR + pos(to) (unit_cell.h:268-273)
    Vector3d displacement_of_hopping(const Hopping<Dim>& h) const {
        const auto& [from, to, lv] = h;
        Vector3d L = displacement_of_lattice(lv);
        const Vector3d& pf = atom_pos.at(from);
        const Vector3d& pt = atom_pos.at(to);
        Vector3d r;
        for (int i = 0; i < 3; ++i) r(i) = ((-pf(i)) + L(i)) + pt(i);
        return r;
    }
    double distance_of_hopping(const Hopping<Dim>& h) const { return displacement_of_hopping(h).norm(); }
    double distance_between_lattices(LatticeVector<Dim> a) const { return displacement_of_lattice(a).norm(); }

    static Hopping<Dim> hopping_hc(const Hopping<Dim>& h) {
        return Hopping<Dim>{std::get<1>(h), std::get<0>(h), -std::get<2>(h)};
    }

    Vector3d atom_position(const std::string& atom) const { return atom_pos.at(atom_no.at(atom)); }
    Vector3d atom_position(int a, const LatticeVector<Dim>& lat_v) const {
        const Vector3d& p = atom_pos.at(a);
        Vector3d L = displacement_of_lattice(lat_v);
        return {p(0) + L(0), p(1) + L(1), p(2) + L(2)};
    }
    Vector3d atom_position(const std::string& atom, const LatticeVector<Dim>& lat_v) const {
        return atom_position(atom_no.at(atom), lat_v);
    }

    int get_atom_number() const { return static_cast<int>(atom_pos.size()); }
    int get_atom_index(const std::string& name) const { return atom_no.at(name); }
    bool atom_exists(const std::string& name) const { return atom_no.count(name) != 0; }
    const std::string& get_atom_name(int i) const { return atom_name.at(i); }
    const Vector3d& get_atom_pos(int i) const { return atom_pos.at(i); }
    const Vector3d& basis(int j) const { return bases_[j]; }

    // (Hopping, t) pairs in insertion order
    const std::vector<std::pair<Hopping<Dim>, Complex>>& hopping_view() const { return hoppings; }

    bool hopping_exists(const Hopping<Dim>& h) const {
        return index.count(h) != 0 || index.count(hopping_hc(h)) != 0;
    }
    // canonical insertion (unit_cell.h:533-547)
    void add_hopping_direct(Hopping<Dim> h, Complex t) {
        if (hopping_exists(h)) {
            // expected for the +-dR pairs of a shell; reported only on request
            if (auto_show) std::cerr << "Hopping already exists" << std::endl;
            return;
        }
        const auto& [from, to, lv] = h;
        if (from == to && lv.isZero() && std::imag(t) != 0) {
            std::cerr << "On-site hopping must be real" << std::endl;
            return;
        }
        Hopping<Dim> stored = (from < to) ? hopping_hc(h) : h;
        Complex ts = (from < to) ? std::conj(t) : t;
        index.emplace(stored, hoppings.size());
        hoppings.emplace_back(stored, ts);
    }

    // Bloch Hamiltonian H(k) (unit_cell.h:305-319), used by the tests as a
    // closed-form spectrum for periodic lattices.
    std::vector<Complex> Hamiltonian_k(const Vector3d& k) const {
        const int n = get_atom_number();
        std::vector<Complex> H(static_cast<size_t>(n) * n, Complex(0, 0));
        for (const auto& [hp, t] : hoppings) {
            const auto& [row, col, lv] = hp;
            Vector3d r = displacement_of_hopping(hp);
            double kr = k.dot(r);
            if (row == col) {
                H[row * n + col] += t.real() * 2. * std::cos(kr);
            } else {
                Complex e = t * std::exp(Complex(0, 1) * kr);
                H[row * n + col] += e;
                H[col * n + row] += std::conj(e);
            }
        }
        return H;
    }

    // 3^Dim - 1 neighbouring lattice vectors, last index fastest
    static const std::vector<LatticeVector<Dim>>& NN_vectors() {
        static const std::vector<LatticeVector<Dim>> v = [] {
            std::vector<LatticeVector<Dim>> out;
            int total = 1;
            for (int i = 0; i < Dim; ++i) total *= 3;
            for (int c = 0; c < total; ++c) {
                LatticeVector<Dim> lv;
                int r = c;
                for (int i = Dim - 1; i >= 0; --i) {
                    lv(i) = (r % 3) - 1;
                    r /= 3;
                }
                if (!lv.isZero()) out.push_back(lv);
            }
            return out;
        }();
        return v;
    }

    // diagnostics switch (the reference's UnitCell::auto_show, unit_cell.h:112)
    inline static bool auto_show = false;

    static LatticeVector<Dim> Intra_Cell() { return LatticeVector<Dim>::Zero(); }
    static LatticeVector<Dim> Left_NN() { return -LatticeVector<Dim>::Unit(0); }
    static LatticeVector<Dim> Right_NN() { return LatticeVector<Dim>::Unit(0); }

private:
    Vector3d bases_[Dim];
    std::unordered_map<std::string, int> atom_no;
    std::vector<std::string> atom_name;
    std::vector<Vector3d> atom_pos;
    std::vector<std::pair<Hopping<Dim>, Complex>> hoppings;
    std::map<Hopping<Dim>, size_t, HoppingLess<Dim>> index;
    std::map<int, UnitCellNN<Dim>> hopping_gen;
};

using UnitCell_1D = UnitCell<1>;
using UnitCell_2D = UnitCell<2>;
using UnitCell_3D = UnitCell<3>;

}  // namespace tb
