// Lattice presets with the reference's geometry (models.h:6-150).  Each
// preset is returned by a function so that callers get their own copy, as
// the reference's `UnitCell<2> p = models::Honeycomb;` idiom does.
#pragma once

#include <cmath>

#include "unit_cell.h"

namespace tb::models {

inline UnitCell<2> Honeycomb() {
    return UnitCell<2>{{{3. / 2, -std::sqrt(3.) / 2, 0}, {3. / 2, std::sqrt(3.) / 2, 0}},
                       {{"A", {0, 0, 0}}, {"B", {1, 0, 0}}}};
}
inline UnitCell<1> stub_1d_X() {
    return UnitCell<1>{{{1, 0, 0}}, {{"A", {0, 0, 0}}, {"C", {0.5, 0, 0}}, {"B", {0, 0.6, 0}}}};
}
inline UnitCell<1> sawtooth_chain_X() {
    return UnitCell<1>{{{1, 0, 0}}, {{"A", {0, 0, 0}}, {"B", {0, 1, 0}}}};
}
inline UnitCell<2> Lieb() {
    return UnitCell<2>{{{1, 0, 0}, {0, 1, 0}},
                       {{"A", {0, 0, 0}}, {"B", {0.5, 0, 0}}, {"C", {0, 0.5, 0}}}};
}
inline UnitCell<2> Kagome() {
    return UnitCell<2>{{{1, 0, 0}, {0.5, std::sqrt(3.) / 2, 0}},
                       {{"A", {0, 0, 0}}, {"B", {0.5, 0, 0}}, {"C", {0.25, std::sqrt(3.) / 4, 0}}}};
}
inline UnitCell<2> Square() { return UnitCell<2>{{{1, 0, 0}, {0, 1, 0}}, {{"A", {0, 0, 0}}}}; }
inline UnitCell<1> Chain() { return UnitCell<1>{{{1, 0, 0}}, {{"A", {0, 0, 0}}}}; }
inline UnitCell<3> FCC() {
    return UnitCell<3>{{{0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}}, {{"A", {0, 0, 0}}}};
}
inline UnitCell<3> Cubic() {
    return UnitCell<3>{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, {{"A", {0, 0, 0}}}};
}

// two-orbital s-p model (models.h:104-127)
inline UnitCell<2> TI_sp() {
    UnitCell<2> p{{{1, 0, 0}, {0, 1, 0}}, {{"s", {0, 0, 0}}, {"p", {0, 0, 0}}}};
    Complex ep = 0.6, es = -ep, t_ss = 0.4, t_pp = -t_ss, t_sp = 0.2;
    const Complex I(0, 1);
    p.add_hopping("s", "s", p.Intra_Cell(), es);
    p.add_hopping("p", "p", p.Intra_Cell(), ep);
    p.add_hopping("s", "s", {1, 0}, t_ss);
    p.add_hopping("s", "s", {0, 1}, t_ss);
    p.add_hopping("p", "p", {1, 0}, t_pp);
    p.add_hopping("p", "p", {0, 1}, t_pp);
    p.add_hopping("s", "p", {1, 0}, 1. * t_sp);
    p.add_hopping("s", "p", {0, 1}, I * t_sp);
    p.add_hopping("s", "p", {-1, 0}, -1. * t_sp);
    p.add_hopping("s", "p", {0, -1}, -I * t_sp);
    return p;
}

// Honeycomb with nearest-neighbour t = +1 (models.h:129-133)
inline UnitCell<2> graphene() {
    UnitCell<2> p = Honeycomb();
    p.add_hopping("A", 1_NN, 1.0);
    return p;
}

inline UnitCell<1> stub_1d_p(double t, double t_) {
    UnitCell<1> p = stub_1d_X();
    p.add_hopping("A", "C", p.Intra_Cell(), t);
    p.add_hopping("A", "C", p.Left_NN(), t);
    p.add_hopping("A", "B", p.Intra_Cell(), t_);
    return p;
}
inline UnitCell<1> sawtooth_chain_p(double t, double t_) {
    UnitCell<1> p = sawtooth_chain_X();
    p.add_hopping("A", "A", p.Right_NN(), t);
    p.add_hopping("A", "B", p.Intra_Cell(), t_);
    p.add_hopping("B", "A", p.Right_NN(), t_);
    return p;
}

}  // namespace tb::models
