// C ABI over the host model layer (tight_binding_amd/host): builds lattices
// and their sparse Hamiltonians for callers in other languages.
#include <cstring>
#include <memory>
#include <string>
#include <variant>
#include <vector>

#include "../host/lattice.h"
#include "../host/models.h"
#include "tbhost.h"
#include "tbkpm.h"

using namespace tb;

struct tbh_cell {
    std::variant<UnitCell<1>, UnitCell<2>, UnitCell<3>> v;
};
struct tbh_lattice {
    std::variant<PartialLattice<1>, PartialLattice<2>, PartialLattice<3>> v;
    // fields set through this interface (the device builder reproduces them)
    int32_t gauge = 0;
    double B = 0.0;
    bool disorder = false;
    double W = 0.0;
    uint64_t seed = 0;
};
struct tbh_csc {
    std::variant<SparseMatrix<double>, SparseMatrix<Complex>> v;
};

namespace {
thread_local std::string g_err;

struct Err {
    int code;
    std::string msg;
};

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return TBK_OK;
    } catch (const Err& e) {
        g_err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return TBK_E_OOM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return TBK_E_INVALID;
    } catch (...) {
        g_err = "unknown error";
        return TBK_E_INVALID;
    }
}
[[noreturn]] void bad(const std::string& m) { throw Err{TBK_E_INVALID, m}; }

Vector3d v3(const double* p) { return {p[0], p[1], p[2]}; }

template <int Dim>
LatticeVector<Dim> lv(const int32_t* p) {
    LatticeVector<Dim> r;
    for (int i = 0; i < Dim; ++i) r(i) = p[i];
    return r;
}

template <int Dim>
UnitCell<Dim> make_cell(const double* lvec, int32_t n_atoms, const char* const* names, const double* pos) {
    std::vector<Vector3d> L;
    for (int i = 0; i < Dim; ++i) L.push_back(v3(lvec + 3 * i));
    std::vector<std::pair<std::string, Vector3d>> atoms;
    for (int i = 0; i < n_atoms; ++i) atoms.emplace_back(names[i], v3(pos + 3 * i));
    return UnitCell<Dim>(std::move(L), std::move(atoms));
}
}  // namespace

extern "C" {

const char* tbh_last_error(void) { return g_err.c_str(); }

int tbh_cell_create(int32_t dim, const double* lattice_vectors, int32_t n_atoms, const char* const* names,
                    const double* positions, tbh_cell** out) {
    return guarded([&] {
        if (!lattice_vectors || !names || !positions || !out || n_atoms <= 0) bad("invalid argument");
        auto c = std::make_unique<tbh_cell>(tbh_cell{UnitCell<1>({{1, 0, 0}}, {{"A", {0, 0, 0}}})});
        if (dim == 1)
            c->v = make_cell<1>(lattice_vectors, n_atoms, names, positions);
        else if (dim == 2)
            c->v = make_cell<2>(lattice_vectors, n_atoms, names, positions);
        else if (dim == 3)
            c->v = make_cell<3>(lattice_vectors, n_atoms, names, positions);
        else
            bad("dimension must be 1, 2 or 3");
        *out = c.release();
    });
}

int tbh_cell_model(const char* name, tbh_cell** out) {
    return guarded([&] {
        if (!name || !out) bad("invalid argument");
        std::string n(name);
        auto c = std::make_unique<tbh_cell>(tbh_cell{models::Chain()});
        if (n == "honeycomb") c->v = models::Honeycomb();
        else if (n == "lieb") c->v = models::Lieb();
        else if (n == "kagome") c->v = models::Kagome();
        else if (n == "square") c->v = models::Square();
        else if (n == "chain") c->v = models::Chain();
        else if (n == "fcc") c->v = models::FCC();
        else if (n == "cubic") c->v = models::Cubic();
        else if (n == "graphene") c->v = models::graphene();
        else if (n == "ti_sp") c->v = models::TI_sp();
        else bad("unknown model " + n);
        *out = c.release();
    });
}

int tbh_cell_destroy(tbh_cell* c) {
    delete c;
    return TBK_OK;
}

int tbh_cell_dim(const tbh_cell* c, int32_t* dim) {
    return guarded([&] {
        if (!c || !dim) bad("invalid argument");
        *dim = static_cast<int32_t>(c->v.index()) + 1;
    });
}

int tbh_cell_add_hopping(tbh_cell* c, const char* from, const char* to, const int32_t* lattice_vector,
                         double t_re, double t_im) {
    return guarded([&] {
        if (!c || !from || !to || !lattice_vector) bad("invalid argument");
        std::visit(
            [&](auto& cell) {
                constexpr int D = std::remove_reference_t<decltype(cell)>::dim;
                cell.add_hopping(from, to, lv<D>(lattice_vector), Complex(t_re, t_im));
            },
            c->v);
    });
}

int tbh_cell_add_hopping_nn(tbh_cell* c, const char* from, uint64_t k, double t_re, double t_im) {
    return guarded([&] {
        if (!c || !from) bad("invalid argument");
        std::visit([&](auto& cell) { cell.add_hopping(from, static_cast<HoppingType>(k), Complex(t_re, t_im)); },
                   c->v);
    });
}

int tbh_cell_hopping_count(const tbh_cell* c, int32_t* n) {
    return guarded([&] {
        if (!c || !n) bad("invalid argument");
        std::visit([&](const auto& cell) { *n = static_cast<int32_t>(cell.hopping_view().size()); }, c->v);
    });
}

int tbh_cell_hoppings(const tbh_cell* c, int32_t* from, int32_t* to, int32_t* lattice_vector, double* t) {
    return guarded([&] {
        if (!c || !from || !to || !lattice_vector || !t) bad("invalid argument");
        std::visit(
            [&](const auto& cell) {
                constexpr int D = std::remove_reference_t<decltype(cell)>::dim;
                size_t i = 0;
                for (const auto& [hp, tv] : cell.hopping_view()) {
                    from[i] = std::get<0>(hp);
                    to[i] = std::get<1>(hp);
                    for (int d = 0; d < D; ++d) lattice_vector[i * D + d] = std::get<2>(hp)(d);
                    t[2 * i] = tv.real();
                    t[2 * i + 1] = tv.imag();
                    ++i;
                }
            },
            c->v);
    });
}

int tbh_lattice_create(const tbh_cell* c, const int32_t* span, tbh_lattice** out) {
    return guarded([&] {
        if (!c || !span || !out) bad("invalid argument");
        std::visit(
            [&](const auto& cell) {
                constexpr int D = std::remove_reference_t<decltype(cell)>::dim;
                *out = new tbh_lattice{PartialLattice<D>(cell, lv<D>(span))};
            },
            c->v);
    });
}

int tbh_lattice_destroy(tbh_lattice* l) {
    delete l;
    return TBK_OK;
}

int tbh_lattice_total_number(const tbh_lattice* l, int32_t* n) {
    return guarded([&] {
        if (!l || !n) bad("invalid argument");
        std::visit([&](const auto& lat) { *n = lat.get_total_number(); }, l->v);
    });
}

int tbh_lattice_set_gauge(tbh_lattice* l, int32_t kind, double B) {
    return guarded([&] {
        if (!l) bad("invalid argument");
        std::visit(
            [&](auto& lat) {
                if (kind == 0)
                    lat.gauge_field = nullptr;
                else if (kind == 1)
                    lat.gauge_field = [B](Vector3d r) { return Vector3d{0.0, B * r(0), 0.0}; };
                else if (kind == 2)
                    lat.gauge_field = [B](Vector3d r) { return B * Vector3d{-r(1), r(0), 0} / 2; };
                else
                    bad("unknown gauge kind");
                l->gauge = kind;
                l->B = B;
            },
            l->v);
    });
}

int tbh_lattice_set_disorder(tbh_lattice* l, double W, uint64_t seed) {
    return guarded([&] {
        if (!l) bad("invalid argument");
        std::visit([&](auto& lat) { lat.set_scalar_field_indexed(uniform_disorder(W, seed)); }, l->v);
        l->disorder = true;
        l->W = W;
        l->seed = seed;
    });
}

int tbh_lattice_info_get(const tbh_lattice* l, tbh_lattice_info* out) {
    return guarded([&] {
        if (!l || !out) bad("invalid argument");
        *out = tbh_lattice_info{};
        std::visit(
            [&](const auto& lat) {
                constexpr int D = std::remove_reference_t<decltype(lat)>::dim;
                out->dim = D;
                out->n_atoms = lat.cell().get_atom_number();
                out->n_hops = static_cast<int32_t>(lat.cell().hopping_view().size());
                for (int d = 0; d < 3; ++d) out->span[d] = d < D ? lat.span()(d) : 1;
                out->cell_number = lat.get_cell_number();
                out->total_number = lat.get_total_number();
                out->removed_sites = lat.removed_count();
                out->position_field = lat.has_position_field() ? 1 : 0;
            },
            l->v);
        out->gauge = l->gauge;
        out->B = l->B;
        out->has_disorder = l->disorder ? 1 : 0;
        out->disorder_W = l->W;
        out->disorder_seed = l->seed;
    });
}

int tbh_lattice_hoppings(const tbh_lattice* l, int32_t* from, int32_t* to, int32_t* lattice_vector, double* t) {
    return guarded([&] {
        if (!l || !from || !to || !lattice_vector || !t) bad("invalid argument");
        std::visit(
            [&](const auto& lat) {
                constexpr int D = std::remove_reference_t<decltype(lat)>::dim;
                size_t i = 0;
                for (const auto& [hp, tv] : lat.cell().hopping_view()) {
                    from[i] = std::get<0>(hp);
                    to[i] = std::get<1>(hp);
                    for (int d = 0; d < D; ++d) lattice_vector[i * D + d] = std::get<2>(hp)(d);
                    t[2 * i] = tv.real();
                    t[2 * i + 1] = tv.imag();
                    ++i;
                }
            },
            l->v);
    });
}

int tbh_lattice_geometry(const tbh_lattice* l, double* bases, double* atom_pos) {
    return guarded([&] {
        if (!l || !bases || !atom_pos) bad("invalid argument");
        std::visit(
            [&](const auto& lat) {
                constexpr int D = std::remove_reference_t<decltype(lat)>::dim;
                const auto& c = lat.cell();
                for (int k = 0; k < 9; ++k) bases[k] = 0.0;
                for (int d = 0; d < D; ++d)
                    for (int k = 0; k < 3; ++k) bases[3 * d + k] = c.basis(d)(k);
                for (int a = 0; a < c.get_atom_number(); ++a)
                    for (int k = 0; k < 3; ++k) atom_pos[3 * a + k] = c.get_atom_pos(a)(k);
            },
            l->v);
    });
}

int tbh_lattice_phase_table(const tbh_lattice* l, double* out) {
    return guarded([&] {
        if (!l || !out) bad("invalid argument");
        if (l->gauge == 0) bad("no gauge field is set");
        std::visit(
            [&](const auto& lat) {
                const std::vector<Complex> p = lat.phase_table();
                for (size_t i = 0; i < p.size(); ++i) {
                    out[2 * i] = p[i].real();
                    out[2 * i + 1] = p[i].imag();
                }
            },
            l->v);
    });
}

int tbh_lattice_remove_atoms(tbh_lattice* l, const char* atom, const int32_t* cells, int32_t n) {
    return guarded([&] {
        if (!l || !atom || (n > 0 && !cells)) bad("invalid argument");
        std::visit(
            [&](auto& lat) {
                constexpr int D = std::remove_reference_t<decltype(lat)>::dim;
                std::vector<LatticeVector<D>> v;
                for (int i = 0; i < n; ++i) v.push_back(lv<D>(cells + D * i));
                lat.remove_atoms(atom, v);
            },
            l->v);
    });
}

int tbh_lattice_hamiltonian(tbh_lattice* l, int32_t builder, const int32_t* extend, int32_t is_complex,
                            tbh_csc** out) {
    return guarded([&] {
        if (!l || !out) bad("invalid argument");
        std::visit(
            [&](auto& lat) {
                constexpr int D = std::remove_reference_t<decltype(lat)>::dim;
                auto build = [&](auto tag) {
                    using S = decltype(tag);
                    if (builder == 0) return lat.template get_Hamiltonian_sparse<S>();
                    if (builder == 1) {
                        if (!extend) bad("periodic builder needs extend");
                        return lat.template get_Hamiltonian_periodic_extend_bare_sparse<S>(lv<D>(extend));
                    }
                    if (builder == 2) return lat.template get_Hamiltonian_bare_sparse<S>();
                    bad("unknown builder");
                };
                if (is_complex)
                    *out = new tbh_csc{build(Complex())};
                else
                    *out = new tbh_csc{build(0.0)};
            },
            l->v);
    });
}

int tbh_lattice_op_position(tbh_lattice* l, int32_t axis, tbh_csc** out) {
    return guarded([&] {
        if (!l || !out) bad("invalid argument");
        std::visit([&](auto& lat) { *out = new tbh_csc{lat.op_position(axis)}; }, l->v);
    });
}

int tbh_lattice_op_velocity(tbh_lattice* l, int32_t axis, int32_t is_complex, tbh_csc** out) {
    return guarded([&] {
        if (!l || !out) bad("invalid argument");
        std::visit(
            [&](auto& lat) {
                if (is_complex)
                    *out = new tbh_csc{lat.template op_velocity<Complex>(axis)};
                else
                    *out = new tbh_csc{lat.template op_velocity<double>(axis)};
            },
            l->v);
    });
}

int tbh_csc_create(int32_t n, int64_t nnz, const int32_t* colptr, const int32_t* rowidx, const double* values,
                   int32_t is_complex, tbh_csc** out) {
    return guarded([&] {
        if (n < 0 || nnz < 0 || !colptr || !out || (nnz > 0 && (!rowidx || !values))) bad("invalid argument");
        if (colptr[0] != 0 || colptr[n] != nnz) bad("column pointers inconsistent with nnz");
        for (int32_t j = 0; j < n; ++j)
            if (colptr[j + 1] < colptr[j]) bad("column pointers not monotone");
        for (int64_t p = 0; p < nnz; ++p)
            if (rowidx[p] < 0 || rowidx[p] >= n) bad("row index out of range");
        auto fill = [&](auto& M) {
            using S = typename std::remove_reference_t<decltype(M)>::value_type;
            M.outer().assign(colptr, colptr + n + 1);
            M.inner().assign(rowidx, rowidx + nnz);
            M.values().resize(static_cast<size_t>(nnz));
            std::memcpy(static_cast<void*>(M.values().data()), values, static_cast<size_t>(nnz) * sizeof(S));
        };
        if (is_complex) {
            SparseMatrix<Complex> M(n, n);
            fill(M);
            *out = new tbh_csc{std::move(M)};
        } else {
            SparseMatrix<double> M(n, n);
            fill(M);
            *out = new tbh_csc{std::move(M)};
        }
    });
}

int tbh_csc_rescale(const tbh_csc* m, double a, double b, tbh_csc** out) {
    return guarded([&] {
        if (!m || !out) bad("invalid argument");
        if (a == 0.0) bad("rescaling by a = 0");
        std::visit([&](const auto& M) { *out = new tbh_csc{rescale(M, a, b)}; }, m->v);
    });
}

int tbh_csc_commutator_diag(const tbh_csc* x, const tbh_csc* m, tbh_csc** out) {
    return guarded([&] {
        if (!x || !m || !out) bad("invalid argument");
        const auto* X = std::get_if<SparseMatrix<double>>(&x->v);
        if (!X) bad("position operator must be real");
        std::visit(
            [&](const auto& M) { *out = new tbh_csc{PartialLattice<1>::commutator_with_diagonal(*X, M)}; },
            m->v);
    });
}

int tbh_csc_info(const tbh_csc* m, int32_t* n, int64_t* nnz, int32_t* is_complex) {
    return guarded([&] {
        if (!m) bad("invalid argument");
        std::visit(
            [&](const auto& M) {
                if (n) *n = M.rows();
                if (nnz) *nnz = M.nonZeros();
            },
            m->v);
        if (is_complex) *is_complex = static_cast<int32_t>(m->v.index());
    });
}

int tbh_csc_copy(const tbh_csc* m, int32_t* colptr, int32_t* rowidx, double* values) {
    return guarded([&] {
        if (!m || !colptr || !rowidx || !values) bad("invalid argument");
        std::visit(
            [&](const auto& M) {
                using S = typename std::remove_reference_t<decltype(M)>::value_type;
                std::memcpy(colptr, M.outerIndexPtr(), (M.cols() + 1) * sizeof(int32_t));
                std::memcpy(rowidx, M.innerIndexPtr(), M.nonZeros() * sizeof(int32_t));
                std::memcpy(values, M.valuePtr(), M.nonZeros() * sizeof(S));
            },
            m->v);
    });
}

int tbh_csc_copy_csr(const tbh_csc* m, int32_t* rowptr, int32_t* colidx, double* values) {
    return guarded([&] {
        if (!m || !rowptr || !colidx || !values) bad("invalid argument");
        std::visit(
            [&](const auto& M) {
                using S = typename std::remove_reference_t<decltype(M)>::value_type;
                std::vector<int32_t> rp, ci;
                std::vector<S> v;
                M.to_csr(rp, ci, v);
                std::memcpy(rowptr, rp.data(), rp.size() * sizeof(int32_t));
                std::memcpy(colidx, ci.data(), ci.size() * sizeof(int32_t));
                std::memcpy(values, v.data(), v.size() * sizeof(S));
            },
            m->v);
    });
}

int tbh_csc_destroy(tbh_csc* m) {
    delete m;
    return TBK_OK;
}

}  // extern "C"
