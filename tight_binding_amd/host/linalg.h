// Minimal dense/sparse containers with the accessor surface the reference's
// hot path uses from Eigen (Eigen is not a dependency of this library).
//
//   SparseMatrix<Scalar>   column-major compressed (CSC) storage, as Eigen's
//                          default ColMajor SparseMatrix used by the reference
//                          (lattice.h:253-254, numerics_cuda.cu:52-64):
//                          outerIndexPtr() = column pointers,
//                          innerIndexPtr() = row indices, valuePtr() = values.
//   Triplet<Scalar>        (row, col, value) as Eigen::Triplet.
//   MatrixXcd              column-major dense complex block (Eigen::MatrixXcd),
//                          the layout of bras_T / kets in numerics_cuda.cuh:10-13.
//   Vector3d, LatticeVector<Dim>  fixed-size vectors (unit_cell.h:29,46).
//
// setFromTriplets reproduces Eigen's documented semantics exactly: duplicate
// (row, col) entries are summed in insertion order into the first occurrence,
// and the result is sorted by (col, row).  This fixes every stored bit of the
// assembled Hamiltonian (lattice.h:253-254, 300-301).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <initializer_list>
#include <stdexcept>
#include <type_traits>
#include <vector>

namespace tb {

using Complex = std::complex<double>;

template <typename T>
inline constexpr bool is_complex_v = false;
template <typename T>
inline constexpr bool is_complex_v<std::complex<T>> = true;

// ---------------------------------------------------------------- Vector3d
struct Vector3d {
    double v[3] = {0.0, 0.0, 0.0};
    Vector3d() = default;
    Vector3d(double x, double y, double z) : v{x, y, z} {}
    double& operator()(int i) { return v[i]; }
    double operator()(int i) const { return v[i]; }
    double& operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
    // Eigen's fixed-size redux sums left to right: ((x0*y0 + x1*y1) + x2*y2).
    double dot(const Vector3d& o) const {
        double s = v[0] * o.v[0];
        s = s + v[1] * o.v[1];
        s = s + v[2] * o.v[2];
        return s;
    }
    double squaredNorm() const { return dot(*this); }
    double norm() const { return std::sqrt(squaredNorm()); }
    bool isZero() const { return v[0] == 0.0 && v[1] == 0.0 && v[2] == 0.0; }
    static Vector3d Zero() { return {}; }
    friend Vector3d operator+(const Vector3d& a, const Vector3d& b) {
        return {a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]};
    }
    friend Vector3d operator-(const Vector3d& a, const Vector3d& b) {
        return {a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]};
    }
    friend Vector3d operator-(const Vector3d& a) { return {-a.v[0], -a.v[1], -a.v[2]}; }
    friend Vector3d operator*(double s, const Vector3d& a) {
        return {s * a.v[0], s * a.v[1], s * a.v[2]};
    }
    friend Vector3d operator*(const Vector3d& a, double s) {
        return {a.v[0] * s, a.v[1] * s, a.v[2] * s};
    }
    friend Vector3d operator/(const Vector3d& a, double s) {
        return {a.v[0] / s, a.v[1] / s, a.v[2] / s};
    }
    Vector3d& operator+=(const Vector3d& o) {
        for (int i = 0; i < 3; ++i) v[i] += o.v[i];
        return *this;
    }
};

// ------------------------------------------------------- LatticeVector<Dim>
template <int Dim>
struct LatticeVector {
    static_assert(Dim >= 1 && Dim <= 3, "lattice dimension must be 1, 2 or 3");
    int v[Dim] = {};
    LatticeVector() = default;
    LatticeVector(std::initializer_list<int> l) {
        if (static_cast<int>(l.size()) != Dim)
            throw std::invalid_argument("LatticeVector: wrong number of components");
        int i = 0;
        for (int x : l) v[i++] = x;
    }
    int& operator()(int i) { return v[i]; }
    int operator()(int i) const { return v[i]; }
    int& operator[](int i) { return v[i]; }
    int operator[](int i) const { return v[i]; }
    bool isZero() const {
        for (int i = 0; i < Dim; ++i)
            if (v[i] != 0) return false;
        return true;
    }
    int dot(const LatticeVector& o) const {
        int s = 0;
        for (int i = 0; i < Dim; ++i) s += v[i] * o.v[i];
        return s;
    }
    static LatticeVector Zero() { return LatticeVector(); }
    static LatticeVector Unit(int i) {
        LatticeVector r;
        r.v[i] = 1;
        return r;
    }
    friend LatticeVector operator+(const LatticeVector& a, const LatticeVector& b) {
        LatticeVector r;
        for (int i = 0; i < Dim; ++i) r.v[i] = a.v[i] + b.v[i];
        return r;
    }
    friend LatticeVector operator-(const LatticeVector& a, const LatticeVector& b) {
        LatticeVector r;
        for (int i = 0; i < Dim; ++i) r.v[i] = a.v[i] - b.v[i];
        return r;
    }
    friend LatticeVector operator-(const LatticeVector& a) {
        LatticeVector r;
        for (int i = 0; i < Dim; ++i) r.v[i] = -a.v[i];
        return r;
    }
    friend bool operator==(const LatticeVector& a, const LatticeVector& b) {
        for (int i = 0; i < Dim; ++i)
            if (a.v[i] != b.v[i]) return false;
        return true;
    }
    friend bool operator!=(const LatticeVector& a, const LatticeVector& b) { return !(a == b); }
};

// ------------------------------------------------------------------ Triplet
template <typename Scalar>
struct Triplet {
    int32_t r, c;
    Scalar v;
    Triplet(int32_t row, int32_t col, Scalar value) : r(row), c(col), v(value) {}
    int32_t row() const { return r; }
    int32_t col() const { return c; }
    Scalar value() const { return v; }
};

// ------------------------------------------------------ SparseMatrix (CSC)
template <typename Scalar>
class SparseMatrix {
public:
    using value_type = Scalar;
    SparseMatrix() = default;
    SparseMatrix(int32_t rows, int32_t cols)
        : rows_(rows), cols_(cols), outer_(static_cast<size_t>(cols) + 1, 0) {}

    int32_t rows() const { return rows_; }
    int32_t cols() const { return cols_; }
    int32_t outerSize() const { return cols_; }
    int64_t nonZeros() const { return static_cast<int64_t>(inner_.size()); }
    bool isCompressed() const { return true; }
    void makeCompressed() {}

    const int32_t* outerIndexPtr() const { return outer_.data(); }
    const int32_t* innerIndexPtr() const { return inner_.data(); }
    const Scalar* valuePtr() const { return values_.data(); }
    int32_t* outerIndexPtr() { return outer_.data(); }
    int32_t* innerIndexPtr() { return inner_.data(); }
    Scalar* valuePtr() { return values_.data(); }

    std::vector<int32_t>& outer() { return outer_; }
    std::vector<int32_t>& inner() { return inner_; }
    std::vector<Scalar>& values() { return values_; }
    const std::vector<int32_t>& outer() const { return outer_; }
    const std::vector<int32_t>& inner() const { return inner_; }
    const std::vector<Scalar>& values() const { return values_; }

    Scalar coeff(int32_t i, int32_t j) const {
        auto b = inner_.begin() + outer_[j], e = inner_.begin() + outer_[j + 1];
        auto it = std::lower_bound(b, e, i);
        if (it != e && *it == i) return values_[it - inner_.begin()];
        return Scalar(0);
    }

    void setIdentity() {
        outer_.assign(static_cast<size_t>(cols_) + 1, 0);
        int32_t d = std::min(rows_, cols_);
        inner_.resize(d);
        values_.assign(d, Scalar(1));
        for (int32_t j = 0; j < cols_; ++j) outer_[j + 1] = std::min(j + 1, d);
        for (int32_t j = 0; j < d; ++j) inner_[j] = j;
    }

    // Eigen setFromTriplets: bucket by row keeping insertion order, merge
    // duplicates into the first occurrence (sum in insertion order), then
    // transpose into column-major storage with ascending row indices.
    template <typename It>
    void setFromTriplets(It begin, It end) {
        const size_t nt = static_cast<size_t>(std::distance(begin, end));
        std::vector<int64_t> row_start(static_cast<size_t>(rows_) + 1, 0);
        for (It it = begin; it != end; ++it) {
            if (it->row() < 0 || it->row() >= rows_ || it->col() < 0 || it->col() >= cols_)
                throw std::out_of_range("setFromTriplets: index out of range");
            row_start[it->row() + 1]++;
        }
        for (int32_t i = 0; i < rows_; ++i) row_start[i + 1] += row_start[i];
        std::vector<int32_t> tcol(nt);
        std::vector<Scalar> tval(nt);
        {
            std::vector<int64_t> fill(row_start.begin(), row_start.end() - 1);
            for (It it = begin; it != end; ++it) {
                int64_t p = fill[it->row()]++;
                tcol[p] = it->col();
                tval[p] = it->value();
            }
        }
        // collapse duplicates row by row (insertion order kept)
        std::vector<int64_t> last_pos(static_cast<size_t>(cols_), -1);
        std::vector<int64_t> row_new_start(static_cast<size_t>(rows_) + 1, 0);
        int64_t w = 0;
        for (int32_t i = 0; i < rows_; ++i) {
            int64_t rs = w;
            row_new_start[i] = rs;
            for (int64_t p = row_start[i]; p < row_start[i + 1]; ++p) {
                int32_t c = tcol[p];
                if (last_pos[c] >= rs) {
                    tval[last_pos[c]] = tval[last_pos[c]] + tval[p];
                } else {
                    last_pos[c] = w;
                    tcol[w] = c;
                    tval[w] = tval[p];
                    ++w;
                }
            }
        }
        row_new_start[rows_] = w;
        // transpose to CSC: rows visited in ascending order => sorted inner
        outer_.assign(static_cast<size_t>(cols_) + 1, 0);
        for (int64_t p = 0; p < w; ++p) outer_[tcol[p] + 1]++;
        for (int32_t j = 0; j < cols_; ++j) outer_[j + 1] += outer_[j];
        inner_.resize(w);
        values_.resize(w);
        std::vector<int32_t> fill(outer_.begin(), outer_.end() - 1);
        for (int32_t i = 0; i < rows_; ++i) {
            for (int64_t p = row_new_start[i]; p < row_new_start[i + 1]; ++p) {
                int32_t q = fill[tcol[p]]++;
                inner_[q] = i;
                values_[q] = tval[p];
            }
        }
    }

    // Row-compressed copy (exact reordering, no arithmetic): row pointers,
    // column indices ascending within each row, values of M itself.
    void to_csr(std::vector<int32_t>& row_ptr, std::vector<int32_t>& col_idx,
                std::vector<Scalar>& val) const {
        row_ptr.assign(static_cast<size_t>(rows_) + 1, 0);
        for (int32_t r : inner_) row_ptr[r + 1]++;
        for (int32_t i = 0; i < rows_; ++i) row_ptr[i + 1] += row_ptr[i];
        col_idx.resize(inner_.size());
        val.resize(inner_.size());
        std::vector<int32_t> fill(row_ptr.begin(), row_ptr.end() - 1);
        for (int32_t j = 0; j < cols_; ++j)
            for (int32_t p = outer_[j]; p < outer_[j + 1]; ++p) {
                int32_t q = fill[inner_[p]]++;
                col_idx[q] = j;
                val[q] = values_[p];
            }
    }

private:
    int32_t rows_ = 0, cols_ = 0;
    std::vector<int32_t> outer_{0};
    std::vector<int32_t> inner_;
    std::vector<Scalar> values_;
};

// ----------------------------------------------------- MatrixXcd (col-major)
class MatrixXcd {
public:
    MatrixXcd() = default;
    MatrixXcd(int64_t rows, int64_t cols)
        : rows_(rows), cols_(cols), data_(static_cast<size_t>(rows * cols)) {}
    int64_t rows() const { return rows_; }
    int64_t cols() const { return cols_; }
    int64_t size() const { return rows_ * cols_; }
    Complex* data() { return data_.data(); }
    const Complex* data() const { return data_.data(); }
    Complex& operator()(int64_t i, int64_t j) { return data_[static_cast<size_t>(j * rows_ + i)]; }
    const Complex& operator()(int64_t i, int64_t j) const {
        return data_[static_cast<size_t>(j * rows_ + i)];
    }
    void setZero() { std::fill(data_.begin(), data_.end(), Complex(0.0, 0.0)); }
    void setIdentity() {
        setZero();
        for (int64_t i = 0; i < std::min(rows_, cols_); ++i) (*this)(i, i) = 1.0;
    }

private:
    int64_t rows_ = 0, cols_ = 0;
    std::vector<Complex> data_;
};

}  // namespace tb
