"""CPU oracle for Hamiltonian assembly -- TEST INFRASTRUCTURE ONLY.

A restatement of the reference's model and assembly path.  The tests, the
smoke check and the benchmark's CPU leg use it as the checker; the product
library never imports it.

Restated (reference file:line):
  UnitCell ctor / add_hopping / add_hopping_direct   unit_cell.h:188-256, 527-547
  UnitCellNN.generate_hoppings_to / add_cell          unit_cell.h:680-707
    (std::priority_queue restated with the push_heap / pop_heap steps of
     libstdc++'s bits/stl_heap.h, so equal-distance ties pop in the same order)
  displacement_of_lattice / displacement_of_hopping   unit_cell.h:268-297
  PartialLattice indexing                             lattice.h:27-35, 106-117
  get_Hamiltonian_sparse                              lattice.h:262-305
  get_Hamiltonian_periodic_extend_bare_sparse         lattice.h:216-258
  get_Hamiltonian_bare_sparse                         lattice.h:347-377
  op_position / op_velocity                           lattice.h:314-345
  set_scalar_field                                    lattice.h:496-505
  GreenFunction::scale  H~ = (H - bI)/a               lattice.h:550-559
  Eigen::SparseMatrix::setFromTriplets: duplicates summed in insertion order,
  column-major result with ascending row indices.

Two choices replace behaviour the reference leaves unspecified (both are the
same in the product): hoppings iterate in insertion order, and disorder comes
from a counter-based generator keyed by the bare site index
(uniform_disorder).  Peierls phases use Python's math.cos / math.sin, i.e.
the C library's, to match std::exp of a pure-imaginary argument bit for bit.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

MASK64 = (1 << 64) - 1


def _cplx(re, im):
    """complex128 array from real and imaginary parts, bits kept (re + 1j*im
    would turn an imaginary -0.0 into +0.0)."""
    re = np.asarray(re, dtype=np.float64)
    out = np.empty(re.shape, dtype=np.complex128)
    out.real = re
    out.imag = im
    return out


# ------------------------------------------------------------------ generators
def splitmix64(x):
    """splitmix64 finaliser on uint64 numpy arrays (wrapping arithmetic)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return x


def counter_uniform(seed: int, counter):
    key = splitmix64(np.uint64(seed & MASK64))
    h = splitmix64(key ^ np.asarray(counter, dtype=np.uint64))
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform_disorder(W: float, seed: int):
    """V(site) = W*(u - 1/2), u = counter_uniform(seed, site)."""

    def f(site_index):
        return W * (counter_uniform(seed, site_index) - 0.5)

    return f


def random_phase_kets(rows: int, R: int, seed: int):
    """e^{i 2 pi u}/sqrt(rows), u = counter_uniform(seed, i*R + r); column-major
    complex128 (rows, R).  Same generator as the device kernel; the phases are
    evaluated with numpy here, so tests pass these arrays to both sides."""
    i = np.arange(rows, dtype=np.uint64)[:, None]
    r = np.arange(R, dtype=np.uint64)[None, :]
    u = counter_uniform(seed, i * np.uint64(R) + r)
    th = 2.0 * math.pi * u
    s = math.sqrt(rows)
    return _cplx(np.cos(th) / s, np.sin(th) / s)


# ------------------------------------------------------- libstdc++ binary heap
class StdPriorityQueue:
    """std::priority_queue<T, vector<T>, Cmp> with libstdc++'s heap steps.
    less(a, b) is the comparator (true if a has lower priority than b)."""

    def __init__(self, less):
        self.c = []
        self.less = less

    def _push_heap_hole(self, hole, top, value):
        c, less = self.c, self.less
        parent = (hole - 1) // 2
        while hole > top and less(c[parent], value):
            c[hole] = c[parent]
            hole = parent
            parent = (hole - 1) // 2
        c[hole] = value

    def push(self, value):
        self.c.append(value)
        self._push_heap_hole(len(self.c) - 1, 0, value)

    def top(self):
        return self.c[0]

    def pop(self):
        c, less = self.c, self.less
        n = len(c)
        if n > 1:
            last = n - 1
            value = c[last]
            c[last] = c[0]
            # __adjust_heap(first, 0, len = last, value)
            length = last
            hole = 0
            top = 0
            second = 0
            while second < (length - 1) // 2:
                second = 2 * (second + 1)
                if less(c[second], c[second - 1]):
                    second -= 1
                c[hole] = c[second]
                hole = second
            if (length & 1) == 0 and second == (length - 2) // 2:
                second = 2 * (second + 1)
                c[hole] = c[second - 1]
                hole = second - 1
            self._push_heap_hole(hole, top, value)
        c.pop()

    def __len__(self):
        return len(self.c)


# ----------------------------------------------------------------- unit cell
def _lat_disp(bases, a):
    """bases * a with left-to-right accumulation (scalar)."""
    out = []
    for i in range(3):
        s = bases[0][i] * float(a[0])
        for j in range(1, len(a)):
            s = s + bases[j][i] * float(a[j])
        out.append(s)
    return out


class UnitCell:
    def __init__(self, dim, lattice_vectors, atoms):
        if len(lattice_vectors) != dim:
            raise ValueError("Number of lattice vectors must be equal to the dimension")
        if not atoms:
            raise ValueError("Unit cell must contain at least one atom")
        self.dim = dim
        self.bases = [tuple(float(x) for x in v) for v in lattice_vectors]
        self.atom_no = {}
        self.atom_pos = []
        for i, (name, pos) in enumerate(atoms):
            if name in self.atom_no:
                raise ValueError("Atom %s is already defined" % name)
            self.atom_no[name] = i
            self.atom_pos.append(tuple(float(x) for x in pos))
        self.hoppings = []  # [((from, to, dR), t)] in insertion order
        self._index = {}
        self._gen = {}

    def copy(self):
        c = UnitCell.__new__(UnitCell)
        c.dim, c.bases, c.atom_no, c.atom_pos = self.dim, self.bases, dict(self.atom_no), list(self.atom_pos)
        c.hoppings = list(self.hoppings)
        c._index = dict(self._index)
        c._gen = {}
        return c

    @property
    def n_atoms(self):
        return len(self.atom_pos)

    def displacement_of_lattice(self, a):
        return _lat_disp(self.bases, a)

    def displacement_of_hopping(self, hp):
        frm, to, lv = hp
        L = _lat_disp(self.bases, lv)
        pf, pt = self.atom_pos[frm], self.atom_pos[to]
        return [((-pf[i]) + L[i]) + pt[i] for i in range(3)]

    def distance_of_hopping(self, hp):
        d = self.displacement_of_hopping(hp)
        s = d[0] * d[0]
        s = s + d[1] * d[1]
        s = s + d[2] * d[2]
        return math.sqrt(s)

    @staticmethod
    def hopping_hc(hp):
        frm, to, lv = hp
        return (to, frm, tuple(-x for x in lv))

    def hopping_exists(self, hp):
        return hp in self._index or self.hopping_hc(hp) in self._index

    def add_hopping_direct(self, hp, t):
        hp = (hp[0], hp[1], tuple(hp[2]))
        t = complex(t)
        if self.hopping_exists(hp):
            return False
        frm, to, lv = hp
        if frm == to and not any(lv) and t.imag != 0:
            return False
        if frm < to:
            hp, t = self.hopping_hc(hp), t.conjugate()
        self._index[hp] = len(self.hoppings)
        self.hoppings.append((hp, t))
        return True

    def add_hopping(self, frm, to, lv, t):
        t = complex(t)
        if frm == to and not any(lv) and t.imag != 0:
            return self
        self.add_hopping_direct((self.atom_no[frm], self.atom_no[to], tuple(lv)), t)
        return self

    def add_hopping_nn(self, frm, k, t):
        t = complex(t)
        a = self.atom_no[frm]
        if k == 0:
            if t.imag != 0:
                return self
            self.add_hopping_direct((a, a, (0,) * self.dim), t)
            return self
        gen = self._gen.setdefault(a, UnitCellNN(a))
        gen.generate_hoppings_to(self, k)
        for hp in gen.NN_solved[k]:
            self.add_hopping_direct(hp, t)
        return self

    @staticmethod
    def nn_vectors(dim):
        out = []
        for c in range(3 ** dim):
            v, r = [0] * dim, c
            for i in range(dim - 1, -1, -1):
                v[i] = (r % 3) - 1
                r //= 3
            if any(v):
                out.append(tuple(v))
        return out


class UnitCellNN:
    def __init__(self, atom):
        self.A = atom
        # comparator of the reference: n > m on distance => min-heap
        self.heap = StdPriorityQueue(lambda n, m: n[0] > m[0])
        self.registered = set()
        self.NN_solved = []

    def add_cell(self, cell, a):
        a = tuple(a)
        if a in self.registered:
            return
        self.registered.add(a)
        for B in range(cell.n_atoms):
            hp = (self.A, B, a)
            self.heap.push((cell.distance_of_hopping(hp), hp))

    def generate_hoppings_to(self, cell, NN):
        if not self.registered:
            self.add_cell(cell, (0,) * cell.dim)
        nnv = UnitCell.nn_vectors(cell.dim)
        while len(self.NN_solved) <= NN:
            d, hp = self.heap.top()
            d0 = d
            shell = []
            while abs(d - d0) < 1e-7:
                for dx in nnv:
                    self.add_cell(cell, tuple(x + y for x, y in zip(hp[2], dx)))
                shell.append(hp)
                self.heap.pop()
                d, hp = self.heap.top()
            self.NN_solved.append(shell)


# ---------------------------------------------------------------- presets
def model(name):
    s3 = math.sqrt(3.0)
    if name == "honeycomb":
        return UnitCell(2, [(3.0 / 2, -s3 / 2, 0), (3.0 / 2, s3 / 2, 0)], [("A", (0, 0, 0)), ("B", (1, 0, 0))])
    if name == "lieb":
        return UnitCell(2, [(1, 0, 0), (0, 1, 0)], [("A", (0, 0, 0)), ("B", (0.5, 0, 0)), ("C", (0, 0.5, 0))])
    if name == "kagome":
        return UnitCell(2, [(1, 0, 0), (0.5, s3 / 2, 0)],
                        [("A", (0, 0, 0)), ("B", (0.5, 0, 0)), ("C", (0.25, s3 / 4, 0))])
    if name == "square":
        return UnitCell(2, [(1, 0, 0), (0, 1, 0)], [("A", (0, 0, 0))])
    if name == "chain":
        return UnitCell(1, [(1, 0, 0)], [("A", (0, 0, 0))])
    if name == "fcc":
        return UnitCell(3, [(0.5, 0.5, 0), (0.5, 0, 0.5), (0, 0.5, 0.5)], [("A", (0, 0, 0))])
    if name == "cubic":
        return UnitCell(3, [(1, 0, 0), (0, 1, 0), (0, 0, 1)], [("A", (0, 0, 0))])
    raise KeyError(name)


# --------------------------------------------------------------- gauges
def landau_gauge(B):
    """A(r) = (0, B*x, 0)."""

    def f(pos):
        x = pos[:, 0]
        z = np.zeros_like(x)
        return np.stack([z, B * x, z], axis=1)

    return f


def symmetric_gauge(B):
    """A(r) = B*(-y, x, 0)/2 evaluated as (B*(-y))/2, (B*x)/2, (B*0)/2."""

    def f(pos):
        x, y = pos[:, 0], pos[:, 1]
        return np.stack([(B * (-y)) / 2, (B * x) / 2, (B * np.zeros_like(x)) / 2], axis=1)

    return f


# -------------------------------------------------------------- sparse CSC
@dataclass
class CSC:
    n: int
    colptr: np.ndarray
    rowidx: np.ndarray
    values: np.ndarray  # float64 or complex128

    @property
    def nnz(self):
        return int(self.rowidx.size)

    def to_dense(self):
        D = np.zeros((self.n, self.n), dtype=self.values.dtype)
        cols = np.repeat(np.arange(self.n), np.diff(self.colptr))
        D[self.rowidx, cols] = self.values
        return D

    def to_csr(self):
        """Exact reordering into CSR (row ptr, col idx ascending, values of M)."""
        cols = np.repeat(np.arange(self.n, dtype=np.int32), np.diff(self.colptr))
        order = np.lexsort((cols, self.rowidx))
        rp = np.zeros(self.n + 1, dtype=np.int32)
        np.cumsum(np.bincount(self.rowidx, minlength=self.n), out=rp[1:])
        return rp, cols[order].astype(np.int32), self.values[order]


def set_from_triplets(n, rows, cols, vals):
    """Eigen setFromTriplets on triplets given in insertion order."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals)
    ins = np.arange(rows.size)
    order = np.lexsort((ins, cols, rows))  # by row, then col, then insertion
    r, c, v = rows[order], cols[order], vals[order]
    if r.size:
        new = np.ones(r.size, dtype=bool)
        new[1:] = (r[1:] != r[:-1]) | (c[1:] != c[:-1])
        starts = np.flatnonzero(new)
        sizes = np.diff(np.append(starts, r.size))
        out = v[starts].copy()
        for k in range(1, int(sizes.max())):
            g = sizes > k
            out[g] = out[g] + v[starts[g] + k]
        r, c, v = r[starts], c[starts], out
    o2 = np.lexsort((r, c))
    r, c, v = r[o2], c[o2], v[o2]
    colptr = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(np.bincount(c, minlength=n), out=colptr[1:])
    return CSC(n, colptr, r.astype(np.int32), v)


def _cos_sin_exact(theta):
    """(cos t, sin t) from the C library, one call per distinct bit pattern
    (so that -0.0 and +0.0 keep their own sine)."""
    bits = np.ascontiguousarray(theta, dtype=np.float64).view(np.uint64)
    ub, inv = np.unique(bits, return_inverse=True)
    u = ub.view(np.float64)
    cs = np.array([math.cos(t) for t in u.tolist()], dtype=np.float64)
    sn = np.array([math.sin(t) for t in u.tolist()], dtype=np.float64)
    return cs[inv], sn[inv]


class PartialLattice:
    def __init__(self, cell: UnitCell, span):
        self.cell = cell.copy()
        self.dim = cell.dim
        self.span_v = [int(x) for x in span]
        self.cell_number = int(np.prod(self.span_v))
        self.total_number = self.cell_number * self.cell.n_atoms
        ss = [0] * self.dim
        ss[-1] = self.cell.n_atoms
        for i in range(self.dim - 2, -1, -1):
            ss[i] = ss[i + 1] * self.span_v[i + 1]
        self.span_s = ss
        self.gauge_field = None
        self.scalar_field = None  # function of (positions array) -> V
        self.site_field = None  # function of (bare index array) -> V
        self.removed = []
        self.index_c = None

    def _ensure_onsite(self):
        for i in range(self.cell.n_atoms):
            hp = (i, i, (0,) * self.dim)
            if not self.cell.hopping_exists(hp):
                self.cell.add_hopping_direct(hp, 0.0)

    def set_scalar_field(self, f):
        self.scalar_field = f
        self._ensure_onsite()

    def set_scalar_field_indexed(self, f):
        self.site_field = f
        self._ensure_onsite()

    def remove_atoms(self, atom_name, cells):
        a = self.cell.atom_no[atom_name]
        for v in cells:
            self.removed.append(int(np.dot(v, self.span_s)) + a)
        self._index_comp_construct()

    def _index_comp_construct(self):
        ic = np.zeros(self.total_number, dtype=np.int64)
        ic[np.asarray(self.removed, dtype=np.int64)] = -1
        keep = ic != -1
        ic[keep] = np.arange(int(keep.sum()))
        self.index_c = ic

    def _cells(self):
        c = np.arange(self.cell_number, dtype=np.int64)
        iv = np.zeros((self.cell_number, self.dim), dtype=np.int64)
        for d in range(self.dim - 1, -1, -1):
            iv[:, d] = c % self.span_v[d]
            c = c // self.span_v[d]
        return iv

    def _positions(self, iv, atom):
        """atom_pos[atom] + bases*i_v, per component, left-to-right."""
        B = self.cell.bases
        p = self.cell.atom_pos[atom]
        out = np.empty((iv.shape[0], 3))
        for i in range(3):
            s = B[0][i] * iv[:, 0].astype(np.float64)
            for j in range(1, self.dim):
                s = s + B[j][i] * iv[:, j].astype(np.float64)
            out[:, i] = p[i] + s
        return out

    def _assemble(self, mode, is_complex, extend=None):
        if mode == "open" and self.index_c is None:
            self.index_c = np.arange(self.total_number, dtype=np.int64)
        iv = self._cells()
        ncell = self.cell_number
        nh = len(self.cell.hoppings)
        span = np.asarray(self.span_v, dtype=np.int64)
        ss = np.asarray(self.span_s, dtype=np.int64)
        R, C, V, K = [], [], [], []
        fields = mode != "bare"
        for h, (hp, t0) in enumerate(self.cell.hoppings):
            frm, to, lv = hp
            to_v = iv + np.asarray(lv, dtype=np.int64)[None, :]
            inside = np.all((to_v >= 0) & (to_v < span[None, :]), axis=1)
            cidx = np.arange(ncell, dtype=np.int64)
            if mode == "periodic":
                ext = np.asarray(extend, dtype=bool)
                keep = np.all(ext[None, :] | ((to_v >= 0) & (to_v < span[None, :])), axis=1) | inside
                m = to_v - span * np.trunc(to_v / span).astype(np.int64)  # C++ truncating mod
                m = m + span
                m = m - span * (m // span)
                row = iv @ ss + frm
                col = m @ ss + to
            else:
                keep = inside
                row = iv @ ss + frm
                col = to_v @ ss + to
                if mode == "open":
                    rc = self.index_c[row]
                    cc = self.index_c[np.where(keep, col, 0)]
                    keep = keep & (rc != -1) & (cc != -1)
                    row, col = rc, np.where(keep, cc, 0)
            row, col, cidx = row[keep], col[keep], cidx[keep]
            ivk = iv[keep]
            tre = np.full(row.size, t0.real)
            tim = np.full(row.size, t0.imag)
            if fields and self.gauge_field is not None:
                d = self.cell.displacement_of_hopping(hp)
                A = self.gauge_field(self._positions(ivk, frm))
                theta = (A[:, 0] * d[0] + A[:, 1] * d[1]) + A[:, 2] * d[2]
                cs, sn = _cos_sin_exact(theta)
                tre, tim = tre * cs - tim * sn, tre * sn + tim * cs
            if fields:
                diag = row == col
                if self.scalar_field is not None and diag.any():
                    tre = tre.copy()
                    tre[diag] = tre[diag] + self.scalar_field(self._positions(ivk[diag], frm))
                if self.site_field is not None and diag.any():
                    tre = tre.copy()
                    bare = ivk[diag] @ ss + frm
                    tre[diag] = tre[diag] + self.site_field(bare)
            tval = _cplx(tre, tim) if is_complex else tre.copy()
            if not is_complex and np.any(np.abs(tim) > 1e-10):
                pass  # the reference warns and keeps the real part
            base = cidx * (2 * nh) + 2 * h
            R.append(row)
            C.append(col)
            V.append(tval)
            K.append(base)
            off = row != col
            R.append(col[off])
            C.append(row[off])
            V.append(np.conj(tval[off]) if is_complex else tval[off])
            K.append(base[off] + 1)
        dt = np.complex128 if is_complex else np.float64
        if R:
            rows, cols, vals, keys = (np.concatenate(R), np.concatenate(C),
                                      np.concatenate(V).astype(dt), np.concatenate(K))
        else:
            rows = cols = keys = np.zeros(0, dtype=np.int64)
            vals = np.zeros(0, dtype=dt)
        order = np.argsort(keys, kind="stable")
        return set_from_triplets(self.total_number, rows[order], cols[order], vals[order])

    def get_Hamiltonian_sparse(self, is_complex=False):
        return self._assemble("open", is_complex)

    def get_Hamiltonian_periodic_extend_bare_sparse(self, extend, is_complex=False):
        return self._assemble("periodic", is_complex, extend)

    def get_Hamiltonian_bare_sparse(self, is_complex=False):
        return self._assemble("bare", is_complex)

    def op_position(self, axis):
        if self.index_c is None:
            self.index_c = np.arange(self.total_number, dtype=np.int64)
        iv = self._cells()
        rows, vals, keys = [], [], []
        for a in range(self.cell.n_atoms):
            bare = iv @ np.asarray(self.span_s) + a
            r = self.index_c[bare]
            ok = r != -1
            # a_pos(axis) + displacement_of_lattice(i)(axis)
            s = self.cell.bases[0][axis] * iv[:, 0].astype(np.float64)
            for j in range(1, self.dim):
                s = s + self.cell.bases[j][axis] * iv[:, j].astype(np.float64)
            pos = self.cell.atom_pos[a][axis] + s
            rows.append(r[ok])
            vals.append(pos[ok])
            keys.append(np.arange(self.cell_number)[ok] * self.cell.n_atoms + a)
        rows, vals, keys = np.concatenate(rows), np.concatenate(vals), np.concatenate(keys)
        o = np.argsort(keys, kind="stable")
        return set_from_triplets(self.total_number, rows[o], rows[o], vals[o])

    def op_velocity(self, axis, is_complex=False):
        X = self.op_position(axis)
        H = self.get_Hamiltonian_sparse(is_complex)
        return commutator_with_diagonal(X, H)


def commutator_with_diagonal(X: CSC, M: CSC):
    """X*M - M*X for diagonal X: entries x_i*M_ij - M_ij*x_j on M's pattern."""
    xd = np.zeros(M.n)
    cols = np.repeat(np.arange(X.n), np.diff(X.colptr))
    xd[X.rowidx] = X.values  # diagonal: rowidx == cols
    assert np.array_equal(X.rowidx, cols)
    mc = np.repeat(np.arange(M.n), np.diff(M.colptr))
    xi, xj = xd[M.rowidx], xd[mc]
    v = M.values
    if np.iscomplexobj(v):
        re = xi * v.real - v.real * xj
        im = xi * v.imag - v.imag * xj
        out = _cplx(re, im)
    else:
        out = xi * v - v * xj
    return CSC(M.n, M.colptr.copy(), M.rowidx.copy(), out)


def rescale(H: CSC, a: float, b: float) -> CSC:
    """H~ = (H - bI)/a with the union pattern of H and I (lattice.h:550-559)."""
    n = H.n
    cols = np.repeat(np.arange(n), np.diff(H.colptr))
    rows = H.rowidx.astype(np.int64)
    has_diag = np.zeros(n, dtype=bool)
    has_diag[cols[rows == cols]] = True
    miss = np.flatnonzero(~has_diag)
    allr = np.concatenate([rows, miss])
    allc = np.concatenate([cols, miss])
    cplx = np.iscomplexobj(H.values)
    if cplx:
        vr = np.concatenate([H.values.real, np.zeros(miss.size)])
        vi = np.concatenate([H.values.imag, np.zeros(miss.size)])
        sub = np.where(allr == allc, b, 0.0)
        vr = (vr - sub) / a
        vi = (vi - 0.0) / a
        vals = _cplx(vr, vi)
    else:
        vr = np.concatenate([H.values, np.zeros(miss.size)])
        sub = np.where(allr == allc, b, 0.0)
        vals = (vr - sub) / a
    o = np.lexsort((allr, allc))
    colptr = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(np.bincount(allc, minlength=n), out=colptr[1:])
    return CSC(n, colptr, allr[o].astype(np.int32), vals[o])
