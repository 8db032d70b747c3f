// Host side of the MI355X kernel-polynomial engine: contexts, resident
// matrices and vector blocks, the moment driver, and the C ABI of tbkpm.h.
//
// The drop-in entry tbk_cheb_first_kind_seq_braket replaces the two
// overloads of cuda::Cheb_first_kind_seq_braket (numerics_cuda.cu:190-323):
// same inputs (rescaled CSC matrix, bras_T, kets, n) and output (n moments,
// not divided by R), with the per-step host synchronisation, per-step
// printing and cuSPARSE/cuBLAS calls of the reference replaced by one fused
// HIP kernel per step whose moment partials stay on the device until the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "aux_kernels.h"
#include "cheb_kernels.h"
#include "gamma_launch.h"
#include "engine.h"
#include "step_launch.h"
#include "tbkpm.h"

namespace engine {

// --------------------------------------------------------- kernel dispatch

void launch_any(int precision, bool cplx, int width, int kch, int cpl, int mode, dim3 grid, hipStream_t st,
                const StepArgs& a) {
    bool ok;
    if (precision == TBK_FP64)
        ok = cplx ? tbk::launch_f64_cplx(width, kch, cpl, mode, grid, st, a)
                  : tbk::launch_f64_real(width, kch, cpl, mode, grid, st, a);
    else
        ok = cplx ? tbk::launch_f32_cplx(width, kch, cpl, mode, grid, st, a)
                  : tbk::launch_f32_real(width, kch, cpl, mode, grid, st, a);
    if (!ok) raise(TBK_E_UNSUPPORTED, "no step kernel for this block width");
}

KernelChoice choose_kernel(const tbk_tuning& tu, int precision, bool cplx, int Rp, int max_row_nnz,
                           int bandwidth, bool force_row, bool row_slabs) {
    const int hbytes = (precision == TBK_FP64 ? 8 : 4) * (cplx ? 2 : 1);
    const int ebytes = precision == TBK_FP64 ? 16 : 8;
    int kch = max_row_nnz <= 4 ? 1 : max_row_nnz <= 8 ? 2 : max_row_nnz <= 16 ? 4 : 0;
    if (force_row) kch = 0;
    int sw = Rp;
    if (kch > 0) {
        if (tu.slab_cols > 0) {
            sw = 1;
            while (sw * 2 <= std::min<int>(tu.slab_cols, Rp)) sw *= 2;
        } else {
            // narrowest slab of >= 16 columns (or Rp) whose gather window
            // 3 * 2D * SW elements fits the L2 budget
            const double budget = tu.l2_budget_bytes > 0 ? tu.l2_budget_bytes : 3.0 * 1024 * 1024;
            while (sw > 16 && 6.0 * bandwidth * sw * ebytes > budget) sw /= 2;
        }
        const int lpr = std::min(sw, tbk::kWave);
        if (!tbk::tile_fits(lpr, sw / lpr, kch, hbytes)) {
            kch = 0;
            sw = Rp;
        }
    }
    if (kch == 0) {
        // row kernel: cpl complex values per lane (at least Rp/64), LPR = Rp / cpl;
        // lanes of row_lane_bytes, 16 by default (one complex128 or two
        // complex64): on gfx950 that is 0.61-0.83x the fp32 step time of one
        // complex64 per lane at R = 8..64, and wider lanes lose in fp64.  A
        // width in bytes means the same in both precisions, so a tuning found
        // on one carries over to the other.
        // A sweep covers W columns: Rp, or row_slab_cols for the moment
        // recursions, whose grid then visits every row once per sweep.
        int W = Rp;
        if (row_slabs && tu.row_slab_cols >= 2 && tu.row_slab_cols < Rp &&
            (tu.row_slab_cols & (tu.row_slab_cols - 1)) == 0)
            W = tu.row_slab_cols;
        const int cmin = W > tbk::kWave ? W / tbk::kWave : 1;
        const int lane_bytes = tu.row_lane_bytes > 0 ? tu.row_lane_bytes : 16;
        int cpl = std::max({cmin, 1, lane_bytes / ebytes});
        if (cpl != 1 && cpl != 2 && cpl != 4) cpl = cmin;
        if (cpl > W || W % cpl != 0 || (cpl > cmin && W * 8 + cpl > 128 * 8 + 4)) cpl = cmin;
        const int lpr = W / cpl;
        return {0, W, tbk::kWavesPerBlock * (tbk::kWave / lpr), cpl};
    }
    const int lpr = std::min(sw, tbk::kWave);
    const int tr = tbk::tile_rows(lpr, sw / lpr, kch, hbytes);
    return {kch, sw, tr, sw / lpr};
}

// ---------------------------------------------------------- matrix upload
// p-quantile of |row - col| over the stored entries (the distance in rows at
// which a gathered neighbour row is used again)
int bandwidth_percentile(const std::vector<int32_t>& rp, const std::vector<int32_t>& ci, double p) {
    if (ci.empty()) return 0;
    std::vector<int32_t> d(ci.size());
    for (size_t i = 0; i + 1 < rp.size(); ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) d[k] = std::abs(ci[k] - static_cast<int32_t>(i));
    const size_t at = std::min(d.size() - 1, static_cast<size_t>(p * static_cast<double>(d.size())));
    std::nth_element(d.begin(), d.begin() + static_cast<std::ptrdiff_t>(at), d.end());
    return d[at];
}

// Gershgorin interval of a resident matrix, formed on the device on first use
// (the union over rows of [Re h_ii - r_i, Re h_ii + r_i], r_i = sum_{j != i} |h_ij|)
void matrix_gershgorin(const tbk_matrix* M, hipStream_t st) {
    auto* m = const_cast<tbk_matrix*>(M);
    if (m->have_gersh) return;
    const unsigned grid =
        static_cast<unsigned>(std::min<int64_t>((static_cast<int64_t>(m->rows) + tbk::kBlock - 1) / tbk::kBlock, 1024));
    DevBuf part;
    part.reserve(static_cast<size_t>(grid) * 16);
    hipLaunchKernelGGL(tbk::gershgorin_kernel, dim3(grid), dim3(tbk::kBlock), 0, st, m->rp.as<const int32_t>(),
                       m->ci.as<const int32_t>(), m->val64.as<const double>(), m->is_complex ? 1 : 0, m->rows,
                       part.as<double2>());
    HIPC(hipGetLastError());
    std::vector<double2> h(grid);
    HIPC(hipMemcpyAsync(h.data(), part.p, static_cast<size_t>(grid) * 16, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    double lo = INFINITY, hi = -INFINITY;
    for (const double2& g : h) {
        lo = std::min(lo, g.x);
        hi = std::max(hi, g.y);
    }
    m->gersh_min = std::isfinite(lo) ? lo : 0.0;
    m->gersh_max = std::isfinite(hi) ? hi : 0.0;
    m->have_gersh = true;
}

// 64-bit fingerprint of a resident matrix's pattern (rows, row pointers,
// columns), formed on the device on first use; never 0
uint64_t matrix_pattern(const tbk_matrix* M, hipStream_t st) {
    auto* m = const_cast<tbk_matrix*>(M);
    if (m->pattern) return m->pattern;
    const int64_t work = std::max<int64_t>(static_cast<int64_t>(m->rows) + 1, m->nnz);
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>((work + tbk::kBlock - 1) / tbk::kBlock, 1024));
    DevBuf part;
    part.reserve(static_cast<size_t>(grid) * 8);
    hipLaunchKernelGGL(tbk::pattern_hash_kernel, dim3(grid), dim3(tbk::kBlock), 0, st, m->rp.as<const int32_t>(),
                       m->ci.as<const int32_t>(), m->rows, m->nnz, part.as<unsigned long long>());
    HIPC(hipGetLastError());
    std::vector<unsigned long long> h(grid);
    HIPC(hipMemcpyAsync(h.data(), part.p, static_cast<size_t>(grid) * 8, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    uint64_t v = static_cast<uint64_t>(m->rows) * 0x9E3779B97F4A7C15ull ^ static_cast<uint64_t>(m->nnz);
    for (unsigned long long x : h) v += x;
    m->pattern = v ? v : 1;
    return m->pattern;
}

void build_csr(const tbk_sparse* M, std::vector<int32_t>& rp, std::vector<int32_t>& ci,
               std::vector<double>& val) {
    const int vw = M->is_complex ? 2 : 1;
    const int64_t nnz = M->nnz;
    if (M->layout == TBK_CSR) {
        rp.assign(M->outer, M->outer + M->rows + 1);
        ci.assign(M->inner, M->inner + nnz);
        val.assign(M->values, M->values + nnz * vw);
        for (int32_t i = 0; i < M->rows; ++i)
            if (rp[i + 1] < rp[i]) raise(TBK_E_INVALID, "row pointers not monotone");
        if (rp[0] != 0 || rp[M->rows] != nnz) raise(TBK_E_INVALID, "row pointers inconsistent with nnz");
        for (int64_t p = 0; p < nnz; ++p)
            if (ci[p] < 0 || ci[p] >= M->cols) raise(TBK_E_INVALID, "column index out of range");
        return;
    }
    // CSC -> CSR: exact stable reordering (columns visited ascending)
    const int32_t* cp = M->outer;
    const int32_t* ri = M->inner;
    if (cp[0] != 0 || cp[M->cols] != nnz) raise(TBK_E_INVALID, "column pointers inconsistent with nnz");
    rp.assign(static_cast<size_t>(M->rows) + 1, 0);
    for (int32_t j = 0; j < M->cols; ++j) {
        if (cp[j + 1] < cp[j]) raise(TBK_E_INVALID, "column pointers not monotone");
    }
    for (int64_t p = 0; p < nnz; ++p) {
        if (ri[p] < 0 || ri[p] >= M->rows) raise(TBK_E_INVALID, "row index out of range");
        rp[ri[p] + 1]++;
    }
    for (int32_t i = 0; i < M->rows; ++i) rp[i + 1] += rp[i];
    ci.resize(nnz);
    val.resize(static_cast<size_t>(nnz) * vw);
    std::vector<int32_t> fill(rp.begin(), rp.end() - 1);
    for (int32_t j = 0; j < M->cols; ++j)
        for (int32_t p = cp[j]; p < cp[j + 1]; ++p) {
            const int32_t q = fill[ri[p]]++;
            ci[q] = j;
            for (int k = 0; k < vw; ++k) val[static_cast<size_t>(q) * vw + k] = M->values[static_cast<size_t>(p) * vw + k];
        }
}

// M == M^H exactly (values compared bit for bit after conjugation), from the
// CSR arrays: the counting transpose of (rp, ci, conj val), rows visited in
// order, equals (rp, ci, val).  A row whose columns are not ascending makes
// the two orders differ and reads as not Hermitian, which only costs the
// faster path.
bool csr_is_hermitian(int32_t n, const std::vector<int32_t>& rp, const std::vector<int32_t>& ci,
                      const std::vector<double>& val, bool cplx) {
    const int vw = cplx ? 2 : 1;
    const size_t nnz = ci.size();
    std::vector<int32_t> tp(static_cast<size_t>(n) + 1, 0);
    for (size_t p = 0; p < nnz; ++p) tp[ci[p] + 1]++;
    for (int32_t i = 0; i < n; ++i) tp[i + 1] += tp[i];
    if (tp != rp) return false;
    std::vector<int32_t> fill(tp.begin(), tp.end() - 1);
    for (int32_t r = 0; r < n; ++r)
        for (int32_t p = rp[r]; p < rp[r + 1]; ++p) {
            const int32_t q = fill[ci[p]]++;
            if (ci[q] != r) return false;
            const double* a = &val[static_cast<size_t>(p) * vw];
            const double* b = &val[static_cast<size_t>(q) * vw];
            if (a[0] != b[0]) return false;
            if (cplx && a[1] != -b[1]) return false;
        }
    return true;
}

// M^H = M to within tol of the largest magnitude, pattern included (a
// structural mismatch or an asymmetric pair beyond tol: false).  Rows are
// put in column order first, so any order of a row's entries is accepted.
bool csr_hermitian_within(int32_t n, const std::vector<int32_t>& rp, const std::vector<int32_t>& ci_in,
                          const std::vector<double>& val_in, bool cplx, double tol) {
    const int vw = cplx ? 2 : 1;
    const size_t nnz = ci_in.size();
    std::vector<int32_t> ci(ci_in);
    std::vector<double> val(val_in);
    {
        std::vector<std::pair<int32_t, int64_t>> row;
        for (int32_t r = 0; r < n; ++r) {
            row.clear();
            for (int64_t p = rp[r]; p < rp[r + 1]; ++p) row.emplace_back(ci_in[p], p);
            std::stable_sort(row.begin(), row.end(),
                             [](const auto& x, const auto& y) { return x.first < y.first; });
            for (size_t k = 0; k < row.size(); ++k) {
                const int64_t dst = rp[r] + static_cast<int64_t>(k), src = row[k].second;
                ci[dst] = row[k].first;
                for (int c = 0; c < vw; ++c) val[dst * vw + c] = val_in[src * vw + c];
            }
        }
    }
    double amax = 0.0;
    for (double v : val) amax = std::max(amax, std::fabs(v));
    const double lim = tol * amax;
    std::vector<int32_t> tp(static_cast<size_t>(n) + 1, 0);
    for (size_t p = 0; p < nnz; ++p) tp[ci[p] + 1]++;
    for (int32_t i = 0; i < n; ++i) tp[i + 1] += tp[i];
    if (tp != rp) return false;
    std::vector<int32_t> fill(tp.begin(), tp.end() - 1);
    for (int32_t r = 0; r < n; ++r)
        for (int32_t p = rp[r]; p < rp[r + 1]; ++p) {
            const int32_t q = fill[ci[p]]++;
            if (ci[q] != r) return false;
            const double* a = &val[static_cast<size_t>(p) * vw];
            const double* b = &val[static_cast<size_t>(q) * vw];
            if (!(std::fabs(a[0] - b[0]) <= lim)) return false;
            if (cplx && !(std::fabs(a[1] + b[1]) <= lim)) return false;
        }
    return true;
}

void matrix_fp32(tbk_matrix* m) {
    if (m->have32) return;
    const int64_t n = m->nnz * (m->is_complex ? 2 : 1);
    m->val32.reserve_placed(std::max<size_t>(static_cast<size_t>(n) * 4, 4));
    if (n > 0) {
        const int grid = static_cast<int>(std::min<int64_t>((n + tbk::kBlock - 1) / tbk::kBlock, 65536));
        hipLaunchKernelGGL(tbk::to_fp32_kernel, dim3(grid), dim3(tbk::kBlock), 0, m->ctx->stream,
                           m->val64.as<const double>(), m->val32.as<float>(), n);
        HIPC(hipGetLastError());
    }
    m->ctx->held_bytes += static_cast<int64_t>(m->val32.bytes);
    m->have32 = true;
}

StepPlan make_plan(tbk_context* ctx, const tbk_matrix* M, int prec, int Rp, int rows, double b, bool force_row,
                   const tbk_tuning& tu, bool row_slabs) {
    StepPlan pl{};
    pl.precision = prec;
    pl.cplx = M->is_complex;
    if (prec == TBK_FP32) matrix_fp32(const_cast<tbk_matrix*>(M));
    pl.kc = choose_kernel(tu, prec, M->is_complex, Rp, M->max_row_nnz, M->bandwidth95,
                          force_row || tu.use_tiled <= 0, row_slabs);
    const KernelChoice& kc = pl.kc;
    const int TR = kc.rows_per_tile;
    const int64_t row_tiles = (static_cast<int64_t>(rows) + TR - 1) / TR;
    // chunked visiting order: chunks of ~chunk_rows rows, transposed with the
    // period P = (matrix bandwidth) / (chunk rows) when that exceeds 1
    tbk::TileOrder order{};
    {
        const int64_t chunk_rows = tu.chunk_rows > 0 ? tu.chunk_rows : 256;
        const int64_t tpc = std::max<int64_t>(1, chunk_rows / TR);
        int64_t P = 1;
        if (tu.chunk_period > 0)
            P = tu.chunk_period;
        else if (tu.chunk_period == 0)
            P = std::max<int64_t>(1, (M->bandwidth95 + (tpc * TR) / 2) / (tpc * TR));
        const int64_t nchunks = (row_tiles + tpc - 1) / tpc;
        P = std::min<int64_t>(P, nchunks);
        const int64_t Q = (nchunks + P - 1) / P;
        if (P * Q * tpc * (Rp / kc.width) >= (int64_t(1) << 31)) raise(TBK_E_UNSUPPORTED, "too many row tiles");
        order.row_tiles = static_cast<uint32_t>(row_tiles);
        order.tpc = static_cast<uint32_t>(tpc);
        order.P = static_cast<uint32_t>(P);
        order.Q = static_cast<uint32_t>(Q);
    }
    const int64_t ntiles = static_cast<int64_t>(order.P) * order.Q * order.tpc * (Rp / kc.width);
    const int bpc = tu.blocks_per_cu > 0 ? tu.blocks_per_cu : 4;
    int64_t cap = static_cast<int64_t>(ctx->num_cu) * bpc;
    int64_t nb = std::min<int64_t>(ntiles, cap);
    int xcd_split = 0;
    if ((nb >= 64 && tu.xcd_split >= 0) || (tu.xcd_split > 0 && nb >= 8)) {
        nb = (nb / 8) * 8;
        xcd_split = 1;
    }
    pl.nb = std::max<int64_t>(nb, 1);
    StepArgs& A = pl.args;
    A.rp = M->rp.as<const int32_t>();
    A.ci = M->ci.as<const int32_t>();
    A.hv = prec == TBK_FP64 ? M->val64.p : M->val32.p;
    A.rows = rows;
    A.Rp = Rp;
    A.opts = ((b != 0.0) ? tbk::kOptShift : 0) | (tu.stream_nt >= 0 ? tbk::kOptStream : 0);
    A.ntiles = ntiles;
    A.order = order;
    A.xcd_split = xcd_split;
    return pl;
}

StepPlan make_plan(tbk_context* ctx, const tbk_matrix* M, int prec, int Rp, int rows, double b, bool force_row) {
    return make_plan(ctx, M, prec, Rp, rows, b, force_row, ctx->tuning, false);
}

StepPlan make_plan(tbk_context* ctx, const tbk_matrix* M, int prec, int Rp, int rows, double b, bool force_row,
                   const tbk_tuning& tu) {
    return make_plan(ctx, M, prec, Rp, rows, b, force_row, tu, false);
}

// One step: y = alpha*(M x) + beta*x - z (mode kFirst: no z), partial sums of
// the mode into partial[0 .. plan.nb).
void launch_plan(const StepPlan& pl, hipStream_t st, int mode, double alpha, double beta, const void* x,
                 const void* z, void* y, const void* bra, double2* partial) {
    launch_plan(pl, st, mode, alpha, beta, x, z, y, bra, partial, nullptr, nullptr);
}

void launch_plan(const StepPlan& pl, hipStream_t st, int mode, double alpha, double beta, const void* x,
                 const void* z, void* y, const void* bra, double2* partial, const void* hv2, void* y2) {
    launch_plan(pl, st, mode, alpha, beta, x, z, y, bra, partial, hv2, y2, nullptr, 0);
}

void launch_plan(const StepPlan& pl, hipStream_t st, int mode, double alpha, double beta, const void* x,
                 const void* z, void* y, const void* bra, double2* partial, const void* hv2, void* y2, void* y3,
                 int64_t wbs) {
    StepArgs A = pl.args;
    A.alpha = alpha;
    A.beta = beta;
    A.x = x;
    A.z = z;
    A.y = y;
    A.bra = bra;
    A.partial = partial;
    A.hv2 = hv2;
    A.y2 = y2;
    A.y3 = y3;
    A.wbs = wbs;
    launch_any(pl.precision, pl.cplx, pl.kc.width, pl.kc.kch, pl.kc.cpl, mode, dim3(static_cast<unsigned>(pl.nb)), st,
               A);
    HIPC(hipGetLastError());
}

}  // namespace engine

namespace {

using namespace engine;

// ------------------------------------------------------- tuning cache
// Tuned launch shapes by problem statistics, process-wide, optionally saved
// to and loaded from a text file (one entry per line: the key's fields, then
// the tuning's fields, then the time per step in ms).
struct TuningKey {
    int32_t num_cu, rows, max_row_nnz, bandwidth95, Rp, precision, cplx, plain;
    int64_t nnz;
    bool operator==(const TuningKey& o) const {
        return num_cu == o.num_cu && rows == o.rows && max_row_nnz == o.max_row_nnz &&
               bandwidth95 == o.bandwidth95 && Rp == o.Rp && precision == o.precision && cplx == o.cplx &&
               plain == o.plain && nnz == o.nnz;
    }
};
struct TuningEntry {
    tbk_tuning tuning;
    double ms_per_step;
};
std::mutex g_tuning_mutex;
std::vector<std::pair<TuningKey, TuningEntry>> g_tuning_cache;

TuningKey tuning_key(const tbk_context* ctx, const tbk_matrix* M, const tbk_block* kets, bool plain) {
    return TuningKey{ctx->num_cu, M->rows, M->max_row_nnz, M->bandwidth95, kets->Rp, kets->precision,
                     M->is_complex ? 1 : 0, plain ? 1 : 0, M->nnz};
}

bool tuning_cache_find(const TuningKey& k, TuningEntry& out) {
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    for (const auto& e : g_tuning_cache)
        if (e.first == k) {
            out = e.second;
            return true;
        }
    return false;
}

void tuning_cache_store(const TuningKey& k, const TuningEntry& v) {
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    for (auto& e : g_tuning_cache)
        if (e.first == k) {
            e.second = v;
            return;
        }
    g_tuning_cache.emplace_back(k, v);
}

void tuning_cache_clear() {
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    g_tuning_cache.clear();
}

constexpr int kTuningFields = static_cast<int>(sizeof(tbk_tuning) / sizeof(int32_t));

void tuning_cache_save(const char* path) {
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    FILE* f = std::fopen(path, "w");
    if (!f) raise(TBK_E_INVALID, std::string("cannot write ") + path);
    std::fprintf(f, "tbkpm-tuning 1 %d\n", kTuningFields);
    for (const auto& e : g_tuning_cache) {
        const TuningKey& k = e.first;
        std::fprintf(f, "%d %d %d %d %d %d %d %d %lld", k.num_cu, k.rows, k.max_row_nnz, k.bandwidth95, k.Rp,
                     k.precision, k.cplx, k.plain, static_cast<long long>(k.nnz));
        const int32_t* t = reinterpret_cast<const int32_t*>(&e.second.tuning);
        for (int i = 0; i < kTuningFields; ++i) std::fprintf(f, " %d", t[i]);
        std::fprintf(f, " %.9g\n", e.second.ms_per_step);
    }
    const bool ok = std::fflush(f) == 0;
    std::fclose(f);
    if (!ok) raise(TBK_E_INVALID, std::string("cannot write ") + path);
}

int32_t tuning_cache_load(const char* path) {
    FILE* f = std::fopen(path, "r");
    if (!f) raise(TBK_E_INVALID, std::string("cannot read ") + path);
    struct Closer {
        FILE* f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    int version = 0, fields = 0;
    if (std::fscanf(f, "tbkpm-tuning %d %d", &version, &fields) != 2 || version != 1 || fields != kTuningFields)
        raise(TBK_E_INVALID, std::string(path) + ": not a tuning cache of this library version");
    int32_t count = 0;
    for (;;) {
        TuningKey k{};
        long long nnz = 0;
        const int got = std::fscanf(f, "%d %d %d %d %d %d %d %d %lld", &k.num_cu, &k.rows, &k.max_row_nnz,
                                    &k.bandwidth95, &k.Rp, &k.precision, &k.cplx, &k.plain, &nnz);
        if (got == EOF) break;
        if (got != 9) raise(TBK_E_INVALID, std::string(path) + ": malformed entry");
        k.nnz = nnz;
        TuningEntry v{};
        int32_t* t = reinterpret_cast<int32_t*>(&v.tuning);
        for (int i = 0; i < kTuningFields; ++i)
            if (std::fscanf(f, "%d", &t[i]) != 1) raise(TBK_E_INVALID, std::string(path) + ": malformed entry");
        if (std::fscanf(f, "%lf", &v.ms_per_step) != 1) raise(TBK_E_INVALID, std::string(path) + ": malformed entry");
        tuning_cache_store(k, v);
        ++count;
    }
    return count;
}

void run_moments(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* bras,
                 const tbk_block* kets, int32_t n, int32_t flags, double* mu_out, double* last_out,
                 tbk_stats* stats) {
    auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !M || !kets || !mu_out) raise(TBK_E_INVALID, "null argument");
    if (n < 2) raise(TBK_E_INVALID, "n must be greater than or equal to 2");
    if (!(a != 0.0) || !std::isfinite(a) || !std::isfinite(b)) raise(TBK_E_INVALID, "invalid rescaling a, b");
    if (M->ctx != ctx || kets->ctx != ctx || (bras && bras->ctx != ctx))
        raise(TBK_E_INVALID, "objects belong to another context");
    if (kets->rows != M->rows) raise(TBK_E_INVALID, "kets rows do not match the matrix");
    if (bras && (bras->rows != kets->rows || bras->R != kets->R || bras->precision != kets->precision))
        raise(TBK_E_INVALID, "bras shape or precision does not match kets");
    set_device(ctx);

    const int prec = kets->precision;
    const int Rp = kets->Rp;
    const int rows = kets->rows;
    const size_t block_bytes = static_cast<size_t>(rows) * Rp * kets->elem_bytes();
    ctx->work[0].reserve(block_bytes);
    ctx->work[1].reserve(block_bytes);

    const bool doubling = (bras == nullptr || bras == kets) && !(flags & TBK_FLAG_PLAIN);
    const tbk_block* bra = bras ? bras : kets;
    const int64_t steps = doubling ? (n + 1) / 2 : (n - 1);

    const StepPlan plan = make_plan(ctx, M, prec, Rp, rows, b, (flags & TBK_FLAG_ROW_KERNEL) != 0, ctx->tuning, true);
    const KernelChoice& kc = plan.kc;
    const int TR = kc.rows_per_tile;
    const tbk::TileOrder order = plan.args.order;
    const int64_t nb = plan.nb;
    ctx->partial.reserve(static_cast<size_t>(steps) * nb * sizeof(double2));
    ctx->sums.reserve(static_cast<size_t>(steps) * sizeof(double2));

    const bool time_each = (flags & TBK_FLAG_TIME_LAUNCHES) != 0;
    ensure_events(ctx, time_each ? static_cast<size_t>(2 * steps) : 2);
    double2* partial = ctx->partial.as<double2>();

    auto buf = [&](int64_t k) -> void* {  // storage of T_k
        if (k == 0) return kets->data.p;
        return ctx->work[(k & 1) ? 1 : 0].p;
    };
    if (!time_each) HIPC(hipEventRecord(ctx->events[0], ctx->stream));
    for (int64_t k = 0; k < steps; ++k) {
        if (time_each) HIPC(hipEventRecord(ctx->events[2 * k], ctx->stream));
        if (k == 0)
            launch_plan(plan, ctx->stream, tbk::kFirst, 1.0 / a, -b / a, buf(0), nullptr, buf(1), bra->data.p,
                        partial);
        else
            launch_plan(plan, ctx->stream, doubling ? tbk::kRecDbl : tbk::kRec, 2.0 / a, -2.0 * b / a, buf(k),
                        buf(k - 1), buf(k + 1), bra->data.p, partial + static_cast<size_t>(k) * nb);
        if (time_each) HIPC(hipEventRecord(ctx->events[2 * k + 1], ctx->stream));
    }
    if (!time_each) HIPC(hipEventRecord(ctx->events[1], ctx->stream));
    hipLaunchKernelGGL(tbk::reduce_partials_kernel, dim3(static_cast<unsigned>(steps)), dim3(tbk::kBlock), 0,
                       ctx->stream, static_cast<const double2*>(partial), static_cast<int32_t>(nb),
                       ctx->sums.as<double2>());
    HIPC(hipGetLastError());
    std::vector<double2> sums(static_cast<size_t>(steps));
    HIPC(hipMemcpyAsync(sums.data(), ctx->sums.p, sums.size() * sizeof(double2), hipMemcpyDeviceToHost,
                        ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));

    // moments
    const double mu0 = sums[0].x, mu1 = sums[0].y;
    mu_out[0] = mu0;
    mu_out[1] = mu1;
    if (doubling) {
        for (int64_t k = 1; k < steps; ++k) {
            if (2 * k < n) mu_out[2 * k] = 2.0 * sums[k].x - mu0;
            if (2 * k + 1 < n) mu_out[2 * k + 1] = 2.0 * sums[k].y - mu1;
        }
    } else {
        for (int64_t k = 1; k < steps; ++k) mu_out[k + 1] = sums[k].x;
    }

    if (last_out) {
        // download T_steps and transpose to column-major complex128
        const void* src = buf(steps);
        const int R = kets->R;
        if (prec == TBK_FP64) {
            std::vector<double2> h(static_cast<size_t>(rows) * Rp);
            HIPC(hipMemcpy(h.data(), src, h.size() * sizeof(double2), hipMemcpyDeviceToHost));
            for (int r = 0; r < R; ++r)
                for (int i = 0; i < rows; ++i) {
                    last_out[2 * (static_cast<size_t>(r) * rows + i)] = h[static_cast<size_t>(i) * Rp + r].x;
                    last_out[2 * (static_cast<size_t>(r) * rows + i) + 1] = h[static_cast<size_t>(i) * Rp + r].y;
                }
        } else {
            std::vector<float2> h(static_cast<size_t>(rows) * Rp);
            HIPC(hipMemcpy(h.data(), src, h.size() * sizeof(float2), hipMemcpyDeviceToHost));
            for (int r = 0; r < R; ++r)
                for (int i = 0; i < rows; ++i) {
                    last_out[2 * (static_cast<size_t>(r) * rows + i)] = h[static_cast<size_t>(i) * Rp + r].x;
                    last_out[2 * (static_cast<size_t>(r) * rows + i) + 1] = h[static_cast<size_t>(i) * Rp + r].y;
                }
        }
    }

    if (stats) {
        float ms = 0.f;
        double kernel_ms = 0.0;
        if (time_each) {
            for (int64_t k = 0; k < steps; ++k) {
                HIPC(hipEventElapsedTime(&ms, ctx->events[2 * k], ctx->events[2 * k + 1]));
                kernel_ms += ms;
            }
        } else {
            HIPC(hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
            kernel_ms = ms;
        }
        const double sval = (prec == TBK_FP64 ? 8.0 : 4.0) * (M->is_complex ? 2.0 : 1.0);
        const double svec = static_cast<double>(kets->elem_bytes());
        const double mat = static_cast<double>(M->nnz) * (sval + 4.0) + (static_cast<double>(rows) + 1.0) * 4.0;
        const double blk = static_cast<double>(rows) * kets->R * svec;
        const double k_rec = doubling ? 3.0 : 4.0;     // x, z, (bra), y
        const double k_first = doubling ? 2.0 : 3.0;   // x, (bra), y
        stats->kernel_ms = kernel_ms;
        stats->steps = steps;
        stats->moments = n;
        stats->bytes_per_step = mat + k_rec * blk;
        stats->bytes_total = (mat + k_first * blk) + static_cast<double>(steps - 1) * (mat + k_rec * blk);
        stats->doubling = doubling ? 1 : 0;
        stats->blocks = static_cast<int32_t>(nb);
        stats->rows_per_tile = TR;
        stats->padded_R = Rp;
        stats->kernel_kind = kc.kch;
        stats->slab_cols = kc.width;
        stats->chunk_period = static_cast<int32_t>(order.P);
        stats->chunk_rows = static_cast<int32_t>(order.tpc) * TR;
        stats->wall_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
}


// ------------------------------------------------- stored Chebyshev blocks
// Device block (rows x Rp, row-major) to column-major complex128 on the host.
void block_to_colmajor(const void* src, int prec, int rows, int R, int Rp, double* out) {
    const size_t elems = static_cast<size_t>(rows) * Rp;
    auto put = [&](const auto& h) {
        for (int r = 0; r < R; ++r)
            for (int i = 0; i < rows; ++i) {
                const size_t s = 2 * (static_cast<size_t>(r) * rows + i);
                out[s] = h[static_cast<size_t>(i) * Rp + r].x;
                out[s + 1] = h[static_cast<size_t>(i) * Rp + r].y;
            }
    };
    if (prec == TBK_FP64) {
        std::vector<double2> h(elems);
        HIPC(hipMemcpy(h.data(), src, elems * sizeof(double2), hipMemcpyDeviceToHost));
        put(h);
    } else {
        std::vector<float2> h(elems);
        HIPC(hipMemcpy(h.data(), src, elems * sizeof(float2), hipMemcpyDeviceToHost));
        put(h);
    }
}

void check_on_bases_args(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* v,
                         int32_t n) {
    if (!ctx || !M || !v) raise(TBK_E_INVALID, "null argument");
    if (n < 1) raise(TBK_E_INVALID, "n must be at least 1");
    if (!(a != 0.0) || !std::isfinite(a) || !std::isfinite(b)) raise(TBK_E_INVALID, "invalid rescaling a, b");
    if (M->ctx != ctx || v->ctx != ctx) raise(TBK_E_INVALID, "objects belong to another context");
    if (v->rows != M->rows) raise(TBK_E_INVALID, "block rows do not match the matrix");
}

// Stored blocks T_k(M~) v, M~ = (M - b)/a, k = 0..n-1
// (Cheb_first_kind_seq_on_bases, numerical.h:334-344): dst(k) is the device
// memory of T_k, laid out as v.  dst(0) gets a copy of v (skipped when it is
// v's own memory); each k >= 1 is one launch, T_1 = M~ T_0 and
// T_k = 2 M~ T_{k-1} - T_{k-2}, the arithmetic of run_moments's steps.
// after(k) runs on the host once T_k is enqueued (the host-buffer drop-in
// downloads there).  Returns the milliseconds between the first and the
// last launch's events.
template <typename Dst, typename After>
double enqueue_on_bases(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* v, int32_t n,
                        Dst dst, After after) {
    set_device(ctx);
    const StepPlan plan = make_plan(ctx, M, v->precision, v->Rp, v->rows, b, false);
    ctx->partial.reserve(static_cast<size_t>(plan.nb) * sizeof(double2));
    double2* part = ctx->partial.as<double2>();
    const size_t block_bytes = static_cast<size_t>(v->rows) * v->Rp * v->elem_bytes();
    ensure_events(ctx, 2);
    if (dst(0) != v->data.p)
        HIPC(hipMemcpyAsync(dst(0), v->data.p, block_bytes, hipMemcpyDeviceToDevice, ctx->stream));
    after(0);
    HIPC(hipEventRecord(ctx->events[0], ctx->stream));
    for (int64_t k = 1; k < n; ++k) {
        if (k == 1)
            launch_plan(plan, ctx->stream, tbk::kFirst, 1.0 / a, -b / a, dst(0), nullptr, dst(1), dst(0), part);
        else
            launch_plan(plan, ctx->stream, tbk::kRecDbl, 2.0 / a, -2.0 * b / a, dst(k - 1), dst(k - 2), dst(k),
                        nullptr, part);
        after(k);
    }
    HIPC(hipEventRecord(ctx->events[1], ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
    return ms;
}

// ------------------------------------------------------ Kubo-Bastin Gamma
// Split-K shape of the contraction: enough workgroups to fill the device,
// chunks a multiple of 16 complex columns.
void contraction_shape(const tbk_context* ctx, int N, int64_t K, int64_t& kc, int& nchunks) {
    const int64_t tiles = static_cast<int64_t>((N + tbk::kGammaTile - 1) / tbk::kGammaTile);
    const int64_t want = std::max<int64_t>(1, (4 * static_cast<int64_t>(ctx->num_cu) + tiles * tiles - 1) / (tiles * tiles));
    kc = (K + want - 1) / want;
    kc = std::max<int64_t>(256, ((kc + 15) / 16) * 16);
    nchunks = static_cast<int>((K + kc - 1) / kc);
}

// Large-N kernels: about one 128 x 128 tile work item per CU, chunks a
// multiple of the LDS stage (16 complex columns covers both kernels), and the
// item count a multiple of the 8 XCDs where the chunk count allows it.
// max_cols > 0 caps the chunk length (the fp32-accumulating kernel).
constexpr int64_t kSplitChunkCols = 32768;
void contraction_shape_big(const tbk_context* ctx, int N, int64_t K, int64_t max_cols, int64_t& kc,
                           tbk::GammaMap& map) {
    map.tiles_n = (N + tbk::kBigTile - 1) / tbk::kBigTile;
    map.tiles_m = map.tiles_n;
    const int64_t tiles = static_cast<int64_t>(map.tiles_n) * map.tiles_m;
    int64_t chunks = std::max<int64_t>(1, (static_cast<int64_t>(ctx->num_cu) + tiles / 2) / tiles);
    while ((tiles * chunks) % 8 != 0 && chunks < 8) ++chunks;
    if (max_cols > 0) chunks = std::max<int64_t>(chunks, (K + max_cols - 1) / max_cols);
    kc = (K + chunks - 1) / chunks;
    kc = std::max<int64_t>(64, ((kc + 15) / 16) * 16);
    map.chunks = static_cast<int32_t>((K + kc - 1) / kc);
    map.per_xcd = (tiles * map.chunks + 7) / 8;
}

}  // namespace

namespace engine {

int contraction_precision(int prec) {
    if (prec == TBK_GAMMA_FP64) return TBK_GAMMA_FP64;
    if (prec == TBK_GAMMA_BF16X3 || prec == TBK_GAMMA_BF16X3_FP32) return TBK_GAMMA_BF16X3;
    raise(TBK_E_INVALID, "unknown contraction precision");
}

ContractionShape contraction_plan(const tbk_context* ctx, int N, int64_t K, int prec) {
    prec = contraction_precision(prec);
    ContractionShape cs{};
    cs.big = N >= 2 * tbk::kBigTile;
    if (cs.big) {
        contraction_shape_big(ctx, N, K, prec == TBK_GAMMA_FP64 ? 0 : kSplitChunkCols, cs.kc, cs.map);
        cs.nchunks = cs.map.chunks;
        // pre-split words: 256 x 128 tiles over the same k chunks (so the
        // same partial sums), unless the context asks for 128 x 128
        cs.wide = prec == TBK_GAMMA_BF16X3 && ctx->tuning.gamma_tile_rows != 128;
        cs.wide_map.tiles_n = (N + tbk::kWideTile - 1) / tbk::kWideTile;
        cs.wide_map.tiles_m = (N + tbk::kBigTile - 1) / tbk::kBigTile;
        cs.wide_map.chunks = cs.map.chunks;
        cs.wide_map.per_xcd =
            (static_cast<int64_t>(cs.wide_map.tiles_n) * cs.wide_map.tiles_m * cs.wide_map.chunks + 7) / 8;
    } else {
        contraction_shape(ctx, N, K, cs.kc, cs.nchunks);
    }
    return cs;
}

// split-K partial tiles into P (nchunks x N x N), then their ordered sum into out
void enqueue_contraction(const ContractionShape& cs, const double2* A, const double2* B, int N, int64_t K, int64_t ldk,
                         int prec, double2* P, double2* out, hipStream_t st) {
    prec = contraction_precision(prec);
    if (prec == TBK_GAMMA_FP64 && cs.big)
        tbk::launch_gamma_f64_big(A, B, N, K, ldk, cs.kc, cs.map, P, st);
    else if (prec == TBK_GAMMA_FP64)
        tbk::launch_gamma_f64(A, B, N, K, ldk, cs.kc, cs.nchunks, P, st);
    else if (cs.big)
        tbk::launch_gamma_bf16x3_big(A, B, N, K, ldk, cs.kc, cs.map, P, st);
    else
        tbk::launch_gamma_bf16x3(A, B, N, K, ldk, cs.kc, cs.nchunks, P, st);
    HIPC(hipGetLastError());
    tbk::launch_gamma_reduce(P, cs.nchunks, static_cast<int64_t>(N) * N, out, st);
    HIPC(hipGetLastError());
}

// Both Kubo stacks: L[n] = T_n(H~)|r>, R[m] = v_x T_m(H~) v_y|r> (n, m < N),
// blocks of blk_bytes each, entry n at n * ldk elements (stack_stride); T_m
// of the right side rotates through W0, W1.
// v_x = [X, H~] stores H~'s pattern; when the uploaded arrays agree (same
// fingerprint) and H~ uses the row kernel, v_x T_{m-1} is formed in the step
// that gathers T_{m-1} for T_m (kDual), which saves a read of every right
// block and a launch per step.  Each product is the same sum in the same
// order as a separate v_x launch, so the stack is the same bits.
bool fused_velocity(const tbk_matrix* H, const tbk_matrix* vx, const StepPlan& pH, hipStream_t st) {
    if (!H->is_complex || !vx->is_complex || H->rows != vx->rows || H->nnz != vx->nnz || pH.kc.kch != 0) return false;
    return matrix_pattern(H, st) == matrix_pattern(vx, st);
}

// Row stride of the stacks: K rounded up to 64 elements, so every stack
// entry starts 1 KB aligned and the contraction's 16-byte column-pair loads
// apply whatever K is (K = sites x columns is odd for one column and an odd
// site count).
int64_t stack_stride(int64_t K) { return ((K + 63) / 64) * 64; }

void enqueue_kubo_stacks(const StepPlan& pH, const StepPlan& pX, const StepPlan& pY, double a, double b,
                         const void* kets, void* L, void* Rs, void* W0, void* W1, size_t blk_bytes, int64_t ldk,
                         int N, double2* part, hipStream_t st, const void* vx_values) {
    const size_t entry = static_cast<size_t>(ldk) * sizeof(double2);
    auto Lk = [&](int64_t k) { return static_cast<char*>(L) + static_cast<size_t>(k) * entry; };
    auto Rk = [&](int64_t k) { return static_cast<char*>(Rs) + static_cast<size_t>(k) * entry; };
    auto W = [&](int64_t k) { return (k & 1) ? W1 : W0; };
    HIPC(hipMemcpyAsync(Lk(0), kets, blk_bytes, hipMemcpyDeviceToDevice, st));
    if (N > 1) launch_plan(pH, st, tbk::kFirst, 1.0 / a, -b / a, Lk(0), nullptr, Lk(1), Lk(0), part);
    for (int64_t n = 1; n + 1 < N; ++n)
        launch_plan(pH, st, tbk::kRecDbl, 2.0 / a, -2.0 * b / a, Lk(n), Lk(n - 1), Lk(n + 1), nullptr, part);
    launch_plan(pY, st, tbk::kFirst, 1.0, 0.0, kets, nullptr, W(0), kets, part);
    launch_plan(pX, st, tbk::kFirst, 1.0, 0.0, W(0), nullptr, Rk(0), W(0), part);
    if (N > 1) {
        launch_plan(pH, st, tbk::kFirst, 1.0 / a, -b / a, W(0), nullptr, W(1), W(0), part);
        if (!vx_values) launch_plan(pX, st, tbk::kFirst, 1.0, 0.0, W(1), nullptr, Rk(1), W(1), part);
    }
    for (int64_t m = 2; m < N; ++m) {
        if (vx_values) {  // T_m and v_x T_{m-1} from one pass over T_{m-1}
            launch_plan(pH, st, tbk::kDual, 2.0 / a, -2.0 * b / a, W(m - 1), W(m), W(m), nullptr, part, vx_values,
                        Rk(m - 1));
        } else {
            launch_plan(pH, st, tbk::kRecDbl, 2.0 / a, -2.0 * b / a, W(m - 1), W(m), W(m), nullptr, part);
            launch_plan(pX, st, tbk::kFirst, 1.0, 0.0, W(m), nullptr, Rk(m), W(m), part);
        }
    }
    if (vx_values && N > 1) launch_plan(pX, st, tbk::kFirst, 1.0, 0.0, W(N - 1), nullptr, Rk(N - 1), W(N - 1), part);
}

// enqueue_kubo_stacks with the stacks holding rows [r0, r1) of every entry
// only: both chains run in full rolling blocks (B0, B1 for T_n; W0, W1 for
// T_m v_y|r>; S for v_x T_m v_y|r>) and each entry's rows, contiguous in the
// row-major block, are copied into its stack slot (entries ldk elements
// apart).  Gamma over all rows is the sum of the row ranges' Gammas, so a
// problem whose stacks exceed the device runs the chains once per range.
void enqueue_kubo_stacks_rows(const StepPlan& pH, const StepPlan& pX, const StepPlan& pY, double a, double b,
                              const void* kets, void* L, void* Rs, void* B0, void* B1, void* W0, void* W1, void* S,
                              int rows, int Rp, int r0, int r1, int64_t ldk, int N, double2* part, hipStream_t st) {
    const size_t row_bytes = static_cast<size_t>(Rp) * sizeof(double2);
    const size_t blk = static_cast<size_t>(rows) * row_bytes;
    const size_t off = static_cast<size_t>(r0) * row_bytes, len = static_cast<size_t>(r1 - r0) * row_bytes;
    const size_t entry = static_cast<size_t>(ldk) * sizeof(double2);
    auto Lk = [&](int64_t k) { return static_cast<char*>(L) + static_cast<size_t>(k) * entry; };
    auto Rk = [&](int64_t k) { return static_cast<char*>(Rs) + static_cast<size_t>(k) * entry; };
    auto B = [&](int64_t k) { return (k & 1) ? B1 : B0; };
    auto W = [&](int64_t k) { return (k & 1) ? W1 : W0; };
    auto keep = [&](void* dst, const void* src) {
        HIPC(hipMemcpyAsync(dst, static_cast<const char*>(src) + off, len, hipMemcpyDeviceToDevice, st));
    };
    HIPC(hipMemcpyAsync(B0, kets, blk, hipMemcpyDeviceToDevice, st));
    keep(Lk(0), B0);
    if (N > 1) {
        launch_plan(pH, st, tbk::kFirst, 1.0 / a, -b / a, B(0), nullptr, B(1), B(0), part);
        keep(Lk(1), B(1));
    }
    for (int64_t n = 1; n + 1 < N; ++n) {  // T_{n+1} over T_{n-1} in place
        launch_plan(pH, st, tbk::kRecDbl, 2.0 / a, -2.0 * b / a, B(n), B(n - 1), B(n + 1), nullptr, part);
        keep(Lk(n + 1), B(n + 1));
    }
    launch_plan(pY, st, tbk::kFirst, 1.0, 0.0, kets, nullptr, W(0), kets, part);
    launch_plan(pX, st, tbk::kFirst, 1.0, 0.0, W(0), nullptr, S, W(0), part);
    keep(Rk(0), S);
    if (N > 1) {
        launch_plan(pH, st, tbk::kFirst, 1.0 / a, -b / a, W(0), nullptr, W(1), W(0), part);
        launch_plan(pX, st, tbk::kFirst, 1.0, 0.0, W(1), nullptr, S, W(1), part);
        keep(Rk(1), S);
    }
    for (int64_t m = 2; m < N; ++m) {
        launch_plan(pH, st, tbk::kRecDbl, 2.0 / a, -2.0 * b / a, W(m - 1), W(m), W(m), nullptr, part);
        launch_plan(pX, st, tbk::kFirst, 1.0, 0.0, W(m), nullptr, S, W(m), part);
        keep(Rk(m), S);
    }
}

// Split-BF16x3 solves with N >= 256 on complex operators keep their stacks
// pre-split: the recursion stores every stack entry as the (hi, lo) bf16
// words the contraction would form from it (split_bf16.h), 8 bytes per
// complex, and the contraction stages those words directly.  Gamma is the
// same bits as from fp64 stacks.
bool kubo_presplit(int prec, int N, bool h_complex, bool vx_complex) {
    return (prec == TBK_GAMMA_BF16X3 || prec == TBK_GAMMA_BF16X3_FP32) && N >= 2 * tbk::kBigTile && h_complex &&
           vx_complex;
}

// The left chain runs in two rolling fp64 blocks B0, B1 (B0 = a copy of the
// kets) and stores each T_n's words at Ls[n]; the right chain is
// enqueue_kubo_stacks's with R_m stored only as words at Rs[m] (the fp64
// results of separate v_x launches go to the scratch block S).  K: complex
// elements per block.
void enqueue_kubo_stacks_presplit(const StepPlan& pH, const StepPlan& pX, const StepPlan& pY, double a, double b,
                                  const void* kets, uint2* Ls, uint2* Rs, void* B0, void* B1, void* S, void* W0,
                                  void* W1, int64_t K, int N, double2* part, hipStream_t st, const void* vx_values) {
    if (pH.kc.kch != 0 || pX.kc.kch != 0 || !pH.cplx || !pX.cplx)
        raise(TBK_E_UNSUPPORTED, "pre-split stacks need the complex row kernel");
    // k-blocked word stacks (word_offset): entry k's words start at k * kWordBlock
    const int64_t bs = static_cast<int64_t>(N) * tbk::kWordBlock;
    auto Lk = [&](int64_t k) { return Ls + k * tbk::kWordBlock; };
    auto Rk = [&](int64_t k) { return Rs + k * tbk::kWordBlock; };
    auto B = [&](int64_t k) { return (k & 1) ? B1 : B0; };
    auto W = [&](int64_t k) { return (k & 1) ? W1 : W0; };
    HIPC(hipMemcpyAsync(B0, kets, static_cast<size_t>(K) * sizeof(double2), hipMemcpyDeviceToDevice, st));
    tbk::launch_split_block(static_cast<const double2*>(kets), Lk(0), K, bs, st);
    HIPC(hipGetLastError());
    if (N > 1)
        launch_plan(pH, st, tbk::kFirstSplit, 1.0 / a, -b / a, B(0), nullptr, B(1), B(0), part, nullptr, Lk(1),
                    nullptr, bs);
    for (int64_t n = 1; n + 1 < N; ++n)  // T_{n+1} over T_{n-1} in place
        launch_plan(pH, st, tbk::kRecDblSplit, 2.0 / a, -2.0 * b / a, B(n), B(n + 1), B(n + 1), nullptr, part, nullptr,
                    Lk(n + 1), nullptr, bs);
    launch_plan(pY, st, tbk::kFirst, 1.0, 0.0, kets, nullptr, W(0), kets, part);
    launch_plan(pX, st, tbk::kFirstSplit, 1.0, 0.0, W(0), nullptr, S, W(0), part, nullptr, Rk(0), nullptr, bs);
    if (N > 1) {
        launch_plan(pH, st, tbk::kFirst, 1.0 / a, -b / a, W(0), nullptr, W(1), W(0), part);
        if (!vx_values)
            launch_plan(pX, st, tbk::kFirstSplit, 1.0, 0.0, W(1), nullptr, S, W(1), part, nullptr, Rk(1), nullptr, bs);
    }
    for (int64_t m = 2; m < N; ++m) {
        if (vx_values) {
            launch_plan(pH, st, tbk::kDualSplit, 2.0 / a, -2.0 * b / a, W(m - 1), W(m), W(m), nullptr, part, vx_values,
                        Rk(m - 1), nullptr, bs);
        } else {
            launch_plan(pH, st, tbk::kRecDbl, 2.0 / a, -2.0 * b / a, W(m - 1), W(m), W(m), nullptr, part);
            launch_plan(pX, st, tbk::kFirstSplit, 1.0, 0.0, W(m), nullptr, S, W(m), part, nullptr, Rk(m), nullptr, bs);
        }
    }
    if (vx_values && N > 1)
        launch_plan(pX, st, tbk::kFirstSplit, 1.0, 0.0, W(N - 1), nullptr, S, W(N - 1), part, nullptr, Rk(N - 1),
                    nullptr, bs);
}

// Both chains of the pre-split stacks as one block of 2 Rp columns, the left
// chain in columns [0, Rp) and the right one in [Rp, 2 Rp): X_0 = [|r> |
// v_y |r>], X_1 = H~ X_0, X_{k+1} = 2 H~ X_k - X_{k-1} in two rolling blocks.
// Each step stores the left half's X_{k+1} as words at Ls[k + 1] and the
// right half's v_x X_k at Rs[k] (kPairFirst, kPairRec): the matrix is read
// once per index for both chains, one launch instead of two.  The words are
// enqueue_kubo_stacks_presplit's with the fused v_x product, bit for bit
// (the same sums per lane).  The last step's left output is not a stack
// entry and is not stored (null target).
void enqueue_kubo_stacks_paired(const StepPlan& pP, const StepPlan& pY, double a, double b, const void* kets,
                                uint2* Ls, uint2* Rs, void* X0, void* X1, void* S, int64_t rows, int Rp, int N,
                                double2* part, hipStream_t st, const void* vx_values) {
    if (pP.kc.kch != 0 || !pP.cplx || !vx_values) raise(TBK_E_UNSUPPORTED, "paired stacks need the fused row kernel");
    // fp32 plan: X blocks of complex64, H's and v_x's fp32 values (vx_values
    // must be v_x's val32 then); the words of X_{k+1} and v_x X_k are split
    // from the fp32 values
    const bool f32 = pP.precision == TBK_FP32;
    if (f32 && Rp < 2) raise(TBK_E_UNSUPPORTED, "fp32 paired stacks need at least two columns per chain");
    const int64_t K = rows * Rp;
    // k-blocked word stacks (word_offset): entry k's words start at k * kWordBlock
    const int64_t bs = static_cast<int64_t>(N) * tbk::kWordBlock;
    auto Lk = [&](int64_t k) { return Ls + k * tbk::kWordBlock; };
    auto Rk = [&](int64_t k) { return Rs + k * tbk::kWordBlock; };
    auto X = [&](int64_t k) { return (k & 1) ? X1 : X0; };
    tbk::launch_split_block(static_cast<const double2*>(kets), Lk(0), K, bs, st);
    HIPC(hipGetLastError());
    launch_plan(pY, st, tbk::kFirst, 1.0, 0.0, kets, nullptr, S, kets, part);  // v_y |r>
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>((2 * K + tbk::kBlock - 1) / tbk::kBlock, 65536));
    if (f32)
        hipLaunchKernelGGL(tbk::pair_blocks_kernel<float2>, dim3(grid), dim3(tbk::kBlock), 0, st,
                           static_cast<const double2*>(kets), static_cast<const double2*>(S), static_cast<float2*>(X0),
                           rows, Rp);
    else
        hipLaunchKernelGGL(tbk::pair_blocks_kernel<double2>, dim3(grid), dim3(tbk::kBlock), 0, st,
                           static_cast<const double2*>(kets), static_cast<const double2*>(S),
                           static_cast<double2*>(X0), rows, Rp);
    HIPC(hipGetLastError());
    launch_plan(pP, st, tbk::kPairFirst, 1.0 / a, -b / a, X(0), nullptr, X(1), X(0), part, vx_values,
                N > 1 ? Lk(1) : nullptr, Rk(0), bs);
    for (int64_t k = 1; k < N; ++k)  // X_{k+1} over X_{k-1}
        launch_plan(pP, st, tbk::kPairRec, 2.0 / a, -2.0 * b / a, X(k), X(k - 1), X(k + 1), nullptr, part, vx_values,
                    k + 1 < N ? Lk(k + 1) : nullptr, Rk(k), bs);
}

void enqueue_contraction_presplit(const ContractionShape& cs, const uint2* A, const uint2* B, int N, int64_t K,
                                  double2* P, double2* out, hipStream_t st) {
    if (!cs.big) raise(TBK_E_UNSUPPORTED, "pre-split contraction needs N >= 256");
    const int64_t wbs = static_cast<int64_t>(N) * tbk::kWordBlock;
    if (cs.wide)
        tbk::launch_gamma_bf16x3_wide_presplit(A, B, N, K, wbs, cs.kc, cs.wide_map, P, st);
    else
        tbk::launch_gamma_bf16x3_big_presplit(A, B, N, K, wbs, cs.kc, cs.map, P, st);
    HIPC(hipGetLastError());
    tbk::launch_gamma_reduce(P, cs.nchunks, static_cast<int64_t>(N) * N, out, st);
    HIPC(hipGetLastError());
}

double kubo_recursion_bytes(const tbk_matrix* H, const tbk_matrix* vx, const tbk_matrix* vy, int rows, int R,
                            int64_t N, bool fused, bool presplit, bool paired, bool paired_f32) {
    // per launch: the matrix arrays once; vector blocks x (+ z) read and y
    // written; split words are half a block
    auto mat = [&](const tbk_matrix* m) {
        return static_cast<double>(m->nnz) * ((m->is_complex ? 16.0 : 8.0) + 4.0) +
               (static_cast<double>(rows) + 1.0) * 4.0;
    };
    const double blk = static_cast<double>(rows) * R * 16.0, half = 0.5 * blk;
    const double split = presplit ? half : 0.0;
    const double vals = static_cast<double>(vx->nnz) * 16.0;
    const double Nm2 = static_cast<double>(std::max<int64_t>(0, N - 2));
    if (paired && paired_f32) {
        // as below with the paired block in complex64 (2 R columns of 8 bytes:
        // one blk) and the matrix values in fp32
        const double mat32 = static_cast<double>(H->nnz) * (8.0 + 4.0) + (static_cast<double>(rows) + 1.0) * 4.0;
        const double vals32 = static_cast<double>(vx->nnz) * 8.0;
        double bytes = blk + half + mat(vy) + 2.0 * blk + 3.0 * blk;
        bytes += mat32 + vals32 + 2.0 * blk + blk;
        bytes += static_cast<double>(N - 1) * (mat32 + vals32 + 3.0 * blk + blk);
        return bytes;
    }
    if (paired) {
        // kets' words; v_y |r>; the paired block X_0; N launches over 2 R
        // columns, each reading the matrix, H2's values, x (and z) and
        // writing y and two half-blocks of words
        double bytes = blk + half + mat(vy) + 2.0 * blk + 4.0 * blk;
        bytes += mat(H) + vals + 4.0 * blk + blk;
        bytes += static_cast<double>(N - 1) * (mat(H) + vals + 6.0 * blk + blk);
        return bytes;
    }
    double bytes = 2.0 * blk;  // the kets copied (into the stack, or the rolling block)
    // left chain
    if (presplit) bytes += blk + half;  // the kets' words
    if (N > 1) bytes += mat(H) + 2.0 * blk + split;
    bytes += Nm2 * (mat(H) + 3.0 * blk + split);
    // right chain; a separate v_x launch writes fp64 (and the words)
    const double rout = presplit ? blk + half : blk;
    bytes += mat(vy) + 2.0 * blk;        // v_y |r>
    bytes += mat(vx) + blk + rout;       // R_0
    if (N > 1) bytes += mat(H) + 2.0 * blk;
    if (fused && N > 1) {
        bytes += Nm2 * (mat(H) + vals + 3.0 * blk + (presplit ? half : blk));  // T_m and R_{m-1}
        bytes += mat(vx) + blk + rout;                                        // R_{N-1}
    } else {
        if (N > 1) bytes += mat(vx) + blk + rout;  // R_1
        bytes += Nm2 * (mat(H) + 3.0 * blk + mat(vx) + blk + rout);
    }
    return bytes;
}

}  // namespace engine

namespace {

// gamma = A^H B on the device (A, B: N x K complex128), into host gamma_out.
double contract_stacks(tbk_context* ctx, const double2* A, const double2* B, int N, int64_t K, int64_t ldk,
                       int prec, double* gamma_out, int* nchunks_out) {
    const ContractionShape cs = contraction_plan(ctx, N, K, prec);
    const int64_t NN = static_cast<int64_t>(N) * N;
    DevBuf& P = ctx->gamma_part;
    DevBuf& out = ctx->gamma_dev;
    P.reserve(static_cast<size_t>(cs.nchunks) * NN * sizeof(double2));
    out.reserve(static_cast<size_t>(NN) * sizeof(double2));
    ensure_events(ctx, 2);
    HIPC(hipEventRecord(ctx->events[0], ctx->stream));
    enqueue_contraction(cs, A, B, N, K, ldk, prec, P.as<double2>(), out.as<double2>(), ctx->stream);
    HIPC(hipEventRecord(ctx->events[1], ctx->stream));
    HIPC(hipMemcpyAsync(gamma_out, out.p, static_cast<size_t>(NN) * sizeof(double2), hipMemcpyDeviceToHost,
                        ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
    if (nchunks_out) *nchunks_out = cs.nchunks;
    return ms;
}

// gamma = A^H B over pre-split stacks (N x K words each), into host gamma_out.
double contract_stacks_presplit(tbk_context* ctx, const uint2* A, const uint2* B, int N, int64_t K,
                                double* gamma_out, int* nchunks_out) {
    const ContractionShape cs = contraction_plan(ctx, N, K, TBK_GAMMA_BF16X3);
    const int64_t NN = static_cast<int64_t>(N) * N;
    DevBuf& P = ctx->gamma_part;
    DevBuf& out = ctx->gamma_dev;
    P.reserve(static_cast<size_t>(cs.nchunks) * NN * sizeof(double2));
    out.reserve(static_cast<size_t>(NN) * sizeof(double2));
    ensure_events(ctx, 2);
    HIPC(hipEventRecord(ctx->events[0], ctx->stream));
    enqueue_contraction_presplit(cs, A, B, N, K, P.as<double2>(), out.as<double2>(), ctx->stream);
    HIPC(hipEventRecord(ctx->events[1], ctx->stream));
    HIPC(hipMemcpyAsync(gamma_out, out.p, static_cast<size_t>(NN) * sizeof(double2), hipMemcpyDeviceToHost,
                        ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
    if (nchunks_out) *nchunks_out = cs.nchunks;
    return ms;
}

void run_kubo(tbk_context* ctx, const tbk_matrix* H, double a, double b, const tbk_matrix* vx, const tbk_matrix* vy,
              const tbk_block* kets, int32_t N, int32_t prec, const double* g, double* gamma_out,
              tbk_kubo_stats* stats) {
    auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !H || !vx || !vy || !kets || !gamma_out) raise(TBK_E_INVALID, "null argument");
    if (N < 1) raise(TBK_E_INVALID, "N must be at least 1");
    if (!(a != 0.0) || !std::isfinite(a) || !std::isfinite(b)) raise(TBK_E_INVALID, "invalid rescaling a, b");
    if (H->ctx != ctx || vx->ctx != ctx || vy->ctx != ctx || kets->ctx != ctx)
        raise(TBK_E_INVALID, "objects belong to another context");
    if (kets->precision != TBK_FP64) raise(TBK_E_INVALID, "Kubo stacks are built from fp64 kets");
    if (H->rows != kets->rows || vx->rows != kets->rows || vy->rows != kets->rows)
        raise(TBK_E_INVALID, "matrix and kets rows differ");
    set_device(ctx);
    const int rows = kets->rows, Rp0 = kets->Rp;
    // Device bytes of the stacks and work blocks for W columns (rolling: the
    // most fp64 work blocks any stack form uses).
    auto stack_entry = [&](int W, bool split) -> size_t {
        const int64_t Kw = static_cast<int64_t>(rows) * W;
        const int64_t kw = ((Kw + tbk::kWordBlock - 1) / tbk::kWordBlock) * tbk::kWordBlock;
        return split ? static_cast<size_t>(kw) * sizeof(uint2) : static_cast<size_t>(stack_stride(Kw)) * sizeof(double2);
    };
    auto need_bytes = [&](int W) -> double {
        const bool split = kubo_presplit(prec, N, H->is_complex, vx->is_complex);
        return 2.0 * static_cast<double>(N) * static_cast<double>(stack_entry(W, split)) +
               7.0 * static_cast<double>(rows) * W * sizeof(double2);
    };
    // Gamma is a sum over the random vectors: when the stacks of the whole
    // block do not fit, the vectors run in groups of Gw columns (the widest
    // power of two that fits) and the groups' Gammas are added in order.
    DevBuf& L = ctx->stack[0];
    DevBuf& Rs = ctx->stack[1];
    int Gw = Rp0;
    int nrc = 1;  // row ranges of the fp64 stacks
    const size_t whole_stack =
        static_cast<size_t>(N) * stack_entry(Rp0, kubo_presplit(prec, N, H->is_complex, vx->is_complex));
    const double cap = ctx->tuning.kubo_stack_mib > 0 ? static_cast<double>(ctx->tuning.kubo_stack_mib) * 1048576.0
                                                      : std::numeric_limits<double>::infinity();
    if (L.bytes < whole_stack || Rs.bytes < whole_stack || need_bytes(Rp0) > cap) {
        HIPC(hipStreamSynchronize(ctx->stream));
        L.release();
        Rs.release();
        size_t free_b = 0, total_b = 0;
        HIPC(hipMemGetInfo(&free_b, &total_b));
        const double avail = std::min(0.95 * static_cast<double>(free_b), cap);
        while (Gw > 1 && need_bytes(Gw) > avail) Gw /= 2;
        // fp64 stacks of one vector still too large: they hold a range of
        // rows of every entry at a time and the chains run once per range
        const bool split = kubo_presplit(prec, N, H->is_complex, vx->is_complex);
        auto need_rows = [&](int nrc) {
            const int crows = (rows + nrc - 1) / nrc;
            return 2.0 * static_cast<double>(N) * static_cast<double>(stack_stride(crows)) * sizeof(double2) +
                   7.0 * static_cast<double>(rows) * sizeof(double2);
        };
        if (need_bytes(Gw) > avail && !split)
            while (nrc < 64 && need_rows(nrc) > avail) ++nrc;
        if (nrc == 1 ? need_bytes(Gw) > avail : need_rows(nrc) > avail)
            raise(TBK_E_OOM, "Kubo stacks of one random vector need " + std::to_string(need_bytes(1) / 1e9) +
                                 " GB, " + std::to_string(avail / 1e9) + " GB available");
    }
    const int ngroups = (kets->R + Gw - 1) / Gw;
    DevBuf gkets;
    std::vector<double> gsum;
    if (ngroups > 1) {
        gkets.reserve(static_cast<size_t>(rows) * Gw * sizeof(double2));
        gsum.assign(static_cast<size_t>(2) * N * N, 0.0);
    }
    double rec_total = 0.0, con_total = 0.0, flops_total = 0.0, rbytes_total = 0.0, stack_b = 0.0;
    int nchunks = 0;
    bool any_fp32 = false;
    for (int grp = 0; grp < ngroups; ++grp) {
    const int Rp = Gw;
    const void* kp = kets->data.p;
    int Rreal = kets->R;
    if (ngroups > 1) {
        // columns [grp Gw, grp Gw + cols) of the caller's kets, Gw wide
        const int col0 = grp * Gw;
        Rreal = std::min(Gw, kets->R - col0);
        const int64_t total = static_cast<int64_t>(rows) * Gw;
        const unsigned cg = static_cast<unsigned>(std::min<int64_t>((total + tbk::kBlock - 1) / tbk::kBlock, 65536));
        hipLaunchKernelGGL(tbk::column_group_kernel, dim3(cg), dim3(tbk::kBlock), 0, ctx->stream,
                           kets->data.as<const double2>(), rows, Rp0, col0, Rreal, Gw, gkets.as<double2>());
        HIPC(hipGetLastError());
        kp = gkets.p;
    }
    double* gout = ngroups > 1 ? gsum.data() : gamma_out;
    if (nrc > 1) {
        // one vector per group (Rp = 1), rows [r0, r1) of every stack entry
        const int crows = (rows + nrc - 1) / nrc;
        const int64_t ldk_c = stack_stride(crows);
        const size_t blk1 = static_cast<size_t>(rows) * sizeof(double2);
        L.reserve(static_cast<size_t>(N) * ldk_c * sizeof(double2));
        Rs.reserve(static_cast<size_t>(N) * ldk_c * sizeof(double2));
        ctx->work[0].reserve(blk1);
        ctx->work[1].reserve(blk1);
        for (DevBuf& w : ctx->kwork) w.reserve(blk1);
        const StepPlan qH = make_plan(ctx, H, TBK_FP64, 1, rows, b, false);
        const StepPlan qX = make_plan(ctx, vx, TBK_FP64, 1, rows, 0.0, false);
        const StepPlan qY = make_plan(ctx, vy, TBK_FP64, 1, rows, 0.0, false);
        ctx->partial.reserve(static_cast<size_t>(std::max({qH.nb, qX.nb, qY.nb})) * sizeof(double2));
        std::vector<double> gpart(static_cast<size_t>(2) * N * N);
        for (int rc = 0; rc < nrc; ++rc) {
            const int r0 = rc * crows, r1 = std::min(rows, r0 + crows);
            ensure_events(ctx, 2);
            HIPC(hipEventRecord(ctx->events[0], ctx->stream));
            enqueue_kubo_stacks_rows(qH, qX, qY, a, b, kp, L.p, Rs.p, ctx->kwork[0].p, ctx->kwork[1].p,
                                     ctx->work[0].p, ctx->work[1].p, ctx->kwork[2].p, rows, 1, r0, r1, ldk_c, N,
                                     ctx->partial.as<double2>(), ctx->stream);
            HIPC(hipEventRecord(ctx->events[1], ctx->stream));
            HIPC(hipStreamSynchronize(ctx->stream));
            float rms = 0.f;
            HIPC(hipEventElapsedTime(&rms, ctx->events[0], ctx->events[1]));
            const int64_t Kc = r1 - r0;
            const double cms = contract_stacks(ctx, L.as<const double2>(), Rs.as<const double2>(), N, Kc, ldk_c, prec,
                                               gpart.data(), &nchunks);
            double* dst = gout;
            if (rc == 0) std::copy(gpart.begin(), gpart.end(), dst);
            else for (size_t i = 0; i < gpart.size(); ++i) dst[i] = dst[i] + gpart[i];
            rec_total += rms;
            con_total += cms;
            flops_total += 8.0 * static_cast<double>(N) * N * static_cast<double>(Kc);
        }
        stack_b = 2.0 * static_cast<double>(N) * static_cast<double>(crows) * sizeof(double2);
        rbytes_total += static_cast<double>(nrc) * kubo_recursion_bytes(H, vx, vy, rows, Rreal, N, false, false, false);
        if (ngroups > 1) {
            const size_t nn2 = static_cast<size_t>(2) * N * N;
            if (grp == 0) std::copy(gout, gout + nn2, gamma_out);
            else for (size_t i = 0; i < nn2; ++i) gamma_out[i] = gamma_out[i] + gout[i];
        }
        continue;
    }
    const int64_t K = static_cast<int64_t>(rows) * Rp;  // complex columns of one stacked block
    const size_t blk_bytes = static_cast<size_t>(K) * sizeof(double2);
    const StepPlan pH = make_plan(ctx, H, TBK_FP64, Rp, rows, b, false);
    const StepPlan pX = make_plan(ctx, vx, TBK_FP64, Rp, rows, 0.0, false);
    const StepPlan pY = make_plan(ctx, vy, TBK_FP64, Rp, rows, 0.0, false);
    const bool presplit = kubo_presplit(prec, N, H->is_complex, vx->is_complex) && pH.kc.kch == 0 && pX.kc.kch == 0;
    const bool fuse = fused_velocity(H, vx, pH, ctx->stream);
    // pre-split stacks with a fused v_x: both chains as one block of 2 Rp columns
    StepPlan pP{};
    bool paired = false, paired32 = false;
    if (presplit && fuse && 2 * Rp <= 256) {
        // the default launch shape (the context's tuning is chosen for Rp
        // columns) with the context's lane width, else 32-byte lanes: two
        // complex values per lane take the fp64 paired step on the 512 x 512
        // square lattice at R = 32 from 229 us to 211 us
        tbk_tuning pt{};
        pt.row_lane_bytes = ctx->tuning.row_lane_bytes > 0 ? ctx->tuning.row_lane_bytes : 32;
        // the fp32 recursion needs a unit's two complex64 on one side of the
        // block: at least two columns per chain
        paired32 = prec == TBK_GAMMA_BF16X3_FP32 && Rp >= 2;
        if (paired32) matrix_fp32(const_cast<tbk_matrix*>(vx));
        pP = make_plan(ctx, H, paired32 ? TBK_FP32 : TBK_FP64, 2 * Rp, rows, b, true, pt);
        paired = pP.kc.kch == 0;
        paired32 = paired32 && paired;
    }
    // fp64 stacks: entries at a stride of ldk elements; word stacks: k-blocked
    // (word_offset), each entry's share rounded up to whole word blocks
    const int64_t ldk = stack_stride(K);
    const int64_t kw = ((K + tbk::kWordBlock - 1) / tbk::kWordBlock) * tbk::kWordBlock;
    const size_t entry_bytes = presplit ? static_cast<size_t>(kw) * sizeof(uint2) : static_cast<size_t>(ldk) * sizeof(double2);
    const size_t stack_bytes = static_cast<size_t>(N) * entry_bytes;
    const int rolling = paired ? 7 : presplit ? 5 : 2;  // fp64 work blocks
    if (L.bytes < stack_bytes || Rs.bytes < stack_bytes) {
        // the stacks are reallocated at the new size; what the context already
        // holds for them is freed first and counts as available
        HIPC(hipStreamSynchronize(ctx->stream));
        L.release();
        Rs.release();
        size_t free_b = 0, total_b = 0;
        HIPC(hipMemGetInfo(&free_b, &total_b));
        const double need = 2.0 * static_cast<double>(stack_bytes) + rolling * static_cast<double>(blk_bytes);
        if (need > 0.95 * static_cast<double>(free_b))
            raise(TBK_E_OOM, "Kubo stacks need " + std::to_string(need / 1e9) + " GB, " +
                                 std::to_string(static_cast<double>(free_b) / 1e9) + " GB free");
        L.reserve(stack_bytes);
        Rs.reserve(stack_bytes);
    }
    ctx->work[0].reserve(blk_bytes);
    ctx->work[1].reserve(blk_bytes);
    if (presplit) {
        ctx->kwork[0].reserve(paired ? 2 * blk_bytes : blk_bytes);
        ctx->kwork[1].reserve(paired ? 2 * blk_bytes : blk_bytes);
        ctx->kwork[2].reserve(blk_bytes);
    }
    ctx->partial.reserve(static_cast<size_t>(std::max({pH.nb, pX.nb, pY.nb, paired ? pP.nb : int64_t{0}})) *
                         sizeof(double2));
    double2* part = ctx->partial.as<double2>();
    ensure_events(ctx, 2);
    HIPC(hipEventRecord(ctx->events[0], ctx->stream));
    if (paired)
        enqueue_kubo_stacks_paired(pP, pY, a, b, kp, L.as<uint2>(), Rs.as<uint2>(), ctx->kwork[0].p,
                                   ctx->kwork[1].p, ctx->kwork[2].p, rows, Rp, N, part, ctx->stream,
                                   paired32 ? vx->val32.p : vx->val64.p);
    else if (presplit)
        enqueue_kubo_stacks_presplit(pH, pX, pY, a, b, kp, L.as<uint2>(), Rs.as<uint2>(), ctx->kwork[0].p,
                                     ctx->kwork[1].p, ctx->kwork[2].p, ctx->work[0].p, ctx->work[1].p, K, N, part,
                                     ctx->stream, fuse ? vx->val64.p : nullptr);
    else
        enqueue_kubo_stacks(pH, pX, pY, a, b, kp, L.p, Rs.p, ctx->work[0].p, ctx->work[1].p, blk_bytes, ldk,
                            N, part, ctx->stream, fuse ? vx->val64.p : nullptr);
    HIPC(hipEventRecord(ctx->events[1], ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    float rec_ms = 0.f;
    HIPC(hipEventElapsedTime(&rec_ms, ctx->events[0], ctx->events[1]));

    const double con_ms = presplit ? contract_stacks_presplit(ctx, L.as<const uint2>(), Rs.as<const uint2>(), N, K,
                                                              gout, &nchunks)
                                   : contract_stacks(ctx, static_cast<const double2*>(L.p),
                                                     static_cast<const double2*>(Rs.p), N, K, ldk, prec, gout,
                                                     &nchunks);
    if (ngroups > 1) {
        const size_t nn2 = static_cast<size_t>(2) * N * N;
        if (grp == 0) std::copy(gout, gout + nn2, gamma_out);
        else for (size_t i = 0; i < nn2; ++i) gamma_out[i] = gamma_out[i] + gout[i];
    }
    rec_total += rec_ms;
    con_total += con_ms;
    flops_total += 8.0 * static_cast<double>(N) * N * static_cast<double>(K);
    stack_b = 2.0 * static_cast<double>(N) * static_cast<double>(K) *
              static_cast<double>(presplit ? sizeof(uint2) : sizeof(double2));
    rbytes_total += kubo_recursion_bytes(H, vx, vy, rows, Rreal, N, fuse, presplit, paired, paired32);
    any_fp32 = any_fp32 || paired32;
    }  // groups
    if (g) {
        for (int64_t n = 0; n < N; ++n)
            for (int64_t m = 0; m < N; ++m) {
                const double w = g[n] * g[m];
                gamma_out[2 * (n * N + m)] *= w;
                gamma_out[2 * (n * N + m) + 1] *= w;
            }
    }
    if (stats) {
        stats->recursion_ms = rec_total;
        stats->contraction_ms = con_total;
        stats->flops = flops_total;
        // the stacks' entries of one group, without the stride pad
        stats->stack_bytes = stack_b;
        stats->nchunks = nchunks;
        stats->vector_groups = ngroups;
        stats->row_ranges = nrc;
        stats->recursion_fp32 = any_fp32 ? 1 : 0;
        stats->recursion_bytes = rbytes_total;
        stats->wall_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
}

std::mutex g_default_mu;
tbk_context* g_default_ctx = nullptr;

}  // namespace

namespace engine {

// the process-wide context of the one-shot entries, on device 0
tbk_context* default_context() {
    std::lock_guard<std::mutex> lk(g_default_mu);
    if (!g_default_ctx) {
        const int rc = tbk_context_create(0, &g_default_ctx);
        if (rc != TBK_OK) raise(rc, g_last_error);
    }
    return g_default_ctx;
}

}  // namespace engine

// =================================================================== C ABI
extern "C" {

const char* tbk_version(void) { return "tbkpm 0.1 gfx950"; }
const char* tbk_last_error(void) { return g_last_error.c_str(); }

int tbk_device_count(int32_t* count) {
    return guarded([&] {
        if (!count) raise(TBK_E_INVALID, "null argument");
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        *count = (e == hipSuccess) ? c : 0;
    });
}

int tbk_context_create(int32_t device, tbk_context** out) {
    return guarded([&] {
        if (!out) raise(TBK_E_INVALID, "null argument");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) raise(TBK_E_DEVICE, "no HIP device");
        if (device < 0 || device >= count) raise(TBK_E_INVALID, "device index out of range");
        auto ctx = std::make_unique<tbk_context>();
        ctx->device = device;
        HIPC(hipSetDevice(device));
        HIPC(hipDeviceGetAttribute(&ctx->num_cu, hipDeviceAttributeMultiprocessorCount, device));
        HIPC(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        *out = ctx.release();
    });
}

int tbk_context_destroy(tbk_context* ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto e : ctx->events) (void)hipEventDestroy(e);
        ctx->work[0].release();
        ctx->work[1].release();
        ctx->partial.release();
        ctx->sums.release();
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int tbk_context_set_tuning(tbk_context* ctx, const tbk_tuning* t) {
    return guarded([&] {
        if (!ctx || !t) raise(TBK_E_INVALID, "null argument");
        if (t->slab_cols < 0 || t->blocks_per_cu < 0 || t->blocks_per_cu > 64 || t->l2_budget_bytes < 0)
            raise(TBK_E_INVALID, "tuning value out of range");
        if (t->gamma_tile_rows != 0 && t->gamma_tile_rows != 128 && t->gamma_tile_rows != 256)
            raise(TBK_E_INVALID, "gamma_tile_rows must be 0, 128 or 256");
        ctx->tuning = *t;
    });
}

int tbk_context_get_tuning(const tbk_context* ctx, tbk_tuning* t) {
    return guarded([&] {
        if (!ctx || !t) raise(TBK_E_INVALID, "null argument");
        *t = ctx->tuning;
    });
}

int tbk_matrix_info(const tbk_matrix* m, int32_t* rows, int64_t* nnz, int32_t* max_row_nnz, int32_t* bandwidth95) {
    return guarded([&] {
        if (!m) raise(TBK_E_INVALID, "null argument");
        if (rows) *rows = m->rows;
        if (nnz) *nnz = m->nnz;
        if (max_row_nnz) *max_row_nnz = m->max_row_nnz;
        if (bandwidth95) *bandwidth95 = m->bandwidth95;
    });
}

int tbk_context_synchronize(tbk_context* ctx) {
    return guarded([&] {
        if (!ctx) raise(TBK_E_INVALID, "null argument");
        set_device(ctx);
        HIPC(hipStreamSynchronize(ctx->stream));
    });
}

int tbk_context_memory(tbk_context* ctx, int64_t* bytes) {
    return guarded([&] {
        if (!ctx || !bytes) raise(TBK_E_INVALID, "null argument");
        *bytes = ctx->held_bytes + static_cast<int64_t>(ctx->work[0].bytes + ctx->work[1].bytes +
                                                        ctx->partial.bytes + ctx->sums.bytes +
                                                        ctx->stack[0].bytes + ctx->stack[1].bytes +
                                                        ctx->gamma_part.bytes + ctx->gamma_dev.bytes);
    });
}

int tbk_context_release_work(tbk_context* ctx) {
    return guarded([&] {
        if (!ctx) raise(TBK_E_INVALID, "null argument");
        set_device(ctx);
        HIPC(hipStreamSynchronize(ctx->stream));
        for (DevBuf* b : {&ctx->work[0], &ctx->work[1], &ctx->partial, &ctx->sums, &ctx->stack[0], &ctx->stack[1],
                          &ctx->kwork[0], &ctx->kwork[1], &ctx->kwork[2], &ctx->gamma_part, &ctx->gamma_dev})
            b->release();
    });
}

int tbk_matrix_create(tbk_context* ctx, const tbk_sparse* M, tbk_matrix** out) {
    return guarded([&] {
        if (!ctx || !M || !out) raise(TBK_E_INVALID, "null argument");
        if (M->rows <= 0 || M->rows != M->cols) raise(TBK_E_INVALID, "matrix must be square and non-empty");
        if (M->nnz < 0 || M->nnz > INT32_MAX) raise(TBK_E_INVALID, "nnz out of range for int32 indices");
        if (M->layout != TBK_CSC && M->layout != TBK_CSR) raise(TBK_E_INVALID, "unknown layout");
        if (!M->outer || (M->nnz > 0 && (!M->inner || !M->values))) raise(TBK_E_INVALID, "null array");
        std::vector<int32_t> rp, ci;
        std::vector<double> val;
        build_csr(M, rp, ci, val);
        set_device(ctx);
        auto m = std::make_unique<tbk_matrix>();
        m->ctx = ctx;
        m->rows = M->rows;
        m->nnz = M->nnz;
        m->is_complex = M->is_complex != 0;
        for (int32_t i = 0; i < M->rows; ++i) m->max_row_nnz = std::max(m->max_row_nnz, rp[i + 1] - rp[i]);
        m->bandwidth95 = bandwidth_percentile(rp, ci, 0.95);
        m->rp.reserve_placed(rp.size() * 4);
        m->ci.reserve_placed(std::max<size_t>(ci.size() * 4, 4));
        m->val64.reserve_placed(std::max<size_t>(val.size() * 8, 8));
        HIPC(hipMemcpy(m->rp.p, rp.data(), rp.size() * 4, hipMemcpyHostToDevice));
        if (!ci.empty()) {
            HIPC(hipMemcpy(m->ci.p, ci.data(), ci.size() * 4, hipMemcpyHostToDevice));
            HIPC(hipMemcpy(m->val64.p, val.data(), val.size() * 8, hipMemcpyHostToDevice));
        }
        ctx->held_bytes += static_cast<int64_t>(m->rp.bytes + m->ci.bytes + m->val64.bytes);
        *out = m.release();
    });
}

int tbk_matrix_destroy(tbk_matrix* m) {
    return guarded([&] {
        if (!m) return;
        set_device(m->ctx);
        HIPC(hipStreamSynchronize(m->ctx->stream));
        m->ctx->held_bytes -= static_cast<int64_t>(m->rp.bytes + m->ci.bytes + m->val64.bytes + m->val32.bytes);
        delete m;
    });
}

int tbk_block_create(tbk_context* ctx, const double* colmajor, int32_t rows, int32_t R,
                     int32_t precision, tbk_block** out) {
    return guarded([&] {
        if (!ctx || !out) raise(TBK_E_INVALID, "null argument");
        if (rows <= 0 || R <= 0) raise(TBK_E_INVALID, "block must have rows >= 1 and R >= 1");
        if (precision != TBK_FP64 && precision != TBK_FP32) raise(TBK_E_INVALID, "unknown precision");
        if (R > 256) raise(TBK_E_UNSUPPORTED, "vector block wider than 256 columns");
        set_device(ctx);
        auto b = std::make_unique<tbk_block>();
        b->ctx = ctx;
        b->rows = rows;
        b->R = R;
        b->Rp = pow2_at_least(R);
        b->precision = precision;
        const size_t elems = static_cast<size_t>(rows) * b->Rp;
        b->data.reserve(elems * b->elem_bytes());
        if (precision == TBK_FP64) {
            std::vector<double2> h(elems, make_double2(0.0, 0.0));
            if (colmajor)
                for (int r = 0; r < R; ++r)
                    for (int i = 0; i < rows; ++i) {
                        const size_t s = 2 * (static_cast<size_t>(r) * rows + i);
                        h[static_cast<size_t>(i) * b->Rp + r] = make_double2(colmajor[s], colmajor[s + 1]);
                    }
            HIPC(hipMemcpy(b->data.p, h.data(), elems * sizeof(double2), hipMemcpyHostToDevice));
        } else {
            std::vector<float2> h(elems, make_float2(0.f, 0.f));
            if (colmajor)
                for (int r = 0; r < R; ++r)
                    for (int i = 0; i < rows; ++i) {
                        const size_t s = 2 * (static_cast<size_t>(r) * rows + i);
                        h[static_cast<size_t>(i) * b->Rp + r] =
                            make_float2(static_cast<float>(colmajor[s]), static_cast<float>(colmajor[s + 1]));
                    }
            HIPC(hipMemcpy(b->data.p, h.data(), elems * sizeof(float2), hipMemcpyHostToDevice));
        }
        ctx->held_bytes += static_cast<int64_t>(b->data.bytes);
        *out = b.release();
    });
}

int tbk_block_random_phase(tbk_context* ctx, int32_t rows, int32_t R, uint64_t seed, int32_t precision,
                           tbk_block** out) {
    return guarded([&] {
        if (!ctx || !out) raise(TBK_E_INVALID, "null argument");
        if (rows <= 0 || R <= 0) raise(TBK_E_INVALID, "block must have rows >= 1 and R >= 1");
        if (precision != TBK_FP64 && precision != TBK_FP32) raise(TBK_E_INVALID, "unknown precision");
        if (R > 256) raise(TBK_E_UNSUPPORTED, "vector block wider than 256 columns");
        set_device(ctx);
        auto b = std::make_unique<tbk_block>();
        b->ctx = ctx;
        b->rows = rows;
        b->R = R;
        b->Rp = pow2_at_least(R);
        b->precision = precision;
        const size_t elems = static_cast<size_t>(rows) * b->Rp;
        b->data.reserve(elems * b->elem_bytes());
        const int grid = static_cast<int>(std::min<size_t>((elems + tbk::kBlock - 1) / tbk::kBlock, 65536));
        const double inv = 1.0 / std::sqrt(static_cast<double>(rows));
        if (precision == TBK_FP64)
            hipLaunchKernelGGL(tbk::random_phase_kernel<double2>, dim3(grid), dim3(tbk::kBlock), 0, ctx->stream,
                               b->data.as<double2>(), rows, R, b->Rp, seed, inv, R, 0);
        else
            hipLaunchKernelGGL(tbk::random_phase_kernel<float2>, dim3(grid), dim3(tbk::kBlock), 0, ctx->stream,
                               b->data.as<float2>(), rows, R, b->Rp, seed, inv, R, 0);
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(ctx->stream));
        ctx->held_bytes += static_cast<int64_t>(b->data.bytes);
        *out = b.release();
    });
}

int tbk_block_download(const tbk_block* b, double* colmajor_out) {
    return guarded([&] {
        if (!b || !colmajor_out) raise(TBK_E_INVALID, "null argument");
        set_device(b->ctx);
        HIPC(hipStreamSynchronize(b->ctx->stream));
        const size_t elems = static_cast<size_t>(b->rows) * b->Rp;
        if (b->precision == TBK_FP64) {
            std::vector<double2> h(elems);
            HIPC(hipMemcpy(h.data(), b->data.p, elems * sizeof(double2), hipMemcpyDeviceToHost));
            for (int r = 0; r < b->R; ++r)
                for (int i = 0; i < b->rows; ++i) {
                    const size_t s = 2 * (static_cast<size_t>(r) * b->rows + i);
                    colmajor_out[s] = h[static_cast<size_t>(i) * b->Rp + r].x;
                    colmajor_out[s + 1] = h[static_cast<size_t>(i) * b->Rp + r].y;
                }
        } else {
            std::vector<float2> h(elems);
            HIPC(hipMemcpy(h.data(), b->data.p, elems * sizeof(float2), hipMemcpyDeviceToHost));
            for (int r = 0; r < b->R; ++r)
                for (int i = 0; i < b->rows; ++i) {
                    const size_t s = 2 * (static_cast<size_t>(r) * b->rows + i);
                    colmajor_out[s] = h[static_cast<size_t>(i) * b->Rp + r].x;
                    colmajor_out[s + 1] = h[static_cast<size_t>(i) * b->Rp + r].y;
                }
        }
    });
}

int tbk_block_copy(const tbk_block* src, tbk_block* dst, float* kernel_ms) {
    return guarded([&] {
        if (!src || !dst) raise(TBK_E_INVALID, "null argument");
        if (src->ctx != dst->ctx) raise(TBK_E_INVALID, "blocks belong to different contexts");
        if (src->rows != dst->rows || src->Rp != dst->Rp || src->precision != dst->precision)
            raise(TBK_E_INVALID, "block shapes differ");
        tbk_context* ctx = src->ctx;
        set_device(ctx);
        const size_t nbytes = static_cast<size_t>(src->rows) * src->Rp * src->elem_bytes();
        const int64_t n16 = static_cast<int64_t>(nbytes / 16);
        const size_t tail = nbytes % 16;  // one complex64 when rows * Rp is odd
        const int64_t per_block = static_cast<int64_t>(tbk::kBlock) * tbk::kCopyUnroll;
        if ((n16 + per_block - 1) / per_block > INT32_MAX) raise(TBK_E_UNSUPPORTED, "block too large to copy");
        const int grid = static_cast<int>((n16 + per_block - 1) / per_block);
        ensure_events(ctx, 2);
        HIPC(hipEventRecord(ctx->events[0], ctx->stream));
        if (n16 > 0)
            hipLaunchKernelGGL(tbk::copy16_kernel, dim3(std::max(grid, 1)), dim3(tbk::kBlock), 0, ctx->stream,
                               src->data.as<const uint4>(), dst->data.as<uint4>(), n16);
        HIPC(hipGetLastError());
        if (tail)
            HIPC(hipMemcpyAsync(static_cast<char*>(dst->data.p) + nbytes - tail,
                                static_cast<const char*>(src->data.p) + nbytes - tail, tail,
                                hipMemcpyDeviceToDevice, ctx->stream));
        HIPC(hipEventRecord(ctx->events[1], ctx->stream));
        HIPC(hipStreamSynchronize(ctx->stream));
        if (kernel_ms) HIPC(hipEventElapsedTime(kernel_ms, ctx->events[0], ctx->events[1]));
    });
}

int tbk_bandwidth_probe(tbk_context* ctx, int64_t bytes, int32_t repeats, tbk_bandwidth* out) {
    return guarded([&] {
        if (!ctx || !out) raise(TBK_E_INVALID, "null argument");
        if (bytes < 16 * 1024 || repeats < 1) raise(TBK_E_INVALID, "bytes must be >= 16 KiB and repeats >= 1");
        set_device(ctx);
        const int64_t n = bytes / 16;
        DevBuf a, b, y, r;
        a.reserve(static_cast<size_t>(n) * 16);
        b.reserve(static_cast<size_t>(n) * 16);
        y.reserve(static_cast<size_t>(n) * 16);
        const unsigned g4 = static_cast<unsigned>(ctx->num_cu) * 4, g8 = static_cast<unsigned>(ctx->num_cu) * 8;
        r.reserve(static_cast<size_t>(g8) * 8);
        hipStream_t st = ctx->stream;
        HIPC(hipMemsetAsync(a.p, 0, static_cast<size_t>(n) * 16, st));
        HIPC(hipMemsetAsync(b.p, 0, static_cast<size_t>(n) * 16, st));
        ensure_events(ctx, 2);
        auto best_ms = [&](auto launch) {
            launch();  // warm
            float best = 1e30f;
            for (int k = 0; k < repeats; ++k) {
                HIPC(hipEventRecord(ctx->events[0], st));
                launch();
                HIPC(hipEventRecord(ctx->events[1], st));
                HIPC(hipEventSynchronize(ctx->events[1]));
                float ms = 0.f;
                HIPC(hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
                best = std::min(best, ms);
            }
            return static_cast<double>(best);
        };
        using tbk::nt_d2;
        const double tc = best_ms([&] {
            hipLaunchKernelGGL(tbk::stream_copy_kernel, dim3(g4), dim3(tbk::kBlock), 0, st, a.as<const nt_d2>(),
                               y.as<nt_d2>(), n);
        });
        const double t2 = best_ms([&] {
            hipLaunchKernelGGL(tbk::stream_2r1w_kernel, dim3(g4), dim3(tbk::kBlock), 0, st, a.as<const nt_d2>(),
                               b.as<const nt_d2>(), y.as<nt_d2>(), n);
        });
        const double tr = best_ms([&] {
            hipLaunchKernelGGL(tbk::stream_read_kernel, dim3(g8), dim3(tbk::kBlock), 0, st, a.as<const nt_d2>(),
                               r.as<double>(), n);
        });
        HIPC(hipGetLastError());
        const double B = static_cast<double>(n) * 16.0;
        out->copy_GBps = 2.0 * B / (tc * 1e-3) / 1e9;
        out->stream_2r1w_GBps = 3.0 * B / (t2 * 1e-3) / 1e9;
        out->read_GBps = B / (tr * 1e-3) / 1e9;
        out->bytes = B;
    });
}

int tbk_block_destroy(tbk_block* b) {
    return guarded([&] {
        if (!b) return;
        set_device(b->ctx);
        HIPC(hipStreamSynchronize(b->ctx->stream));
        b->ctx->held_bytes -= static_cast<int64_t>(b->data.bytes);
        delete b;
    });
}

int tbk_moments(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* bras,
                const tbk_block* kets, int32_t n, int32_t flags, double* mu_out, double* last_out,
                tbk_stats* stats) {
    return guarded([&] { run_moments(ctx, M, a, b, bras, kets, n, flags, mu_out, last_out, stats); });
}

int tbk_cheb_seq_on_bases(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* v,
                          int32_t n, tbk_block* const* out, float* kernel_ms) {
    return guarded([&] {
        check_on_bases_args(ctx, M, a, b, v, n);
        if (!out) raise(TBK_E_INVALID, "null argument");
        std::vector<const tbk_block*> seen;
        seen.reserve(static_cast<size_t>(n));
        for (int32_t k = 0; k < n; ++k) {
            const tbk_block* o = out[k];
            if (!o) raise(TBK_E_INVALID, "null output block");
            if (o->ctx != ctx) raise(TBK_E_INVALID, "objects belong to another context");
            if (o->rows != v->rows || o->R != v->R || o->Rp != v->Rp || o->precision != v->precision)
                raise(TBK_E_INVALID, "output block shape or precision does not match v");
            if (k > 0 && o == v) raise(TBK_E_INVALID, "only out[0] may be v itself");
            seen.push_back(o);
        }
        std::sort(seen.begin(), seen.end());
        if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
            raise(TBK_E_INVALID, "output blocks must be distinct");
        const double ms = enqueue_on_bases(
            ctx, M, a, b, v, n, [&](int64_t k) -> void* { return out[k]->data.p; }, [](int64_t) {});
        if (kernel_ms) *kernel_ms = static_cast<float>(ms);
    });
}

int tbk_cheb_step(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* x,
                  const tbk_block* z, tbk_block* y) {
    return guarded([&] {
        check_on_bases_args(ctx, M, a, b, x, 1);
        if (!y) raise(TBK_E_INVALID, "null argument");
        auto alike = [&](const tbk_block* o) {
            return o->ctx == ctx && o->rows == x->rows && o->R == x->R && o->Rp == x->Rp &&
                   o->precision == x->precision;
        };
        if (!alike(y) || (z && !alike(z))) raise(TBK_E_INVALID, "block shape, precision or context differs from x");
        if (y == x || z == x) raise(TBK_E_INVALID, "y and z must not be x");
        set_device(ctx);
        const StepPlan plan = make_plan(ctx, M, x->precision, x->Rp, x->rows, b, false);
        ctx->partial.reserve(static_cast<size_t>(plan.nb) * sizeof(double2));
        double2* part = ctx->partial.as<double2>();
        if (!z)
            launch_plan(plan, ctx->stream, tbk::kFirst, 1.0 / a, -b / a, x->data.p, nullptr, y->data.p, x->data.p,
                        part);
        else
            launch_plan(plan, ctx->stream, tbk::kRecDbl, 2.0 / a, -2.0 * b / a, x->data.p, z->data.p, y->data.p,
                        nullptr, part);
    });
}

int tbk_cheb_expand(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* v, int32_t n,
                    const double* coef, tbk_block* out, float* kernel_ms) {
    return guarded([&] {
        check_on_bases_args(ctx, M, a, b, v, n);
        if (!coef || !out) raise(TBK_E_INVALID, "null argument");
        if (out->ctx != ctx || out->rows != v->rows || out->R != v->R || out->Rp != v->Rp ||
            out->precision != v->precision)
            raise(TBK_E_INVALID, "out block shape, precision or context differs from v");
        if (out == v) raise(TBK_E_INVALID, "out must not be v");
        for (int64_t k = 0; k < 2 * static_cast<int64_t>(n); ++k)
            if (!std::isfinite(coef[k])) raise(TBK_E_INVALID, "non-finite coefficient");
        set_device(ctx);
        const size_t block_bytes = static_cast<size_t>(v->rows) * v->Rp * v->elem_bytes();
        ctx->work[0].reserve(block_bytes);
        ctx->work[1].reserve(block_bytes);
        const int64_t elems = static_cast<int64_t>(v->rows) * v->Rp;
        const unsigned grid = static_cast<unsigned>(std::min<int64_t>((elems + tbk::kBlock - 1) / tbk::kBlock, 65536));
        // T_k in v (k = 0) or the rolling work blocks; out += c_k T_k right
        // after T_k is formed, in order of k
        auto buf = [&](int64_t k) -> void* { return k == 0 ? v->data.p : ctx->work[k & 1].p; };
        auto add = [&](int64_t k) {
            const double cr = coef[2 * k], ci = coef[2 * k + 1];
            if (v->precision == TBK_FP64)
                hipLaunchKernelGGL(tbk::block_axpy_kernel<double2>, dim3(grid), dim3(tbk::kBlock), 0, ctx->stream,
                                   out->data.as<double2>(), static_cast<const double2*>(buf(k)), cr, ci, elems,
                                   k == 0 ? 1 : 0);
            else
                hipLaunchKernelGGL(tbk::block_axpy_kernel<float2>, dim3(grid), dim3(tbk::kBlock), 0, ctx->stream,
                                   out->data.as<float2>(), static_cast<const float2*>(buf(k)), cr, ci, elems,
                                   k == 0 ? 1 : 0);
            HIPC(hipGetLastError());
        };
        const double ms = enqueue_on_bases(ctx, M, a, b, v, n, buf, add);
        if (kernel_ms) *kernel_ms = static_cast<float>(ms);
    });
}

int tbk_autotune_flags(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* kets,
                       int32_t steps, int32_t flags, tbk_tuning* best, double* best_ms_per_step) {
    return guarded([&] {
        if (!ctx || !M || !kets) raise(TBK_E_INVALID, "null argument");
        if (steps < 2) raise(TBK_E_INVALID, "autotune needs at least 2 steps per trial");
        if (flags & ~(TBK_FLAG_PLAIN | TBK_AUTOTUNE_REFRESH)) raise(TBK_E_INVALID, "unknown autotune flag");
        const bool plain = (flags & TBK_FLAG_PLAIN) != 0;
        // a shape tuned before for the same matrix statistics, width,
        // precision and mode on this device is installed without trials
        const TuningKey key = tuning_key(ctx, M, kets, plain);
        if (!(flags & TBK_AUTOTUNE_REFRESH)) {
            TuningEntry hit;
            if (tuning_cache_find(key, hit)) {
                ctx->tuning = hit.tuning;
                if (best) *best = hit.tuning;
                if (best_ms_per_step) *best_ms_per_step = hit.ms_per_step;
                return;
            }
        }
        // two moments per step (bras = kets), or one with TBK_FLAG_PLAIN
        const int32_t n = plain ? steps + 1 : 2 * steps;
        const int32_t mode = plain ? TBK_FLAG_PLAIN : 0;
        std::vector<double> mu(static_cast<size_t>(n));
        struct Restore {  // a failed trial leaves the caller's tuning in place
            tbk_context* c;
            tbk_tuning saved;
            bool done = false;
            ~Restore() {
                if (!done) c->tuning = saved;
            }
        } restore{ctx, ctx->tuning};
        auto trial = [&](const tbk_tuning& t) -> double {
            ctx->tuning = t;
            tbk_stats st{};
            run_moments(ctx, M, a, b, nullptr, kets, 4, mode, mu.data(), nullptr, nullptr);  // warm
            run_moments(ctx, M, a, b, nullptr, kets, n, mode | TBK_FLAG_TIME_LAUNCHES, mu.data(), nullptr, &st);
            return st.kernel_ms / static_cast<double>(st.steps);
        };
        tbk_tuning cur = ctx->tuning;
        double cur_ms = trial(cur);
        // coordinate descent: each knob in turn, keep the fastest value
        struct Knob {
            int32_t tbk_tuning::*field;
            std::vector<int32_t> values;
        };
        const std::vector<Knob> knobs = {
            {&tbk_tuning::use_tiled, {0, 1}},
            {&tbk_tuning::blocks_per_cu, {2, 4, 8, 16}},
            {&tbk_tuning::chunk_rows, {64, 128, 256, 512}},
            {&tbk_tuning::chunk_period, {0, -1}},
            {&tbk_tuning::xcd_split, {0, -1}},
            {&tbk_tuning::stream_nt, {1, -1}},
            // lanes narrower than one element are the same kernel as 16 bytes in
            // fp64; offering them only lets noise pick a width that would slow
            // fp32 work sharing the context
            {&tbk_tuning::row_lane_bytes,
             kets->precision == TBK_FP64 ? std::vector<int32_t>{16, 32} : std::vector<int32_t>{8, 16, 32}},
            // column sweeps of the row kernel: a sweep's gather window shrinks
            // with its width (widths below the padded R only)
            {&tbk_tuning::row_slab_cols, {0, 16, 32, 64}},
        };
        for (int pass = 0; pass < 2; ++pass) {
            for (const auto& k : knobs) {
                for (int32_t v : k.values) {
                    if (cur.*(k.field) == v) continue;
                    if (k.field == &tbk_tuning::row_slab_cols && v >= kets->Rp) continue;
                    tbk_tuning t = cur;
                    t.*(k.field) = v;
                    const double ms = trial(t);
                    if (ms < 0.99 * cur_ms) {  // a change must win by more than 1%
                        cur_ms = ms;
                        cur = t;
                    }
                }
            }
        }
        ctx->tuning = cur;
        restore.done = true;
        tuning_cache_store(key, TuningEntry{cur, cur_ms});
        if (best) *best = cur;
        if (best_ms_per_step) *best_ms_per_step = cur_ms;
    });
}

int tbk_autotune(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* kets,
                 int32_t steps, tbk_tuning* best, double* best_ms_per_step) {
    return tbk_autotune_flags(ctx, M, a, b, kets, steps, 0, best, best_ms_per_step);
}

int tbk_tuning_cache_save(const char* path) {
    return guarded([&] {
        if (!path) raise(TBK_E_INVALID, "null argument");
        tuning_cache_save(path);
    });
}

int tbk_tuning_cache_load(const char* path, int32_t* entries) {
    return guarded([&] {
        if (!path) raise(TBK_E_INVALID, "null argument");
        const int32_t k = tuning_cache_load(path);
        if (entries) *entries = k;
    });
}

int tbk_tuning_cache_clear(void) {
    return guarded([&] { tuning_cache_clear(); });
}

int tbk_kubo_gamma(tbk_context* ctx, const tbk_matrix* H, double a, double b, const tbk_matrix* vx,
                   const tbk_matrix* vy, const tbk_block* kets, int32_t N, int32_t gamma_precision,
                   const double* g, double* gamma_out, tbk_kubo_stats* stats) {
    return guarded([&] { run_kubo(ctx, H, a, b, vx, vy, kets, N, gamma_precision, g, gamma_out, stats); });
}

int tbk_stack_contract(tbk_context* ctx, const double* A, const double* B, int32_t N, int64_t K,
                       int32_t gamma_precision, double* gamma_out, double* contraction_ms) {
    return guarded([&] {
        if (!ctx || !A || !B || !gamma_out) raise(TBK_E_INVALID, "null argument");
        if (N < 1 || K < 1) raise(TBK_E_INVALID, "N and K must be positive");
        set_device(ctx);
        const size_t bytes = static_cast<size_t>(N) * static_cast<size_t>(K) * sizeof(double2);
        DevBuf dA, dB;
        dA.reserve(bytes);
        dB.reserve(bytes);
        HIPC(hipMemcpy(dA.p, A, bytes, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(dB.p, B, bytes, hipMemcpyHostToDevice));
        const double ms = contract_stacks(ctx, dA.as<const double2>(), dB.as<const double2>(), N, K, K,
                                          gamma_precision, gamma_out, nullptr);
        if (contraction_ms) *contraction_ms = ms;
    });
}

int tbk_kubo_zeta(tbk_context* ctx, const double* gamma, int32_t N, const double* e, int32_t n_e, double volume,
                  double* zeta_out, double* kernel_ms) {
    return guarded([&] {
        if (!ctx || !gamma || !e || !zeta_out) raise(TBK_E_INVALID, "null argument");
        if (N < 1 || n_e < 1) raise(TBK_E_INVALID, "N and n_e must be positive");
        if (!(volume != 0.0) || !std::isfinite(volume)) raise(TBK_E_INVALID, "invalid volume");
        for (int32_t j = 0; j < n_e; ++j)
            if (!(e[j] > -1.0 && e[j] < 1.0)) raise(TBK_E_INVALID, "energies must lie in (-1, 1)");
        if (tbk::kubo_zeta_energies_per_group(N) == 0)
            raise(TBK_E_UNSUPPORTED, "N too large for the zeta kernel's LDS tables");
        set_device(ctx);
        const size_t gbytes = static_cast<size_t>(N) * N * sizeof(double2);
        DevBuf dG, dE, dZ;
        dG.reserve(gbytes);
        dE.reserve(static_cast<size_t>(n_e) * sizeof(double));
        dZ.reserve(static_cast<size_t>(n_e) * sizeof(double2));
        HIPC(hipMemcpyAsync(dG.p, gamma, gbytes, hipMemcpyHostToDevice, ctx->stream));
        HIPC(hipMemcpyAsync(dE.p, e, static_cast<size_t>(n_e) * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        ensure_events(ctx, 2);
        HIPC(hipEventRecord(ctx->events[0], ctx->stream));
        tbk::launch_kubo_zeta(dG.as<const double2>(), N, dE.as<const double>(), n_e, 1.0 / volume,
                              dZ.as<double2>(), ctx->stream);
        HIPC(hipGetLastError());
        HIPC(hipEventRecord(ctx->events[1], ctx->stream));
        HIPC(hipMemcpyAsync(zeta_out, dZ.p, static_cast<size_t>(n_e) * sizeof(double2), hipMemcpyDeviceToHost,
                            ctx->stream));
        HIPC(hipStreamSynchronize(ctx->stream));
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
        if (kernel_ms) *kernel_ms = ms;
    });
}

int tbk_cheb_first_kind_seq_braket(const tbk_sparse* M, int32_t n, const double* bras_T, const double* kets,
                                   int32_t ket_rows, int32_t ket_cols, double* mu_out) {
    return guarded([&] {
        if (!M || !bras_T || !kets || !mu_out) raise(TBK_E_INVALID, "null argument");
        if (n < 2) raise(TBK_E_INVALID, "n must be greater than or equal to 2");
        if (ket_rows != M->rows || ket_rows != M->cols) raise(TBK_E_INVALID, "kets rows do not match M");
        tbk_context* ctx = default_context();
        struct Guard {
            tbk_matrix* m = nullptr;
            tbk_block* k = nullptr;
            tbk_block* b = nullptr;
            ~Guard() {
                if (b) tbk_block_destroy(b);
                if (k) tbk_block_destroy(k);
                if (m) tbk_matrix_destroy(m);
            }
        } g;
        int rc = tbk_matrix_create(ctx, M, &g.m);
        if (rc != TBK_OK) raise(rc, g_last_error);
        rc = tbk_block_create(ctx, kets, ket_rows, ket_cols, TBK_FP64, &g.k);
        if (rc != TBK_OK) raise(rc, g_last_error);
        // Two moments per step needs bra = ket and a Hermitian M (T_k(M)^H
        // T_k(M) = T_k(M)^2 = (T_2k + T_0)/2); otherwise one moment per step
        // against the bra block, as numerical.h:363-381 computes it.
        bool doubling = bras_T == kets;
        if (doubling) {
            std::vector<int32_t> rp, ci;
            std::vector<double> val;
            build_csr(M, rp, ci, val);
            doubling = csr_is_hermitian(M->rows, rp, ci, val, M->is_complex != 0);
        }
        if (!doubling) {
            rc = tbk_block_create(ctx, bras_T, ket_rows, ket_cols, TBK_FP64, &g.b);
            if (rc != TBK_OK) raise(rc, g_last_error);
        }
        run_moments(ctx, g.m, 1.0, 0.0, g.b, g.k, n, 0, mu_out, nullptr, nullptr);
    });
}

int tbk_cheb_first_kind_seq_on_bases(const tbk_sparse* M, int32_t n, const double* v, int32_t v_rows,
                                     int32_t v_cols, double* out) {
    return guarded([&] {
        if (!M || !v || !out) raise(TBK_E_INVALID, "null argument");
        if (n < 1) raise(TBK_E_INVALID, "n must be at least 1");
        if (v_rows != M->rows || v_rows != M->cols) raise(TBK_E_INVALID, "v rows do not match M");
        tbk_context* ctx = default_context();
        struct Guard {
            tbk_matrix* m = nullptr;
            tbk_block* k = nullptr;
            ~Guard() {
                if (k) tbk_block_destroy(k);
                if (m) tbk_matrix_destroy(m);
            }
        } g;
        int rc = tbk_matrix_create(ctx, M, &g.m);
        if (rc != TBK_OK) raise(rc, g_last_error);
        rc = tbk_block_create(ctx, v, v_rows, v_cols, TBK_FP64, &g.k);
        if (rc != TBK_OK) raise(rc, g_last_error);
        const tbk_block* vb = g.k;
        const size_t block_bytes = static_cast<size_t>(vb->rows) * vb->Rp * vb->elem_bytes();
        ctx->work[0].reserve(block_bytes);
        ctx->work[1].reserve(block_bytes);
        // T_k over T_{k-2} in two rolling blocks, each downloaded once formed
        auto dst = [&](int64_t k) -> void* { return k == 0 ? vb->data.p : ctx->work[k & 1].p; };
        const size_t blk = 2 * static_cast<size_t>(v_rows) * v_cols;
        std::copy(v, v + blk, out);
        enqueue_on_bases(ctx, g.m, 1.0, 0.0, vb, n, dst, [&](int64_t k) {
            if (k == 0) return;
            HIPC(hipStreamSynchronize(ctx->stream));
            block_to_colmajor(dst(k), TBK_FP64, vb->rows, vb->R, vb->Rp, out + static_cast<size_t>(k) * blk);
        });
    });
}

}  // extern "C"
