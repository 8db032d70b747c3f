// Batched ragged moment problems: many independent Hamiltonians of different
// sizes and moment counts run as one block-diagonal recursion, one batched
// step launch per Chebyshev step for all problems still active.
//
// Layout.  Problem p occupies rows [row_begin_p, row_begin_p + rows_p) of the
// concatenated CSR (column indices offset by row_begin_p), padded with empty
// rows to a multiple of kBatchTileRows, so every tile of kBatchTileRows rows
// belongs to one problem.  Problems are laid out in order of decreasing step
// count: the problems still active at step k are a prefix of the layout, and
// a step launch covers the tiles of that prefix only.  Vector blocks are
// row-major [row][Rp] over the concatenated rows (padding rows stay zero).
//
// Moments.  Each wave writes its partial sums of a tile to its own slot;
// batch_moments_kernel adds a problem's slots in a fixed order and applies
// the doubling identities, so a problem's moments do not depend on the other
// problems in the batch.  Its vectors T_n are those of a single-problem run
// (same per-row arithmetic and rescaling).
//
// Graphs.  With TBK_BATCH_GRAPHS the launches of each chunk of 64 steps are
// captured once into a hipGraph (no memset nodes) and replayed on later runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "aux_kernels.h"
#include "engine.h"

namespace {
using namespace engine;

constexpr int32_t kBatchTileRows = 256;  // multiple of every workgroup pass (4 waves x 64/LPR rows)
constexpr int64_t kChunkSteps = 64;

// One workgroup per active problem: s = sum of its wave slots in a fixed
// order; step 0 gives mu_0, mu_1; step k >= 1 gives mu_2k = 2 s.x - mu_0 and
// mu_2k+1 = 2 s.y - mu_1 (indices below the problem's moment count).
__global__ __launch_bounds__(tbk::kBlock) void batch_moments_kernel(const double2* __restrict__ partial,
                                                                    const int64_t* __restrict__ tile_begin,
                                                                    const int64_t* __restrict__ tile_end,
                                                                    const int64_t* __restrict__ mu_off,
                                                                    const int32_t* __restrict__ n_mom, int64_t step,
                                                                    double* __restrict__ mu) {
    const int pos = blockIdx.x;
    const int64_t s0 = 4 * tile_begin[pos], s1 = 4 * tile_end[pos];
    double a0 = 0.0, a1 = 0.0;
    for (int64_t i = s0 + threadIdx.x; i < s1; i += tbk::kBlock) {
        a0 = a0 + partial[i].x;
        a1 = a1 + partial[i].y;
    }
    __shared__ double r0[tbk::kBlock], r1[tbk::kBlock];
    r0[threadIdx.x] = a0;
    r1[threadIdx.x] = a1;
    __syncthreads();
    for (int w = tbk::kBlock / 2; w >= 1; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            r0[threadIdx.x] = r0[threadIdx.x] + r0[threadIdx.x + w];
            r1[threadIdx.x] = r1[threadIdx.x] + r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* m = mu + mu_off[pos];
        const int64_t n = n_mom[pos];
        if (step == 0) {
            m[0] = r0[0];
            m[1] = r1[0];
        } else {
            if (2 * step < n) m[2 * step] = 2.0 * r0[0] - m[0];
            if (2 * step + 1 < n) m[2 * step + 1] = 2.0 * r1[0] - m[1];
        }
    }
}

}  // namespace

struct tbk_batch {
    tbk_context* ctx = nullptr;
    int32_t np = 0, R = 0, Rp = 0, precision = TBK_FP64;
    bool cplx = false;
    int64_t rows_total = 0, ntiles = 0, max_steps = 0;
    // per layout position (problems sorted by decreasing steps)
    std::vector<int32_t> prob;          // problem index of each position
    std::vector<int64_t> row_begin, rows, tile_begin, tile_end, steps, nnz;
    std::vector<uint64_t> seed;
    std::vector<int32_t> nmom;
    std::vector<int64_t> mu_off;        // offset of the problem's moments in mu_out (problem order)
    int64_t mu_total = 0;
    DevBuf rp, ci, val, tile_prob, coef, d_tb, d_te, d_off, d_n, mu, partial, buf[3];
    // launch shape
    int32_t cpl = 1, bpc = 4, xcd = 1;
    std::vector<hipGraphExec_t> graphs;  // per chunk, captured on first use
    size_t elem() const { return precision == TBK_FP64 ? 16 : 8; }

    void drop_graphs() {
        for (hipGraphExec_t g : graphs)
            if (g) (void)hipGraphExecDestroy(g);
        graphs.clear();
    }
    ~tbk_batch() { drop_graphs(); }

    // active positions at step k (prefix length)
    int32_t active(int64_t k) const {
        int32_t lo = 0, hi = np;  // steps is non-increasing
        while (lo < hi) {
            const int32_t mid = (lo + hi) / 2;
            if (steps[mid] > k) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    }

    int64_t grid_blocks(int64_t nt) const {
        int64_t nb = std::min<int64_t>(nt, static_cast<int64_t>(ctx->num_cu) * bpc);
        if (xcd && nb >= 8) nb = (nb / 8) * 8;
        return std::max<int64_t>(nb, 1);
    }

    void launch_step(int64_t k, hipStream_t st) {
        const int32_t na = active(k);
        if (na == 0) return;
        tbk::BatchArgs A{};
        A.rp = rp.as<const int32_t>();
        A.ci = ci.as<const int32_t>();
        A.hv = val.p;
        A.x = buf[k % 3].p;
        A.z = buf[(k + 2) % 3].p;
        A.y = buf[(k + 1) % 3].p;
        A.rows = static_cast<int32_t>(rows_total);
        A.Rp = Rp;
        A.tile_prob = tile_prob.as<const int32_t>();
        A.coef = coef.as<const double>();
        A.tile_rows = kBatchTileRows;
        A.opts = (ctx->tuning.stream_nt >= 0 ? tbk::kOptStream : 0);
        A.ntiles = tile_end[na - 1];
        const int64_t nb = grid_blocks(A.ntiles);
        A.xcd_split = (xcd && nb >= 8) ? 1 : 0;
        A.partial = partial.as<double2>();
        const int mode = k == 0 ? tbk::kFirst : tbk::kRecDbl;
        bool ok;
        const dim3 grid(static_cast<unsigned>(nb));
        if (precision == TBK_FP64)
            ok = cplx ? tbk::launch_batch_f64_cplx(Rp, cpl, mode, grid, st, A)
                      : tbk::launch_batch_f64_real(Rp, cpl, mode, grid, st, A);
        else
            ok = cplx ? tbk::launch_batch_f32_cplx(Rp, cpl, mode, grid, st, A)
                      : tbk::launch_batch_f32_real(Rp, cpl, mode, grid, st, A);
        if (!ok) raise(TBK_E_UNSUPPORTED, "no batched step kernel for this width");
        HIPC(hipGetLastError());
        hipLaunchKernelGGL(batch_moments_kernel, dim3(static_cast<unsigned>(na)), dim3(tbk::kBlock), 0, st,
                           static_cast<const double2*>(partial.p), d_tb.as<const int64_t>(), d_te.as<const int64_t>(),
                           d_off.as<const int64_t>(), d_n.as<const int32_t>(), k, mu.as<double>());
        HIPC(hipGetLastError());
    }

    void init_kets(hipStream_t st) {
        for (int32_t i = 0; i < np; ++i) {
            char* dst = static_cast<char*>(buf[0].p) + static_cast<size_t>(row_begin[i]) * Rp * elem();
            const int64_t total = rows[i] * Rp;
            const unsigned grid = static_cast<unsigned>(std::min<int64_t>((total + tbk::kBlock - 1) / tbk::kBlock, 65536));
            const double inv = 1.0 / std::sqrt(static_cast<double>(rows[i]));
            if (precision == TBK_FP64)
                hipLaunchKernelGGL(tbk::random_phase_kernel<double2>, dim3(grid), dim3(tbk::kBlock), 0, st,
                                   reinterpret_cast<double2*>(dst), static_cast<int32_t>(rows[i]), R, Rp, seed[i], inv);
            else
                hipLaunchKernelGGL(tbk::random_phase_kernel<float2>, dim3(grid), dim3(tbk::kBlock), 0, st,
                                   reinterpret_cast<float2*>(dst), static_cast<int32_t>(rows[i]), R, Rp, seed[i], inv);
            HIPC(hipGetLastError());
        }
    }

    // algorithmic bytes of the whole run (matrix arrays once per step; x, z
    // read and y written per recursion step, x read and y written at T_1)
    double bytes_total() const {
        const double sval = (precision == TBK_FP64 ? 8.0 : 4.0) * (cplx ? 2.0 : 1.0);
        double t = 0.0;
        for (int32_t i = 0; i < np; ++i) {
            const double mat = static_cast<double>(nnz[i]) * (sval + 4.0) + (static_cast<double>(rows[i]) + 1.0) * 4.0;
            const double blk = static_cast<double>(rows[i]) * R * static_cast<double>(elem());
            t += (mat + 2.0 * blk) + static_cast<double>(steps[i] - 1) * (mat + 3.0 * blk);
        }
        return t;
    }
};

namespace {

void batch_create(tbk_context* ctx, int32_t np, const tbk_sparse* mats, const double* a, const double* b,
                  const int32_t* n_moments, int32_t R, int32_t precision, const uint64_t* seeds, tbk_batch** out) {
    if (!ctx || !mats || !a || !b || !n_moments || !seeds || !out) raise(TBK_E_INVALID, "null argument");
    if (np < 1) raise(TBK_E_INVALID, "a batch needs at least one problem");
    if (R < 1) raise(TBK_E_INVALID, "R must be positive");
    if (precision != TBK_FP64 && precision != TBK_FP32) raise(TBK_E_INVALID, "unknown precision");
    const int Rp = pow2_at_least(R);
    if (Rp > 256) raise(TBK_E_UNSUPPORTED, "R above 256 is not supported");
    const bool cplx = mats[0].is_complex != 0;
    for (int32_t i = 0; i < np; ++i) {
        const tbk_sparse& M = mats[i];
        if (M.rows <= 0 || M.rows != M.cols) raise(TBK_E_INVALID, "every matrix must be square and non-empty");
        if ((M.is_complex != 0) != cplx) raise(TBK_E_INVALID, "a batch holds only real or only complex matrices");
        if (n_moments[i] < 2) raise(TBK_E_INVALID, "n_moments must be at least 2");
        if (!(a[i] != 0.0) || !std::isfinite(a[i]) || !std::isfinite(b[i]))
            raise(TBK_E_INVALID, "invalid rescaling a, b");
    }
    set_device(ctx);
    auto B = std::make_unique<tbk_batch>();
    B->ctx = ctx;
    B->np = np;
    B->R = R;
    B->Rp = Rp;
    B->precision = precision;
    B->cplx = cplx;

    std::vector<int32_t> order(np);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        return (n_moments[x] + 1) / 2 > (n_moments[y] + 1) / 2;
    });
    // moment offsets in problem order
    std::vector<int64_t> off(np);
    int64_t acc = 0;
    for (int32_t i = 0; i < np; ++i) {
        off[i] = acc;
        acc += n_moments[i];
    }
    B->mu_total = acc;

    // concatenated, tile-padded CSR
    std::vector<int32_t> rp{0}, ci;
    std::vector<double> val;
    std::vector<int32_t> tile_prob;
    std::vector<double> coef;
    const int vw = cplx ? 2 : 1;
    int64_t row = 0, tile = 0;
    for (int32_t pos = 0; pos < np; ++pos) {
        const int32_t p = order[pos];
        std::vector<int32_t> prp, pci;
        std::vector<double> pval;
        build_csr(&mats[p], prp, pci, pval);
        const int64_t n = mats[p].rows;
        const int64_t tiles = (n + kBatchTileRows - 1) / kBatchTileRows;
        if (row + tiles * kBatchTileRows >= (int64_t(1) << 31) ||
            static_cast<int64_t>(ci.size()) + static_cast<int64_t>(pci.size()) >= (int64_t(1) << 31))
            raise(TBK_E_UNSUPPORTED, "batch exceeds int32 row or entry indexing");
        const int32_t base = static_cast<int32_t>(ci.size());
        for (int64_t i = 1; i <= n; ++i) rp.push_back(base + prp[i]);
        for (int64_t i = n; i < tiles * kBatchTileRows; ++i) rp.push_back(rp.back());
        for (int32_t c : pci) ci.push_back(c + static_cast<int32_t>(row));
        val.insert(val.end(), pval.begin(), pval.end());
        for (int64_t t = 0; t < tiles; ++t) tile_prob.push_back(pos);
        coef.push_back(1.0 / a[p]);
        coef.push_back(-b[p] / a[p]);
        coef.push_back(2.0 / a[p]);
        coef.push_back(-2.0 * b[p] / a[p]);
        B->prob.push_back(p);
        B->row_begin.push_back(row);
        B->rows.push_back(n);
        B->tile_begin.push_back(tile);
        B->tile_end.push_back(tile + tiles);
        B->steps.push_back((n_moments[p] + 1) / 2);
        B->nnz.push_back(static_cast<int64_t>(pci.size()));
        B->seed.push_back(seeds[p]);
        B->nmom.push_back(n_moments[p]);
        B->mu_off.push_back(off[p]);
        row += tiles * kBatchTileRows;
        tile += tiles;
    }
    B->rows_total = row;
    B->ntiles = tile;
    B->max_steps = B->steps[0];

    // device arrays
    auto up = [&](DevBuf& d, const void* src, size_t bytes) {
        d.reserve(std::max<size_t>(bytes, 16));
        if (bytes) HIPC(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    };
    up(B->rp, rp.data(), rp.size() * sizeof(int32_t));
    up(B->ci, ci.data(), ci.size() * sizeof(int32_t));
    std::vector<float> val32;
    if (precision == TBK_FP64) {
        up(B->val, val.data(), val.size() * sizeof(double));
    } else {
        val32.resize(val.size());
        for (size_t i = 0; i < val.size(); ++i) val32[i] = static_cast<float>(val[i]);
        up(B->val, val32.data(), val32.size() * sizeof(float));
    }
    (void)vw;
    up(B->tile_prob, tile_prob.data(), tile_prob.size() * sizeof(int32_t));
    up(B->coef, coef.data(), coef.size() * sizeof(double));
    up(B->d_tb, B->tile_begin.data(), B->tile_begin.size() * sizeof(int64_t));
    up(B->d_te, B->tile_end.data(), B->tile_end.size() * sizeof(int64_t));
    up(B->d_off, B->mu_off.data(), B->mu_off.size() * sizeof(int64_t));
    up(B->d_n, B->nmom.data(), B->nmom.size() * sizeof(int32_t));
    B->mu.reserve(static_cast<size_t>(B->mu_total) * sizeof(double));
    B->partial.reserve(static_cast<size_t>(4 * B->ntiles) * sizeof(double2));
    const size_t blk = static_cast<size_t>(B->rows_total) * Rp * B->elem();
    for (auto& d : B->buf) {
        d.reserve(blk);
        HIPC(hipMemsetAsync(d.p, 0, blk, ctx->stream));  // padding rows and columns stay zero
    }
    HIPC(hipStreamSynchronize(ctx->stream));
    int64_t held = static_cast<int64_t>(B->rp.bytes + B->ci.bytes + B->val.bytes + 3 * blk + B->partial.bytes +
                                        B->mu.bytes);
    ctx->held_bytes += held;
    *out = B.release();
}

void batch_run(tbk_batch* B, int32_t flags, double* mu_out, tbk_batch_stats* stats) {
    auto t0 = std::chrono::steady_clock::now();
    if (!B || !mu_out) raise(TBK_E_INVALID, "null argument");
    tbk_context* ctx = B->ctx;
    set_device(ctx);
    hipStream_t st = ctx->stream;
    const bool use_graphs = (flags & TBK_BATCH_GRAPHS) != 0;
    ensure_events(ctx, 2);
    B->init_kets(st);
    HIPC(hipEventRecord(ctx->events[0], st));
    const int64_t nchunks = (B->max_steps + kChunkSteps - 1) / kChunkSteps;
    int64_t graphs_launched = 0;
    if (use_graphs && static_cast<int64_t>(B->graphs.size()) != nchunks) {
        B->drop_graphs();
        B->graphs.assign(static_cast<size_t>(nchunks), nullptr);
    }
    for (int64_t c = 0; c < nchunks; ++c) {
        const int64_t k0 = c * kChunkSteps, k1 = std::min(B->max_steps, k0 + kChunkSteps);
        if (!use_graphs) {
            for (int64_t k = k0; k < k1; ++k) B->launch_step(k, st);
            continue;
        }
        hipGraphExec_t& ex = B->graphs[static_cast<size_t>(c)];
        if (!ex) {
            hipGraph_t g = nullptr;
            HIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                for (int64_t k = k0; k < k1; ++k) B->launch_step(k, st);
            } catch (...) {
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPC(hipStreamEndCapture(st, &g));
            const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPC(e);
        }
        HIPC(hipGraphLaunch(ex, st));
        ++graphs_launched;
    }
    HIPC(hipEventRecord(ctx->events[1], st));
    std::vector<double> mu(static_cast<size_t>(B->mu_total));
    HIPC(hipMemcpyAsync(mu.data(), B->mu.p, mu.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    std::memcpy(mu_out, mu.data(), mu.size() * sizeof(double));
    if (stats) {
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
        stats->kernel_ms = ms;
        stats->launches = B->max_steps;
        stats->graphs = graphs_launched;
        stats->bytes_total = B->bytes_total();
        stats->rows_total = B->rows_total;
        stats->tile_rows = kBatchTileRows;
        stats->cpl = B->cpl;
        stats->blocks_per_cu = B->bpc;
        stats->xcd_split = B->xcd;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
}

bool shape_valid(int Rp, int cpl) {
    if (cpl != 1 && cpl != 2 && cpl != 4) return false;
    if (Rp % cpl != 0) return false;
    const int lpr = Rp / cpl;
    if (lpr > tbk::kWave) return false;
    return !((cpl == 4 && Rp < 8) || (cpl == 1 && Rp > 64) || (cpl == 2 && Rp > 128));
}

// Time `steps` recursion steps over all tiles for each launch shape
// (complex values per lane, workgroups per CU, XCD split); keep the best.
void batch_autotune(tbk_batch* B, int32_t steps, double* best_ms) {
    if (!B) raise(TBK_E_INVALID, "null argument");
    if (steps < 1) raise(TBK_E_INVALID, "steps must be positive");
    tbk_context* ctx = B->ctx;
    set_device(ctx);
    hipStream_t st = ctx->stream;
    ensure_events(ctx, 2);
    B->init_kets(st);
    B->launch_step(0, st);  // T_1, so that every buffer holds finite values
    const int32_t keep_cpl = B->cpl, keep_bpc = B->bpc, keep_xcd = B->xcd;
    double best = -1.0;
    int32_t bc = keep_cpl, bb = keep_bpc, bx = keep_xcd;
    // steps k = 1, 2, 3, 1, 2, ... continue the recursion in the three
    // rotating blocks (step k reads blocks k%3, (k+2)%3 and writes (k+1)%3)
    // with every problem active; the moments they write are overwritten by
    // the next run
    int64_t cycle = 0;
    auto next_k = [&] { return 1 + (cycle++ % 3); };
    for (int cpl : {1, 2, 4}) {
        if (!shape_valid(B->Rp, cpl)) continue;
        for (int bpc : {2, 4, 8}) {
            for (int xcd : {1, 0}) {
                B->cpl = cpl;
                B->bpc = bpc;
                B->xcd = xcd;
                B->launch_step(next_k(), st);  // warm
                HIPC(hipEventRecord(ctx->events[0], st));
                for (int32_t s = 0; s < steps; ++s) B->launch_step(next_k(), st);
                HIPC(hipEventRecord(ctx->events[1], st));
                HIPC(hipEventSynchronize(ctx->events[1]));
                float ms = 0.f;
                HIPC(hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
                const double per = ms / steps;
                if (best < 0.0 || per < best) {
                    best = per;
                    bc = cpl;
                    bb = bpc;
                    bx = xcd;
                }
            }
        }
    }
    B->cpl = bc;
    B->bpc = bb;
    B->xcd = bx;
    if (bc != keep_cpl || bb != keep_bpc || bx != keep_xcd) B->drop_graphs();
    if (best_ms) *best_ms = best;
}

}  // namespace

extern "C" {

int tbk_batch_create(tbk_context* ctx, int32_t n_problems, const tbk_sparse* mats, const double* a, const double* b,
                     const int32_t* n_moments, int32_t R, int32_t precision, const uint64_t* seeds, tbk_batch** out) {
    return guarded([&] { batch_create(ctx, n_problems, mats, a, b, n_moments, R, precision, seeds, out); });
}

int tbk_batch_run(tbk_batch* batch, int32_t flags, double* mu_out, tbk_batch_stats* stats) {
    return guarded([&] { batch_run(batch, flags, mu_out, stats); });
}

int tbk_batch_set_shape(tbk_batch* batch, int32_t cpl, int32_t blocks_per_cu, int32_t xcd_split) {
    return guarded([&] {
        if (!batch) raise(TBK_E_INVALID, "null argument");
        if (!shape_valid(batch->Rp, cpl)) raise(TBK_E_UNSUPPORTED, "no batched kernel for this complex-per-lane count");
        if (blocks_per_cu < 1 || blocks_per_cu > 32) raise(TBK_E_INVALID, "blocks_per_cu must be in 1..32");
        batch->cpl = cpl;
        batch->bpc = blocks_per_cu;
        batch->xcd = xcd_split ? 1 : 0;
        batch->drop_graphs();
    });
}

int tbk_batch_autotune(tbk_batch* batch, int32_t steps, double* best_ms_per_step) {
    return guarded([&] { batch_autotune(batch, steps, best_ms_per_step); });
}

int tbk_batch_destroy(tbk_batch* batch) {
    return guarded([&] {
        if (!batch) return;
        tbk_context* ctx = batch->ctx;
        set_device(ctx);
        HIPC(hipStreamSynchronize(ctx->stream));
        const size_t blk = static_cast<size_t>(batch->rows_total) * batch->Rp * batch->elem();
        ctx->held_bytes -= static_cast<int64_t>(batch->rp.bytes + batch->ci.bytes + batch->val.bytes + 3 * blk +
                                                batch->partial.bytes + batch->mu.bytes);
        delete batch;
    });
}

}  // extern "C"
