// Batched ragged moment problems: many independent Hamiltonians of different
// sizes and moment counts run together, one launch of cheb_slots_kernel per
// Chebyshev step for all problems still active.
//
// Layout.  The problems' CSR arrays (local column indices) and their vector
// blocks (row-major [row][Rp]) are concatenated; each problem is a slot
// whose descriptor points at its own pieces.  Problems are laid out in order
// of decreasing step count, so the problems still active at step k are a
// prefix, and so are their workgroups: a step launch covers exactly the
// workgroups of that prefix.  Each problem owns a fixed set of workgroups
// that splits its rows as the single-problem step kernel does (tiles dealt
// per XCD in the matrix's chunk order), and each workgroup writes one
// partial; slots_moments_kernel adds a problem's partials in a fixed order
// and applies the doubling identities.  A problem's moments therefore do
// not depend on the other problems in the batch, and its vectors T_n are
// those of a single-problem run.
//
// Graphs.  With TBK_BATCH_GRAPHS the launches of each chunk of 64 steps are
// captured once into a hipGraph (no memset nodes) and replayed on later runs.
//
// Persistent launches.  With TBK_BATCH_PERSISTENT each chunk of 64 steps is
// one launch of cheb_slots_persistent_kernel (plus a reset and a check
// kernel, captured together with TBK_BATCH_GRAPHS): every problem is placed
// on one XCD's queue, its rows cut into shares, and the workgroups on that
// XCD work through the queue's (step, problem, share) items in step order
// (cheb_kernels.h).  The item partials of every step are kept and
// slots_history_moments_kernel sums them per problem in share order at the
// end, so a problem's moments again do not depend on the batch.
//
// Synchronised launches.  With TBK_BATCH_SYNCED each chunk of 64 steps is
// one launch of cheb_slots_sync_kernel over the whole GPU: its resident
// workgroups do the per-step launches' workgroup indices step after step,
// one moment partial per wave, a grid barrier between steps, and form each
// step's moments after its barrier (cheb_kernels.h).
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

constexpr int64_t kChunkSteps = 64;
// persistent launches: XCD queues, rows per share, polls before a wait gives up
constexpr int kPersistQueues = 8;
constexpr int32_t kPersistShareRows = 4096;
constexpr int32_t kPersistSpinLimit = 1 << 22;

}  // namespace

struct tbk_batch {
    tbk_context* ctx = nullptr;
    int32_t np = 0, R = 0, Rp = 0, precision = TBK_FP64;
    bool cplx = false;
    int64_t rows_total = 0, max_steps = 0;
    // per layout position (problems sorted by decreasing steps)
    std::vector<int32_t> prob;  // problem index of each position
    std::vector<int64_t> row_begin, rows, rp_begin, nz_begin, nnz, steps, bandwidth, mu_off;
    std::vector<uint64_t> seed;
    std::vector<int32_t> nmom;
    std::vector<double> coef;     // per position: (1/a, -b/a, 2/a, -2b/a)
    std::vector<int64_t> wg_end;  // cumulative workgroups of positions 0..i
    int64_t mu_total = 0;
    DevBuf rp, ci, val, mu, partial, desc, wg_slot, buf[3];
    // launch shape: complex values per lane (0: 16-byte lanes), workgroups per CU
    int32_t cpl = 0, bpc = 4;
    std::vector<hipGraphExec_t> graphs;  // per chunk, captured on first use
    size_t elem() const { return precision == TBK_FP64 ? 16 : 8; }
    int32_t cpl_used() const { return slot_cpl(cpl, precision, Rp); }

    // Persistent launches (TBK_BATCH_PERSISTENT): every problem on one XCD
    // queue, items (step, problem, share) in step order; built on first use.
    struct Persist {
        bool built = false;
        int32_t nchunks = 0;
        int64_t grid = 0, chunk = 64;
        std::vector<int64_t> step_base;  // per chunk: offset of its per-queue step prefixes
        DevBuf nshare, pbase, q_slots, q_off, q_share_pre, q_step_items, totals, counters, done, error, part;
        std::vector<hipGraphExec_t> graphs;
    } ps;

    // Grid-synchronised persistent launches (TBK_BATCH_SYNCED): per step,
    // the workgroup count of its per-step launch and its active slots; the
    // partials of two steps; the barrier counter and the error flag.
    struct Synced {
        bool built = false;
        int64_t wmax = 0, grid = 0;
        DevBuf step_wg, step_na, part, bar, error;
        std::vector<hipGraphExec_t> graphs;
    } sy;

    void drop_graphs() {
        for (hipGraphExec_t g : graphs)
            if (g) (void)hipGraphExecDestroy(g);
        graphs.clear();
        for (hipGraphExec_t g : ps.graphs)
            if (g) (void)hipGraphExecDestroy(g);
        ps.graphs.clear();
        for (hipGraphExec_t g : sy.graphs)
            if (g) (void)hipGraphExecDestroy(g);
        sy.graphs.clear();
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

    // descriptors and workgroup map for the current launch shape
    void build_descriptors(hipStream_t st) {
        const int lpr = Rp / cpl_used();
        std::vector<tbk::SlotDesc> d(static_cast<size_t>(np));
        std::vector<int32_t> map;
        wg_end.assign(static_cast<size_t>(np), 0);
        const size_t hv_elem = (precision == TBK_FP64 ? 8 : 4) * (cplx ? 2 : 1);
        int64_t w0 = 0;
        for (int32_t i = 0; i < np; ++i) {
            tbk::SlotDesc& s = d[i];
            std::memset(&s, 0, sizeof(s));
            s.rp = rp.as<const int32_t>() + rp_begin[i];
            s.ci = ci.as<const int32_t>() + nz_begin[i];
            s.hv = static_cast<const char*>(val.p) + static_cast<size_t>(nz_begin[i]) * hv_elem;
            for (int q = 0; q < 3; ++q)
                s.blk[q] = static_cast<char*>(buf[q].p) + static_cast<size_t>(row_begin[i]) * Rp * elem();
            for (int c = 0; c < 4; ++c) s.coef[c] = coef[4 * static_cast<size_t>(i) + c];
            s.mu = mu.as<double>() + mu_off[i];
            s.k0 = 0;
            s.steps = steps[i];
            s.rows = static_cast<int32_t>(rows[i]);
            s.n = nmom[i];
            slot_split(ctx->num_cu, bpc, lpr, rows[i], bandwidth[i], s.order, s.nwg);
            s.wg0 = w0;
            map.insert(map.end(), static_cast<size_t>(s.nwg), i);
            w0 += s.nwg;
            wg_end[i] = w0;
        }
        desc.reserve(d.size() * sizeof(tbk::SlotDesc));
        wg_slot.reserve(map.size() * sizeof(int32_t));
        partial.reserve(static_cast<size_t>(w0) * sizeof(double2));
        HIPC(hipMemcpyAsync(desc.p, d.data(), d.size() * sizeof(tbk::SlotDesc), hipMemcpyHostToDevice, st));
        HIPC(hipMemcpyAsync(wg_slot.p, map.data(), map.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIPC(hipStreamSynchronize(st));
    }

    void launch_step(int64_t k, hipStream_t st) {
        const int32_t na = active(k);
        if (na == 0) return;
        tbk::SlotArgs A{};
        A.slots = desc.as<const tbk::SlotDesc>();
        A.wg_slot = wg_slot.as<const int32_t>();
        A.s = static_cast<int32_t>(k);
        A.Rp = Rp;
        A.opts = ctx->tuning.stream_nt >= 0 ? tbk::kOptStream : 0;
        A.partial = partial.as<double2>();
        const dim3 grid(static_cast<unsigned>(wg_end[na - 1]));
        const int c = cpl_used();
        bool ok;
        if (precision == TBK_FP64)
            ok = cplx ? tbk::launch_slots_f64_cplx(Rp, c, grid, st, A) : tbk::launch_slots_f64_real(Rp, c, grid, st, A);
        else
            ok = cplx ? tbk::launch_slots_f32_cplx(Rp, c, grid, st, A) : tbk::launch_slots_f32_real(Rp, c, grid, st, A);
        if (!ok) raise(TBK_E_UNSUPPORTED, "no batched step kernel for this width");
        HIPC(hipGetLastError());
        hipLaunchKernelGGL(tbk::slots_moments_kernel, dim3(static_cast<unsigned>(na)), dim3(tbk::kBlock), 0, st,
                           desc.as<const tbk::SlotDesc>(), static_cast<int32_t>(k), partial.as<const double2>());
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
                                   reinterpret_cast<double2*>(dst), static_cast<int32_t>(rows[i]), R, Rp, seed[i], inv, R, 0);
            else
                hipLaunchKernelGGL(tbk::random_phase_kernel<float2>, dim3(grid), dim3(tbk::kBlock), 0, st,
                                   reinterpret_cast<float2*>(dst), static_cast<int32_t>(rows[i]), R, Rp, seed[i], inv, R, 0);
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

    int64_t device_bytes() const {
        return static_cast<int64_t>(rp.bytes + ci.bytes + val.bytes + mu.bytes + partial.bytes + desc.bytes +
                                    wg_slot.bytes + buf[0].bytes + buf[1].bytes + buf[2].bytes + ps.part.bytes +
                                    ps.q_step_items.bytes + sy.step_wg.bytes + sy.step_na.bytes + sy.part.bytes);
    }

    // The persistent plan: problems placed on the 8 XCD queues (below), each
    // queue's problems in layout order (decreasing steps, so the problems
    // active at a step are a prefix); ceil(rows / kPersistShareRows) shares
    // each; per chunk of kChunkSteps steps and queue, the items before each
    // step.
    void build_persist(hipStream_t st) {
        constexpr int nq = kPersistQueues;
        std::vector<int32_t> nsh(static_cast<size_t>(np));
        std::vector<int64_t> pb(static_cast<size_t>(np));
        int64_t acc = 0;
        for (int32_t i = 0; i < np; ++i) {
            nsh[i] = static_cast<int32_t>((rows[i] + kPersistShareRows - 1) / kPersistShareRows);
            pb[i] = acc;
            acc += steps[i] * nsh[i];
        }
        // placement: problems in decreasing step count (larger rows first
        // among equal counts), each to the queue holding the fewest rows so
        // far.  The problems active at a step are the ones placed before
        // some point of this order, so every step's rows are balanced over
        // the XCDs, not only the total work.
        std::vector<int32_t> byw(static_cast<size_t>(np));
        std::iota(byw.begin(), byw.end(), 0);
        std::stable_sort(byw.begin(), byw.end(), [&](int32_t x, int32_t y) {
            return steps[x] != steps[y] ? steps[x] > steps[y] : rows[x] > rows[y];
        });
        std::vector<std::vector<int32_t>> xq(nq);
        std::vector<double> load(nq, 0.0);
        for (int32_t p : byw) {
            const int q = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
            xq[q].push_back(p);
            load[q] += static_cast<double>(rows[p]);
        }
        std::vector<int32_t> qs, qo(nq + 1, 0);
        std::vector<int64_t> pre;
        for (int q = 0; q < nq; ++q) {
            std::sort(xq[q].begin(), xq[q].end());
            qo[q] = static_cast<int32_t>(qs.size());
            int64_t a = 0;
            pre.push_back(0);
            for (int32_t p : xq[q]) {
                qs.push_back(p);
                a += nsh[p];
                pre.push_back(a);
            }
        }
        qo[nq] = static_cast<int32_t>(qs.size());
        const int64_t pch = kChunkSteps;
        ps.chunk = pch;
        const int32_t nch = static_cast<int32_t>((max_steps + pch - 1) / pch);
        std::vector<int64_t> items, tot;
        ps.step_base.assign(static_cast<size_t>(nch), 0);
        for (int32_t c = 0; c < nch; ++c) {
            const int64_t k0 = static_cast<int64_t>(c) * pch, k1 = std::min(max_steps, k0 + pch);
            ps.step_base[c] = static_cast<int64_t>(items.size());
            for (int q = 0; q < nq; ++q) {
                const int32_t n_q = qo[q + 1] - qo[q];
                int64_t a = 0;
                for (int64_t k = k0; k <= k1; ++k) {
                    items.push_back(a);
                    if (k == k1) break;
                    int32_t act = 0;  // problems of the queue still active at step k (a prefix)
                    while (act < n_q && steps[qs[qo[q] + act]] > k) ++act;
                    a += pre[static_cast<size_t>(qo[q] + q + act)];
                }
                tot.push_back(a);
            }
        }
        auto up = [&](DevBuf& d, const void* src, size_t bytes) {
            d.reserve(std::max<size_t>(bytes, 16));
            if (bytes) HIPC(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, st));
        };
        up(ps.nshare, nsh.data(), nsh.size() * sizeof(int32_t));
        up(ps.pbase, pb.data(), pb.size() * sizeof(int64_t));
        up(ps.q_slots, qs.data(), qs.size() * sizeof(int32_t));
        up(ps.q_off, qo.data(), qo.size() * sizeof(int32_t));
        up(ps.q_share_pre, pre.data(), pre.size() * sizeof(int64_t));
        up(ps.q_step_items, items.data(), items.size() * sizeof(int64_t));
        up(ps.totals, tot.data(), tot.size() * sizeof(int64_t));
        ps.counters.reserve(nq * sizeof(int32_t));
        ps.done.reserve(static_cast<size_t>(np) * sizeof(int32_t));
        ps.error.reserve(sizeof(int32_t));
        ps.part.reserve(std::max<size_t>(static_cast<size_t>(acc) * sizeof(double2), 16));
        HIPC(hipStreamSynchronize(st));
        ps.nchunks = nch;
        ps.grid = static_cast<int64_t>(nq) * std::max(1, ctx->num_cu / nq) * std::max(1, bpc);
        ps.built = true;
    }

    // chunk c: resets, the persistent launch, the check that every queue was served
    void launch_persist_chunk(int32_t c, hipStream_t st) {
        constexpr int nq = kPersistQueues;
        const int64_t k0 = static_cast<int64_t>(c) * ps.chunk, k1 = std::min(max_steps, k0 + ps.chunk);
        hipLaunchKernelGGL(tbk::persist_reset_kernel, dim3(static_cast<unsigned>((np + tbk::kBlock - 1) / tbk::kBlock)),
                           dim3(tbk::kBlock), 0, st, ps.done.as<int32_t>(), np, ps.counters.as<int32_t>(), nq,
                           c == 0 ? ps.error.as<int32_t>() : nullptr);
        HIPC(hipGetLastError());
        tbk::PersistSched S{};
        S.slots = desc.as<const tbk::SlotDesc>();
        S.nshare = ps.nshare.as<const int32_t>();
        S.pbase = ps.pbase.as<const int64_t>();
        S.q_slots = ps.q_slots.as<const int32_t>();
        S.q_off = ps.q_off.as<const int32_t>();
        S.q_share_pre = ps.q_share_pre.as<const int64_t>();
        S.q_step_items = ps.q_step_items.as<const int64_t>() + ps.step_base[c];
        S.counters = ps.counters.as<int32_t>();
        S.done = ps.done.as<int32_t>();
        S.error = ps.error.as<int32_t>();
        S.part = ps.part.as<double2>();
        S.k0 = static_cast<int32_t>(k0);
        S.k1 = static_cast<int32_t>(k1);
        S.share_rows = kPersistShareRows;
        S.spin_limit = kPersistSpinLimit;
        S.Rp = Rp;
        // rows pass between workgroups through the XCD's L2: every access
        // temporal (a non-temporal store does not leave the row in the L2
        // the next step reads it from)
        S.opts = 0;
        for (int i = 0; i < 16; ++i) S.xcc_queue[i] = i < nq ? i : -1;
        const dim3 grid(static_cast<unsigned>(ps.grid));
        const int cp = cpl_used();
        bool ok;
        if (precision == TBK_FP64)
            ok = cplx ? tbk::launch_persist_f64_cplx(Rp, cp, grid, st, S) : tbk::launch_persist_f64_real(Rp, cp, grid, st, S);
        else
            ok = cplx ? tbk::launch_persist_f32_cplx(Rp, cp, grid, st, S) : tbk::launch_persist_f32_real(Rp, cp, grid, st, S);
        if (!ok) raise(TBK_E_UNSUPPORTED, "no persistent batched kernel for this width");
        HIPC(hipGetLastError());
        hipLaunchKernelGGL(tbk::persist_check_kernel, dim3(1), dim3(64), 0, st, ps.counters.as<const int32_t>(),
                           ps.totals.as<const int64_t>() + static_cast<int64_t>(c) * nq, nq, ps.error.as<int32_t>());
        HIPC(hipGetLastError());
    }

    // The synchronised plan: per step, the per-step launch's workgroups (the
    // active prefix) and active slots.
    void build_synced(hipStream_t st) {
        std::vector<int64_t> wg(static_cast<size_t>(max_steps));
        std::vector<int32_t> na(static_cast<size_t>(max_steps));
        for (int64_t k = 0; k < max_steps; ++k) {
            na[k] = active(k);
            wg[k] = na[k] > 0 ? wg_end[na[k] - 1] : 0;
        }
        sy.wmax = wg_end.back();
        sy.step_wg.reserve(wg.size() * sizeof(int64_t));
        sy.step_na.reserve(na.size() * sizeof(int32_t));
        sy.part.reserve(static_cast<size_t>(2 * tbk::kWavesPerBlock * sy.wmax) * sizeof(double2));
        sy.bar.reserve(tbk::kSyncWords * sizeof(int32_t));
        sy.error.reserve(sizeof(int32_t));
        HIPC(hipMemcpyAsync(sy.step_wg.p, wg.data(), wg.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
        HIPC(hipMemcpyAsync(sy.step_na.p, na.data(), na.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIPC(hipStreamSynchronize(st));
        sy.built = true;
    }

    // chunk c: zero the barrier counter (and, for the first chunk, the error
    // flag), then the synchronised launch of the chunk's steps
    void launch_synced_chunk(int64_t c, hipStream_t st) {
        if (!sy.built || sy.wmax != wg_end.back())
            raise(TBK_E_INVALID, "synchronised plan does not match the launch shape");
        const int64_t k0 = c * kChunkSteps, k1 = std::min(max_steps, k0 + kChunkSteps);
        hipLaunchKernelGGL(tbk::persist_reset_kernel, dim3(1), dim3(tbk::kBlock), 0, st, sy.bar.as<int32_t>(),
                           tbk::kSyncWords, nullptr, 0, c == 0 ? sy.error.as<int32_t>() : nullptr);
        HIPC(hipGetLastError());
        tbk::SyncSched S{};
        S.slots = desc.as<const tbk::SlotDesc>();
        S.wg_slot = wg_slot.as<const int32_t>();
        S.step_wg = sy.step_wg.as<const int64_t>() + k0;
        S.step_na = sy.step_na.as<const int32_t>() + k0;
        S.partial = sy.part.as<double2>();
        S.wmax = sy.wmax;
        S.bar = sy.bar.as<int32_t>();
        S.error = sy.error.as<int32_t>();
        S.s0 = static_cast<int32_t>(k0);
        S.nsteps = static_cast<int32_t>(k1 - k0);
        S.Rp = Rp;
        S.opts = ctx->tuning.stream_nt >= 0 ? tbk::kOptStream : 0;
        S.spin_limit = kPersistSpinLimit;
        const int cp = cpl_used();
        int64_t g;
        if (precision == TBK_FP64)
            g = cplx ? tbk::launch_sync_f64_cplx(Rp, cp, bpc, ctx->num_cu, st, S)
                     : tbk::launch_sync_f64_real(Rp, cp, bpc, ctx->num_cu, st, S);
        else
            g = cplx ? tbk::launch_sync_f32_cplx(Rp, cp, bpc, ctx->num_cu, st, S)
                     : tbk::launch_sync_f32_real(Rp, cp, bpc, ctx->num_cu, st, S);
        if (g == 0) raise(TBK_E_UNSUPPORTED, "no synchronised batched kernel for this width");
        HIPC(hipGetLastError());
        sy.grid = g;
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
    std::vector<int64_t> off(np);
    int64_t acc = 0;
    for (int32_t i = 0; i < np; ++i) {
        off[i] = acc;
        acc += n_moments[i];
    }
    B->mu_total = acc;

    // concatenated CSR with local indices
    std::vector<int32_t> rp, ci;
    std::vector<double> val;
    int64_t row = 0;
    for (int32_t pos = 0; pos < np; ++pos) {
        const int32_t p = order[pos];
        std::vector<int32_t> prp, pci;
        std::vector<double> pval;
        build_csr(&mats[p], prp, pci, pval);
        const int64_t n = mats[p].rows;
        B->prob.push_back(p);
        B->row_begin.push_back(row);
        B->rows.push_back(n);
        B->rp_begin.push_back(static_cast<int64_t>(rp.size()));
        B->nz_begin.push_back(static_cast<int64_t>(ci.size()));
        B->nnz.push_back(static_cast<int64_t>(pci.size()));
        B->bandwidth.push_back(bandwidth_percentile(prp, pci, 0.95));
        B->steps.push_back((n_moments[p] + 1) / 2);
        B->seed.push_back(seeds[p]);
        B->nmom.push_back(n_moments[p]);
        B->mu_off.push_back(off[p]);
        B->coef.push_back(1.0 / a[p]);
        B->coef.push_back(-b[p] / a[p]);
        B->coef.push_back(2.0 / a[p]);
        B->coef.push_back(-2.0 * b[p] / a[p]);
        rp.insert(rp.end(), prp.begin(), prp.end());
        ci.insert(ci.end(), pci.begin(), pci.end());
        val.insert(val.end(), pval.begin(), pval.end());
        row += n;
    }
    B->rows_total = row;
    B->max_steps = B->steps[0];

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
    B->mu.reserve(static_cast<size_t>(B->mu_total) * sizeof(double));
    const size_t blk = static_cast<size_t>(B->rows_total) * Rp * B->elem();
    for (auto& d : B->buf) d.reserve(blk);
    HIPC(hipStreamSynchronize(ctx->stream));
    B->build_descriptors(ctx->stream);
    ctx->held_bytes += B->device_bytes();
    *out = B.release();
}

void batch_run(tbk_batch* B, int32_t flags, double* mu_out, tbk_batch_stats* stats) {
    auto t0 = std::chrono::steady_clock::now();
    if (!B || !mu_out) raise(TBK_E_INVALID, "null argument");
    tbk_context* ctx = B->ctx;
    set_device(ctx);
    hipStream_t st = ctx->stream;
    const bool use_graphs = (flags & TBK_BATCH_GRAPHS) != 0;
    const bool persistent = (flags & TBK_BATCH_PERSISTENT) != 0;
    const bool synced = (flags & TBK_BATCH_SYNCED) != 0;
    if (flags & ~(TBK_BATCH_GRAPHS | TBK_BATCH_PERSISTENT | TBK_BATCH_SYNCED)) raise(TBK_E_INVALID, "unknown batch flag");
    if (persistent && synced) raise(TBK_E_INVALID, "TBK_BATCH_PERSISTENT and TBK_BATCH_SYNCED exclude each other");
    ensure_events(ctx, 2);
    if (persistent && !B->ps.built) {
        const int64_t before = B->device_bytes();
        B->build_persist(st);
        ctx->held_bytes += B->device_bytes() - before;
    }
    if (synced && !B->sy.built) {
        const int64_t before = B->device_bytes();
        B->build_synced(st);
        ctx->held_bytes += B->device_bytes() - before;
    }
    B->init_kets(st);
    HIPC(hipEventRecord(ctx->events[0], st));
    if (persistent) {
        int64_t graphs_launched = 0;
        if (use_graphs && static_cast<int32_t>(B->ps.graphs.size()) != B->ps.nchunks)
            B->ps.graphs.assign(static_cast<size_t>(B->ps.nchunks), nullptr);
        for (int32_t c = 0; c < B->ps.nchunks; ++c) {
            if (!use_graphs) {
                B->launch_persist_chunk(c, st);
                continue;
            }
            hipGraphExec_t& ex = B->ps.graphs[static_cast<size_t>(c)];
            if (!ex) {
                hipGraph_t g = nullptr;
                HIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                try {
                    B->launch_persist_chunk(c, st);
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
        hipLaunchKernelGGL(tbk::slots_history_moments_kernel, dim3(static_cast<unsigned>(B->np)), dim3(tbk::kBlock), 0,
                           st, B->desc.as<const tbk::SlotDesc>(), B->ps.nshare.as<const int32_t>(),
                           B->ps.pbase.as<const int64_t>(), B->ps.part.as<const double2>());
        HIPC(hipGetLastError());
        HIPC(hipEventRecord(ctx->events[1], st));
        std::vector<double> mu(static_cast<size_t>(B->mu_total));
        int32_t err = 0;
        HIPC(hipMemcpyAsync(mu.data(), B->mu.p, mu.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPC(hipMemcpyAsync(&err, B->ps.error.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        if (err != 0)
            raise(TBK_E_DEVICE, "persistent batch launch: a wait gave up or an XCD queue was not served");
        std::memcpy(mu_out, mu.data(), mu.size() * sizeof(double));
        if (stats) {
            float ms = 0.f;
            HIPC(hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
            stats->kernel_ms = ms;
            stats->launches = B->ps.nchunks;
            stats->graphs = graphs_launched;
            stats->bytes_total = B->bytes_total();
            stats->rows_total = B->rows_total;
            stats->workgroups = B->ps.grid;
            stats->cpl = B->cpl_used();
            stats->blocks_per_cu = B->bpc;
            stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        return;
    }
    const int64_t nchunks = (B->max_steps + kChunkSteps - 1) / kChunkSteps;
    int64_t graphs_launched = 0;
    std::vector<hipGraphExec_t>& chunk_graphs = synced ? B->sy.graphs : B->graphs;
    if (use_graphs && static_cast<int64_t>(chunk_graphs.size()) != nchunks) {
        for (hipGraphExec_t g : chunk_graphs)
            if (g) (void)hipGraphExecDestroy(g);
        chunk_graphs.assign(static_cast<size_t>(nchunks), nullptr);
    }
    // the launches of chunk c: one per step, or one synchronised launch
    auto launch_chunk = [&](int64_t c) {
        if (synced) {
            B->launch_synced_chunk(c, st);
            return;
        }
        const int64_t k0 = c * kChunkSteps, k1 = std::min(B->max_steps, k0 + kChunkSteps);
        for (int64_t k = k0; k < k1; ++k) B->launch_step(k, st);
    };
    for (int64_t c = 0; c < nchunks; ++c) {
        if (!use_graphs) {
            launch_chunk(c);
            continue;
        }
        hipGraphExec_t& ex = chunk_graphs[static_cast<size_t>(c)];
        if (!ex) {
            hipGraph_t g = nullptr;
            HIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                launch_chunk(c);
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
    int32_t err = 0;
    HIPC(hipMemcpyAsync(mu.data(), B->mu.p, mu.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    if (synced) HIPC(hipMemcpyAsync(&err, B->sy.error.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (err != 0) raise(TBK_E_DEVICE, "synchronised batch launch: a grid barrier gave up");
    std::memcpy(mu_out, mu.data(), mu.size() * sizeof(double));
    if (stats) {
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
        stats->kernel_ms = ms;
        stats->launches = synced ? nchunks : B->max_steps;
        stats->graphs = graphs_launched;
        stats->bytes_total = B->bytes_total();
        stats->rows_total = B->rows_total;
        stats->workgroups = synced ? B->sy.grid : B->wg_end.back();
        stats->cpl = B->cpl_used();
        stats->blocks_per_cu = B->bpc;
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

void set_shape(tbk_batch* B, int32_t cpl, int32_t bpc) {
    const int64_t before = B->device_bytes();
    B->cpl = cpl;
    B->bpc = bpc;
    B->drop_graphs();
    B->ps.grid = static_cast<int64_t>(kPersistQueues) * std::max(1, B->ctx->num_cu / kPersistQueues) * std::max(1, bpc);
    B->build_descriptors(B->ctx->stream);
    B->sy.built = false;  // its per-step workgroup counts and partials follow the descriptors
    B->ctx->held_bytes += B->device_bytes() - before;
}

// Time `steps` recursion steps with every problem active for each launch
// shape (complex values per lane, workgroups per CU); keep the best.
void batch_autotune(tbk_batch* B, int32_t steps, double* best_ms) {
    if (!B) raise(TBK_E_INVALID, "null argument");
    if (steps < 1) raise(TBK_E_INVALID, "steps must be positive");
    tbk_context* ctx = B->ctx;
    set_device(ctx);
    hipStream_t st = ctx->stream;
    ensure_events(ctx, 2);
    B->init_kets(st);
    B->launch_step(0, st);  // T_1, so that every buffer holds finite values
    double best = -1.0;
    int32_t bc = B->cpl_used(), bb = B->bpc;
    // steps k = 1, 2, 3, 1, 2, ... continue the recursion in the three
    // rotating blocks (step k reads blocks k%3, (k+2)%3 and writes (k+1)%3)
    // with every problem active; the moments they write are overwritten by
    // the next run
    int64_t cycle = 0;
    auto next_k = [&] { return 1 + (cycle++ % 3); };
    for (int cpl : {1, 2, 4}) {
        if (!shape_valid(B->Rp, cpl)) continue;
        for (int bpc : {2, 4, 8}) {
            set_shape(B, cpl, bpc);
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
            }
        }
    }
    set_shape(B, bc, bb);
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

int tbk_batch_set_shape(tbk_batch* batch, int32_t cpl, int32_t blocks_per_cu) {
    return guarded([&] {
        if (!batch) raise(TBK_E_INVALID, "null argument");
        if (!shape_valid(batch->Rp, cpl)) raise(TBK_E_UNSUPPORTED, "no batched kernel for this complex-per-lane count");
        if (blocks_per_cu < 1 || blocks_per_cu > 32) raise(TBK_E_INVALID, "blocks_per_cu must be in 1..32");
        set_device(batch->ctx);
        set_shape(batch, cpl, blocks_per_cu);
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
        ctx->held_bytes -= batch->device_bytes();
        delete batch;
    });
}

}  // extern "C"
