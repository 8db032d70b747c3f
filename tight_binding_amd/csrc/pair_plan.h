// Host-side schedule of the two-step launch (cheb_pair_kernel).
//
// One launch advances the recursion by two steps on blocks P (T_{k-1}) and
// Q (T_k), in place:
//   front 1, tile i:  P[i] <- 2 H~ Q - P[i]     (T_{k+1}; gathers Q)
//   front 2, tile j:  Q[j] <- 2 H~ P - Q[j]     (T_{k+2}; gathers P)
// Rows are cut into tiles of tt consecutive rows.  Front 2 of tile j may
// start once front 1 is done on
//   - every tile holding a column of tile j's rows (front 2 gathers T_{k+1}
//     there; its own tile always, since the moment partials read T_{k+1} of
//     the tile's own rows), and
//   - every tile whose rows have a column in tile j (they gather T_k of
//     tile j, which front 2 of j overwrites).
// deps(j) is that set of tiles, stored as ranges of consecutive tiles: a
// banded matrix gives one range, a periodic lattice one more per wrapped
// direction.
//
// The tiles are split into nq contiguous slabs, one per queue (one XCD each
// on the device).  A front-1 tile that a front-2 tile of another slab waits
// on is a band tile: a prologue launch computes it, so that no wait crosses a
// slab.  Queue q then lists the non-band front-1 tiles and all front-2 tiles
// of slab q, each front-2 tile after every front-1 tile of its slab it waits
// on.  A workgroup therefore only ever waits on items of its own queue with
// lower ticket numbers (or on band tiles, done before the launch), and
// front-1 items never wait, so every wait ends.
//
// This header has no device code: it is compiled into the library and into
// the host-side schedule test.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <vector>

namespace tbk {

struct PairSchedule {
    int32_t tt = 0;  // rows per tile
    int32_t ntiles = 0;
    int32_t nq = 0;
    bool ok = false;                 // false: the matrix does not suit two-step launches
    std::vector<int32_t> items;      // queue items: t >= 0 front-1 tile t, ~t (< 0) front-2 tile t
    std::vector<int32_t> qoff;       // queue q's items at [qoff[q], qoff[q + 1])
    std::vector<int32_t> dep_ptr;    // tile j's ranges at [dep_ptr[j], dep_ptr[j + 1])
    std::vector<int32_t> dep_rng;    // 2 per range: first and last tile
    std::vector<int32_t> band;       // band tiles, ascending
    std::vector<int32_t> done_init;  // per tile: 1 for band tiles (done by the prologue), else 0
};

// rp, ci: CSR row pointers and column indices of the rows x rows matrix.
// slack: front-1 tiles listed ahead of the last one a front-2 item needs.
// max_band_frac: the largest share of tiles the prologue may take.
// max_deps: the most dependency tiles per tile, on average, before the
// matrix is judged too scattered for this schedule.
inline PairSchedule build_pair_schedule(const std::vector<int32_t>& rp, const std::vector<int32_t>& ci,
                                        int32_t rows, int32_t tt, int32_t nq, int32_t slack, double max_band_frac,
                                        int32_t max_deps = 64) {
    PairSchedule s;
    s.tt = tt;
    s.nq = nq;
    if (tt <= 0 || nq <= 0 || rows <= 0 || static_cast<int64_t>(rp.size()) != static_cast<int64_t>(rows) + 1)
        return s;
    const int32_t nt = static_cast<int32_t>((static_cast<int64_t>(rows) + tt - 1) / tt);
    s.ntiles = nt;
    if (nt < nq) return s;

    // D(t): tiles holding a column of tile t's rows
    std::vector<std::vector<int32_t>> D(nt);
    int64_t total = 0;
    for (int32_t t = 0; t < nt; ++t) {
        std::vector<int32_t>& d = D[t];
        const int32_t r0 = t * tt, r1 = std::min(rows, r0 + tt);
        for (int32_t p = rp[r0]; p < rp[r1]; ++p) {
            const int32_t c = ci[p];
            if (c < 0 || c >= rows) return s;
            const int32_t ct = c / tt;
            if (d.empty() || d.back() != ct) d.push_back(ct);
        }
        std::sort(d.begin(), d.end());
        d.erase(std::unique(d.begin(), d.end()), d.end());
        total += static_cast<int64_t>(d.size());
        if (total > static_cast<int64_t>(max_deps) * nt) return s;
    }
    // deps(j) = D(j) + {i : j in D(i)} + {j}, as ranges
    std::vector<std::vector<int32_t>> dep(nt);
    for (int32_t t = 0; t < nt; ++t) {
        dep[t].push_back(t);
        for (int32_t c : D[t]) {
            dep[t].push_back(c);
            dep[c].push_back(t);
        }
    }
    s.dep_ptr.assign(static_cast<size_t>(nt) + 1, 0);
    for (int32_t j = 0; j < nt; ++j) {
        std::vector<int32_t>& d = dep[j];
        std::sort(d.begin(), d.end());
        d.erase(std::unique(d.begin(), d.end()), d.end());
        for (size_t k = 0; k < d.size();) {
            size_t e = k;
            while (e + 1 < d.size() && d[e + 1] == d[e] + 1) ++e;
            s.dep_rng.push_back(d[k]);
            s.dep_rng.push_back(d[e]);
            k = e + 1;
        }
        s.dep_ptr[j + 1] = static_cast<int32_t>(s.dep_rng.size() / 2);
    }
    // slabs and band tiles
    std::vector<int32_t> start(static_cast<size_t>(nq) + 1);
    for (int32_t q = 0; q <= nq; ++q) start[q] = static_cast<int32_t>((static_cast<int64_t>(nt) * q) / nq);
    std::vector<int32_t> slab_of(nt);
    for (int32_t q = 0; q < nq; ++q)
        for (int32_t t = start[q]; t < start[q + 1]; ++t) slab_of[t] = q;
    s.done_init.assign(nt, 0);
    for (int32_t j = 0; j < nt; ++j)
        for (int32_t i : dep[j])
            if (slab_of[i] != slab_of[j]) s.done_init[i] = 1;
    for (int32_t t = 0; t < nt; ++t)
        if (s.done_init[t]) s.band.push_back(t);
    if (static_cast<double>(s.band.size()) > max_band_frac * nt) return s;

    // queues: front-1 tiles ascending, each front-2 tile after the front-1
    // tiles of its slab it waits on (and slack more)
    s.qoff.assign(static_cast<size_t>(nq) + 1, 0);
    s.items.reserve(2 * static_cast<size_t>(nt));
    for (int32_t q = 0; q < nq; ++q) {
        s.qoff[q] = static_cast<int32_t>(s.items.size());
        int32_t e = start[q];  // next front-1 tile to list
        for (int32_t j = start[q]; j < start[q + 1]; ++j) {
            int32_t last = -1;
            for (int32_t i : dep[j])
                if (slab_of[i] == q && !s.done_init[i]) last = std::max(last, i);
            const int64_t need = std::min<int64_t>(static_cast<int64_t>(last) + slack, start[q + 1] - 1);
            for (; e <= need; ++e)
                if (!s.done_init[e]) s.items.push_back(e);
            s.items.push_back(~j);
        }
        for (; e < start[q + 1]; ++e)
            if (!s.done_init[e]) s.items.push_back(e);
    }
    s.qoff[nq] = static_cast<int32_t>(s.items.size());
    s.ok = true;
    return s;
}

}  // namespace tbk
