// Host test of the two-step schedule (tight_binding_amd/csrc/pair_plan.h).
//
// For lattice and banded matrices it checks the schedule's invariants, then
// replays the two-step launch on the CPU: the band tiles first (the
// prologue), then workers of each queue drawing tickets in order, one ahead,
// each item reading its inputs when its waits are met and writing its rows
// at a later, random moment.  The result must equal two plain steps bit for
// bit, and no replay may stall with every worker waiting.
//
// Usage: pair_schedule [seed]; prints "ok <cases>" or the first failure.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "pair_plan.h"

namespace {

struct Csr {
    int32_t n = 0;
    std::vector<int32_t> rp, ci;
    std::vector<double> v;
};

Csr from_rows(int32_t n, std::vector<std::vector<std::pair<int32_t, double>>>& rows) {
    Csr m;
    m.n = n;
    m.rp.push_back(0);
    for (auto& r : rows) {
        std::sort(r.begin(), r.end());
        for (auto& e : r) {
            m.ci.push_back(e.first);
            m.v.push_back(e.second);
        }
        m.rp.push_back(static_cast<int32_t>(m.ci.size()));
    }
    return m;
}

// honeycomb-like: L x L cells of 2 sites, three bonds per site, periodic
// in x and, if periodic_y, in y; on-site terms
Csr lattice(int32_t L, bool periodic_y, std::mt19937_64& g) {
    const int32_t n = 2 * L * L;
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<std::vector<std::pair<int32_t, double>>> rows(n);
    auto id = [&](int32_t x, int32_t y, int32_t s) { return 2 * (y * L + x) + s; };
    auto bond = [&](int32_t a, int32_t b) {
        const double t = U(g);
        rows[a].push_back({b, t});
        rows[b].push_back({a, t});
    };
    for (int32_t y = 0; y < L; ++y)
        for (int32_t x = 0; x < L; ++x) {
            const int32_t a = id(x, y, 0);
            rows[a].push_back({a, U(g)});
            rows[a + 1].push_back({a + 1, U(g)});
            bond(a, a + 1);
            bond(a, id((x + 1) % L, y, 1));
            if (y + 1 < L || periodic_y) bond(a, id(x, (y + 1) % L, 1));
        }
    return from_rows(n, rows);
}

// random symmetric pattern within bandwidth bw, some empty rows
Csr banded(int32_t n, int32_t bw, std::mt19937_64& g) {
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::uniform_int_distribution<int32_t> off(1, bw);
    std::vector<std::vector<std::pair<int32_t, double>>> rows(n);
    for (int32_t r = 0; r < n; ++r) {
        if (g() % 7 == 0) continue;
        for (int k = 0; k < 2; ++k) {
            const int32_t c = r + off(g);
            if (c >= n) continue;
            const double t = U(g);
            rows[r].push_back({c, t});
            rows[c].push_back({r, t});
        }
    }
    for (auto& r : rows) {  // merge duplicate columns
        std::sort(r.begin(), r.end());
        std::vector<std::pair<int32_t, double>> u;
        for (auto& e : r)
            if (!u.empty() && u.back().first == e.first)
                u.back().second += e.second;
            else
                u.push_back(e);
        r = u;
    }
    return from_rows(n, rows);
}

// non-symmetric: row r has columns r .. r + bw only (a row's readers are
// then all behind it)
Csr one_sided(int32_t n, int32_t bw, std::mt19937_64& g) {
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<std::vector<std::pair<int32_t, double>>> rows(n);
    for (int32_t r = 0; r < n; ++r)
        for (int32_t c = r; c <= std::min(n - 1, r + bw); c += 1 + static_cast<int32_t>(g() % 7))
            rows[r].push_back({c, U(g)});
    return from_rows(n, rows);
}

// y = alpha*(H x) + beta*x - z on rows [r0, r1)
void rows_step(const Csr& H, double alpha, double beta, const std::vector<double>& x, const std::vector<double>& z,
               std::vector<double>& y, int32_t r0, int32_t r1) {
    for (int32_t r = r0; r < r1; ++r) {
        double acc = 0.0;
        for (int32_t p = H.rp[r]; p < H.rp[r + 1]; ++p) acc = acc + H.v[p] * x[H.ci[p]];
        y[r] = alpha * acc + beta * x[r] - z[r];
    }
}

int fails = 0;
void fail(const std::string& what) {
    if (fails++ < 5) std::printf("FAIL %s\n", what.c_str());
}

bool check_case(const Csr& H, int32_t tt, int32_t nq, int32_t slack, uint64_t seed, bool expect_ok,
                const std::string& name, double max_band = 0.95) {
    const tbk::PairSchedule s = tbk::build_pair_schedule(H.rp, H.ci, H.n, tt, nq, slack, max_band);
    if (s.ok != expect_ok) {
        fail(name + ": ok = " + std::to_string(s.ok));
        return false;
    }
    if (!s.ok) return true;
    const int32_t nt = s.ntiles;
    auto tile = [&](int32_t r) { return r / tt; };
    std::vector<int32_t> start(nq + 1);
    for (int32_t q = 0; q <= nq; ++q) start[q] = static_cast<int32_t>((static_cast<int64_t>(nt) * q) / nq);
    auto slab = [&](int32_t t) {
        int32_t q = 0;
        while (t >= start[q + 1]) ++q;
        return q;
    };
    auto in_deps = [&](int32_t j, int32_t i) {
        for (int32_t g = s.dep_ptr[j]; g < s.dep_ptr[j + 1]; ++g)
            if (s.dep_rng[2 * g] <= i && i <= s.dep_rng[2 * g + 1]) return true;
        return false;
    };
    // deps cover the exact dependencies
    for (int32_t r = 0; r < H.n; ++r)
        for (int32_t p = H.rp[r]; p < H.rp[r + 1]; ++p) {
            const int32_t a = tile(r), b = tile(H.ci[p]);
            if (!in_deps(a, b) || !in_deps(b, a)) fail(name + ": missing dependency");
        }
    for (int32_t j = 0; j < nt; ++j) {
        if (!in_deps(j, j)) fail(name + ": tile not in its own deps");
        for (int32_t g = s.dep_ptr[j]; g < s.dep_ptr[j + 1]; ++g)
            for (int32_t i = s.dep_rng[2 * g]; i <= s.dep_rng[2 * g + 1]; ++i)
                if (slab(i) != slab(j) && !s.done_init[i]) fail(name + ": foreign dependency not in the band");
    }
    // items: each non-band tile once in front 1, each tile once in front 2,
    // in its own slab's queue, each front-2 item after its waits
    std::vector<int> seen1(nt, 0), seen2(nt, 0);
    for (int32_t q = 0; q < nq; ++q) {
        std::vector<int32_t> pos1(nt, -1);
        for (int32_t k = s.qoff[q]; k < s.qoff[q + 1]; ++k) {
            const int32_t it = s.items[k];
            const int32_t t = it >= 0 ? it : ~it;
            if (t < 0 || t >= nt || slab(t) != q) {
                fail(name + ": item outside its slab");
                continue;
            }
            if (it >= 0) {
                seen1[t]++;
                pos1[t] = k;
                if (s.done_init[t]) fail(name + ": band tile queued");
            } else {
                seen2[t]++;
                for (int32_t g = s.dep_ptr[t]; g < s.dep_ptr[t + 1]; ++g)
                    for (int32_t i = s.dep_rng[2 * g]; i <= s.dep_rng[2 * g + 1]; ++i)
                        if (!s.done_init[i] && pos1[i] < 0) fail(name + ": front-2 item before a front-1 it waits on");
            }
        }
    }
    for (int32_t t = 0; t < nt; ++t) {
        if (seen2[t] != 1) fail(name + ": front-2 tile count");
        if (seen1[t] != (s.done_init[t] ? 0 : 1)) fail(name + ": front-1 tile count");
    }
    if (fails) return false;

    // replay against two plain steps
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const double alpha = 0.7, beta = -0.1;
    std::vector<double> P(H.n), Q(H.n);
    for (auto& x : P) x = U(g);
    for (auto& x : Q) x = U(g);
    std::vector<double> T1(H.n), T2(H.n);
    rows_step(H, alpha, beta, Q, P, T1, 0, H.n);
    rows_step(H, alpha, beta, T1, Q, T2, 0, H.n);

    std::vector<int> done(s.done_init.begin(), s.done_init.end());
    for (int32_t t : s.band) rows_step(H, alpha, beta, Q, P, P, t * tt, std::min(H.n, t * tt + tt));
    struct Worker {
        int32_t q;
        int32_t cur = -1, next = -1;  // tickets
        bool running = false;
        std::vector<double> out;  // rows computed at start, written at the end
    };
    std::vector<int32_t> counter(nq, 0);
    std::vector<Worker> W;
    const int per_q = 1 + static_cast<int>(g() % 5);
    for (int32_t q = 0; q < nq; ++q)
        for (int w = 0; w < per_q; ++w) {
            Worker x;
            x.q = q;
            x.cur = counter[q]++;
            W.push_back(x);
        }
    auto len = [&](int32_t q) { return s.qoff[q + 1] - s.qoff[q]; };
    auto ready = [&](int32_t it) {
        if (it >= 0) return true;
        const int32_t j = ~it;
        for (int32_t gi = s.dep_ptr[j]; gi < s.dep_ptr[j + 1]; ++gi)
            for (int32_t i = s.dep_rng[2 * gi]; i <= s.dep_rng[2 * gi + 1]; ++i)
                if (!done[i]) return false;
        return true;
    };
    size_t live = W.size();
    while (live > 0) {
        std::vector<size_t> can;
        for (size_t w = 0; w < W.size(); ++w) {
            Worker& x = W[w];
            if (x.cur < 0) continue;
            if (x.running || ready(s.items[s.qoff[x.q] + x.cur])) can.push_back(w);
        }
        if (can.empty()) {
            fail(name + ": replay stalled with every worker waiting");
            return false;
        }
        Worker& x = W[can[g() % can.size()]];
        const int32_t it = s.items[s.qoff[x.q] + x.cur];
        const int32_t t = it >= 0 ? it : ~it;
        const int32_t r0 = t * tt, r1 = std::min(H.n, r0 + tt);
        if (!x.running) {  // start: draw the next ticket, read the inputs
            x.next = counter[x.q]++;
            std::vector<double> y(it >= 0 ? P : Q);
            if (it >= 0)
                rows_step(H, alpha, beta, Q, P, y, r0, r1);
            else
                rows_step(H, alpha, beta, P, Q, y, r0, r1);
            x.out.assign(y.begin() + r0, y.begin() + r1);
            x.running = true;
        } else {  // finish: write the rows, publish, move to the next ticket
            std::vector<double>& dst = it >= 0 ? P : Q;
            std::copy(x.out.begin(), x.out.end(), dst.begin() + r0);
            if (it >= 0) done[t] = 1;
            x.running = false;
            x.cur = x.next < len(x.q) ? x.next : -1;
            if (x.cur < 0) --live;
        }
    }
    for (int32_t r = 0; r < H.n; ++r)
        if (P[r] != T1[r] || Q[r] != T2[r]) {
            fail(name + ": replay differs from two plain steps at row " + std::to_string(r));
            return false;
        }
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    const uint64_t seed = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1;
    std::mt19937_64 g(seed);
    int cases = 0;
    for (int rep = 0; rep < 3; ++rep) {
        const Csr hp = lattice(48, true, g), ho = lattice(40, false, g);
        for (int32_t tt : {4, 16, 64}) {
            for (int32_t slack : {0, 3}) {
                cases += check_case(hp, tt, 8, slack, g(), true, "periodic tt=" + std::to_string(tt));
                cases += check_case(ho, tt, 8, slack, g(), true, "open tt=" + std::to_string(tt));
            }
        }
        const Csr b = banded(20000 + static_cast<int32_t>(g() % 1000), 60, g);
        cases += check_case(b, 16, 8, 0, g(), true, "banded");
        cases += check_case(b, 16, 3, 2, g(), true, "banded, 3 queues");
        const Csr u = one_sided(12000, 90, g);
        cases += check_case(u, 8, 8, 0, g(), true, "one-sided");
        cases += check_case(u, 32, 4, 1, g(), true, "one-sided, 4 queues");
    }
    // too small for the queues, and too scattered for the band
    const Csr small = lattice(4, true, g);
    cases += check_case(small, 64, 8, 0, 1, false, "small");
    {
        std::uniform_int_distribution<int32_t> col(0, 9999);
        std::vector<std::vector<std::pair<int32_t, double>>> rows(10000);
        for (int32_t r = 0; r < 10000; ++r) {
            const int32_t c = col(g);
            rows[r].push_back({c, 1.0});
            if (c != r) rows[c].push_back({r, 1.0});
        }
        for (auto& r : rows) {
            std::sort(r.begin(), r.end());
            r.erase(std::unique(r.begin(), r.end(), [](auto& a, auto& b) { return a.first == b.first; }), r.end());
        }
        cases += check_case(from_rows(10000, rows), 16, 8, 0, 1, false, "scattered", 0.125);
    }
    if (fails) return 1;
    std::printf("ok %d\n", cases);
    return 0;
}
