// The request server's admission core (tight_binding_amd/csrc/admission.h)
// on the CPU, over a fake device memory that hands out address ranges and
// can refuse allocations: admission against the budget, result rows, the
// waiting-time rule, best fit and hole merging, lazy slabs and their
// retirement, a refused device allocation, and the return of every slab.
// Prints "ok <case>" per case; exits non-zero on the first failed check.
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <map>
#include <mutex>
#include <vector>

#include "admission.h"

namespace {

constexpr size_t GB = size_t{1} << 30;

// a failed check exits at once
#define CHECK(c)                                                           \
    do {                                                                   \
        if (!(c)) {                                                        \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            std::exit(1);                                                  \
        }                                                                  \
    } while (0)

struct FakeMemory : serve::SlabMemory {
    uintptr_t next = uintptr_t{1} << 40;
    std::map<char*, size_t> live;
    int refuse_next = 0;
    size_t allocations = 0;
    char* allocate(size_t bytes) override {
        if (refuse_next > 0) {
            --refuse_next;
            return nullptr;
        }
        char* p = reinterpret_cast<char*>(next);
        next += bytes + GB;
        live[p] = bytes;
        ++allocations;
        return p;
    }
    void release(char* base) override {
        CHECK(live.count(base) == 1);
        live.erase(base);
    }
};

struct Req : serve::Ticket {
    int id = 0;
};

Req* make(std::vector<Req>& pool, int id, size_t bytes, bool dos, double t) {
    Req& r = pool[static_cast<size_t>(id)];
    r.id = id;
    r.footprint = static_cast<int64_t>(bytes);
    r.dos = dos;
    r.t_submit = t;
    return &r;
}

// rows: free result rows per pool (DoS, sigma_xy)
struct Rows {
    int dos = 100, kubo = 100;
    bool has(const Req* r) const { return (r->dos ? dos : kubo) > 0; }
    void take(const Req* r) { --(r->dos ? dos : kubo); }
};

std::vector<Req*> pick(serve::Admission& a, std::deque<Req*>& q, Rows& rows, double now) {
    return a.pick(q, now, [&](Req* r) { return rows.has(r); }, [&](Req* r) { rows.take(r); });
}

void finish(serve::Admission& a, Req* r, bool gives_span = true) {
    if (gives_span && r->slab) a.give(r->slab, r->arena_off, static_cast<size_t>(r->footprint));
    a.finished(*r);
}

bool disjoint(const std::vector<Req*>& rs) {
    for (size_t i = 0; i < rs.size(); ++i)
        for (size_t j = i + 1; j < rs.size(); ++j) {
            const Req* x = rs[i];
            const Req* y = rs[j];
            if (x->slab != y->slab) continue;
            const size_t xe = x->arena_off + serve::Arena::round(static_cast<size_t>(x->footprint));
            const size_t ye = y->arena_off + serve::Arena::round(static_cast<size_t>(y->footprint));
            if (x->arena_off < ye && y->arena_off < xe) return false;
        }
    return true;
}

}  // namespace

int main() {
    std::vector<Req> pool(64);
    std::mutex mu;
    {  // the whole budget as one slab; first fit in arrival order within the budget
        FakeMemory mem;
        serve::Admission a;
        a.budget = static_cast<int64_t>(10 * GB);
        CHECK(a.reserve_whole(mem) && a.arena.slabs.size() == 1 && a.arena.held == 10 * GB);
        CHECK(a.admissible(static_cast<int64_t>(10 * GB)) && !a.admissible(static_cast<int64_t>(10 * GB) + 1));
        Rows rows;
        std::deque<Req*> q = {make(pool, 0, 4 * GB, true, 0.0), make(pool, 1, 4 * GB, false, 0.0),
                              make(pool, 2, 4 * GB, true, 0.0)};
        auto got = pick(a, q, rows, 0.0);
        CHECK(got.size() == 2 && got[0]->id == 0 && got[1]->id == 1 && q.size() == 1);
        CHECK(a.used == static_cast<int64_t>(8 * GB) && disjoint(got));
        CHECK(got[0]->arena_off % serve::Arena::kGrain == 0 && got[1]->arena_off % serve::Arena::kGrain == 0);
        a.place(got, mem, mu);  // spans of the whole-budget slab: nothing allocated
        CHECK(mem.allocations == 1);
        CHECK(pick(a, q, rows, 0.0).empty());
        finish(a, got[0]);
        auto got2 = pick(a, q, rows, 0.0);
        CHECK(got2.size() == 1 && got2[0]->id == 2 && q.empty() && a.used == static_cast<int64_t>(8 * GB));
        finish(a, got[1]);
        finish(a, got2[0]);
        CHECK(a.used == 0 && a.arena.slabs[0]->holes.size() == 1 && a.arena.slabs[0]->holes.begin()->second == 10 * GB);
        a.release_all(mem);
        CHECK(mem.live.empty());
        std::printf("ok whole_budget\n");
    }
    {  // result rows: a request waits for a free row of its own pool
        FakeMemory mem;
        serve::Admission a;
        a.budget = static_cast<int64_t>(10 * GB);
        CHECK(a.reserve_whole(mem));
        Rows rows;
        rows.dos = 1;
        rows.kubo = 0;
        std::deque<Req*> q = {make(pool, 3, GB, true, 0.0), make(pool, 4, GB, true, 0.0), make(pool, 5, GB, false, 0.0)};
        auto got = pick(a, q, rows, 0.0);
        CHECK(got.size() == 1 && got[0]->id == 3 && rows.dos == 0 && a.used == static_cast<int64_t>(GB));
        rows.dos = 1;
        rows.kubo = 1;
        auto got2 = pick(a, q, rows, 0.0);
        CHECK(got2.size() == 2 && got2[0]->id == 4 && got2[1]->id == 5);
        a.release_all(mem);
        std::printf("ok result_rows\n");
    }
    {  // a request waiting past max_wait_bypass_ms stops later ones from passing it
        FakeMemory mem;
        serve::Admission a;
        a.budget = static_cast<int64_t>(10 * GB);
        a.max_wait_bypass_ms = 1000.0;
        CHECK(a.reserve_whole(mem));
        Rows rows;
        Req* busy = make(pool, 6, 6 * GB, true, 0.0);
        std::deque<Req*> q0 = {busy};
        CHECK(pick(a, q0, rows, 0.0).size() == 1);
        std::deque<Req*> q = {make(pool, 7, 6 * GB, true, 0.0), make(pool, 8, 2 * GB, true, 0.5)};
        auto early = pick(a, q, rows, 0.5);  // the big one has waited 0.5 s: the small one passes
        CHECK(early.size() == 1 && early[0]->id == 8 && q.size() == 1);
        std::deque<Req*> q2 = {make(pool, 9, 6 * GB, true, 0.0), make(pool, 10, GB, true, 1.9)};
        auto late = pick(a, q2, rows, 2.0);  // waited 2 s: nothing passes it
        CHECK(late.empty() && q2.size() == 2);
        a.release_all(mem);
        std::printf("ok bypass_order\n");
    }
    {  // best fit over the holes, and neighbouring holes merge on return
        FakeMemory mem;
        serve::Admission a;
        a.budget = static_cast<int64_t>(6 * GB);
        CHECK(a.reserve_whole(mem));
        Rows rows;
        std::deque<Req*> q = {make(pool, 11, 2 * GB, true, 0.0), make(pool, 12, GB, true, 0.0),
                              make(pool, 13, 2 * GB, true, 0.0), make(pool, 14, GB, true, 0.0)};
        auto got = pick(a, q, rows, 0.0);
        CHECK(got.size() == 4 && disjoint(got));
        finish(a, got[0]);  // a 2 GB hole at the start
        finish(a, got[3]);  // a 1 GB hole at the end
        std::deque<Req*> q2 = {make(pool, 15, GB, true, 0.0)};
        auto one = pick(a, q2, rows, 0.0);
        CHECK(one.size() == 1 && one[0]->arena_off == got[3]->arena_off);  // the smallest hole that fits
        finish(a, one[0]);
        finish(a, got[1]);
        finish(a, got[2]);
        const auto& holes = a.arena.slabs[0]->holes;
        CHECK(holes.size() == 1 && holes.begin()->first == 0 && holes.begin()->second == 6 * GB);
        CHECK(a.arena.slabs[0]->live == 0 && a.used == 0);
        a.release_all(mem);
        std::printf("ok best_fit_and_merge\n");
    }
    {  // lazy arena: slabs made as needed (the first 4 GiB, later ones as large
       // as all before), kept, and retired to make room for a larger one
        FakeMemory mem;
        serve::Admission a;
        a.budget = static_cast<int64_t>(12 * GB);
        Rows rows;
        std::deque<Req*> q = {make(pool, 16, GB, true, 0.0)};
        auto got = pick(a, q, rows, 0.0);
        CHECK(got.size() == 1 && got[0]->slab == nullptr && got[0]->new_slab == serve::Arena::kFirstSlab);
        CHECK(a.arena.pending == serve::Arena::kFirstSlab);
        a.place(got, mem, mu);
        CHECK(got[0]->slab != nullptr && got[0]->arena_off == 0 && a.arena.held == 4 * GB && a.arena.pending == 0);
        std::deque<Req*> q2 = {make(pool, 17, 5 * GB, true, 0.0)};
        auto big = pick(a, q2, rows, 0.0);
        CHECK(big.size() == 1 && big[0]->new_slab == 5 * GB);  // max(n, 4 GiB, held)
        a.place(big, mem, mu);
        CHECK(a.arena.held == 9 * GB && mem.live.size() == 2);
        finish(a, got[0]);
        finish(a, big[0]);
        // 7 GB fits no hole: 3 GB of budget left plus the smallest idle slab
        // (4 GiB), retired for a new 7 GB slab; the 5 GB slab stays
        std::deque<Req*> q3 = {make(pool, 18, 7 * GB, true, 0.0)};
        auto larger = pick(a, q3, rows, 0.0);
        CHECK(larger.size() == 1 && larger[0]->slab == nullptr && larger[0]->new_slab == 7 * GB);
        CHECK(a.retired.size() == 1 && a.arena.held == 5 * GB);
        a.place(larger, mem, mu);
        CHECK(larger[0]->slab != nullptr && mem.live.size() == 2 && a.arena.held == 12 * GB);
        a.release_all(mem);
        CHECK(mem.live.empty());
        std::printf("ok lazy_slabs\n");
    }
    {  // retirement: no hole fits and the budget is held by idle slabs
        FakeMemory mem;
        serve::Admission a;
        a.budget = static_cast<int64_t>(10 * GB);
        Rows rows;
        std::deque<Req*> q = {make(pool, 19, GB, true, 0.0)};
        auto s1 = pick(a, q, rows, 0.0);
        a.place(s1, mem, mu);  // a 4 GiB slab
        finish(a, s1[0]);
        std::deque<Req*> q2 = {make(pool, 20, 8 * GB, true, 0.0)};
        auto s2 = pick(a, q2, rows, 0.0);
        CHECK(s2.size() == 1 && s2[0]->slab == nullptr && a.retired.size() == 1);
        CHECK(s2[0]->new_slab >= 8 * GB && s2[0]->new_slab <= 10 * GB);
        a.place(s2, mem, mu);  // frees the retired slab first
        CHECK(a.retired.empty() && mem.live.size() == 1 && a.arena.slabs.size() == 1 && s2[0]->slab != nullptr);
        a.release_all(mem);
        CHECK(mem.live.empty());
        std::printf("ok retire_idle_slab\n");
    }
    {  // a device that refuses the slab: the request is left without a span
       // (the server fails it with TBK_E_OOM) and its bytes return to the budget
        FakeMemory mem;
        serve::Admission a;
        a.budget = static_cast<int64_t>(10 * GB);
        Rows rows;
        mem.refuse_next = 1;
        std::deque<Req*> q = {make(pool, 21, 2 * GB, false, 0.0)};
        auto got = pick(a, q, rows, 0.0);
        CHECK(got.size() == 1 && a.used == static_cast<int64_t>(2 * GB));
        a.place(got, mem, mu);
        CHECK(got[0]->slab == nullptr && a.arena.pending == 0 && a.arena.held == 0 && mem.live.empty());
        finish(a, got[0], false);
        CHECK(a.used == 0);
        std::deque<Req*> q2 = {make(pool, 22, 2 * GB, false, 0.0)};
        auto again = pick(a, q2, rows, 0.0);
        a.place(again, mem, mu);
        CHECK(again[0]->slab != nullptr && mem.live.size() == 1);
        a.release_all(mem);
        CHECK(mem.live.empty());
        // refusal of the whole budget at creation
        FakeMemory no;
        no.refuse_next = 1;
        serve::Admission b;
        b.budget = static_cast<int64_t>(GB);
        CHECK(!b.reserve_whole(no) && b.arena.slabs.empty());
        std::printf("ok device_refusal\n");
    }
    std::printf("all ok\n");
    return 0;
}
