// Admission core of the request server: the device-memory arena and the
// decisions of which queued requests run, against an HBM budget.  Plain C++
// with no HIP: the server (server.hip) drives it with hipMalloc / hipFree
// behind SlabMemory, and the CPU tests drive the same code with a fake
// device memory that can refuse allocations.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

namespace serve {

// One slab of the arena: a device allocation carved into spans.
struct Slab {
    char* base = nullptr;
    size_t cap = 0;
    std::map<size_t, size_t> holes;  // offset -> bytes
    int32_t live = 0;                // spans handed out and not returned
};

// The server's device memory: the whole budget as one slab allocated at
// creation, or (lazy arena) slabs allocated as requests need them (the
// first of 4 GiB, each later one at least as large as all before it, all
// within the budget) and kept, so a server that has run a request of some
// size holds the memory for the next one.  A request gets one span of one
// slab: the smallest hole that fits, in address order.  Slabs with no span
// handed out are freed only to make room for a larger slab.  Under the
// server's mutex; the device allocations themselves happen outside it.
struct Arena {
    static constexpr size_t kGrain = 64 * 1024;
    static constexpr size_t kFirstSlab = size_t{4} << 30;
    std::vector<std::unique_ptr<Slab>> slabs;
    size_t held = 0;     // bytes of the slabs
    size_t pending = 0;  // bytes of slabs being allocated
    static size_t round(size_t n) { return (n + kGrain - 1) / kGrain * kGrain; }
    static void cut(Slab& sl, std::map<size_t, size_t>::iterator it, size_t n, size_t& off) {
        off = it->first;
        const size_t rest = it->second - n;
        sl.holes.erase(it);
        if (rest) sl.holes[off + n] = rest;
        sl.live++;
    }
    // best fit over the holes of every slab
    bool take(size_t n, Slab*& where, size_t& off) {
        n = round(n);
        Slab* best = nullptr;
        std::map<size_t, size_t>::iterator bit;
        for (auto& sl : slabs)
            for (auto it = sl->holes.begin(); it != sl->holes.end(); ++it)
                if (it->second >= n && (!best || it->second < bit->second)) {
                    best = sl.get();
                    bit = it;
                }
        if (!best) return false;
        cut(*best, bit, n, off);
        where = best;
        return true;
    }
    // size of a new slab for a span of n bytes within `limit`, retiring idle
    // slabs into `retired` when that makes the room; 0: none fits now
    size_t plan_slab(size_t n, size_t limit, std::vector<Slab*>& retired) {
        n = round(n);
        auto size_for = [&](size_t room) -> size_t {
            if (room < n) return 0;
            const size_t want = std::max({n, kFirstSlab, held + pending});
            return std::min(room, want) / kGrain * kGrain;
        };
        const size_t room = limit > held + pending ? limit - held - pending : 0;
        if (size_t sz = size_for(room)) return sz;
        size_t idle = 0;
        for (auto& sl : slabs)
            if (sl->live == 0) idle += sl->cap;
        if (room + idle < n) return 0;
        // retire idle slabs, smallest first, until the span fits
        std::vector<Slab*> cand;
        for (auto& sl : slabs)
            if (sl->live == 0) cand.push_back(sl.get());
        std::sort(cand.begin(), cand.end(), [](const Slab* a, const Slab* b) { return a->cap < b->cap; });
        size_t freed = 0;
        for (Slab* c : cand) {
            if (room + freed >= n) break;
            freed += c->cap;
            retired.push_back(c);
        }
        for (Slab* c : retired) {
            held -= c->cap;
            for (auto& sl : slabs)
                if (sl.get() == c) sl.release();
        }
        slabs.erase(std::remove(slabs.begin(), slabs.end(), nullptr), slabs.end());
        return size_for(room + freed);
    }
    void give(Slab* sl, size_t off, size_t n) {
        n = round(n);
        sl->live--;
        auto next = sl->holes.lower_bound(off);
        if (next != sl->holes.end() && off + n == next->first) {
            n += next->second;
            next = sl->holes.erase(next);
        }
        if (next != sl->holes.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += n;
                return;
            }
        }
        sl->holes[off] = n;
    }
};

// Device memory behind the slabs: hipMalloc / hipFree in the server, a fake
// in the CPU tests.  allocate returns null when the device refuses.
struct SlabMemory {
    virtual ~SlabMemory() = default;
    virtual char* allocate(size_t bytes) = 0;
    virtual void release(char* base) = 0;
};

// What admission needs of a request (the server's Request derives from it).
struct Ticket {
    int64_t footprint = 0;
    bool dos = true;          // the DoS or the sigma_xy pool of result rows
    double t_submit = 0.0;    // seconds
    Slab* slab = nullptr;     // its span of the arena while admitted (null: a new slab is being made for it,
                              // or the device refused one)
    size_t arena_off = 0;
    size_t new_slab = 0;      // bytes of that new slab
};

// Admission against an HBM budget.  State under the caller's mutex except
// where a function says otherwise.
struct Admission {
    int64_t budget = 0;
    int64_t used = 0;                   // footprints of admitted, unfinished requests
    double max_wait_bypass_ms = 1000.0;
    Arena arena;
    std::vector<Slab*> retired;         // idle slabs to free (outside the mutex)

    // a request larger than the whole budget can never run
    bool admissible(int64_t footprint) const { return footprint <= budget; }

    // The whole budget as one slab (at creation, no mutex needed); false if
    // the device refuses it.
    bool reserve_whole(SlabMemory& mem) {
        auto sl = std::make_unique<Slab>();
        sl->cap = static_cast<size_t>(budget) / Arena::kGrain * Arena::kGrain;
        sl->base = mem.allocate(sl->cap);
        if (!sl->base) return false;
        sl->holes[0] = sl->cap;
        arena.held = sl->cap;
        arena.slabs.push_back(std::move(sl));
        return true;
    }

    // First fit over the queue in arrival order: a request runs when a
    // result row is free for it (has_row), its footprint fits the rest of
    // the budget, and the arena has a span for it or room for a new slab;
    // take_row then reserves its row.  A request that has waited more than
    // max_wait_bypass_ms stops later ones from passing it.  The admitted
    // tickets leave the queue (a sequence of Ticket-derived pointers).
    template <typename Queue, typename HasRow, typename TakeRow>
    std::vector<typename Queue::value_type> pick(Queue& queue, double now, HasRow has_row, TakeRow take_row) {
        std::vector<typename Queue::value_type> out;
        for (auto it = queue.begin(); it != queue.end();) {
            Ticket* t = *it;
            size_t off = 0, nsl = 0;
            Slab* sl = nullptr;
            const bool fits = has_row(*it) && used + t->footprint <= budget &&
                              (arena.take(static_cast<size_t>(t->footprint), sl, off) ||
                               (nsl = arena.plan_slab(static_cast<size_t>(t->footprint), static_cast<size_t>(budget),
                                                      retired)) > 0);
            if (fits) {
                t->slab = sl;
                t->arena_off = off;
                t->new_slab = nsl;
                arena.pending += nsl;
                used += t->footprint;
                take_row(*it);
                out.push_back(*it);
                it = queue.erase(it);
                continue;
            }
            if (1e3 * (now - t->t_submit) > max_wait_bypass_ms) break;
            ++it;
        }
        return out;
    }

    // Outside the mutex (mu is taken for the state updates): free the
    // retired slabs, then allocate the new slab of every admitted ticket
    // that needs one and place its span at the slab's start.  A refused
    // allocation leaves the ticket without a slab; the caller fails it.
    template <typename Adm>
    void place(Adm& adm, SlabMemory& mem, std::mutex& mu) {
        std::vector<Slab*> dead;
        {
            std::lock_guard<std::mutex> lk(mu);
            dead.swap(retired);
        }
        for (Slab* sl : dead) {
            mem.release(sl->base);
            delete sl;
        }
        for (auto& a : adm) {
            Ticket* t = a;
            if (t->slab) continue;
            auto sl = std::make_unique<Slab>();
            char* p = mem.allocate(t->new_slab);
            std::lock_guard<std::mutex> lk(mu);
            arena.pending -= t->new_slab;
            if (!p) {
                t->slab = nullptr;
                continue;
            }
            sl->base = p;
            sl->cap = t->new_slab;
            sl->holes[0] = sl->cap;
            size_t off = 0;
            Arena::cut(*sl, sl->holes.begin(), Arena::round(static_cast<size_t>(t->footprint)), off);
            t->slab = sl.get();
            t->arena_off = off;
            arena.held += sl->cap;
            arena.slabs.push_back(std::move(sl));
        }
    }

    // a finished (or failed) request's footprint back to the budget
    void finished(const Ticket& t) { used -= t.footprint; }
    // a span back to the arena
    void give(Slab* sl, size_t off, size_t bytes) { arena.give(sl, off, bytes); }

    // every slab back to the device (at destruction, no mutex needed)
    void release_all(SlabMemory& mem) {
        for (auto& sl : arena.slabs) mem.release(sl->base);
        arena.slabs.clear();
        arena.held = 0;
        for (Slab* sl : retired) {
            mem.release(sl->base);
            delete sl;
        }
        retired.clear();
    }
};

}  // namespace serve
