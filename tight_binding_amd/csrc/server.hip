// Request server: DoS and sigma_xy requests of mixed models and sizes,
// continuously batched on one GPU.
//
// A worker thread owns the device work.  The server's device memory is an
// arena: its whole budget allocated at creation (or slabs grown within the
// budget as requests need them, and kept): every admitted request gets one
// contiguous span of it (its footprint), and its matrices, vector blocks,
// stacks and Gamma buffers are carved from that span, so serving allocates
// and frees no device memory for requests (freeing and reallocating
// hundreds of GB costs seconds per request).  Requests wait in
// a queue until their footprint fits the budget and a span of the arena
// (and a slot is free); then they are admitted while other requests are
// running:
//
//  * DoS request -> a slot: its matrix, three rotating vector blocks and a
//    moment buffer stay resident until it finishes.  All slots of one class
//    (precision, real/complex values, padded vector width) advance together:
//    a chunk of S steps is S launches of cheb_slots_kernel over the tiles of
//    every slot of the class, each slot at its own step index, followed by a
//    moment launch per step.  A slot joins at the next chunk after admission
//    and leaves after the chunk in which it finishes.  Its tiles, partial
//    sums and their reduction order depend only on the slot itself, so its
//    moments are the same bits alone and in any mix.
//  * sigma_xy request -> a Kubo job on a second stream: both vector stacks,
//    the Gamma contraction on matrix cores, the kernel weights and zeta at
//    the N Chebyshev nodes, overlapping the DoS chunks on the device.
//
// Two chunks are in flight at most: while the device runs chunk c, the host
// completes chunk c-1 (results to the requests, memory back to the budget)
// and admits new requests for chunk c+1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "admission.h"
#include "aux_kernels.h"
#include "engine.h"
#include "gamma_launch.h"

namespace {
using namespace engine;


// spectral range of auto-bounds requests: Lanczos steps, residual tolerance,
// seed mix (the request's seed xor this), contraction epsilon
constexpr int kAutoIters = 200;
constexpr double kAutoTol = 1e-3;
constexpr uint64_t kAutoSeedMix = 0xB0D5B0D5B0D5B0D5ull;
constexpr double kAutoEpsilon = 0.01;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Host copy of a sparse matrix handed to submit.
struct HostSparse {
    tbk_sparse hdr{};
    std::vector<int32_t> outer, inner;
    std::vector<double> values;
    bool present = false;
    void copy_from(const tbk_sparse* M) {
        if (!M) return;
        if (M->rows <= 0 || M->rows != M->cols) raise(TBK_E_INVALID, "matrix must be square and non-empty");
        if (M->layout != TBK_CSC && M->layout != TBK_CSR) raise(TBK_E_INVALID, "unknown layout");
        if (M->nnz < 0 || M->nnz > INT32_MAX) raise(TBK_E_INVALID, "nnz out of range for int32 indices");
        if (!M->outer || (M->nnz > 0 && (!M->inner || !M->values))) raise(TBK_E_INVALID, "null array");
        hdr = *M;
        const int32_t no = (M->layout == TBK_CSC ? M->cols : M->rows) + 1;
        outer.assign(M->outer, M->outer + no);
        inner.assign(M->inner, M->inner + M->nnz);
        values.assign(M->values, M->values + M->nnz * (M->is_complex ? 2 : 1));
        hdr.outer = outer.data();
        hdr.inner = inner.data();
        hdr.values = values.data();
        present = true;
    }
    void clear() {
        std::vector<int32_t>().swap(outer);
        std::vector<int32_t>().swap(inner);
        std::vector<double>().swap(values);
        present = false;
    }
};

// Host copy of a lattice description handed to submit (assembled on the
// device at admission).
struct HostLattice {
    tbk_lattice_desc desc{};
    std::vector<int32_t> from, to, lv;
    std::vector<double> t, phase, apos;
    bool present = false;
    void copy_from(const tbk_lattice_desc* d) {
        if (d->dim < 1 || d->dim > 3 || d->n_atoms < 1 || d->n_hops < 0 || d->builder < 0 || d->builder > 2)
            raise(TBK_E_INVALID, "invalid lattice description");
        if (d->n_hops > 0 && (!d->hop_from || !d->hop_to || !d->hop_lv || !d->hop_t))
            raise(TBK_E_INVALID, "null hopping table");
        for (int k = 0; k < d->dim; ++k)
            if (d->span[k] < 1) raise(TBK_E_INVALID, "span must be positive");
        desc = *d;
        const size_t nh = static_cast<size_t>(d->n_hops);
        from.assign(d->hop_from, d->hop_from + nh);
        to.assign(d->hop_to, d->hop_to + nh);
        lv.assign(d->hop_lv, d->hop_lv + nh * d->dim);
        t.assign(d->hop_t, d->hop_t + 2 * nh);
        if (d->phase && d->builder != 2) {
            const size_t pn = static_cast<size_t>(engine::lattice_rows(d) / d->n_atoms) * nh * 2;
            phase.assign(d->phase, d->phase + pn);
        }
        desc.hop_from = from.data();
        desc.hop_to = to.data();
        desc.hop_lv = lv.data();
        desc.hop_t = t.data();
        desc.phase = phase.empty() ? nullptr : phase.data();
        if (d->atom_pos) apos.assign(d->atom_pos, d->atom_pos + 3 * static_cast<size_t>(d->n_atoms));
        desc.atom_pos = apos.empty() ? nullptr : apos.data();
        present = true;
    }
    void clear() {
        std::vector<int32_t>().swap(from);
        std::vector<int32_t>().swap(to);
        std::vector<int32_t>().swap(lv);
        std::vector<double>().swap(t);
        std::vector<double>().swap(phase);
        std::vector<double>().swap(apos);
        present = false;
    }
};

using serve::Arena;
using serve::Slab;

// A submitted request; serve::Ticket holds its footprint, submission time
// and arena placement.
struct Request : serve::Ticket {
    int64_t ticket = 0;
    int32_t kind = TBK_REQ_DOS, precision = TBK_FP64, n = 0, R = 0, Rp = 0;
    uint64_t seed = 0;
    double a = 1.0, b = 0.0;
    bool auto_bounds = false;  // DoS submitted with a = 0: a, b from the spectral range at admission
    HostSparse H, vx, vy;      // a lattice request keeps only H.hdr's rows, is_complex and an nnz bound
    HostLattice lat;
    std::vector<double> g;
    int32_t Gw = 0;            // sigma_xy: random vectors per group (a power of two <= Rp; see kubo_footprint)
    int32_t row_chunks = 1;    // sigma_xy, fp64 stacks: row ranges the stacks hold one at a time
    std::atomic<bool> cancel{false};  // tbk_server_cancel on a running request
    int32_t state = TBK_REQ_QUEUED;
    int32_t error_code = 0;
    std::string error;
    double t_admit = 0, t_done = 0;
    std::vector<double> result;
};

}  // namespace
struct tbk_server;
namespace {

// The arena span a request's device buffers are carved from; handed back to
// the server's arena when the slot or job holding it is destroyed, after
// every buffer in it has been released (it is their first member).
struct Lease {
    tbk_server* srv = nullptr;
    Slab* slab = nullptr;
    size_t off = 0;
    Region region;
    Lease() = default;
    Lease(const Lease&) = delete;
    Lease& operator=(const Lease&) = delete;
    ~Lease();
};

// Slabs of the arena from the device: a refused hipMalloc leaves no error
// behind for the next HIP call.
struct HipSlabMemory : serve::SlabMemory {
    char* allocate(size_t bytes) override {
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        return static_cast<char*>(p);
    }
    void release(char* base) override { (void)hipFree(base); }
};

// bytes lost to the 256-byte alignment of the buffers carved from a span
constexpr int64_t kCarveSlack = 64 * 256;

struct Slot {
    Lease lease;
    Request* req = nullptr;
    tbk_matrix* M = nullptr;
    DevBuf blk[3], mu;
    int64_t k = 0, steps = 0;
    tbk::TileOrder order{};  // row tiles of one workgroup pass, chunks of 256 rows, transposed
    int32_t nwg = 8;         // workgroups of the slot in a launch (multiple of 8)
    int32_t pin = -1;        // row of the pinned moment staging
    bool cancelled = false;  // leaves without a result
    ~Slot() {
        if (M) tbk_matrix_destroy(M);
    }
};

struct KuboJob {
    Lease lease;
    Request* req = nullptr;
    tbk_matrix *H = nullptr, *vx = nullptr, *vy = nullptr;
    DevBuf kets, L, Rs, W0, W1, B0, B1, S, part, P, G, Gacc, g, e, z;
    hipEvent_t done = nullptr;
    int32_t pin = -1;
    ~KuboJob() {
        if (done) (void)hipEventDestroy(done);
        for (tbk_matrix* m : {H, vx, vy})
            if (m) tbk_matrix_destroy(m);
    }
};

struct ClassKey {
    int32_t prec, cplx, Rp;
    bool operator<(const ClassKey& o) const {
        return std::tie(prec, cplx, Rp) < std::tie(o.prec, o.cplx, o.Rp);
    }
};

// Per-class device arrays of a chunk (descriptors, tile map, partials) and
// their pinned staging, one staging set per in-flight chunk parity.
struct ClassState {
    DevBuf desc, tile_slot, partial;
    tbk::SlotDesc* stage_desc[2] = {nullptr, nullptr};
    int32_t* stage_tiles[2] = {nullptr, nullptr};
    size_t cap_desc[2] = {0, 0}, cap_tiles[2] = {0, 0};
};

struct Chunk {
    hipEvent_t ev = nullptr;
    std::vector<Slot*> finishing;
    std::vector<void*> graveyard;  // device allocations replaced while this chunk may use them
    int parity = 0;
};

}  // namespace

struct tbk_server {
    tbk_context* ctx = nullptr;
    hipStream_t kstream = nullptr;
    hipStream_t bstream = nullptr;  // spectral-range Lanczos of auto-bounds requests (prep thread)
    tbk_server_config cfg{};
    serve::Admission adm;            // budget, bytes in use, the arena and its retired slabs (under mu)
    HipSlabMemory slab_memory;

    // shared state, under mu
    std::mutex mu;
    std::condition_variable cv_prep, cv_work, done_cv;
    std::deque<Request*> queue;          // submitted, waiting for admission
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    std::vector<Slot*> ready;            // prepared slots, joining at the next chunk
    std::vector<Slot*> trash;            // finished slots, freed by the preparation thread
    std::vector<void*> trash_raw;        // replaced per-class arrays, freed by the preparation thread
    int64_t prep_gen = 0;                // bumped by every change the preparation thread acts on
    std::vector<int32_t> free_pin, free_kpin;
    int64_t next_ticket = 1;
    int32_t preparing = 0;               // admitted requests still being prepared
    bool stop = false, worker_done = false;
    std::string fatal;
    tbk_server_stats st{};

    std::thread worker, prep;
    // worker-only state
    std::vector<Slot*> slots;
    std::map<ClassKey, ClassState> classes;
    std::deque<Chunk> inflight;
    std::vector<hipEvent_t> free_events;
    int64_t chunk_seq = 0;
    std::vector<void*> pinned_extra;     // grown staging buffers, freed at destroy
    // preparation-thread-only state
    std::vector<KuboJob*> kjobs;
    double* pin_mu = nullptr;    // max_slots x max_moments doubles
    double* pin_kubo = nullptr;  // max_kubo x (2 max_kubo_N + 2 max_kubo_N^2) doubles

    static constexpr size_t kMaxInflight = 2;

    int64_t dos_footprint(const Request& r) const;
    int64_t kubo_footprint(const Request& r) const;
    void size_kubo(Request& r) const;
    void run();
    void run_prep();
    std::vector<std::pair<Request*, int32_t>> pick_admissions();
    Slot* start_dos(Request* r, int32_t pin);
    void start_kubo(Request* r, int32_t pin);
    void enqueue_chunk();
    void complete_chunk(Chunk& c);
    void poll_kubo(bool block);
    void finish(Request* r, int32_t code, const std::string& msg);
    void fail_all(const std::string& msg);
    hipEvent_t get_event();
    void grow(DevBuf& b, size_t bytes);
    void give_back(Slab* sl, size_t off, size_t bytes);
    void make_slabs(std::vector<std::pair<Request*, int32_t>>& adm);
    void lease_to(Lease& l, const Request* r);
    template <typename T>
    T* stage(T*& p, size_t& cap, size_t n);
};

namespace {

// acc = g (first) or acc + g, over n complex values
__global__ __launch_bounds__(256) void gamma_accumulate_kernel(double2* __restrict__ acc,
                                                               const double2* __restrict__ g, int64_t n,
                                                               int32_t first) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * 256) {
        const double2 v = g[i];
        if (first) {
            acc[i] = v;
        } else {
            double2 a = acc[i];
            a.x = a.x + v.x;
            a.y = a.y + v.y;
            acc[i] = a;
        }
    }
}

int64_t mat_bytes(const HostSparse& m, bool fp32_copy) {
    const int64_t vw = m.hdr.is_complex ? 2 : 1;
    int64_t b = (static_cast<int64_t>(m.hdr.rows) + 1) * 4 + m.hdr.nnz * 4 + m.hdr.nnz * 8 * vw;
    if (fp32_copy) b += m.hdr.nnz * 4 * vw;
    return b;
}

}  // namespace

int64_t tbk_server::dos_footprint(const Request& r) const {
    const int64_t rows = r.H.hdr.rows;
    const int64_t elem = r.precision == TBK_FP64 ? 16 : 8;
    int64_t vec = 3 * rows * r.Rp * elem;
    // the Lanczos vectors and partials of an auto-bounds request live before the blocks
    if (r.auto_bounds) vec = std::max<int64_t>(vec, 3 * rows * 16 + 64 * 1024);
    // assembly scratch of a lattice request (row counts, scan, phase table) lives before them too
    if (r.lat.present)
        vec = std::max<int64_t>(vec, (rows + 1) * 4 + static_cast<int64_t>(r.lat.phase.size()) * 8 + (1 << 20));
    return mat_bytes(r.H, r.precision == TBK_FP32) + vec + static_cast<int64_t>(r.n) * 8 + kCarveSlack;
}

// Device memory of a sigma_xy request whose random vectors run in groups of
// r.Gw: Gamma is a sum over vectors, so the two N-entry stacks only ever hold
// one group's columns, and the groups' Gammas are added up.
int64_t tbk_server::kubo_footprint(const Request& r) const {
    const int64_t rows = r.H.hdr.rows;
    const int64_t blk = rows * r.Gw * 16;
    const int64_t N = r.n;
    // fp64 stacks may hold one of row_chunks row ranges at a time
    const int64_t crows = (rows + r.row_chunks - 1) / r.row_chunks;
    const ContractionShape cs = contraction_plan(ctx, r.n, crows * r.Gw, r.precision);
    // pre-split stacks: half-size entries and three more rolling blocks
    const bool ps = kubo_presplit(r.precision, r.n, r.H.hdr.is_complex != 0, r.vx.hdr.is_complex != 0);
    // one stack entry: fp64 at the stack stride, words in whole word blocks
    const int64_t K = crows * r.Gw;
    const int64_t entry = ps ? ((K + tbk::kWordBlock - 1) / tbk::kWordBlock) * tbk::kWordBlock * 8 : stack_stride(K) * 16;
    const int64_t stacks = 2 * N * entry, rolling = ps ? 5 * blk : r.row_chunks > 1 ? 5 * blk : 2 * blk;
    // a lattice request also holds the assembled H while H~ is formed
    // (H~ has the size of a velocity operator: one more entry per row)
    int64_t mats = r.lat.present ? mat_bytes(r.H, false) + 3 * mat_bytes(r.vx, false)
                                 : mat_bytes(r.H, false) + mat_bytes(r.vx, false) + mat_bytes(r.vy, false);
    // the fp32 recursion's copies of H~'s and v_x's values (complex64)
    if (r.precision == TBK_GAMMA_BF16X3_FP32) mats += 2 * (r.vx.hdr.nnz + 256) * 8;
    return mats + blk /* kets */ +
           stacks + rolling + static_cast<int64_t>(ctx->num_cu) * 32 * 16 /* step partials */ +
           static_cast<int64_t>(cs.nchunks) * N * N * 16 + 2 * N * N * 16 /* Gamma, its sum */ + N * 8 * 2 + N * 16 +
           kCarveSlack;
}

// The widest group of random vectors (a power of two, at most the padded R)
// whose footprint fits the budget, set with its footprint; the choice
// depends on the request and the budget only, so a request's groups and
// therefore its result are the same whatever else the server runs.
void tbk_server::size_kubo(Request& r) const {
    r.Gw = r.Rp;
    r.row_chunks = 1;
    r.footprint = kubo_footprint(r);
    while (r.footprint > adm.budget && r.Gw > 1) {
        r.Gw /= 2;
        r.footprint = kubo_footprint(r);
    }
    // fp64 stacks of one vector still too large: hold them a row range at a
    // time (the chains then run once per range)
    const bool ps = kubo_presplit(r.precision, r.n, r.H.hdr.is_complex != 0, r.vx.hdr.is_complex != 0);
    while (r.footprint > adm.budget && !ps && r.row_chunks < 64) {
        ++r.row_chunks;
        r.footprint = kubo_footprint(r);
    }
}

hipEvent_t tbk_server::get_event() {
    if (!free_events.empty()) {
        hipEvent_t e = free_events.back();
        free_events.pop_back();
        return e;
    }
    hipEvent_t e;
    HIPC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
}

// grow a per-class device array; the old allocation may still be read by an
// in-flight chunk, so it is released when the newest in-flight chunk completes
void tbk_server::grow(DevBuf& b, size_t bytes) {
    if (bytes <= b.bytes) return;
    if (b.p && b.owned) {
        if (!inflight.empty()) {
            inflight.back().graveyard.push_back(b.p);
        } else {
            std::lock_guard<std::mutex> lk(mu);
            trash_raw.push_back(b.p);
            prep_gen++;
            cv_prep.notify_all();
        }
        b.p = nullptr;
        b.bytes = 0;
    }
    b.reserve(std::max(bytes, static_cast<size_t>(4096)));
}

// a span back to the arena (not under mu); queued requests may now fit
void tbk_server::give_back(Slab* sl, size_t off, size_t bytes) {
    std::lock_guard<std::mutex> lk(mu);
    adm.give(sl, off, bytes);
    prep_gen++;
    cv_prep.notify_all();
}

void tbk_server::lease_to(Lease& l, const Request* r) {
    if (!r->slab) raise(TBK_E_OOM, "the device refused a slab of " + std::to_string(r->new_slab / 1e9) + " GB");
    l.srv = this;
    l.slab = r->slab;
    l.off = r->arena_off;
    l.region.base = r->slab->base + r->arena_off;
    l.region.cap = static_cast<size_t>(r->footprint);
    l.region.off = 0;
}

template <typename T>
T* tbk_server::stage(T*& p, size_t& cap, size_t n) {
    if (n > cap) {
        // staging of the chunk that used this parity has completed; the old
        // buffer is kept until destroy (pinned frees may synchronise)
        if (p) pinned_extra.push_back(p);
        const size_t c = std::max<size_t>(n, 2 * cap);
        void* q = nullptr;
        HIPC(hipHostMalloc(&q, c * sizeof(T), hipHostMallocDefault));
        p = static_cast<T*>(q);
        cap = c;
    }
    return p;
}

void tbk_server::finish(Request* r, int32_t code, const std::string& msg) {
    std::lock_guard<std::mutex> lk(mu);
    r->t_done = now_s();
    r->state = code == TBK_OK ? TBK_REQ_DONE : TBK_REQ_FAILED;
    r->error_code = code;
    r->error = msg;
    adm.finished(*r);
    if (code == TBK_OK)
        st.completed++;
    else
        st.failed++;
    if (code == TBK_E_CANCELLED) st.cancelled++;
    if (r->kind == TBK_REQ_DOS) st.running_dos--;
    else st.running_kubo--;
    r->H.clear();
    r->vx.clear();
    r->vy.clear();
    r->lat.clear();
    done_cv.notify_all();
    prep_gen++;
    cv_prep.notify_all();
}

// First fit over the queue in arrival order (under mu,
// serve::Admission::pick); a request that has waited more than
// max_wait_bypass_ms stops later ones from passing it.  Each admitted
// request takes its pinned result row.
std::vector<std::pair<Request*, int32_t>> tbk_server::pick_admissions() {
    std::vector<int32_t> rows;
    auto pool = [&](const Request* r) -> std::vector<int32_t>& { return r->dos ? free_pin : free_kpin; };
    const std::vector<Request*> picked =
        adm.pick(queue, now_s(), [&](Request* r) { return !pool(r).empty(); },
                 [&](Request* r) {
                     rows.push_back(pool(r).back());
                     pool(r).pop_back();
                 });
    std::vector<std::pair<Request*, int32_t>> out;
    const double t = now_s();
    for (size_t i = 0; i < picked.size(); ++i) {
        Request* r = picked[i];
        r->state = TBK_REQ_RUNNING;
        r->t_admit = t;
        if (r->dos)
            st.running_dos++;
        else
            st.running_kubo++;
        out.emplace_back(r, rows[i]);
    }
    preparing += static_cast<int32_t>(out.size());
    st.queued = static_cast<int64_t>(queue.size());
    st.used_bytes = adm.used;
    return out;
}

// Outside mu: free the retired slabs, then allocate the new slab of every
// admission that needs one and place its span at the slab's start
// (serve::Admission::place).  An allocation the device refuses fails its
// request (lease_to).
void tbk_server::make_slabs(std::vector<std::pair<Request*, int32_t>>& picked) {
    std::vector<Request*> reqs_;
    for (auto& pr : picked) reqs_.push_back(pr.first);
    adm.place(reqs_, slab_memory, mu);
}

Slot* tbk_server::start_dos(Request* r, int32_t pin) {
    auto s = std::make_unique<Slot>();
    lease_to(s->lease, r);
    s->req = r;
    s->pin = pin;
    Region* rg = &s->lease.region;
    RegionScope placed(rg);  // the matrix arrays too
    if (r->lat.present) {
        assemble_matrix(ctx, &r->lat.desc, bstream, &s->M, nullptr);
    } else {
        const int rc = tbk_matrix_create(ctx, &r->H.hdr, &s->M);
        if (rc != TBK_OK) raise(rc, g_last_error);
    }
    if (r->auto_bounds) {
        // E_min, E_max from the Lanczos range estimate, contracted as
        // GreenFunction(N, HE, ref_HE, epsilon = 0.01) does (lattice.h:560-569)
        tbk_bounds bd{};
        spectral_bounds(ctx, s->M, kAutoIters, kAutoTol, r->seed ^ kAutoSeedMix, bstream, &bd);
        const double a = (bd.upper - bd.lower) / (2.0 - kAutoEpsilon), b = 0.5 * (bd.upper + bd.lower);
        if (!(a > 0.0) || !std::isfinite(a) || !std::isfinite(b))
            raise(TBK_E_INVALID, "the spectrum is a single point: give a, b");
        std::lock_guard<std::mutex> lk(mu);
        r->a = a;
        r->b = b;
    }
    if (r->precision == TBK_FP32) matrix_fp32(s->M);
    const int64_t rows = r->H.hdr.rows;
    const size_t elem = r->precision == TBK_FP64 ? 16 : 8;
    for (auto& b : s->blk) b.reserve(static_cast<size_t>(rows) * r->Rp * elem, rg);
    s->mu.reserve(static_cast<size_t>(r->n) * 8, rg);
    s->steps = (r->n + 1) / 2;
    slot_split(ctx->num_cu, 4, r->Rp / slot_cpl(cfg.cpl, r->precision, r->Rp), rows, s->M->bandwidth95, s->order,
               s->nwg);
    const size_t total = static_cast<size_t>(rows) * r->Rp;
    const unsigned grid = static_cast<unsigned>(std::min<size_t>((total + tbk::kBlock - 1) / tbk::kBlock, 65536));
    const double inv = 1.0 / std::sqrt(static_cast<double>(rows));
    if (r->precision == TBK_FP64)
        hipLaunchKernelGGL(tbk::random_phase_kernel<double2>, dim3(grid), dim3(tbk::kBlock), 0, ctx->stream,
                           s->blk[0].as<double2>(), static_cast<int32_t>(rows), r->R, r->Rp, r->seed, inv, r->R, 0);
    else
        hipLaunchKernelGGL(tbk::random_phase_kernel<float2>, dim3(grid), dim3(tbk::kBlock), 0, ctx->stream,
                           s->blk[0].as<float2>(), static_cast<int32_t>(rows), r->R, r->Rp, r->seed, inv, r->R, 0);
    HIPC(hipGetLastError());
    return s.release();
}

void tbk_server::start_kubo(Request* r, int32_t pin) {
    auto j = std::make_unique<KuboJob>();
    lease_to(j->lease, r);
    j->req = r;
    j->pin = pin;
    Region* rg = &j->lease.region;
    RegionScope placed(rg);  // the matrix arrays too
    double a = r->a, b = r->b;
    if (r->lat.present) {
        // assemble H, form H~ (the steps then run it as given) and the
        // velocity operators of H~ along x and y
        tbk_matrix* Hraw = nullptr;
        assemble_matrix(ctx, &r->lat.desc, bstream, &Hraw, nullptr);
        std::unique_ptr<tbk_matrix, int (*)(tbk_matrix*)> hold(Hraw, tbk_matrix_destroy);
        rescale_matrix(ctx, Hraw, a, b, bstream, &j->H);
        velocity_matrix(ctx, j->H, &r->lat.desc, 0, bstream, &j->vx);
        velocity_matrix(ctx, j->H, &r->lat.desc, 1, bstream, &j->vy);
        a = 1.0;
        b = 0.0;
    } else {
        int rc = tbk_matrix_create(ctx, &r->H.hdr, &j->H);
        if (rc == TBK_OK) rc = tbk_matrix_create(ctx, &r->vx.hdr, &j->vx);
        if (rc == TBK_OK) rc = tbk_matrix_create(ctx, &r->vy.hdr, &j->vy);
        if (rc != TBK_OK) raise(rc, g_last_error);
    }
    // the R random vectors run in groups of Gw columns (the whole block when
    // it fits): Rp below is a group's padded width
    const int rows = r->H.hdr.rows, Rp = r->Gw, N = r->n;
    const int ngroups = (r->R + Rp - 1) / Rp;
    // row ranges of the fp64 stacks (one unless the stacks of a group exceed the budget)
    const int nrc = r->row_chunks;
    const int crows = (rows + nrc - 1) / nrc;
    const int64_t K = static_cast<int64_t>(nrc > 1 ? crows : rows) * Rp;  // stack entry: complex elements
    const size_t blk = static_cast<size_t>(rows) * Rp * sizeof(double2);  // a full vector block
    const StepPlan pH = make_plan(ctx, j->H, TBK_FP64, Rp, rows, b, false);
    const StepPlan pX = make_plan(ctx, j->vx, TBK_FP64, Rp, rows, 0.0, false);
    const StepPlan pY = make_plan(ctx, j->vy, TBK_FP64, Rp, rows, 0.0, false);
    // the footprint counted the pre-split form whenever kubo_presplit holds;
    // the server's default tuning always takes the row kernel it needs
    const bool presplit = kubo_presplit(r->precision, N, j->H->is_complex, j->vx->is_complex);
    if (presplit && (pH.kc.kch != 0 || pX.kc.kch != 0)) raise(TBK_E_UNSUPPORTED, "pre-split stacks need the row kernel");
    if (presplit && nrc > 1) raise(TBK_E_UNSUPPORTED, "row ranges of pre-split stacks");
    const void* vxv = fused_velocity(j->H, j->vx, pH, bstream) ? j->vx->val64.p : nullptr;
    // pre-split with a fused v_x: both chains as one block of 2 Rp columns in
    // W0, W1 (two blocks each) and S -- the five blocks the footprint counts
    // TBK_GAMMA_BF16X3_FP32 with at least two columns per chain: the paired
    // chains run in fp32 (fp32 copies of H's and v_x's values, in the lease)
    StepPlan pP{};
    bool paired = false;
    if (presplit && vxv && 2 * Rp <= 256) {
        const bool f32 = r->precision == TBK_GAMMA_BF16X3_FP32 && Rp >= 2;
        if (f32) {
            matrix_fp32(j->vx);
            vxv = j->vx->val32.p;
        }
        pP = make_plan(ctx, j->H, f32 ? TBK_FP32 : TBK_FP64, 2 * Rp, rows, b, true, tbk_tuning{});
        paired = pP.kc.kch == 0;
        if (!paired) vxv = j->vx->val64.p;
    }
    // fp64 stacks at a stride of ldk elements; word stacks k-blocked
    const int64_t ldk = stack_stride(K);
    const int64_t kw = ((K + tbk::kWordBlock - 1) / tbk::kWordBlock) * tbk::kWordBlock;
    const size_t entry = presplit ? static_cast<size_t>(kw) * sizeof(uint2) : static_cast<size_t>(ldk) * sizeof(double2);
    j->kets.reserve(blk, rg);
    j->L.reserve(static_cast<size_t>(N) * entry, rg);
    j->Rs.reserve(static_cast<size_t>(N) * entry, rg);
    j->W0.reserve(paired ? 2 * blk : blk, rg);
    j->W1.reserve(paired ? 2 * blk : blk, rg);
    if (nrc > 1) {
        j->B0.reserve(blk, rg);
        j->B1.reserve(blk, rg);
        j->S.reserve(blk, rg);
    }
    if (presplit) {
        if (!paired) {
            j->B0.reserve(blk, rg);
            j->B1.reserve(blk, rg);
        }
        j->S.reserve(blk, rg);
    }
    j->part.reserve(static_cast<size_t>(std::max({pH.nb, pX.nb, pY.nb, paired ? pP.nb : int64_t{0}})) *
                    sizeof(double2), rg);
    const ContractionShape cs = contraction_plan(ctx, N, K, r->precision);
    const int64_t NN = static_cast<int64_t>(N) * N;
    j->P.reserve(static_cast<size_t>(cs.nchunks) * NN * sizeof(double2), rg);
    j->G.reserve(static_cast<size_t>(NN) * sizeof(double2), rg);
    if (ngroups * nrc > 1) j->Gacc.reserve(static_cast<size_t>(NN) * sizeof(double2), rg);
    j->g.reserve(static_cast<size_t>(N) * sizeof(double), rg);
    j->e.reserve(static_cast<size_t>(N) * sizeof(double), rg);
    j->z.reserve(static_cast<size_t>(N) * sizeof(double2), rg);
    const int64_t Ncap = cfg.max_kubo_moments;
    double* pz = pin_kubo + static_cast<size_t>(j->pin) * static_cast<size_t>(2 * Ncap + 2 * Ncap * Ncap);
    double* pg = pz + 2 * N;  // Gamma region; the weights and nodes are staged there first
    for (int i = 0; i < N; ++i) {
        pg[i] = r->g.empty() ? 1.0 : r->g[i];
        pg[N + i] = std::cos((i + 0.5) / N * M_PI);
    }
    hipStream_t st_ = kstream;
    HIPC(hipMemcpyAsync(j->g.p, pg, static_cast<size_t>(N) * 8, hipMemcpyHostToDevice, st_));
    HIPC(hipMemcpyAsync(j->e.p, pg + N, static_cast<size_t>(N) * 8, hipMemcpyHostToDevice, st_));
    const double inv = 1.0 / std::sqrt(static_cast<double>(rows));
    const int64_t kets_elems = static_cast<int64_t>(rows) * Rp;
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>((kets_elems + tbk::kBlock - 1) / tbk::kBlock, 65536));
    const int nparts = ngroups * nrc;
    for (int part = 0; part < nparts; ++part) {
        const int grp = part / nrc, rc = part % nrc;
        if (rc == 0) {
            // columns [grp Gw, grp Gw + cols) of the request's R-wide random block
            const int col0 = grp * Rp, cols = std::min(Rp, r->R - col0);
            hipLaunchKernelGGL(tbk::random_phase_kernel<double2>, dim3(grid), dim3(tbk::kBlock), 0, st_,
                               j->kets.as<double2>(), rows, cols, Rp, r->seed, inv, r->R, col0);
            HIPC(hipGetLastError());
        }
        if (nrc > 1) {
            // rows [r0, r1) of every stack entry; the last range may be shorter
            const int r0 = rc * crows, r1 = std::min(rows, r0 + crows);
            const int64_t Kc = static_cast<int64_t>(r1 - r0) * Rp;
            enqueue_kubo_stacks_rows(pH, pX, pY, a, b, j->kets.p, j->L.p, j->Rs.p, j->B0.p, j->B1.p, j->W0.p,
                                     j->W1.p, j->S.p, rows, Rp, r0, r1, ldk, N, j->part.as<double2>(), st_);
            const ContractionShape csr = contraction_plan(ctx, N, Kc, r->precision);
            enqueue_contraction(csr, j->L.as<const double2>(), j->Rs.as<const double2>(), N, Kc, ldk, r->precision,
                                j->P.as<double2>(), j->G.as<double2>(), st_);
        } else if (paired) {
            enqueue_kubo_stacks_paired(pP, pY, a, b, j->kets.p, j->L.as<uint2>(), j->Rs.as<uint2>(), j->W0.p,
                                       j->W1.p, j->S.p, rows, Rp, N, j->part.as<double2>(), st_, vxv);
            enqueue_contraction_presplit(cs, j->L.as<const uint2>(), j->Rs.as<const uint2>(), N, K,
                                         j->P.as<double2>(), j->G.as<double2>(), st_);
        } else if (presplit) {
            enqueue_kubo_stacks_presplit(pH, pX, pY, a, b, j->kets.p, j->L.as<uint2>(), j->Rs.as<uint2>(), j->B0.p,
                                         j->B1.p, j->S.p, j->W0.p, j->W1.p, K, N, j->part.as<double2>(), st_, vxv);
            enqueue_contraction_presplit(cs, j->L.as<const uint2>(), j->Rs.as<const uint2>(), N, K,
                                         j->P.as<double2>(), j->G.as<double2>(), st_);
        } else {
            enqueue_kubo_stacks(pH, pX, pY, a, b, j->kets.p, j->L.p, j->Rs.p, j->W0.p, j->W1.p, blk, ldk, N,
                                j->part.as<double2>(), st_, vxv);
            enqueue_contraction(cs, j->L.as<const double2>(), j->Rs.as<const double2>(), N, K, ldk, r->precision,
                                j->P.as<double2>(), j->G.as<double2>(), st_);
        }
        if (nparts > 1) {
            // the parts' Gammas (vector groups x row ranges) added in order
            const unsigned ag = static_cast<unsigned>(std::min<int64_t>((NN + 255) / 256, 4096));
            hipLaunchKernelGGL(gamma_accumulate_kernel, dim3(ag), dim3(256), 0, st_, j->Gacc.as<double2>(),
                               j->G.as<const double2>(), NN, part == 0 ? 1 : 0);
            HIPC(hipGetLastError());
        }
    }
    double2* Gsum = nparts > 1 ? j->Gacc.as<double2>() : j->G.as<double2>();
    tbk::launch_gamma_weight(Gsum, j->g.as<const double>(), N, st_);
    HIPC(hipGetLastError());
    tbk::launch_kubo_zeta(Gsum, N, j->e.as<const double>(), N, 1.0, j->z.as<double2>(), st_);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(pz, j->z.p, static_cast<size_t>(N) * 16, hipMemcpyDeviceToHost, st_));
    HIPC(hipMemcpyAsync(pg, Gsum, static_cast<size_t>(NN) * 16, hipMemcpyDeviceToHost, st_));
    HIPC(hipEventCreateWithFlags(&j->done, hipEventDisableTiming));
    HIPC(hipEventRecord(j->done, st_));
    kjobs.push_back(j.release());
}

void tbk_server::poll_kubo(bool block) {
    for (size_t i = 0; i < kjobs.size();) {
        KuboJob* j = kjobs[i];
        hipError_t q = block ? hipEventSynchronize(j->done) : hipEventQuery(j->done);
        if (q == hipErrorNotReady) {
            ++i;
            continue;
        }
        HIPC(q);
        Request* r = j->req;
        const int64_t N = r->n, Ncap = cfg.max_kubo_moments;
        const double* pz = pin_kubo + static_cast<size_t>(j->pin) * static_cast<size_t>(2 * Ncap + 2 * Ncap * Ncap);
        const bool dropped = r->cancel.load();
        if (!dropped) r->result.assign(pz, pz + 2 * N + 2 * N * N);
        const int32_t pin = j->pin;
        kjobs.erase(kjobs.begin() + static_cast<std::ptrdiff_t>(i));
        delete j;
        {
            std::lock_guard<std::mutex> lk(mu);
            free_kpin.push_back(pin);
            st.kubo_done++;
        }
        if (dropped)
            finish(r, TBK_E_CANCELLED, "cancelled");
        else
            finish(r, TBK_OK, "");
    }
}

void tbk_server::enqueue_chunk() {
    const int32_t S = cfg.chunk_steps;
    Chunk c;
    c.parity = static_cast<int>(chunk_seq & 1);
    // cancelled slots leave without running: freed once this chunk (and so
    // the one before it, which may still use them) has completed
    for (size_t i = 0; i < slots.size();) {
        if (slots[i]->req->cancel.load()) {
            slots[i]->cancelled = true;
            c.finishing.push_back(slots[i]);
            slots.erase(slots.begin() + static_cast<std::ptrdiff_t>(i));
        } else {
            ++i;
        }
    }
    hipStream_t st_ = ctx->stream;
    // group the slots by class
    std::map<ClassKey, std::vector<Slot*>> groups;
    for (Slot* s : slots) {
        const Request* r = s->req;
        groups[ClassKey{r->precision, r->H.hdr.is_complex ? 1 : 0, r->Rp}].push_back(s);
    }
    for (auto& kv : groups) {
        const ClassKey& key = kv.first;
        std::vector<Slot*>& v = kv.second;
        ClassState& cs = classes[key];
        int64_t nwg = 0;
        for (Slot* x : v) nwg += x->nwg;
        tbk::SlotDesc* hd = stage(cs.stage_desc[c.parity], cs.cap_desc[c.parity], v.size());
        int32_t* ht = stage(cs.stage_tiles[c.parity], cs.cap_tiles[c.parity], static_cast<size_t>(nwg));
        int64_t w0 = 0;
        for (size_t i = 0; i < v.size(); ++i) {
            Slot* x = v[i];
            const Request* r = x->req;
            tbk::SlotDesc& d = hd[i];
            std::memset(&d, 0, sizeof(d));
            d.rp = x->M->rp.as<const int32_t>();
            d.ci = x->M->ci.as<const int32_t>();
            d.hv = r->precision == TBK_FP64 ? x->M->val64.p : x->M->val32.p;
            for (int q = 0; q < 3; ++q) d.blk[q] = x->blk[q].p;
            d.coef[0] = 1.0 / r->a;
            d.coef[1] = -r->b / r->a;
            d.coef[2] = 2.0 / r->a;
            d.coef[3] = -2.0 * r->b / r->a;
            d.mu = x->mu.as<double>();
            d.k0 = x->k;
            d.steps = x->steps;
            d.wg0 = w0;
            d.nwg = x->nwg;
            d.rows = r->H.hdr.rows;
            d.n = r->n;
            d.order = x->order;
            for (int32_t t = 0; t < x->nwg; ++t) ht[w0 + t] = static_cast<int32_t>(i);
            w0 += x->nwg;
        }
        grow(cs.desc, v.size() * sizeof(tbk::SlotDesc));
        grow(cs.tile_slot, static_cast<size_t>(nwg) * sizeof(int32_t));
        grow(cs.partial, static_cast<size_t>(nwg) * sizeof(double2));
        HIPC(hipMemcpyAsync(cs.desc.p, hd, v.size() * sizeof(tbk::SlotDesc), hipMemcpyHostToDevice, st_));
        HIPC(hipMemcpyAsync(cs.tile_slot.p, ht, static_cast<size_t>(nwg) * sizeof(int32_t), hipMemcpyHostToDevice,
                            st_));
        tbk::SlotArgs A{};
        A.slots = cs.desc.as<const tbk::SlotDesc>();
        A.wg_slot = cs.tile_slot.as<const int32_t>();
        A.Rp = key.Rp;
        A.opts = ctx->tuning.stream_nt >= 0 ? tbk::kOptStream : 0;
        A.partial = cs.partial.as<double2>();
        const int cpl = slot_cpl(cfg.cpl, key.prec, key.Rp);
        const dim3 grid(static_cast<unsigned>(nwg));
        for (int32_t s = 0; s < S; ++s) {
            A.s = s;
            bool ok;
            if (key.prec == TBK_FP64)
                ok = key.cplx ? tbk::launch_slots_f64_cplx(key.Rp, cpl, grid, st_, A)
                              : tbk::launch_slots_f64_real(key.Rp, cpl, grid, st_, A);
            else
                ok = key.cplx ? tbk::launch_slots_f32_cplx(key.Rp, cpl, grid, st_, A)
                              : tbk::launch_slots_f32_real(key.Rp, cpl, grid, st_, A);
            if (!ok) raise(TBK_E_UNSUPPORTED, "no slot kernel for this width");
            HIPC(hipGetLastError());
            hipLaunchKernelGGL(tbk::slots_moments_kernel, dim3(static_cast<unsigned>(v.size())), dim3(tbk::kBlock), 0, st_,
                               cs.desc.as<const tbk::SlotDesc>(), s, cs.partial.as<const double2>());
            HIPC(hipGetLastError());
        }
    }
    // slots finishing in this chunk: moments to pinned staging, then leave
    for (size_t i = 0; i < slots.size();) {
        Slot* s = slots[i];
        s->k += S;
        if (s->k >= s->steps) {
            double* dst = pin_mu + static_cast<size_t>(s->pin) * static_cast<size_t>(cfg.max_moments);
            HIPC(hipMemcpyAsync(dst, s->mu.p, static_cast<size_t>(s->req->n) * 8, hipMemcpyDeviceToHost, st_));
            c.finishing.push_back(s);
            slots.erase(slots.begin() + static_cast<std::ptrdiff_t>(i));
        } else {
            ++i;
        }
    }
    c.ev = get_event();
    HIPC(hipEventRecord(c.ev, st_));
    inflight.push_back(std::move(c));
    ++chunk_seq;
    std::lock_guard<std::mutex> lk(mu);
    st.chunks++;
}

void tbk_server::complete_chunk(Chunk& c) {
    HIPC(hipEventSynchronize(c.ev));
    free_events.push_back(c.ev);
    for (Slot* s : c.finishing) {
        Request* r = s->req;
        if (s->cancelled) {
            finish(r, TBK_E_CANCELLED, "cancelled");
            continue;
        }
        const double* src = pin_mu + static_cast<size_t>(s->pin) * static_cast<size_t>(cfg.max_moments);
        r->result.assign(src, src + r->n);
        finish(r, TBK_OK, "");
    }
    {
        // device memory goes back through the preparation thread, whose frees
        // may wait for the device
        std::lock_guard<std::mutex> lk(mu);
        for (Slot* s : c.finishing) {
            free_pin.push_back(s->pin);
            trash.push_back(s);
        }
        for (void* p : c.graveyard) trash_raw.push_back(p);
        prep_gen++;
        cv_prep.notify_all();
    }
}

void tbk_server::fail_all(const std::string& msg) {
    std::lock_guard<std::mutex> lk(mu);
    if (fatal.empty()) fatal = msg;
    for (auto& kv : reqs) {
        Request* r = kv.second.get();
        if (r->state == TBK_REQ_QUEUED || r->state == TBK_REQ_RUNNING) {
            r->state = TBK_REQ_FAILED;
            r->error_code = TBK_E_DEVICE;
            r->error = fatal;
            r->t_done = now_s();
            st.failed++;
        }
    }
    queue.clear();
    done_cv.notify_all();
    prep_gen++;
    cv_prep.notify_all();
    cv_work.notify_all();
}

// Worker: batched chunks of the DoS slots, at most kMaxInflight in flight.
void tbk_server::run() {
    try {
        set_device(ctx);
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    for (Slot* x : ready) slots.push_back(x);
                    ready.clear();
                    if (!slots.empty() || !inflight.empty()) break;
                    if (!fatal.empty() || (stop && queue.empty() && preparing == 0)) {
                        worker_done = true;
                        prep_gen++;
                        cv_prep.notify_all();
                        return;
                    }
                    cv_work.wait(lk);
                }
            }
            if (!slots.empty() && inflight.size() < kMaxInflight) enqueue_chunk();
            if (inflight.size() >= kMaxInflight || (slots.empty() && !inflight.empty())) {
                complete_chunk(inflight.front());
                inflight.pop_front();
            }
        }
    } catch (const Error& e) {
        fail_all(e.msg);
    } catch (const std::exception& e) {
        fail_all(e.what());
    }
    std::lock_guard<std::mutex> lk(mu);
    worker_done = true;
    prep_gen++;
    cv_prep.notify_all();
}

// Preparation thread: admission, uploads and kets of new slots, the whole
// enqueue of sigma_xy jobs and their completion, and every device free.
void tbk_server::run_prep() {
    int64_t seen_gen = -1;
    try {
        set_device(ctx);
        for (;;) {
            std::vector<std::pair<Request*, int32_t>> adm;
            std::vector<Slot*> dead;
            std::vector<void*> dead_raw;
            {
                std::unique_lock<std::mutex> lk(mu);
                // act on news (submission, completion, freed memory, stop);
                // poll the sigma_xy jobs every 200 us while any is running;
                // once stopping with the worker gone, run to the exit test
                auto news = [&] { return prep_gen != seen_gen || (stop && worker_done); };
                if (kjobs.empty())
                    cv_prep.wait(lk, news);
                else
                    cv_prep.wait_for(lk, std::chrono::microseconds(200), news);
                seen_gen = prep_gen;
                if (!fatal.empty()) break;
                dead.swap(trash);
                dead_raw.swap(trash_raw);
                adm = pick_admissions();
                if (stop && worker_done && queue.empty() && kjobs.empty() && adm.empty() && dead.empty() &&
                    dead_raw.empty())
                    break;
            }
            for (Slot* x : dead) delete x;
            for (void* p : dead_raw) (void)hipFree(p);
            make_slabs(adm);
            for (auto& pr : adm) {
                Request* r = pr.first;
                try {
                    if (r->kind == TBK_REQ_DOS) {
                        Slot* x = start_dos(r, pr.second);
                        std::lock_guard<std::mutex> lk(mu);
                        ready.push_back(x);
                        preparing--;
                        cv_work.notify_all();
                    } else {
                        start_kubo(r, pr.second);
                        std::lock_guard<std::mutex> lk(mu);
                        preparing--;
                        cv_work.notify_all();
                    }
                } catch (const Error& e) {
                    {
                        std::lock_guard<std::mutex> lk(mu);
                        (r->kind == TBK_REQ_DOS ? free_pin : free_kpin).push_back(pr.second);
                        preparing--;
                        cv_work.notify_all();
                    }
                    finish(r, e.code, e.msg);
                }
            }
            poll_kubo(false);
        }
    } catch (const Error& e) {
        fail_all(e.msg);
    } catch (const std::exception& e) {
        fail_all(e.what());
    }
}

namespace {

void check_values(const HostSparse& m, const char* name) {
    for (double v : m.values)
        if (!std::isfinite(v)) raise(TBK_E_INVALID, std::string("non-finite value in ") + name);
}

void check_lattice(const HostLattice& l) {
    auto finite = [](const std::vector<double>& v) {
        for (double x : v)
            if (!std::isfinite(x)) return false;
        return true;
    };
    if (!finite(l.t) || !finite(l.phase) || !finite(l.apos) || !std::isfinite(l.desc.disorder_W))
        raise(TBK_E_INVALID, "non-finite hopping, phase, position or disorder value in the lattice");
    for (double b : l.desc.bases)
        if (!std::isfinite(b)) raise(TBK_E_INVALID, "non-finite lattice basis");
}

Lease::~Lease() {
    if (srv) srv->give_back(slab, off, region.cap);
}

void fill_info(const Request& r, tbk_request_info* info) {
    if (!info) return;
    info->state = r.state;
    info->error_code = r.error_code;
    const double t = now_s();
    info->queue_ms = 1e3 * ((r.t_admit > 0 ? r.t_admit : t) - r.t_submit);
    info->run_ms = r.t_admit > 0 ? 1e3 * ((r.t_done > 0 ? r.t_done : t) - r.t_admit) : 0.0;
    info->latency_ms = 1e3 * ((r.t_done > 0 ? r.t_done : t) - r.t_submit);
    info->footprint_bytes = r.footprint;
    info->result_doubles = static_cast<int64_t>(r.result.size());
    info->a = r.a;
    info->b = r.b;
    const bool kubo = r.kind == TBK_REQ_KUBO && r.Gw > 0;
    info->vector_groups = kubo ? (r.R + r.Gw - 1) / r.Gw : 1;
    info->row_ranges = kubo ? r.row_chunks : 1;
}

}  // namespace

extern "C" {

int tbk_server_create(int32_t device, const tbk_server_config* cfg, tbk_server** out) {
    return guarded([&] {
        if (!out) raise(TBK_E_INVALID, "null argument");
        auto s = std::make_unique<tbk_server>();
        if (cfg) s->cfg = *cfg;
        tbk_server_config& c = s->cfg;
        if (c.max_slots <= 0) c.max_slots = 512;
        if (c.max_kubo <= 0) c.max_kubo = 1;
        if (c.chunk_steps <= 0) c.chunk_steps = 64;
        if (c.max_moments <= 0) c.max_moments = 4096;
        if (c.max_kubo_moments <= 0) c.max_kubo_moments = 1024;
        if (c.max_wait_bypass_ms <= 0) c.max_wait_bypass_ms = 1000.0;
        if (c.cpl != 0 && c.cpl != 1 && c.cpl != 2 && c.cpl != 4) raise(TBK_E_INVALID, "cpl must be 0, 1, 2 or 4");
        int rc = tbk_context_create(device, &s->ctx);
        if (rc != TBK_OK) raise(rc, g_last_error);
        set_device(s->ctx);
        HIPC(hipStreamCreateWithFlags(&s->kstream, hipStreamNonBlocking));
        HIPC(hipStreamCreateWithFlags(&s->bstream, hipStreamNonBlocking));
        size_t free_b = 0, total_b = 0;
        HIPC(hipMemGetInfo(&free_b, &total_b));
        s->adm.budget =
            c.hbm_budget_bytes > 0 ? c.hbm_budget_bytes : static_cast<int64_t>(0.85 * static_cast<double>(free_b));
        s->adm.max_wait_bypass_ms = c.max_wait_bypass_ms;
        if (c.lazy_arena != 0 && c.lazy_arena != 1) raise(TBK_E_INVALID, "lazy_arena must be 0 or 1");
        // the whole budget as one slab, now (else slabs as requests need them)
        if (!c.lazy_arena && !s->adm.reserve_whole(s->slab_memory))
            raise(TBK_E_OOM, "could not allocate the server's budget of " + std::to_string(s->adm.budget / 1e9) +
                                 " GB (" + std::to_string(static_cast<double>(free_b) / 1e9) + " GB free)");

        void* p = nullptr;
        HIPC(hipHostMalloc(&p, static_cast<size_t>(c.max_slots) * c.max_moments * sizeof(double), hipHostMallocDefault));
        s->pin_mu = static_cast<double*>(p);
        for (int32_t i = c.max_slots - 1; i >= 0; --i) s->free_pin.push_back(i);
        const size_t kb = static_cast<size_t>(2 * c.max_kubo_moments + 2 * static_cast<int64_t>(c.max_kubo_moments) *
                                                                             c.max_kubo_moments);
        HIPC(hipHostMalloc(&p, static_cast<size_t>(c.max_kubo) * kb * sizeof(double), hipHostMallocDefault));
        s->pin_kubo = static_cast<double*>(p);
        for (int32_t i = c.max_kubo - 1; i >= 0; --i) s->free_kpin.push_back(i);
        s->st.budget_bytes = s->adm.budget;
        tbk_server* raw = s.get();
        s->worker = std::thread([raw] { raw->run(); });
        s->prep = std::thread([raw] { raw->run_prep(); });
        *out = s.release();
    });
}

int tbk_server_submit(tbk_server* s, const tbk_request* q, int64_t* ticket) {
    return guarded([&] {
        if (!s || !q || !ticket) raise(TBK_E_INVALID, "null argument");
        if (q->kind != TBK_REQ_DOS && q->kind != TBK_REQ_KUBO) raise(TBK_E_INVALID, "unknown request kind");
        if (!q->H && !q->lattice) raise(TBK_E_INVALID, "request without a Hamiltonian");
        if (q->R < 1) raise(TBK_E_INVALID, "R must be positive");
        if (!std::isfinite(q->a) || !std::isfinite(q->b)) raise(TBK_E_INVALID, "invalid rescaling a, b");
        if (q->a == 0.0 && q->kind != TBK_REQ_DOS) raise(TBK_E_INVALID, "sigma_xy requests need a, b");
        auto r = std::make_unique<Request>();
        r->kind = q->kind;
        r->dos = q->kind == TBK_REQ_DOS;
        r->precision = q->precision;
        r->n = q->n_moments;
        r->R = q->R;
        r->Rp = pow2_at_least(q->R);
        r->seed = q->seed;
        r->a = q->a;
        r->b = q->b;
        r->auto_bounds = q->kind == TBK_REQ_DOS && q->a == 0.0;
        if (r->Rp > 256) raise(TBK_E_UNSUPPORTED, "R above 256 is not supported");
        if (q->kind == TBK_REQ_DOS) {
            if (q->precision != TBK_FP64 && q->precision != TBK_FP32) raise(TBK_E_INVALID, "unknown precision");
            if (q->n_moments < 2) raise(TBK_E_INVALID, "n_moments must be at least 2");
            if (q->n_moments > s->cfg.max_moments) raise(TBK_E_UNSUPPORTED, "more moments than the server allows");
        } else {
            if (q->precision != TBK_GAMMA_FP64 && q->precision != TBK_GAMMA_BF16X3 &&
                q->precision != TBK_GAMMA_BF16X3_FP32)
                raise(TBK_E_INVALID, "unknown contraction precision");
            if (q->H && (!q->vx || !q->vy)) raise(TBK_E_INVALID, "sigma_xy request without velocity operators");
            if (!q->H && (q->vx || q->vy))
                raise(TBK_E_INVALID, "a lattice sigma_xy request forms its velocity operators itself");
            if (!q->H && (!q->lattice->atom_pos || q->lattice->dim < 2))
                raise(TBK_E_INVALID, "a lattice sigma_xy request needs atom positions and two axes");
            if (q->n_moments < 1) raise(TBK_E_INVALID, "n_moments must be at least 1");
            if (q->n_moments > s->cfg.max_kubo_moments) raise(TBK_E_UNSUPPORTED, "more moments than the server allows");
            if (tbk::kubo_zeta_energies_per_group(q->n_moments) == 0) raise(TBK_E_UNSUPPORTED, "N too large for zeta");
        }
        if (q->H) {
            r->H.copy_from(q->H);
            check_values(r->H, "H");
            std::vector<int32_t> rp, ci;
            std::vector<double> val;
            build_csr(&r->H.hdr, rp, ci, val);
            if (!csr_hermitian_within(r->H.hdr.rows, rp, ci, val, r->H.hdr.is_complex != 0, 1e-12))
                raise(TBK_E_INVALID, "H is not Hermitian (M^H differs from M beyond 1e-12 of its largest value)");
        } else {
            // a lattice: the device assembles it at admission; H.hdr carries
            // the size, the value type and a bound of the stored entries
            r->lat.copy_from(q->lattice);
            check_lattice(r->lat);
            const int64_t rows = lattice_rows(&r->lat.desc);
            if (rows > INT32_MAX) raise(TBK_E_UNSUPPORTED, "more than 2^31 - 1 sites");
            r->H.hdr.rows = r->H.hdr.cols = static_cast<int32_t>(rows);
            r->H.hdr.nnz = lattice_nnz_bound(&r->lat.desc);
            r->H.hdr.is_complex = q->lattice->is_complex;
        }
        if (q->kind == TBK_REQ_KUBO && !q->H) {
            // H~ = (H - b)/a over the union pattern and v_x, v_y on its
            // pattern, all formed on the device at admission
            r->vx.hdr = r->vy.hdr = r->H.hdr;
            r->vx.hdr.nnz = r->vy.hdr.nnz = r->H.hdr.nnz + r->H.hdr.rows;
            if (q->g) r->g.assign(q->g, q->g + q->n_moments);
            s->size_kubo(*r);
        } else if (q->kind == TBK_REQ_KUBO) {
            r->vx.copy_from(q->vx);
            r->vy.copy_from(q->vy);
            check_values(r->vx, "v_x");
            check_values(r->vy, "v_y");
            if (r->vx.hdr.rows != r->H.hdr.rows || r->vy.hdr.rows != r->H.hdr.rows)
                raise(TBK_E_INVALID, "velocity operators and H differ in size");
            if (q->g) r->g.assign(q->g, q->g + q->n_moments);
            s->size_kubo(*r);
        } else {
            r->footprint = s->dos_footprint(*r);
        }
        if (!s->adm.admissible(r->footprint))
            raise(TBK_E_OOM, "request needs more device memory than the server budget");
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) raise(TBK_E_INVALID, "server is shutting down");
        if (!s->fatal.empty()) raise(TBK_E_DEVICE, s->fatal);
        r->ticket = s->next_ticket++;
        r->t_submit = now_s();
        *ticket = r->ticket;
        s->queue.push_back(r.get());
        s->reqs[r->ticket] = std::move(r);
        s->st.submitted++;
        s->st.queued = static_cast<int64_t>(s->queue.size());
        s->prep_gen++;
        s->cv_prep.notify_all();
    });
}

int tbk_server_info(tbk_server* s, int64_t ticket, tbk_request_info* info) {
    return guarded([&] {
        if (!s || !info) raise(TBK_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) raise(TBK_E_INVALID, "unknown ticket");
        fill_info(*it->second, info);
    });
}

int tbk_server_wait(tbk_server* s, int64_t ticket, double timeout_s, tbk_request_info* info) {
    return guarded([&] {
        if (!s) raise(TBK_E_INVALID, "null argument");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) raise(TBK_E_INVALID, "unknown ticket");
        Request* r = it->second.get();
        auto pred = [&] { return r->state == TBK_REQ_DONE || r->state == TBK_REQ_FAILED; };
        if (timeout_s < 0)
            s->done_cv.wait(lk, pred);
        else
            s->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
        fill_info(*r, info);
    });
}

int tbk_server_take(tbk_server* s, int64_t ticket, double* out, int64_t capacity, tbk_request_info* info) {
    return guarded([&] {
        if (!s) raise(TBK_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) raise(TBK_E_INVALID, "unknown ticket");
        Request* r = it->second.get();
        if (r->state == TBK_REQ_QUEUED || r->state == TBK_REQ_RUNNING) raise(TBK_E_INVALID, "request not finished");
        fill_info(*r, info);
        if (r->state == TBK_REQ_FAILED) {
            const int code = r->error_code;
            const std::string msg = r->error;
            s->reqs.erase(it);
            raise(code ? code : TBK_E_DEVICE, msg);
        }
        if (static_cast<int64_t>(r->result.size()) > capacity) raise(TBK_E_INVALID, "output buffer too small");
        if (out && !r->result.empty()) std::memcpy(out, r->result.data(), r->result.size() * sizeof(double));
        s->reqs.erase(it);
    });
}

int tbk_server_cancel(tbk_server* s, int64_t ticket) {
    return guarded([&] {
        if (!s) raise(TBK_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) raise(TBK_E_INVALID, "unknown ticket");
        Request* r = it->second.get();
        if (r->state == TBK_REQ_QUEUED) {
            auto q = std::find(s->queue.begin(), s->queue.end(), r);
            if (q != s->queue.end()) s->queue.erase(q);
            r->state = TBK_REQ_FAILED;
            r->error_code = TBK_E_CANCELLED;
            r->error = "cancelled";
            r->t_done = now_s();
            r->H.clear();
            r->vx.clear();
            r->vy.clear();
            r->lat.clear();
            s->st.failed++;
            s->st.cancelled++;
            s->st.queued = static_cast<int64_t>(s->queue.size());
            s->done_cv.notify_all();
            s->prep_gen++;
            s->cv_prep.notify_all();
            return;
        }
        if (r->state == TBK_REQ_RUNNING) {
            r->cancel.store(true);  // the worker or the preparation thread ends it
            return;
        }
        raise(TBK_E_INVALID, "request already finished");
    });
}

int tbk_server_get_stats(tbk_server* s, tbk_server_stats* out) {
    return guarded([&] {
        if (!s || !out) raise(TBK_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        *out = s->st;
        out->arena_bytes = static_cast<int64_t>(s->adm.arena.held);
        out->arena_slabs = static_cast<int64_t>(s->adm.arena.slabs.size());
        out->used_bytes = s->adm.used;
        out->queued = static_cast<int64_t>(s->queue.size());
    });
}

int tbk_server_destroy(tbk_server* s) {
    return guarded([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
            s->prep_gen++;
            s->cv_prep.notify_all();
            s->cv_work.notify_all();
        }
        if (s->worker.joinable()) s->worker.join();
        if (s->prep.joinable()) s->prep.join();
        set_device(s->ctx);
        (void)hipDeviceSynchronize();
        for (Slot* x : s->slots) delete x;
        for (Slot* x : s->ready) delete x;
        for (Slot* x : s->trash) delete x;
        for (void* p : s->trash_raw) (void)hipFree(p);
        for (KuboJob* x : s->kjobs) delete x;
        for (Chunk& c : s->inflight) {
            for (Slot* x : c.finishing) delete x;
            for (void* p : c.graveyard) (void)hipFree(p);
        }
        for (auto& kv : s->classes)
            for (int q = 0; q < 2; ++q) {
                if (kv.second.stage_desc[q]) (void)hipHostFree(kv.second.stage_desc[q]);
                if (kv.second.stage_tiles[q]) (void)hipHostFree(kv.second.stage_tiles[q]);
            }
        s->classes.clear();
        for (void* p : s->pinned_extra) (void)hipHostFree(p);
        for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
        if (s->pin_mu) (void)hipHostFree(s->pin_mu);
        if (s->pin_kubo) (void)hipHostFree(s->pin_kubo);
        s->adm.release_all(s->slab_memory);
        if (s->kstream) (void)hipStreamDestroy(s->kstream);
        if (s->bstream) (void)hipStreamDestroy(s->bstream);
        tbk_context* ctx = s->ctx;
        delete s;
        tbk_context_destroy(ctx);
    });
}

}  // extern "C"
