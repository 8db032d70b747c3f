// Internal declarations shared by the translation units of libtbkpm.so:
// error handling, device buffers, the context / matrix / block objects of
// tbkpm.h, and the step-launch planner.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "gamma_launch.h"
#include "step_launch.h"
#include "tbkpm.h"

namespace engine {

inline thread_local std::string g_last_error;

struct Error {
    int code;
    std::string msg;
};

[[noreturn]] inline void raise(int code, const std::string& msg) { throw Error{code, msg}; }

inline void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        raise(TBK_E_OOM, std::string(what) + ": " + hipGetErrorString(e));
    raise(TBK_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPC(expr) hip_check((expr), #expr)

template <typename F>
inline int guarded(F&& f) {
    try {
        f();
        return TBK_OK;
    } catch (const Error& e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return TBK_E_OOM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return TBK_E_INVALID;
    } catch (...) {
        g_last_error = "unknown error";
        return TBK_E_INVALID;
    }
}

// A span of device memory that buffers are carved from front to back
// (256-byte aligned); the owner of the span frees it, the buffers never do.
struct Region {
    char* base = nullptr;
    size_t cap = 0, off = 0;
    void* take(size_t n) {
        const size_t o = (off + 255) & ~static_cast<size_t>(255);
        if (!base || o > cap || n > cap - o) return nullptr;
        off = o + n;
        return base + o;
    }
};

// The region that matrix arrays are placed in on this thread (null: their
// own allocations); set by RegionScope.
inline thread_local Region* tl_matrix_region = nullptr;

struct RegionScope {
    Region* prev;
    explicit RegionScope(Region* r) : prev(tl_matrix_region) { tl_matrix_region = r; }
    ~RegionScope() { tl_matrix_region = prev; }
    RegionScope(const RegionScope&) = delete;
    RegionScope& operator=(const RegionScope&) = delete;
};

// Grow-only device buffer: its own allocation, or a piece of a Region when
// one is given and has room.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool owned = true;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        owned = true;
    }
    void reserve(size_t n, Region* rg = nullptr) {
        if (n <= bytes) return;
        release();
        if (n == 0) return;
        if (rg) {
            if (void* q = rg->take(n)) {
                p = q;
                bytes = n;
                owned = false;
                return;
            }
        }
        HIPC(hipMalloc(&p, n));
        bytes = n;
    }
    // a matrix array: in the thread's matrix region when one is set
    void reserve_placed(size_t n) { reserve(n, tl_matrix_region); }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

inline int pow2_at_least(int v) {
    int r = 1;
    while (r < v) r <<= 1;
    return r;
}

}  // namespace engine

struct tbk_context {
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    engine::DevBuf work[2];
    engine::DevBuf partial;
    engine::DevBuf sums;
    engine::DevBuf stack[2];                 // Kubo left/right vector stacks, kept between calls
    engine::DevBuf kwork[3];                 // rolling fp64 blocks of pre-split Kubo stacks
    engine::DevBuf gamma_part, gamma_dev;    // split-K Gamma partials and the reduced Gamma
    std::vector<hipEvent_t> events;
    int64_t held_bytes = 0;  // matrices + blocks owned through this context
    tbk_tuning tuning{};     // zeros: automatic choices
};

struct tbk_matrix {
    tbk_context* ctx = nullptr;
    int32_t rows = 0;
    int64_t nnz = 0;
    bool is_complex = false;
    int32_t max_row_nnz = 0;
    int32_t bandwidth95 = 0;  // 95th percentile of |row - col| over the entries
    // formed on the device on first use (matrix_gershgorin, matrix_pattern)
    bool have_gersh = false;
    double gersh_min = 0.0, gersh_max = 0.0;  // Gershgorin interval (Hermitian reading)
    uint64_t pattern = 0;                     // fingerprint of (rows, row pointers, columns); 0: not formed
    engine::DevBuf rp, ci, val64, val32;
    bool have32 = false;
};

struct tbk_block {
    tbk_context* ctx = nullptr;
    int32_t rows = 0, R = 0, Rp = 0, precision = TBK_FP64;
    engine::DevBuf data;
    size_t elem_bytes() const { return precision == TBK_FP64 ? 16 : 8; }
};

namespace engine {

inline void set_device(const tbk_context* ctx) { HIPC(hipSetDevice(ctx->device)); }

inline void ensure_events(tbk_context* ctx, size_t n) {
    while (ctx->events.size() < n) {
        hipEvent_t e;
        HIPC(hipEventCreate(&e));
        ctx->events.push_back(e);
    }
}

using tbk::StepArgs;

// Kernel family, slab width and rows per tile for a matrix and block shape.
struct KernelChoice {
    int kch;            // 0: row kernel
    int width;          // columns one sweep covers (the slab width; Rp for the row kernel)
    int rows_per_tile;
    int cpl;            // complex values per lane
};

// ------------------------------------------------------------ the driver
// Launch plan of one step kernel for a matrix and a block shape: kernel
// family, slab width, visiting order and grid.  Shared by every recursion
// driver (moments, Kubo stacks).
struct StepPlan {
    KernelChoice kc;
    StepArgs args;  // x, z, y, bra, alpha, beta, partial filled per launch
    int64_t nb;
    int precision;
    bool cplx;
};

void launch_any(int precision, bool cplx, int width, int kch, int cpl, int mode, dim3 grid, hipStream_t st,
                const StepArgs& a);
KernelChoice choose_kernel(const tbk_tuning& tu, int precision, bool cplx, int Rp, int max_row_nnz,
                           int bandwidth, bool force_row, bool row_slabs = false);
int bandwidth_percentile(const std::vector<int32_t>& rp, const std::vector<int32_t>& ci, double p);
void build_csr(const tbk_sparse* M, std::vector<int32_t>& rp, std::vector<int32_t>& ci, std::vector<double>& val);
bool csr_hermitian_within(int32_t n, const std::vector<int32_t>& rp, const std::vector<int32_t>& ci,
                          const std::vector<double>& val, bool cplx, double tol);
void matrix_gershgorin(const tbk_matrix* M, hipStream_t st);
uint64_t matrix_pattern(const tbk_matrix* M, hipStream_t st);
tbk_context* default_context();
// Hamiltonian assembly on the device (assemble.hip)
int64_t lattice_rows(const tbk_lattice_desc* d);
int64_t lattice_nnz_bound(const tbk_lattice_desc* d);
void assemble_matrix(tbk_context* ctx, const tbk_lattice_desc* d, hipStream_t st, tbk_matrix** out, double* ms);
int32_t distance_percentile95(const int32_t* rp, const int32_t* ci, int32_t n, int64_t nnz, hipStream_t st);
// Operators formed from resident matrices (operators.hip)
void rescale_matrix(tbk_context* ctx, const tbk_matrix* M, double a, double b, hipStream_t st, tbk_matrix** out);
void velocity_matrix(tbk_context* ctx, const tbk_matrix* M, const tbk_lattice_desc* d, int axis, hipStream_t st,
                     tbk_matrix** out);
// FP64 Lanczos estimate of the spectral range of a Hermitian matrix (bounds.hip)
void spectral_bounds(tbk_context* ctx, const tbk_matrix* M, int max_iters, double tol, uint64_t seed,
                     hipStream_t st, tbk_bounds* out);
void matrix_fp32(tbk_matrix* m);
StepPlan make_plan(tbk_context* ctx, const tbk_matrix* M, int prec, int Rp, int rows, double b, bool force_row);
StepPlan make_plan(tbk_context* ctx, const tbk_matrix* M, int prec, int Rp, int rows, double b, bool force_row,
                   const tbk_tuning& tu);
// row_slabs: the row kernel may sweep tu.row_slab_cols columns at a time (the
// moment recursions; the Kubo stack modes store whole rows)
StepPlan make_plan(tbk_context* ctx, const tbk_matrix* M, int prec, int Rp, int rows, double b, bool force_row,
                   const tbk_tuning& tu, bool row_slabs);
void launch_plan(const StepPlan& pl, hipStream_t st, int mode, double alpha, double beta, const void* x,
                 const void* z, void* y, const void* bra, double2* partial, const void* hv2, void* y2);
// y3: the paired modes' right-half words; wbs: block stride of k-blocked
// word stacks (word_offset)
void launch_plan(const StepPlan& pl, hipStream_t st, int mode, double alpha, double beta, const void* x,
                 const void* z, void* y, const void* bra, double2* partial, const void* hv2, void* y2, void* y3,
                 int64_t wbs);
void launch_plan(const StepPlan& pl, hipStream_t st, int mode, double alpha, double beta, const void* x,
                 const void* z, void* y, const void* bra, double2* partial);

// Work split of one problem in a cheb_slots_kernel launch, the split that
// cheb_step_kernel uses for a single matrix: row tiles of one workgroup pass
// (lpr lanes per row), chunks of 256 rows transposed with period P =
// bandwidth / chunk rows, and nwg workgroups (a multiple of 8, at most
// blocks_per_cu per CU) dealing the tiles per XCD.  Depends only on the
// problem, so its partial sums are the same in any batch.
inline void slot_split(int num_cu, int blocks_per_cu, int lpr, int64_t rows, int64_t bandwidth95,
                       tbk::TileOrder& order, int32_t& nwg) {
    const int64_t TR = static_cast<int64_t>(tbk::kWavesPerBlock) * (tbk::kWave / lpr);
    const int64_t row_tiles = (rows + TR - 1) / TR;
    const int64_t tpc = std::max<int64_t>(1, 256 / TR);
    const int64_t chunks = (row_tiles + tpc - 1) / tpc;
    int64_t P = std::max<int64_t>(1, (bandwidth95 + (tpc * TR) / 2) / (tpc * TR));
    P = std::min<int64_t>(P, chunks);
    const int64_t Q = (chunks + P - 1) / P;
    order.row_tiles = static_cast<uint32_t>(row_tiles);
    order.tpc = static_cast<uint32_t>(tpc);
    order.P = static_cast<uint32_t>(P);
    order.Q = static_cast<uint32_t>(Q);
    const int64_t visits = P * Q * tpc;
    const int64_t cap = std::max<int64_t>(8, (static_cast<int64_t>(num_cu) * blocks_per_cu / 8) * 8);
    nwg = static_cast<int32_t>(std::min<int64_t>(cap, ((visits + 7) / 8) * 8));
}

// Complex values per lane of a slot class: 16-byte lanes (one complex128 or
// two complex64) unless requested, within 64 lanes per row.
inline int slot_cpl(int requested, int prec, int Rp) {
    int cpl = requested > 0 ? requested : (prec == TBK_FP64 ? 1 : std::min(2, Rp));
    if (Rp / cpl > tbk::kWave || (cpl == 4 && Rp < 8) || cpl > Rp) cpl = Rp > tbk::kWave ? Rp / tbk::kWave : 1;
    return cpl;
}

// Kubo-Bastin contraction shape and launches (gamma.hip kernels).
struct ContractionShape {
    bool big;
    int64_t kc;
    int nchunks;
    tbk::GammaMap map;
    bool wide;               // pre-split words: the 256 x 128 kernel on wide_map
    tbk::GammaMap wide_map;  // (same chunks as map)
};
ContractionShape contraction_plan(const tbk_context* ctx, int N, int64_t K, int prec);
void enqueue_contraction(const ContractionShape& cs, const double2* A, const double2* B, int N, int64_t K, int64_t ldk,
                         int prec,
                         double2* P, double2* out, hipStream_t st);
// vx_values: v_x's values when v_x has H's pattern (fused_velocity), else null
// stack entry n of L and Rs at n * ldk elements (stack_stride)
int64_t stack_stride(int64_t K);
void enqueue_kubo_stacks(const StepPlan& pH, const StepPlan& pX, const StepPlan& pY, double a, double b,
                         const void* kets, void* L, void* Rs, void* W0, void* W1, size_t blk_bytes, int64_t ldk,
                         int N, double2* part, hipStream_t st, const void* vx_values);
bool fused_velocity(const tbk_matrix* H, const tbk_matrix* vx, const StepPlan& pH, hipStream_t st);
// fp64 stacks of rows [r0, r1) of every entry, the chains in full rolling
// blocks B0, B1, W0, W1 and the scratch block S
void enqueue_kubo_stacks_rows(const StepPlan& pH, const StepPlan& pX, const StepPlan& pY, double a, double b,
                              const void* kets, void* L, void* Rs, void* B0, void* B1, void* W0, void* W1, void* S,
                              int rows, int Rp, int r0, int r1, int64_t ldk, int N, double2* part, hipStream_t st);
bool kubo_presplit(int prec, int N, bool h_complex, bool vx_complex);
void enqueue_kubo_stacks_presplit(const StepPlan& pH, const StepPlan& pX, const StepPlan& pY, double a, double b,
                                  const void* kets, uint2* Ls, uint2* Rs, void* B0, void* B1, void* S, void* W0,
                                  void* W1, int64_t K, int N, double2* part, hipStream_t st, const void* vx_values);
// A, B: k-blocked word stacks of N entries (word_offset, block stride N * kWordBlock)
void enqueue_contraction_presplit(const ContractionShape& cs, const uint2* A, const uint2* B, int N, int64_t K,
                                  double2* P, double2* out, hipStream_t st);
// pre-split stacks with both chains in one block of 2 Rp columns (fused v_x
// only; pP: H's plan at 2 Rp columns, row kernel, fp64 or fp32 -- then
// vx_values are v_x's fp32 values); X0, X1: 2 fp64 blocks each (one in
// fp32), S: one block of scratch
void enqueue_kubo_stacks_paired(const StepPlan& pP, const StepPlan& pY, double a, double b, const void* kets,
                                uint2* Ls, uint2* Rs, void* X0, void* X1, void* S, int64_t rows, int Rp, int N,
                                double2* part, hipStream_t st, const void* vx_values);
double kubo_recursion_bytes(const tbk_matrix* H, const tbk_matrix* vx, const tbk_matrix* vy, int rows, int R,
                            int64_t N, bool fused, bool presplit, bool paired = false, bool paired_f32 = false);
// the contraction's kernel family of a gamma precision (TBK_GAMMA_BF16X3_FP32
// contracts as TBK_GAMMA_BF16X3)
int contraction_precision(int prec);

}  // namespace engine
