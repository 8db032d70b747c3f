/*
 * tbkpm.h - C ABI of the MI355X (gfx950) kernel-polynomial engine.
 *
 * Plain pointers and sizes only.  Every function returns TBK_OK (0) or a
 * negative TBK_E_* code; tbk_last_error() returns the message of the last
 * failure on the calling thread.  No C++ exception crosses this boundary.
 *
 * Dense vector blocks cross the boundary as the reference passes them:
 * column-major rows x R, complex128 interleaved (re, im) -- the layout of
 * Eigen::MatrixXcd bras_T / kets in numerics_cuda.cuh:10-13.  On the device a
 * block is stored row-major [site][Rp] (Rp = R rounded up to a power of two,
 * padding columns zero) so one wavefront reads a site's whole row of vectors
 * with one coalesced access.
 *
 * Sparse matrices cross as compressed arrays with int32 indices, either
 * column-compressed (CSC: Eigen's default ColMajor storage that the reference
 * hands to cuSPARSE, numerics_cuda.cu:52-64) or row-compressed (CSR).
 *
 * Threading: a tbk_context owns one HIP stream on one device.  Calls on one
 * context must not overlap (callers serialise them); distinct contexts may be
 * used from distinct threads.
 */
#ifndef TBKPM_H
#define TBKPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TBK_OK = 0,
    TBK_E_INVALID = -1,     /* bad argument: n < 2, R < 1, null pointer, shape mismatch */
    TBK_E_DEVICE = -2,      /* HIP runtime or launch failure, or no device */
    TBK_E_OOM = -3,         /* device allocation failed */
    TBK_E_UNSUPPORTED = -4, /* valid request outside what this build supports */
    TBK_E_CANCELLED = -5    /* a server request cancelled by tbk_server_cancel */
};

/* precision of a vector block and of the recursion run on it */
enum {
    TBK_FP64 = 0, /* complex128 vectors, fp64 matrix values, fp64 moments */
    TBK_FP32 = 1  /* complex64 vectors, fp32 matrix values, moments summed in fp64 */
};

enum { TBK_CSC = 0, TBK_CSR = 1 };

/* flags for tbk_moments */
enum {
    TBK_FLAG_PLAIN = 1,         /* bras == kets: still run one step per moment (no doubling) */
    TBK_FLAG_TIME_LAUNCHES = 2, /* time every step launch with its own HIP event pair */
    TBK_FLAG_ROW_KERNEL = 4     /* use the row kernel even where the tiled kernel applies */
};

typedef struct tbk_sparse {
    int32_t rows, cols;
    int64_t nnz;
    const int32_t* outer; /* CSC: column pointers (cols+1); CSR: row pointers (rows+1) */
    const int32_t* inner; /* CSC: row indices; CSR: column indices */
    const double* values; /* nnz doubles, or 2*nnz interleaved (re, im) if is_complex */
    int32_t is_complex;
    int32_t layout; /* TBK_CSC or TBK_CSR */
} tbk_sparse;

typedef struct tbk_stats {
    double kernel_ms;       /* device time of the recursion step launches (HIP events) */
    double wall_ms;         /* host wall time of the whole call */
    int64_t steps;          /* step launches (applications of H~ to the block) */
    int64_t moments;        /* moments produced */
    double bytes_per_step;  /* algorithmic HBM bytes of one recursion step launch */
    double bytes_total;     /* algorithmic HBM bytes of all step launches */
    int32_t doubling;       /* 1 if the Chebyshev product identity produced two moments/step */
    int32_t blocks;         /* workgroups per step launch */
    int32_t rows_per_tile;  /* rows one workgroup handles per tile */
    int32_t padded_R;       /* vector columns stored per site on the device */
    int32_t kernel_kind;    /* 0: row kernel; 1, 2, 4: tiled kernel for rows of <= 4*kind entries */
    int32_t slab_cols;      /* vector columns one sweep over the rows covers */
    int32_t chunk_rows;     /* rows per chunk of the visiting order */
    int32_t chunk_period;   /* period P of the transposed chunk order (1: row order) */
} tbk_stats;

/* Launch-shape choices of a context (tbk_autotune sets these; zero fields
   mean the defaults, which are the measured best on the 2.1M-site
   honeycomb with R = 64). */
typedef struct tbk_tuning {
    int32_t slab_cols;       /* columns per sweep of the tiled kernel (power of two <= R) */
    int32_t blocks_per_cu;   /* workgroups per CU in a step launch (default 4) */
    int32_t xcd_split;       /* 1: deal row ranges per XCD, -1: never, 0: when >= 64 workgroups */
    int32_t l2_budget_bytes; /* per-XCD L2 bytes the slab heuristic may plan for (default 3 MiB) */
    int32_t stream_nt;       /* non-temporal loads/stores for rows used once per step: 0/1 on, -1 off */
    int32_t use_tiled;       /* 1: LDS-staged tiled kernel where rows have <= 16 entries; 0: row kernel */
    int32_t chunk_rows;      /* rows per chunk of the visiting order (default 256) */
    int32_t chunk_period;    /* chunk-grid period P; 0: bandwidth / chunk rows, -1: row order */
    int32_t row_lane_bytes;  /* row kernel: vector bytes per lane, 8, 16 or 32 (0: 16), i.e. 1 or 2
                                complex128 / 1, 2 or 4 complex64 (at least R/64 complex); the
                                paired sigma_xy block (both chains, 2R columns) uses this width,
                                or 32 when it is 0 */
    int32_t row_slab_cols;   /* moment recursions with the row kernel: sweep the rows this many
                                vector columns at a time (a power of two below the padded R; 0:
                                whole rows), so that the gathered rows of a sweep take less L2 */
    int32_t kubo_stack_mib;  /* tbk_kubo_gamma: the most device memory (MiB) its vector stacks and
                                work blocks may take; beyond it the random vectors run in groups
                                (0: 95% of the free memory) */
    int32_t gamma_tile_rows; /* split-bf16 contraction over pre-split stacks: Gamma rows per
                                workgroup tile, 256 (0: default; 64 x 64 per wave) or 128 (64 x 32
                                per wave); both give the same bits */
} tbk_tuning;

const char* tbk_version(void);
const char* tbk_last_error(void);
int tbk_device_count(int32_t* count);

/*
 * Drop-in for cuda::Cheb_first_kind_seq_braket (numerics_cuda.cuh:10-13,
 * numerics_cuda.cu:190-323; CPU twin numerical.h:363-381): both overloads
 * (real and complex M) in one entry selected by M->is_complex.
 *   mu[i] = sum over sites and columns of Re(conj(bras_T) * T_i(M) kets),
 *   T_0 = kets, T_1 = M kets, T_i = 2 M T_{i-1} - T_{i-2}, i < n.
 * Not divided by the number of columns.  M is used as given (already
 * rescaled).  n < 2 fails with TBK_E_INVALID, as numerics_cuda.cu:194-197.
 * bras_T and kets: ket_rows x ket_cols column-major complex128, both of that
 * size.  When bras_T is the same pointer as kets and M is exactly Hermitian
 * (checked on the host: M^H equals M value for value), two moments are formed
 * per step by the doubling identities; any other input takes one moment per
 * step against the bra block.  Uses a process-wide context on device 0.
 */
int tbk_cheb_first_kind_seq_braket(const tbk_sparse* M, int32_t n, const double* bras_T,
                                   const double* kets, int32_t ket_rows, int32_t ket_cols,
                                   double* mu_out);

/*
 * Host-buffer drop-in for Chebyshev::Cheb_first_kind_seq_on_bases
 * (numerical.h:334-344) with a sparse M (used as given, already rescaled):
 *   out block k = T_k(M) v, k = 0 .. n-1,
 * n consecutive v_rows x v_cols column-major complex128 blocks.  FP64 on the
 * device in two rolling blocks, each T_k copied to out once formed, so device
 * memory stays at three blocks whatever n is (tbk_cheb_seq_on_bases keeps
 * every block resident instead).  n >= 1; n = 1 gives v.  Uses the
 * process-wide context on device 0.
 */
int tbk_cheb_first_kind_seq_on_bases(const tbk_sparse* M, int32_t n, const double* v, int32_t v_rows,
                                     int32_t v_cols, double* out);

/* ------------------------------------------------ device-resident objects */
typedef struct tbk_context tbk_context;
typedef struct tbk_matrix tbk_matrix;
typedef struct tbk_block tbk_block;

int tbk_context_create(int32_t device, tbk_context** out);
int tbk_context_destroy(tbk_context* ctx);
int tbk_context_synchronize(tbk_context* ctx);
int tbk_context_set_tuning(tbk_context* ctx, const tbk_tuning* t);
int tbk_context_get_tuning(const tbk_context* ctx, tbk_tuning* t);
/* bytes of device memory the context currently holds (matrices, blocks, work) */
int tbk_context_memory(tbk_context* ctx, int64_t* bytes);
/* free the work buffers the context keeps between calls (recursion blocks,
   partial sums, Kubo stacks and Gamma buffers); matrices and blocks stay */
int tbk_context_release_work(tbk_context* ctx);

/* upload a square sparse matrix; CSC input is reordered to CSR exactly */
int tbk_matrix_create(tbk_context* ctx, const tbk_sparse* M, tbk_matrix** out);
int tbk_matrix_destroy(tbk_matrix* m);
/* shape statistics used by the kernel choice: longest row and the 95th
   percentile of |row - col| over the entries */
int tbk_matrix_info(const tbk_matrix* m, int32_t* rows, int64_t* nnz, int32_t* max_row_nnz,
                    int32_t* bandwidth95);

/* copy a resident matrix back row-compressed: row pointers (rows + 1),
   column indices (nnz, ascending within a row), values (nnz doubles, or
   2 nnz interleaved if complex) */
int tbk_matrix_download(const tbk_matrix* m, int32_t* rowptr, int32_t* colidx, double* values);

/*
 * A lattice Hamiltonian described by its tables, for assembly on the device.
 * builder 0: get_Hamiltonian_sparse (open, lattice.h:262-305) without removed
 * sites; 1: get_Hamiltonian_periodic_extend_bare_sparse(extend)
 * (lattice.h:216-258); 2: get_Hamiltonian_bare_sparse (lattice.h:347-377,
 * no fields).  Sites: bare index = cell . strides + atom, last lattice index
 * fastest (lattice.h:28, 106-117).  The hopping table is the lattice's own, in
 * table order (including the zero on-site entries a scalar field adds).
 */
typedef struct tbk_lattice_desc {
    int32_t dim;              /* 1, 2 or 3 */
    int32_t span[3];          /* cells per axis (first dim used) */
    int32_t extend[3];        /* builder 1: wrap along axes with extend != 0 */
    int32_t n_atoms, n_hops, builder, is_complex;
    const int32_t* hop_from;  /* n_hops atom indices */
    const int32_t* hop_to;
    const int32_t* hop_lv;    /* n_hops * dim lattice vectors */
    const double* hop_t;      /* 2 n_hops: (re, im) */
    const double* phase;      /* NULL, or exp(+i A(r_from).d) per (cell, hopping) as the host builders
                                 compute it, cell_number * n_hops complex (builders 0, 1) */
    int32_t has_disorder;     /* on-site W (u(seed, site) - 1/2), the site-keyed counter generator */
    double disorder_W;
    uint64_t disorder_seed;
    /* geometry, for position and velocity operators (tbk_matrix_velocity;
       tbh_lattice_geometry gives both): bases[3 d + k] = component k of
       lattice vector d; atom_pos[3 a + k] = component k of atom a's
       position, or NULL where no position operator is formed */
    double bases[9];
    const double* atom_pos;
} tbk_lattice_desc;

/*
 * Assemble on the device the matrix the host builder produces (the same
 * entries, order and bits) and keep it resident; no host-side matrix is
 * formed.  At most 32 hoppings may touch one atom (TBK_E_UNSUPPORTED).
 */
int tbk_matrix_assemble(tbk_context* ctx, const tbk_lattice_desc* d, tbk_matrix** out, double* wall_ms);

/*
 * H~ = (M - b I)/a as a new resident matrix over the union pattern of M and
 * I, every diagonal entry explicit: rescale / GreenFunction::scale
 * (lattice.h:550-559).  The entries and bits of the host rescale of the same
 * matrix: (m_ij - b)/a on the diagonal, (m_ij - 0)/a elsewhere, (0 - b)/a
 * for a diagonal M lacks.
 */
int tbk_matrix_rescale(tbk_context* ctx, const tbk_matrix* M, double a, double b, tbk_matrix** out);

/*
 * The velocity operator v = X M - M X without the factor i
 * (commutator_with_diagonal, lattice.h:338-345; snippets.h:316-317) for the
 * position operator X along `axis` of the lattice's sites (op_position,
 * lattice.h:314-333):
 *   x_s = atom_pos[a][axis] + (sum_d bases[d][axis] i_d)  for site s of
 *   cell i, atom a (the builders' bare index; no removed sites).
 * M's pattern; values x_i m_ij - m_ij x_j, the host's bits.  M's rows must
 * equal the lattice's sites and d->atom_pos must be set.
 */
int tbk_matrix_velocity(tbk_context* ctx, const tbk_matrix* M, const tbk_lattice_desc* d, int32_t axis,
                        tbk_matrix** out);

/* upload rows x R column-major complex128 (NULL: zeros) as a block */
int tbk_block_create(tbk_context* ctx, const double* colmajor, int32_t rows, int32_t R,
                     int32_t precision, tbk_block** out);
/* random-phase block e^{i theta}/sqrt(rows), theta = 2 pi u(seed, site*R + r),
   generated on the device */
int tbk_block_random_phase(tbk_context* ctx, int32_t rows, int32_t R, uint64_t seed,
                           int32_t precision, tbk_block** out);
/* Attainable-bandwidth references on buffers of `bytes` each, best of
   `repeats`, all with non-temporal accesses: a copy (read + write), the
   two-reads-one-write mix of a recursion step without its gathers
   (y = a + b), and a read-only reduction.  GB/s counts every byte moved. */
typedef struct tbk_bandwidth {
    double copy_GBps, stream_2r1w_GBps, read_GBps;
    double bytes; /* per buffer */
} tbk_bandwidth;
int tbk_bandwidth_probe(tbk_context* ctx, int64_t bytes, int32_t repeats, tbk_bandwidth* out);

/* copy a block back as rows x R column-major complex128 */
int tbk_block_download(const tbk_block* b, double* colmajor_out);
/* device copy dst = src (same shape and precision); kernel_ms (optional)
   receives the copy's device time */
int tbk_block_copy(const tbk_block* src, tbk_block* dst, float* kernel_ms);
int tbk_block_destroy(tbk_block* b);

/*
 * Chebyshev moments of H~ = (M - b I)/a, rescaling fused into the step
 * kernel (a = 1, b = 0 uses M as given):
 *   mu[i] = Re <bras| T_i(H~) |kets>, i = 0 .. n-1.
 * bras == NULL means bras = kets; then, unless TBK_FLAG_PLAIN, moments come
 * two per step from T_{m+k} = 2 T_m T_k - T_{|m-k|}:
 *   mu_{2k} = 2 <T_k|T_k> - mu_0,  mu_{2k+1} = 2 <T_{k+1}|T_k> - mu_1.
 * The recursion runs in the precision of `kets` (bras must match).
 * last_out (optional, NULL to skip): the last vector block T_{steps} written by
 * the recursion, rows x R column-major complex128.
 */
int tbk_moments(tbk_context* ctx, const tbk_matrix* M, double a, double b,
                const tbk_block* bras, const tbk_block* kets, int32_t n, int32_t flags,
                double* mu_out, double* last_out, tbk_stats* stats);

/*
 * tbk_moments without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.
 *
 * Same moments as tbk_moments (the same bits for the same context tuning),
 * written to mu_dev, a device array of n doubles.  The two rolling vector
 * blocks are the caller's: work0 and work1, device buffers of at least
 * kets' rows x Rp elements each (Rp = R rounded up to a power of two;
 * elements complex128 in TBK_FP64, complex64 in TBK_FP32; see
 * tbk_block_device_ptr), distinct from each other and from kets and bras.  On
 * return T_steps is in work1 when steps is odd and in work0 when even
 * (steps = (n+1)/2 with doubling, n-1 otherwise), T_{steps-1} in the other.
 * The per-step moment partials live in a fixed 2.5 MiB device array, one
 * per GPU, kept until the process ends: at most 32768 steps and 4096
 * workgroups per step launch (TBK_E_UNSUPPORTED beyond).  Every call on a
 * GPU shares that array, so calls that overlap on one GPU (two streams, two
 * threads) race; calls ordered on one stream do not.  Work is enqueued on
 * `stream` (a hipStream_t; NULL: the context's stream) and the call returns
 * without synchronising.  Nothing is allocated: in TBK_FP32 the matrix's
 * fp32 values must already exist (formed by its first fp32 tbk_moments).
 * TBK_FLAG_TIME_LAUNCHES is rejected (timing needs events).
 */
int tbk_moments_noalloc(tbk_context* ctx, const tbk_matrix* M, double a, double b,
                        const tbk_block* bras, const tbk_block* kets, int32_t n, int32_t flags,
                        void* work0, void* work1, double* mu_dev, void* stream);
/* the device address of a block's storage: rows x Rp elements, row-major */
int tbk_block_device_ptr(const tbk_block* b, void** out);

/*
 * Stored Chebyshev blocks, device-resident: replaces
 * Chebyshev::Cheb_first_kind_seq_on_bases (numerical.h:334-344) for sparse M.
 *   out[0] = v,  out[1] = H~ v,  out[k] = 2 H~ out[k-1] - out[k-2],  k < n,
 * H~ = (M - b I)/a.  out holds n blocks made by the caller with v's rows, R
 * and precision, pairwise distinct; out[0] may be v itself (no copy), no
 * other may.  The blocks have the bits of the vectors tbk_moments forms.
 * kernel_ms (optional): device time of the n - 1 launches.  n >= 1 (the
 * reference's loop writes past its vector for n < 2).
 */
int tbk_cheb_seq_on_bases(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* v,
                          int32_t n, tbk_block* const* out, float* kernel_ms);

/*
 * Autotuner: times `steps` recursion steps of tbk_moments for M and kets under
 * candidate launch shapes (coordinate descent over kernel family, workgroups
 * per CU, chunk size, XCD split and streaming stores, from the current
 * tuning), installs the fastest in the context and returns it with its time
 * per step.  Results do not depend on the tuning beyond rounding of the
 * moment sums (the recursion vectors are bit-identical across shapes).
 */
/*
 * One Chebyshev step on resident blocks, the step of the recursions above
 * (the device side of Chebyshev::Cheb_first_kind_seq_view::next,
 * numerical.h:346-361):
 *   z == NULL:  y = H~ x,           z != NULL:  y = 2 H~ x - z,
 * H~ = (M - b I)/a.  y may be z (in place); neither y nor z may be x.  All
 * blocks share x's shape and precision.  Enqueued on the context's stream
 * (tbk_block_download waits for it).
 */
int tbk_cheb_step(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* x,
                  const tbk_block* z, tbk_block* y);

/*
 * Chebyshev expansion applied to a block:  out = sum_{k<n} c_k T_k(H~) v,
 * c_k = coef[2k] + i coef[2k+1].  The terms are added in order of k, each as
 * soon as T_k is formed (the recursion of tbk_cheb_seq_on_bases in two
 * rolling blocks).  With v = columns of the identity and c_k = Im alpha_k(z)
 * this is GreenFunction::direct_eval's sum (lattice.h:579-587) without
 * storing the T_k.  out: v's shape and precision, not v.  kernel_ms
 * (optional): device time of the steps and additions after the first term.
 */
int tbk_cheb_expand(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* v, int32_t n,
                    const double* coef, tbk_block* out, float* kernel_ms);

int tbk_autotune(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* kets,
                 int32_t steps, tbk_tuning* best, double* best_ms_per_step);
/* flags of tbk_autotune_flags, besides TBK_FLAG_PLAIN (tune the
   one-moment-per-step recursion instead of the two-moments-per-step one) */
enum { TBK_AUTOTUNE_REFRESH = 8 /* run the trials even when the cache holds this problem's shape */ };
/*
 * tbk_autotune for one recursion mode, with a process-wide cache: the shape
 * found for a problem is kept under the matrix's rows, entries, longest row
 * and 95th-percentile bandwidth, the padded R, precision, value type, mode
 * and the device's CU count, and a later call with the same key installs it
 * without trials (best_ms_per_step then is the time measured when it was
 * found).  tbk_autotune is this call with flags 0.
 */
int tbk_autotune_flags(tbk_context* ctx, const tbk_matrix* M, double a, double b, const tbk_block* kets,
                       int32_t steps, int32_t flags, tbk_tuning* best, double* best_ms_per_step);
/* the tuning cache as a text file (one line per problem shape), read back by
   a later process; entries (optional) receives the number of lines read */
int tbk_tuning_cache_save(const char* path);
int tbk_tuning_cache_load(const char* path, int32_t* entries);
int tbk_tuning_cache_clear(void);

/* ----------------------------------------------- Kubo-Bastin Gamma_nm */
enum {
    TBK_GAMMA_FP64 = 0,   /* fp64 matrix-core contraction */
    TBK_GAMMA_BF16X3 = 1, /* split-bf16 three-term products, fp32 accumulation per chunk, fp64 across */
    /* TBK_GAMMA_BF16X3's contraction over stacks from an fp32 recursion
     * (complex64 vectors, fp32 H and v_x values, words split from the fp32
     * values) where the paired stacks apply: complex H, v_x with H's pattern,
     * N >= 256 and 2 <= R <= 128.  Other problems take TBK_GAMMA_BF16X3's
     * fp64 recursion; tbk_kubo_stats.recursion_fp32 says which ran.  The
     * split words carry 16 significant bits, so the fp32 recursion's rounding
     * stays below the contraction's own. */
    TBK_GAMMA_BF16X3_FP32 = 2
};

typedef struct tbk_kubo_stats {
    double recursion_ms;   /* device time building both vector stacks */
    double contraction_ms; /* device time of the Gamma contraction and its chunk reduction */
    double wall_ms;
    double flops;          /* 8 N^2 K: real flops of the complex contraction, K = sites x columns */
    double stack_bytes;    /* bytes of the two stacks' entries (each entry padded to a multiple of 64 elements) */
    int32_t nchunks;       /* split-K chunks of the contraction */
    int32_t vector_groups; /* groups the random vectors ran in (1: the whole block; more when the
                              whole block's stacks exceed the free device memory) */
    double recursion_bytes; /* algorithmic HBM bytes of the stack-building launches and copies */
    int32_t row_ranges;     /* row ranges the fp64 stacks held one at a time (1: whole entries; more
                               when the stacks of one random vector exceed the memory available) */
    int32_t recursion_fp32; /* 1 when the stacks came from the fp32 recursion (TBK_GAMMA_BF16X3_FP32) */
} tbk_kubo_stats;

/*
 * Kubo-Bastin moments (snippets.h:306-342 with the double-sum definition of
 * snippets.h:337-342; stored-block recursion numerical.h:334-344):
 *   aL_n = g_n T_n(H~) |r>,   aR_m = g_m v_x T_m(H~) v_y |r>,
 *   gamma[n*N + m] = sum over sites and columns of conj(aL_n) * aR_m,
 * H~ = (H - b I)/a fused as in tbk_moments, v_x and v_y used as given (the
 * reference builds them as X H~ - H~ X, lattice.h:338-345 / snippets.h:316-317),
 * g (length N) the kernel weights or NULL for ones.  gamma_out: N*N complex128
 * row-major (re, im interleaved).  kets must be fp64.  Both stacks (2 N blocks)
 * are held in device memory during the call; TBK_E_OOM if they do not fit.
 */
int tbk_kubo_gamma(tbk_context* ctx, const tbk_matrix* H, double a, double b, const tbk_matrix* vx,
                   const tbk_matrix* vy, const tbk_block* kets, int32_t N, int32_t gamma_precision,
                   const double* g, double* gamma_out, tbk_kubo_stats* stats);

/*
 * The contraction alone: gamma[n*N + m] = sum_k conj(A[n*K + k]) * B[m*K + k]
 * for host arrays A, B of N x K complex128 (interleaved).
 */
int tbk_stack_contract(tbk_context* ctx, const double* A, const double* B, int32_t N, int64_t K,
                       int32_t gamma_precision, double* gamma_out, double* contraction_ms);

/*
 * Kubo-Bastin integrand at rescaled energies e[0..n_e), |e_j| < 1
 * (snippets.h:343-349 g_mn, normalised as snippets.h:400-405):
 *   zeta_j = sum_{n,m} gamma[n][m] g_mn(e_j) / (1 - e_j^2)^2 / volume,
 *   g_mn(e) = a_n a_m [(e - i m s) e^{i m theta} T_n(e) + (e + i n s) e^{-i n theta} T_m(e)],
 * e = cos(theta), s = sqrt(1 - e^2), a_0 = 1, a_{k>0} = 2; computed on the
 * device.  gamma: N x N complex128 row-major (kernel weights applied, as
 * tbk_kubo_gamma returns it); zeta_out: n_e complex128.  N <= 2730 (the
 * per-energy tables must fit in LDS), else TBK_E_UNSUPPORTED.
 */
int tbk_kubo_zeta(tbk_context* ctx, const double* gamma, int32_t N, const double* e, int32_t n_e, double volume,
                  double* zeta_out, double* kernel_ms);

/* ------------------------------------------------ spectral reconstruction */
/*
 * Chebyshev reconstruction at energies e[0..n_e) in rescaled units, on the
 * device (one workgroup per energy, fixed-order sums):
 *   out_j = scale * [g_0 mu_0 + 2 sum_{n=1}^{N-1} g_n mu_n cos(n theta_j)] / sin(theta_j),
 *   e_j = cos(theta_j).
 * scale = 1: the values of GreenFunctionImag::fft() (lattice.h:625-640) -- at
 * the nodes e_k = cos((k + 1/2) pi / N) its FFTW REDFT01 of g_n mu_n times
 * 1/sqrt(1 - e_k^2) -- at any energies; scale = -1: GreenFunctionImag's
 * operator() (lattice.h:617-623); scale = 1/(R pi): the DoS of
 * density_of_states_seq (functions.h:27-30).  g may be NULL (all ones); ctx
 * may be NULL (the process-wide context of device 0).  |e_j| >= 1 gives a
 * non-finite value, as the reference's at e = -1.
 * e = NULL: the n_e nodes theta_k = (k + 1/2) pi / n_e, with theta_k and
 * sin(theta_k) formed directly (no rounding of e_k enters near the edges):
 * fft(n) with n_e = n N, the moments zero-padded past N.
 */
int tbk_cheb_reconstruct(tbk_context* ctx, const double* mu, const double* g, int32_t N, const double* e,
                         int32_t n_e, double scale, double* out, double* kernel_ms);

/* ------------------------------------------------------ spectral range */
/*
 * Extreme eigenvalues of a Hermitian matrix, the E_min, E_max of the
 * rescaling.  Replaces numerical::linalg::matrix_spectral_range_sparse
 * (numerical.h:23-46), whose Spectra solver refuses more than 5000 rows and
 * then returns {0, 0}.  FP64 Lanczos on the device from a seeded
 * random-phase unit vector, no reorthogonalisation; after every 16 steps the
 * extreme Ritz values theta of the tridiagonal matrix (Sturm bisection) and
 * their residual bounds r = beta_{k+1} |y_k| are formed on the host.  Stops
 * when both r <= tol * max(|theta_min|, |theta_max|, theta_max - theta_min),
 * at an invariant subspace, or after min(max_iters, rows) steps.
 * An eigenvalue of M lies within resid_min of ritz_min (and within
 * resid_max of ritz_max), and ritz_min >= E_min, ritz_max <= E_max; the
 * Gershgorin interval holds the whole spectrum.  lower/upper =
 * max(gersh_min, ritz_min - resid_min), min(gersh_max, ritz_max + resid_max).
 */
typedef struct tbk_bounds {
    double ritz_min, ritz_max;    /* extreme Ritz values */
    double resid_min, resid_max;  /* their residual bounds */
    double gersh_min, gersh_max;  /* Gershgorin interval */
    double lower, upper;          /* the estimate of [E_min, E_max] (see above) */
    int32_t iters;                /* Lanczos steps taken */
    int32_t converged;            /* 1: both residuals met tol (or an invariant subspace) */
} tbk_bounds;

int tbk_spectral_bounds(tbk_context* ctx, const tbk_matrix* M, int32_t max_iters, double tol, uint64_t seed,
                        tbk_bounds* out);

/*
 * One-shot form with the reference's signature semantics
 * (matrix_spectral_range_sparse, numerical.h:23-46, both overloads): the
 * extreme Ritz values after at most 1000 steps at tol 1e-8, on the
 * process-wide context of device 0; any size.
 */
int tbk_spectral_range(const tbk_sparse* M, double* e_min, double* e_max);

/* ------------------------------------------- batched ragged moment problems */
typedef struct tbk_batch tbk_batch;

enum {
    TBK_BATCH_GRAPHS = 1,    /* capture each chunk of 64 steps in a hipGraph, replay on later runs */
    /* each chunk of 64 steps as one persistent launch: every problem is placed
     * on one XCD, its rows cut into shares of 4096; the workgroups on an XCD
     * draw (step, problem, share) items from that XCD's queue in step order,
     * an item waiting (bounded) until its problem's previous step is done.
     * The moments are summed in a fixed order after the run, so a problem's
     * moments do not depend on the rest of the batch; they differ from the
     * per-step launches' only in the grouping of the partial sums.  A wait
     * that gives up, or a queue left unserved, fails the run with
     * TBK_E_DEVICE. */
    TBK_BATCH_PERSISTENT = 2,
    /* each chunk of 64 steps as one persistent launch over the whole GPU: the
     * launch's workgroups (all resident) do the work of the per-step launches
     * step after step, each wave on its own, with a grid-wide barrier
     * between steps, and form each step's moments after its barrier from one
     * partial per wave.  The vectors are the per-step launches'; the moments
     * differ from theirs only in the grouping of the partial sums and do not
     * depend on the rest of the batch.  A barrier that gives up fails the run
     * with TBK_E_DEVICE.  Not combined with TBK_BATCH_PERSISTENT. */
    TBK_BATCH_SYNCED = 4
};

typedef struct tbk_batch_stats {
    double kernel_ms;      /* device time from the first step launch to the last moment launch */
    double wall_ms;        /* host wall time of the call */
    int64_t launches;      /* batched step launches (the largest step count in the batch), or
                              persistent launches (one per chunk of 64 steps) */
    int64_t graphs;        /* chunk graphs launched (0 without TBK_BATCH_GRAPHS) */
    double bytes_total;    /* algorithmic HBM bytes of every problem's steps */
    int64_t rows_total;    /* rows of all problems */
    int64_t workgroups;    /* workgroups of a step launch with every problem active */
    int32_t cpl;           /* complex values per lane (lanes per row = padded R / cpl) */
    int32_t blocks_per_cu; /* workgroups per CU a large problem gets */
} tbk_batch_stats;

/*
 * A batch of independent moment problems run as one block-diagonal
 * recursion (no reference counterpart: the reference solves one Hamiltonian
 * per call, numerics_cuda.cu:190-323).  Problem p: matrix mats[p] (CSC or CSR,
 * as tbk_matrix_create), rescaling H~ = (H - b[p])/a[p] fused into the step,
 * n_moments[p] >= 2 moments, R random-phase vectors generated on the device
 * from seeds[p] exactly as tbk_block_random_phase(rows_p, R, seeds[p]).  All
 * problems share R, the precision and real-or-complex values.  Bras = kets:
 * two moments per step.  Each problem's moments are independent of the other
 * problems in the batch (bit for bit) and its vectors equal those of a
 * single-problem run.
 */
int tbk_batch_create(tbk_context* ctx, int32_t n_problems, const tbk_sparse* mats, const double* a,
                     const double* b, const int32_t* n_moments, int32_t R, int32_t precision,
                     const uint64_t* seeds, tbk_batch** out);
/* run the whole batch from fresh kets; mu_out: the problems' moments
   concatenated in problem order (sum of n_moments doubles) */
int tbk_batch_run(tbk_batch* batch, int32_t flags, double* mu_out, tbk_batch_stats* stats);
/* launch shape: complex values per lane (1, 2 or 4) and the workgroups per CU
   a large problem gets (each problem splits its rows over its own workgroups,
   dealt per XCD, as the single-problem step does) */
int tbk_batch_set_shape(tbk_batch* batch, int32_t cpl, int32_t blocks_per_cu);
/* time `steps` steps per launch shape, keep the fastest, return its time per step */
int tbk_batch_autotune(tbk_batch* batch, int32_t steps, double* best_ms_per_step);
int tbk_batch_destroy(tbk_batch* batch);

/* ------------------------------------------------------ request server */
/*
 * A serving loop on one device (no reference counterpart: the reference
 * runs one solve per call).  Requests are copied at submit and queued; a
 * worker thread admits them while others run, as soon as their device-memory
 * footprint fits the budget and a slot is free; DoS requests of one class
 * (precision, real/complex, padded R) advance together in batched chunks,
 * sigma_xy requests run on a second stream beside them.  A DoS result is the
 * same bits alone and in any mix of requests.  All functions are thread-safe.
 */
typedef struct tbk_server tbk_server;

enum { TBK_REQ_DOS = 0, TBK_REQ_KUBO = 1 };
enum { TBK_REQ_QUEUED = 0, TBK_REQ_RUNNING = 1, TBK_REQ_DONE = 2, TBK_REQ_FAILED = 3 };

typedef struct tbk_server_config {
    int64_t hbm_budget_bytes;  /* device memory for requests (0: 85% of the free memory at creation): the
                                  arena that every admitted request's matrices, vector blocks, stacks and
                                  Gamma buffers are carved from (see lazy_arena) */
    int32_t max_slots;         /* DoS requests on the device at once (0: 512) */
    int32_t max_kubo;          /* sigma_xy requests on the device at once (0: 1) */
    int32_t chunk_steps;       /* steps per batched chunk (0: 32) */
    int32_t cpl;               /* complex values per lane of the batched launches (0: automatic) */
    int32_t max_moments;       /* largest DoS moment count accepted (0: 4096) */
    int32_t max_kubo_moments;  /* largest sigma_xy N accepted (0: 1024) */
    double max_wait_bypass_ms; /* a request queued longer than this is not overtaken (0: 1000) */
    int32_t lazy_arena;        /* 0: the whole budget is allocated at creation (one allocation; serving then
                                  allocates nothing); 1: slabs are allocated as requests need them (the
                                  first of 4 GiB, each later one at least as large as all before), kept,
                                  and freed only to make room for a larger one */
    int32_t reserved;
} tbk_server_config;

typedef struct tbk_request {
    int32_t kind;      /* TBK_REQ_DOS or TBK_REQ_KUBO */
    int32_t precision; /* DoS: TBK_FP64 or TBK_FP32; sigma_xy: TBK_GAMMA_FP64 or TBK_GAMMA_BF16X3 */
    int32_t n_moments; /* DoS: moments (>= 2); sigma_xy: N of the N x N Gamma */
    int32_t R;         /* random-phase vectors, generated as tbk_block_random_phase(rows, R, seed) */
    uint64_t seed;
    double a, b;       /* H~ = (H - b)/a, fused into the steps.  DoS: a = 0 asks the server for the
                          bounds: Lanczos (tbk_spectral_bounds, 200 steps, tol 1e-3) gives
                          [lower, upper], then a = (upper - lower)/1.99, b = (upper + lower)/2,
                          the contraction of GreenFunction(N, HE, ref_HE, 0.01) (lattice.h:560-569) */
    const tbk_sparse* H;
    const tbk_sparse* vx; /* sigma_xy: velocity operators, used as given */
    const tbk_sparse* vy;
    const double* g;   /* sigma_xy: kernel weights g_n (N of them) or NULL for ones */
    const tbk_lattice_desc* lattice; /* with H = NULL: the Hamiltonian as a lattice description,
                                        assembled on the device at admission (tbk_matrix_assemble).
                                        sigma_xy: vx = vy = NULL; the server forms
                                        H~ = (H - b)/a (tbk_matrix_rescale) and v_x, v_y of H~
                                        along axes 0 and 1 (tbk_matrix_velocity; atom_pos set) */
} tbk_request;

typedef struct tbk_request_info {
    int32_t state;          /* TBK_REQ_* */
    int32_t error_code;     /* TBK_E_* of a failed request */
    double queue_ms;        /* submit to admission */
    double run_ms;          /* admission to completion */
    double latency_ms;      /* submit to completion (so far, if not finished) */
    int64_t footprint_bytes;
    int64_t result_doubles; /* DoS: n moments; sigma_xy: 2N (zeta at the N Chebyshev nodes) + 2N^2 (Gamma) */
    double a, b;            /* the rescaling used (estimated at admission for a DoS request with a = 0) */
    int32_t vector_groups;  /* sigma_xy: groups the random vectors run in (1: the whole block at once) */
    int32_t row_ranges;     /* sigma_xy, fp64 stacks: row ranges the stacks hold one at a time (1: whole entries) */
} tbk_request_info;

typedef struct tbk_server_stats {
    int64_t submitted, completed, failed, queued;
    int64_t running_dos, running_kubo, kubo_done, chunks;
    int64_t used_bytes, budget_bytes;
    int64_t arena_bytes, arena_slabs; /* device memory the arena holds, in how many slabs */
    int64_t cancelled;                /* requests ended by tbk_server_cancel (counted in failed too) */
} tbk_server_stats;

int tbk_server_create(int32_t device, const tbk_server_config* cfg, tbk_server** out);
/* copy the request's inputs and queue it; TBK_E_OOM if its footprint exceeds the budget.
   Uploaded matrices are checked: every value finite, and H (DoS and sigma_xy) Hermitian to
   1e-12 of its largest magnitude, same pattern (TBK_E_INVALID otherwise; the moments' doubling
   identities and the spectral range need a Hermitian H).  A lattice's hopping table, phase
   table and disorder must be finite; its Hermiticity is the lattice's own. */
int tbk_server_submit(tbk_server* s, const tbk_request* r, int64_t* ticket);
/* Cancel a queued or running request: it ends TBK_REQ_FAILED with TBK_E_CANCELLED.  A queued
   request leaves the queue at once, a running DoS request at the end of the chunk in flight,
   a running sigma_xy request when its enqueued work ends (its result is dropped).  TBK_E_INVALID
   for an unknown ticket or a request already finished. */
int tbk_server_cancel(tbk_server* s, int64_t ticket);
int tbk_server_info(tbk_server* s, int64_t ticket, tbk_request_info* info);
/* block until the request is done or failed, or timeout_s passes (< 0: no limit) */
int tbk_server_wait(tbk_server* s, int64_t ticket, double timeout_s, tbk_request_info* info);
/* copy a finished request's result (DoS: the n moments; sigma_xy: zeta_k at
   e_k = cos((k + 1/2) pi / N), k < N, as complex128, then the weighted N x N
   Gamma row-major complex128) and forget the ticket; a failed request returns
   its error code */
int tbk_server_take(tbk_server* s, int64_t ticket, double* out, int64_t capacity, tbk_request_info* info);
int tbk_server_get_stats(tbk_server* s, tbk_server_stats* st);
/* finish every accepted request, then stop */
int tbk_server_destroy(tbk_server* s);

#ifdef __cplusplus
}
#endif

#endif /* TBKPM_H */
