/*
 * CPU oracle for the Chebyshev moment path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates the reference's CPU algorithm so that the HIP engine can
 * be checked against it (tests, the smoke check, and the CPU-baseline leg of
 * the benchmark).  It is never linked into or called by the product library.
 *
 * Restated functions (reference file:line):
 *   oracle_cheb_braket     numerical.h:363-381  Cheb_first_kind_seq_braket
 *   oracle_spmm_csc        numerical.h:368,374 via Eigen's sparse(ColMajor) x
 *                          dense product: res = 0; for each block column c,
 *                          for each matrix column j ascending, rhs_j =
 *                          (1+0i)*x(j,c); res(row,c) += value*rhs_j, with
 *                          value = 2*M_ij inside the recursion (2.0 * M * T)
 *   oracle_trace_braket    numerical.h:127-135  Re tr(bras^H kets), rows
 *                          outer, columns inner, products rounded then added
 *   oracle_cheb_on_bases   numerical.h:334-344  stored blocks T_n(M) V
 *   oracle_cheb_steps_threads  the same recursion with each step's product
 *                          and trace split over the block columns across
 *                          OpenMP threads (all-cores timing variant; the
 *                          moments differ from the single-thread loop only
 *                          in the summation order of the trace)
 *   oracle_kernel_*        numerical.h:393-426  Jackson / Lorentz / Fejer
 *
 * Arithmetic is written out in real operations in the order the reference's
 * std::complex / Eigen expressions evaluate, and the file is compiled with
 * -ffp-contract=off so no product is fused into an add.
 *
 * Blocks are column-major rows x R, complex128 interleaved (re, im).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double re, im;
} cplx;

static inline cplx cmul(cplx a, cplx b) {
    cplx r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

/* res = scale * M x over a column-major block (Eigen sparse x dense, ColMajor) */
static void oracle_spmm_csc(int is_complex, int32_t rows, const int32_t* colptr, const int32_t* rowidx,
                            const double* val, double scale, const cplx* x, int32_t R, cplx* res) {
    const cplx one = {1.0, 0.0};
    memset(res, 0, sizeof(cplx) * (size_t)rows * R);
    for (int32_t c = 0; c < R; ++c) {
        const cplx* xc = x + (size_t)c * rows;
        cplx* rc = res + (size_t)c * rows;
        for (int32_t j = 0; j < rows; ++j) {
            const cplx rhs_j = cmul(one, xc[j]);
            for (int32_t p = colptr[j]; p < colptr[j + 1]; ++p) {
                const int32_t i = rowidx[p];
                cplx prod;
                if (is_complex) {
                    cplx v;
                    v.re = scale * val[2 * (size_t)p];
                    v.im = scale * val[2 * (size_t)p + 1];
                    prod = cmul(v, rhs_j);
                } else {
                    const double v = scale * val[p];
                    prod.re = v * rhs_j.re;
                    prod.im = v * rhs_j.im;
                }
                rc[i].re = rc[i].re + prod.re;
                rc[i].im = rc[i].im + prod.im;
            }
        }
    }
}

double oracle_trace_braket(const double* bras, const double* kets, int32_t rows, int32_t R) {
    const cplx* b = (const cplx*)bras;
    const cplx* k = (const cplx*)kets;
    double trace = 0.0;
    for (int32_t i = 0; i < rows; ++i)
        for (int32_t j = 0; j < R; ++j) {
            const cplx bb = b[(size_t)j * rows + i], kk = k[(size_t)j * rows + i];
            trace += bb.re * kk.re + bb.im * kk.im;
        }
    return trace;
}

/*
 * mu[i] = Re tr(bras^H T_i(M) kets), i < n; n >= 2 (the reference writes
 * mu[1] without a check, numerical.h:366-371).  If last != NULL the block
 * T_{last_index} (0 <= last_index < n) is copied there.
 * Returns 0, or -1 on bad arguments / allocation failure.
 */
int oracle_cheb_braket(int is_complex, int32_t rows, const int32_t* colptr, const int32_t* rowidx,
                       const double* val, int32_t n, const double* bras, const double* kets, int32_t R,
                       double* mu, double* last, int32_t last_index) {
    if (n < 2 || rows <= 0 || R <= 0) return -1;
    const size_t blk = (size_t)rows * R;
    cplx* Tn_2 = (cplx*)malloc(sizeof(cplx) * blk);
    cplx* Tn_1 = (cplx*)malloc(sizeof(cplx) * blk);
    cplx* Tn = (cplx*)malloc(sizeof(cplx) * blk);
    if (!Tn_2 || !Tn_1 || !Tn) {
        free(Tn_2);
        free(Tn_1);
        free(Tn);
        return -1;
    }
    memcpy(Tn_2, kets, sizeof(cplx) * blk);
    oracle_spmm_csc(is_complex, rows, colptr, rowidx, val, 1.0, Tn_2, R, Tn_1);
    mu[0] = oracle_trace_braket(bras, kets, rows, R);
    mu[1] = oracle_trace_braket(bras, (const double*)Tn_1, rows, R);
    if (last && last_index == 0) memcpy(last, kets, sizeof(cplx) * blk);
    if (last && last_index == 1) memcpy(last, Tn_1, sizeof(cplx) * blk);
    for (int32_t i = 2; i < n; ++i) {
        /* Tn = 2.0 * M * Tn_1 - Tn_2: product into a temporary, then subtract */
        oracle_spmm_csc(is_complex, rows, colptr, rowidx, val, 2.0, Tn_1, R, Tn);
        for (size_t p = 0; p < blk; ++p) {
            Tn[p].re = Tn[p].re - Tn_2[p].re;
            Tn[p].im = Tn[p].im - Tn_2[p].im;
        }
        mu[i] = oracle_trace_braket(bras, (const double*)Tn, rows, R);
        if (last && last_index == i) memcpy(last, Tn, sizeof(cplx) * blk);
        cplx* t = Tn_2;
        Tn_2 = Tn_1;
        Tn_1 = Tn;
        Tn = t;
    }
    free(Tn_2);
    free(Tn_1);
    free(Tn);
    return 0;
}

/*
 * Timing leg of the CPU baseline: runs `steps` recursion steps of the same
 * loop (fresh temporary per step as numerical.h:374 allocates one) and
 * returns the moments of those steps.  steps >= 1.
 */
int oracle_cheb_steps(int is_complex, int32_t rows, const int32_t* colptr, const int32_t* rowidx,
                      const double* val, int32_t steps, const double* kets, int32_t R, double* mu) {
    return oracle_cheb_braket(is_complex, rows, colptr, rowidx, val, steps + 2, kets, kets, R, mu, NULL, 0);
}

/*
 * All-cores timing variant of oracle_cheb_steps: every step's product,
 * subtraction and trace run over the block columns in parallel (one column's
 * work is the single-thread loop's for that column); the per-column traces
 * are added in column order.  nthreads <= 0: the OpenMP default.
 */
int oracle_cheb_steps_threads(int is_complex, int32_t rows, const int32_t* colptr, const int32_t* rowidx,
                              const double* val, int32_t steps, const double* kets, int32_t R, double* mu,
                              int32_t nthreads) {
    const int32_t n = steps + 2;
    if (steps < 1 || rows <= 0 || R <= 0) return -1;
    const size_t blk = (size_t)rows * R;
    cplx* T0 = (cplx*)malloc(sizeof(cplx) * blk);
    cplx* T1 = (cplx*)malloc(sizeof(cplx) * blk);
    cplx* T2 = (cplx*)malloc(sizeof(cplx) * blk);
    double* part = (double*)malloc(sizeof(double) * (size_t)R);
    if (!T0 || !T1 || !T2 || !part) {
        free(T0);
        free(T1);
        free(T2);
        free(part);
        return -1;
    }
    if (nthreads > 0) omp_set_num_threads(nthreads);
    memcpy(T0, kets, sizeof(cplx) * blk);
    cplx* Tn_2 = T0;
    cplx* Tn_1 = T1;
    cplx* Tn = T2;
    const cplx* k = (const cplx*)kets;
    for (int32_t i = 1; i < n; ++i) {
        const double scale = i == 1 ? 1.0 : 2.0;
        const cplx* src = i == 1 ? Tn_2 : Tn_1;
        cplx* dst = i == 1 ? Tn_1 : Tn;
#pragma omp parallel for schedule(static)
        for (int32_t c = 0; c < R; ++c) {
            oracle_spmm_csc(is_complex, rows, colptr, rowidx, val, scale, src + (size_t)c * rows, 1,
                            dst + (size_t)c * rows);
            cplx* d = dst + (size_t)c * rows;
            if (i > 1) {
                const cplx* z = Tn_2 + (size_t)c * rows;
                for (int32_t r = 0; r < rows; ++r) {
                    d[r].re = d[r].re - z[r].re;
                    d[r].im = d[r].im - z[r].im;
                }
            }
            double t = 0.0;
            const cplx* kc = k + (size_t)c * rows;
            for (int32_t r = 0; r < rows; ++r) t += kc[r].re * d[r].re + kc[r].im * d[r].im;
            part[c] = t;
        }
        double s = 0.0;
        for (int32_t c = 0; c < R; ++c) s += part[c];
        mu[i] = s;
        if (i > 1) {
            cplx* t = Tn_2;
            Tn_2 = Tn_1;
            Tn_1 = Tn;
            Tn = t;
        }
    }
    mu[0] = oracle_trace_braket(kets, kets, rows, R);
    free(T0);
    free(T1);
    free(T2);
    free(part);
    return 0;
}

/* Stored blocks T_0..T_{n-1} of M applied to V (numerical.h:334-344); out is
   n consecutive column-major blocks. */
int oracle_cheb_on_bases(int is_complex, int32_t rows, const int32_t* colptr, const int32_t* rowidx,
                         const double* val, int32_t n, const double* V, int32_t R, double* out) {
    if (n < 1 || rows <= 0 || R <= 0) return -1;
    const size_t blk = (size_t)rows * R;
    cplx* T = (cplx*)out;
    memcpy(T, V, sizeof(cplx) * blk);
    if (n == 1) return 0;
    oracle_spmm_csc(is_complex, rows, colptr, rowidx, val, 1.0, T, R, T + blk);
    for (int32_t i = 2; i < n; ++i) {
        cplx* Ti = T + (size_t)i * blk;
        oracle_spmm_csc(is_complex, rows, colptr, rowidx, val, 2.0, T + (size_t)(i - 1) * blk, R, Ti);
        const cplx* Tm2 = T + (size_t)(i - 2) * blk;
        for (size_t p = 0; p < blk; ++p) {
            Ti[p].re = Ti[p].re - Tm2[p].re;
            Ti[p].im = Ti[p].im - Tm2[p].im;
        }
    }
    return 0;
}

/* kernels (numerical.h:400-426) */
void oracle_kernel_jackson(int32_t N, double* g) {
    const int32_t N1 = N + 1;
    const double pi = 3.14159265358979323846;
    for (int32_t n = 0; n < N; ++n) {
        const double s1 = pi / N1;
        const double s2 = n * s1;
        g[n] = ((N1 - n) * cos(s2) + sin(s2) / tan(s1)) / N1;
    }
}
void oracle_kernel_lorentz(int32_t N, double lambda, double* g) {
    for (int32_t n = 0; n < N; ++n) g[n] = sinh(lambda * (1.0 - (double)n / N)) / sinh(lambda);
}
void oracle_kernel_fejer(int32_t N, double* g) {
    for (int32_t n = 0; n < N; ++n) g[n] = 1.0 - (double)n / N;
}
