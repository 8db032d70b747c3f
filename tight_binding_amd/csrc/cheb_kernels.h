// Device kernels of the Chebyshev moment recursion (gfx950, wave64).
//
// One step applies H~ = (H - bI)/a to a whole block of R vectors stored
// row-major [site][Rp] and fuses, per site row r:
//     acc  = sum_{p in row r, ascending column} H[p] * x[col[p]]
//     y_r  = alpha*acc + beta*x_r - z_r          (three-term update)
//     partial moments from y_r, x_r and (optionally) a bra row
// A row is handled by LPR lanes, each owning CPL columns in 16-byte units
// interleaved across the lanes (lane_unit_elems), so the gather of a
// neighbour row is one coalesced access per unit.
// Workgroups take tiles of consecutive rows; with xcd_split the tiles are
// dealt so that each XCD (blocks b with equal b % 8) walks one contiguous
// eighth of the rows, keeping a lattice neighbourhood inside one L2.
//
// Rounding: this file is compiled without floating-point contraction, so
// each product is rounded and then added -- the operation order of the
// CSC scatter-axpy in numerical.h:368-374 (Eigen sparse x dense, columns
// ascending).  With alpha = 2, beta = 0 (a = 1, b = 0) the fp64 vectors are
// bit-identical to that CPU recursion.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "split_bf16.h"

namespace tbk {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

enum StepMode : int {
    kFirst = 0,   // T1 = H~ T0;           partials (Re<bra|T0>, Re<bra|T1>)
    kRec = 1,     // T_{k+1} = 2H~T_k - T_{k-1}; partial Re<bra|T_{k+1}>
    kRecDbl = 2,  // same update;          partials (Re<T_k|T_k>, Re<T_{k+1}|T_k>)
    kDual = 3,    // kRecDbl, and y2 = H2 T_k for a second matrix H2 with the same
                  // pattern (values hv2), formed from the same gathered rows
    // pre-split stacks (fp64 complex only): the output also, or instead,
    // stored as split-bf16 words (split_word) at y2 viewed as uint2
    kFirstSplit = 4,   // kFirst, and y's split words at y2
    kRecDblSplit = 5,  // kRecDbl, and y's split words at y2
    kDualSplit = 6,    // kDual with H2 T_k stored only as split words at y2
    // paired chains (complex matrices; fp64 or fp32 vectors): columns
    // [0, Rp/2) and [Rp/2, Rp) of the block are two recursions on the same
    // matrix; the left half's y is stored as split words at y2, the right
    // half's H2 x as split words at y3, both in blocks of Rp/2 columns
    kPairFirst = 7,  // kFirst form
    kPairRec = 8,    // kRecDbl form
};

template <typename VT>
struct Vec2;
template <>
struct Vec2<double> {
    using type = double2;
};
template <>
struct Vec2<float> {
    using type = float2;
};

template <typename VT, bool CH>
struct HVal {
    using type = VT;
};
template <typename VT>
struct HVal<VT, true> {
    using type = typename Vec2<VT>::type;
};

__device__ __forceinline__ void madd(double2& acc, double h, double2 x) {
    acc.x = acc.x + h * x.x;
    acc.y = acc.y + h * x.y;
}
__device__ __forceinline__ void madd(float2& acc, float h, float2 x) {
    acc.x = acc.x + h * x.x;
    acc.y = acc.y + h * x.y;
}
// (h.re x.re - h.im x.im, h.re x.im + h.im x.re), then added
__device__ __forceinline__ void madd(double2& acc, double2 h, double2 x) {
    double re = h.x * x.x - h.y * x.y;
    double im = h.x * x.y + h.y * x.x;
    acc.x = acc.x + re;
    acc.y = acc.y + im;
}
__device__ __forceinline__ void madd(float2& acc, float2 h, float2 x) {
    float re = h.x * x.x - h.y * x.y;
    float im = h.x * x.y + h.y * x.x;
    acc.x = acc.x + re;
    acc.y = acc.y + im;
}
// Re(conj(a) b) accumulated in fp64
__device__ __forceinline__ double redot(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ double redot(float2 a, float2 b) {
    return static_cast<double>(a.x) * static_cast<double>(b.x) +
           static_cast<double>(a.y) * static_cast<double>(b.y);
}

// Step options (bit flags of the kernels' `opts` argument).
constexpr int kOptShift = 1;   // beta*x_r term present (b != 0)
constexpr int kOptStream = 2;  // non-temporal access for the once-used rows

// Rows read or written once per step (T_{k-1}, bra, T_{k+1}) may bypass
// cache retention so that L2 keeps the gathered neighbour rows.
template <typename V2>
__device__ __forceinline__ V2 ld_stream(const V2* p, int opts) {
    using E = decltype(p->x);
    typedef E vec2 __attribute__((ext_vector_type(2)));
    if (opts & kOptStream) {
        const vec2 v = __builtin_nontemporal_load(reinterpret_cast<const vec2*>(p));
        V2 r;
        r.x = v.x;
        r.y = v.y;
        return r;
    }
    return *p;
}
template <typename V2>
__device__ __forceinline__ void st_stream(V2* p, V2 v, int opts) {
    using E = decltype(v.x);
    typedef E vec2 __attribute__((ext_vector_type(2)));
    if (opts & kOptStream) {
        vec2 w;
        w.x = v.x;
        w.y = v.y;
        __builtin_nontemporal_store(w, reinterpret_cast<vec2*>(p));
    } else {
        *p = v;
    }
}

// CPL consecutive values at p (the row kernel passes one 16-byte unit of a
// lane, lane_unit_elems values).  8-byte elements (complex64) go in pairs
// through 16-byte accesses: p is 16-byte aligned whenever CPL is even, since
// unit offsets are multiples of CPL and rows multiples of Rp.
template <int CPL, typename V2>
__device__ __forceinline__ void ld_lane(const V2* p, V2* out) {
    if constexpr (sizeof(V2) == 8 && CPL % 2 == 0) {
        typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int c = 0; c < CPL; c += 2) {
            const f4 w = *reinterpret_cast<const f4*>(p + c);
            out[c].x = w.x;
            out[c].y = w.y;
            out[c + 1].x = w.z;
            out[c + 1].y = w.w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < CPL; ++c) out[c] = p[c];
    }
}
template <int CPL, typename V2>
__device__ __forceinline__ void ld_lane_stream(const V2* p, V2* out, int opts) {
    if constexpr (sizeof(V2) == 8 && CPL % 2 == 0) {
        typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int c = 0; c < CPL; c += 2) {
            const f4* q = reinterpret_cast<const f4*>(p + c);
            const f4 w = (opts & kOptStream) ? __builtin_nontemporal_load(q) : *q;
            out[c].x = w.x;
            out[c].y = w.y;
            out[c + 1].x = w.z;
            out[c + 1].y = w.w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < CPL; ++c) out[c] = ld_stream(p + c, opts);
    }
}
template <int CPL, typename V2>
__device__ __forceinline__ void st_lane_stream(V2* p, const V2* v, int opts) {
    if constexpr (sizeof(V2) == 8 && CPL % 2 == 0) {
        typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int c = 0; c < CPL; c += 2) {
            f4 w;
            w.x = v[c].x;
            w.y = v[c].y;
            w.z = v[c + 1].x;
            w.w = v[c + 1].y;
            f4* q = reinterpret_cast<f4*>(p + c);
            if (opts & kOptStream)
                __builtin_nontemporal_store(w, q);
            else
                *q = w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < CPL; ++c) st_stream(p + c, v[c], opts);
    }
}

// split-bf16 words of a complex value (fp64 or fp32 vectors), stored
// non-temporally (the stack entries are read only by the contraction)
template <typename V2>
__device__ __forceinline__ void st_split(uint2* p, V2 v) {
    typedef uint32_t u2v __attribute__((ext_vector_type(2)));
    const uint2 w = split_word(v);
    u2v u;
    u.x = w.x;
    u.y = w.y;
    __builtin_nontemporal_store(u, reinterpret_cast<u2v*>(p));
}
// two consecutive values' words as one 16-byte store (p 16-byte aligned)
template <typename V2>
__device__ __forceinline__ void st_split2(uint2* p, V2 v0, V2 v1) {
    typedef uint32_t u4v __attribute__((ext_vector_type(4)));
    const uint2 w0 = split_word(v0), w1 = split_word(v1);
    u4v u;
    u.x = w0.x;
    u.y = w0.y;
    u.z = w1.x;
    u.w = w1.y;
    __builtin_nontemporal_store(u, reinterpret_cast<u4v*>(p));
}
// one value per lane: the words of this lane's value and of its row
// neighbour's (lane ^ 1, by a quad permutation) as one 16-byte store from
// the even lane (p 16-byte aligned there).  Both lanes of a pair must be
// active: they hold one row.
template <typename V2>
__device__ __forceinline__ void st_split_lanes(uint2* p, V2 v, bool even) {
    typedef uint32_t u4v __attribute__((ext_vector_type(4)));
    const uint2 w = split_word(v);
    constexpr int kSwapNeighbour = 0xB1;  // quad_perm(1, 0, 3, 2)
    const uint32_t nx = static_cast<uint32_t>(
        __builtin_amdgcn_mov_dpp(static_cast<int>(w.x), kSwapNeighbour, 0xF, 0xF, false));
    const uint32_t ny = static_cast<uint32_t>(
        __builtin_amdgcn_mov_dpp(static_cast<int>(w.y), kSwapNeighbour, 0xF, 0xF, false));
    if (even) {
        u4v u;
        u.x = w.x;
        u.y = w.y;
        u.z = nx;
        u.w = ny;
        __builtin_nontemporal_store(u, reinterpret_cast<u4v*>(p));
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, kWave);
    return v;
}

// Visiting order of row tiles.  Row tiles are grouped in chunks of tpc
// tiles; with P > 1 the chunk grid (Q rows of P chunks) is walked column by
// column -- chunk q*P + p in the order p-major -- so that chunks P apart in
// the row numbering, which a matrix of bandwidth ~ P chunks couples, are
// visited one after the other: a neighbour row fetched for one chunk is
// reused by the next instead of ~2D rows later.  P = 1 is row order.  Visit
// indices past the last row tile are empty.  Slabs repeat the order.
struct TileOrder {
    uint32_t row_tiles;  // tiles covering the rows once
    uint32_t tpc;        // tiles per chunk
    uint32_t P, Q;       // chunk grid: P columns, Q rows (P*Q*tpc >= row_tiles)
};

__device__ __forceinline__ int64_t visit_row_tile(const TileOrder& o, uint32_t t, uint32_t& slab) {
    const uint32_t per = o.P * o.Q * o.tpc;
    slab = t / per;
    const uint32_t u = t - slab * per;
    const uint32_t cv = u / o.tpc, off = u - cv * o.tpc;
    const uint32_t p = cv / o.Q, q = cv - p * o.Q;
    const uint32_t rt = (q * o.P + p) * o.tpc + off;
    return rt < o.row_tiles ? static_cast<int64_t>(rt) : -1;
}

// Tile range of this workgroup (see file comment).
__device__ __forceinline__ void tile_range(int64_t ntiles, int xcd_split, int64_t& first,
                                           int64_t& end, int64_t& stride) {
    const int64_t nb = gridDim.x;
    const int64_t b = blockIdx.x;
    if (xcd_split) {
        const int64_t xcd = b & 7, j = b >> 3, per = nb >> 3;
        first = (ntiles * xcd) / 8 + j;
        end = (ntiles * (xcd + 1)) / 8;
        stride = per;
    } else {
        first = b;
        end = ntiles;
        stride = nb;
    }
}

// Columns of a row held by a lane.  A lane's CPL values form NU units of UE
// consecutive values, one 16-byte access each (a single unit when the lane
// holds at most 16 bytes); unit u of lane q starts at column
// (q + u * LPR) * UE.  Each access instruction of a wave then covers whole
// 16-byte-aligned runs of the row.  With a lane's units side by side instead,
// every store instruction would write half of each 32-byte sector, and the
// memory side sees about 1.5x the bytes.
template <typename V2, int CPL>
constexpr int lane_unit_elems() {
    return CPL * static_cast<int>(sizeof(V2)) > 16 ? 16 / static_cast<int>(sizeof(V2)) : CPL;
}
// Offset of unit u from the lane's first unit.
template <typename V2, int LPR, int CPL>
constexpr int unit_step(int u) {
    return u * LPR * lane_unit_elems<V2, CPL>();
}
// p: the lane's first unit in the row (row offset + q * UE)
template <int LPR, int CPL, typename V2>
__device__ __forceinline__ void ld_row(const V2* p, V2* out) {
    constexpr int UE = lane_unit_elems<V2, CPL>();
#pragma unroll
    for (int u = 0; u < CPL / UE; ++u) ld_lane<UE>(p + unit_step<V2, LPR, CPL>(u), out + u * UE);
}
template <int LPR, int CPL, typename V2>
__device__ __forceinline__ void ld_row_stream(const V2* p, V2* out, int opts) {
    constexpr int UE = lane_unit_elems<V2, CPL>();
#pragma unroll
    for (int u = 0; u < CPL / UE; ++u) ld_lane_stream<UE>(p + unit_step<V2, LPR, CPL>(u), out + u * UE, opts);
}
template <int LPR, int CPL, typename V2>
__device__ __forceinline__ void st_row_stream(V2* p, const V2* v, int opts) {
    constexpr int UE = lane_unit_elems<V2, CPL>();
#pragma unroll
    for (int u = 0; u < CPL / UE; ++u) st_lane_stream<UE>(p + unit_step<V2, LPR, CPL>(u), v + u * UE, opts);
}

// One site row r of a step (q: this lane's column group): the neighbour
// gathers, y_r = alpha*acc + beta*x_r - z_r, and the moment partials of the
// mode added to p0, p1.
// Units whose second product (the fused operator hv2) is used: all of them,
// except in the paired modes with two or more units per lane, whose first
// CPL / 2 units are left-half columns that store only the first product.
// Value c lies in unit c / UE of the lane's CPL / UE units; with one unit
// per lane a lane's values are all on one side and all form both products.
template <bool PAIR, int CPL, int UE>
__device__ __forceinline__ constexpr bool dual_unit(int c) {
    return !PAIR || CPL / UE < 2 || c / UE >= (CPL / UE) / 2;
}

template <typename VT, bool CH, int LPR, int CPL, int MODE>
__device__ __forceinline__ void step_row(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                         const typename HVal<VT, CH>::type* __restrict__ hv,
                                         const typename Vec2<VT>::type* __restrict__ x,
                                         const typename Vec2<VT>::type* z, typename Vec2<VT>::type* y,
                                         const typename Vec2<VT>::type* __restrict__ bra, int r, int q, int32_t Rp,
                                         VT alpha, VT beta, int32_t opts, double& p0, double& p1,
                                         const typename HVal<VT, CH>::type* __restrict__ hv2 = nullptr,
                                         typename Vec2<VT>::type* y2 = nullptr,
                                         typename Vec2<VT>::type* y3 = nullptr, int64_t wbs = 0) {
    using V2 = typename Vec2<VT>::type;
    using H = typename HVal<VT, CH>::type;
    // the update and moment form of the mode, and its extra outputs
    constexpr bool FIRST = MODE == kFirst || MODE == kFirstSplit || MODE == kPairFirst;
    constexpr bool PAIR = MODE == kPairFirst || MODE == kPairRec;
    constexpr bool DUAL = MODE == kDual || MODE == kDualSplit || PAIR;
    constexpr bool SPLIT_Y = MODE == kFirstSplit || MODE == kRecDblSplit;
    constexpr int UE = lane_unit_elems<V2, CPL>();
    // this lane's first unit in row r
    const size_t own = static_cast<size_t>(r) * Rp + static_cast<size_t>(q) * UE;

    V2 zv[CPL], bv[CPL], acc[CPL], acc2[CPL], xs[CPL], sv[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        acc[c].x = 0;
        acc[c].y = 0;
        acc2[c].x = 0;
        acc2[c].y = 0;
        xs[c].x = 0;
        xs[c].y = 0;
    }
    if (!FIRST) ld_row_stream<LPR, CPL>(z + own, zv, opts);
    // the first step with bra = kets (two moments per step) takes the bra row
    // from the gathered own row instead of reading the block a second time
    const bool bra_is_x = FIRST && bra == x;
    if ((FIRST || MODE == kRec) && !bra_is_x) ld_row_stream<LPR, CPL>(bra + own, bv, opts);
    const int s = rp[r], e = rp[r + 1];
    bool have_self = false;
    int p = s;
    for (; p + 4 <= e; p += 4) {
        int cc[4];
        H hh[4], hh2[4];
        V2 xv[4][CPL];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            cc[k] = ci[p + k];
            hh[k] = hv[p + k];
            if (DUAL) hh2[k] = hv2[p + k];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ld_row<LPR, CPL>(x + static_cast<size_t>(cc[k]) * Rp + static_cast<size_t>(q) * UE, xv[k]);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) madd(acc[c], hh[k], xv[k][c]);
            if (DUAL) {
#pragma unroll
                for (int c = 0; c < CPL; ++c)
                    if (dual_unit<PAIR, CPL, UE>(c)) madd(acc2[c], hh2[k], xv[k][c]);
            }
            if (cc[k] == r) {
                have_self = true;
#pragma unroll
                for (int c = 0; c < CPL; ++c) xs[c] = xv[k][c];
            }
        }
    }
    for (; p < e; ++p) {
        const int cc = ci[p];
        const H hh = hv[p];
        V2 xv[CPL];
        ld_row<LPR, CPL>(x + static_cast<size_t>(cc) * Rp + static_cast<size_t>(q) * UE, xv);
#pragma unroll
        for (int c = 0; c < CPL; ++c) madd(acc[c], hh, xv[c]);
        if (DUAL) {
            const H hh2 = hv2[p];
#pragma unroll
            for (int c = 0; c < CPL; ++c)
                if (dual_unit<PAIR, CPL, UE>(c)) madd(acc2[c], hh2, xv[c]);
        }
        if (cc == r) {
            have_self = true;
#pragma unroll
            for (int c = 0; c < CPL; ++c) xs[c] = xv[c];
        }
    }
    if (!have_self) ld_row<LPR, CPL>(x + own, xs);
    if (bra_is_x) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) bv[c] = xs[c];
    }

#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        V2 v;
        v.x = alpha * acc[c].x;
        v.y = alpha * acc[c].y;
        if (opts & kOptShift) {
            v.x = v.x + beta * xs[c].x;
            v.y = v.y + beta * xs[c].y;
        }
        if (!FIRST) {
            v.x = v.x - zv[c].x;
            v.y = v.y - zv[c].y;
        }
        sv[c] = v;
        if (FIRST) {
            p0 = p0 + redot(bv[c], xs[c]);
            p1 = p1 + redot(bv[c], v);
        } else if (MODE == kRec) {
            p0 = p0 + redot(bv[c], v);
        } else {
            p0 = p0 + redot(xs[c], xs[c]);
            p1 = p1 + redot(v, xs[c]);
        }
    }
    st_row_stream<LPR, CPL>(y + own, sv, opts);
    if constexpr (MODE == kDual) st_row_stream<LPR, CPL>(y2 + own, acc2, opts);
    // split words of the pre-split stacks (fp64: one value per unit).  With
    // two or more lanes per row, lanes q and q ^ 1 hold adjacent columns of
    // every unit, and the even lane stores both values' words as one 16-byte
    // store: rows and units start at even columns, so a pair never straddles
    // a 256-word block.
    if constexpr (SPLIT_Y || MODE == kDualSplit) {
        uint2* ws = reinterpret_cast<uint2*>(y2);
        if constexpr (LPR >= 2) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const int64_t w = static_cast<int64_t>(own) + unit_step<V2, LPR, CPL>(c);
                st_split_lanes(ws + word_offset(w, wbs), SPLIT_Y ? sv[c] : acc2[c], (q & 1) == 0);
            }
        } else {
            // one lane per row: its values are the row's consecutive columns,
            // and a row (CPL values, CPL dividing 256) lies in one word block
            ws += word_offset(static_cast<int64_t>(own), wbs);
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                if (CPL % 2 == 0 && (c & 1)) continue;
                const V2& a0 = SPLIT_Y ? sv[c] : acc2[c];
                if (CPL % 2 == 0) {
                    const V2& a1 = SPLIT_Y ? sv[c + 1] : acc2[c + 1];
                    st_split2(ws + c, a0, a1);
                } else {
                    st_split(ws + c, a0);
                }
            }
        }
    }
    // paired chains: the left half of the columns stores y, the right half
    // H2 x, as words in blocks of Rp / 2 columns
    if constexpr (PAIR) {
        const int half = Rp >> 1;
        if constexpr (CPL == 1) {
            const bool left = q < half;
            // the last step's left output is no stack entry: y2 null, not stored
            if (left && y2 == nullptr) return;
            const int64_t k = static_cast<int64_t>(r) * half + (left ? q : q - half);
            uint2* ws = (left ? reinterpret_cast<uint2*>(y2) : reinterpret_cast<uint2*>(y3)) + word_offset(k, wbs);
            if ((half & 1) == 0) {  // even halves: neighbour lanes' words in one 16-byte store
                st_split_lanes(ws, left ? sv[0] : acc2[0], (q & 1) == 0);
            } else {
                st_split(ws, left ? sv[0] : acc2[0]);
            }
        } else if constexpr (UE >= 2) {
            // complex64 pairs: unit u of a lane holds the adjacent columns
            // q UE + u LPR UE (+1), on one side of the block since the half
            // width is even; both values' words go out as one 16-byte store
            // (k even, so the pair lies in one word block)
            static_assert(UE == 2, "split words of 16-byte units of two values");
#pragma unroll
            for (int u = 0; u < CPL / UE; ++u) {
                const int col = q * UE + unit_step<V2, LPR, CPL>(u);
                const bool left = col < half;
                if (left && y2 == nullptr) continue;
                const int64_t k = static_cast<int64_t>(r) * half + (left ? col : col - half);
                uint2* ws =
                    (left ? reinterpret_cast<uint2*>(y2) : reinterpret_cast<uint2*>(y3)) + word_offset(k, wbs);
                const V2* src = (left ? sv : acc2) + u * UE;
                st_split2(ws, src[0], src[1]);
            }
        } else {
            // unit c of a lane sits at column q + c * LPR: the first CPL / 2
            // units of every lane are left-half columns, the others right
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const bool left = c < CPL / 2;
                if (left && y2 == nullptr) continue;
                const int col = q + unit_step<V2, LPR, CPL>(c);
                const int64_t k = static_cast<int64_t>(r) * half + (left ? col : col - half);
                uint2* ws =
                    (left ? reinterpret_cast<uint2*>(y2) : reinterpret_cast<uint2*>(y3)) + word_offset(k, wbs);
                if constexpr (LPR >= 2) {
                    st_split_lanes(ws, left ? sv[c] : acc2[c], (q & 1) == 0);
                } else {
                    st_split(ws, left ? sv[c] : acc2[c]);
                }
            }
        }
    }
}

template <typename VT, bool CH, int LPR, int CPL, int MODE>
__global__ __launch_bounds__(kBlock) void cheb_step_kernel(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
    const typename HVal<VT, CH>::type* __restrict__ hv, const typename Vec2<VT>::type* __restrict__ x,
    const typename Vec2<VT>::type* z, typename Vec2<VT>::type* y,
    const typename Vec2<VT>::type* __restrict__ bra, int32_t rows, int32_t Rp, VT alpha, VT beta,
    int32_t opts, int64_t ntiles, TileOrder order, int32_t xcd_split, double2* __restrict__ partial,
    const typename HVal<VT, CH>::type* __restrict__ hv2, typename Vec2<VT>::type* y2,
    typename Vec2<VT>::type* y3, int64_t wbs) {
    constexpr int RPW = kWave / LPR;
    constexpr int TR = kWavesPerBlock * RPW;
    static_assert(LPR >= 1 && LPR <= kWave && (kWave % LPR) == 0, "bad lanes per row");

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int sub = lane / LPR;
    const int q = lane % LPR;

    int64_t first, end, stride;
    tile_range(ntiles, xcd_split, first, end, stride);

    // a sweep covers LPR * CPL columns; with Rp wider, slab s of the visiting
    // order is columns [s LPR CPL, (s + 1) LPR CPL), i.e. lanes shifted by
    // s LPR CPL / UE units of UE values
    using V2 = typename Vec2<VT>::type;
    constexpr int UE = lane_unit_elems<V2, CPL>();
    constexpr int kSlabUnits = LPR * CPL / UE;
    double p0 = 0.0, p1 = 0.0;
    for (int64_t t = first; t < end; t += stride) {
        uint32_t slab;
        const int64_t rt = visit_row_tile(order, static_cast<uint32_t>(t), slab);
        if (rt < 0) continue;
        int r = static_cast<int>(rt * TR) + wave * RPW + sub;
        if (LPR == kWave) r = __builtin_amdgcn_readfirstlane(r);
        if (r >= rows) continue;
        step_row<VT, CH, LPR, CPL, MODE>(rp, ci, hv, x, z, y, bra, r, q + static_cast<int>(slab) * kSlabUnits, Rp,
                                         alpha, beta, opts, p0, p1, hv2, y2, y3, wbs);
    }

    // deterministic block reduction: butterfly in the wave, then waves in order
    __shared__ double red[2][kWavesPerBlock];
    p0 = wave_sum(p0);
    p1 = wave_sum(p1);
    if (lane == 0) {
        red[0][wave] = p0;
        red[1][wave] = p1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = red[0][0], s1 = red[1][0];
        for (int w = 1; w < kWavesPerBlock; ++w) {
            s0 = s0 + red[0][w];
            s1 = s1 + red[1][w];
        }
        partial[blockIdx.x] = make_double2(s0, s1);
    }
}

// Resident problems of a serving loop ("slots"), each with its own CSR
// arrays and three rotating vector blocks (T_k in blk[k % 3]); the slots
// of one class (precision, value type, padded width) advance together, one
// launch per step.  k0 is the slot's step index at the chunk's first launch,
// so the launch with offset s performs step k0 + s of every slot that has
// not finished (k0 + s < steps).  Step 0 is the first step (T_1, moments 0
// and 1).  Slot d owns the nwg workgroups wg0 .. wg0 + nwg - 1 of a launch
// (wg0 and nwg multiples of 8, so its workgroup j runs on XCD j % 8) and
// splits its row tiles over them exactly as cheb_step_kernel splits a
// matrix: tile ranges dealt per XCD, visited in the slot's chunk order.
struct SlotDesc {
    const int32_t* rp;
    const int32_t* ci;
    const void* hv;
    void* blk[3];
    double coef[4];  // (1/a, -b/a, 2/a, -2b/a)
    double* mu;      // n moments
    int64_t k0;
    int64_t steps;   // (n + 1) / 2
    int64_t wg0;
    int32_t nwg;
    int32_t rows;
    int32_t n;
    int32_t pad_;
    TileOrder order;
};

// Rows of workgroup index w of a slot launch at step offset s (step
// d.k0 + s of the slot) that wave `wave` owns: the step on those rows, and
// the wave's moment sums (all lanes hold them).  False (for the whole
// workgroup) once the slot is done.  No barrier.
template <typename VT, bool CH, int LPR, int CPL>
__device__ __forceinline__ bool slots_wave_rows(const SlotDesc* __restrict__ slots,
                                                const int32_t* __restrict__ wg_slot, int64_t w, int wave, int32_t s,
                                                int32_t Rp, int32_t opts, double& p0, double& p1) {
    using V2 = typename Vec2<VT>::type;
    using H = typename HVal<VT, CH>::type;
    constexpr int RPW = kWave / LPR;
    constexpr int TR = kWavesPerBlock * RPW;
    const int lane = threadIdx.x & (kWave - 1);
    const int sub = lane / LPR;
    const int q = lane % LPR;

    const SlotDesc& d = slots[wg_slot[w]];
    const int64_t k = d.k0 + s;
    if (k >= d.steps) return false;
    const V2* x = static_cast<const V2*>(d.blk[k % 3]);
    const V2* z = static_cast<const V2*>(d.blk[(k + 2) % 3]);
    V2* y = static_cast<V2*>(d.blk[(k + 1) % 3]);
    const int32_t* rp = d.rp;
    const int32_t* ci = d.ci;
    const H* hv = static_cast<const H*>(d.hv);
    const double ca = d.coef[k == 0 ? 0 : 2], cb = d.coef[k == 0 ? 1 : 3];
    const VT alpha = static_cast<VT>(ca), beta = static_cast<VT>(cb);
    const int32_t o = (cb != 0.0) ? (opts | kOptShift) : (opts & ~kOptShift);
    const TileOrder order = d.order;
    const int32_t rows = d.rows;

    // this workgroup's tiles: as tile_range with xcd_split, within the slot
    const int64_t j = w - d.wg0;
    const int64_t ntiles = static_cast<int64_t>(order.P) * order.Q * order.tpc;
    const int64_t xcd = j & 7, jj = j >> 3, per = d.nwg >> 3;
    const int64_t first = (ntiles * xcd) / 8 + jj, end = (ntiles * (xcd + 1)) / 8;

    p0 = 0.0;
    p1 = 0.0;
    for (int64_t t = first; t < end; t += per) {
        uint32_t slab_unused;
        const int64_t rt = visit_row_tile(order, static_cast<uint32_t>(t), slab_unused);
        if (rt < 0) continue;
        int r = static_cast<int>(rt * TR) + wave * RPW + sub;
        if (LPR == kWave) r = __builtin_amdgcn_readfirstlane(r);
        if (r >= rows) continue;
        if (k == 0)
            step_row<VT, CH, LPR, CPL, kFirst>(rp, ci, hv, x, z, y, x, r, q, Rp, alpha, beta, o, p0, p1);
        else
            step_row<VT, CH, LPR, CPL, kRecDbl>(rp, ci, hv, x, z, y, x, r, q, Rp, alpha, beta, o, p0, p1);
    }
    p0 = wave_sum(p0);
    p1 = wave_sum(p1);
    return true;
}

// Workgroup w of a slot launch at step offset s: its share of the slot's rows
// for step d.k0 + s, and the workgroup's moment partials in partial[w]
// (waves' sums added in wave order).  Nothing (the whole workgroup, before
// any barrier) once the slot is done.
template <typename VT, bool CH, int LPR, int CPL>
__device__ __forceinline__ void slots_workgroup_step(const SlotDesc* __restrict__ slots,
                                                     const int32_t* __restrict__ wg_slot, int64_t w, int32_t s,
                                                     int32_t Rp, int32_t opts, double2* __restrict__ partial) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    double p0, p1;
    if (!slots_wave_rows<VT, CH, LPR, CPL>(slots, wg_slot, w, wave, s, Rp, opts, p0, p1)) return;
    __shared__ double red[2][kWavesPerBlock];
    if (lane == 0) {
        red[0][wave] = p0;
        red[1][wave] = p1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = red[0][0], s1 = red[1][0];
        for (int i = 1; i < kWavesPerBlock; ++i) {
            s0 = s0 + red[0][i];
            s1 = s1 + red[1][i];
        }
        partial[w] = make_double2(s0, s1);
    }
}

template <typename VT, bool CH, int LPR, int CPL>
__global__ __launch_bounds__(kBlock) void cheb_slots_kernel(const SlotDesc* __restrict__ slots,
                                                            const int32_t* __restrict__ wg_slot, int32_t s,
                                                            int32_t Rp, int32_t opts, double2* __restrict__ partial) {
    slots_workgroup_step<VT, CH, LPR, CPL>(slots, wg_slot, blockIdx.x, s, Rp, opts, partial);
}

// Moments of slot `slot` at step d.k0 + s: its workgroups' partials
// (per_wg of them per workgroup, from partial[d.wg0 * per_wg]) summed in a
// fixed order, then mu_0, mu_1 (step 0) or
// the doubling identities mu_2k = 2 s.x - mu_0, mu_2k+1 = 2 s.y - mu_1.
// Nothing once the slot is done; ends with a workgroup barrier.
__device__ __forceinline__ void slot_step_moments(const SlotDesc* __restrict__ slots, int32_t slot, int32_t s,
                                                  const double2* __restrict__ partial, int per_wg = 1) {
    const SlotDesc& d = slots[slot];
    const int64_t k = d.k0 + s;
    if (k >= d.steps) return;  // the whole workgroup
    const double2* p = partial + d.wg0 * per_wg;
    double a0 = 0.0, a1 = 0.0;
    for (int64_t i = threadIdx.x; i < static_cast<int64_t>(d.nwg) * per_wg; i += kBlock) {
        a0 = a0 + p[i].x;
        a1 = a1 + p[i].y;
    }
    __shared__ double r0[kBlock], r1[kBlock];
    r0[threadIdx.x] = a0;
    r1[threadIdx.x] = a1;
    __syncthreads();
    for (int w = kBlock / 2; w >= 1; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            r0[threadIdx.x] = r0[threadIdx.x] + r0[threadIdx.x + w];
            r1[threadIdx.x] = r1[threadIdx.x] + r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* m = d.mu;
        if (k == 0) {
            m[0] = r0[0];
            m[1] = r1[0];
        } else {
            if (2 * k < d.n) m[2 * k] = 2.0 * r0[0] - m[0];
            if (2 * k + 1 < d.n) m[2 * k + 1] = 2.0 * r1[0] - m[1];
        }
    }
    __syncthreads();
}

// XCC_ID hardware register (bits 3:0): the XCD the wave runs on
__device__ __forceinline__ int xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 15; }

// ------------------------------------------- grid-synchronised batched recursion
// One launch runs steps [k0, k0 + nsteps) of every slot with the work of the
// per-step launches: at each step the launch's G workgroups (all resident:
// the host sizes G from the kernel's occupancy) take the step's workgroup
// indices w = blockIdx.x, blockIdx.x + G, ... below step_wg[i] (the active
// slots' workgroups), so each slot's rows are split and visited exactly as
// in cheb_slots_kernel; with G a multiple of 8, index w runs on XCD w % 8 as
// it does there.  The waves of a workgroup go through those indices on their
// own, each writing its own moment partial (no workgroup barrier between
// indices), so the moments group the same products per wave rather than per
// workgroup.  A grid barrier ends the step; after it the step's moments are
// formed (slot j by workgroup j mod G) while the next step runs.  The
// partials of step i are in half i % 2 of `partial`, so the moments of step
// i are read before step i + 2 (after the next barrier) overwrites them.
//
// Barriers (grid_barrier below): the workgroups of an XCD count their
// arrivals, the last of them writes the XCD's L2 back once and signals the
// grid counter, and every workgroup waits for all XCDs and then invalidates
// its caches, so rows written on other XCDs are read from memory.  Polls are
// bounded: past spin_limit polls, or once another wait has given up, a wait
// sets *error and every workgroup leaves the loop; the host then discards
// the results.
struct SyncSched {
    const SlotDesc* slots;
    const int32_t* wg_slot;
    const int64_t* step_wg;  // per step of the launch: workgroup indices of that step (active slots' prefix)
    const int32_t* step_na;  // per step of the launch: active slots
    double2* partial;        // two halves of kWavesPerBlock * wmax partials (one per wave)
    int64_t wmax;
    int32_t* bar;            // kSyncWords counters of this launch (zeroed before it): [0] leaders'
                             // arrivals, [1 + x] workgroups on XCD x, [1 + kXcds + x] arrivals on XCD x
    int32_t* error;
    int32_t s0, nsteps;      // step offsets s0 .. s0 + nsteps - 1 (slot step d.k0 + s)
    int32_t Rp, opts, spin_limit;
};

constexpr int kXcds = 16;                 // XCC_ID values
constexpr int kSyncWords = 1 + 2 * kXcds;  // counters of a synchronised launch

// Bounded poll of *ctr until it reaches target (thread 0); false, with
// *error set, past spin_limit polls or once another wait has given up.
__device__ __forceinline__ bool sync_poll(const int32_t* ctr, int32_t target, int32_t spin_limit, int32_t* error) {
    int32_t polls = 0;
    while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++polls > spin_limit ||
            ((polls & 63) == 0 && __hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
            __hip_atomic_store(error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    return true;
}

// The launch's first barrier: every workgroup counts itself on its XCD,
// then a plain grid barrier (each workgroup's thread 0 adds to bar[0]).
// Afterwards *per_xcd is the number of workgroups on this workgroup's XCD
// and *nxcd the number of XCDs with workgroups.
__device__ __forceinline__ bool sync_register(int32_t* bar, int xcd, int32_t G, int32_t spin_limit, int32_t* error,
                                              int32_t* per_xcd, int32_t* nxcd) {
    __shared__ int32_t sh[3];
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(bar + 1 + xcd, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(bar, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        const bool ok = sync_poll(bar, G, spin_limit, error);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        int32_t n = 0;
        for (int x = 0; x < kXcds; ++x)
            n += __hip_atomic_load(bar + 1 + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 0 ? 1 : 0;
        sh[0] = ok ? 1 : 0;
        sh[1] = __hip_atomic_load(bar + 1 + xcd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sh[2] = n;
    }
    __syncthreads();
    *per_xcd = sh[1];
    *nxcd = sh[2];
    return sh[0] != 0;
}

// Grid barrier i (1-based, after the first) of the launch; false if it gave
// up.  The workgroups of an XCD count their arrivals on that XCD (their
// stores are in its L2: the workgroup barrier waited for them); the last of
// them writes the XCD's L2 back once (agent-scope release) and adds to
// bar[0]; every workgroup waits until all XCDs have, then takes an
// agent-scope acquire (its CU's caches and its XCD's L2 invalidated).
__device__ __forceinline__ bool grid_barrier(int32_t* bar, int xcd, int32_t per_xcd, int32_t nxcd, int32_t G,
                                             int32_t i, int32_t spin_limit, int32_t* error) {
    __shared__ int32_t passed;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int32_t old =
            __hip_atomic_fetch_add(bar + 1 + kXcds + xcd, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (old + 1 == i * per_xcd) __hip_atomic_fetch_add(bar, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        const bool ok = sync_poll(bar, G + i * nxcd, spin_limit, error);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        passed = ok ? 1 : 0;
    }
    __syncthreads();
    return passed != 0;
}

template <typename VT, bool CH, int LPR, int CPL>
__global__ __launch_bounds__(kBlock) void cheb_slots_sync_kernel(SyncSched a) {
    const int64_t G = gridDim.x;
    const int wave = threadIdx.x / kWave;
    const int xcd = xcc_id();
    int32_t per_xcd, nxcd;
    if (!sync_register(a.bar, xcd, static_cast<int32_t>(G), a.spin_limit, a.error, &per_xcd, &nxcd)) return;
    for (int32_t i = 0; i < a.nsteps; ++i) {
        const int32_t s = a.s0 + i;
        double2* part = a.partial + (s & 1) * (kWavesPerBlock * a.wmax);
        const int64_t W = a.step_wg[i];
        // each wave goes through its workgroup indices on its own (no
        // workgroup barrier between them); one partial per wave
        for (int64_t w = blockIdx.x; w < W; w += G) {
            double p0, p1;
            if (slots_wave_rows<VT, CH, LPR, CPL>(a.slots, a.wg_slot, w, wave, s, a.Rp, a.opts, p0, p1) &&
                (threadIdx.x & (kWave - 1)) == 0)
                part[kWavesPerBlock * w + wave] = make_double2(p0, p1);
        }
        if (!grid_barrier(a.bar, xcd, per_xcd, nxcd, static_cast<int32_t>(G), i + 1, a.spin_limit, a.error)) return;
        for (int32_t j = static_cast<int32_t>(blockIdx.x); j < a.step_na[i]; j += static_cast<int32_t>(G))
            slot_step_moments(a.slots, j, s, part, kWavesPerBlock);
    }
}

// ------------------------------------------- persistent batched recursion
// One launch runs steps [k0, k1) of every slot.  Each slot (problem) lives
// on one XCD: its row tiles are dealt round robin to ns = ceil(rows /
// share_rows) shares, and the XCD's queue lists the items (step k, slot,
// share) step by step -- all
// items of step k of the queue's slots, then those of step k + 1.  The
// workgroups on an XCD (the XCC_ID hardware register) draw tickets from
// its queue counter in order, one ahead of the item they work on; so a
// slot's vectors, its counter and every reader and writer of them share one
// L2.  An item of step k > k0 waits until every share of its slot's step
// k - 1 is done (done[slot] counts the shares finished in this launch): all
// of those have lower tickets of the same queue, drawn by workgroups that
// never wait on higher ones, so every wait ends whatever the residency.  A
// wait is still bounded: past spin_limit polls, or once another wait has
// given up, it sets *error and goes on, and the host discards the results.
//
// The rows are loaded and stored with temporal accesses only (opts without
// kOptStream): a non-temporal store does not leave the row in the L2 that
// the next step gathers it from.
//
// Ordering within one XCD: after an item every wave waits until L2 has
// acknowledged its row stores, the workgroup meets at a barrier, and thread
// 0 adds to done[slot] (an L2 atomic); a waiting item polls with L2 loads
// and takes an agent-scope acquire (its CU's L1 is invalidated) before the
// barrier that releases its waves to gather.  Every item takes it, since the
// CU's L1 may hold lines of the block from an earlier step (a workgroup-scope
// invalidate, buffer_inv sc0, left stale rows in the moments).  Partial moments of item (k, slot, share) go to
// part[pbase[slot] + k * nshare[slot] + share]; slots_history_moments_kernel
// sums them in share order after the launch.
struct PersistSched {
    const SlotDesc* slots;         // per slot (layout position)
    const int32_t* nshare;         // shares per slot
    const int64_t* pbase;          // first partial of each slot
    const int32_t* q_slots;        // slots of queue q at [q_off[q], q_off[q + 1]), decreasing step counts
    const int32_t* q_off;
    const int64_t* q_share_pre;    // per queue: prefix sums of nshare over its slots (length + 1 each),
                                   // at q_off[q] + q
    const int64_t* q_step_items;   // per queue: items before step k0 + i, i = 0 .. k1 - k0 (k1 - k0 + 1
                                   // entries), at q * (k1 - k0 + 1)
    int32_t* counters;             // ticket counter per queue
    int32_t* done;                 // shares finished per slot in this launch
    int32_t* error;
    double2* part;
    int32_t k0, k1, share_rows, spin_limit;
    int32_t Rp, opts;
    int32_t xcc_queue[16];         // queue of each XCC_ID value, -1: none
};


// largest i in [0, n) with a[i] <= v (a ascending, a[0] <= v)
__device__ __forceinline__ int32_t last_at_most(const int64_t* a, int32_t n, int64_t v) {
    int32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) / 2;
        if (a[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename VT, bool CH, int LPR, int CPL>
__global__ __launch_bounds__(kBlock) void cheb_slots_persistent_kernel(PersistSched s) {
    using V2 = typename Vec2<VT>::type;
    using H = typename HVal<VT, CH>::type;
    constexpr int RPW = kWave / LPR;
    constexpr int TR = kWavesPerBlock * RPW;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int sub = lane / LPR;
    const int q = lane % LPR;
    const int qid = s.xcc_queue[xcc_id()];
    if (qid < 0) return;  // an XCD without a queue (the whole workgroup, before any barrier)
    const int32_t nk = s.k1 - s.k0;
    const int64_t* step_items = s.q_step_items + static_cast<int64_t>(qid) * (nk + 1);
    const int64_t total = step_items[nk];
    const int32_t qs0 = s.q_off[qid], nqs = s.q_off[qid + 1] - qs0;
    const int64_t* share_pre = s.q_share_pre + qs0 + qid;
    int32_t* counter = s.counters + qid;

    __shared__ double red[2][kWavesPerBlock];
    __shared__ int32_t ticket[2];
    if (threadIdx.x == 0)
        ticket[0] = __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    int64_t cur = ticket[0];
    while (cur < total) {
        if (threadIdx.x == 0)
            ticket[1] = __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // the item: step k0 + kk, the a-th slot of the queue, share j
        const int32_t kk = last_at_most(step_items, nk, cur);
        const int64_t i = cur - step_items[kk];
        const int32_t a = last_at_most(share_pre, nqs, i);
        const int32_t j = static_cast<int32_t>(i - share_pre[a]);
        const int32_t slot = s.q_slots[qs0 + a];
        const int64_t k = s.k0 + kk;
        const int32_t ns = s.nshare[slot];
        if (wave == 0) {
            if (kk > 0) {
                const int32_t need = kk * ns;
                int32_t spins = 0;
                for (;;) {
                    const int32_t v = __hip_atomic_load(s.done + slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (v >= need) break;
                    if (++spins > s.spin_limit ||
                        __hip_atomic_load(s.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                        if (lane == 0) __hip_atomic_store(s.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();

        const SlotDesc& d = s.slots[slot];
        const V2* x = static_cast<const V2*>(d.blk[k % 3]);
        const V2* z = static_cast<const V2*>(d.blk[(k + 2) % 3]);
        V2* y = static_cast<V2*>(d.blk[(k + 1) % 3]);
        const double ca = d.coef[k == 0 ? 0 : 2], cb = d.coef[k == 0 ? 1 : 3];
        const VT alpha = static_cast<VT>(ca), beta = static_cast<VT>(cb);
        const int32_t o = (cb != 0.0) ? (s.opts | kOptShift) : (s.opts & ~kOptShift);
        const int32_t rows = d.rows;
        // share j: row tiles j, j + ns, j + 2 ns, ... (TR rows each), so the
        // workgroups on a slot sweep its rows side by side and a row's
        // neighbours one lattice row away are gathered while still in L2
        const int64_t ntiles = (static_cast<int64_t>(rows) + TR - 1) / TR;
        double p0 = 0.0, p1 = 0.0;
        for (int64_t t = j; t < ntiles; t += ns) {
            const int64_t rr = t * TR + wave * RPW + sub;
            if (rr >= rows) continue;
            int r = static_cast<int>(rr);
            if (LPR == kWave) r = __builtin_amdgcn_readfirstlane(r);
            if (k == 0)
                step_row<VT, CH, LPR, CPL, kFirst>(d.rp, d.ci, static_cast<const H*>(d.hv), x, z, y, x, r, q, s.Rp,
                                                   alpha, beta, o, p0, p1);
            else
                step_row<VT, CH, LPR, CPL, kRecDbl>(d.rp, d.ci, static_cast<const H*>(d.hv), x, z, y, x, r, q,
                                                    s.Rp, alpha, beta, o, p0, p1);
        }
        p0 = wave_sum(p0);
        p1 = wave_sum(p1);
        if (lane == 0) {
            red[0][wave] = p0;
            red[1][wave] = p1;
        }
        // every wave waits until L2 has acknowledged its row stores; the
        // count then goes to the same L2
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (threadIdx.x == 0) {
            double s0 = red[0][0], s1 = red[1][0];
            for (int w = 1; w < kWavesPerBlock; ++w) {
                s0 = s0 + red[0][w];
                s1 = s1 + red[1][w];
            }
            s.part[s.pbase[slot] + k * ns + j] = make_double2(s0, s1);
            __hip_atomic_fetch_add(s.done + slot, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();  // ticket[1] stored; red free again
        cur = ticket[1];
        __syncthreads();  // everyone has read ticket[1] before thread 0 overwrites it
    }
}

// Tiled variant of cheb_step_kernel for matrices with at most KMAX = 4*KCH
// entries per row (every lattice Hamiltonian of the presets: 4 to 13).
// Per tile of TR consecutive rows the workgroup first stages the row
// pointers, column indices and values of the tile into LDS with coalesced
// loads; then every lane issues the loads of all RW rows it owns -- the
// T_{k-1} row, the bra row and KMAX neighbour-row gathers per row (slots past
// a row's end read the row itself and are discarded) -- before any of them
// is consumed, so a wave keeps RW*(KMAX+1) row loads in flight.  Results and
// accumulation order are those of cheb_step_kernel.
//
// Column slabs: the Rp columns of a block are processed as Rp/SW slabs of
// SW = LPR*CPL columns, each a sweep over all rows (tile t -> slab t /
// row_tiles, row tile t % row_tiles).  A neighbour row gathered for row i is
// needed again up to the matrix bandwidth D rows later; a narrow slab keeps
// the rows touched in that window (about 3*2D*SW elements) inside one XCD's
// L2, so the gathers hit in cache instead of re-reading memory.  The CSR
// arrays are then read once per slab.
//
// Rows per lane-slot and tile size of cheb_tile_kernel: as many rows as the
// register budget allows (4 rows of 4 gathers of one complex per lane), while
// the staged CSR of a tile stays within kTileLds bytes of LDS.
constexpr int kTileLds = 16384;
constexpr int tile_rw(int lpr, int cpl, int kch, int hbytes) {
    int rw = 4 / (kch * cpl);
    if (rw < 1) rw = 1;
    while (rw > 1 && kWavesPerBlock * (kWave / lpr) * rw * 4 * kch * (4 + hbytes) > kTileLds) rw /= 2;
    return rw;
}
constexpr int tile_rows(int lpr, int cpl, int kch, int hbytes) {
    return kWavesPerBlock * (kWave / lpr) * tile_rw(lpr, cpl, kch, hbytes);
}
constexpr bool tile_fits(int lpr, int cpl, int kch, int hbytes) {
    return tile_rows(lpr, cpl, kch, hbytes) * 4 * kch * (4 + hbytes) <= kTileLds;
}

template <typename VT, bool CH, int LPR, int CPL, int MODE, int KCH>
__global__ __launch_bounds__(kBlock) void cheb_tile_kernel(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
    const typename HVal<VT, CH>::type* __restrict__ hv, const typename Vec2<VT>::type* __restrict__ x,
    const typename Vec2<VT>::type* z, typename Vec2<VT>::type* y,
    const typename Vec2<VT>::type* __restrict__ bra, int32_t rows, int32_t Rp, VT alpha, VT beta,
    int32_t opts, int64_t ntiles, TileOrder order, int32_t xcd_split, double2* __restrict__ partial) {
    using V2 = typename Vec2<VT>::type;
    using H = typename HVal<VT, CH>::type;
    constexpr int RPW = kWave / LPR;
    constexpr int SW = LPR * CPL;
    constexpr int KMAX = 4 * KCH;
    constexpr int RW = tile_rw(LPR, CPL, KCH, sizeof(H));
    constexpr int TR = kWavesPerBlock * RPW * RW;
    constexpr int CAP = TR * KMAX;
    static_assert(tile_fits(LPR, CPL, KCH, sizeof(H)), "staged tile exceeds its LDS budget");

    __shared__ int32_t s_rp[TR + 1];
    __shared__ int32_t s_ci[CAP];
    __shared__ H s_hv[CAP];

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int sub = lane / LPR;
    const int q = lane % LPR;

    int64_t first, end, stride;
    tile_range(ntiles, xcd_split, first, end, stride);

    double p0 = 0.0, p1 = 0.0;
    for (int64_t t = first; t < end; t += stride) {
        uint32_t slab;
        const int64_t rt = visit_row_tile(order, static_cast<uint32_t>(t), slab);
        if (rt < 0) continue;  // uniform over the workgroup
        const int64_t r0 = rt * TR;
        // value c of lane q at column col0 + c * LPR: each access instruction
        // of a wave covers whole runs of the row (no partial sectors)
        const size_t col0 = static_cast<size_t>(slab) * SW + static_cast<size_t>(q);
        for (int i = threadIdx.x; i <= TR; i += kBlock) {
            const int64_t rr = r0 + i;
            s_rp[i] = rp[rr < rows ? rr : rows];
        }
        __syncthreads();
        const int base = s_rp[0];
        const int cnt = s_rp[TR] - base;  // <= CAP: every row has <= KMAX entries
        for (int j = threadIdx.x; j < cnt; j += kBlock) {
            s_ci[j] = ci[base + j];
            s_hv[j] = hv[base + j];
        }
        __syncthreads();

        int loc[RW], r[RW];
        bool live[RW];
        size_t own[RW];
#pragma unroll
        for (int k = 0; k < RW; ++k) {
            loc[k] = (k * kWavesPerBlock + wave) * RPW + sub;
            const int64_t rr = r0 + loc[k];
            live[k] = rr < rows;
            r[k] = static_cast<int>(live[k] ? rr : rows - 1);
            own[k] = static_cast<size_t>(r[k]) * Rp + col0;
        }

        V2 zv[RW][CPL], bv[RW][CPL], xv[RW][KMAX][CPL];
        H hh[RW][KMAX];
        int cc[RW][KMAX];
        int ne[RW];
#pragma unroll
        for (int k = 0; k < RW; ++k) {
            if (MODE != kFirst) {
#pragma unroll
                for (int c = 0; c < CPL; ++c) zv[k][c] = ld_stream(z + own[k] + c * LPR, opts);
            }
            if (MODE != kRecDbl) {
#pragma unroll
                for (int c = 0; c < CPL; ++c) bv[k][c] = ld_stream(bra + own[k] + c * LPR, opts);
            }
        }
#pragma unroll
        for (int k = 0; k < RW; ++k) {
            const int s = s_rp[loc[k]] - base;
            ne[k] = s_rp[loc[k] + 1] - base - s;
#pragma unroll
            for (int j = 0; j < KMAX; ++j) {
                const bool ok = j < ne[k];
                const int idx = ok ? s + j : 0;  // s + j < cnt <= CAP whenever ok
                cc[k][j] = ok ? s_ci[idx] : r[k];
                hh[k][j] = ok ? s_hv[idx] : H{};
                const V2* xr = x + static_cast<size_t>(cc[k][j]) * Rp + col0;
#pragma unroll
                for (int c = 0; c < CPL; ++c) xv[k][j][c] = xr[c * LPR];
            }
        }

#pragma unroll
        for (int k = 0; k < RW; ++k) {
            V2 acc[CPL], xs[CPL];
            bool have_self = false;
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                acc[c].x = 0;
                acc[c].y = 0;
                xs[c].x = 0;
                xs[c].y = 0;
            }
#pragma unroll
            for (int j = 0; j < KMAX; ++j) {
                if (j < ne[k]) {
#pragma unroll
                    for (int c = 0; c < CPL; ++c) madd(acc[c], hh[k][j], xv[k][j][c]);
                    if (cc[k][j] == r[k]) {
                        have_self = true;
#pragma unroll
                        for (int c = 0; c < CPL; ++c) xs[c] = xv[k][j][c];
                    }
                }
            }
            if (!have_self) {
#pragma unroll
                for (int c = 0; c < CPL; ++c) xs[c] = x[own[k] + c * LPR];
            }
            if (!live[k]) continue;
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                V2 v;
                v.x = alpha * acc[c].x;
                v.y = alpha * acc[c].y;
                if (opts & kOptShift) {
                    v.x = v.x + beta * xs[c].x;
                    v.y = v.y + beta * xs[c].y;
                }
                if (MODE != kFirst) {
                    v.x = v.x - zv[k][c].x;
                    v.y = v.y - zv[k][c].y;
                }
                st_stream(y + own[k] + c * LPR, v, opts);
                if (MODE == kFirst) {
                    p0 = p0 + redot(bv[k][c], xs[c]);
                    p1 = p1 + redot(bv[k][c], v);
                } else if (MODE == kRec) {
                    p0 = p0 + redot(bv[k][c], v);
                } else {
                    p0 = p0 + redot(xs[c], xs[c]);
                    p1 = p1 + redot(v, xs[c]);
                }
            }
        }
        __syncthreads();  // the next tile overwrites the staged CSR
    }

    __shared__ double red[2][kWavesPerBlock];
    p0 = wave_sum(p0);
    p1 = wave_sum(p1);
    if (lane == 0) {
        red[0][wave] = p0;
        red[1][wave] = p1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = red[0][0], s1 = red[1][0];
        for (int w = 1; w < kWavesPerBlock; ++w) {
            s0 = s0 + red[0][w];
            s1 = s1 + red[1][w];
        }
        partial[blockIdx.x] = make_double2(s0, s1);
    }
}

}  // namespace tbk
