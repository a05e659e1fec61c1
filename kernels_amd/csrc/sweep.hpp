// The stencil sweep kernel for gfx950.
//
// One sweep computes, for every interior point p of an n x n grid,
//     out[p] += S(in)[p]                       (Cxx11/stencil_seq.hpp starR/gridR)
// and, when fused, writes in_next = in + 1 for every point of the grid
// (the refresh of Cxx11/stencil.cc:189, moved into the sweep; in_next is the
// other buffer of a ping-pong pair, so in stays read-only while neighbours
// still read it).
//
// Mapping.  A wavefront owns a strip of 64*V columns (V elements of 16 bytes
// per lane) and a block of rows, and streams down the rows.  Each input row
// is read from HBM once per strip with one 16-byte load per lane (plus a
// handful of lanes loading the R-wide column halo), parked in a ring of 2R+1
// rows in LDS, and every tap of the output row R rows above is then read
// from LDS (star stencils at large R keep their vertical taps in a register
// ring instead and park only R+1 rows in LDS: wave_tile_star).  FP64 lanes
// read their LDS windows as aligned 16-byte column pairs, which keeps the
// LDS banks free of conflicts at odd R.  Rows are software-pipelined: the
// loads of the next rows are in flight while the current row is computed.  A
// workgroup is four waves on adjacent strips.  The workgroup -> tile order is a launch parameter
// (tile_of): row-major over the grid measured best on MI355X; the orders
// that give each of the eight XCDs a contiguous run of tiles, so that a
// tile's vertical neighbours (whose rows it re-reads as halo) share its L2,
// are kept as alternatives for the autotuner and for other shapes.  The
// FP32/BF16 forms at R <= 3 may instead keep the widened windows of the last
// 2R+1 rows in registers, neighbour columns arriving by DPP lane shifts
// (wave_tile_rw, no LDS).
//
// Frame.  out's frame (its first and last R rows and columns) is never
// written.  The one lane per edge strip and row whose vector straddles a
// frame boundary parks its result in an LDS stash, and after the row loop
// the wave writes those rows' interior elements (edge_store,
// flush_edge_stash).  Grids the library owns (batched and served solves)
// store the frame back with the bits it held instead (FRAME_REWRITE).
//
// Arithmetic.  FP64: the products are summed in the generator's order and
// the sum is added to out last, with contraction off, so results are
// bitwise those of the reference's serial operator.  FP32 / BF16 storage:
// FP32 accumulation, out = T(float(out) + acc); for stars over antisymmetric
// pairs, acc = w1*(x[p+d1]-x[p-d1]), then fma for the remaining pairs; for
// grids from ring sums (ring_step).
#pragma once

#include "prk_common.hpp"

#include <utility>

namespace prk {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Build-time tuning knobs (defaults are the measured best on MI355X).
#ifndef PRK_SWEEP_PREFETCH
#define PRK_SWEEP_PREFETCH 2  // rows in the register pipeline
#endif
#ifndef PRK_SWEEP_WAVES_PER_EU
#define PRK_SWEEP_WAVES_PER_EU 0  // >0: ask the compiler for this occupancy
#endif

constexpr int kWavesPerBlock = 4;
constexpr int kThreads = 64 * kWavesPerBlock;
constexpr int kPrefetch = PRK_SWEEP_PREFETCH;

struct SweepArgs {
    const void *in;
    void *in_next;  // nullptr when not fused
    void *out;
    long long pitch;  // elements between rows
    int n;              // columns
    RowSpan span;
    int rows_per_wave;  // row block height
    int nstrips;        // ceil(n / (64*V))
    int ngroups;        // ceil(nstrips / 4)
    int nrowblocks;     // ceil(span.nrows / rows_per_wave)
    int ntiles;         // ngroups * nrowblocks
    int chunk;          // tiles per XCD (order 0) / column groups per XCD (order 2) /
                        // column groups padded to a multiple of 8 (orders 3..7)
    int pad_writable;   // columns n..pitch-1 may be overwritten
    int order;          // tile order, see tile_of()
    int stash;          // dynamic LDS holds the edge stash (kStashBytes)
};

// Workgroup -> tile.  Workgroup b runs on XCD b%8 (round-robin dispatch).
//   order 0: each XCD takes a contiguous run of tiles in column-major order
//            (a column group top to bottom, then the next);
//   order 1: row-major over the whole grid, no XCD grouping;
//   order 2: each XCD owns a contiguous range of column groups and walks it
//            row-major, so the workgroups running at once on an XCD cover a
//            narrow band of rows and a tile's vertical neighbour (whose rows
//            it re-reads as halo) runs on the same XCD shortly after it;
//   order 3: row-major with the column-group count padded to a multiple of 8
//            (chunk), so column group g always runs on XCD g%8: a tile and
//            the one below it share an L2 and run one row band apart.  When
//            the group count is not a multiple of 8 some XCDs get one group
//            more than others in every band;
//   orders 4..7: order 3 with the group -> XCD map rotated by one group every
//            2^(order-4) row bands (group g of band rb on XCD
//            (g + rb / 2^(order-4)) % 8): a tile and the one below share an
//            L2 except across a rotation, and over eight rotations every XCD
//            gets every group, which evens out their loads.
__device__ inline bool tile_of(const SweepArgs &a, int b, int &group, int &rb)
{
    if (a.order >= 4) {
        rb = b / a.chunk;
        const int shift = (rb >> (a.order - 4)) & 7;
        group = b - rb * a.chunk - shift;
        group = group < 0 ? group + a.chunk : group;
        return rb < a.nrowblocks && group < a.ngroups;
    }
    if (a.order == 1) {
        if (b >= a.ntiles) return false;
        rb = b / a.ngroups;
        group = b - rb * a.ngroups;
        return true;
    }
    if (a.order == 3) {
        rb = b / a.chunk;
        group = b - rb * a.chunk;
        return rb < a.nrowblocks && group < a.ngroups;
    }
    const int x = b & 7, k = b >> 3;
    if (a.order == 2) {
        rb = k / a.chunk;
        group = x * a.chunk + (k - rb * a.chunk);
        return rb < a.nrowblocks && (k - rb * a.chunk) < a.chunk && group < a.ngroups;
    }
    const int tile = x * a.chunk + k;
    if (k >= a.chunk || tile >= a.ntiles) return false;
    group = tile / a.nrowblocks;
    rb = tile - group * a.nrowblocks;
    return true;
}

template <int DT, int R>
struct SweepGeom {
    typedef typename DTypeTraits<DT>::T T;
    static constexpr int V = DTypeTraits<DT>::V;
    static constexpr int SW = 64 * V;                      // strip width (columns)
    static constexpr int HP = ((R + V - 1) / V) * V;       // halo pad, whole vectors
    static constexpr int NH = HP / V;                      // halo vectors per side
    static constexpr int LR = SW + 2 * HP;                 // LDS row length (elements)
    static constexpr int S = 2 * R + 1;                    // rows a point reads
};

template <typename T>
__device__ inline float to_f32(T x);
template <>
__device__ inline float to_f32<float>(float x) { return x; }
template <>
__device__ inline float to_f32<bf16_t>(bf16_t x) { return bf16_to_f32(x); }

// V floats -> V storage values.  BF16 pairs go through gfx950's packed
// conversion (v_cvt_pk_bf16_f32), which rounds to nearest even like
// f32_to_bf16 for every finite value.
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

template <int V>
__device__ inline void from_f32_vec(const float (&f)[V], float *y)
{
#pragma unroll
    for (int v = 0; v < V; ++v) y[v] = f[v];
}
template <int V>
__device__ inline void from_f32_vec(const float (&f)[V], bf16_t *y)
{
#pragma unroll
    for (int v = 0; v < V; v += 2) {
        const float hi = v + 1 < V ? f[v + 1] : f[v];
        const uint32_t w = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t){f[v], hi}, bf16x2_t));
        y[v] = (bf16_t)w;
        if (v + 1 < V) y[v + 1] = (bf16_t)(w >> 16);
    }
}

// One row of input (main vector + halo vector) and one row of out, in flight.
template <typename T, int V>
struct RowRegs {
    union {
        u32x4 v;
        T e[V];
    } main, halo, outv;
};

// in + 1 for one vector, in storage precision.
template <int DT, int V>
__device__ inline void plus_one(const typename DTypeTraits<DT>::T *x, typename DTypeTraits<DT>::T *y)
{
#pragma clang fp contract(off)
    if constexpr (DT == F64) {
#pragma unroll
        for (int v = 0; v < V; ++v) y[v] = x[v] + 1.0;
    } else {
        typedef typename DTypeTraits<DT>::T T;
        float f[V];
#pragma unroll
        for (int v = 0; v < V; ++v) f[v] = to_f32<T>(x[v]) + 1.0f;
        from_f32_vec<V>(f, y);
    }
}

// Strict-order FP64 sums of the taps for the lane's V output columns (folds
// over the compile-time tap list, so every index and weight is an
// immediate).  Each chain keeps the generator's order.  At R = 1 the chains
// run one column after the other; from R = 2 on the two chains advance
// together, fed column pair by column pair (pair_sums_f64 below), which the
// compiler schedules with fewer live registers than one chain after the
// other (measured on MI355X: grid R=4 FP64 1.47x faster, star R=2 at
// n=30000 1.2 % faster; R = 1 0.3 % slower, so it keeps the single chains).
template <int P, int R, size_t... I>
__device__ inline double tap_sum_f64(const double *const (&rows)[2 * R + 1], int v, std::index_sequence<I...>)
{
#pragma clang fp contract(off)
    constexpr const TapList &L = Stencil<P, R>::taps;
    double s = rows[L.t[0].di + R][v + L.t[0].dj] * L.t[0].w;
    ((s = s + rows[L.t[I + 1].di + R][v + L.t[I + 1].dj] * L.t[I + 1].w), ...);
    return s;
}

// The same two chains (V = 2) fed from 16-byte-aligned column pairs.  For an
// odd R the columns a lane needs start at an odd offset, and reading them as
// unaligned 8-byte pairs (ds_read2_b64 at offsets 8 and 16 from each lane's
// 16-byte slot) conflicted in the LDS banks for about 45 % of the kernel's
// time (grid R=3 and R=5, rocprof SQ_LDS_BANK_CONFLICT).  Here the lane reads
// whole aligned pairs (c, c+1), c even, with ds_read_b128, and each pair is
// used where the chains need it: chain 1 takes its taps of column offsets c-1
// and c (columns c and c+1), chain 0 those of c and c+1.  The pairs go left to
// right, so each chain still adds its taps in the generator's order.
template <int P, int R>
constexpr int tap_begin(int dj)  // first tap (generator order) with column offset >= dj
{
    int k = 0;
    while (k < Stencil<P, R>::taps.count && Stencil<P, R>::taps.t[k].dj < dj) ++k;
    return k;
}

template <int P, int R, int C>
constexpr bool pair_needs_row(int di)  // some tap of columns C-1..C+1 reads row offset di
{
    for (int k = tap_begin<P, R>(C - 1); k < tap_begin<P, R>(C + 2); ++k)
        if (Stencil<P, R>::taps.t[k].di == di) return true;
    return false;
}

template <int P, int R, int C, int B>
__device__ inline void load_pair_rows(const double *const (&rows)[2 * R + 1], double (&x)[2 * R + 1][2])
{
    if constexpr (pair_needs_row<P, R, C>(B - R)) {
        union { u32x4 v; double e[2]; } u;
        u.v = *(const u32x4 *)(rows[B] + C);
        x[B][0] = u.e[0];
        x[B][1] = u.e[1];
    }
    if constexpr (B < 2 * R) load_pair_rows<P, R, C, B + 1>(rows, x);
}

// Tap K of chain CH (output column CH), whose column CH + dj lies in the pair at C.
template <int P, int R, int CH, int C, int K>
__device__ inline void pair_tap_f64(const double (&x)[2 * R + 1][2], double &s)
{
#pragma clang fp contract(off)
    constexpr Tap t = Stencil<P, R>::taps.t[K];
    constexpr int e = CH + t.dj - C;
    static_assert(e == 0 || e == 1, "tap outside the column pair");
    if constexpr (K == 0)
        s = x[t.di + R][e] * t.w;
    else
        s = s + x[t.di + R][e] * t.w;
}

template <int P, int R, int CH, int C, int B, size_t... I>
__device__ inline void pair_taps_f64(const double (&x)[2 * R + 1][2], double &s, std::index_sequence<I...>)
{
    (pair_tap_f64<P, R, CH, C, B + (int)I>(x, s), ...);
}

template <int P, int R, int C>
__device__ inline void pair_sums_f64(const double *const (&rows)[2 * R + 1], double (&s)[2])
{
    constexpr int c_hi = (R & 1) ? R + 1 : R;  // last pair: chain 1's column R + 1
    double x[2 * R + 1][2];
    load_pair_rows<P, R, C, 0>(rows, x);
    constexpr int b1 = tap_begin<P, R>(C - 1), e1 = tap_begin<P, R>(C + 1);
    constexpr int b0 = tap_begin<P, R>(C), e0 = tap_begin<P, R>(C + 2);
    pair_taps_f64<P, R, 1, C, b1>(x, s[1], std::make_index_sequence<e1 - b1>{});
    pair_taps_f64<P, R, 0, C, b0>(x, s[0], std::make_index_sequence<e0 - b0>{});
    // Grid stencils from R = 3: a scheduling fence after each pair keeps the
    // compiler from hoisting every pair's loads to the top of the row (it
    // held 212 VGPRs at R=4 and 2 waves per SIMD; with the fence 98 and the
    // LDS-bound 4; measured on MI355X: grid R=3 +5 %, R=4 +10 %, R=5 +11 %,
    // R=6..8 +42-48 %).  The star stencils read full columns only in the
    // centre pair, and lost 1-5 % with it.
    if constexpr (P == GRID && R >= 3) __builtin_amdgcn_sched_barrier(0);
    if constexpr (C + 2 <= c_hi) pair_sums_f64<P, R, C + 2>(rows, s);
}

// Reduced-precision forms.  The window of a row is the columns [-HP, V+HP)
// around the lane's own V columns (window index i = column i - HP); a row
// whose taps reach MAXD columns to either side needs the indices
// [HP-MAXD, HP+V-1+MAXD].  Each 16-byte vector holding a needed column is
// read from LDS and widened to float once; the sums then read the register
// window.  With WHOLE, an empty asm keeps every read whole: left to itself
// the compiler narrows the partly used edge vectors to 2- and 4-byte LDS
// reads at a 16-byte lane stride, which conflict in the LDS banks (19.8 % of
// the BF16 grid R=3 sweep's cycles).  Measured on MI355X with the grid ring
// sums, alternating, at fixed tilings: whole reads make BF16 grid R=5 37 %
// and R=7 21 % faster, but BF16 grid R=2, 3, 4 and 6 0-3.5 % and R=8 10 %
// slower (the asm also constrains the scheduler), so only the BF16 grid R=5
// and R=7 sums use them.
template <typename T, int V, int HP, int MAXD, bool WHOLE>
__device__ inline void load_window(const T *row, float (&w)[V + 2 * HP])
{
    constexpr int NV = (V + 2 * HP) / V;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        if (q * V + V - 1 < HP - MAXD || q * V > HP + V - 1 + MAXD) continue;  // no needed column
        union { u32x4 v; T e[V]; } x;
        x.v = *(const u32x4 *)(row + (q * V - HP));
        if constexpr (WHOLE) __asm__ volatile("" : "+v"(x.v));
#pragma unroll
        for (int e = 0; e < V; ++e) w[q * V + e] = to_f32<T>(x.e[e]);
    }
}

// Widest column offset of the pairs on rows +-a.
template <int P, int R>
constexpr int pair_max_dj(int a)
{
    int m = 0;
    for (int k = 0; k < Stencil<P, R>::pairs.count; ++k) {
        const Tap t = Stencil<P, R>::pairs.t[k];
        const int ad = t.di < 0 ? -t.di : t.di, aj = t.dj < 0 ? -t.dj : t.dj;
        if (ad == a && aj > m) m = aj;
    }
    return m;
}

// Stars: acc = w1*(x[+d1]-x[-d1]), then fma over the other pairs, in the
// pair order of make_pairs (by row-offset magnitude |di|, generator order
// within one).
// The pairs of one magnitude a read only the rows +a and -a, so the sum walks
// the magnitudes 0..R with two widened rows live at a time.
template <int P, int R>
constexpr int pair_begin(int a)
{
    int k = 0;
    while (k < Stencil<P, R>::pairs.count &&
           (Stencil<P, R>::pairs.t[k].di < 0 ? -Stencil<P, R>::pairs.t[k].di : Stencil<P, R>::pairs.t[k].di) < a)
        ++k;
    return k;
}

// Pair K for the lane's V columns; wp holds row +|di|, wm row -|di|.
template <int P, int R, int V, int HP, int K>
__device__ inline void pair_step(const float (&wp)[V + 2 * HP], const float (&wm)[V + 2 * HP], float (&acc)[V])
{
#pragma clang fp contract(off)
    constexpr Tap t = Stencil<P, R>::pairs.t[K];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const float xp = t.di >= 0 ? wp[HP + v + t.dj] : wm[HP + v + t.dj];
        const float xm = t.di >= 0 ? wm[HP + v - t.dj] : wp[HP + v - t.dj];
        if constexpr (K == 0)
            acc[v] = (float)t.w * (xp - xm);
        else
            acc[v] = __builtin_fmaf((float)t.w, xp - xm, acc[v]);
    }
}

template <int P, int R, int V, int HP, int B, size_t... I>
__device__ inline void pair_range(const float (&wp)[V + 2 * HP], const float (&wm)[V + 2 * HP], float (&acc)[V],
                                  std::index_sequence<I...>)
{
    (pair_step<P, R, V, HP, B + (int)I>(wp, wm, acc), ...);
}

template <typename T, int P, int R, int V, int HP, int A>
__device__ inline void pair_sum_from(const T *const (&rows)[2 * R + 1], float (&acc)[V])
{
    constexpr int b = pair_begin<P, R>(A), e = pair_begin<P, R>(A + 1);
    if constexpr (e > b) {
        constexpr int md = pair_max_dj<P, R>(A);
        float wp[V + 2 * HP];
        load_window<T, V, HP, md, false>(rows[R + A], wp);
        if constexpr (A == 0) {
            pair_range<P, R, V, HP, b>(wp, wp, acc, std::make_index_sequence<e - b>{});
        } else {
            float wm[V + 2 * HP];
            load_window<T, V, HP, md, false>(rows[R - A], wm);
            pair_range<P, R, V, HP, b>(wp, wm, acc, std::make_index_sequence<e - b>{});
        }
    }
    if constexpr (A < R) pair_sum_from<T, P, R, V, HP, A + 1>(rows, acc);
}

// ---------------------------------------------------------------------------
// Grid stencils in the reduced-precision forms: ring sums.  Ring j of a grid
// stencil (the offsets at Chebyshev distance j) has one weight on its edges,
// +ce_j on column +j and on row +j and -ce_j on column -j and on row -j (the
// corners excluded), and +-cc_j on the corners (j, j) and (-j, -j)
// (weight_ab).  So, x(a, b) the input at row offset a, column offset b,
//   S = sum_j ce_j * (H_j + V_j) + cc_j * C_j,
//   V_j = VS_{j-1}(j) - VS_{j-1}(-j), VS_k(b) the sum of column b over rows
//         -k..k: VS_0(b) = x(0, b), VS_k(b) = VS_{k-1}(b) + (x(k,b) + x(-k,b));
//   H_j = D_j(0) + D_j(-1) + D_j(1) + ... + D_j(-(j-1)) + D_j(j-1), left to
//         right, D_j(b) = x(j, b) - x(-j, b);
//   C_j = x(j, j) - x(-j, -j);
//   acc = ce_1 * (H_1 + V_1), then fmaf(cc_j, C_j, acc) and
//   fmaf(ce_{j+1}, H_{j+1} + V_{j+1}, acc) in ring order.
// The column sums VS and the differences D are formed once per column and
// shared by the lane's V points.  Per point at R=3 that is 32 FP32
// operations (BF16, V=8) or 35 (FP32, V=4) against the 42 of the paired
// form (a subtraction and an fma per tap pair), at R=8 at most 153 or 188
// against 272 (column sums no point uses are never formed).  The BF16 grid sweeps and the large-R grids are bound by their VALU
// issue and run faster for it (BF16 grid R=3 +2-3 %, R=5 +25 %, R=8 +46 %;
// FP32 grid R=5 +14 %, R=8 +28 %); the HBM-bound ones do not notice.
// ---------------------------------------------------------------------------
template <int R, int J>
constexpr float ring_edge_w() { return (float)weight_ab(GRID, R, 0, J); }   // tap (J, 0)
template <int R, int J>
constexpr float ring_corner_w() { return (float)weight_ab(GRID, R, J, J); } // tap (J, J)

// Ring J for the lane's V points from the widened windows of rows +J (wp)
// and -J (wm); vs holds VS_{J-1} on entry and VS_J on return.  Window index
// HP + v + b is column b of point v.
template <int R, int V, int HP, int J>
__device__ inline void ring_step(const float (&wp)[V + 2 * HP], const float (&wm)[V + 2 * HP],
                                 float (&vs)[V + 2 * HP], float (&acc)[V])
{
#pragma clang fp contract(off)
    static_assert(HP >= R, "the window holds columns -R..V-1+R");
    constexpr float ce = ring_edge_w<R, J>(), cc = ring_corner_w<R, J>();
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int c = HP + v;
        const float vj = vs[c + J] - vs[c - J];
        float hj = wp[c] - wm[c];
#pragma unroll
        for (int m = 1; m < J; ++m) {
            hj = hj + (wp[c - m] - wm[c - m]);
            hj = hj + (wp[c + m] - wm[c + m]);
        }
        const float tj = hj + vj;
        const float cj = wp[c + J] - wm[c - J];
        if constexpr (J == 1)
            acc[v] = ce * tj;
        else
            acc[v] = __builtin_fmaf(ce, tj, acc[v]);
        acc[v] = __builtin_fmaf(cc, cj, acc[v]);
    }
    if constexpr (J < R) {
#pragma unroll
        for (int i = HP - R; i < HP + V + R; ++i) vs[i] = vs[i] + (wp[i] + wm[i]);
    }
}

template <typename T, int R, int V, int HP, int J>
__device__ inline void ring_sum_from(const T *const (&rows)[2 * R + 1], float (&vs)[V + 2 * HP], float (&acc)[V])
{
    constexpr bool whole = sizeof(T) == 2 && (R == 5 || R == 7);
    float wp[V + 2 * HP], wm[V + 2 * HP];
    load_window<T, V, HP, R, whole>(rows[R + J], wp);
    load_window<T, V, HP, R, whole>(rows[R - J], wm);
    ring_step<R, V, HP, J>(wp, wm, vs, acc);
    if constexpr (J < R) ring_sum_from<T, R, V, HP, J + 1>(rows, vs, acc);
}

// The grid sum from the LDS ring (rows as in compute_points).
template <typename T, int R, int V, int HP>
__device__ inline void ring_sum(const T *const (&rows)[2 * R + 1], float (&acc)[V])
{
    constexpr bool whole = sizeof(T) == 2 && (R == 5 || R == 7);
    float vs[V + 2 * HP];
    load_window<T, V, HP, R, whole>(rows[R], vs);
    ring_sum_from<T, R, V, HP, 1>(rows, vs, acc);
}

// The V outputs of one lane for output row o.  rows[k] points at the LDS row
// holding input row o-R+k, already offset to this lane's first column.
template <int DT, int P, int R, int V, int HP>
__device__ inline void compute_points(const typename DTypeTraits<DT>::T *const (&rows)[2 * R + 1],
                                      const typename DTypeTraits<DT>::T *outv,
                                      typename DTypeTraits<DT>::T *res)
{
#pragma clang fp contract(off)
    typedef typename DTypeTraits<DT>::T T;
    typedef Stencil<P, R> St;
    if constexpr (DT == F64) {
        double sum[V];
        if constexpr (R >= 2) {
            static_assert(V == 2, "FP64 column pairs");
            pair_sums_f64<P, R, (R & 1) ? -R - 1 : -R>(rows, sum);
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v)
                sum[v] = tap_sum_f64<P, R>(rows, v, std::make_index_sequence<St::taps.count - 1>{});
        }
#pragma unroll
        for (int v = 0; v < V; ++v) res[v] = outv[v] + sum[v];
    } else {
        float acc[V];
        if constexpr (P == GRID)
            ring_sum<T, R, V, HP>(rows, acc);
        else
            pair_sum_from<T, P, R, V, HP, 0>(rows, acc);
        float f[V];
#pragma unroll
        for (int v = 0; v < V; ++v) f[v] = to_f32<T>(outv[v]) + acc[v];
        from_f32_vec<V>(f, res);
    }
}

// ---------------------------------------------------------------------------
// Star stencils from a register ring.  A star point reads its own columns in
// rows o-R..o+R and its own row in columns -R..R.  The sweep for the larger
// star radii keeps the lane's own columns of the last 2R+1 rows in registers,
// so its LDS ring holds only the R+1 rows that will be, or now are, the
// centre row: half the LDS per wave, so more waves fit on a CU (the 2R+1-row
// ring held star R=5 and R=8 to 3 and 2 workgroups per CU).  The arithmetic
// is that of compute_points, operand for operand.  Measured on MI355X at
// n=16384 against the LDS ring: FP64 R=6/7/8 +2/+2/+4 %, FP32 R=4 +1.5 %,
// FP32 R=5..8 +2..+6 %, BF16 R=6/7/8 +13/+15/+17 %; below those radii (FP64
// R=4 and 5: -3 % tuned; BF16 R=4 and 5: -1.6 % and -2 %) it lost 1-4 % (the
// unrolled ring costs more than the occupancy gains), so they keep the LDS
// ring.
// ---------------------------------------------------------------------------
template <int DT, int P, int R>
constexpr bool star_reg_ring() { return P == STAR && R >= (DT == F32 ? 4 : 6); }

// LDS ring rows per wave of the sweep kernel, and its static LDS per workgroup.
template <int DT, int P, int R>
constexpr int sweep_ring_rows() { return star_reg_ring<DT, P, R>() ? R + 1 : 2 * R + 1; }

template <int DT, int P, int R>
constexpr size_t sweep_lds_bytes()
{
    return (size_t)kWavesPerBlock * sweep_ring_rows<DT, P, R>() * SweepGeom<DT, R>::LR *
           sizeof(typename DTypeTraits<DT>::T);
}

// FP64 tap K of both chains: horizontal taps read the centre-row window wc
// (columns C0 .. C0+NWC-1 of the lane), vertical taps the register column.
template <int R, int C0, int NWC, int K>
__device__ inline void star_tap_f64(const double (&wc)[NWC], const double (&col)[2 * R + 1][2], double (&s)[2])
{
#pragma clang fp contract(off)
    constexpr Tap t = Stencil<STAR, R>::taps.t[K];
#pragma unroll
    for (int v = 0; v < 2; ++v) {
        double x;
        if constexpr (t.di == 0)
            x = wc[v + t.dj - C0];
        else
            x = col[t.di + R][v];
        if constexpr (K == 0)
            s[v] = x * t.w;
        else
            s[v] = s[v] + x * t.w;
    }
}

template <int R, int C0, int NWC, size_t... I>
__device__ inline void star_taps_f64(const double (&wc)[NWC], const double (&col)[2 * R + 1][2], double (&s)[2],
                                     std::index_sequence<I...>)
{
    (star_tap_f64<R, C0, NWC, (int)I>(wc, col, s), ...);
}

// Reduced-precision pair K (pair order of make_pairs): horizontal pairs from
// the widened centre-row window, vertical pairs from the register columns.
template <typename T, int R, int V, int HP, int K>
__device__ inline void star_pair_step(const float (&wc)[V + 2 * HP], const T (&col)[2 * R + 1][V], float (&acc)[V])
{
#pragma clang fp contract(off)
    constexpr Tap t = Stencil<STAR, R>::pairs.t[K];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        float xp, xm;
        if constexpr (t.di == 0) {
            xp = wc[HP + v + t.dj];
            xm = wc[HP + v - t.dj];
        } else {
            xp = to_f32<T>(col[R + t.di][v]);
            xm = to_f32<T>(col[R - t.di][v]);
        }
        if constexpr (K == 0)
            acc[v] = (float)t.w * (xp - xm);
        else
            acc[v] = __builtin_fmaf((float)t.w, xp - xm, acc[v]);
    }
}

template <typename T, int R, int V, int HP, size_t... I>
__device__ inline void star_pairs(const float (&wc)[V + 2 * HP], const T (&col)[2 * R + 1][V], float (&acc)[V],
                                  std::index_sequence<I...>)
{
    (star_pair_step<T, R, V, HP, (int)I>(wc, col, acc), ...);
}

// The V outputs of one lane for output row o: crow points at the LDS copy of
// row o at the lane's first column, col[k] holds the lane's columns of row
// o-R+k.
template <int DT, int R, int V, int HP>
__device__ inline void star_points(const typename DTypeTraits<DT>::T *crow,
                                   const typename DTypeTraits<DT>::T (&col)[2 * R + 1][V],
                                   const typename DTypeTraits<DT>::T *outv, typename DTypeTraits<DT>::T *res)
{
#pragma clang fp contract(off)
    typedef typename DTypeTraits<DT>::T T;
    if constexpr (DT == F64) {
        static_assert(V == 2, "FP64 column pairs");
        constexpr int C0 = (R & 1) ? -R - 1 : -R;          // aligned pairs cover columns -R .. R+1
        constexpr int NWC = ((R & 1) ? R + 2 : R + 1) - C0 + 1;
        double wc[NWC];
#pragma unroll
        for (int q = 0; q < NWC / 2; ++q) {
            union { u32x4 v; double e[2]; } u;
            u.v = *(const u32x4 *)(crow + C0 + 2 * q);
            wc[2 * q] = u.e[0];
            wc[2 * q + 1] = u.e[1];
        }
        double sum[2];
        star_taps_f64<R, C0, NWC>(wc, col, sum, std::make_index_sequence<Stencil<STAR, R>::taps.count>{});
#pragma unroll
        for (int v = 0; v < V; ++v) res[v] = outv[v] + sum[v];
    } else {
        float wc[V + 2 * HP];
        load_window<T, V, HP, R, false>(crow, wc);
        float acc[V];
        star_pairs<T, R, V, HP>(wc, col, acc, std::make_index_sequence<Stencil<STAR, R>::pairs.count>{});
        float f[V];
#pragma unroll
        for (int v = 0; v < V; ++v) f[v] = to_f32<T>(outv[v]) + acc[v];
        from_f32_vec<V>(f, res);
    }
}

// ---------------------------------------------------------------------------
// Raw buffer access.  Every HBM access of the sweep is a buffer instruction on
// a per-row resource: an offset at or beyond the resource's byte range reads
// zero and drops the write without touching memory.  Invalid rows, columns
// and pipeline slots are therefore expressed as out-of-range offsets rather
// than branches, which keeps the row loop straight-line code and lets the
// compiler wait only for the loads a row actually needs.
// ---------------------------------------------------------------------------
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x2 __attribute__((ext_vector_type(2)));
constexpr unsigned kOOB = 0x80000000u;         // offset beyond every row resource
constexpr int kRsrcWord3 = 0x00020000;          // gfx9 raw buffer, 32-bit data format

__device__ inline __amdgpu_buffer_rsrc_t row_rsrc(const void *row, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(row), (short)0, (int)bytes, kRsrcWord3);
}

template <int AUX>
__device__ inline u32x4 bload16(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, AUX));
}

template <int AUX>
__device__ inline void bstore16(__amdgpu_buffer_rsrc_t r, unsigned off, u32x4 v)
{
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(i32x4, v), r, (int)off, 0, AUX);
}

// One element of storage type T.
template <typename T, int AUX>
__device__ inline T bload_elem(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, AUX));
    else if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, AUX));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, AUX));
}

template <typename T, int AUX>
__device__ inline void bstore_elem(__amdgpu_buffer_rsrc_t r, unsigned off, T v)
{
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(i32x2, v), r, (int)off, 0, AUX);
    else if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)off, 0, AUX);
    else
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(short, v), r, (int)off, 0, AUX);
}

// Access modes of the sweep.
//   MODE_VEC    rows 16-byte aligned and every written vector whole: n % V == 0,
//               or the caller lets columns n..pitch-1 be overwritten.
//   MODE_VEC_LD 16-byte vector loads, element stores (aligned rows, ragged n).
//   MODE_ELEM   element loads and stores (rows not 16-byte aligned).
enum { MODE_ELEM = 0, MODE_VEC = 1, MODE_VEC_LD = 2 };

// One wavefront's tile: columns [c0, c0+64V) of rows [r0, r1) of an array of
// sp.nrows rows and n columns.  ring points at this wave's (2R+1) x LR
// element LDS ring.  LDAUX adds cache policy bits to every HBM load (the
// batch kernel bypasses L1 with it, since neighbouring tiles written by
// other CUs are read there).
// Store offsets of this lane's V elements of out (frame columns and columns
// past n out of range) and whether all V are interior.  Computed where a
// store needs them, from lane_now(): hoisted out of the row loop they raised
// the register count of the sweep kernels by 2-8 VGPRs and put some
// large-R ones past an occupancy step.
// The wave's lane, counted by the hardware where it is needed: a volatile
// asm is neither hoisted out of the row loop nor merged with an earlier
// count, so nothing derived from it stays in a register across the loop.
__device__ inline int lane_now()
{
    int lane;
    __asm__ volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(lane));
    return lane;
}

template <int V, int ES>
__device__ inline bool out_offsets(int c0, int n, int R, unsigned (&off)[V])
{
    const int cm = c0 + lane_now() * V;
#pragma unroll
    for (int v = 0; v < V; ++v) off[v] = (cm + v >= R) && (cm + v < n - R) ? (unsigned)((cm + v) * ES) : kOOB;
    return cm >= R && cm + V <= n - R;
}

// Deferred edge stores (FRAME_KEEP).  A lane whose vector holds both frame
// and interior columns (the vector holding column R, and the one holding
// column n-R-1 and n-R) can only store its interior elements one by one,
// into a line whose frame bytes stay unwritten.  Such a partial store is
// acknowledged late, and loads and stores share one counter (vmcnt), so a
// wave that made one every row waited at its next rows' loads: the edge
// workgroups ran up to twice as long as the others and the sweep lost up to
// 15 % (FP32 grid R=6, BF16 star R=6 at n=16384).  Instead the edge lane parks its
// result vector for output row j of the tile in an LDS stash (slots 2j and
// 2j+1 for the left and right sides: one tile has at most one lane per side
// that straddles a frame boundary), and lane j writes the
// elements of row j, all rows at once, after the tile's last output row.
constexpr int kStashRows = 64;
constexpr size_t kStashBytes = 2 * kStashRows * sizeof(u32x4);

// In the row loop, edge waves: lanes whose vector is interior store it whole;
// the others either park it in the stash (`stash` non-null) or store their
// interior elements now.
template <typename T, int V, int ES, int AUX>
__device__ inline void edge_store(__amdgpu_buffer_rsrc_t rout, unsigned off_main, const u32x4 &resv, int c0,
                                  int n, int R, u32x4 *stash, int row)
{
    if (stash) {
        // [cm, cm+V) against the interior [R, n-R); a straddling vector
        // starts in the frame on the left side (slot 2 row), in the interior
        // on the right (slot 2 row + 1)
        const int cm = c0 + lane_now() * V;
        const bool whole = cm >= R && cm + V <= n - R;
        bstore16<AUX>(rout, whole ? off_main : kOOB, resv);
        if (!whole && cm < n - R && cm + V > R) stash[2 * row + (cm < R ? 0 : 1)] = resv;
    } else {
        unsigned off[V];
        const bool whole = out_offsets<V, ES>(c0, n, R, off);
        bstore16<AUX>(rout, whole ? off_main : kOOB, resv);
        union { u32x4 v; T e[V]; } r;
        r.v = resv;
#pragma unroll
        for (int v = 0; v < V; ++v) bstore_elem<T, AUX>(rout, whole ? kOOB : off[v], r.e[v]);
    }
}

// After the row loop: lane j writes the parked interior elements of output
// row olo + j (nout <= kStashRows rows; the caller keeps kStashRows rows of
// out under 2 GiB, so every offset fits one resource).
template <typename T, int V, int ES, int AUX>
__device__ inline void flush_edge_stash(const u32x4 *stash, T *out, long long pitch, int n, int R, int c0, int sw,
                                        int olo, int nout)
{
    const int lane = lane_now();
    // the stash writes of this wave land before its reads (in order per
    // wave); the fences keep the compiler from moving them
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const __amdgpu_buffer_rsrc_t r =
        row_rsrc(out + (long long)olo * pitch, (unsigned)(((long long)(nout - 1) * pitch + n) * ES));
    const int cb_left = R - R % V;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int col = side == 0 ? R : n - R;  // first interior / first right frame column
        const int cb = col - col % V;           // the vector holding it
        if (col % V == 0 || cb < c0 || cb >= c0 + sw) continue;
        if (side == 1 && cb == cb_left) continue;  // one vector holds both boundaries
        if (lane < nout) {
            union { u32x4 v; T e[V]; } x;
            x.v = stash[2 * lane + side];
            const unsigned row_off = (unsigned)((long long)lane * pitch * ES);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const int c = cb + v;
                bstore_elem<T, AUX>(r, c >= R && c < n - R ? row_off + (unsigned)(c * ES) : kOOB, x.e[v]);
            }
        }
    }
}

// How a wave stores out (whole-vector mode).  In the first and last strips
// of a row (`edge`, wave-uniform) a lane's vector may mix interior and frame
// columns:
//   FRAME_KEEP     such lanes store their interior elements one by one
//                  (frame elements out of range): the frame is never
//                  written, as the reference's operators promise.  Waves of
//                  the other strips store whole vectors; the test is one
//                  scalar branch at the store.
//   FRAME_REWRITE  every lane stores the whole vector with the frame
//                  elements as they were loaded (grids the library owns,
//                  where nothing else writes the frame: one store instead
//                  of V).
enum { FRAME_KEEP = 1, FRAME_REWRITE = 2 };

template <int DT, int P, int R, bool FUSE, int MODE, int PF, int NT, int LDAUX, int FRAME>
__device__ inline void wave_tile_impl(const void *in_, void *in_next_, void *out_, long long pitch, int n,
                                      const RowSpan &sp, int c0, int r0, int r1, int pad_writable,
                                      typename DTypeTraits<DT>::T *ring_base, int lane, bool edge,
                                      u32x4 *stash_lds)
{
    typedef SweepGeom<DT, R> G;
    typedef typename G::T T;
    constexpr int V = G::V, SW = G::SW, HP = G::HP, NH = G::NH, LR = G::LR, S = G::S;
    constexpr int ES = (int)sizeof(T);
    constexpr int AUX_STREAM = (NT & 1) ? 2 : 0;  // nt bit on out loads and all stores
    constexpr int AUX_IN = LDAUX;
    constexpr int AUX_OUT_LD = AUX_STREAM | LDAUX;
    (void)SW;

    const int lo = max(r0 - R, 0), hi = min(r1 + R, sp.nrows);
    const int olo = max(r0, R), ohi = min(r1, sp.nrows - R);
    // edge waves park their straddling vectors (tiles of at most kStashRows output rows)
    u32x4 *const stash = FRAME == FRAME_KEEP && MODE == MODE_VEC && ohi - olo <= kStashRows ? stash_lds : nullptr;
    const int ulo = max(r0, sp.upd_lo), uhi = min(r1, sp.upd_hi);

    const T *in = (const T *)in_;
    T *in_next = (T *)in_next_;
    T *out = (T *)out_;
    // Loads may read a whole padded row; stores stay inside [0, n) unless the
    // caller allows the padding to be written.
    const unsigned ld_bytes = (unsigned)(pitch * ES);
    const unsigned st_bytes = pad_writable ? ld_bytes : (unsigned)(n * ES);

    const int cm = c0 + lane * V;  // first column of this lane's main vector
    // Halo vector of this lane (lanes 0..NH-1 left, NH..2NH-1 right); other
    // lanes point their halo load out of range.
    const int ch = lane < NH ? c0 - HP + lane * V : c0 + SW + (lane - NH) * V;
    const int lh = lane < NH ? lane * V : HP + SW + (lane - NH) * V;  // LDS column
    const unsigned off_main = (unsigned)(cm * ES);
    const unsigned off_halo = lane < 2 * NH ? (unsigned)(ch * ES) : kOOB;  // ch<0 wraps out of range

    T(*ring)[LR] = (T(*)[LR])ring_base;

    // Row pipeline: buf[k] holds input row (main + halo) and the out row R
    // above it, loaded PF steps ahead.
    RowRegs<T, V> buf[PF];
    auto issue = [&](RowRegs<T, V> &r, int ri) {
        const bool row_ok = ri < hi;
        const int rr = row_ok ? ri : lo;
        const __amdgpu_buffer_rsrc_t rin = row_rsrc(in + (long long)rr * pitch, ld_bytes);
        const int o = ri - R;
        const bool out_ok = o >= olo && o < ohi;
        const __amdgpu_buffer_rsrc_t rout = row_rsrc(out + (long long)(out_ok ? o : olo) * pitch, ld_bytes);
        if constexpr (MODE != MODE_ELEM) {
            r.main.v = bload16<AUX_IN>(rin, row_ok ? off_main : kOOB);
            r.halo.v = bload16<AUX_IN>(rin, row_ok ? off_halo : kOOB);
            r.outv.v = bload16<AUX_OUT_LD>(rout, out_ok ? off_main : kOOB);
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                r.main.e[v] = bload_elem<T, AUX_IN>(rin, row_ok ? off_main + v * ES : kOOB);
                r.halo.e[v] = bload_elem<T, AUX_IN>(rin, (row_ok && off_halo != kOOB) ? off_halo + v * ES : kOOB);
                r.outv.e[v] = bload_elem<T, AUX_OUT_LD>(rout, out_ok ? off_main + v * ES : kOOB);
            }
        }
    };

#pragma unroll
    for (int k = 0; k < PF; ++k) issue(buf[k], lo + k);

    int slot = 0;  // ring slot of row ri: (ri - lo) % S
    for (int base = lo; base < hi; base += PF) {
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const int ri = base + k;
            RowRegs<T, V> cur = buf[k];
            issue(buf[k], ri + PF);

            // park row ri in the ring
            *(u32x4 *)&ring[slot][HP + lane * V] = cur.main.v;
            if (lane < 2 * NH) *(u32x4 *)&ring[slot][lh] = cur.halo.v;

            if constexpr (FUSE) {
                const bool nx_ok = ri >= ulo && ri < uhi;
                const __amdgpu_buffer_rsrc_t rnx =
                    row_rsrc(in_next + (long long)(nx_ok ? ri : r0) * pitch, st_bytes);
                union { u32x4 v; T e[V]; } nx;
                plus_one<DT, V>(cur.main.e, nx.e);
                if constexpr (MODE == MODE_VEC) {
                    bstore16<AUX_STREAM>(rnx, nx_ok ? off_main : kOOB, nx.v);
                } else {
#pragma unroll
                    for (int v = 0; v < V; ++v)
                        bstore_elem<T, AUX_STREAM>(rnx, nx_ok ? off_main + v * ES : kOOB, nx.e[v]);
                }
            }

            // LDS writes of this wave land before its reads (in-order per
            // wave); the fences keep the compiler from moving them.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

            // Output row o = ri - R: rows o-R .. o+R live in slots
            // slot+1 .. slot+2R+1 (mod S).  The first 2R rows of a tile only
            // fill the ring (their output rows belong to the tile above), as
            // do rows past the interior: the whole wave skips their
            // arithmetic (o is the same in every lane).
            const int o = ri - R;
            if (o >= olo && o < ohi) {
                const T *rows[S];
#pragma unroll
                for (int k2 = 0; k2 < S; ++k2) {
                    int s2 = slot + 1 + k2;
                    s2 = s2 >= S ? s2 - S : s2;
                    rows[k2] = &ring[s2][HP + lane * V];
                }
                union { u32x4 v; T e[V]; } res;
                compute_points<DT, P, R, V, HP>(rows, cur.outv.e, res.e);
                const __amdgpu_buffer_rsrc_t rout = row_rsrc(out + (long long)o * pitch, st_bytes);
                if constexpr (MODE == MODE_VEC && FRAME == FRAME_REWRITE) {
                    unsigned off[V];
                    out_offsets<V, ES>(c0, n, R, off);
#pragma unroll
                    for (int v = 0; v < V; ++v) res.e[v] = off[v] != kOOB ? res.e[v] : cur.outv.e[v];
                    bstore16<AUX_STREAM>(rout, off_main, res.v);
                } else if constexpr (MODE == MODE_VEC) {
                    if (!edge) {
                        bstore16<AUX_STREAM>(rout, off_main, res.v);
                    } else {
                        edge_store<T, V, ES, AUX_STREAM>(rout, off_main, res.v, c0, n, R, stash, o - olo);
                        if (stash && o + 1 == ohi)
                            flush_edge_stash<T, V, ES, AUX_STREAM>(stash, out, pitch, n, R, c0, SW, olo, ohi - olo);
                    }
                } else {
                    unsigned off[V];
                    out_offsets<V, ES>(c0, n, R, off);
#pragma unroll
                    for (int v = 0; v < V; ++v) bstore_elem<T, AUX_STREAM>(rout, off[v], res.e[v]);
                }
            }
            slot = (slot + 1 == S) ? 0 : slot + 1;
        }
    }
}

// KEEP_FRAME false: the grid belongs to the library (batched and served
// solves), so every wave stores whole vectors, the frame elements of out as
// they were loaded (FRAME_REWRITE).
template <int DT, int P, int R, bool FUSE, int MODE, int PF, int NT, int LDAUX, bool KEEP_FRAME = true>
__device__ inline void wave_tile(const void *in_, void *in_next_, void *out_, long long pitch, int n,
                                 const RowSpan &sp, int c0, int r0, int r1, int pad_writable,
                                 typename DTypeTraits<DT>::T *ring_base, int lane, u32x4 *stash_lds = nullptr)
{
    // wave-uniform: only the first and last strips of a row hold frame columns
    const bool edge = c0 < R || c0 + SweepGeom<DT, R>::SW > n - R;
    wave_tile_impl<DT, P, R, FUSE, MODE, PF, NT, LDAUX, KEEP_FRAME ? FRAME_KEEP : FRAME_REWRITE>(
        in_, in_next_, out_, pitch, n, sp, c0, r0, r1, pad_writable, ring_base, lane, edge, stash_lds);
}

// wave_tile for the star radii of star_reg_ring: the same loads, stores and
// arithmetic, with the rows' main vectors held in a register ring qm of M
// slots (M >= 2R+1+PF, a multiple of PF; row ri in slot (ri - lo) % M; the
// row loop is unrolled by M so every slot index is a constant) and an LDS
// ring of R+1 rows for the centre rows.
template <int DT, int R, bool FUSE, int MODE, int PF, int NT>
__device__ inline void wave_tile_star(const void *in_, void *in_next_, void *out_, long long pitch, int n,
                                      const RowSpan &sp, int c0, int r0, int r1, int pad_writable,
                                      typename DTypeTraits<DT>::T *ring_base, int lane, u32x4 *stash_lds)
{
    // wave-uniform: only the first and last strips of a row hold frame columns
    const bool edge = c0 < R || c0 + SweepGeom<DT, R>::SW > n - R;
    typedef SweepGeom<DT, R> G;
    typedef typename G::T T;
    constexpr int V = G::V, HP = G::HP, NH = G::NH, LR = G::LR, S = G::S;
    constexpr int L = R + 1;                         // LDS ring rows
    constexpr int M = (S + PF + PF - 1) / PF * PF;   // register ring slots, >= S + PF
    constexpr int ES = (int)sizeof(T);
    constexpr int AUX_STREAM = (NT & 1) ? 2 : 0;

    const int lo = max(r0 - R, 0), hi = min(r1 + R, sp.nrows);
    const int olo = max(r0, R), ohi = min(r1, sp.nrows - R);
    // edge waves park their straddling vectors (tiles of at most kStashRows output rows)
    u32x4 *const stash = MODE == MODE_VEC && ohi - olo <= kStashRows ? stash_lds : nullptr;
    const int ulo = max(r0, sp.upd_lo), uhi = min(r1, sp.upd_hi);

    const T *in = (const T *)in_;
    T *in_next = (T *)in_next_;
    T *out = (T *)out_;
    const unsigned ld_bytes = (unsigned)(pitch * ES);
    const unsigned st_bytes = pad_writable ? ld_bytes : (unsigned)(n * ES);

    const int cm = c0 + lane * V;
    const int ch = lane < NH ? c0 - HP + lane * V : c0 + G::SW + (lane - NH) * V;
    const int lh = lane < NH ? lane * V : HP + G::SW + (lane - NH) * V;
    const unsigned off_main = (unsigned)(cm * ES);
    const unsigned off_halo = lane < 2 * NH ? (unsigned)(ch * ES) : kOOB;

    T(*ring)[LR] = (T(*)[LR])ring_base;

    // Row ri: its main vector in qm[(ri - lo) % M], its halo vector and the
    // out row ri - R in buf[(ri - lo) % PF] (M is a multiple of PF).
    union Vec { u32x4 v; T e[V]; };
    Vec qm[M];
    struct Side { Vec halo, outv; } buf[PF];
    auto issue = [&](Vec &m, Side &sd, int ri) {
        const bool row_ok = ri < hi;
        const int rr = row_ok ? ri : lo;
        const __amdgpu_buffer_rsrc_t rin = row_rsrc(in + (long long)rr * pitch, ld_bytes);
        const int o = ri - R;
        const bool out_ok = o >= olo && o < ohi;
        const __amdgpu_buffer_rsrc_t rout = row_rsrc(out + (long long)(out_ok ? o : olo) * pitch, ld_bytes);
        if constexpr (MODE != MODE_ELEM) {
            m.v = bload16<0>(rin, row_ok ? off_main : kOOB);
            sd.halo.v = bload16<0>(rin, row_ok ? off_halo : kOOB);
            sd.outv.v = bload16<AUX_STREAM>(rout, out_ok ? off_main : kOOB);
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                m.e[v] = bload_elem<T, 0>(rin, row_ok ? off_main + v * ES : kOOB);
                sd.halo.e[v] = bload_elem<T, 0>(rin, (row_ok && off_halo != kOOB) ? off_halo + v * ES : kOOB);
                sd.outv.e[v] = bload_elem<T, AUX_STREAM>(rout, out_ok ? off_main + v * ES : kOOB);
            }
        }
    };

#pragma unroll
    for (int k = 0; k < PF; ++k) issue(qm[k], buf[k], lo + k);

    int lslot = 0;  // LDS ring slot of row ri: (ri - lo) % L
    for (int base = lo; base < hi; base += M) {
#pragma clang loop unroll(full)
        for (int k = 0; k < M; ++k) {
            const int ri = base + k;
            if (ri >= hi) return;  // uniform: the same row in every lane
            const Vec cur = qm[k];
            const Side sd = buf[k % PF];
            // qm slot (k+PF)%M held row ri+PF-M <= ri-2R-1, no longer needed.
            issue(qm[(k + PF) % M], buf[k % PF], ri + PF);

            *(u32x4 *)&ring[lslot][HP + lane * V] = cur.v;
            if (lane < 2 * NH) *(u32x4 *)&ring[lslot][lh] = sd.halo.v;

            if constexpr (FUSE) {
                const bool nx_ok = ri >= ulo && ri < uhi;
                const __amdgpu_buffer_rsrc_t rnx =
                    row_rsrc(in_next + (long long)(nx_ok ? ri : r0) * pitch, st_bytes);
                Vec nx;
                plus_one<DT, V>(cur.e, nx.e);
                if constexpr (MODE == MODE_VEC) {
                    bstore16<AUX_STREAM>(rnx, nx_ok ? off_main : kOOB, nx.v);
                } else {
#pragma unroll
                    for (int v = 0; v < V; ++v)
                        bstore_elem<T, AUX_STREAM>(rnx, nx_ok ? off_main + v * ES : kOOB, nx.e[v]);
                }
            }

            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

            // Output row o = ri - R (rows o-R .. o+R are all >= lo here, see
            // olo): its LDS slot is (lslot - R) mod (R+1) = lslot + 1 mod L,
            // overwritten only by the next row; row o-R+j is in register
            // slot (k - 2R + j) mod M.
            const int o = ri - R;
            if (o >= olo && o < ohi) {
                const int cs = lslot + 1 == L ? 0 : lslot + 1;
                T col[S][V];
#pragma unroll
                for (int j = 0; j < S; ++j)
#pragma unroll
                    for (int v = 0; v < V; ++v) col[j][v] = qm[(k + M - 2 * R + j) % M].e[v];
                Vec res;
                star_points<DT, R, V, HP>(&ring[cs][HP + lane * V], col, sd.outv.e, res.e);
                const __amdgpu_buffer_rsrc_t rout = row_rsrc(out + (long long)o * pitch, st_bytes);
                if constexpr (MODE == MODE_VEC) {
                    if (!edge) {
                        bstore16<AUX_STREAM>(rout, off_main, res.v);
                    } else {
                        edge_store<T, V, ES, AUX_STREAM>(rout, off_main, res.v, c0, n, R, stash, o - olo);
                        if (stash && o + 1 == ohi)
                            flush_edge_stash<T, V, ES, AUX_STREAM>(stash, out, pitch, n, R, c0, G::SW, olo, ohi - olo);
                    }
                } else {
                    unsigned off[V];
                    out_offsets<V, ES>(c0, n, R, off);
#pragma unroll
                    for (int v = 0; v < V; ++v) bstore_elem<T, AUX_STREAM>(rout, off[v], res.e[v]);
                }
            }
            lslot = lslot + 1 == L ? 0 : lslot + 1;
        }
    }
}


// ---------------------------------------------------------------------------
// Reduced-precision sweeps at small R from a register ring of widened row
// windows.  The LDS-ring sweep widens every value it reads, so each input
// row is widened once for every one of the 2R+1 output rows that read it,
// and the partly used edge vectors of its windows are read from LDS in
// pieces that conflict in the banks.  Here each input row is widened once,
// when it arrives: the lane's own V columns from its load, the R columns on
// either side from the neighbouring lanes' loads by whole-wave DPP shifts
// (lane 0 and lane 63 take them from the strip's halo vectors, which lanes 0
// and 1 load).  The windows of the last 2R+1 rows stay in registers (the row
// loop is unrolled by 2R+1, so every slot is a constant register); no LDS is
// used.  The arithmetic is that of compute_points, operand for operand, so
// results are bitwise those of the LDS-ring sweep.
// ---------------------------------------------------------------------------
template <int DT, int P, int R>
constexpr bool reg_window_ok() { return DT != F64 && R <= 3; }

// Whole-wave lane shifts: lane i receives lane i-1's value (from_left) or
// lane i+1's (from_right); lane 0, resp. lane 63, has no source and keeps
// `edge`.
__device__ inline unsigned from_left(unsigned x, unsigned edge)
{
    return (unsigned)__builtin_amdgcn_update_dpp((int)edge, (int)x, 0x138 /* wave_shr:1 */, 0xF, 0xF, false);
}
__device__ inline unsigned from_right(unsigned x, unsigned edge)
{
    return (unsigned)__builtin_amdgcn_update_dpp((int)edge, (int)x, 0x130 /* wave_shl:1 */, 0xF, 0xF, false);
}

// The window of one input row as loaded: the lane's own vector and the
// dwords of its neighbours' vectors that hold the R columns on either side.
// main is the lane's own vector; lhalo the vector of columns -V..-1 (used by
// lane 0: the strip's left halo vector); rhalo the vector of columns
// 64V..65V-1 of the strip (used by lane 63: the strip's right halo vector).
template <typename T, int V, int R>
struct PackedWindow {
    static_assert(R <= V, "neighbour columns come from the adjacent lane only");
    static constexpr int E = 4 / (int)sizeof(T);  // elements per dword
    static constexpr int LD0 = (V - R) / E;       // first dword of the left neighbour used
    static constexpr int NL = V / E - LD0;        // left neighbour dwords
    static constexpr int NR = (R - 1) / E + 1;    // right neighbour dwords
    u32x4 m;
    unsigned l[NL], r[NR];

    __device__ inline void put(const u32x4 &main, const u32x4 &lhalo, const u32x4 &rhalo)
    {
        m = main;
#pragma unroll
        for (int d = 0; d < NL; ++d) l[d] = from_left(main[LD0 + d], lhalo[LD0 + d]);
#pragma unroll
        for (int d = 0; d < NR; ++d) r[d] = from_right(main[d], rhalo[d]);
    }
    // w[i] = column i - R from the lane's first column, i in [0, V + 2R)
    __device__ inline void widen(float (&w)[V + 2 * R]) const
    {
        union { u32x4 v; T e[V]; } mm, ll, rr;
        mm.v = m;
        ll.v = u32x4{0u, 0u, 0u, 0u};
        rr.v = u32x4{0u, 0u, 0u, 0u};
#pragma unroll
        for (int d = 0; d < NL; ++d) ll.v[LD0 + d] = l[d];
#pragma unroll
        for (int d = 0; d < NR; ++d) rr.v[d] = r[d];
#pragma unroll
        for (int i = 0; i < R; ++i) w[i] = to_f32<T>(ll.e[V - R + i]);
#pragma unroll
        for (int i = 0; i < V; ++i) w[R + i] = to_f32<T>(mm.e[i]);
#pragma unroll
        for (int i = 0; i < R; ++i) w[R + V + i] = to_f32<T>(rr.e[i]);
    }
};

// A ring slot: the window widened once, when its row arrives.  (Keeping the
// BF16 windows as loaded and widening them where they are used does not
// shrink the ring: the compiler keeps each slot's widened values live across
// the unrolled steps, or, forced to widen at every use, still needs over 190
// VGPRs at grid R=3.)
template <typename T, int V, int R>
struct RingSlot {
    float w[V + 2 * R];
    __device__ inline void put(const u32x4 &main, const u32x4 &lhalo, const u32x4 &rhalo)
    {
        PackedWindow<T, V, R> pw;
        pw.put(main, lhalo, rhalo);
        pw.widen(w);
    }
};

// Star paired sums over the ring: the window of row o + a is in slot
// (k + 1 + R + a) % S for the output row of ring step k (rows o-R..o+R sit
// in slots k+1..k+S).  k is the step of the fully unrolled row loop, so every
// index is a constant once the loop is unrolled and the ring stays in
// registers.
template <int P, int R, int V, int S, int A, typename Slot>
__device__ inline void rw_pair_sum(const Slot (&W)[S], int k, float (&acc)[V])
{
    constexpr int b = pair_begin<P, R>(A), e = pair_begin<P, R>(A + 1);
    if constexpr (e > b)
        pair_range<P, R, V, R, b>(W[(k + 1 + R + A) % S].w, W[(k + 1 + R - A) % S].w, acc,
                                  std::make_index_sequence<e - b>{});
    if constexpr (A < R) rw_pair_sum<P, R, V, S, A + 1>(W, k, acc);
}

// Grid ring sums over the ring slots (rows as in rw_pair_sum).
template <int R, int V, int S, int J, typename Slot>
__device__ inline void rw_ring_from(const Slot (&W)[S], int k, float (&vs)[V + 2 * R], float (&acc)[V])
{
    ring_step<R, V, R, J>(W[(k + 1 + R + J) % S].w, W[(k + 1 + R - J) % S].w, vs, acc);
    if constexpr (J < R) rw_ring_from<R, V, S, J + 1>(W, k, vs, acc);
}

template <int P, int R, int V, int S, typename Slot>
__device__ inline void rw_sum(const Slot (&W)[S], int k, float (&acc)[V])
{
    if constexpr (P == GRID) {
        float vs[V + 2 * R];
#pragma unroll
        for (int i = 0; i < V + 2 * R; ++i) vs[i] = W[(k + 1 + R) % S].w[i];
        rw_ring_from<R, V, S, 1>(W, k, vs, acc);
    } else {
        rw_pair_sum<P, R, V, S, 0>(W, k, acc);
    }
}

template <int DT, int P, int R, bool FUSE, int MODE, int PF, int NT>
__device__ inline void wave_tile_rw(const void *in_, void *in_next_, void *out_, long long pitch, int n,
                                    const RowSpan &sp, int c0, int r0, int r1, int pad_writable, int lane,
                                    u32x4 *stash_lds)
{
    // wave-uniform: only the first and last strips of a row hold frame columns
    const bool edge = c0 < R || c0 + SweepGeom<DT, R>::SW > n - R;
    typedef SweepGeom<DT, R> G;
    typedef typename G::T T;
    constexpr int V = G::V, HP = G::HP, NH = G::NH, S = G::S;
    constexpr int ES = (int)sizeof(T);
    constexpr int AUX_STREAM = (NT & 1) ? 2 : 0;
    static_assert(HP == V && NH == 1, "one halo vector per side");
    static_assert(PF <= S, "rows in flight fit the ring");

    const int lo = max(r0 - R, 0), hi = min(r1 + R, sp.nrows);
    const int olo = max(r0, R), ohi = min(r1, sp.nrows - R);
    // edge waves park their straddling vectors (tiles of at most kStashRows output rows)
    u32x4 *const stash = MODE == MODE_VEC && ohi - olo <= kStashRows ? stash_lds : nullptr;
    const int ulo = max(r0, sp.upd_lo), uhi = min(r1, sp.upd_hi);

    const T *in = (const T *)in_;
    T *in_next = (T *)in_next_;
    T *out = (T *)out_;
    const unsigned ld_bytes = (unsigned)(pitch * ES);
    const unsigned st_bytes = pad_writable ? ld_bytes : (unsigned)(n * ES);

    const int cm = c0 + lane * V;
    // halo vectors: lane 0 the V columns left of the strip, lane 1 the V
    // columns right of it; other lanes load nothing
    const int ch = lane == 0 ? c0 - V : c0 + G::SW;
    const unsigned off_main = (unsigned)(cm * ES);
    const unsigned off_halo = lane < 2 ? (unsigned)(ch * ES) : kOOB;  // ch < 0 wraps out of range

    // Row ri's loads (main, halo, out row ri - R) in raw[(ri - lo) % S]; its
    // window in W[(ri - lo) % S].
    RowRegs<T, V> raw[S];
    RingSlot<T, V, R> W[S];
    auto issue = [&](RowRegs<T, V> &r, int ri) {
        const bool row_ok = ri < hi;
        const int rr = row_ok ? ri : lo;
        const __amdgpu_buffer_rsrc_t rin = row_rsrc(in + (long long)rr * pitch, ld_bytes);
        const int o = ri - R;
        const bool out_ok = o >= olo && o < ohi;
        const __amdgpu_buffer_rsrc_t rout = row_rsrc(out + (long long)(out_ok ? o : olo) * pitch, ld_bytes);
        if constexpr (MODE != MODE_ELEM) {
            r.main.v = bload16<0>(rin, row_ok ? off_main : kOOB);
            r.halo.v = bload16<0>(rin, row_ok ? off_halo : kOOB);
            r.outv.v = bload16<AUX_STREAM>(rout, out_ok ? off_main : kOOB);
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                r.main.e[v] = bload_elem<T, 0>(rin, row_ok ? off_main + v * ES : kOOB);
                r.halo.e[v] = bload_elem<T, 0>(rin, (row_ok && off_halo != kOOB) ? off_halo + v * ES : kOOB);
                r.outv.e[v] = bload_elem<T, AUX_STREAM>(rout, out_ok ? off_main + v * ES : kOOB);
            }
        }
    };

#pragma unroll
    for (int k = 0; k < PF; ++k) issue(raw[k], lo + k);

    for (int base = lo; base < hi; base += S) {
#pragma clang loop unroll(full)
        for (int k = 0; k < S; ++k) {
            const int ri = base + k;
            if (ri >= hi) return;  // uniform: the same row in every lane
            const RowRegs<T, V> cur = raw[k];
            issue(raw[(k + PF) % S], ri + PF);

            if constexpr (FUSE) {
                const bool nx_ok = ri >= ulo && ri < uhi;
                const __amdgpu_buffer_rsrc_t rnx =
                    row_rsrc(in_next + (long long)(nx_ok ? ri : r0) * pitch, st_bytes);
                union { u32x4 v; T e[V]; } nx;
                plus_one<DT, V>(cur.main.e, nx.e);
                if constexpr (MODE == MODE_VEC) {
                    bstore16<AUX_STREAM>(rnx, nx_ok ? off_main : kOOB, nx.v);
                } else {
#pragma unroll
                    for (int v = 0; v < V; ++v)
                        bstore_elem<T, AUX_STREAM>(rnx, nx_ok ? off_main + v * ES : kOOB, nx.e[v]);
                }
            }

            // lane 63's right neighbours: the right halo vector, loaded by lane 1
            u32x4 rhalo;
#pragma unroll
            for (int d = 0; d < 4; ++d) rhalo[d] = (unsigned)__builtin_amdgcn_readlane((int)cur.halo.v[d], 1);
            W[k].put(cur.main.v, cur.halo.v, rhalo);

            const int o = ri - R;
            if (o >= olo && o < ohi) {
                float acc[V];
                rw_sum<P, R, V, S>(W, k, acc);
                float f[V];
#pragma unroll
                for (int v = 0; v < V; ++v) f[v] = to_f32<T>(cur.outv.e[v]) + acc[v];
                union { u32x4 v; T e[V]; } res;
                from_f32_vec<V>(f, res.e);
                const __amdgpu_buffer_rsrc_t rout = row_rsrc(out + (long long)o * pitch, st_bytes);
                if constexpr (MODE == MODE_VEC) {
                    if (!edge) {
                        bstore16<AUX_STREAM>(rout, off_main, res.v);
                    } else {
                        edge_store<T, V, ES, AUX_STREAM>(rout, off_main, res.v, c0, n, R, stash, o - olo);
                        if (stash && o + 1 == ohi)
                            flush_edge_stash<T, V, ES, AUX_STREAM>(stash, out, pitch, n, R, c0, G::SW, olo, ohi - olo);
                    }
                } else {
                    unsigned off[V];
                    out_offsets<V, ES>(c0, n, R, off);
#pragma unroll
                    for (int v = 0; v < V; ++v) bstore_elem<T, AUX_STREAM>(rout, off[v], res.e[v]);
                }
            }
        }
    }
}


// Dynamic LDS of the sweep kernels: the launch's occupancy padding, whose
// first kStashBytes hold the edge stash when SweepArgs::stash is set.
extern __shared__ __attribute__((aligned(16))) u32x4 sweep_dyn_lds[];

// The register-window sweep kernel: the launch geometry of sweep_kernel, no LDS.
template <int DT, int P, int R, bool FUSE, int MODE, int PF, int NT>
__global__ __launch_bounds__(kThreads) void sweep_rw_kernel(SweepArgs a)
{
    int group, rb;
    if (!tile_of(a, blockIdx.x, group, rb)) return;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int strip = group * kWavesPerBlock + wave;
    if (strip >= a.nstrips) return;
    const int r0 = rb * a.rows_per_wave;
    const int r1 = min(r0 + a.rows_per_wave, a.span.nrows);
    // no edge stash: the register-window sweep's edge waves showed no cost
    // from storing their straddling vectors at once
    wave_tile_rw<DT, P, R, FUSE, MODE, PF, NT>(a.in, a.in_next, a.out, a.pitch, a.n, a.span,
                                               strip * SweepGeom<DT, R>::SW, r0, r1, a.pad_writable, lane, nullptr);
}

template <int DT, int P, int R, bool FUSE, int MODE, int PF, int NT>
__global__ __launch_bounds__(kThreads)
#if PRK_SWEEP_WAVES_PER_EU > 0
__attribute__((amdgpu_waves_per_eu(PRK_SWEEP_WAVES_PER_EU)))
#endif
void sweep_kernel(SweepArgs a)
{
    typedef SweepGeom<DT, R> G;
    typedef typename G::T T;
    __shared__ __attribute__((aligned(16))) T lds[kWavesPerBlock][sweep_ring_rows<DT, P, R>()][G::LR];

    int group, rb;
    if (!tile_of(a, blockIdx.x, group, rb)) return;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int strip = group * kWavesPerBlock + wave;
    if (strip >= a.nstrips) return;
    const int r0 = rb * a.rows_per_wave;
    const int r1 = min(r0 + a.rows_per_wave, a.span.nrows);
    if constexpr (star_reg_ring<DT, P, R>())
        wave_tile_star<DT, R, FUSE, MODE, PF, NT>(a.in, a.in_next, a.out, a.pitch, a.n, a.span, strip * G::SW, r0,
                                                  r1, a.pad_writable, &lds[wave][0][0], lane,
                                                  a.stash ? sweep_dyn_lds : nullptr);
    else
        wave_tile<DT, P, R, FUSE, MODE, PF, NT, 0>(a.in, a.in_next, a.out, a.pitch, a.n, a.span, strip * G::SW,
                                                   r0, r1, a.pad_writable, &lds[wave][0][0], lane,
                                                   a.stash ? sweep_dyn_lds : nullptr);
}

}  // namespace prk
