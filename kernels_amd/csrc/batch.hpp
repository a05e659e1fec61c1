// Batched solves: many independent solves (grids of any size, star/grid,
// radius 1..8, their own iteration counts) in a few launches.
//
// The host cuts the schedule into rounds; a round holds at most one stage of
// each solve, so its tiles are independent and one plain launch runs them all
// (a table gives each workgroup its solve, stage and tiles); stage order
// within a solve comes from launch order.  Solves are grouped in classes of
// one (storage type, pattern, radius), each with its own kernel and schedule
// on its own branch of one captured hipGraph.  No wave ever waits on
// another.  (A persistent form, one launch whose waves took tiles from
// queues and waited on per-solve counters, hung on MI355X; it was removed and
// the host refuses PRK_BATCH_PERSISTENT.)
//
// A solve is a sequence of stages over its tiles: stage 0 initialises in and
// out, stages 1..K+1 are the K+1 fused sweeps (sweep k reads buffer k%2 and
// writes buffer (k+1)%2), stage K+2 computes the L1 norm partials and stage
// K+3 reduces them.  A tile is one wavefront's share of a stage: 64*V
// columns of a block of rows, processed by the same wave_tile() as the
// single-grid sweep, so a solve's arithmetic does not depend on what else is
// in the launch.
#pragma once

#include "prk_stencil.h"
#include "sweep.hpp"

namespace prk {

constexpr int kBatchMaxRadius = kMaxRadius;  // batched and served solves: radii 1..8

// Row pitch (elements) of the grids the library allocates itself (batched,
// served and one-grid-per-launch solves): rows padded to 128 bytes, so that
// every row starts on a cache line.  With rows padded only to 16 bytes, a
// wave's 1 KiB row segment straddles nine lines whenever n is not a multiple
// of 128 bytes: measured on MI355X, batches of equal FP32 grids ran 15-20 %
// slower at n = 600, 1025, 1100, 2100 than at n = 512, 1024, 2048, and
// 11-17 % faster with 128-byte rows (no change at the aligned sizes).
inline long long lib_pitch(int n, int dtype)
{
    const long long per = 128 / (long long)dtype_size(dtype);
    return ((long long)n + per - 1) / per * per;
}

struct DDv {
    double hi, lo;
};

// 32-byte aligned (96 bytes): every workgroup reads its grid's entry with
// scalar loads at the start of a tile.
struct alignas(32) BGrid {
    void *buf0, *buf1, *out;
    long long pitch;         // elements, 128-byte multiple (lib_pitch); padding may be written
    int n, pattern, radius;
    int iters;               // K: the solve runs K+1 sweeps
    int rows;                // rows per tile
    int strips;              // ceil(n / 64V)
    int ntiles;              // strips * ceil(n / rows)
    int init_kind;
    unsigned long long seed;
    DDv *partials;           // ntiles norm partials
    double *norm;            // result: L1 norm / (n-2R)^2
};

__device__ inline DDv dd_add_dev(DDv a, DDv b)
{
#pragma clang fp contract(off)
    const double s = a.hi + b.hi;
    const double bb = s - a.hi;
    const double e = (a.hi - (s - bb)) + (b.hi - bb);
    const double lo = e + a.lo + b.lo;
    const double hi = s + lo;
    return DDv{hi, lo - (hi - s)};
}

// Fixed-order butterfly over the 64 lanes: every lane ends with the total.
__device__ inline DDv wave_reduce_dd(DDv v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        DDv o;
        o.hi = __shfl_xor(v.hi, m, 64);
        o.lo = __shfl_xor(v.lo, m, 64);
        // combine in lane order so both partners compute the same value
        const bool low = (__lane_id() & m) == 0;
        v = low ? dd_add_dev(v, o) : dd_add_dev(o, v);
    }
    return v;
}

// Stage 0: in = initial values, out = 0 over the tile (padding columns 0).
template <int DT>
__device__ inline void init_tile(const BGrid &g, int c0, int r0, int r1, int lane)
{
    typedef typename DTypeTraits<DT>::T T;
    constexpr int V = DTypeTraits<DT>::V, ES = (int)sizeof(T);
    const unsigned row_bytes = (unsigned)(g.pitch * ES);
    const int cm = c0 + lane * V;
    for (int r = r0; r < r1; ++r) {
        union { u32x4 v; T e[V]; } x, z;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int c = cm + v;
            double val = 0.0;
            if (c < g.n)
                val = g.init_kind == 0 ? (double)(r + c) : seeded_value(g.seed, (uint64_t)r * (uint64_t)g.n + c);
            if constexpr (DT == F64) {
                x.e[v] = val;
                z.e[v] = 0.0;
            } else if constexpr (DT == F32) {
                x.e[v] = (float)val;
                z.e[v] = 0.0f;
            } else {
                x.e[v] = f32_to_bf16((float)val);
                z.e[v] = 0;
            }
        }
        const unsigned off = (unsigned)(cm * ES);
        bstore16<0>(row_rsrc((T *)g.buf0 + (long long)r * g.pitch, row_bytes), off, x.v);
        bstore16<0>(row_rsrc((T *)g.out + (long long)r * g.pitch, row_bytes), off, z.v);
    }
}

// Stage K+2: compensated partial L1 sum of the tile's interior points.
template <int DT>
__device__ inline DDv norm_tile(const BGrid &g, int c0, int r0, int r1, int lane)
{
    typedef typename DTypeTraits<DT>::T T;
    constexpr int V = DTypeTraits<DT>::V, ES = (int)sizeof(T);
    const unsigned row_bytes = (unsigned)(g.pitch * ES);
    const int cm = c0 + lane * V;
    const int R = g.radius;
    double sum = 0.0, comp = 0.0;
    for (int r = max(r0, R); r < min(r1, g.n - R); ++r) {
        union { u32x4 v; T e[V]; } x;
        x.v = bload16<1>(row_rsrc((const T *)g.out + (long long)r * g.pitch, row_bytes), (unsigned)(cm * ES));
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int c = cm + v;
            if (c >= R && c < g.n - R) {
                double a;
                if constexpr (DT == F64) a = x.e[v];
                else if constexpr (DT == F32) a = (double)x.e[v];
                else a = (double)bf16_to_f32(x.e[v]);
                a = fabs(a);
                const double t = sum + a;
                if (fabs(sum) >= a) comp += (sum - t) + a;
                else comp += (a - t) + sum;
                sum = t;
            }
        }
    }
    return wave_reduce_dd(DDv{sum, comp});
}

template <int DT, int P, int R>
__device__ inline void sweep_tile(const BGrid &g, int k, int c0, int r0, int r1, typename DTypeTraits<DT>::T *ring,
                                  int lane)
{
    const void *in = (k & 1) ? g.buf1 : g.buf0;
    void *nx = (k & 1) ? g.buf0 : g.buf1;
    // the grids are the library's own: out's frame may be rewritten with
    // the values it holds (FRAME_REWRITE)
    wave_tile<DT, P, R, true, MODE_VEC, 4, 1, 1, false>(in, nx, g.out, g.pitch, g.n, whole_grid(g.n), c0, r0, r1, 1,
                                                        ring, lane);
}

// ---------------------------------------------------------------------------
// Round form.  One launch = one round.  Every workgroup of the launch has a
// work entry (solve, stage, first tile) and its four waves run four
// consecutive tiles of that stage.  Stages: 0 init, 1..K+1 sweeps, K+2 norm
// partials, K+3 the solve's final norm (one tile: reduces the partials in
// tile order).
// ---------------------------------------------------------------------------
struct RWork {
    int grid;
    int stage;
    int tile0;
};

struct RoundArgs {
    const BGrid *grids;
    const RWork *work;  // one entry per workgroup
};

// One kernel per (storage type, pattern, radius): each keeps the registers
// and LDS ring its own stencil needs.
template <int DT, int P, int R>
__global__ __launch_bounds__(kThreads) void batch_round_kernel(RoundArgs a)
{
    typedef SweepGeom<DT, R> G;
    typedef typename G::T T;
    constexpr int SW = G::SW;
    __shared__ __attribute__((aligned(16))) T lds[kWavesPerBlock][G::S * G::LR];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const RWork wk = a.work[blockIdx.x];
    const BGrid &g = a.grids[wk.grid];
    const int norm_stage = g.iters + 2;
    const int tile = wk.tile0 + wave;
    if (tile >= (wk.stage == norm_stage + 1 ? 1 : g.ntiles)) return;
    const int rb = tile / g.strips;
    const int c0 = (tile - rb * g.strips) * SW;
    const int r0 = rb * g.rows;
    const int r1 = min(r0 + g.rows, g.n);
    if (wk.stage == 0) {
        init_tile<DT>(g, c0, r0, r1, lane);
    } else if (wk.stage < norm_stage) {
        sweep_tile<DT, P, R>(g, wk.stage - 1, c0, r0, r1, lds[wave], lane);
    } else if (wk.stage == norm_stage) {
        const DDv part = norm_tile<DT>(g, c0, r0, r1, lane);
        if (lane == 0) g.partials[tile] = part;
    } else {
        DDv acc{0.0, 0.0};
        for (int k = lane; k < g.ntiles; k += 64) acc = dd_add_dev(acc, g.partials[k]);
        acc = wave_reduce_dd(acc);
        const long long wd = g.n - 2 * g.radius;
        if (lane == 0) *g.norm = (acc.hi + acc.lo) / (double)(wd * wd);
    }
}

// Host helpers shared by the batch and the server (batch.hip).
int batch_class_of(const prk_solve &s);  // (dtype, pattern, radius) class index
constexpr int kBatchClasses = 3 * 2 * kBatchMaxRadius;
int batch_launch_round(int cls, const RoundArgs &a, unsigned blocks, hipStream_t s);
int batch_validate_solve(const prk_solve &s);

// Error of a reduced-precision result x (FP32 or BF16, pitch xp) against the
// same solve in FP64 y (pitch yp) over the interior R <= i, j < n-R:
// out[0] = max |x-y|, out[1] = sum |x-y|, out[2] = sum (x-y)^2,
// out[3] = sum y^2, each summed in a fixed order (kErrBlocks partials of
// four doubles in `partials`, then one block), on stream s.
constexpr int kErrBlocks = 256;
int error_vs_f64_launch(int dtype, const void *x, long long xp, const double *y, long long yp, int n, int R,
                        double *partials, double *out, hipStream_t s);

}  // namespace prk
