// Batched solves: many independent solves (grids of any size, star/grid,
// radius 1..8, their own iteration counts) in a few launches.
//
// Two execution forms share the tile code below.
//   Round form (batch_round_kernel, the one in use): the host cuts the
//   schedule into rounds; a round holds at most one stage of each solve, so
//   its tiles are independent and one plain launch runs them all (a table
//   gives each workgroup its solve, stage and tiles); stage order within a
//   solve comes from launch order.  Solves are grouped in classes of one
//   (storage type, pattern, radius), each with its own kernel and schedule
//   on its own branch of one captured hipGraph.  No wave ever waits on
//   another.
//   Persistent form (batch_kernel): one launch, waves pull tiles from per-XCD
//   ticket queues and spin on per-solve counters.  It hung on MI355X and is
//   disabled; the host refuses it (see prk_batch_create).
//
// A solve is a sequence of stages over its tiles: stage 0 initialises in and
// out, stages 1..K+1 are the K+1 fused sweeps (sweep k reads buffer k%2 and
// writes buffer (k+1)%2), stage K+2 computes the L1 norm.  A tile is one
// wavefront's share of a stage: 64*V columns of a block of rows, processed
// by the same wave_tile() as the single-grid sweep, so a solve's arithmetic
// does not depend on what else is in the launch.
//
// Scheduling.  The host orders (solve, stage) work items into one queue per
// XCD and gives every solve to exactly one XCD, so all tiles of a solve run
// under one L2: a tile's writes are visible to the tiles of the next stage
// once they have left the wave (s_waitcnt), and loads bypass the per-CU L1
// (sc0) so they never see a stale line.  Waves take tickets from their
// XCD's queue in order; a ticket of stage s waits (bounded spin) until the
// solve's completed-tile counter shows every tile of stage s-1 done.  Since
// every ticket a wave waits on was taken earlier by a running wave, the
// queue always drains.  A wait that exceeds the spin limit sets the error
// flag, skips its tile and lets every wave exit; the host reports it.
#pragma once

#include "prk_stencil.h"
#include "sweep.hpp"

namespace prk {

constexpr int kBatchMaxRadius = kMaxRadius;  // batched and served solves: radii 1..8
constexpr int kPersistentMaxRadius = 4;      // the (disabled) persistent kernel's stencils
constexpr int kMaxQueues = 8;

struct DDv {
    double hi, lo;
};

struct BGrid {
    void *buf0, *buf1, *out;
    long long pitch;         // elements, 16-byte multiple; padding may be written
    int n, pattern, radius;
    int iters;               // K: the solve runs K+1 sweeps
    int rows;                // rows per tile
    int strips;              // ceil(n / 64V)
    int ntiles;              // strips * ceil(n / rows)
    int stage_begin;         // first stage of this launch (windowed runs)
    int init_kind;
    unsigned long long seed;
    DDv *partials;           // ntiles norm partials
    double *norm;            // result: L1 norm / (n-2R)^2
};

struct BItem {
    int grid;
    int stage;
    long long first;  // first ticket of this item in its queue
};

struct BQueue {
    const BItem *items;
    int nitems;
    long long total;  // tickets in the queue
};

struct BatchArgs {
    const BGrid *grids;
    BQueue q[kMaxQueues];
    int nq;
    unsigned *tile_done;            // per grid, tiles completed in this launch
    unsigned *norm_done;            // per grid
    unsigned long long *tickets;    // per queue
    int *error;                     // nonzero: a wait timed out
    unsigned spin_limit;
    unsigned zero;                  // always 0 (see poll_counter)
};

// Reads a counter with an atomic add of a kernel argument that is 0.  An add
// of a literal 0 would be turned into an atomic load by the compiler, and on
// gfx950 an agent-scope atomic load (sc1) reads around the XCD's L2, where
// the increments (atomic adds) are performed: the poll would never see them.
// Polls and increments must take the same path.
__device__ inline unsigned poll_counter(unsigned *p, unsigned zero)
{
    return __hip_atomic_fetch_add(p, zero, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void bump_counter(unsigned *p)
{
    __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline int xcc_id()
{
    // XCC_ID hardware register (id 20), bits 3:0.
    return (int)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);
}

__device__ inline unsigned long long wave_broadcast64(unsigned long long v)
{
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

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
    wave_tile<DT, P, R, true, MODE_VEC, 2, 1, 1>(in, nx, g.out, g.pitch, g.n, whole_grid(g.n), c0, r0, r1, 1,
                                                 ring, lane);
}

// Two waves per SIMD: the largest stencil (grid, R=4) would otherwise take
// more than 256 VGPRs and leave one wave per SIMD for every solve.
template <int DT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void batch_kernel(BatchArgs a)
{
    typedef SweepGeom<DT, kPersistentMaxRadius> GM;
    typedef typename GM::T T;
    constexpr int SW = GM::SW;
    __shared__ __attribute__((aligned(16))) T lds[kWavesPerBlock][GM::S * GM::LR];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int qi = xcc_id() % a.nq;
    const BQueue q = a.q[qi];
    T *ring = lds[wave];

    for (;;) {
        unsigned long long t = 0;
        if (lane == 0) t = atomicAdd(&a.tickets[qi], 1ull);
        t = wave_broadcast64(t);
        if ((long long)t >= q.total) break;
        // item holding ticket t: last item with first <= t
        int lo = 0, hi = q.nitems - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (q.items[mid].first <= (long long)t) lo = mid;
            else hi = mid - 1;
        }
        const BItem it = q.items[lo];
        const BGrid &g = a.grids[it.grid];
        const int tile = (int)((long long)t - it.first);

        // wait for the previous stage of this solve
        const unsigned need = (unsigned)(it.stage - g.stage_begin) * (unsigned)g.ntiles;
        bool ok = true;
        if (need) {
            unsigned spins = 0;
            for (;;) {
                unsigned done = 0;
                int failed = 0;
                if (lane == 0) {
                    done = poll_counter(&a.tile_done[it.grid], a.zero);
                    failed = (int)poll_counter((unsigned *)a.error, a.zero);
                }
                done = __builtin_amdgcn_readfirstlane(done);
                failed = __builtin_amdgcn_readfirstlane(failed);
                if (done >= need) break;
                // give up once any wave has timed out (its tile will never count)
                if (failed || ++spins > a.spin_limit) {
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        if (!ok) {
            if (lane == 0) bump_counter((unsigned *)a.error);
            continue;
        }

        const int rb = tile / g.strips;
        const int c0 = (tile - rb * g.strips) * SW;
        const int r0 = rb * g.rows;
        const int r1 = min(r0 + g.rows, g.n);
        const int norm_stage = g.iters + 2;
        if (it.stage == 0) {
            init_tile<DT>(g, c0, r0, r1, lane);
        } else if (it.stage < norm_stage) {
            const int k = it.stage - 1;
            const int pr = g.pattern * 8 + g.radius;
            switch (pr) {
            case 1: sweep_tile<DT, STAR, 1>(g, k, c0, r0, r1, ring, lane); break;
            case 2: sweep_tile<DT, STAR, 2>(g, k, c0, r0, r1, ring, lane); break;
            case 3: sweep_tile<DT, STAR, 3>(g, k, c0, r0, r1, ring, lane); break;
            case 4: sweep_tile<DT, STAR, 4>(g, k, c0, r0, r1, ring, lane); break;
            case 9: sweep_tile<DT, GRID, 1>(g, k, c0, r0, r1, ring, lane); break;
            case 10: sweep_tile<DT, GRID, 2>(g, k, c0, r0, r1, ring, lane); break;
            case 11: sweep_tile<DT, GRID, 3>(g, k, c0, r0, r1, ring, lane); break;
            case 12: sweep_tile<DT, GRID, 4>(g, k, c0, r0, r1, ring, lane); break;
            default: break;
            }
        } else {
            const DDv part = norm_tile<DT>(g, c0, r0, r1, lane);
            if (lane == 0) g.partials[tile] = part;
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the partial is in L2
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            unsigned prev = 0;
            if (lane == 0) prev = __hip_atomic_fetch_add(&a.norm_done[it.grid], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            prev = __builtin_amdgcn_readfirstlane(prev);
            if (prev == (unsigned)g.ntiles - 1) {
                // last partial of this solve: reduce all partials in tile order
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                DDv acc{0.0, 0.0};
                const __amdgpu_buffer_rsrc_t rp = row_rsrc(g.partials, (unsigned)(g.ntiles * sizeof(DDv)));
                for (int k = lane; k < g.ntiles; k += 64) {
                    const DDv p = __builtin_bit_cast(DDv, bload16<1>(rp, (unsigned)(k * sizeof(DDv))));
                    acc = dd_add_dev(acc, p);
                }
                acc = wave_reduce_dd(acc);
                const long long w = g.n - 2 * g.radius;
                if (lane == 0) *g.norm = (acc.hi + acc.lo) / (double)(w * w);
            }
        }
        // publish: every store of this tile has reached L2 before the count
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) bump_counter(&a.tile_done[it.grid]);
    }
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

// Clears the per-launch counters (a kernel, so the launch can live in a graph).
static __global__ __launch_bounds__(256) void batch_reset_kernel(unsigned *tile_done, unsigned *norm_done, int ngrids,
                                                          unsigned long long *tickets, int ntickets, int *error)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < ngrids) {
        tile_done[i] = 0;
        norm_done[i] = 0;
    }
    if (i < ntickets) tickets[i] = 0;
    if (i == 0) *error = 0;
}

}  // namespace prk
