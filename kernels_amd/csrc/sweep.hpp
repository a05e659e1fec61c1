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
// from LDS.  Rows are software-pipelined: the loads of the next rows are in
// flight while the current row is computed.  A workgroup is four waves on
// adjacent strips; workgroups are mapped so that the eight XCDs each take a
// contiguous run of tiles in column-major order, which keeps a tile's
// vertical neighbours (whose rows it re-reads as halo) in the same L2.
//
// Arithmetic.  FP64: the products are summed in the generator's order and
// the sum is added to out last, with contraction off, so results are
// bitwise those of the reference's serial operator.  FP32 / BF16 storage:
// FP32 accumulation over antisymmetric pairs, acc = w1*(x[p+d1]-x[p-d1]),
// then fma for the remaining pairs, out = T(float(out) + acc).
#pragma once

#include "prk_common.hpp"

#include <utility>

namespace prk {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kWavesPerBlock = 4;
constexpr int kThreads = 64 * kWavesPerBlock;
constexpr int kPrefetch = 3;  // rows in the register pipeline

struct SweepArgs {
    const void *in;
    void *in_next;  // nullptr when not fused
    void *out;
    long long pitch;  // elements between rows
    int n;
    int rows_per_wave;  // row block height
    int nstrips;        // ceil(n / (64*V))
    int ngroups;        // ceil(nstrips / 4)
    int nrowblocks;     // ceil(n / rows_per_wave)
    int ntiles;         // ngroups * nrowblocks
    int chunk;          // ceil(ntiles / 8): tiles per XCD
};

template <int DT, int R>
struct SweepGeom {
    typedef typename DTypeTraits<DT>::T T;
    static constexpr int V = DTypeTraits<DT>::V;
    static constexpr int SW = 64 * V;                      // strip width (columns)
    static constexpr int HP = ((R + V - 1) / V) * V;       // halo pad, whole vectors
    static constexpr int NH = HP / V;                      // halo vectors per side
    static constexpr int LR = SW + 2 * HP;                 // LDS row length (elements)
    static constexpr int S = 2 * R + 1;                    // ring rows
    static constexpr size_t lds_bytes = (size_t)kWavesPerBlock * S * LR * sizeof(T);
};

template <typename T>
__device__ inline float to_f32(T x);
template <>
__device__ inline float to_f32<float>(float x) { return x; }
template <>
__device__ inline float to_f32<bf16_t>(bf16_t x) { return bf16_to_f32(x); }

template <typename T>
__device__ inline T from_f32(float x);
template <>
__device__ inline float from_f32<float>(float x) { return x; }
template <>
__device__ inline bf16_t from_f32<bf16_t>(float x) { return f32_to_bf16(x); }

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
#pragma unroll
        for (int v = 0; v < V; ++v) y[v] = from_f32<T>(to_f32<T>(x[v]) + 1.0f);
    }
}

// Strict-order FP64 sum of the taps for output column v (fold over the
// compile-time tap list, so every index and weight is an immediate).
template <int P, int R, size_t... I>
__device__ inline double tap_sum_f64(const double *const (&rows)[2 * R + 1], int v,
                                     std::index_sequence<I...>)
{
#pragma clang fp contract(off)
    constexpr const TapList &L = Stencil<P, R>::taps;
    double s = rows[L.t[0].di + R][v + L.t[0].dj] * L.t[0].w;
    ((s = s + rows[L.t[I + 1].di + R][v + L.t[I + 1].dj] * L.t[I + 1].w), ...);
    return s;
}

// FP32 paired sum: acc = w1*(x[+d1]-x[-d1]), then fma over the other pairs.
template <typename T, int P, int R, size_t... I>
__device__ inline float pair_sum_f32(const T *const (&rows)[2 * R + 1], int v,
                                     std::index_sequence<I...>)
{
#pragma clang fp contract(off)
    constexpr const TapList &L = Stencil<P, R>::pairs;
    float acc = (float)L.t[0].w * (to_f32<T>(rows[R + L.t[0].di][v + L.t[0].dj]) -
                                   to_f32<T>(rows[R - L.t[0].di][v - L.t[0].dj]));
    ((acc = __builtin_fmaf((float)L.t[I + 1].w,
                           to_f32<T>(rows[R + L.t[I + 1].di][v + L.t[I + 1].dj]) -
                               to_f32<T>(rows[R - L.t[I + 1].di][v - L.t[I + 1].dj]),
                           acc)),
     ...);
    return acc;
}

// The V outputs of one lane for output row o.  rows[k] points at the LDS row
// holding input row o-R+k, already offset to this lane's first column.
template <int DT, int P, int R, int V>
__device__ inline void compute_points(const typename DTypeTraits<DT>::T *const (&rows)[2 * R + 1],
                                      const typename DTypeTraits<DT>::T *outv,
                                      typename DTypeTraits<DT>::T *res)
{
#pragma clang fp contract(off)
    typedef typename DTypeTraits<DT>::T T;
    typedef Stencil<P, R> St;
    if constexpr (DT == F64) {
#pragma unroll
        for (int v = 0; v < V; ++v)
            res[v] = outv[v] + tap_sum_f64<P, R>(rows, v, std::make_index_sequence<St::taps.count - 1>{});
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const float acc = pair_sum_f32<T, P, R>(rows, v, std::make_index_sequence<St::pairs.count - 1>{});
            res[v] = from_f32<T>(to_f32<T>(outv[v]) + acc);
        }
    }
}

template <int DT, int P, int R, bool FUSE, bool VEC>
__global__ __launch_bounds__(kThreads) void sweep_kernel(SweepArgs a)
{
    typedef SweepGeom<DT, R> G;
    typedef typename G::T T;
    constexpr int V = G::V, SW = G::SW, HP = G::HP, NH = G::NH, LR = G::LR, S = G::S;
    __shared__ __attribute__((aligned(16))) T lds[kWavesPerBlock][S][LR];

    // XCD-aware tile order: workgroup b runs on XCD b%8; give each XCD a
    // contiguous run of tiles, tiles column-major (row blocks fastest).
    const int b = blockIdx.x;
    const int tile = (b & 7) * a.chunk + (b >> 3);
    if ((b >> 3) >= a.chunk || tile >= a.ntiles) return;
    const int group = tile / a.nrowblocks;
    const int rb = tile - group * a.nrowblocks;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int strip = group * kWavesPerBlock + wave;
    if (strip >= a.nstrips) return;

    const int n = a.n;
    const long long pitch = a.pitch;
    const int c0 = strip * SW;
    const int r0 = rb * a.rows_per_wave;
    const int r1 = min(r0 + a.rows_per_wave, n);
    const int lo = max(r0 - R, 0), hi = min(r1 + R, n);
    const int olo = max(r0, R), ohi = min(r1, n - R);

    const T *in = (const T *)a.in;
    T *in_next = (T *)a.in_next;
    T *out = (T *)a.out;

    const int cm = c0 + lane * V;  // first column of this lane's main vector
    // Halo vector of this lane (lanes 0..NH-1 left, NH..2NH-1 right).
    const bool has_halo = lane < 2 * NH;
    const int ch = lane < NH ? c0 - HP + lane * V : c0 + SW + (lane - NH) * V;
    const int lh = lane < NH ? lane * V : HP + SW + (lane - NH) * V;  // LDS column
    const bool main_ok = cm < n;
    const bool halo_ok = has_halo && ch >= 0 && ch < n;

    auto load_row = [&](RowRegs<T, V> &r, int ri) {
        const T *row = in + (long long)ri * pitch;
        if constexpr (VEC) {
            r.main.v = main_ok ? *(const u32x4 *)(row + cm) : u32x4{0, 0, 0, 0};
            r.halo.v = halo_ok ? *(const u32x4 *)(row + ch) : u32x4{0, 0, 0, 0};
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                r.main.e[v] = (cm + v < n) ? row[cm + v] : T(0);
                r.halo.e[v] = (has_halo && ch + v >= 0 && ch + v < n) ? row[ch + v] : T(0);
            }
        }
        const int o = ri - R;
        if (o >= olo && o < ohi) {
            const T *orow = out + (long long)o * pitch;
            if constexpr (VEC) {
                r.outv.v = main_ok ? *(const u32x4 *)(orow + cm) : u32x4{0, 0, 0, 0};
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v) r.outv.e[v] = (cm + v < n) ? orow[cm + v] : T(0);
            }
        }
    };

    T(*ring)[LR] = lds[wave];
    RowRegs<T, V> buf[kPrefetch];
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k)
        if (lo + k < hi) load_row(buf[k], lo + k);

    int slot = 0;  // ring slot of row ri: (ri - lo) % S
    for (int base = lo; base < hi; base += kPrefetch) {
#pragma unroll
        for (int k = 0; k < kPrefetch; ++k) {
            const int ri = base + k;
            if (ri < hi) {
                RowRegs<T, V> &r = buf[k];
                // park row ri in the ring
                *(u32x4 *)&ring[slot][HP + lane * V] = r.main.v;
                if (has_halo) *(u32x4 *)&ring[slot][lh] = r.halo.v;

                if (FUSE && ri >= r0 && ri < r1) {
                    union { u32x4 v; T e[V]; } nx;
                    plus_one<DT, V>(r.main.e, nx.e);
                    T *dst = in_next + (long long)ri * pitch + cm;
                    if (VEC && cm + V <= n) {
                        *(u32x4 *)dst = nx.v;
                    } else {
#pragma unroll
                        for (int v = 0; v < V; ++v)
                            if (cm + v < n) dst[v] = nx.e[v];
                    }
                }

                const int o = ri - R;
                const bool do_out = (o >= olo && o < ohi);
                RowRegs<T, V> cur = r;
                if (ri + kPrefetch < hi) load_row(r, ri + kPrefetch);

                // LDS writes of this wave land before its reads (in-order
                // per wave); the fences keep the compiler from moving them.
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

                if (do_out) {
                    // rows o-R .. o+R live in slots slot+1 .. slot+2R+1 (mod S)
                    const T *rows[S];
#pragma unroll
                    for (int k2 = 0; k2 < S; ++k2) {
                        int s2 = slot + 1 + k2;
                        s2 = s2 >= S ? s2 - S : s2;
                        rows[k2] = &ring[s2][HP + lane * V];
                    }
                    union { u32x4 v; T e[V]; } res;
                    compute_points<DT, P, R, V>(rows, cur.outv.e, res.e);
                    T *dst = out + (long long)o * pitch + cm;
                    if (VEC && cm >= R && cm + V <= n - R) {
                        *(u32x4 *)dst = res.v;
                    } else {
#pragma unroll
                        for (int v = 0; v < V; ++v)
                            if (cm + v >= R && cm + v < n - R) dst[v] = res.e[v];
                    }
                }
                slot = (slot + 1 == S) ? 0 : slot + 1;
            }
        }
    }
}

}  // namespace prk
