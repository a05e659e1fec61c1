// Shared definitions for the gfx950 stencil kernels: storage types, the
// stencil weight tables and tap order, the seeded initial-value generator.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace prk {

enum Pattern : int { STAR = 0, GRID = 1 };
enum DType : int { F64 = 0, F32 = 1, BF16 = 2 };

// Radii 1..5 are the reference's generated operators (Cxx11/stencil.cc:142-159);
// 6..8 follow the same weight formulas (the generator's loop bound, not its
// formulas, stops at 5: Cxx11/generate-cxx-stencil.py:245).
constexpr int kMaxRadius = 8;
constexpr int kRefMaxRadius = 5;

// ---------------------------------------------------------------------------
// Weights.  Cxx11/generate-cxx-stencil.py:209-232 assigns W[a+R][b+R], where a
// is the column offset and b the row offset of the emitted term
// in[(i+b)*n+(j+a)] (bodygen, :15-36):
//   star: W(0,±k) = W(±k,0) = ±1/(2kR)
//   grid, ring m = max(|a|,|b|): |a|<m -> sign(b)/(4m(2m-1)R);
//         |b|<m -> sign(a)/(4m(2m-1)R); a==b -> sign(a)/(4mR); a==-b -> 0.
// Evaluated in double at compile time; the division is IEEE, so each value is
// the double the generator printed (its literals round-trip).
// ---------------------------------------------------------------------------
__host__ __device__ constexpr double weight_ab(int pattern, int R, int a, int b)
{
    if (a == 0 && b == 0) return 0.0;
    if (pattern == STAR) {
        if (a == 0) return (b > 0 ? 1.0 : -1.0) / (2.0 * (b > 0 ? b : -b) * R);
        if (b == 0) return (a > 0 ? 1.0 : -1.0) / (2.0 * (a > 0 ? a : -a) * R);
        return 0.0;
    }
    const int aa = a > 0 ? a : -a, bb = b > 0 ? b : -b;
    const int m = aa > bb ? aa : bb;
    if (aa < m) return (b > 0 ? 1.0 : -1.0) / (double)(4 * m * (2 * m - 1) * R);
    if (bb < m) return (a > 0 ? 1.0 : -1.0) / (double)(4 * m * (2 * m - 1) * R);
    if (a == b) return (a > 0 ? 1.0 : -1.0) / (double)(4 * m * R);
    return 0.0;
}

// Rows of the array a sweep reads and the rows it refreshes.  Outputs are
// always rows [R, nrows-R) of the array.  A whole grid is rows [0, n), every
// row refreshed.  A slab (one share of a row-strip decomposition) is its own
// rows with up to R halo rows on each side: its first and last R rows are
// halo rows or the grid's frame, so its outputs are rows [R, nrows-R) too,
// and only its own rows [upd_lo, upd_hi) are refreshed.
struct RowSpan {
    int nrows;           // rows in the array
    int upd_lo, upd_hi;  // rows that get in_next = in + 1
};

__host__ __device__ inline RowSpan whole_grid(int n) { return RowSpan{n, 0, n}; }

struct Tap {
    int di, dj;  // row offset, column offset
    double w;
};

constexpr int kMaxTaps = (2 * kMaxRadius + 1) * (2 * kMaxRadius + 1);

struct TapList {
    int count;
    Tap t[kMaxTaps];
};

// Generator order (bodygen, :15-36): column offset outer, row offset inner,
// zero weights skipped.  The FP64 kernels sum the products in exactly this
// order, left to right, and add the sum to out last.
__host__ __device__ constexpr TapList make_taps(int pattern, int R)
{
    TapList L{};
    L.count = 0;
    for (int a = -R; a <= R; ++a)
        for (int b = -R; b <= R; ++b) {
            const double w = weight_ab(pattern, R, a, b);
            if (w != 0.0) L.t[L.count++] = Tap{b, a, w};
        }
    return L;
}

// Taps of positive weight.  Every stencil here is antisymmetric
// (W(-d) = -W(d)), so S(x)[p] = sum_k w_k (x[p+d_k] - x[p-d_k]) over these
// taps; the reduced-precision star forms evaluate that paired form.  Order: by
// row-offset magnitude |di| = 0..R, generator order within one magnitude, so
// a kernel needs only the two rows p±|di| at a time.
__host__ __device__ constexpr TapList make_pairs(int pattern, int R)
{
    const TapList all = make_taps(pattern, R);
    TapList P{};
    P.count = 0;
    for (int a = 0; a <= R; ++a)
        for (int k = 0; k < all.count; ++k) {
            const int ad = all.t[k].di < 0 ? -all.t[k].di : all.t[k].di;
            if (all.t[k].w > 0.0 && ad == a) P.t[P.count++] = all.t[k];
        }
    return P;
}

template <int P, int R>
struct Stencil {
    static constexpr TapList taps = make_taps(P, R);
    static constexpr TapList pairs = make_pairs(P, R);
    // Nominal stencil size of the reference's rate report (Cxx11/stencil.cc:221).
    static constexpr int nominal_size = (P == STAR) ? 4 * R + 1 : (2 * R + 1) * (2 * R + 1);
    // Row offsets a point reads beyond its own column (grid: every row).
    static constexpr bool needs_row_span(int di) { return P == GRID || di == 0; }
};

// ---------------------------------------------------------------------------
// Storage types.  Every grid row is moved in 16-byte vectors of V elements.
// ---------------------------------------------------------------------------
typedef uint16_t bf16_t;  // raw bfloat16 bit pattern

__host__ __device__ inline float bf16_to_f32(bf16_t h)
{
    return __uint_as_float((uint32_t)h << 16);
}

// Round to nearest even; NaN stays NaN.  The host oracle uses the same
// integer rounding, so the BF16 forms match it bit for bit.
__host__ __device__ inline bf16_t f32_to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) {
        uint16_t h = (uint16_t)(u >> 16);
        if (u & 0x007fffffu) h |= 0x0040u;
        return h;
    }
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

template <int DT> struct DTypeTraits;
template <> struct DTypeTraits<F64> {
    typedef double T;
    typedef double Acc;
    static constexpr int V = 2;
};
template <> struct DTypeTraits<F32> {
    typedef float T;
    typedef float Acc;
    static constexpr int V = 4;
};
template <> struct DTypeTraits<BF16> {
    typedef bf16_t T;
    typedef float Acc;
    static constexpr int V = 8;
};

inline size_t dtype_size(int dt) { return dt == F64 ? 8 : dt == F32 ? 4 : 2; }

// ---------------------------------------------------------------------------
// Seeded initial values: SplitMix64 of (seed, idx), top 53 bits as a uniform
// u in [0,1), value 2u-1 (exact in double).  Same on host and device.
// ---------------------------------------------------------------------------
__host__ __device__ inline double seeded_value(uint64_t seed, uint64_t idx)
{
    uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    const double u = (double)(z >> 11) * 0x1.0p-53;
    return 2.0 * u - 1.0;
}

}  // namespace prk
