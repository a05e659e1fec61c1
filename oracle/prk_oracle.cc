// CPU oracle for the Cxx11 stencil path.  TEST INFRASTRUCTURE ONLY: the
// checker, never the thing measured or shipped.  Built with g++ -O2
// -ffp-contract=off -fno-fast-math so every sum keeps the IEEE order
// written here.  Declarations and the reference citations are in
// prk_oracle.h.
#include "prk_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Weights.  Cxx11/generate-cxx-stencil.py:209-232 fills W[a+R][b+R] where the
// first index a is the COLUMN offset and the second index b the ROW offset of
// the emitted term in[(i+b)*n+(j+a)] (bodygen, :15-36).
//   star: W(0,±k) = W(±k,0) = ±1/(2kR)
//   grid, ring m = max(|a|,|b|):
//     |a| < m:          sign(b) / (4m(2m-1)R)    (:224-225)
//     |b| < m:          sign(a) / (4m(2m-1)R)    (:226-227)
//     a == b, |a| = m:  sign(a) / (4mR)          (:229-230)
//     a == -b:          0 (never assigned)
// ---------------------------------------------------------------------------
constexpr double weight_ab(int pattern, int R, int a, int b)
{
    if (a == 0 && b == 0) return 0.0;
    if (pattern == PRKO_STAR) {
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

struct Tap { int di, dj; double w; };

// radii 1..5: the reference's operators; 6..8: the same formulas extended
constexpr int max_radius = 8;
constexpr int max_taps = (2 * max_radius + 1) * (2 * max_radius + 1);

struct TapList {
    int count;
    Tap t[max_taps];
};

// Generator order: outer loop over the column offset (first W index), inner
// over the row offset, skipping zero weights (bodygen, :15-36).
constexpr TapList make_taps(int pattern, int R)
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

bool valid_pr(int pattern, int R)
{
    return (pattern == PRKO_STAR || pattern == PRKO_GRID) && R >= 1 && R <= max_radius;
}

// ---------------------------------------------------------------------------
// Strict-order point update shared by the serial and OpenMP forms.
// out = out + ((((t1 + t2) + t3) + ...)), tk = in[...] * wk
// ---------------------------------------------------------------------------
template <int P, int R>
struct Kernel {
    static constexpr TapList L = make_taps(P, R);

    template <size_t... I>
    static inline double sum(const double *__restrict in, long p, long n, std::index_sequence<I...>)
    {
        double s = in[p + L.t[0].di * n + L.t[0].dj] * L.t[0].w;
        ((s = s + in[p + L.t[I + 1].di * n + L.t[I + 1].dj] * L.t[I + 1].w), ...);
        return s;
    }

    static inline double point(const double *__restrict in, long p, long n)
    {
        return sum(in, p, n, std::make_index_sequence<L.count - 1>{});
    }

    // Tiled loop nest of Cxx11/stencil_seq.hpp / stencil_openmp.hpp.  The
    // OpenMP variant is orphaned: it must run inside a parallel region.
    static void tiled(int n, int t, const double *__restrict in, double *__restrict out, bool omp)
    {
        if (omp) {
#ifdef _OPENMP
#pragma omp for collapse(2)
#endif
            for (int it = R; it < n - R; it += t)
                for (int jt = R; jt < n - R; jt += t)
                    for (int i = it; i < std::min(n - R, it + t); ++i) {
#ifdef _OPENMP
#pragma omp simd
#endif
                        for (int j = jt; j < std::min(n - R, jt + t); ++j) {
                            const long p = (long)i * n + j;
                            out[p] = out[p] + point(in, p, n);
                        }
                    }
        } else {
            for (int it = R; it < n - R; it += t)
                for (int jt = R; jt < n - R; jt += t)
                    for (int i = it; i < std::min(n - R, it + t); ++i)
                        for (int j = jt; j < std::min(n - R, jt + t); ++j) {
                            const long p = (long)i * n + j;
                            out[p] = out[p] + point(in, p, n);
                        }
        }
    }
};

using TiledFn = void (*)(int, int, const double *, double *, bool);

// The (pattern, radius) -> operator switch of Cxx11/stencil.cc:142-159.
TiledFn select(int pattern, int R)
{
    if (pattern == PRKO_STAR) {
        switch (R) {
        case 1: return Kernel<PRKO_STAR, 1>::tiled;
        case 2: return Kernel<PRKO_STAR, 2>::tiled;
        case 3: return Kernel<PRKO_STAR, 3>::tiled;
        case 4: return Kernel<PRKO_STAR, 4>::tiled;
        case 5: return Kernel<PRKO_STAR, 5>::tiled;
        case 6: return Kernel<PRKO_STAR, 6>::tiled;
        case 7: return Kernel<PRKO_STAR, 7>::tiled;
        case 8: return Kernel<PRKO_STAR, 8>::tiled;
        }
    } else if (pattern == PRKO_GRID) {
        switch (R) {
        case 1: return Kernel<PRKO_GRID, 1>::tiled;
        case 2: return Kernel<PRKO_GRID, 2>::tiled;
        case 3: return Kernel<PRKO_GRID, 3>::tiled;
        case 4: return Kernel<PRKO_GRID, 4>::tiled;
        case 5: return Kernel<PRKO_GRID, 5>::tiled;
        case 6: return Kernel<PRKO_GRID, 6>::tiled;
        case 7: return Kernel<PRKO_GRID, 7>::tiled;
        case 8: return Kernel<PRKO_GRID, 8>::tiled;
        }
    }
    return nullptr;
}

// ---------------------------------------------------------------------------
// Reduced-precision storage forms.
// ---------------------------------------------------------------------------
inline uint16_t f32_to_bf16(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) {  // inf / nan: truncate, keep nan quiet
        uint16_t h = (uint16_t)(u >> 16);
        if (u & 0x007fffffu) h |= 0x0040u;
        return h;
    }
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

inline float bf16_to_f32(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

template <typename T> inline float ld(const T *p, long i);
template <> inline float ld<float>(const float *p, long i) { return p[i]; }
template <> inline float ld<uint16_t>(const uint16_t *p, long i) { return bf16_to_f32(p[i]); }
template <typename T> inline T st(float x);
template <> inline float st<float>(float x) { return x; }
template <> inline uint16_t st<uint16_t>(float x) { return f32_to_bf16(x); }

struct Pair { int di, dj; float w; };

// Star stencils: the taps of positive weight, weights rounded to float,
// ordered by row-offset magnitude |di| = 0..R and in generator order within
// one magnitude.
std::vector<Pair> make_pairs(int pattern, int R)
{
    const TapList L = make_taps(pattern, R);
    std::vector<Pair> v;
    for (int a = 0; a <= R; ++a)
        for (int k = 0; k < L.count; ++k)
            if (L.t[k].w > 0.0 && std::abs(L.t[k].di) == a)
                v.push_back(Pair{L.t[k].di, L.t[k].dj, (float)L.t[k].w});
    return v;
}

// The reduced-precision operator.  x(a, b) is the input at row offset a,
// column offset b from the point, widened to float.
struct LpOp {
    int pattern, R;
    std::vector<Pair> pairs;  // star
    float ce[max_radius + 1], cc[max_radius + 1];  // grid: edge and corner weight of ring j

    LpOp(int p, int r) : pattern(p), R(r), pairs(make_pairs(p, r))
    {
        for (int j = 1; j <= R; ++j) {
            ce[j] = (float)weight_ab(PRKO_GRID, R, 0, j);  // taps (+-j, |dj| < j) and (|di| < j, +-j)
            cc[j] = (float)weight_ab(PRKO_GRID, R, j, j);  // taps (j, j) and (-j, -j)
        }
    }

    template <typename X>
    float point(const X &x) const
    {
        float acc = 0.0f;
        if (pattern == PRKO_STAR) {
            for (size_t k = 0; k < pairs.size(); ++k) {
                const float diff = x(pairs[k].di, pairs[k].dj) - x(-pairs[k].di, -pairs[k].dj);
                acc = (k == 0) ? pairs[k].w * diff : std::fmaf(pairs[k].w, diff, acc);
            }
            return acc;
        }
        // Grid: ring j = max(|di|, |dj|) has one weight on its edges,
        // +ce[j] on column +j and row +j, -ce[j] on column -j and row -j
        // (the corners excluded), and +-cc[j] on the corners (j, j) and
        // (-j, -j).  So S = sum_j ce[j] * (V_j + H_j) + cc[j] * C_j with
        //   V_j = VS_{j-1}(+j) - VS_{j-1}(-j), VS_k(b) the sum of column b
        //         over rows -k..k: VS_0(b) = x(0, b),
        //         VS_k(b) = VS_{k-1}(b) + (x(k, b) + x(-k, b));
        //   H_j = D_j(0), then + D_j(-m), + D_j(+m) for m = 1..j-1, with
        //         D_j(b) = x(j, b) - x(-j, b);
        //   C_j = x(j, j) - x(-j, -j);
        // acc = ce[1] * (H_1 + V_1), then fmaf(cc[j], C_j, acc) and
        // fmaf(ce[j+1], H_{j+1} + V_{j+1}, acc) in ring order.
        float vs[2 * max_radius + 1];
        for (int b = -R; b <= R; ++b) vs[b + R] = x(0, b);
        for (int j = 1; j <= R; ++j) {
            const float vj = vs[R + j] - vs[R - j];
            float hj = x(j, 0) - x(-j, 0);
            for (int m = 1; m < j; ++m) {
                hj = hj + (x(j, -m) - x(-j, -m));
                hj = hj + (x(j, m) - x(-j, m));
            }
            const float tj = hj + vj;
            const float cj = x(j, j) - x(-j, -j);
            acc = (j == 1) ? ce[j] * tj : std::fmaf(ce[j], tj, acc);
            acc = std::fmaf(cc[j], cj, acc);
            if (j < R)
                for (int b = -R; b <= R; ++b) vs[b + R] = vs[b + R] + (x(j, b) + x(-j, b));
        }
        return acc;
    }
};

template <typename T>
void lp_sweep(const LpOp &op, int R, int n, int row_lo, int row_hi, long row0, const T *in, T *out)
{
    // rows [row_lo,row_hi) of the global grid; arrays start at global row row0
    for (int i = std::max(row_lo, R); i < std::min(row_hi, n - R); ++i)
        for (int j = R; j < n - R; ++j) {
            const long p = (long)(i - row0) * n + j;
            const float acc = op.point([&](int a, int b) { return ld(in, p + (long)a * n + b); });
            out[p] = st<T>(ld(out, p) + acc);
        }
}

template <typename T>
void lp_init_row(int init, uint64_t seed, int n, int i, T *in_row, T *out_row)
{
    for (int j = 0; j < n; ++j) {
        const float v = (init == PRKO_INIT_ANALYTIC)
                            ? (float)(i + j)
                            : (float)prko_seeded_value(seed, (uint64_t)i * n + j);
        in_row[j] = st<T>(v);
        if (out_row) out_row[j] = st<T>(0.0f);
    }
}

template <typename T>
int lp_band(int pattern, int R, int n, int K, int init, uint64_t seed, int a, int b, T *out_band,
            T *in_final_band)
{
    const int lo = std::max(a - R, 0), hi = std::min(b + R, n);
    const long rows = hi - lo;
    std::vector<T> in((size_t)rows * n), out((size_t)rows * n);
    for (int i = lo; i < hi; ++i)
        lp_init_row<T>(init, seed, n, i, &in[(size_t)(i - lo) * n], &out[(size_t)(i - lo) * n]);
    const LpOp op(pattern, R);
    for (int it = 0; it <= K; ++it) {
        lp_sweep<T>(op, R, n, a, b, lo, in.data(), out.data());
        for (auto &x : in) x = st<T>(ld(&x, 0) + 1.0f);
    }
    std::memcpy(out_band, &out[(size_t)(a - lo) * n], sizeof(T) * (size_t)(b - a) * n);
    if (in_final_band)
        std::memcpy(in_final_band, &in[(size_t)(a - lo) * n], sizeof(T) * (size_t)(b - a) * n);
    return 0;
}

}  // namespace

extern "C" {

double prko_weight(int pattern, int radius, int di, int dj)
{
    if (!valid_pr(pattern, radius) || di < -radius || di > radius || dj < -radius || dj > radius)
        return 0.0;
    return weight_ab(pattern, radius, dj, di);
}

int prko_taps(int pattern, int radius, int *di, int *dj, double *w)
{
    if (!valid_pr(pattern, radius)) return -1;
    const TapList L = make_taps(pattern, radius);
    for (int k = 0; k < L.count; ++k) {
        if (di) di[k] = L.t[k].di;
        if (dj) dj[k] = L.t[k].dj;
        if (w) w[k] = L.t[k].w;
    }
    return L.count;
}

double prko_seeded_value(uint64_t seed, uint64_t idx)
{
    uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    const double u = (double)(z >> 11) * 0x1.0p-53;
    return 2.0 * u - 1.0;
}

int prko_init(int init, uint64_t seed, int n, double *in, double *out)
{
    if (n < 1 || (init != PRKO_INIT_ANALYTIC && init != PRKO_INIT_SEEDED)) return -1;
    for (long i = 0; i < n; ++i)
        for (long j = 0; j < n; ++j) {
            in[i * n + j] = (init == PRKO_INIT_ANALYTIC) ? (double)(i + j)
                                                         : prko_seeded_value(seed, (uint64_t)(i * n + j));
            if (out) out[i * n + j] = 0.0;
        }
    return 0;
}

int prko_stencil(int pattern, int radius, int n, int t, const double *in, double *out)
{
    TiledFn f = select(pattern, radius);
    if (!f || n < 1 || t < 1 || 2 * radius + 1 > n) return -1;
    f(n, t, in, out, false);
    return 0;
}

int prko_add(int n, double *in)
{
    if (n < 1) return -1;
    const long nn = (long)n * n;
    for (long p = 0; p < nn; ++p) in[p] = in[p] + 1.0;
    return 0;
}

double prko_l1norm(int n, int radius, const double *out)
{
    const long active = (long)(n - 2 * radius) * (long)(n - 2 * radius);
    double norm = 0.0;
    for (long i = radius; i < n - radius; ++i)
        for (long j = radius; j < n - radius; ++j) norm += std::fabs(out[i * n + j]);
    return norm / (double)active;
}

int prko_run(int pattern, int radius, int n, int t, int iterations, int init, uint64_t seed,
             double *out, double *in_final)
{
    TiledFn f = select(pattern, radius);
    if (!f || n < 1 || t < 1 || iterations < 0 || 2 * radius + 1 > n) return -1;
    std::vector<double> in((size_t)n * n);
    if (prko_init(init, seed, n, in.data(), out)) return -1;
    for (int it = 0; it <= iterations; ++it) {
        f(n, t, in.data(), out, false);
        prko_add(n, in.data());
    }
    if (in_final) std::memcpy(in_final, in.data(), sizeof(double) * in.size());
    return 0;
}

int prko_band(int pattern, int radius, int n, int iterations, int init, uint64_t seed, int a,
              int b, double *out_band)
{
    TiledFn f = select(pattern, radius);
    if (!f || n < 1 || iterations < 0 || 2 * radius + 1 > n || a < 0 || b > n || a >= b) return -1;
    if (init != PRKO_INIT_ANALYTIC && init != PRKO_INIT_SEEDED) return -1;
    // Local sub-grid: input rows [lo,hi), all n columns.  The band's output
    // rows are interior rows of the local grid exactly when they are
    // interior rows of the global one, so the local sweep is the global
    // sweep restricted to these rows.
    const int lo = std::max(a - radius, 0), hi = std::min(b + radius, n);
    const int rows = hi - lo;
    std::vector<double> in((size_t)rows * n), out((size_t)rows * n, 0.0);
    for (long i = lo; i < hi; ++i)
        for (long j = 0; j < n; ++j)
            in[(size_t)(i - lo) * n + j] = (init == PRKO_INIT_ANALYTIC)
                                               ? (double)(i + j)
                                               : prko_seeded_value(seed, (uint64_t)(i * n + j));
    const TapList *L = nullptr;
    static const TapList tables[2][max_radius] = {
        {make_taps(0, 1), make_taps(0, 2), make_taps(0, 3), make_taps(0, 4), make_taps(0, 5), make_taps(0, 6),
         make_taps(0, 7), make_taps(0, 8)},
        {make_taps(1, 1), make_taps(1, 2), make_taps(1, 3), make_taps(1, 4), make_taps(1, 5), make_taps(1, 6),
         make_taps(1, 7), make_taps(1, 8)}};
    L = &tables[pattern][radius - 1];
    const int r0 = std::max(a, radius), r1 = std::min(b, n - radius);
    for (int it = 0; it <= iterations; ++it) {
        for (long i = r0; i < r1; ++i)
            for (long j = radius; j < n - radius; ++j) {
                const long p = (i - lo) * n + j;
                double s = in[p + L->t[0].di * (long)n + L->t[0].dj] * L->t[0].w;
                for (int k = 1; k < L->count; ++k)
                    s = s + in[p + L->t[k].di * (long)n + L->t[k].dj] * L->t[k].w;
                out[p] = out[p] + s;
            }
        for (auto &x : in) x = x + 1.0;
    }
    std::memcpy(out_band, &out[(size_t)(a - lo) * n], sizeof(double) * (size_t)(b - a) * n);
    return 0;
}

int prko_run_lp(int dtype, int pattern, int radius, int n, int iterations, int init,
                uint64_t seed, void *out, void *in_final)
{
    if (!valid_pr(pattern, radius) || n < 1 || iterations < 0 || 2 * radius + 1 > n) return -1;
    if (init != PRKO_INIT_ANALYTIC && init != PRKO_INIT_SEEDED) return -1;
    if (dtype == PRKO_F32)
        return lp_band<float>(pattern, radius, n, iterations, init, seed, 0, n, (float *)out,
                              (float *)in_final);
    if (dtype == PRKO_BF16)
        return lp_band<uint16_t>(pattern, radius, n, iterations, init, seed, 0, n,
                                 (uint16_t *)out, (uint16_t *)in_final);
    return -1;
}

int prko_band_lp(int dtype, int pattern, int radius, int n, int iterations, int init,
                 uint64_t seed, int a, int b, void *out_band)
{
    if (!valid_pr(pattern, radius) || n < 1 || iterations < 0 || 2 * radius + 1 > n) return -1;
    if (init != PRKO_INIT_ANALYTIC && init != PRKO_INIT_SEEDED) return -1;
    if (a < 0 || b > n || a >= b) return -1;
    if (dtype == PRKO_F32)
        return lp_band<float>(pattern, radius, n, iterations, init, seed, a, b, (float *)out_band,
                              nullptr);
    if (dtype == PRKO_BF16)
        return lp_band<uint16_t>(pattern, radius, n, iterations, init, seed, a, b,
                                 (uint16_t *)out_band, nullptr);
    return -1;
}

int prko_sweep_lp(int dtype, int pattern, int radius, int n, const void *in, void *out)
{
    if (!valid_pr(pattern, radius) || n < 1 || 2 * radius + 1 > n) return -1;
    const LpOp op(pattern, radius);
    if (dtype == PRKO_F32) {
        lp_sweep<float>(op, radius, n, 0, n, 0, (const float *)in, (float *)out);
        return 0;
    }
    if (dtype == PRKO_BF16) {
        lp_sweep<uint16_t>(op, radius, n, 0, n, 0, (const uint16_t *)in, (uint16_t *)out);
        return 0;
    }
    return -1;
}

double prko_l1norm_lp(int dtype, int n, int radius, const void *out)
{
    const long active = (long)(n - 2 * radius) * (long)(n - 2 * radius);
    double norm = 0.0;
    for (long i = radius; i < n - radius; ++i)
        for (long j = radius; j < n - radius; ++j) {
            const double v = (dtype == PRKO_F32) ? (double)((const float *)out)[i * n + j]
                                                 : (double)bf16_to_f32(((const uint16_t *)out)[i * n + j]);
            norm += std::fabs(v);
        }
    return norm / (double)active;
}

uint16_t prko_f32_to_bf16(float x) { return f32_to_bf16(x); }
float prko_bf16_to_f32(uint16_t h) { return bf16_to_f32(h); }

int prko_omp_run(int pattern, int radius, int n, int t, int iterations, int threads, double *avg_s,
                 double *norm_out, int *threads_used)
{
    TiledFn f = select(pattern, radius);
    if (!f || n < 1 || t < 1 || iterations < 1 || 2 * radius + 1 > n) return -1;
    double *in = new double[(size_t)n * n];
    double *out = new double[(size_t)n * n];
    double t0 = 0.0, t1 = 0.0;
    int team = 1;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
#endif
    {
#ifdef _OPENMP
#pragma omp master
        team = omp_get_num_threads();
#pragma omp for collapse(2)
#endif
        for (int it = 0; it < n; it += t)
            for (int jt = 0; jt < n; jt += t)
                for (int i = it; i < std::min(n, it + t); i++)
                    for (int j = jt; j < std::min(n, jt + t); j++) {
                        in[(long)i * n + j] = (double)(i + j);
                        out[(long)i * n + j] = 0.0;
                    }
        for (int iter = 0; iter <= iterations; iter++) {
            if (iter == 1) {
#ifdef _OPENMP
#pragma omp barrier
#pragma omp master
                t0 = omp_get_wtime();
#endif
            }
            f(n, t, in, out, true);
#ifdef _OPENMP
#pragma omp for collapse(2)
#endif
            for (int it = 0; it < n; it += t)
                for (int jt = 0; jt < n; jt += t)
                    for (int i = it; i < std::min(n, it + t); i++)
                        for (int j = jt; j < std::min(n, jt + t); j++) in[(long)i * n + j] += 1.0;
        }
#ifdef _OPENMP
#pragma omp barrier
#pragma omp master
        t1 = omp_get_wtime();
#endif
    }
    double norm = 0.0;
    const long active = (long)(n - 2 * radius) * (long)(n - 2 * radius);
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : norm)
#endif
    for (int i = radius; i < n - radius; i++)
        for (int j = radius; j < n - radius; j++) norm += std::fabs(out[(long)i * n + j]);
    norm /= (double)active;
    delete[] in;
    delete[] out;
    if (avg_s) *avg_s = (t1 - t0) / iterations;
    if (norm_out) *norm_out = norm;
    if (threads_used) *threads_used = team;
    return 0;
}

int prko_transpose(int order, int tile, double *A, double *B)
{
    if (order < 1 || !A || !B || A == B) return -1;
    if (tile <= 0 || tile >= order) tile = order;
    for (long it = 0; it < order; it += tile)
        for (long jt = 0; jt < order; jt += tile)
            for (long i = it; i < std::min<long>(order, it + tile); i++)
                for (long j = jt; j < std::min<long>(order, jt + tile); j++) {
                    B[i * order + j] += A[j * order + i];
                    A[j * order + i] += 1.0;
                }
    return 0;
}

int prko_transpose_run(int order, int tile, int iterations, double *A, double *B)
{
    if (order < 1 || iterations < 0 || !A || !B || A == B) return -1;
    const long nn = (long)order * order;
    for (long k = 0; k < nn; ++k) {
        A[k] = (double)k;
        B[k] = 0.0;
    }
    for (int iter = 0; iter <= iterations; ++iter)
        if (prko_transpose(order, tile, A, B)) return -1;
    return 0;
}

double prko_transpose_abserr(int order, int iterations, const double *B)
{
    const double addit = (iterations + 1.) * (iterations / 2.);
    double abserr = 0.0;
    for (long j = 0; j < order; j++)
        for (long i = 0; i < order; i++) {
            const long ij = i * order + j, ji = j * order + i;
            abserr += std::fabs(B[ji] - ((double)ij * (1. + iterations) + addit));
        }
    return abserr;
}

}  // extern "C"
