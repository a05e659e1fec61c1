// Autotuner for the fused sweep: for a (pattern, radius, dtype, n) it times
// every candidate tiling on the device -- pipeline variant (depth 1..4 rows,
// temporal or non-temporal streams, and for the FP32/BF16 forms at R <= 3
// the register-window sweep), row-block height per wavefront and
// resident workgroups per CU -- and records the fastest in a process-wide
// table that plan_sweep() consults.  The table can be saved to and loaded
// from a text file.  Tiling never changes results, only speed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "prk_internal.hpp"
#include "sweep.hpp"

namespace prk {

namespace {

typedef std::tuple<int, int, int, int> Key;  // pattern, radius, dtype, n
std::mutex g_mu;
std::map<Key, prk_tuning> g_table;

const int kCandVariants[] = {1, 2, 4, 5, 6, 7, 8};
const int kCandWgs[] = {2, 3, 4};
const int kCandRows[] = {4, 6, 8, 12, 16, 24, 32, 48, 64};
// Row blocks beyond 32 only for R >= 4, where the 2R rows each block re-reads
// are a large share (star R=8 FP64: 48 rows 3 % faster than 24); below that
// they never won and would only lengthen the search.
constexpr int kTallRowsMinRadius = 4;
constexpr int kFinalists = 4;  // candidates timed again in the second stage

}  // namespace

// Measured defaults on MI355X (n = 16384 and 30000): a 3-row non-temporal
// pipeline, two workgroups per CU (four for BF16, whose eight columns per
// lane make each row step longer), row blocks growing with the radius so
// the 2R halo rows each block re-reads stay a small fraction.
prk_tuning heuristic_tuning(int pattern, int radius, int dtype, int n)
{
    (void)pattern;
    prk_tuning t;
    t.variant = 5;
    t.order = 1;
    t.wgs_per_cu = dtype == BF16 ? 4 : 2;
    int rows = dtype == BF16 ? 12 : (radius <= 2 ? 8 : radius == 3 ? 12 : radius == 4 ? 16 : 24);
    t.rows_per_wave = std::max(1, std::min(rows, n));
    t.gbs = 0.0;
    return t;
}

prk_tuning choose_tuning(int pattern, int radius, int dtype, int n)
{
    {
        std::lock_guard<std::mutex> lock(g_mu);
        auto it = g_table.find(Key(pattern, radius, dtype, n));
        if (it != g_table.end()) return it->second;
        // nearest tuned size of the same stencil and dtype within a factor 1.5
        const prk_tuning *best = nullptr;
        double bestd = std::log(1.5);
        for (const auto &kv : g_table) {
            if (std::get<0>(kv.first) != pattern || std::get<1>(kv.first) != radius ||
                std::get<2>(kv.first) != dtype)
                continue;
            const double d = std::fabs(std::log((double)std::get<3>(kv.first) / n));
            if (d <= bestd) {
                bestd = d;
                best = &kv.second;
            }
        }
        if (best) {
            prk_tuning t = *best;
            t.rows_per_wave = std::max(1, std::min(t.rows_per_wave, n));
            return t;
        }
    }
    return heuristic_tuning(pattern, radius, dtype, n);
}

}  // namespace prk

using namespace prk;

extern "C" {

int prk_tuning_get(int pattern, int radius, int dtype, int n, prk_tuning *t)
{
    if (!t) return set_error(PRK_EINVAL, "null tuning");
    *t = choose_tuning(pattern, radius, dtype, n);
    return PRK_OK;
}

int prk_tuning_set(int pattern, int radius, int dtype, int n, const prk_tuning *t)
{
    if (!t || t->variant < 0 || t->variant >= kSweepVariants || t->rows_per_wave < 1 || t->wgs_per_cu < 0 ||
        t->wgs_per_cu > 16 || t->order < 0 || t->order > 7)
        return set_error(PRK_EINVAL, "bad tuning");
    std::lock_guard<std::mutex> lock(g_mu);
    g_table[Key(pattern, radius, dtype, n)] = *t;
    return PRK_OK;
}

int prk_tuning_clear(void)
{
    std::lock_guard<std::mutex> lock(g_mu);
    g_table.clear();
    return PRK_OK;
}

int prk_tuning_save(const char *path)
{
    FILE *f = path ? std::fopen(path, "w") : nullptr;
    if (!f) return set_error(PRK_EINVAL, "cannot write %s", path ? path : "(null)");
    std::lock_guard<std::mutex> lock(g_mu);
    std::fprintf(f, "# pattern radius dtype n variant rows_per_wave wgs_per_cu order gbs\n");
    for (const auto &kv : g_table)
        std::fprintf(f, "%d %d %d %d %d %d %d %d %.1f\n", std::get<0>(kv.first), std::get<1>(kv.first),
                     std::get<2>(kv.first), std::get<3>(kv.first), kv.second.variant, kv.second.rows_per_wave,
                     kv.second.wgs_per_cu, kv.second.order, kv.second.gbs);
    std::fclose(f);
    return PRK_OK;
}

int prk_tuning_load(const char *path)
{
    FILE *f = path ? std::fopen(path, "r") : nullptr;
    if (!f) return set_error(PRK_EINVAL, "cannot read %s", path ? path : "(null)");
    char line[256];
    int loaded = 0;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        int p, r, d, n;
        prk_tuning t;
        if (std::sscanf(line, "%d %d %d %d %d %d %d %d %lf", &p, &r, &d, &n, &t.variant, &t.rows_per_wave,
                        &t.wgs_per_cu, &t.order, &t.gbs) == 9 &&
            prk_tuning_set(p, r, d, n, &t) == PRK_OK)
            ++loaded;
    }
    std::fclose(f);
    return loaded;
}

}  // extern "C"

namespace {

// The search on three caller grids of n rows at `pitch` (contents
// overwritten), on stream s; records the winner in the table.
int autotune_on(int pattern, int radius, int dtype, int n, long long pitch, void *b0, void *b1, void *out, int reps,
                hipStream_t s, prk_tuning *best)
{
    const size_t es = dtype_size(dtype);
    void *bufs[2] = {b0, b1};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = check_hip(hipEventCreate(&e0), "autotune event");
    if (rc == PRK_OK) rc = check_hip(hipEventCreate(&e1), "autotune event");
    if (rc == PRK_OK) rc = prk_init(PRK_INIT_ANALYTIC, 0, dtype, n, pitch, b0, out, s);
    const double bytes_iter = 2.0 * es * (double)n * n + 2.0 * es * (double)(n - 2 * radius) * (n - 2 * radius);
    int cur = 0;
    // Time `r` sweeps of one tiling (after one untimed sweep); GB/s or < 0.
    auto time_tiling = [&](const prk_tuning &t, int r, double *gbs) -> int {
        SweepPlan p;
        int e = plan_sweep(pattern, radius, dtype, n, pitch, b0, b1, out, t.rows_per_wave, &p);
        if (e) return e;
        p.variant = t.variant;
        p.wgs_per_cu = t.wgs_per_cu;
        p.order = t.order;
        p.rows_per_wave = t.rows_per_wave;
        e = launch_sweep(p, bufs[cur], bufs[cur ^ 1], out, s);  // warm-up
        cur ^= 1;
        if (!e) e = check_hip(hipEventRecord(e0, s), "event");
        for (int k = 0; k < r && !e; ++k) {
            e = launch_sweep(p, bufs[cur], bufs[cur ^ 1], out, s);
            cur ^= 1;
        }
        if (!e) e = check_hip(hipEventRecord(e1, s), "event");
        if (!e) e = check_hip(hipEventSynchronize(e1), "event sync");
        float ms = 0.0f;
        if (!e) e = check_hip(hipEventElapsedTime(&ms, e0, e1), "event time");
        *gbs = (!e && ms > 0.0f) ? bytes_iter * r / (ms * 1e-3) / 1e9 : -1.0;
        return e;
    };
    // Stage 1: every candidate, `reps` sweeps each.
    std::vector<prk_tuning> cands;
    for (int v : kCandVariants) {
        // beyond the reference's radii one variant is built (see launch_pr);
        // 6 and 7 exist only where the register-window sweep applies
        if (radius > kRefMaxRadius && v != 4) continue;
        if (v >= 6 && !register_window_variant(v, radius, dtype)) continue;
        for (int w : kCandWgs) {
            for (int rows : kCandRows) {
                if (rc != PRK_OK) break;
                if (rows > n || (rows > 32 && radius < kTallRowsMinRadius)) continue;
                prk_tuning t{v, rows, w, 1, 0.0};
                rc = time_tiling(t, reps, &t.gbs);
                if (rc == PRK_OK && t.gbs > 0.0) cands.push_back(t);
            }
        }
    }
    // Stage 2: the fastest few again, over four times as many sweeps, twice
    // each in turn; a short sample alone picks noise as often as a tiling.
    std::sort(cands.begin(), cands.end(), [](const prk_tuning &a, const prk_tuning &b) { return a.gbs > b.gbs; });
    if (cands.size() > (size_t)kFinalists) cands.resize(kFinalists);
    std::vector<double> sum(cands.size(), 0.0);
    for (int pass = 0; pass < 2 && rc == PRK_OK; ++pass)
        for (size_t i = 0; i < cands.size() && rc == PRK_OK; ++i) {
            double g = 0.0;
            rc = time_tiling(cands[i], 4 * reps, &g);
            sum[i] += g;
        }
    prk_tuning top{};
    top.gbs = -1.0;
    for (size_t i = 0; i < cands.size() && rc == PRK_OK; ++i)
        if (sum[i] / 2 > top.gbs) {
            top = cands[i];
            top.gbs = sum[i] / 2;
        }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (top.gbs <= 0.0) return set_error(PRK_EINVAL, "no tiling candidate fits n=%d", n);
    rc = prk_tuning_set(pattern, radius, dtype, n, &top);
    if (rc) return rc;
    if (best) *best = top;
    return PRK_OK;
}

}  // namespace

extern "C" {

int prk_autotune(int pattern, int radius, int dtype, int n, int reps, prk_tuning *best)
{
    if (reps < 1) return set_error(PRK_EINVAL, "reps must be >= 1");
    if (dtype != PRK_F64 && dtype != PRK_F32 && dtype != PRK_BF16) return set_error(PRK_EINVAL, "unknown dtype %d", dtype);
    if (n < 1) return set_error(PRK_EINVAL, "bad n %d", n);
    const size_t es = dtype_size(dtype);
    const long long pitch = ((long long)n * (long long)es + 15) / 16 * 16 / (long long)es;
    const size_t bytes = (size_t)n * (size_t)pitch * es;
    void *buf[3] = {nullptr, nullptr, nullptr};
    int rc = PRK_OK;
    for (int k = 0; k < 3 && rc == PRK_OK; ++k) rc = check_hip(hipMalloc(&buf[k], bytes), "autotune alloc");
    hipStream_t s = nullptr;
    if (rc == PRK_OK) rc = check_hip(hipStreamCreate(&s), "autotune stream");
    if (rc == PRK_OK) rc = autotune_on(pattern, radius, dtype, n, pitch, buf[0], buf[1], buf[2], reps, s, best);
    if (s) (void)hipStreamDestroy(s);
    for (void *b : buf)
        if (b) (void)hipFree(b);
    return rc;
}

int prk_autotune_buffers(int pattern, int radius, int dtype, int n, long long pitch, void *buf0, void *buf1,
                         void *out, int reps, void *stream, prk_tuning *best)
{
    if (reps < 1) return set_error(PRK_EINVAL, "reps must be >= 1");
    if (!buf0 || !buf1 || !out || buf0 == buf1 || buf0 == out || buf1 == out)
        return set_error(PRK_EINVAL, "autotune needs three separate grids");
    if (dtype != PRK_F64 && dtype != PRK_F32 && dtype != PRK_BF16) return set_error(PRK_EINVAL, "unknown dtype %d", dtype);
    if (n < 1 || pitch < n) return set_error(PRK_EINVAL, "bad grid (n=%d pitch=%lld)", n, pitch);
    return autotune_on(pattern, radius, dtype, n, pitch, buf0, buf1, out, reps, (hipStream_t)stream, best);
}

}  // extern "C"
