// Autotuner for the fused sweep: for a (pattern, radius, dtype, n) it times
// every candidate tiling on the device -- pipeline variant (depth 1..3 rows,
// temporal or non-temporal streams), row-block height per wavefront and
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

const int kCandVariants[] = {1, 2, 4, 5};
const int kCandWgs[] = {2, 3, 4};
const int kCandRows[] = {4, 6, 8, 12, 16, 24, 32};

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

int prk_autotune(int pattern, int radius, int dtype, int n, int reps, prk_tuning *best)
{
    if (reps < 1) return set_error(PRK_EINVAL, "reps must be >= 1");
    const size_t es = dtype_size(dtype);
    if (n < 1) return set_error(PRK_EINVAL, "bad n %d", n);
    const long long pitch = ((long long)n * (long long)es + 15) / 16 * 16 / (long long)es;
    const size_t bytes = (size_t)n * (size_t)pitch * es;
    void *buf[3] = {nullptr, nullptr, nullptr};
    int rc = PRK_OK;
    for (int k = 0; k < 3 && rc == PRK_OK; ++k) rc = check_hip(hipMalloc(&buf[k], bytes), "autotune alloc");
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == PRK_OK) rc = check_hip(hipStreamCreate(&s), "autotune stream");
    if (rc == PRK_OK) rc = check_hip(hipEventCreate(&e0), "autotune event");
    if (rc == PRK_OK) rc = check_hip(hipEventCreate(&e1), "autotune event");
    if (rc == PRK_OK) rc = prk_init(PRK_INIT_ANALYTIC, 0, dtype, n, pitch, buf[0], buf[2], s);
    const double bytes_iter = 2.0 * es * (double)n * n + 2.0 * es * (double)(n - 2 * radius) * (n - 2 * radius);
    prk_tuning top{};
    top.gbs = -1.0;
    int cur = 0;
    for (int v : kCandVariants) {
        // beyond the reference's radii one variant is built (see launch_pr)
        if (radius > kRefMaxRadius && v != 4) continue;
        for (int w : kCandWgs) {
            for (int rows : kCandRows) {
                if (rc != PRK_OK) break;
                if (rows > n) continue;
                SweepPlan p;
                rc = plan_sweep(pattern, radius, dtype, n, pitch, buf[0], buf[1], buf[2], rows, &p);
                if (rc) break;
                p.variant = v;
                p.wgs_per_cu = w;
                p.order = 1;
                p.rows_per_wave = rows;
                rc = launch_sweep(p, buf[cur], buf[cur ^ 1], buf[2], s);  // warm-up
                cur ^= 1;
                if (rc == PRK_OK) rc = check_hip(hipEventRecord(e0, s), "event");
                for (int r = 0; r < reps && rc == PRK_OK; ++r) {
                    rc = launch_sweep(p, buf[cur], buf[cur ^ 1], buf[2], s);
                    cur ^= 1;
                }
                if (rc == PRK_OK) rc = check_hip(hipEventRecord(e1, s), "event");
                if (rc == PRK_OK) rc = check_hip(hipEventSynchronize(e1), "event sync");
                float ms = 0.0f;
                if (rc == PRK_OK) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "event time");
                if (rc == PRK_OK && ms > 0.0f) {
                    const double gbs = bytes_iter * reps / (ms * 1e-3) / 1e9;
                    if (gbs > top.gbs) {
                        top.variant = v;
                        top.wgs_per_cu = w;
                        top.rows_per_wave = rows;
                        top.order = 1;
                        top.gbs = gbs;
                    }
                }
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    for (void *b : buf)
        if (b) (void)hipFree(b);
    if (rc) return rc;
    if (top.gbs <= 0.0) return set_error(PRK_EINVAL, "no tiling candidate fits n=%d", n);
    rc = prk_tuning_set(pattern, radius, dtype, n, &top);
    if (rc) return rc;
    if (best) *best = top;
    return PRK_OK;
}

}  // extern "C"
