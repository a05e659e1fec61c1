// Test infrastructure: a CPU stand-in for the HIP calls the solve server
// (kernels_amd/csrc/server.hip) makes, so its host logic -- admission, the
// pool, window building, retirement, the worker thread -- can run under
// AddressSanitizer / ThreadSanitizer without a GPU.
//
// Device memory is host memory (mmap, reserve-only); stream capture records
// each round launch into the graph; hipGraphLaunch replays the records: it
// checks every work entry against its grid table (grid index, class, stage
// range, tile range, that every pointer lies inside a live device allocation)
// and performs the final-norm stage by writing 2(K+1), the analytic norm.
//
// Device time.  With PRK_MOCK_GBS unset everything completes at once.  With
// PRK_MOCK_GBS=<GB/s> the stand-in keeps a virtual device clock: a graph
// launch (or an async copy) finishes at max(now, device busy until) + its
// bytes / rate, an event completes when the work enqueued before it does, and
// hipEventSynchronize sleeps until then.  While work is in flight the memory
// it uses (the window's tables, the copy's source and destination, the graph
// exec) is registered, and the stand-in aborts when the host frees it
// (hipFree, hipHostFree, hipGraphExecDestroy) or reads device memory it
// writes (hipMemcpy) before the work is done -- hazards that an instant
// device never shows.
#include <sys/mman.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
#include <algorithm>

#include "batch.hpp"
#include "prk_internal.hpp"
#include "prk_stencil.h"

namespace prk {

static thread_local std::string g_err;
int set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
int check_hip(hipError_t e, const char *what)
{
    return e == hipSuccess ? 0 : set_error(PRK_EHIP, "%s: hip error %d", what, (int)e);
}
int batch_class_of(const prk_solve &s) { return (s.dtype * 2 + s.pattern) * kBatchMaxRadius + (s.radius - 1); }
int batch_validate_solve(const prk_solve &s)
{
    if (s.pattern < 0 || s.pattern > 1 || s.radius < 1 || s.radius > kBatchMaxRadius || s.dtype < 0 || s.dtype > 2 ||
        s.n < 2 * s.radius + 1 || s.iterations < 0)
        return set_error(PRK_EINVAL, "bad solve");
    return PRK_OK;
}

struct Launch {
    int cls;
    RoundArgs a;
    unsigned blocks;
};
static thread_local std::vector<Launch> *g_cap = nullptr;
static std::mutex g_mu;
static std::map<const char *, size_t> g_allocs;
long long g_items = 0, g_final = 0;
int g_check_all = 1;  // 0: check only each stage's first work entry
long long g_hazards = 0;

// virtual device clock (seconds on the steady clock)
static double mock_now()
{
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}
static double g_gbs = -1;        // simulated device rate in GB/s; 0: instant
static double g_busy_until = 0;  // all enqueued work done at this time
struct InUse {
    const char *lo;
    size_t bytes;
    double until;
    const char *what;
};
static std::vector<InUse> g_inuse;  // memory in use by enqueued work (guarded by g_mu)

static double device_gbs()
{
    if (g_gbs < 0) {
        const char *e = std::getenv("PRK_MOCK_GBS");
        g_gbs = e ? std::atof(e) : 0.0;
    }
    return g_gbs;
}

// Enqueue work of `bytes` on the (single, in-order) simulated device; returns
// its completion time.  Caller holds g_mu.
static double enqueue_locked(double bytes)
{
    const double now = mock_now();
    if (device_gbs() <= 0) return now;
    g_busy_until = std::max(now, g_busy_until) + bytes / (device_gbs() * 1e9);
    return g_busy_until;
}

static void mark_in_use_locked(const void *p, size_t bytes, double until, const char *what)
{
    if (until <= mock_now()) return;
    g_inuse.push_back(InUse{(const char *)p, bytes, until, what});
}

// Aborts if [p, p+bytes) overlaps memory that in-flight work still uses.
static void check_idle_locked(const void *p, size_t bytes, const char *op)
{
    const double now = mock_now();
    size_t k = 0;
    for (size_t i = 0; i < g_inuse.size(); ++i)
        if (g_inuse[i].until > now) g_inuse[k++] = g_inuse[i];
    g_inuse.resize(k);
    const char *lo = (const char *)p, *hi = lo + (bytes ? bytes : 1);
    for (const InUse &u : g_inuse) {
        if (lo < u.lo + (u.bytes ? u.bytes : 1) && u.lo < hi) {
            ++g_hazards;
            std::fprintf(stderr, "mock hazard: %s of %p (+%zu) while in-flight work uses it (%s, %.3f ms left)\n", op,
                         p, bytes, u.what, (u.until - now) * 1e3);
            std::abort();
        }
    }
}
// Failure injection: the instantiation with this ordinal (1-based) fails.
long long g_instantiations = 0, g_fail_instantiation = 0;
long long g_updates = 0;  // exec updates in place
long long g_error_launches = 0;  // shadow comparisons

#define MOCK_CHECK(c, ...)                                                            \
    do {                                                                              \
        if (!(c)) {                                                                   \
            std::fprintf(stderr, "mock check failed (%s:%d): ", __FILE__, __LINE__);  \
            std::fprintf(stderr, __VA_ARGS__);                                        \
            std::fprintf(stderr, "\n");                                               \
            std::abort();                                                             \
        }                                                                             \
    } while (0)

static bool is_dev(const void *p, size_t bytes)
{
    std::lock_guard<std::mutex> l(g_mu);
    auto it = g_allocs.upper_bound((const char *)p);
    if (it == g_allocs.begin()) return false;
    --it;
    return (const char *)p >= it->first && (const char *)p + bytes <= it->first + it->second;
}

static double replay(const Launch &l, std::vector<const void *> &norms)
{
    double bytes = 0;
    const int R = l.cls % kBatchMaxRadius + 1, P = (l.cls / kBatchMaxRadius) % 2, DT = l.cls / (2 * kBatchMaxRadius);
    MOCK_CHECK(is_dev(l.a.work, sizeof(RWork) * l.blocks), "work table outside device memory");
    for (unsigned b = 0; b < l.blocks; ++b) {
        const RWork w = l.a.work[b];
        ++g_items;
        if (device_gbs() > 0) {
            // bytes this entry's (up to four) tiles move: init writes two
            // arrays, a sweep reads in and out and writes in_next and out,
            // the norm reads out
            const BGrid &gg = l.a.grids[w.grid];
            const int es = DT == 0 ? 8 : DT == 1 ? 4 : 2;
            const int tiles = std::min(4, (w.stage == gg.iters + 3 ? 1 : gg.ntiles) - w.tile0);
            const int mult = w.stage == 0 ? 2 : w.stage <= gg.iters + 1 ? 4 : 1;
            bytes += (double)tiles * gg.rows * 64.0 * (16 / es) * es * mult;
        }
        if (!g_check_all && w.tile0 != 0) continue;
        MOCK_CHECK(w.grid >= 0 && is_dev(l.a.grids + w.grid, sizeof(BGrid)), "grid %d outside table", w.grid);
        const BGrid g = l.a.grids[w.grid];
        MOCK_CHECK(g.radius == R && g.pattern == P, "grid in the wrong class");
        const int last = g.iters + 3;
        MOCK_CHECK(w.stage >= 0 && w.stage <= last, "stage %d of %d", w.stage, last);
        MOCK_CHECK(w.tile0 >= 0 && w.tile0 < (w.stage == last ? 1 : g.ntiles), "tile %d", w.tile0);
        const size_t es = DT == 0 ? 8 : DT == 1 ? 4 : 2;
        const size_t arr = es * (size_t)g.n * (size_t)g.pitch;
        MOCK_CHECK(is_dev(g.buf0, arr) && is_dev(g.buf1, arr) && is_dev(g.out, arr), "arrays outside the pool");
        MOCK_CHECK(is_dev(g.norm, 8) && is_dev(g.partials, sizeof(DDv) * g.ntiles), "norm outside the pool");
        if (w.stage == 0 && w.tile0 == 0) *g.norm = -1.0;
        if (w.stage == last) {
            *g.norm = 2.0 * (g.iters + 1);
            norms.push_back(g.norm);
            ++g_final;
        }
    }
    return bytes;
}

// The error of a shadowed request against its FP64 twin: the stand-in checks
// that both grids and the buffers lie in live device memory and reports no
// error (its replay does not compute the grids).
int error_vs_f64_launch(int dtype, const void *x, long long xp, const double *y, long long yp, int n, int R,
                        double *partials, double *out, hipStream_t)
{
    const size_t es = dtype == 1 ? 4 : 2;
    MOCK_CHECK(dtype == 1 || dtype == 2, "error of a dtype %d result", dtype);
    MOCK_CHECK(R >= 1 && n > 2 * R && xp >= n && yp >= n, "error shape");
    MOCK_CHECK(is_dev(x, es * (size_t)n * (size_t)xp) && is_dev(y, 8 * (size_t)n * (size_t)yp),
               "error grids outside the pool");
    MOCK_CHECK(is_dev(partials, sizeof(double) * 4 * kErrBlocks) && is_dev(out, 4 * sizeof(double)),
               "error buffers outside device memory");
    for (int k = 0; k < 4; ++k) out[k] = k == 3 ? 1.0 : 0.0;
    ++g_error_launches;
    return 0;
}

int batch_launch_round(int cls, const RoundArgs &a, unsigned blocks, hipStream_t)
{
    MOCK_CHECK(g_cap, "round launched outside stream capture");
    MOCK_CHECK(blocks > 0, "launch of zero workgroups");
    g_cap->push_back(Launch{cls, a, blocks});
    return PRK_OK;
}
}  // namespace prk

using namespace prk;

extern "C" {
const char *prk_last_error(void) { return prk::g_err.c_str(); }
hipError_t hipMalloc(void **p, size_t n)
{
    void *m = mmap(nullptr, n ? n : 1, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (m == MAP_FAILED) return hipErrorOutOfMemory;
    std::lock_guard<std::mutex> l(g_mu);
    g_allocs[(const char *)m] = n;
    *p = m;
    return hipSuccess;
}
hipError_t hipFree(void *p)
{
    if (!p) return hipSuccess;
    std::lock_guard<std::mutex> l(g_mu);
    auto it = g_allocs.find((const char *)p);
    MOCK_CHECK(it != g_allocs.end(), "hipFree of an unknown pointer");
    check_idle_locked(p, it->second, "hipFree");
    munmap(p, it->second ? it->second : 1);
    g_allocs.erase(it);
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned)
{
    *p = std::malloc(n);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void *p)
{
    {
        std::lock_guard<std::mutex> l(g_mu);
        if (p) check_idle_locked(p, 1, "hipHostFree");
    }
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k)
{
    {
        // a synchronous copy on the null stream does not wait for the
        // server's non-blocking streams: what it reads must be finished
        std::lock_guard<std::mutex> l(g_mu);
        if (k == hipMemcpyDeviceToHost) check_idle_locked(s, n, "hipMemcpy (device to host)");
    }
    std::memcpy(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t)
{
    MOCK_CHECK(!g_cap, "copy inside stream capture");
    std::memcpy(d, s, n);
    std::lock_guard<std::mutex> l(g_mu);
    const double until = enqueue_locked((double)n);
    mark_in_use_locked(s, n, until, "async copy source");
    mark_in_use_locked(d, n, until, "async copy destination");
    return hipSuccess;
}
hipError_t hipMemcpy2D(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h, hipMemcpyKind)
{
    for (size_t r = 0; r < h; ++r) std::memcpy((char *)d + r * dp, (const char *)s + r * sp, w);
    return hipSuccess;
}
hipError_t hipGetDevice(int *d)
{
    *d = 0;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned)
{
    *s = (hipStream_t)std::malloc(16);
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s)
{
    std::free(s);
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned)
{
    *e = (hipEvent_t)std::calloc(1, 16);
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e)
{
    std::free(e);
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t)
{
    if (g_cap) return hipSuccess;  // a graph node, not a timed event
    std::lock_guard<std::mutex> l(g_mu);
    *(double *)e = device_gbs() > 0 ? std::max(mock_now(), g_busy_until) : 0.0;
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e)
{
    double until;
    {
        std::lock_guard<std::mutex> l(g_mu);
        until = *(double *)e;
    }
    const double d = until - mock_now();
    if (d > 0) std::this_thread::sleep_for(std::chrono::duration<double>(d));
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode)
{
    MOCK_CHECK(!g_cap, "nested stream capture");
    g_cap = new std::vector<Launch>();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g)
{
    MOCK_CHECK(g_cap, "end of capture without a begin");
    *g = (hipGraph_t)g_cap;
    g_cap = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t)
{
    if (++g_instantiations == g_fail_instantiation) {
        *e = nullptr;
        return hipErrorInvalidValue;
    }
    *e = (hipGraphExec_t) new std::vector<Launch>(*(std::vector<Launch> *)g);
    return hipSuccess;
}
// An exec takes a graph's arguments in place when the graph has its
// launches' classes in the same order (the same topology).  While a failure
// is armed for an instantiation every update is refused, so that each window
// instantiates and the armed ordinal is reached.
static thread_local hipError_t g_last_error = hipSuccess;
hipError_t hipGraphExecUpdate(hipGraphExec_t e, hipGraph_t g, hipGraphNode_t *bad, hipGraphExecUpdateResult *res)
{
    {
        std::lock_guard<std::mutex> l(g_mu);
        check_idle_locked(e, 1, "hipGraphExecUpdate");
    }
    auto &ex = *(std::vector<Launch> *)e;
    const auto &ng = *(const std::vector<Launch> *)g;
    bool same = g_fail_instantiation == 0 && ex.size() == ng.size();
    for (size_t i = 0; same && i < ng.size(); ++i) same = ex[i].cls == ng[i].cls;
    if (bad) *bad = nullptr;
    if (!same) {
        if (res) *res = hipGraphExecUpdateErrorTopologyChanged;
        g_last_error = hipErrorGraphExecUpdateFailure;
        return hipErrorGraphExecUpdateFailure;
    }
    ex = ng;
    ++prk::g_updates;
    if (res) *res = hipGraphExecUpdateSuccess;
    return hipSuccess;
}
hipError_t hipGetLastError(void)
{
    const hipError_t e = g_last_error;
    g_last_error = hipSuccess;
    return e;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t)
{
    double bytes = 0;
    const auto &ls = *(std::vector<Launch> *)e;
    std::vector<const void *> norms;
    for (const Launch &l : ls) bytes += replay(l, norms);
    std::lock_guard<std::mutex> lk(g_mu);
    const double until = enqueue_locked(bytes);
    mark_in_use_locked(e, 1, until, "graph exec");
    for (const void *nrm : norms) mark_in_use_locked(nrm, sizeof(double), until, "final norm written by the window");
    for (const Launch &l : ls) {
        mark_in_use_locked(l.a.work, sizeof(RWork) * l.blocks, until, "window work table");
        mark_in_use_locked(l.a.grids, sizeof(BGrid), until, "window grid table");
    }
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e)
{
    {
        std::lock_guard<std::mutex> l(g_mu);
        check_idle_locked(e, 1, "hipGraphExecDestroy");
    }
    delete (std::vector<Launch> *)e;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete (std::vector<Launch> *)g;
    return hipSuccess;
}
}  // extern "C"
