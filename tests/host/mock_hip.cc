// Test infrastructure: a CPU stand-in for the HIP calls the solve server
// (kernels_amd/csrc/server.hip) makes, so its host logic -- admission, the
// pool, window building, retirement, the worker thread -- can run under
// AddressSanitizer / ThreadSanitizer without a GPU.
//
// Device memory is host memory (mmap, reserve-only); streams and events are
// no-ops (everything completes at once); stream capture records each round
// launch into the graph; hipGraphLaunch replays the records: it checks every
// work entry against its grid table (grid index, class, stage range, tile
// range, that every pointer lies inside a live device allocation) and
// performs the final-norm stage by writing 2(K+1), the analytic norm.
#include <sys/mman.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

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
// Failure injection: the instantiation with this ordinal (1-based) fails.
long long g_instantiations = 0, g_fail_instantiation = 0;

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

static void replay(const Launch &l)
{
    const int R = l.cls % kBatchMaxRadius + 1, P = (l.cls / kBatchMaxRadius) % 2, DT = l.cls / (2 * kBatchMaxRadius);
    MOCK_CHECK(is_dev(l.a.work, sizeof(RWork) * l.blocks), "work table outside device memory");
    for (unsigned b = 0; b < l.blocks; ++b) {
        const RWork w = l.a.work[b];
        ++g_items;
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
            ++g_final;
        }
    }
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
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind)
{
    std::memcpy(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t)
{
    MOCK_CHECK(!g_cap, "copy inside stream capture");
    std::memcpy(d, s, n);
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
    *e = (hipEvent_t)std::malloc(16);
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e)
{
    std::free(e);
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
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
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t)
{
    for (const Launch &l : *(std::vector<Launch> *)e) replay(l);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e)
{
    delete (std::vector<Launch> *)e;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete (std::vector<Launch> *)g;
    return hipSuccess;
}
}  // extern "C"
