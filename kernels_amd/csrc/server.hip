// Continuous-batching solve server.
//
// Requests (solves) arrive at any time through prk_server_submit.  A worker
// thread admits them first-come first-served into one device memory pool
// bounded by the HBM budget, keeps their grids resident until they finish,
// and runs all admitted requests together in windows: a window is a few
// rounds, a round holds the next stage of every active request (init, the
// fused sweeps, norm partials, final norm -- the stages of the batched
// path), and each round launches one tile kernel per (storage type,
// pattern, radius) class over all of its requests.  A window's launches are
// captured into a hipGraph on the host while the previous window runs on
// the GPU.  Requests that arrive while a window runs join at the next one;
// requests that finish free their memory for the queue behind them.  A
// request's arithmetic is that of the single-grid sweep, so its result does
// not depend on what it was batched with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.hpp"
#include "prk_internal.hpp"

namespace prk {
namespace {

double now_s()
{
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int vec_of(int dt) { return dt == F64 ? 2 : dt == F32 ? 4 : 8; }

// First-fit sub-allocator over one device allocation, with coalescing.
class Pool {
  public:
    void init(size_t size) { free_[0] = size; }
    bool alloc(size_t bytes, size_t *off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t at = it->first + bytes;
                free_.erase(it);
                if (rest) free_[at] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }

  private:
    std::map<size_t, size_t> free_;
    size_t used_ = 0;
};

struct Req {
    uint64_t id;
    prk_solve s;
    int keep;
    double t_submit = 0, t_admit = 0, t_done = 0;
    size_t off = 0, bytes = 0;
    BGrid g{};
    int stage = 0;       // next stage to schedule
    int final_window = -1;
    double norm = 0;
    bool done = false;
    int pins = 0;  // fetches in progress (release waits for them)
    // FP64 shadow solve: `twin` of a reduced-precision request, `owner` of
    // the internal twin; the owner is answered once both have finished, with
    // its error against the twin
    Req *twin = nullptr, *owner = nullptr;
    bool internal = false, finished = false;
    int err_valid = 0;
    double err_max_abs = 0, err_mean_abs = 0, err_normwise = 0;
};

constexpr int kRoundsPerWindow = 8;

}  // namespace
}  // namespace prk

using namespace prk;

struct prk_server {
    prk_server_config cfg{};
    int device = 0;
    char *pool_base = nullptr;
    Pool pool;
    hipStream_t stream = nullptr, capture = nullptr;
    hipStream_t branch[kBatchClasses] = {};
    hipEvent_t fork = nullptr, join[kBatchClasses] = {};
    hipEvent_t window_done[2] = {nullptr, nullptr};
    // per-window device tables (double-buffered) and pinned host staging
    void *d_tab[2] = {nullptr, nullptr};
    size_t d_tab_cap[2] = {0, 0};
    void *h_tab[2] = {nullptr, nullptr};
    size_t h_tab_cap[2] = {0, 0};
    size_t norm_off[2] = {0, 0};  // finishing requests' norms in the slot's tables
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    // FP64 shadow solves: error kernels on their own stream, partials and the
    // result on the device, the result in pinned memory, a count per class
    hipStream_t aux = nullptr;
    hipEvent_t err_done = nullptr;
    double *d_err = nullptr;  // kErrBlocks * 4 partials, then 4 results
    double *h_err = nullptr;
    long long shadow_seq[kBatchClasses] = {};

    std::mutex mu;
    std::condition_variable cv_work, cv_idle;
    std::deque<Req *> pending;
    std::vector<Req *> active;  // admitted, still has stages to schedule
    std::vector<Req *> inflight[2];  // finishing in the window of that slot
    std::vector<Req *> awaiting;     // finished, waiting for their shadow partner
    std::unordered_map<uint64_t, Req *> kept;
    std::vector<prk_server_result> completed;
    uint64_t next_id = 1;
    bool stop = false;
    int error = PRK_OK;
    char error_msg[256] = {0};
    prk_server_stats st{};
    std::thread worker;
};

namespace {

int fail(prk_server *s, int rc)
{
    std::lock_guard<std::mutex> lock(s->mu);
    if (!s->error) {
        s->error = rc;
        std::snprintf(s->error_msg, sizeof s->error_msg, "%s", prk_last_error());
    }
    s->cv_idle.notify_all();
    return rc;
}

size_t request_bytes(const prk_solve &s, BGrid &g)
{
    const int V = vec_of(s.dtype);
    std::memset(&g, 0, sizeof g);
    g.pitch = lib_pitch(s.n, s.dtype);
    g.n = s.n;
    g.pattern = s.pattern;
    g.radius = s.radius;
    g.iters = s.iterations;
    g.rows = std::min(s.n, 32);
    g.strips = (s.n + 64 * V - 1) / (64 * V);
    g.ntiles = g.strips * ((s.n + g.rows - 1) / g.rows);
    g.init_kind = s.init;
    g.seed = s.seed;
    const size_t arr = align_up((size_t)s.n * (size_t)g.pitch * dtype_size(s.dtype), 256);
    return 3 * arr + align_up(sizeof(DDv) * (size_t)g.ntiles, 256) + 256;
}

void place_request(prk_server *s, Req *r)
{
    const size_t es = dtype_size(r->s.dtype);
    const size_t arr = align_up((size_t)r->g.n * (size_t)r->g.pitch * es, 256);
    char *base = s->pool_base + r->off;
    r->g.buf0 = base;
    r->g.buf1 = base + arr;
    r->g.out = base + 2 * arr;
    r->g.partials = (DDv *)(base + 3 * arr);
    r->g.norm = (double *)(base + 3 * arr + align_up(sizeof(DDv) * (size_t)r->g.ntiles, 256));
}

int ensure_capacity(prk_server *s, int slot, size_t bytes)
{
    if (s->d_tab_cap[slot] < bytes) {
        if (s->d_tab[slot]) (void)hipFree(s->d_tab[slot]);
        s->d_tab[slot] = nullptr;
        const size_t cap = std::max(bytes, s->d_tab_cap[slot] * 2);
        int rc = check_hip(hipMalloc(&s->d_tab[slot], cap), "server tables");
        if (rc) return rc;
        s->d_tab_cap[slot] = cap;
    }
    if (s->h_tab_cap[slot] < bytes) {
        if (s->h_tab[slot]) (void)hipHostFree(s->h_tab[slot]);
        s->h_tab[slot] = nullptr;
        const size_t cap = std::max(bytes, s->h_tab_cap[slot] * 2);
        int rc = check_hip(hipHostMalloc(&s->h_tab[slot], cap, hipHostMallocDefault), "server staging");
        if (rc) return rc;
        s->h_tab_cap[slot] = cap;
    }
    return PRK_OK;
}

// Answer a finished user request (lock held): its result, then its memory
// kept for fetches or returned to the pool.
void complete_locked(prk_server *s, Req *r)
{
    prk_server_result res;
    std::memset(&res, 0, sizeof res);
    res.id = r->id;
    res.norm = r->norm;
    res.t_submit = r->t_submit;
    res.t_admit = r->t_admit;
    res.t_done = r->t_done;
    res.status = PRK_OK;
    res.err_valid = r->err_valid;
    res.err_max_abs = r->err_max_abs;
    res.err_mean_abs = r->err_mean_abs;
    res.err_normwise = r->err_normwise;
    s->completed.push_back(res);
    s->st.served += 1;
    if (r->keep) {
        s->kept[r->id] = r;
    } else {
        s->pool.release(r->off, r->bytes);
        delete r;
    }
}

// Error of a shadowed request against its FP64 twin, both finished: a
// fixed-order reduction on the error stream (the window stream may already
// run the next window; these grids are no longer written).
int shadow_error(prk_server *s, Req *r, Req *t)
{
    const int wd = r->g.n - 2 * r->g.radius;
    int rc = error_vs_f64_launch(r->s.dtype, r->g.out, r->g.pitch, (const double *)t->g.out, t->g.pitch, r->g.n,
                                 r->g.radius, s->d_err, s->d_err + 4 * kErrBlocks, s->aux);
    if (!rc)
        rc = check_hip(hipMemcpyAsync(s->h_err, s->d_err + 4 * kErrBlocks, 4 * sizeof(double), hipMemcpyDeviceToHost,
                                      s->aux),
                       "server error copy");
    if (!rc) rc = check_hip(hipEventRecord(s->err_done, s->aux), "server error event");
    if (!rc) rc = check_hip(hipEventSynchronize(s->err_done), "server error wait");
    if (rc) return rc;
    r->err_valid = 1;
    r->err_max_abs = s->h_err[0];
    r->err_mean_abs = wd > 0 ? s->h_err[1] / ((double)wd * (double)wd) : 0.0;
    r->err_normwise = s->h_err[3] > 0 ? std::sqrt(s->h_err[2] / s->h_err[3]) : 0.0;
    return PRK_OK;
}

// Retire the requests whose final stage ran in the window of `slot`.
void retire(prk_server *s, int slot)
{
    std::vector<Req *> fin;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        fin.swap(s->inflight[slot]);
    }
    const double t = now_s();
    // the window's one copy brought its finishing requests' norms into the
    // slot's pinned staging, in the order of `fin`
    const double *norms = (const double *)((const char *)s->h_tab[slot] + s->norm_off[slot]);
    for (size_t i = 0; i < fin.size(); ++i) {
        Req *r = fin[i];
        r->norm = norms[i];
        r->t_done = t;
        r->done = true;
    }
    // shadowed requests and their twins pair up; a pair is compared once
    // both have finished
    std::vector<Req *> pairs;  // owners whose twin has finished too
    {
        std::lock_guard<std::mutex> lock(s->mu);
        for (Req *r : fin) {
            Req *owner = r->internal ? r->owner : r;
            if (!owner->twin) {
                complete_locked(s, r);
                continue;
            }
            r->finished = true;
            Req *other = r->internal ? owner : owner->twin;
            if (other->finished) {
                s->awaiting.erase(std::remove(s->awaiting.begin(), s->awaiting.end(), other), s->awaiting.end());
                pairs.push_back(owner);
            } else {
                s->awaiting.push_back(r);
            }
        }
    }
    for (Req *r : pairs) {
        Req *tw = r->twin;
        const int rc = shadow_error(s, r, tw);
        if (rc) fail(s, rc);
        std::lock_guard<std::mutex> lock(s->mu);
        r->t_done = std::max(r->t_done, tw->t_done);
        r->twin = nullptr;
        s->pool.release(tw->off, tw->bytes);
        delete tw;
        complete_locked(s, r);
    }
    std::lock_guard<std::mutex> lock(s->mu);
    s->st.pool_used = (long long)s->pool.used();
    s->cv_idle.notify_all();
}

// Admit pending requests first-come first-served while memory and the
// active-work target allow (called with the lock held).
void admit(prk_server *s)
{
    long long tiles = 0;
    for (Req *r : s->active) tiles += r->g.ntiles;
    const double t = now_s();
    while (!s->pending.empty()) {
        Req *r = s->pending.front();
        if (!s->active.empty() && tiles >= s->cfg.round_tiles) break;
        if ((int)s->active.size() >= s->cfg.max_active) break;
        size_t off = 0;
        if (!s->pool.alloc(r->bytes, &off)) break;  // head waits for memory
        s->pending.pop_front();
        r->off = off;
        place_request(s, r);
        r->t_admit = t;
        s->active.push_back(r);
        tiles += r->g.ntiles;
    }
    s->st.pool_used = (long long)s->pool.used();
}

// Build window `w` into table slot w%2 and capture its launches.
int build_window(prk_server *s, long long w)
{
    const double t_build = now_s();
    const int slot = (int)(w & 1);
    std::vector<Req *> act;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        act = s->active;
    }
    // grids table + one work list per (round, class)
    std::vector<BGrid> grids(act.size());
    for (size_t i = 0; i < act.size(); ++i) grids[i] = act[i]->g;
    std::vector<RWork> work;
    struct Launch {
        int cls;
        size_t first;
        int blocks;
    };
    std::vector<Launch> launches[kBatchClasses];
    std::vector<Req *> finishing;
    std::vector<size_t> finishing_grid;  // their entries in `grids`
    for (int round = 0; round < kRoundsPerWindow; ++round) {
        std::vector<RWork> per[kBatchClasses];
        bool any = false;
        for (size_t i = 0; i < act.size(); ++i) {
            Req *r = act[i];
            const int last = r->g.iters + 3;
            if (r->stage > last) continue;
            const int tiles = r->stage == last ? 1 : r->g.ntiles;
            const int c = batch_class_of(r->s);
            for (int t0 = 0; t0 < tiles; t0 += kWavesPerBlock) per[c].push_back(RWork{(int)i, r->stage, t0});
            if (r->stage == last) {
                r->final_window = (int)w;
                finishing.push_back(r);
                finishing_grid.push_back(i);
            }
            ++r->stage;
            any = true;
        }
        if (!any) break;
        for (int c = 0; c < kBatchClasses; ++c) {
            if (per[c].empty()) continue;
            launches[c].push_back(Launch{c, work.size(), (int)per[c].size()});
            work.insert(work.end(), per[c].begin(), per[c].end());
        }
    }
    {
        std::lock_guard<std::mutex> lock(s->mu);
        std::vector<Req *> keep;
        for (Req *r : s->active)
            if (r->stage <= r->g.iters + 3) keep.push_back(r);
        s->active.swap(keep);
        s->inflight[slot] = finishing;
    }
    // tables: grids | work | the finishing requests' norms, which their
    // final-norm stage writes there so that one copy brings them all back
    const size_t gbytes = align_up(sizeof(BGrid) * grids.size(), 256);
    const size_t wbytes = align_up(sizeof(RWork) * work.size(), 256);
    const size_t nbytes = sizeof(double) * finishing.size();
    int rc = ensure_capacity(s, slot, gbytes + wbytes + nbytes + 256);
    if (rc) return rc;
    double *d_norms = (double *)((char *)s->d_tab[slot] + gbytes + wbytes);
    for (size_t k = 0; k < finishing_grid.size(); ++k) grids[finishing_grid[k]].norm = d_norms + k;
    s->norm_off[slot] = gbytes + wbytes;
    std::memcpy(s->h_tab[slot], grids.data(), sizeof(BGrid) * grids.size());
    std::memcpy((char *)s->h_tab[slot] + gbytes, work.data(), sizeof(RWork) * work.size());
    rc = check_hip(hipMemcpyAsync(s->d_tab[slot], s->h_tab[slot], gbytes + wbytes, hipMemcpyHostToDevice, s->stream),
                   "server table upload");
    if (rc) return rc;
    const BGrid *d_grids = (const BGrid *)s->d_tab[slot];
    const RWork *d_work = (const RWork *)((char *)s->d_tab[slot] + gbytes);

    // capture: fork one branch per class, its rounds in order, join
    rc = check_hip(hipStreamBeginCapture(s->capture, hipStreamCaptureModeThreadLocal), "server capture");
    if (rc) return rc;
    rc = check_hip(hipEventRecord(s->fork, s->capture), "server fork");
    for (int c = 0; c < kBatchClasses && !rc; ++c) {
        if (launches[c].empty()) continue;
        rc = check_hip(hipStreamWaitEvent(s->branch[c], s->fork, 0), "server fork wait");
        for (const Launch &l : launches[c]) {
            if (rc) break;
            RoundArgs a{d_grids, d_work + l.first};
            rc = batch_launch_round(c, a, (unsigned)l.blocks, s->branch[c]);
        }
        if (!rc) rc = check_hip(hipEventRecord(s->join[c], s->branch[c]), "server join");
        if (!rc) rc = check_hip(hipStreamWaitEvent(s->capture, s->join[c], 0), "server join wait");
    }
    hipGraph_t g = nullptr;
    const int rc2 = check_hip(hipStreamEndCapture(s->capture, &g), "server end capture");
    if (!rc) rc = rc2;
    // The slot's exec last ran window w-2, which has finished.  Where the new
    // graph has its topology (the same classes, the same rounds each) the
    // exec takes the new kernel arguments in place; otherwise it is replaced.
    bool updated = false;
    if (!rc && s->exec[slot]) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        updated = hipGraphExecUpdate(s->exec[slot], g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
        if (!updated) {
            (void)hipGetLastError();  // a refused update is not an error of the next call
            (void)hipGraphExecDestroy(s->exec[slot]);
            s->exec[slot] = nullptr;
        }
    }
    if (!rc && !updated) rc = check_hip(hipGraphInstantiate(&s->exec[slot], g, nullptr, nullptr, 0), "server instantiate");
    if (g) (void)hipGraphDestroy(g);
    if (!rc) rc = check_hip(hipGraphLaunch(s->exec[slot], s->stream), "server graph launch");
    if (!rc && nbytes)
        rc = check_hip(hipMemcpyAsync((char *)s->h_tab[slot] + s->norm_off[slot], d_norms, nbytes,
                                      hipMemcpyDeviceToHost, s->stream),
                       "server norm copy");
    if (!rc) rc = check_hip(hipEventRecord(s->window_done[slot], s->stream), "server window event");
    size_t nl = 0;
    for (auto &v : launches) nl += v.size();
    std::lock_guard<std::mutex> lock(s->mu);
    s->st.windows += 1;
    s->st.launches += (long long)nl;
    if (updated) s->st.graph_updates += 1;
    s->st.build_seconds += now_s() - t_build;
    return rc;
}

void worker_main(prk_server *s)
{
    if (hipSetDevice(s->device) != hipSuccess) {
        fail(s, set_error(PRK_EHIP, "server: hipSetDevice failed"));
        return;
    }
    long long w = 0;
    bool prev_inflight = false;
    for (;;) {
        {
            std::unique_lock<std::mutex> lock(s->mu);
            admit(s);
            if (s->active.empty() && !prev_inflight) {
                if (s->stop) break;
                s->cv_idle.notify_all();
                if (s->pending.empty()) {
                    s->cv_work.wait(lock, [&] { return s->stop || !s->pending.empty(); });
                } else {
                    // queued work that does not fit: kept outputs hold the
                    // memory until released; look again shortly
                    s->cv_work.wait_for(lock, std::chrono::milliseconds(5));
                }
                admit(s);
                if (s->active.empty()) {
                    if (s->stop && s->pending.empty()) break;
                    continue;
                }
            }
        }
        bool launched = false;
        {
            std::unique_lock<std::mutex> lock(s->mu);
            launched = !s->active.empty();
        }
        if (launched) {
            if (build_window(s, w) != PRK_OK) {
                fail(s, PRK_EHIP);
                break;
            }
        }
        // wait for the previous window, retire what finished in it
        if (prev_inflight) {
            const int pslot = (int)((w - 1) & 1);
            const double t_wait = now_s();
            if (hipEventSynchronize(s->window_done[pslot]) != hipSuccess) {
                fail(s, set_error(PRK_EHIP, "server: window failed"));
                break;
            }
            {
                std::lock_guard<std::mutex> lock(s->mu);
                s->st.wait_seconds += now_s() - t_wait;
            }
            retire(s, pslot);
        }
        if (launched) {
            prev_inflight = true;
            ++w;
        } else {
            prev_inflight = false;
        }
    }
    if (prev_inflight) {
        const int pslot = (int)((w - 1) & 1);
        (void)hipEventSynchronize(s->window_done[pslot]);
        retire(s, pslot);
    }
}

void server_free(prk_server *s)
{
    if (!s) return;
    for (int k = 0; k < 2; ++k) {
        if (s->exec[k]) (void)hipGraphExecDestroy(s->exec[k]);
        if (s->window_done[k]) (void)hipEventDestroy(s->window_done[k]);
        if (s->d_tab[k]) (void)hipFree(s->d_tab[k]);
        if (s->h_tab[k]) (void)hipHostFree(s->h_tab[k]);
    }
    for (int c = 0; c < kBatchClasses; ++c) {
        if (s->branch[c]) (void)hipStreamDestroy(s->branch[c]);
        if (s->join[c]) (void)hipEventDestroy(s->join[c]);
    }
    if (s->fork) (void)hipEventDestroy(s->fork);
    if (s->capture) (void)hipStreamDestroy(s->capture);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    if (s->pool_base) (void)hipFree(s->pool_base);
    // A request (or shadow twin) is in exactly one of these: queued, active,
    // finishing in a window that was never retired (the worker stopped on an
    // error), finished and waiting for its shadow partner, kept.
    for (Req *r : s->pending) delete r;
    for (Req *r : s->active) delete r;
    for (auto &v : s->inflight)
        for (Req *r : v) delete r;
    for (Req *r : s->awaiting) delete r;
    for (auto &kv : s->kept) delete kv.second;
    if (s->err_done) (void)hipEventDestroy(s->err_done);
    if (s->aux) (void)hipStreamDestroy(s->aux);
    if (s->d_err) (void)hipFree(s->d_err);
    if (s->h_err) (void)hipHostFree(s->h_err);
    delete s;
}

}  // namespace

extern "C" {

int prk_server_create(const prk_server_config *cfg, prk_server **out)
{
    if (!cfg || !out) return set_error(PRK_EINVAL, "null server config");
    if (cfg->budget_bytes < (1 << 20) || cfg->max_pending < 1 || cfg->max_active < 1 || cfg->round_tiles < 1 ||
        cfg->shadow_every < 0)
        return set_error(PRK_EINVAL, "bad server config");
    prk_server *s = new prk_server();
    s->cfg = *cfg;
    int rc = check_hip(hipGetDevice(&s->device), "hipGetDevice");
    if (!rc) rc = check_hip(hipMalloc((void **)&s->pool_base, (size_t)cfg->budget_bytes), "server pool");
    if (rc) {
        server_free(s);
        return rc == PRK_EHIP ? set_error(PRK_ENOMEM, "cannot reserve %lld bytes", cfg->budget_bytes) : rc;
    }
    s->pool.init((size_t)cfg->budget_bytes);
    s->st.budget = cfg->budget_bytes;
    rc = check_hip(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "server stream");
    if (!rc) rc = check_hip(hipStreamCreateWithFlags(&s->capture, hipStreamNonBlocking), "server capture stream");
    if (!rc) rc = check_hip(hipEventCreateWithFlags(&s->fork, hipEventDisableTiming), "server event");
    for (int c = 0; c < kBatchClasses && !rc; ++c) {
        rc = check_hip(hipStreamCreateWithFlags(&s->branch[c], hipStreamNonBlocking), "server branch");
        if (!rc) rc = check_hip(hipEventCreateWithFlags(&s->join[c], hipEventDisableTiming), "server event");
    }
    for (int k = 0; k < 2 && !rc; ++k)
        rc = check_hip(hipEventCreateWithFlags(&s->window_done[k], hipEventDisableTiming), "server event");
    if (!rc && cfg->shadow_every > 0) {
        rc = check_hip(hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking), "server error stream");
        if (!rc) rc = check_hip(hipEventCreateWithFlags(&s->err_done, hipEventDisableTiming), "server event");
        if (!rc)
            rc = check_hip(hipMalloc((void **)&s->d_err, sizeof(double) * (4 * kErrBlocks + 4)), "server error buffer");
        if (!rc) rc = check_hip(hipHostMalloc((void **)&s->h_err, 4 * sizeof(double), hipHostMallocDefault),
                                "server error staging");
    }
    if (rc) {
        server_free(s);
        return rc;
    }
    s->worker = std::thread(worker_main, s);
    *out = s;
    return PRK_OK;
}

int prk_server_submit(prk_server *s, const prk_solve *solve, int keep_output, uint64_t *id)
{
    if (!s || !solve) return set_error(PRK_EINVAL, "null server or solve");
    int rc = batch_validate_solve(*solve);
    if (rc) return rc;
    Req *r = new Req();
    r->s = *solve;
    r->keep = keep_output;
    r->bytes = request_bytes(*solve, r->g);
    r->t_submit = now_s();
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->error) {
        delete r;
        return set_error(s->error, "server failed: %s", s->error_msg);
    }
    if ((long long)r->bytes > s->cfg.budget_bytes) {
        delete r;
        s->st.rejected += 1;
        return set_error(PRK_ENOMEM, "request needs %zu bytes, budget %lld", r->bytes, s->cfg.budget_bytes);
    }
    if ((int)s->pending.size() >= s->cfg.max_pending) {
        delete r;
        s->st.rejected += 1;
        return set_error(PRK_EBUSY, "queue full (%d pending)", s->cfg.max_pending);
    }
    r->id = s->next_id++;
    s->pending.push_back(r);
    s->st.submitted += 1;
    // every shadow_every-th reduced-precision request of its class gets an
    // internal FP64 twin (queued right behind it) to measure its error
    if (s->cfg.shadow_every > 0 && solve->dtype != PRK_F64 &&
        s->shadow_seq[batch_class_of(*solve)]++ % s->cfg.shadow_every == 0) {
        Req *t = new Req();
        t->s = *solve;
        t->s.dtype = PRK_F64;
        t->bytes = request_bytes(t->s, t->g);
        if ((long long)(t->bytes + r->bytes) <= s->cfg.budget_bytes) {
            t->internal = true;
            t->owner = r;
            t->t_submit = r->t_submit;
            r->twin = t;
            s->pending.push_back(t);
            s->st.shadow_solves += 1;
        } else {
            delete t;  // the pair could never be resident at once
        }
    }
    if (id) *id = r->id;
    s->cv_work.notify_one();
    return PRK_OK;
}

int prk_server_poll(prk_server *s, prk_server_result *results, int max_results, int *count)
{
    if (!s || !count || (max_results > 0 && !results)) return set_error(PRK_EINVAL, "bad poll arguments");
    std::lock_guard<std::mutex> lock(s->mu);
    const int n = std::min<int>(max_results, (int)s->completed.size());
    for (int i = 0; i < n; ++i) results[i] = s->completed[i];
    s->completed.erase(s->completed.begin(), s->completed.begin() + n);
    *count = n;
    if (s->error) return set_error(s->error, "server failed: %s", s->error_msg);
    return PRK_OK;
}

int prk_server_drain(prk_server *s, double timeout_s)
{
    if (!s) return set_error(PRK_EINVAL, "null server");
    std::unique_lock<std::mutex> lock(s->mu);
    // every accepted request has been answered
    const bool ok = s->cv_idle.wait_for(lock, std::chrono::duration<double>(timeout_s),
                                        [&] { return s->error || s->st.served >= s->st.submitted; });
    if (s->error) return set_error(s->error, "server failed: %s", s->error_msg);
    return ok ? PRK_OK : set_error(PRK_ETIMEOUT, "server did not drain in %.1f s", timeout_s);
}

int prk_server_fetch(prk_server *s, uint64_t id, int which, void *host, size_t bytes)
{
    if (!s || !host || which < 0 || which > 1) return set_error(PRK_EINVAL, "bad fetch arguments");
    Req *r = nullptr;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        auto it = s->kept.find(id);
        if (it == s->kept.end()) return set_error(PRK_ENOTFOUND, "no kept output for request %llu", (unsigned long long)id);
        r = it->second;
        ++r->pins;  // a concurrent release of this id waits until the copy is done
    }
    const size_t es = dtype_size(r->s.dtype);
    int rc;
    if (bytes < (size_t)r->g.n * r->g.n * es) {
        rc = set_error(PRK_EINVAL, "host buffer too small");
    } else {
        const void *src = which == 0 ? r->g.out : ((r->s.iterations + 1) % 2 ? r->g.buf1 : r->g.buf0);
        rc = check_hip(hipMemcpy2D(host, (size_t)r->g.n * es, src, (size_t)r->g.pitch * es, (size_t)r->g.n * es,
                                   (size_t)r->g.n, hipMemcpyDeviceToHost),
                       "server fetch");
    }
    {
        std::lock_guard<std::mutex> lock(s->mu);
        --r->pins;
    }
    s->cv_idle.notify_all();
    return rc;
}

int prk_server_release(prk_server *s, uint64_t id)
{
    if (!s) return set_error(PRK_EINVAL, "null server");
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->kept.find(id);
    if (it == s->kept.end()) return set_error(PRK_ENOTFOUND, "no kept output for request %llu", (unsigned long long)id);
    // fetches of this id finish first (look the id up again on every wake:
    // another release may have taken it meanwhile)
    s->cv_idle.wait(lock, [&] {
        auto j = s->kept.find(id);
        return j == s->kept.end() || j->second->pins == 0;
    });
    it = s->kept.find(id);
    if (it == s->kept.end()) return set_error(PRK_ENOTFOUND, "no kept output for request %llu", (unsigned long long)id);
    s->pool.release(it->second->off, it->second->bytes);
    delete it->second;
    s->kept.erase(it);
    s->st.pool_used = (long long)s->pool.used();
    s->cv_work.notify_one();
    return PRK_OK;
}

double prk_server_clock(void) { return now_s(); }

int prk_server_stats_get(prk_server *s, prk_server_stats *out)
{
    if (!s || !out) return set_error(PRK_EINVAL, "bad stats arguments");
    std::lock_guard<std::mutex> lock(s->mu);
    *out = s->st;
    out->pending = (long long)s->pending.size();
    out->active = (long long)s->active.size();
    return PRK_OK;
}

void prk_server_destroy(prk_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
        for (Req *r : s->pending) {
            delete r;
            s->st.rejected += 1;
        }
        s->pending.clear();
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    server_free(s);
}

}  // extern "C"
