// Host side of the batched solve path: one device pool for every solve's
// grids, a sliding-window schedule cut into rounds (a round holds at most one
// stage of each solve), one launch per round, all rounds of a storage type on
// one stream and the storage types on parallel branches of one captured
// hipGraph.  Also the one-grid-per-launch baseline that runs the same solves
// through the single-grid sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "batch.hpp"
#include "prk_internal.hpp"

namespace prk {

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int vec_of(int dt) { return dt == F64 ? 2 : dt == F32 ? 4 : 8; }

// Rows per tile: each tile re-reads 2R halo rows and pays a pipeline
// prologue, so batched tiles are taller than the single-grid row blocks
// (32 measured best for the 4096-grid mix on MI355X).
int default_tile_rows(int) { return 32; }

// Tuned tile heights, keyed on (pattern, radius, dtype, size class); the
// size class of n is floor(log2 n).  Filled by prk_batch_autotune.
typedef std::tuple<int, int, int, int> BKey;
std::mutex g_brows_mu;
std::map<BKey, int> g_brows;

int size_class(int n)
{
    int k = 0;
    while ((2 << k) <= n) ++k;
    return k;
}

BKey batch_key(const prk_solve &s) { return BKey(s.pattern, s.radius, s.dtype, size_class(s.n)); }

int tile_rows_for(const prk_solve &s, int rows_per_tile)
{
    if (rows_per_tile > 0) return rows_per_tile;
    std::lock_guard<std::mutex> lock(g_brows_mu);
    auto it = g_brows.find(batch_key(s));
    return it != g_brows.end() ? it->second : default_tile_rows(s.radius);
}

const int kBatchCandRows[] = {16, 32, 64, 128};

// Tiles a round aims for: several waves per CU on all 256 CUs, so a round's
// tail (its slowest waves) is a small part of it.
long long g_round_tiles = 16384;

struct Round {
    size_t first_work;  // first workgroup entry
    int blocks;         // workgroups in the launch
    long long tiles;
};

}  // namespace

}  // namespace prk

using namespace prk;

struct prk_batch {
    std::vector<prk_solve> solves;
    std::vector<BGrid> grids;  // host copies
    void *pool = nullptr;      // grids
    size_t pool_bytes = 0;
    void *meta = nullptr;      // descriptors, items, partials, norms
    BGrid *d_grids = nullptr;
    RWork *d_work = nullptr;
    double *d_norms = nullptr;
    // classes: storage type x pattern x radius (1..8)
    static constexpr int kClasses = 3 * 2 * kBatchMaxRadius;
    std::vector<Round> rounds[kClasses];
    long long total_tiles = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture = nullptr;
    hipStream_t branch[kClasses] = {};
    hipEvent_t fork = nullptr, join[kClasses] = {};
    hipEvent_t done = nullptr;
    bool launched = false;
};

namespace {

int validate_solve(const prk_solve &s)
{
    if (s.pattern != PRK_STAR && s.pattern != PRK_GRID) return set_error(PRK_EINVAL, "bad pattern %d", s.pattern);
    if (s.radius < 1 || s.radius > kBatchMaxRadius)
        return set_error(PRK_ENOSTENCIL, "batched solves support radius 1..%d, got %d", kBatchMaxRadius, s.radius);
    if (s.dtype != PRK_F64 && s.dtype != PRK_F32 && s.dtype != PRK_BF16)
        return set_error(PRK_EINVAL, "bad dtype %d", s.dtype);
    if (s.n < 2 * s.radius + 1) return set_error(PRK_EINVAL, "n=%d too small for radius %d", s.n, s.radius);
    if (s.n > 92680) return set_error(PRK_EINVAL, "n=%d too large", s.n);
    if (s.iterations < 0) return set_error(PRK_EINVAL, "iterations must be >= 0");
    if (s.init != PRK_INIT_ANALYTIC && s.init != PRK_INIT_SEEDED) return set_error(PRK_EINVAL, "bad init");
    return PRK_OK;
}

int stage_tiles(const BGrid &g, int stage) { return stage == g.iters + 3 ? 1 : g.ntiles; }

// Sliding window: keep enough solves active that a round of their next
// stages holds about `target` tiles; each round takes one stage of every
// active solve; a solve leaves after its final-norm stage and the next one
// in submission order takes its place.
void schedule_rounds(const std::vector<int> &ids, const std::vector<BGrid> &grids, long long target,
                     std::vector<RWork> &work, std::vector<Round> &rounds)
{
    std::vector<int> active;
    std::vector<int> stage(grids.size(), 0);
    size_t next = 0;
    long long in_flight = 0;
    while (!active.empty() || next < ids.size()) {
        while (next < ids.size() && (active.empty() || in_flight < target)) {
            active.push_back(ids[next]);
            in_flight += grids[ids[next]].ntiles;
            ++next;
        }
        Round r{work.size(), 0, 0};
        std::vector<int> keep;
        for (int g : active) {
            const int tiles = stage_tiles(grids[g], stage[g]);
            for (int t0 = 0; t0 < tiles; t0 += kWavesPerBlock) work.push_back(RWork{g, stage[g], t0});
            r.blocks += (tiles + kWavesPerBlock - 1) / kWavesPerBlock;
            r.tiles += tiles;
            if (++stage[g] <= grids[g].iters + 3) keep.push_back(g);
            else in_flight -= grids[g].ntiles;
        }
        rounds.push_back(r);
        active.swap(keep);
    }
}

void batch_free(prk_batch *b)
{
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (int d = 0; d < prk_batch::kClasses; ++d) {
        if (b->branch[d]) (void)hipStreamDestroy(b->branch[d]);
        if (b->join[d]) (void)hipEventDestroy(b->join[d]);
    }
    if (b->capture) (void)hipStreamDestroy(b->capture);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->pool) (void)hipFree(b->pool);
    if (b->meta) (void)hipFree(b->meta);
    delete b;
}


template <int DT, int P>
int launch_round_pr(int radius, const RoundArgs &a, unsigned blocks, hipStream_t s)
{
    switch (radius) {
    case 1: hipLaunchKernelGGL((batch_round_kernel<DT, P, 1>), dim3(blocks), dim3(kThreads), 0, s, a); break;
    case 2: hipLaunchKernelGGL((batch_round_kernel<DT, P, 2>), dim3(blocks), dim3(kThreads), 0, s, a); break;
    case 3: hipLaunchKernelGGL((batch_round_kernel<DT, P, 3>), dim3(blocks), dim3(kThreads), 0, s, a); break;
    case 4: hipLaunchKernelGGL((batch_round_kernel<DT, P, 4>), dim3(blocks), dim3(kThreads), 0, s, a); break;
    case 5: hipLaunchKernelGGL((batch_round_kernel<DT, P, 5>), dim3(blocks), dim3(kThreads), 0, s, a); break;
    case 6: hipLaunchKernelGGL((batch_round_kernel<DT, P, 6>), dim3(blocks), dim3(kThreads), 0, s, a); break;
    case 7: hipLaunchKernelGGL((batch_round_kernel<DT, P, 7>), dim3(blocks), dim3(kThreads), 0, s, a); break;
    default: hipLaunchKernelGGL((batch_round_kernel<DT, P, 8>), dim3(blocks), dim3(kThreads), 0, s, a); break;
    }
    return check_hip(hipGetLastError(), "batch round launch");
}

}  // namespace

namespace prk {

int batch_class_of(const prk_solve &s) { return (s.dtype * 2 + s.pattern) * kBatchMaxRadius + (s.radius - 1); }

// Fixed-order block reduction of four values (max, then three sums).
__device__ inline void err_block_reduce(double (&v)[4], double *out)
{
    __shared__ double red[4][256];
#pragma unroll
    for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + w]);
#pragma unroll
            for (int k = 1; k < 4; ++k) red[k][threadIdx.x] = red[k][threadIdx.x] + red[k][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) out[threadIdx.x] = red[threadIdx.x][0];
}

template <int DT>
__global__ __launch_bounds__(256) void error_partials_kernel(const void *x_, long long xp, const double *y,
                                                             long long yp, int n, int R, double *partials)
{
    typedef typename DTypeTraits<DT>::T T;
    const T *x = (const T *)x_;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = R + (int)blockIdx.x; i < n - R; i += (int)gridDim.x)
        for (int j = R + (int)threadIdx.x; j < n - R; j += (int)blockDim.x) {
            const double yv = y[(long long)i * yp + j];
            const double d = fabs(to_f32<T>(x[(long long)i * xp + j]) - yv);
            v[0] = fmax(v[0], d);
            v[1] += d;
            v[2] += d * d;
            v[3] += yv * yv;
        }
    err_block_reduce(v, partials + 4 * blockIdx.x);
}

__global__ __launch_bounds__(256) void error_final_kernel(const double *partials, double *out)
{
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = (int)threadIdx.x; b < kErrBlocks; b += (int)blockDim.x) {
        v[0] = fmax(v[0], partials[4 * b]);
#pragma unroll
        for (int k = 1; k < 4; ++k) v[k] += partials[4 * b + k];
    }
    err_block_reduce(v, out);
}

int error_vs_f64_launch(int dtype, const void *x, long long xp, const double *y, long long yp, int n, int R,
                        double *partials, double *out, hipStream_t s)
{
    if (dtype == F32)
        hipLaunchKernelGGL(error_partials_kernel<F32>, dim3(kErrBlocks), dim3(256), 0, s, x, xp, y, yp, n, R,
                           partials);
    else if (dtype == BF16)
        hipLaunchKernelGGL(error_partials_kernel<BF16>, dim3(kErrBlocks), dim3(256), 0, s, x, xp, y, yp, n, R,
                           partials);
    else
        return set_error(PRK_EINVAL, "error against FP64 needs an FP32 or BF16 result");
    hipLaunchKernelGGL(error_final_kernel, dim3(1), dim3(256), 0, s, partials, out);
    return check_hip(hipGetLastError(), "error kernel launch");
}

int batch_launch_round(int cls, const RoundArgs &a, unsigned blocks, hipStream_t s)
{
    const int radius = cls % kBatchMaxRadius + 1;
    const int pattern = (cls / kBatchMaxRadius) % 2;
    const int dt = cls / (2 * kBatchMaxRadius);
    switch (dt * 2 + pattern) {
    case 0: return launch_round_pr<F64, STAR>(radius, a, blocks, s);
    case 1: return launch_round_pr<F64, GRID>(radius, a, blocks, s);
    case 2: return launch_round_pr<F32, STAR>(radius, a, blocks, s);
    case 3: return launch_round_pr<F32, GRID>(radius, a, blocks, s);
    case 4: return launch_round_pr<BF16, STAR>(radius, a, blocks, s);
    default: return launch_round_pr<BF16, GRID>(radius, a, blocks, s);
    }
}

int batch_validate_solve(const prk_solve &s) { return validate_solve(s); }

}  // namespace prk

using namespace prk;

extern "C" {

int prk_batch_create(const prk_solve *solves, int count, int rows_per_tile, int flags, prk_batch **out)
{
    if (!solves || count < 1 || !out) return set_error(PRK_EINVAL, "need at least one solve");
    if (rows_per_tile < 0) return set_error(PRK_EINVAL, "rows_per_tile must be >= 0");
    if (flags & PRK_BATCH_PERSISTENT)
        return set_error(PRK_EINVAL, "the persistent tile-queue form is disabled: it hung on MI355X");
    if (flags & ~PRK_BATCH_PERSISTENT) return set_error(PRK_EINVAL, "unknown batch flags 0x%x", flags);
    int rc;
    for (int i = 0; i < count; ++i)
        if ((rc = validate_solve(solves[i]))) return rc;

    prk_batch *b = new prk_batch();
    b->solves.assign(solves, solves + count);
    b->grids.resize(count);

    // grid pool layout
    size_t off = 0;
    std::vector<size_t> offs(count);
    size_t partial_total = 0;
    for (int i = 0; i < count; ++i) {
        const prk_solve &s = solves[i];
        const int V = vec_of(s.dtype);
        BGrid &g = b->grids[i];
        std::memset(&g, 0, sizeof g);
        g.pitch = lib_pitch(s.n, s.dtype);
        g.n = s.n;
        g.pattern = s.pattern;
        g.radius = s.radius;
        g.iters = s.iterations;
        g.rows = std::min(s.n, tile_rows_for(s, rows_per_tile));
        g.strips = (s.n + 64 * V - 1) / (64 * V);
        g.ntiles = g.strips * ((s.n + g.rows - 1) / g.rows);
        g.init_kind = s.init;
        g.seed = s.seed;
        offs[i] = off;
        off += 3 * align_up((size_t)s.n * (size_t)g.pitch * dtype_size(s.dtype), 256);
        partial_total += (size_t)g.ntiles;
    }
    b->pool_bytes = off;
    rc = check_hip(hipMalloc(&b->pool, off), "batch grid pool");
    if (rc) {
        batch_free(b);
        return set_error(PRK_ENOMEM, "batch needs %zu bytes of device memory", off);
    }
    for (int i = 0; i < count; ++i) {
        BGrid &g = b->grids[i];
        const size_t arr = align_up((size_t)g.n * (size_t)g.pitch * dtype_size(solves[i].dtype), 256);
        g.buf0 = (char *)b->pool + offs[i];
        g.buf1 = (char *)g.buf0 + arr;
        g.out = (char *)g.buf1 + arr;
    }

    // one schedule per class; the classes share the round budget by work
    std::vector<RWork> work;
    std::vector<int> ids[prk_batch::kClasses];
    std::vector<long long> cwork(prk_batch::kClasses, 0);
    long long all_work = 0;
    for (int i = 0; i < count; ++i) {
        const int c = batch_class_of(solves[i]);
        ids[c].push_back(i);
        const long long w = (long long)(solves[i].iterations + 1) * b->grids[i].ntiles;
        cwork[c] += w;
        all_work += w;
    }
    for (int c = 0; c < prk_batch::kClasses; ++c) {
        if (ids[c].empty()) continue;
        const long long target = std::max<long long>(1024, g_round_tiles * cwork[c] / std::max<long long>(all_work, 1));
        schedule_rounds(ids[c], b->grids, target, work, b->rounds[c]);
        for (const Round &r : b->rounds[c]) b->total_tiles += r.tiles;
    }

    // metadata block
    size_t m = 0;
    const size_t o_grids = m;
    m = align_up(m + sizeof(BGrid) * count, 256);
    const size_t o_work = m;
    m = align_up(m + sizeof(RWork) * work.size(), 256);
    const size_t o_norms = m;
    m = align_up(m + sizeof(double) * count, 256);
    const size_t o_part = m;
    m = align_up(m + sizeof(DDv) * partial_total, 256);
    rc = check_hip(hipMalloc(&b->meta, m), "batch metadata");
    if (rc) {
        batch_free(b);
        return rc;
    }
    char *mb = (char *)b->meta;
    b->d_grids = (BGrid *)(mb + o_grids);
    b->d_work = (RWork *)(mb + o_work);
    b->d_norms = (double *)(mb + o_norms);
    DDv *parts = (DDv *)(mb + o_part);
    size_t pk = 0;
    for (int i = 0; i < count; ++i) {
        b->grids[i].partials = parts + pk;
        pk += (size_t)b->grids[i].ntiles;
        b->grids[i].norm = b->d_norms + i;
    }
    rc = check_hip(hipMemcpy(b->d_grids, b->grids.data(), sizeof(BGrid) * count, hipMemcpyHostToDevice), "grids");
    if (!rc)
        rc = check_hip(hipMemcpy(b->d_work, work.data(), sizeof(RWork) * work.size(), hipMemcpyHostToDevice),
                       "work table");
    if (rc) {
        batch_free(b);
        return rc;
    }

    // capture: fork one branch per class present, its rounds in order, join
    rc = check_hip(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking), "batch stream");
    if (!rc) rc = check_hip(hipEventCreateWithFlags(&b->done, hipEventDisableTiming), "batch event");
    if (!rc) rc = check_hip(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "batch event");
    for (int c = 0; c < prk_batch::kClasses && !rc; ++c) {
        if (b->rounds[c].empty()) continue;
        rc = check_hip(hipStreamCreateWithFlags(&b->branch[c], hipStreamNonBlocking), "batch branch stream");
        if (!rc) rc = check_hip(hipEventCreateWithFlags(&b->join[c], hipEventDisableTiming), "batch event");
    }
    if (!rc) rc = check_hip(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal), "begin capture");
    if (!rc) {
        rc = check_hip(hipEventRecord(b->fork, b->capture), "fork");
        for (int c = 0; c < prk_batch::kClasses && !rc; ++c) {
            if (b->rounds[c].empty()) continue;
            rc = check_hip(hipStreamWaitEvent(b->branch[c], b->fork, 0), "fork wait");
            for (const Round &r : b->rounds[c]) {
                if (rc) break;
                RoundArgs a{b->d_grids, b->d_work + r.first_work};
                rc = batch_launch_round(c, a, (unsigned)r.blocks, b->branch[c]);
            }
            if (!rc) rc = check_hip(hipEventRecord(b->join[c], b->branch[c]), "join");
            if (!rc) rc = check_hip(hipStreamWaitEvent(b->capture, b->join[c], 0), "join wait");
        }
        hipGraph_t g = nullptr;
        const int rc2 = check_hip(hipStreamEndCapture(b->capture, &g), "end capture");
        b->graph = g;
        if (!rc) rc = rc2;
    }
    if (!rc) rc = check_hip(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
    if (rc) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return PRK_OK;
}

int prk_batch_run(prk_batch *b, void *stream)
{
    if (!b) return set_error(PRK_EINVAL, "null batch");
    hipStream_t s = (hipStream_t)stream;
    int rc = check_hip(hipGraphLaunch(b->exec, s), "batch graph launch");
    if (!rc) rc = check_hip(hipEventRecord(b->done, s), "batch event");
    if (!rc) b->launched = true;
    return rc;
}

int prk_batch_wait(prk_batch *b, double *norms)
{
    if (!b || !b->launched) return set_error(PRK_EINVAL, "batch has not been run");
    int rc = check_hip(hipEventSynchronize(b->done), "batch wait");
    if (!rc && norms)
        rc = check_hip(hipMemcpy(norms, b->d_norms, sizeof(double) * b->solves.size(), hipMemcpyDeviceToHost),
                       "batch norms");
    return rc;
}

int prk_batch_output(const prk_batch *b, int idx, int which, void *host, size_t bytes)
{
    if (!b || idx < 0 || idx >= (int)b->solves.size() || !host || which < 0 || which > 1)
        return set_error(PRK_EINVAL, "bad batch output request");
    const BGrid &g = b->grids[idx];
    const size_t es = dtype_size(b->solves[idx].dtype);
    if (bytes < (size_t)g.n * g.n * es) return set_error(PRK_EINVAL, "host buffer too small");
    const void *src = which == 0 ? g.out : ((b->solves[idx].iterations + 1) % 2 ? g.buf1 : g.buf0);
    return check_hip(hipMemcpy2D(host, (size_t)g.n * es, src, (size_t)g.pitch * es, (size_t)g.n * es, (size_t)g.n,
                                 hipMemcpyDeviceToHost),
                     "batch output copy");
}

int prk_batch_info(const prk_batch *b, long long *device_bytes, long long *tiles, int *launches)
{
    if (!b) return set_error(PRK_EINVAL, "null batch");
    if (device_bytes) *device_bytes = (long long)b->pool_bytes;
    if (tiles) *tiles = b->total_tiles;
    if (launches) {
        size_t l = 0;
        for (const auto &r : b->rounds) l += r.size();
        *launches = (int)l;
    }
    return PRK_OK;
}

int prk_batch_errors(const prk_batch *b, const prk_batch *ref, double *err)
{
    if (!b || !ref || !err) return set_error(PRK_EINVAL, "bad batch errors request");
    const size_t count = b->solves.size();
    if (ref->solves.size() != count) return set_error(PRK_EINVAL, "the batches hold different solve counts");
    for (size_t i = 0; i < count; ++i) {
        const prk_solve &x = b->solves[i], &y = ref->solves[i];
        if (y.dtype != PRK_F64 || x.pattern != y.pattern || x.radius != y.radius || x.n != y.n ||
            x.iterations != y.iterations || x.init != y.init || x.seed != y.seed)
            return set_error(PRK_EINVAL, "solve %zu of the reference batch is not the same solve in FP64", i);
    }
    // partials (kErrBlocks x 4) then 4 results per solve
    double *d = nullptr;
    const size_t doubles = (size_t)kErrBlocks * 4 + 4 * count;
    if (int rc = check_hip(hipMalloc((void **)&d, doubles * sizeof(double)), "batch errors allocation")) return rc;
    std::vector<double> h(4 * count, 0.0);
    hipStream_t s = nullptr;
    int rc = check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "batch errors stream");
    // the batches ran on their callers' streams: let that work finish first
    if (!rc) rc = check_hip(hipDeviceSynchronize(), "batch errors wait");
    for (size_t i = 0; i < count && !rc; ++i) {
        if (b->solves[i].dtype == PRK_F64) continue;
        const BGrid &gx = b->grids[i], &gy = ref->grids[i];
        rc = error_vs_f64_launch(b->solves[i].dtype, gx.out, gx.pitch, (const double *)gy.out, gy.pitch, gx.n,
                                 gx.radius, d, d + (size_t)kErrBlocks * 4 + 4 * i, s);
    }
    if (!rc)
        rc = check_hip(hipMemcpyAsync(h.data(), d + (size_t)kErrBlocks * 4, 4 * count * sizeof(double),
                                      hipMemcpyDeviceToHost, s),
                       "batch errors copy");
    if (!rc) rc = check_hip(hipStreamSynchronize(s), "batch errors wait");
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(d);
    if (rc) return rc;
    for (size_t i = 0; i < count; ++i) {
        const double *v = &h[4 * i];
        const double m = (double)(b->solves[i].n - 2 * b->solves[i].radius);
        err[3 * i] = b->solves[i].dtype == PRK_F64 ? 0.0 : v[0];
        err[3 * i + 1] = b->solves[i].dtype == PRK_F64 ? 0.0 : v[1] / (m * m);
        err[3 * i + 2] = (b->solves[i].dtype == PRK_F64 || v[3] == 0.0) ? 0.0 : std::sqrt(v[2] / v[3]);
    }
    return PRK_OK;
}

void prk_batch_destroy(prk_batch *b) { batch_free(b); }

int prk_batch_autotune(const prk_solve *solves, int count, int sample, int reps, int *keys_tuned)
{
    if (!solves || count < 1 || sample < 1 || reps < 1) return set_error(PRK_EINVAL, "bad batch autotune arguments");
    int rc;
    for (int i = 0; i < count; ++i)
        if ((rc = validate_solve(solves[i]))) return rc;
    std::map<BKey, std::vector<prk_solve>> picked;
    for (int i = 0; i < count; ++i) {
        std::vector<prk_solve> &g = picked[batch_key(solves[i])];
        if ((int)g.size() < sample) {
            prk_solve t = solves[i];
            t.iterations = std::min(t.iterations, 10);  // the ranking needs a few sweeps, not all
            g.push_back(t);
        }
    }
    // Each group's sample is repeated until a round of it fills the GPU as
    // the whole batch's rounds do (twice the round target at the default
    // height): timed alone, a handful of small grids would leave most CUs
    // idle and reward short tiles for their parallelism alone.
    std::map<BKey, std::vector<prk_solve>> groups;
    for (auto &kv : picked) {
        std::vector<prk_solve> &g = groups[kv.first];
        long long tiles = 0;
        for (size_t k = 0; tiles < 2 * g_round_tiles && g.size() < 4096; ++k) {
            const prk_solve &t = kv.second[k % kv.second.size()];
            g.push_back(t);
            const int V = vec_of(t.dtype), rows = std::min(t.n, default_tile_rows(t.radius));
            tiles += (long long)((t.n + 64 * V - 1) / (64 * V)) * ((t.n + rows - 1) / rows);
        }
    }
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rc = check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "autotune stream");
    if (!rc) rc = check_hip(hipEventCreate(&e0), "event");
    if (!rc) rc = check_hip(hipEventCreate(&e1), "event");
    int tuned = 0;
    for (auto &kv : groups) {
        if (rc) break;
        const std::vector<prk_solve> &sub = kv.second;
        const int nmin = std::min_element(sub.begin(), sub.end(), [](const prk_solve &a, const prk_solve &b) {
                             return a.n < b.n;
                         })->n;
        // The default height is timed first and kept unless a candidate is
        // at least 5 % faster: timed alone, a class runs on the whole GPU,
        // while in a mixed batch it shares the GPU with the other classes'
        // branches, and small gains measured alone do not carry over.
        const int def_rows = default_tile_rows(sub[0].radius);
        int best_rows = 0;
        float best_ms = 0.0f, def_ms = 0.0f;
        int cand[1 + sizeof kBatchCandRows / sizeof kBatchCandRows[0]];
        int ncand = 0;
        cand[ncand++] = def_rows;
        for (int rows : kBatchCandRows)
            if (rows != def_rows) cand[ncand++] = rows;
        for (int ci = 0; ci < ncand; ++ci) {
            const int rows = cand[ci];
            if (rc) break;
            if (rows > 2 * nmin && rows != def_rows) continue;
            prk_batch *b = nullptr;
            rc = prk_batch_create(sub.data(), (int)sub.size(), rows, 0, &b);
            if (rc) break;
            rc = prk_batch_run(b, st);  // warm-up
            if (!rc) rc = check_hip(hipEventRecord(e0, st), "event");
            for (int r = 0; r < reps && !rc; ++r) rc = prk_batch_run(b, st);
            if (!rc) rc = check_hip(hipEventRecord(e1, st), "event");
            if (!rc) rc = check_hip(hipEventSynchronize(e1), "event sync");
            float ms = 0.0f;
            if (!rc) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "event time");
            prk_batch_destroy(b);
            if (rc) break;
            if (ci == 0) {
                def_ms = ms;
                best_rows = rows;
                best_ms = ms;
            } else if (ms < best_ms && ms < 0.95f * def_ms) {
                best_rows = rows;
                best_ms = ms;
            }
        }
        if (!rc && best_rows > 0) {
            std::lock_guard<std::mutex> lock(g_brows_mu);
            g_brows[kv.first] = best_rows;
            ++tuned;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (keys_tuned) *keys_tuned = tuned;
    return rc;
}

int prk_batch_tuning_get(int pattern, int radius, int dtype, int n, int *rows)
{
    if (!rows) return set_error(PRK_EINVAL, "null rows");
    prk_solve s{};
    s.pattern = pattern;
    s.radius = radius;
    s.dtype = dtype;
    s.n = n;
    s.iterations = 0;
    s.init = PRK_INIT_ANALYTIC;
    int rc = validate_solve(s);
    if (rc) return rc;
    *rows = std::min(n, tile_rows_for(s, 0));
    return PRK_OK;
}

int prk_batch_tuning_set(int pattern, int radius, int dtype, int n, int rows)
{
    prk_solve s{};
    s.pattern = pattern;
    s.radius = radius;
    s.dtype = dtype;
    s.n = n;
    s.iterations = 0;
    s.init = PRK_INIT_ANALYTIC;
    int rc = validate_solve(s);
    if (rc) return rc;
    if (rows < 1) return set_error(PRK_EINVAL, "tile height must be >= 1, got %d", rows);
    std::lock_guard<std::mutex> lock(g_brows_mu);
    g_brows[batch_key(s)] = rows;
    return PRK_OK;
}

int prk_batch_tuning_clear(void)
{
    std::lock_guard<std::mutex> lock(g_brows_mu);
    g_brows.clear();
    return PRK_OK;
}

int prk_set_batch_round_tiles(long long tiles)
{
    if (tiles < 1) return set_error(PRK_EINVAL, "round size must be >= 1 tile");
    g_round_tiles = tiles;
    return PRK_OK;
}

int prk_solve_one_by_one(const prk_solve *solves, int count, int dtype_override, double *norms, void *stream)
{
    if (!solves || count < 1) return set_error(PRK_EINVAL, "need at least one solve");
    int rc;
    std::vector<prk_solve> ss(solves, solves + count);
    size_t need = 0;
    for (auto &s : ss) {
        if (dtype_override >= 0) s.dtype = dtype_override;
        if ((rc = validate_solve(s))) return rc;
        need = std::max(need, (size_t)s.n * (size_t)lib_pitch(s.n, s.dtype) * dtype_size(s.dtype));
    }
    need = align_up(need, 256);
    void *mem = nullptr;
    double *d_norms = nullptr;
    if ((rc = check_hip(hipMalloc(&mem, 3 * need), "one-by-one buffers"))) return rc;
    // the norms, then this call's own norm partials (no workspace shared
    // with other calls)
    const size_t norms_bytes = align_up(sizeof(double) * count, 256);
    if ((rc = check_hip(hipMalloc(&d_norms, norms_bytes + kNormPartialsBytes), "one-by-one norms"))) {
        (void)hipFree(mem);
        return rc;
    }
    void *partials = (char *)d_norms + norms_bytes;
    hipStream_t st = (hipStream_t)stream;
    void *b0 = mem, *b1 = (char *)mem + need, *out = (char *)mem + 2 * need;
    for (int i = 0; i < count && rc == PRK_OK; ++i) {
        const prk_solve &s = ss[i];
        const long long pitch = lib_pitch(s.n, s.dtype);
        rc = prk_init(s.init, s.seed, s.dtype, s.n, pitch, b0, out, st);
        SweepPlan p;
        if (!rc) rc = plan_sweep(s.pattern, s.radius, s.dtype, s.n, pitch, b0, b1, out, 0, &p);
        void *bufs[2] = {b0, b1};
        for (int k = 0; k <= s.iterations && rc == PRK_OK; ++k)
            rc = launch_sweep(p, bufs[k & 1], bufs[(k + 1) & 1], out, st);
        if (!rc) rc = l1norm_async(s.dtype, s.n, pitch, s.radius, out, d_norms + i, partials, st);
    }
    if (!rc) rc = check_hip(hipStreamSynchronize(st), "one-by-one sync");
    if (!rc && norms) rc = check_hip(hipMemcpy(norms, d_norms, sizeof(double) * count, hipMemcpyDeviceToHost), "norms");
    (void)hipFree(d_norms);
    (void)hipFree(mem);
    return rc;
}

}  // extern "C"
