// C ABI of libprkstencil.so: the drop-in operator table, the fused sweep,
// initialisation, the device L1 norm, the driver loop and the HBM
// calibration kernels.  Declarations and reference citations: prk_stencil.h.
#include <execinfo.h>
#include <hip/hip_runtime.h>
#include <signal.h>
#include <unistd.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "prk_internal.hpp"
#include "sweep.hpp"

namespace prk {

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

int check_hip(hipError_t e, const char *what)
{
    if (e == hipSuccess) return PRK_OK;
    return set_error(PRK_EHIP, "%s: %s", what, hipGetErrorString(e));
}

// Process-wide overrides of the tuned tiling (tests and tuning probes).
static int g_rows_per_wave = 0;  // 0: tuned / default
static int g_variant_override = -1;
static int g_order = -1;
static int g_wgs_per_cu = -1;

static bool valid_dtype(int dt) { return dt == F64 || dt == F32 || dt == BF16; }

static int vec_width(int dt) { return dt == F64 ? 2 : dt == F32 ? 4 : 8; }

int plan_sweep(int pattern, int radius, int dtype, int n, long long pitch, const void *in,
               const void *in_next, const void *out, int rows_per_wave, SweepPlan *plan)
{
    if (pattern != STAR && pattern != GRID)
        return set_error(PRK_EINVAL, "pattern must be 0 (star) or 1 (grid), got %d", pattern);
    if (radius < 1 || radius > kMaxRadius)
        return set_error(PRK_ENOSTENCIL, "no stencil for radius %d (supported 1..%d)", radius,
                         kMaxRadius);
    if (!valid_dtype(dtype)) return set_error(PRK_EINVAL, "unknown dtype %d", dtype);
    if (n < 1) return set_error(PRK_EINVAL, "grid dimension must be positive, got %d", n);
    if (2 * radius + 1 > n)
        return set_error(PRK_EINVAL, "stencil radius %d too large for n=%d", radius, n);
    if (pitch < n) return set_error(PRK_EINVAL, "pitch %lld < n %d", pitch, n);
    if (!in || !out) return set_error(PRK_EINVAL, "null grid pointer");
    if (in_next && (in_next == in || in_next == out))
        return set_error(PRK_EINVAL, "in_next must be a separate buffer");
    const size_t es = dtype_size(dtype);
    if ((unsigned long long)pitch * es >= (1ull << 31))
        return set_error(PRK_EINVAL, "row of %lld elements exceeds 2 GiB", pitch);
    const bool aligned = ((uintptr_t)in % 16 == 0) && ((uintptr_t)out % 16 == 0) &&
                         (!in_next || (uintptr_t)in_next % 16 == 0) && ((pitch * es) % 16 == 0);
    const int V = vec_width(dtype);
    plan->dtype = dtype;
    plan->pattern = pattern;
    plan->radius = radius;
    plan->n = n;
    plan->span = whole_grid(n);
    plan->pitch = pitch;
    plan->pad_writable = false;
    plan->mode = !aligned ? MODE_ELEM : (n % V == 0 ? MODE_VEC : MODE_VEC_LD);
    const prk_tuning t = choose_tuning(pattern, radius, dtype, n);
    plan->variant = g_variant_override >= 0 ? g_variant_override : t.variant;
    plan->order = g_order >= 0 ? g_order : t.order;
    plan->wgs_per_cu = g_wgs_per_cu >= 0 ? g_wgs_per_cu : t.wgs_per_cu;
    int rows = rows_per_wave > 0 ? rows_per_wave : (g_rows_per_wave > 0 ? g_rows_per_wave : t.rows_per_wave);
    plan->rows_per_wave = std::max(1, std::min(rows, n));
    return PRK_OK;
}

int launch_sweep(const SweepPlan &p, const void *in, void *in_next, void *out, hipStream_t s)
{
    switch (p.dtype) {
    case F64: return launch_sweep_f64(p, in, in_next, out, s);
    case F32: return launch_sweep_f32(p, in, in_next, out, s);
    case BF16: return launch_sweep_bf16(p, in, in_next, out, s);
    }
    return set_error(PRK_EINVAL, "unknown dtype %d", p.dtype);
}

// ---------------------------------------------------------------------------
// in += 1 over a pitched grid (elementwise; 16-byte vectors when aligned).
// ---------------------------------------------------------------------------
template <int DT, bool VEC>
__global__ __launch_bounds__(256) void add_one_kernel(void *buf, int n, long long pitch)
{
    typedef typename DTypeTraits<DT>::T T;
    constexpr int V = DTypeTraits<DT>::V;
    for (int row = blockIdx.y; row < n; row += gridDim.y) {
        T *p = (T *)buf + (long long)row * pitch;
        for (int c = (blockIdx.x * 256 + threadIdx.x) * V; c < n; c += gridDim.x * 256 * V) {
            if (VEC && c + V <= n) {
                union { u32x4 v; T e[V]; } x, y;
                x.v = *(const u32x4 *)(p + c);
                plus_one<DT, V>(x.e, y.e);
                *(u32x4 *)(p + c) = y.v;
            } else {
                for (int v = 0; v < V && c + v < n; ++v) {
                    T y;
                    plus_one<DT, 1>(&p[c + v], &y);
                    p[c + v] = y;
                }
            }
        }
    }
}

static int launch_add(int dtype, int n, long long pitch, void *in, hipStream_t s)
{
    if (!valid_dtype(dtype)) return set_error(PRK_EINVAL, "unknown dtype %d", dtype);
    if (n < 1 || pitch < n || !in) return set_error(PRK_EINVAL, "bad grid (n=%d pitch=%lld)", n, pitch);
    const bool vec = ((uintptr_t)in % 16 == 0) && ((pitch * dtype_size(dtype)) % 16 == 0);
    const int V = vec_width(dtype);
    const unsigned bx = std::min<unsigned>(8, (unsigned)((n + 256 * V - 1) / (256 * V)));
    dim3 grid(bx, std::min(n, 65535));
    switch (dtype) {
    case F64:
        if (vec) hipLaunchKernelGGL((add_one_kernel<F64, true>), grid, dim3(256), 0, s, in, n, pitch);
        else hipLaunchKernelGGL((add_one_kernel<F64, false>), grid, dim3(256), 0, s, in, n, pitch);
        break;
    case F32:
        if (vec) hipLaunchKernelGGL((add_one_kernel<F32, true>), grid, dim3(256), 0, s, in, n, pitch);
        else hipLaunchKernelGGL((add_one_kernel<F32, false>), grid, dim3(256), 0, s, in, n, pitch);
        break;
    default:
        if (vec) hipLaunchKernelGGL((add_one_kernel<BF16, true>), grid, dim3(256), 0, s, in, n, pitch);
        else hipLaunchKernelGGL((add_one_kernel<BF16, false>), grid, dim3(256), 0, s, in, n, pitch);
        break;
    }
    return check_hip(hipGetLastError(), "add launch");
}

// ---------------------------------------------------------------------------
// Initial values
// ---------------------------------------------------------------------------
template <int DT>
__global__ __launch_bounds__(256) void init_kernel(int kind, uint64_t seed, int n, long long pitch,
                                                   void *in, void *out)
{
    typedef typename DTypeTraits<DT>::T T;
    for (int row = blockIdx.y; row < n; row += gridDim.y) {
        T *pi = (T *)in + (long long)row * pitch;
        T *po = out ? (T *)out + (long long)row * pitch : nullptr;
        for (int c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256) {
            double v = kind == PRK_INIT_ANALYTIC ? (double)(row + c)
                                                 : seeded_value(seed, (uint64_t)row * (uint64_t)n + c);
            if constexpr (DT == F64) {
                pi[c] = v;
                if (po) po[c] = 0.0;
            } else if constexpr (DT == F32) {
                pi[c] = (float)v;
                if (po) po[c] = 0.0f;
            } else {
                pi[c] = f32_to_bf16((float)v);
                if (po) po[c] = 0;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// L1 norm: compensated per-thread sums (Neumaier), double-double combination
// across the block and, in a second single-block pass, across blocks in a
// fixed order, so the result does not depend on scheduling.
// ---------------------------------------------------------------------------
struct DD {
    double hi, lo;
};

__device__ inline DD dd_add(DD a, DD b)
{
#pragma clang fp contract(off)
    const double s = a.hi + b.hi;
    const double bb = s - a.hi;
    const double e = (a.hi - (s - bb)) + (b.hi - bb);
    const double lo = e + a.lo + b.lo;
    const double hi = s + lo;
    return DD{hi, lo - (hi - s)};
}

constexpr int kNormThreads = 256;
static_assert(sizeof(DD) * kNormBlocks == kNormPartialsBytes, "norm partials layout");

template <int DT>
__global__ __launch_bounds__(kNormThreads) void norm_partial_kernel(const void *out, int n,
                                                                    long long pitch, int R, DD *partials)
{
#pragma clang fp contract(off)
    typedef typename DTypeTraits<DT>::T T;
    const T *o = (const T *)out;
    double sum = 0.0, comp = 0.0;
    for (int i = R + blockIdx.x; i < n - R; i += gridDim.x) {
        const T *row = o + (long long)i * pitch;
        for (int j = R + threadIdx.x; j < n - R; j += kNormThreads) {
            double x;
            if constexpr (DT == F64) x = row[j];
            else if constexpr (DT == F32) x = (double)row[j];
            else x = (double)bf16_to_f32(row[j]);
            x = fabs(x);
            const double t = sum + x;
            if (fabs(sum) >= x) comp += (sum - t) + x;
            else comp += (x - t) + sum;
            sum = t;
        }
    }
    __shared__ DD red[kNormThreads];
    red[threadIdx.x] = DD{sum, comp};
    __syncthreads();
    for (int s = kNormThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = dd_add(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kNormThreads) void norm_final_kernel(const DD *partials, int nparts,
                                                                  double active, double *result)
{
#pragma clang fp contract(off)
    __shared__ DD red[kNormThreads];
    DD acc{0.0, 0.0};
    for (int k = threadIdx.x; k < nparts; k += kNormThreads) acc = dd_add(acc, partials[k]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kNormThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = dd_add(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) result[0] = (red[0].hi + red[0].lo) / active;
}

// Per-device workspace for the norm (partials + result), allocated once.
struct NormWorkspace {
    DD *partials = nullptr;
    double *result = nullptr;
};
static std::mutex g_ws_mutex;
static std::vector<NormWorkspace> g_ws;

static int norm_workspace(NormWorkspace **ws)
{
    int dev = 0;
    int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    if ((int)g_ws.size() <= dev) g_ws.resize(dev + 1);
    NormWorkspace &w = g_ws[dev];
    if (!w.partials) {
        void *p = nullptr;
        rc = check_hip(hipMalloc(&p, sizeof(DD) * kNormBlocks + 64), "norm workspace");
        if (rc) return rc;
        w.partials = (DD *)p;
        w.result = (double *)((char *)p + sizeof(DD) * kNormBlocks);
    }
    *ws = &w;
    return PRK_OK;
}

// Blocks of the norm's first pass (one partial each): one per interior row,
// at most kNormBlocks.
static int norm_blocks(int n, int radius)
{
    return (int)std::max<long long>(1, std::min<long long>(kNormBlocks, (long long)n - 2 * radius));
}

int enqueue_l1norm(int dtype, int n, long long pitch, int radius, const void *out, double *d_result,
                   DD *partials, hipStream_t s)
{
    const long long w = n - 2 * radius;
    const int blocks = norm_blocks(n, radius);
    switch (dtype) {
    case F64:
        hipLaunchKernelGGL(norm_partial_kernel<F64>, dim3(blocks), dim3(kNormThreads), 0, s, out, n, pitch, radius, partials);
        break;
    case F32:
        hipLaunchKernelGGL(norm_partial_kernel<F32>, dim3(blocks), dim3(kNormThreads), 0, s, out, n, pitch, radius, partials);
        break;
    default:
        hipLaunchKernelGGL(norm_partial_kernel<BF16>, dim3(blocks), dim3(kNormThreads), 0, s, out, n, pitch, radius, partials);
        break;
    }
    hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(kNormThreads), 0, s, (const DD *)partials, blocks,
                       (double)(w * w), d_result);
    return check_hip(hipGetLastError(), "norm launch");
}

// ---------------------------------------------------------------------------
// HBM calibration
// ---------------------------------------------------------------------------
// One 16-byte vector per thread and one short-lived workgroup per 4 KiB, with
// non-temporal hints: the fastest streaming shape measured on MI355X.
__global__ __launch_bounds__(256) void triad_kernel(long long len, double scalar, double *A,
                                                    const double *B, const double *C)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    const long long nv = len / 2;
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k < nv) {
        d2 a = __builtin_nontemporal_load(&((d2 *)A)[k]);
        const d2 b = __builtin_nontemporal_load(&((const d2 *)B)[k]);
        const d2 c = __builtin_nontemporal_load(&((const d2 *)C)[k]);
        a += b + scalar * c;
        __builtin_nontemporal_store(a, &((d2 *)A)[k]);
    }
    if (k == 0 && (len & 1)) A[len - 1] += B[len - 1] + scalar * C[len - 1];
}

// The sweep's two read and two write streams, with no stencil between them
// (in through the cache as the sweep reads it, out and the stores
// non-temporal: the fastest of the hint combinations measured on MI355X).
__global__ __launch_bounds__(256) void stream_mix_kernel(long long len, const double *in, double *in_next,
                                                         double *out)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k < len / 2) {
        const d2 x = ((const d2 *)in)[k];
        d2 o = __builtin_nontemporal_load(&((d2 *)out)[k]);
        o += x;
        __builtin_nontemporal_store(x + 1.0, &((d2 *)in_next)[k]);
        __builtin_nontemporal_store(o, &((d2 *)out)[k]);
    }
    if (k == 0 && (len & 1)) {
        in_next[len - 1] = in[len - 1] + 1.0;
        out[len - 1] += in[len - 1];
    }
}

__global__ __launch_bounds__(256) void copy_kernel(long long nvec, u32x4 *dst, const u32x4 *src)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k < nvec) __builtin_nontemporal_store(__builtin_nontemporal_load(&src[k]), &dst[k]);
}

// Copy in a chosen shape: a workgroup's chunk is VPL rows of 256 vectors
// (lane t takes vector t of each row), all VPL loads issued before the
// stores; chunks are taken in turn by `gridDim.x` workgroups.
template <int VPL, bool NT>
__global__ __launch_bounds__(256) void copy_shape_kernel(long long nvec, u32x4 *dst, const u32x4 *src)
{
    const long long chunk = 256LL * VPL;
    for (long long base = (long long)blockIdx.x * chunk; base < nvec; base += (long long)gridDim.x * chunk) {
        u32x4 v[VPL];
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
            const long long k = base + j * 256 + threadIdx.x;
            if (k < nvec) v[j] = NT ? __builtin_nontemporal_load(&src[k]) : src[k];
        }
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
            const long long k = base + j * 256 + threadIdx.x;
            if (k < nvec) {
                if (NT) __builtin_nontemporal_store(v[j], &dst[k]);
                else dst[k] = v[j];
            }
        }
    }
}

template <int VPL>
static void launch_copy_shape(int nt, unsigned blocks, long long nvec, u32x4 *dst, const u32x4 *src, hipStream_t s)
{
    if (nt) hipLaunchKernelGGL((copy_shape_kernel<VPL, true>), dim3(blocks), dim3(256), 0, s, nvec, dst, src);
    else hipLaunchKernelGGL((copy_shape_kernel<VPL, false>), dim3(blocks), dim3(256), 0, s, nvec, dst, src);
}

// ---------------------------------------------------------------------------
// The norm's partials in a static device array, for the allocation-free
// entry point.  In host code `scratch` names a host-side stand-in, not device
// memory: kernels get the device address looked up for the current device
// (each device holds its own copy of the array).
// ---------------------------------------------------------------------------
constexpr int kScratchPartials = kNormBlocks;  // 16 KiB per device
static __device__ __align__(256) DD scratch[kScratchPartials];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device

static int scratch_address(void **ptr)
{
    int dev = 0;
    int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= kMaxDevices) return set_error(PRK_EINVAL, "device %d beyond the %d supported", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "norm scratch address");
        if (rc) return rc;
        if (!p) return set_error(PRK_EHIP, "norm scratch address is null");
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return PRK_OK;
}

int l1norm_async(int dtype, int n, long long pitch, int radius, const void *out, double *d_result, void *partials,
                 hipStream_t s)
{
    if (!valid_dtype(dtype)) return set_error(PRK_EINVAL, "unknown dtype %d", dtype);
    if (n < 1 || pitch < n || !out || !d_result || !partials || radius < 1 || 2 * radius + 1 > n)
        return set_error(PRK_EINVAL, "bad norm arguments (n=%d radius=%d)", n, radius);
    return enqueue_l1norm(dtype, n, pitch, radius, out, d_result, (DD *)partials, s);
}

// ---------------------------------------------------------------------------
// Crash backtraces.  With the handlers installed, a fatal signal in any
// thread (a host fault in the server's worker thread, say) first writes the
// faulting thread's native frames to stderr, then the handler that was
// there before (Python's faulthandler under pytest, or the default action)
// runs.  Only async-signal-safe calls are made in the handler; backtrace()
// is called once at install so that its first use, which may load libgcc,
// is not in a handler.
// ---------------------------------------------------------------------------
constexpr int kCrashSignals[] = {SIGSEGV, SIGBUS, SIGILL, SIGFPE, SIGABRT};
constexpr int kNumCrashSignals = (int)(sizeof(kCrashSignals) / sizeof(kCrashSignals[0]));
struct sigaction g_prev_action[kNumCrashSignals];
bool g_crash_installed = false;
std::mutex g_crash_mu;

void crash_write(const char *s) { (void)!write(2, s, strlen(s)); }

void crash_handler(int sig)
{
    char digits[12], num[12];
    int nd = 0, k = 0;
    for (int v = sig; v > 0 || nd == 0; v /= 10) digits[nd++] = (char)('0' + v % 10);
    while (nd) num[k++] = digits[--nd];
    num[k] = 0;
    crash_write("prk: fatal signal ");
    crash_write(num);
    crash_write(", native frames of the faulting thread:\n");
    void *frames[64];
    backtrace_symbols_fd(frames, backtrace(frames, 64), 2);
    for (int i = 0; i < kNumCrashSignals; ++i)
        if (kCrashSignals[i] == sig) sigaction(sig, &g_prev_action[i], nullptr);
    raise(sig);  // delivered to the previous handler once this one returns
}

}  // namespace prk

using namespace prk;

// ---------------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------------
extern "C" {

int prk_crash_backtrace(int enable)
{
    std::lock_guard<std::mutex> lock(g_crash_mu);
    if (enable && !g_crash_installed) {
        void *f[1];
        (void)backtrace(f, 1);
        struct sigaction sa;
        std::memset(&sa, 0, sizeof sa);
        sa.sa_handler = crash_handler;
        sigemptyset(&sa.sa_mask);
        for (int i = 0; i < kNumCrashSignals; ++i)
            if (sigaction(kCrashSignals[i], &sa, &g_prev_action[i]) != 0) {
                for (int j = 0; j < i; ++j) sigaction(kCrashSignals[j], &g_prev_action[j], nullptr);
                return set_error(PRK_EINVAL, "sigaction failed for signal %d", kCrashSignals[i]);
            }
        g_crash_installed = true;
    } else if (!enable && g_crash_installed) {
        for (int i = 0; i < kNumCrashSignals; ++i) sigaction(kCrashSignals[i], &g_prev_action[i], nullptr);
        g_crash_installed = false;
    }
    return PRK_OK;
}

const char *prk_last_error(void) { return g_last_error.c_str(); }
int prk_abi_version(void) { return PRK_STENCIL_ABI_VERSION; }

int prk_set_rows_per_wave(int rows)
{
    if (rows < 0) return set_error(PRK_EINVAL, "rows per wave must be >= 0");
    g_rows_per_wave = rows;
    return PRK_OK;
}

int prk_set_sweep_variant(int variant)
{
    if (variant < -1 || variant >= kSweepVariants)
        return set_error(PRK_EINVAL, "sweep variant must be -1..%d", kSweepVariants - 1);
    g_variant_override = variant;
    return PRK_OK;
}

int prk_set_sweep_order(int order, int wgs_per_cu)
{
    if (order < -1 || order > 7 || wgs_per_cu < -1 || wgs_per_cu > 16)
        return set_error(PRK_EINVAL, "order must be -1..7 and wgs_per_cu -1..16 (-1: tuned)");
    g_order = order;
    g_wgs_per_cu = wgs_per_cu;
    return PRK_OK;
}

int prk_stencil(int pattern, int radius, int n, int t, const double *in, double *out)
{
    if (t < 1) return set_error(PRK_EINVAL, "tile size must be >= 1, got %d", t);
    SweepPlan p;
    int rc = plan_sweep(pattern, radius, F64, n, n, in, nullptr, out, 0, &p);
    if (rc) return rc;
    return launch_sweep(p, in, nullptr, out, 0);
}

#define PRK_DROPIN(name, pat, r)                                            \
    int name(int n, int t, const double *in, double *out)                   \
    {                                                                       \
        return prk_stencil(pat, r, n, t, in, out);                          \
    }
PRK_DROPIN(prk_star1, STAR, 1)
PRK_DROPIN(prk_star2, STAR, 2)
PRK_DROPIN(prk_star3, STAR, 3)
PRK_DROPIN(prk_star4, STAR, 4)
PRK_DROPIN(prk_star5, STAR, 5)
PRK_DROPIN(prk_grid1, GRID, 1)
PRK_DROPIN(prk_grid2, GRID, 2)
PRK_DROPIN(prk_grid3, GRID, 3)
PRK_DROPIN(prk_grid4, GRID, 4)
PRK_DROPIN(prk_grid5, GRID, 5)
#undef PRK_DROPIN

int prk_add(int n, int t, double *in)
{
    if (t < 1) return set_error(PRK_EINVAL, "tile size must be >= 1, got %d", t);
    return launch_add(F64, n, n, in, 0);
}

int prk_add_ex(int dtype, int n, long long pitch, void *in, void *stream)
{
    return launch_add(dtype, n, pitch, in, (hipStream_t)stream);
}

int prk_sweep(int pattern, int radius, int dtype, int n, long long pitch, const void *in, void *in_next,
              void *out, void *stream)
{
    if (!in_next) return set_error(PRK_EINVAL, "fused sweep needs in_next");
    SweepPlan p;
    int rc = plan_sweep(pattern, radius, dtype, n, pitch, in, in_next, out, 0, &p);
    if (rc) return rc;
    return launch_sweep(p, in, in_next, out, (hipStream_t)stream);
}

int prk_sweep_slab(int pattern, int radius, int dtype, int n, int row0, int rows, long long pitch,
                   const void *in, void *in_next, void *out, void *stream)
{
    if (!in_next) return set_error(PRK_EINVAL, "fused sweep needs in_next");
    SweepPlan p;
    int rc = plan_sweep(pattern, radius, dtype, n, pitch, in, in_next, out, 0, &p);
    if (rc) return rc;
    if (row0 < 0 || rows < 1 || rows > n - row0)
        return set_error(PRK_EINVAL, "slab of rows [%d, %lld) is not inside the %d rows of the grid", row0,
                         (long long)row0 + rows, n);
    // Local row 0 is global row row0 - top.  The first R local rows are halo
    // rows or frame rows (top < R only when row0 < R), and so are the last R,
    // so the outputs, global rows [max(row0, R), min(row0+rows, n-R)), are
    // local rows [R, nrows-R) as for a whole grid (empty when nrows < 2R+1).
    const int top = std::min(radius, row0), bot = std::min(radius, n - row0 - rows);
    p.span.nrows = top + rows + bot;
    p.span.upd_lo = top;
    p.span.upd_hi = top + rows;
    return launch_sweep(p, in, in_next, out, (hipStream_t)stream);
}

int prk_stencil_ex(int pattern, int radius, int dtype, int n, long long pitch, const void *in, void *out,
                   void *stream)
{
    SweepPlan p;
    int rc = plan_sweep(pattern, radius, dtype, n, pitch, in, nullptr, out, 0, &p);
    if (rc) return rc;
    return launch_sweep(p, in, nullptr, out, (hipStream_t)stream);
}

int prk_init(int init, uint64_t seed, int dtype, int n, long long pitch, void *in, void *out, void *stream)
{
    if (init != PRK_INIT_ANALYTIC && init != PRK_INIT_SEEDED)
        return set_error(PRK_EINVAL, "unknown init kind %d", init);
    if (!valid_dtype(dtype)) return set_error(PRK_EINVAL, "unknown dtype %d", dtype);
    if (n < 1 || pitch < n || !in) return set_error(PRK_EINVAL, "bad grid (n=%d pitch=%lld)", n, pitch);
    dim3 grid(std::min(8, (n + 255) / 256), std::min(n, 65535));
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
    case F64: hipLaunchKernelGGL(init_kernel<F64>, grid, dim3(256), 0, s, init, seed, n, pitch, in, out); break;
    case F32: hipLaunchKernelGGL(init_kernel<F32>, grid, dim3(256), 0, s, init, seed, n, pitch, in, out); break;
    default: hipLaunchKernelGGL(init_kernel<BF16>, grid, dim3(256), 0, s, init, seed, n, pitch, in, out); break;
    }
    return check_hip(hipGetLastError(), "init launch");
}

int prk_l1norm(int dtype, int n, long long pitch, int radius, const void *out, double *norm, void *stream)
{
    if (!valid_dtype(dtype)) return set_error(PRK_EINVAL, "unknown dtype %d", dtype);
    if (n < 1 || pitch < n || !out || !norm || radius < 1 || 2 * radius + 1 > n)
        return set_error(PRK_EINVAL, "bad norm arguments (n=%d radius=%d)", n, radius);
    NormWorkspace *ws = nullptr;
    int rc = norm_workspace(&ws);
    if (rc) return rc;
    // the workspace is shared per device: one blocking norm at a time
    static std::mutex norm_mu;
    std::lock_guard<std::mutex> lock(norm_mu);
    hipStream_t s = (hipStream_t)stream;
    rc = enqueue_l1norm(dtype, n, pitch, radius, out, ws->result, ws->partials, s);
    if (rc) return rc;
    rc = check_hip(hipMemcpyAsync(norm, ws->result, sizeof(double), hipMemcpyDeviceToHost, s), "norm copy");
    if (rc) return rc;
    return check_hip(hipStreamSynchronize(s), "norm sync");
}

int prk_l1norm_noalloc(int dtype, int n, long long pitch, int radius, const void *out, double *d_norm, void *stream)
{
    if (!valid_dtype(dtype)) return set_error(PRK_EINVAL, "unknown dtype %d", dtype);
    if (n < 1 || pitch < n || !out || !d_norm || radius < 1 || 2 * radius + 1 > n)
        return set_error(PRK_EINVAL, "bad norm arguments (n=%d radius=%d)", n, radius);
    const int blocks = norm_blocks(n, radius);
    if (blocks > kScratchPartials)
        return set_error(PRK_EINVAL, "norm of n=%d needs %d partials; the static scratch holds %d", n, blocks,
                         kScratchPartials);
    void *ptr = nullptr;
    int rc = scratch_address(&ptr);
    if (rc) return rc;
    return enqueue_l1norm(dtype, n, pitch, radius, out, d_norm, (DD *)ptr, (hipStream_t)stream);
}

int prk_run(int pattern, int radius, int dtype, int n, long long pitch, int iterations, void *buf0, void *buf1,
            void *out, void *stream, int *final_in_buf, double *avg_seconds)
{
    if (iterations < 1) return set_error(PRK_EINVAL, "iterations must be >= 1, got %d", iterations);
    if (!buf1 || buf0 == buf1) return set_error(PRK_EINVAL, "run needs two distinct in buffers");
    SweepPlan p;
    int rc = plan_sweep(pattern, radius, dtype, n, pitch, buf0, buf1, out, 0, &p);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if ((rc = check_hip(hipEventCreate(&e0), "event create"))) return rc;
    if ((rc = check_hip(hipEventCreate(&e1), "event create"))) {
        (void)hipEventDestroy(e0);
        return rc;
    }
    void *bufs[2] = {buf0, buf1};
    int cur = 0;
    for (int it = 0; it <= iterations && rc == PRK_OK; ++it) {
        if (it == 1) rc = check_hip(hipEventRecord(e0, s), "event record");
        if (rc == PRK_OK) rc = launch_sweep(p, bufs[cur], bufs[cur ^ 1], out, s);
        cur ^= 1;
    }
    if (rc == PRK_OK) rc = check_hip(hipEventRecord(e1, s), "event record");
    if (rc == PRK_OK) rc = check_hip(hipEventSynchronize(e1), "event sync");
    float ms = 0.0f;
    if (rc == PRK_OK) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "event elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (final_in_buf) *final_in_buf = cur;
    if (avg_seconds) *avg_seconds = 1e-3 * (double)ms / iterations;
    return PRK_OK;
}

int prk_triad(long long len, double scalar, double *A, const double *B, const double *C, void *stream)
{
    if (len < 1 || !A || !B || !C) return set_error(PRK_EINVAL, "bad triad arguments");
    if (((uintptr_t)A | (uintptr_t)B | (uintptr_t)C) % 16)
        return set_error(PRK_EINVAL, "triad needs 16-byte aligned arrays");
    const long long nv = len / 2;
    if ((nv + 255) / 256 >= (1LL << 32)) return set_error(PRK_EINVAL, "triad length too large");
    const unsigned blocks = (unsigned)std::max<long long>(1, (nv + 255) / 256);
    hipLaunchKernelGGL(triad_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, len, scalar, A, B, C);
    return check_hip(hipGetLastError(), "triad launch");
}

int prk_stream_mix(long long len, const double *in, double *in_next, double *out, void *stream)
{
    if (len < 1 || !in || !in_next || !out || in == in_next || in == out || in_next == out)
        return set_error(PRK_EINVAL, "stream mix needs three separate arrays of len >= 1");
    if (((uintptr_t)in | (uintptr_t)in_next | (uintptr_t)out) % 16)
        return set_error(PRK_EINVAL, "stream mix needs 16-byte aligned arrays");
    const long long nv = len / 2;
    if ((nv + 255) / 256 >= (1LL << 32)) return set_error(PRK_EINVAL, "stream mix length too large");
    const unsigned blocks = (unsigned)std::max<long long>(1, (nv + 255) / 256);
    hipLaunchKernelGGL(stream_mix_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, len, in, in_next, out);
    return check_hip(hipGetLastError(), "stream mix launch");
}

int prk_copy_shape(long long bytes, void *dst, const void *src, int vpl, int nt, int blocks, void *stream)
{
    if (bytes < 16 || bytes % 16 || !dst || !src || ((uintptr_t)dst | (uintptr_t)src) % 16)
        return set_error(PRK_EINVAL, "copy needs a positive multiple of 16 bytes, 16-byte aligned");
    if (vpl != 1 && vpl != 2 && vpl != 4 && vpl != 8) return set_error(PRK_EINVAL, "vpl must be 1, 2, 4 or 8");
    if (nt != 0 && nt != 1) return set_error(PRK_EINVAL, "nt must be 0 or 1");
    if (blocks < 0) return set_error(PRK_EINVAL, "blocks must be >= 0");
    const long long nvec = bytes / 16;
    const long long chunks = (nvec + 256LL * vpl - 1) / (256LL * vpl);
    if (chunks >= (1LL << 31)) return set_error(PRK_EINVAL, "copy too large");
    const unsigned grid = blocks > 0 ? (unsigned)std::min<long long>(blocks, chunks) : (unsigned)chunks;
    hipStream_t s = (hipStream_t)stream;
    switch (vpl) {
    case 1: launch_copy_shape<1>(nt, grid, nvec, (u32x4 *)dst, (const u32x4 *)src, s); break;
    case 2: launch_copy_shape<2>(nt, grid, nvec, (u32x4 *)dst, (const u32x4 *)src, s); break;
    case 4: launch_copy_shape<4>(nt, grid, nvec, (u32x4 *)dst, (const u32x4 *)src, s); break;
    default: launch_copy_shape<8>(nt, grid, nvec, (u32x4 *)dst, (const u32x4 *)src, s); break;
    }
    return check_hip(hipGetLastError(), "copy launch");
}

int prk_copy(long long bytes, void *dst, const void *src, void *stream)
{
    if (bytes < 16 || bytes % 16 || !dst || !src || ((uintptr_t)dst | (uintptr_t)src) % 16)
        return set_error(PRK_EINVAL, "copy needs a positive multiple of 16 bytes, 16-byte aligned");
    const long long nv = bytes / 16;
    if ((nv + 255) / 256 >= (1LL << 32)) return set_error(PRK_EINVAL, "copy too large");
    const unsigned blocks = (unsigned)((nv + 255) / 256);
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, nv, (u32x4 *)dst,
                       (const u32x4 *)src);
    return check_hip(hipGetLastError(), "copy launch");
}

}  // extern "C"
