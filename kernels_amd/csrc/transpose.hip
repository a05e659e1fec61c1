// Matrix transpose for gfx950: one iteration is B += A^T, A += 1
// (Cxx11/transpose.cc:118-131; the HIP driver's transposeNoBankConflict,
// Cxx11/transpose-hip.cc:64-87), both matrices order x order FP64, row-major.
//
// Mapping.  A workgroup (four waves) owns one tile of A, 64 rows x 128
// columns by default.  It reads the tile's rows with 16-byte buffer loads
// (one 1 KiB row segment per wave and pass), writes A + 1 back to the same
// addresses, and parks the tile in LDS (rows padded by one element so that
// reading a column spreads over the banks).  After a barrier it reads the
// tile's columns from LDS and adds them to the matching rows of B (16-byte
// loads and stores again, 512-byte row segments).  Every
// element of A and B is read once and written once: 32 bytes per element,
// the algorithmic traffic of the reference's Rate (4 * 8 * order^2 bytes).
// All eight loads of a phase are issued before the first is consumed.
// Edges: rows and columns beyond `order` are out-of-range buffer offsets
// (reads give 0, writes are dropped).  Odd orders use 64 x 64 tiles and
// 8-byte accesses.
//
// Values stay integers below 2^53 (order^2 * (iterations + 1) plus the
// added ones), so every sum is exact and results equal the reference's bit
// for bit in any order of operations.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "prk_internal.hpp"
#include "prk_stencil.h"
#include "sweep.hpp"

namespace prk {
namespace {

constexpr int kT = 64;          // tile edge (the odd-order kernel, and the tile-count unit)
constexpr int kTThreads = 256;  // four waves

struct TransposeArgs {
    double *A;
    double *B;
    int order;
    int ntr, ntc;  // tiles down / across A
    int mode;      // tile order: 0 row-major, 1 column-major, 2 groups of 8 tile rows
};

// Workgroup -> tile (tr, tc) of A.  Mode 2 walks groups of eight tile rows
// column by column, so the workgroups in flight read eight row bands of A
// and write eight adjacent column bands of each B row band.
__device__ inline void tile_of_transpose(const TransposeArgs &a, int b, int &tr, int &tc)
{
    if (a.mode == 1) {
        tc = b / a.ntr;
        tr = b - tc * a.ntr;
    } else if (a.mode == 2) {
        const int per = 8 * a.ntc;            // tiles in a full group
        const int g = b / per, k = b - g * per;
        const int rows = min(8, a.ntr - g * 8);  // tile rows in this group
        tc = k / rows;
        tr = g * 8 + (k - tc * rows);
    } else {
        tr = b / a.ntc;
        tc = b - tr * a.ntc;
    }
}

// One TR x TC tile of A (TR rows, TC columns): VE doubles per lane access
// (2: 16 bytes; 1: odd orders).  NT: non-temporal stores.
template <int TR, int TC, int VE, int NT, int NTH>
__global__ __launch_bounds__(NTH) void transpose_kernel(TransposeArgs a)
{
    constexpr int LPR1 = TC / VE, RPP1 = NTH / LPR1, NP1 = TR / RPP1;  // A rows
    constexpr int LPR2 = TR / VE, RPP2 = NTH / LPR2, NP2 = TC / RPP2;  // B rows
    static_assert(NP1 >= 1 && NP2 >= 1 && RPP1 * NP1 == TR && RPP2 * NP2 == TC, "tile shape");
    constexpr int AUX_ST = NT ? 2 : 0;
    __shared__ double tile[TR][TC + 1];

    int tr, tc;
    tile_of_transpose(a, blockIdx.x, tr, tc);
    const int order = a.order;
    const unsigned row_bytes = (unsigned)order * 8u;
    union V { u32x4 v; double d[2]; };

    // phase 1: A tile rows -> registers -> (A + 1, LDS)
    {
        const int lr = threadIdx.x / LPR1, lc = (threadIdx.x - lr * LPR1) * VE;
        const int c = tc * TC + lc;
        const unsigned off = c < order ? (unsigned)c * 8u : kOOB;
        V x[NP1];
        __amdgpu_buffer_rsrc_t rs[NP1];
#pragma unroll
        for (int p = 0; p < NP1; ++p) {
            const int r = tr * TR + p * RPP1 + lr;
            const bool ok = r < order;
            rs[p] = row_rsrc(a.A + (long long)(ok ? r : 0) * order, ok ? row_bytes : 0u);
            if constexpr (VE == 2)
                x[p].v = bload16<0>(rs[p], off);
            else
                x[p].d[0] = bload_elem<double, 0>(rs[p], off);
        }
#pragma unroll
        for (int p = 0; p < NP1; ++p) {
            V y;
#pragma unroll
            for (int e = 0; e < VE; ++e) {
                y.d[e] = x[p].d[e] + 1.0;
                tile[p * RPP1 + lr][lc + e] = x[p].d[e];
            }
            if constexpr (VE == 2)
                bstore16<AUX_ST>(rs[p], off, y.v);
            else
                bstore_elem<double, AUX_ST>(rs[p], off, y.d[0]);
        }
    }
    __syncthreads();
    // phase 2: B rows tc*TC + q, columns tr*TR + lc: B[y][x] += A[x][y]
    {
        const int lr = threadIdx.x / LPR2, lc = (threadIdx.x - lr * LPR2) * VE;
        const int c = tr * TR + lc;
        const unsigned off = c < order ? (unsigned)c * 8u : kOOB;
        V b[NP2];
        __amdgpu_buffer_rsrc_t rs[NP2];
#pragma unroll
        for (int p = 0; p < NP2; ++p) {
            const int y = tc * TC + p * RPP2 + lr;
            const bool ok = y < order;
            rs[p] = row_rsrc(a.B + (long long)(ok ? y : 0) * order, ok ? row_bytes : 0u);
            if constexpr (VE == 2)
                b[p].v = bload16<0>(rs[p], off);
            else
                b[p].d[0] = bload_elem<double, 0>(rs[p], off);
        }
#pragma unroll
        for (int p = 0; p < NP2; ++p) {
            const int q = p * RPP2 + lr;
#pragma unroll
            for (int e = 0; e < VE; ++e) b[p].d[e] = b[p].d[e] + tile[lc + e][q];
            if constexpr (VE == 2)
                bstore16<AUX_ST>(rs[p], off, b[p].v);
            else
                bstore_elem<double, AUX_ST>(rs[p], off, b[p].d[0]);
        }
    }
}

// One 128 x 128 tile of A on 1024 threads, in two halves through 66 KiB of
// LDS, so that both matrices move in 1 KiB row segments (the 64 x 128 tile
// writes B in 512-byte segments) and two workgroups fit on a CU.  Each
// thread loads 16 doubles of A (eight 16-byte loads, rows p*16 + t/64,
// columns 2(t%64), 2(t%64)+1 of the tile) and writes A + 1 back at once.
// Half h = 0, 1: the threads holding tile columns [64h, 64h+64) park them in
// LDS (128 rows x 64 columns, rows padded by one element); after a barrier
// every thread adds four 16-byte pieces of B rows 64h + q*16 + t/64 (tile
// rows 2(t%64), 2(t%64)+1 as B columns) and stores them; a barrier ends the
// half before the LDS is reused.
template <int NT>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void transpose_kernel_2h(TransposeArgs a)
{
    constexpr int T = 128, H = 64, NTH = 1024, RPP = NTH / 64;  // 16 A rows per pass
    constexpr int AUX_ST = NT ? 2 : 0;
    __shared__ double half[T][H + 1];

    int tr, tc;
    tile_of_transpose(a, blockIdx.x, tr, tc);
    const int order = a.order;
    const unsigned row_bytes = (unsigned)order * 8u;
    union V { u32x4 v; double d[2]; };
    // lr: the wave's index, uniform (row resources stay in scalar registers)
    const int lr = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l2 = (threadIdx.x & 63) * 2;

    // A: tile rows p*16 + lr, columns l2, l2+1
    V x[T / RPP];
    {
        const int c = tc * T + l2;
        const unsigned off = c < order ? (unsigned)c * 8u : kOOB;
        __amdgpu_buffer_rsrc_t rs[T / RPP];
#pragma unroll
        for (int p = 0; p < T / RPP; ++p) {
            const int r = tr * T + p * RPP + lr;
            const bool ok = r < order;
            rs[p] = row_rsrc(a.A + (long long)(ok ? r : 0) * order, ok ? row_bytes : 0u);
            x[p].v = bload16<0>(rs[p], off);
        }
#pragma unroll
        for (int p = 0; p < T / RPP; ++p) {
            V y;
            y.d[0] = x[p].d[0] + 1.0;
            y.d[1] = x[p].d[1] + 1.0;
            bstore16<AUX_ST>(rs[p], off, y.v);
        }
    }
    // B rows tc*T + 64h + q*16 + lr, columns tr*T + l2, l2+1
    const int cb = tr * T + l2;
    const unsigned offb = cb < order ? (unsigned)cb * 8u : kOOB;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if ((l2 >> 6) == h) {
#pragma unroll
            for (int p = 0; p < T / RPP; ++p) {
                half[p * RPP + lr][l2 - H * h] = x[p].d[0];
                half[p * RPP + lr][l2 - H * h + 1] = x[p].d[1];
            }
        }
        __syncthreads();
        V b[H / RPP];
        __amdgpu_buffer_rsrc_t rs[H / RPP];
#pragma unroll
        for (int q = 0; q < H / RPP; ++q) {
            const int y = tc * T + H * h + q * RPP + lr;
            const bool ok = y < order;
            rs[q] = row_rsrc(a.B + (long long)(ok ? y : 0) * order, ok ? row_bytes : 0u);
            b[q].v = bload16<0>(rs[q], offb);
        }
#pragma unroll
        for (int q = 0; q < H / RPP; ++q) {
            const int j = q * RPP + lr;  // column of the half
            b[q].d[0] = b[q].d[0] + half[l2][j];
            b[q].d[1] = b[q].d[1] + half[l2 + 1][j];
            bstore16<AUX_ST>(rs[q], offb, b[q].v);
        }
        __syncthreads();
    }
}

// A = 0, 1, ..., order^2 - 1 (row-major), B = 0 (Cxx11/transpose.cc:110-112).
__global__ __launch_bounds__(256) void transpose_init_kernel(double *A, double *B, long long nn)
{
    for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < nn; k += (long long)gridDim.x * 256) {
        A[k] = (double)k;
        B[k] = 0.0;
    }
}

// Partial sums of |B[j][i] - ((i*order + j)(K+1) + K(K+1)/2)| in a fixed order
// (Cxx11/transpose.cc:137-146).
constexpr int kChkBlocks = 1024;
constexpr int kChkThreads = 256;

__global__ __launch_bounds__(kChkThreads) void transpose_check_kernel(const double *B, int order, int iterations,
                                                                     double *partials)
{
#pragma clang fp contract(off)
    const double addit = (iterations + 1.0) * (iterations / 2.0);
    const long long nn = (long long)order * order;
    double s = 0.0;
    for (long long k = (long long)blockIdx.x * kChkThreads + threadIdx.x; k < nn;
         k += (long long)kChkBlocks * kChkThreads) {
        const long long j = k / order, i = k - j * order;  // k = j*order + i
        const double ref = (double)(i * order + j) * (1.0 + iterations) + addit;
        s += fabs(B[k] - ref);
    }
    __shared__ double red[kChkThreads];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kChkThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kChkThreads) void transpose_check_final(const double *partials, double *out)
{
    __shared__ double red[kChkThreads];
    double s = 0.0;
    for (int k = threadIdx.x; k < kChkBlocks; k += kChkThreads) s += partials[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kChkThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

std::mutex g_chk_mu;
std::vector<double *> g_chk_ws;  // per device: kChkBlocks partials + 1 result

int check_workspace(double **ws)
{
    int dev = 0;
    int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_chk_mu);
    if ((int)g_chk_ws.size() <= dev) g_chk_ws.resize(dev + 1, nullptr);
    if (!g_chk_ws[dev]) {
        void *p = nullptr;
        rc = check_hip(hipMalloc(&p, sizeof(double) * (kChkBlocks + 1)), "transpose check workspace");
        if (rc) return rc;
        g_chk_ws[dev] = (double *)p;
    }
    *ws = g_chk_ws[dev];
    return PRK_OK;
}

int transpose_args(int order, const double *A, const double *B)
{
    if (order < 1) return set_error(PRK_EINVAL, "matrix order must be positive, got %d", order);
    if (order > 46340) return set_error(PRK_EINVAL, "matrix order %d too large (overflow risk)", order);
    if (!A || !B) return set_error(PRK_EINVAL, "null matrix pointer");
    if (A == B) return set_error(PRK_EINVAL, "A and B must be separate matrices");
    if ((uintptr_t)A % 8 || (uintptr_t)B % 8) return set_error(PRK_EINVAL, "matrices must be 8-byte aligned");
    return PRK_OK;
}

int g_tr_shape = -1, g_tr_mode = -1;  // -1: the defaults below

template <int TR, int TC, int VE, int NTH = kTThreads>
int launch_shape(TransposeArgs a, hipStream_t s)
{
    a.ntr = (a.order + TR - 1) / TR;
    a.ntc = (a.order + TC - 1) / TC;
    const unsigned blocks = (unsigned)a.ntr * (unsigned)a.ntc;
    hipLaunchKernelGGL((transpose_kernel<TR, TC, VE, 1, NTH>), dim3(blocks), dim3(NTH), 0, s, a);
    return check_hip(hipGetLastError(), "transpose launch");
}

int launch_2h(TransposeArgs a, hipStream_t s)
{
    a.ntr = (a.order + 127) / 128;
    a.ntc = a.ntr;
    const unsigned blocks = (unsigned)a.ntr * (unsigned)a.ntc;
    hipLaunchKernelGGL(transpose_kernel_2h<1>, dim3(blocks), dim3(1024), 0, s, a);
    return check_hip(hipGetLastError(), "transpose launch");
}

int launch_transpose(int order, double *A, double *B, hipStream_t s)
{
    TransposeArgs a{A, B, order, 0, 0, g_tr_mode < 0 ? 0 : g_tr_mode};
    const bool vec = order % 2 == 0 && (uintptr_t)A % 16 == 0 && (uintptr_t)B % 16 == 0;
    if (!vec) return launch_shape<kT, kT, 1>(a, s);
    // default: 64 x 128 tiles in row-major order, measured best on MI355X at
    // order 32768 (5.24 TB/s; 64 x 64: 4.95, every other shape/order <= 5.21)
    switch (g_tr_shape < 0 ? 2 : g_tr_shape) {
    case 1: return launch_shape<128, 64, 2>(a, s);
    case 2: return launch_shape<64, 128, 2>(a, s);
    case 3: return launch_shape<32, 64, 2>(a, s);
    case 4: return launch_shape<128, 128, 2, 1024>(a, s);
    case 5: return launch_shape<32, 256, 2>(a, s);
    case 6: return launch_shape<256, 32, 2>(a, s);
    case 7: return launch_shape<64, 128, 2, 512>(a, s);
    case 8: return launch_2h(a, s);
    default: return launch_shape<64, 64, 2>(a, s);
    }
}

}  // namespace
}  // namespace prk

using namespace prk;

extern "C" {

int prk_transpose(int order, int tile, double *A, double *B)
{
    (void)tile;  // a tiling hint in the reference; results do not depend on it
    int rc = transpose_args(order, A, B);
    if (rc) return rc;
    return launch_transpose(order, A, B, nullptr);
}

int prk_transpose_ex(int order, double *A, double *B, void *stream)
{
    int rc = transpose_args(order, A, B);
    if (rc) return rc;
    return launch_transpose(order, A, B, (hipStream_t)stream);
}

int prk_set_transpose_variant(int shape, int mode)
{
    if (shape < -1 || shape > 8 || mode < -1 || mode > 2)
        return set_error(PRK_EINVAL, "transpose shape must be -1..8 and tile order -1..2");
    g_tr_shape = shape;
    g_tr_mode = mode;
    return PRK_OK;
}

int prk_transpose_init(int order, double *A, double *B, void *stream)
{
    int rc = transpose_args(order, A, B);
    if (rc) return rc;
    const long long nn = (long long)order * order;
    const unsigned blocks = (unsigned)std::min<long long>((nn + 255) / 256, 16384);
    hipLaunchKernelGGL(transpose_init_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, A, B, nn);
    return check_hip(hipGetLastError(), "transpose init launch");
}

int prk_transpose_run(int order, int iterations, double *A, double *B, void *stream_, double *avg_seconds)
{
    hipStream_t stream = (hipStream_t)stream_;
    int rc = transpose_args(order, A, B);
    if (rc) return rc;
    if (iterations < 1) return set_error(PRK_EINVAL, "iterations must be >= 1");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rc = check_hip(hipEventCreate(&e0), "event");
    if (!rc) rc = check_hip(hipEventCreate(&e1), "event");
    for (int it = 0; it <= iterations && !rc; ++it) {
        if (it == 1) rc = check_hip(hipEventRecord(e0, stream), "event record");
        if (!rc) rc = launch_transpose(order, A, B, stream);
    }
    if (!rc) rc = check_hip(hipEventRecord(e1, stream), "event record");
    if (!rc) rc = check_hip(hipEventSynchronize(e1), "event sync");
    float ms = 0.0f;
    if (!rc) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "event time");
    if (!rc && avg_seconds) *avg_seconds = (double)ms * 1e-3 / iterations;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

int prk_transpose_check(int order, int iterations, const double *B, double *abserr, void *stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    if (order < 1 || !B || !abserr || iterations < 0) return set_error(PRK_EINVAL, "bad transpose check arguments");
    double *ws = nullptr;
    int rc = check_workspace(&ws);
    if (rc) return rc;
    // the workspace is shared per device: one blocking check at a time
    static std::mutex check_mu;
    std::lock_guard<std::mutex> lock(check_mu);
    hipLaunchKernelGGL(transpose_check_kernel, dim3(kChkBlocks), dim3(kChkThreads), 0, stream, B, order, iterations,
                       ws);
    hipLaunchKernelGGL(transpose_check_final, dim3(1), dim3(kChkThreads), 0, stream, (const double *)ws,
                       ws + kChkBlocks);
    rc = check_hip(hipGetLastError(), "transpose check launch");
    if (rc) return rc;
    rc = check_hip(hipMemcpyAsync(abserr, ws + kChkBlocks, sizeof(double), hipMemcpyDeviceToHost, stream),
                   "transpose check copy");
    if (!rc) rc = check_hip(hipStreamSynchronize(stream), "transpose check sync");
    return rc;
}

}  // extern "C"
