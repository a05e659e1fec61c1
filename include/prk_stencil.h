/*
 * prk_stencil.h -- C ABI of the MI355X (gfx950) stencil library
 * (libprkstencil.so).
 *
 * The reference path is the Cxx11 stencil of the Parallel Research Kernels:
 * a table of generated operators
 *     void starR(const int n, const int t, const double* in, double* out)
 * selected by (pattern, radius) (Cxx11/stencil.cc:142-159), the in += 1
 * refresh (Cxx11/stencil.cc:189, HIP `add` Cxx11/stencil-hip.cc:75-83) and
 * the L1-norm validation (Cxx11/stencil.cc:199-216).
 *
 * Conventions (all entry points):
 *   - plain C types only; grid pointers are DEVICE pointers (hipMalloc'd or
 *     any device-accessible allocation), row-major;
 *   - every function returns 0 on success or a negative PRK_E* code, never
 *     aborts; prk_last_error() gives the message of the calling thread's
 *     last failure;
 *   - `stream` arguments are hipStream_t values passed as void* (NULL = the
 *     null stream); work is enqueued asynchronously unless stated otherwise;
 *   - dtype: PRK_F64 (double), PRK_F32 (float), PRK_BF16 (raw bfloat16 bits,
 *     uint16_t); the storage forms accumulate in FP32.
 */
#ifndef PRK_STENCIL_H
#define PRK_STENCIL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRK_STENCIL_ABI_VERSION 1

enum prk_pattern { PRK_STAR = 0, PRK_GRID = 1 };
enum prk_dtype { PRK_F64 = 0, PRK_F32 = 1, PRK_BF16 = 2 };
enum prk_init { PRK_INIT_ANALYTIC = 0, PRK_INIT_SEEDED = 1 };

enum prk_status {
    PRK_OK = 0,
    PRK_EINVAL = -1,   /* bad argument (size, radius, pattern, dtype, alignment) */
    PRK_ENOSTENCIL = -2, /* no operator for this (pattern, radius): the
                            reference's `nothing` (Cxx11/stencil.cc:66-74) */
    PRK_EHIP = -3,     /* HIP runtime error (launch, copy, allocation) */
    PRK_ENOMEM = -4,   /* admission/budget refused or device allocation failed */
    PRK_EBUSY = -5,    /* queue full / slot unavailable */
    PRK_ENOTFOUND = -6,/* unknown request / handle */
    PRK_ETIMEOUT = -7  /* a bounded device wait expired */
};

/* Message of the calling thread's last failure ("" if none). */
const char *prk_last_error(void);
int prk_abi_version(void);

/* ------------------------------------------------------------------------
 * Drop-in operators: the reference's generated table, one entry per
 * (pattern, radius).  Replace
 *   void starR(const int n, const int t, const double* RESTRICT in,
 *              double* RESTRICT out)         Cxx11/stencil_openmp.hpp:18
 *   __global__ void starR(const int n, const prk_float* in,
 *              prk_float* out)               Cxx11/stencil_cuda.hpp:12
 * Semantics: out[i*n+j] += S(in) on the interior R <= i,j < n-R, with the
 * terms in the generator's order (Cxx11/generate-cxx-stencil.py:8-39), so
 * results are bitwise those of Cxx11/stencil_seq.hpp.  in/out are n*n
 * doubles, pitch n; the frame of out is not touched.  t is the reference's
 * tile size: a tiling hint that never changes results (must be >= 1).
 * Enqueued on the null stream, like the reference's HIP launches
 * (Cxx11/stencil-hip.cc:213).
 * ------------------------------------------------------------------------ */
int prk_star1(int n, int t, const double *in, double *out);
int prk_star2(int n, int t, const double *in, double *out);
int prk_star3(int n, int t, const double *in, double *out);
int prk_star4(int n, int t, const double *in, double *out);
int prk_star5(int n, int t, const double *in, double *out);
int prk_grid1(int n, int t, const double *in, double *out);
int prk_grid2(int n, int t, const double *in, double *out);
int prk_grid3(int n, int t, const double *in, double *out);
int prk_grid4(int n, int t, const double *in, double *out);
int prk_grid5(int n, int t, const double *in, double *out);

/* The (pattern, radius) switch of Cxx11/stencil.cc:142-159 as one call.
 * Radii 1..5 are the reference's operators; 6..8 extend the table with the
 * generator's own weight formulas (Cxx11/generate-cxx-stencil.py:209-232).
 * A radius outside 1..8 returns PRK_ENOSTENCIL (the reference aborts on the
 * CPU and silently does nothing on HIP beyond its 5). */
int prk_stencil(int pattern, int radius, int n, int t, const double *in, double *out);

/* in[i*n+j] += 1 on all n*n points.  Replaces the HIP `add` kernel,
 * Cxx11/stencil-hip.cc:75-83 (CPU: Cxx11/stencil.cc:189). Null stream. */
int prk_add(int n, int t, double *in);

/* ------------------------------------------------------------------------
 * Fused sweep: out += S(in) on the interior and in_next = in + 1 on all
 * n*n points, in one pass (in_next must not alias in).  Rows are `pitch`
 * elements apart (pitch >= n); columns n..pitch-1 are never written, nor
 * is the frame of out (its first and last R rows and columns).
 * With pitch*size and the pointers 16-byte aligned the kernel moves whole
 * 16-byte vectors; otherwise it falls back to element accesses.
 * FP64 results are bitwise those of the unfused drop-in pair
 * (prk_starR then prk_add).  Replaces the per-iteration pair of launches of
 * Cxx11/stencil-hip.cc:213,216.
 * ------------------------------------------------------------------------ */
int prk_sweep(int pattern, int radius, int dtype, int n, long long pitch, const void *in,
              void *in_next, void *out, void *stream);

/* Fused sweep of one slab: the share of one GPU when the n x n grid is cut
 * into horizontal slabs, the row-strip form of the decomposition of
 * MPI1/Stencil/stencil.c (its ranks exchange R ghost rows with their
 * neighbours, :338-370, apply the stencil to their own interior points,
 * :406-417, and add 1 to their own points only, :419-420).  The slab owns
 * global rows [row0, row0+rows).  in, in_next and out hold global rows
 * [row0-top, row0+rows+bot) at `pitch` elements apart, top = min(R, row0)
 * and bot = min(R, n-row0-rows): the owned rows with up to R halo rows on
 * each side.  The call does out += S(in) on the owned rows inside the
 * interior R <= i,j < n-R and in_next = in + 1 on the owned rows; halo rows
 * of in_next and out are not written (in_next's halo rows are the
 * neighbours' to send before the next sweep).  Each point's arithmetic is
 * that of prk_sweep, so slabs whose halo rows are refreshed from their
 * neighbours every sweep reproduce the whole-grid results bit for bit.
 * Errors: PRK_EINVAL when the slab leaves [0, n). */
int prk_sweep_slab(int pattern, int radius, int dtype, int n, int row0, int rows, long long pitch,
                   const void *in, void *in_next, void *out, void *stream);

/* Unfused sweep in any dtype: out += S(in) only. */
int prk_stencil_ex(int pattern, int radius, int dtype, int n, long long pitch, const void *in,
                   void *out, void *stream);

/* in += 1 on n*n points of any dtype (storage rounding for FP32/BF16). */
int prk_add_ex(int dtype, int n, long long pitch, void *in, void *stream);

/* Initial values (Cxx11/stencil.cc:170-181): PRK_INIT_ANALYTIC gives
 * in[i][j] = i+j; PRK_INIT_SEEDED gives 2u-1 with u the top 53 bits of
 * SplitMix64(seed, i*n+j) / 2^53.  out is zeroed (may be NULL).  Values are
 * rounded once to the storage dtype (BF16 via FP32). */
int prk_init(int init, uint64_t seed, int dtype, int n, long long pitch, void *in, void *out,
             void *stream);

/* L1 norm of the interior divided by (n-2R)^2 (Cxx11/stencil.cc:199-206),
 * computed on the device in FP64 with compensated per-thread sums and a
 * deterministic double-double cross-block reduction.  Synchronous: waits on
 * `stream` and writes the result to *norm (host). */
int prk_l1norm(int dtype, int n, long long pitch, int radius, const void *out, double *norm,
               void *stream);

/* The norm of prk_l1norm, allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same kernels, with their 1024 double-double block partials in a static
 * device array of 16 KiB per GPU instead of the library's workspace; the
 * result (bitwise prk_l1norm's) goes to *d_norm, device memory the caller
 * owns.  Enqueued on `stream` and not waited for: no allocation, no copy, no
 * synchronization.  Every n fits (the norm uses at most 1024 partials; a
 * request beyond the array would return PRK_EINVAL).  Calls on one GPU share
 * its array: calls that overlap on one GPU (on different streams, or from
 * different threads without ordering) race and give wrong norms. */
int prk_l1norm_noalloc(int dtype, int n, long long pitch, int radius, const void *out, double *d_norm,
                       void *stream);

/* The driver loop of Cxx11/stencil.cc:183-191 on the device: iterations+1
 * fused sweeps over the ping-pong pair (buf0 holds the initial in; buf1 is
 * scratch), sweep 0 untimed, sweeps 1..iterations timed with HIP events on
 * `stream`.  *final_in_buf receives 0 or 1 (the buffer holding in after the
 * last refresh), *avg_seconds the mean time of one timed sweep.
 * Synchronous. */
int prk_run(int pattern, int radius, int dtype, int n, long long pitch, int iterations,
            void *buf0, void *buf1, void *out, void *stream, int *final_in_buf,
            double *avg_seconds);

/* Tiling overrides for the autotuner and tests (process-wide; they never
 * change results).  rows: row-block height per wavefront, 0 = automatic.
 * variant: fused-sweep pipeline variant 0..5 (pipeline depth 1..3 x
 * temporal / non-temporal streams), 6 and 7 the register-window form of the
 * FP32/BF16 sweeps at R <= 3 (depth 2 and 3, non-temporal; other shapes run
 * them as 4 and 5), 8 depth 4, non-temporal; -1 = the tuned or default
 * choice. */
int prk_set_rows_per_wave(int rows);
int prk_set_sweep_variant(int variant);
/* order: workgroup-to-tile order (0 column-major runs per XCD, 1 row-major,
 * 2 column ranges per XCD walked row-major, 3 row-major with the column
 * groups padded to a multiple of the 8 XCDs, 4..7 order 3 with the group ->
 * XCD map rotated every 2^(order-4) row bands); wgs_per_cu: cap on resident
 * workgroups per CU (0 = none).  -1 for either: the tuned choice. */
int prk_set_sweep_order(int order, int wgs_per_cu);

/* ------------------------------------------------------------------------
 * Autotuner.  A tiling of the fused sweep: pipeline variant, row-block
 * height per wavefront, resident workgroups per CU (0 = no cap) and tile
 * order; gbs is the algorithmic bandwidth it measured (0 if not measured).
 * prk_autotune times every candidate for (pattern, radius, dtype, n) on the
 * current device with `reps` sweeps each, then the four fastest twice more
 * with 4*reps sweeps each (it allocates three n x n scratch grids), records
 * the fastest of those in a process-wide table that later sweeps of that
 * shape use, and returns it.  prk_tuning_get returns what a sweep of
 * that shape would use now (tuned entry, nearest tuned size within 1.5x, or
 * the built-in default).  Save/load keep the table in a text file;
 * prk_tuning_load returns the number of entries read.
 * ------------------------------------------------------------------------ */
typedef struct prk_tuning {
    int variant;
    int rows_per_wave;
    int wgs_per_cu;
    int order;
    double gbs;
} prk_tuning;

int prk_autotune(int pattern, int radius, int dtype, int n, int reps, prk_tuning *best);
/* The same search on three caller grids of n rows at `pitch` (two ping-pong
 * buffers and out; their contents are overwritten), on `stream`; allocates
 * no device memory.  (Measured at n = 30000 FP64 on MI355X: grids
 * allocated after prk_autotune had freed its scratch grids swept 3 % slower,
 * with the same tiling, than grids allocated before any tuning.) */
int prk_autotune_buffers(int pattern, int radius, int dtype, int n, long long pitch, void *buf0, void *buf1,
                         void *out, int reps, void *stream, prk_tuning *best);
int prk_tuning_get(int pattern, int radius, int dtype, int n, prk_tuning *t);
int prk_tuning_set(int pattern, int radius, int dtype, int n, const prk_tuning *t);
int prk_tuning_clear(void);
int prk_tuning_save(const char *path);
int prk_tuning_load(const char *path);

/* ------------------------------------------------------------------------
 * Batched solves (no reference counterpart: the reference runs one grid per
 * process).  A solve is the whole driver of Cxx11/stencil.cc:170-216 for one
 * grid: initial values, iterations+1 fused sweeps, L1 norm.  A batch holds
 * many solves of any sizes, patterns, radii (1..8), iteration counts and
 * storage types; prk_batch_create allocates all their grids in one device
 * pool (rows padded to 128 bytes; padding columns are scratch), cuts a
 * sliding-window schedule into rounds (one stage of each active solve per
 * round, one launch per round and (storage type, pattern, radius) class
 * over the tiles of its solves) and captures the whole run in a hipGraph,
 * one branch per class.  Every solve's arithmetic is that of the
 * single-grid sweep, so each result is the same whatever else is in the
 * batch.  rows_per_tile: tile height (0 = the height prk_batch_autotune
 * chose for the solve's key, else 32).  flags: 0, or PRK_BATCH_PERSISTENT for the
 * single-launch tile-queue form, which is disabled (it hung on MI355X) and
 * returns PRK_EINVAL.
 *   prk_batch_run     enqueue one run of every solve on `stream`
 *   prk_batch_wait    wait for it; norms[count] receives the L1 norms
 *   prk_batch_output  copy solve idx's out (which=0) or final in (which=1)
 *                     grid to host, dense n*n
 *   prk_batch_info    device bytes, total tiles, launches per run
 * prk_solve_one_by_one runs the same solves one grid per launch through the
 * single-grid path (init, fused sweeps, norm), optionally forcing the
 * storage type (dtype_override >= 0): the unbatched baseline.
 * ------------------------------------------------------------------------ */
typedef struct prk_solve {
    int pattern;
    int radius;
    int dtype;
    int n;
    int iterations;
    int init;
    uint64_t seed;
} prk_solve;

typedef struct prk_batch prk_batch;

#define PRK_BATCH_PERSISTENT 1

int prk_batch_create(const prk_solve *solves, int count, int rows_per_tile, int flags,
                     prk_batch **out);
int prk_batch_run(prk_batch *b, void *stream);
int prk_batch_wait(prk_batch *b, double *norms);
int prk_batch_output(const prk_batch *b, int idx, int which, void *host, size_t bytes);
int prk_batch_info(const prk_batch *b, long long *device_bytes, long long *tiles, int *launches);
/* Error against FP64 of every result of a batch (no reference counterpart;
 * the reference validates each run's norm, Cxx11/stencil.cc:209-216).
 * `ref` holds the same solves (pattern, radius, n, iterations, init, seed,
 * in the same order) stored in FP64; both batches have run and been waited
 * for.  For solve i, err[3i], err[3i+1], err[3i+2] receive max |x - y|,
 * mean |x - y| and ||x - y||_2 / ||y||_2 over the interior, x the solve's
 * out in `b` widened to double and y its out in `ref` (zeros when the solve
 * is stored in FP64 in `b`, or when ||y|| is 0 for the last).  Computed on
 * the device in a fixed order (the same results on every call); blocks. */
int prk_batch_errors(const prk_batch *b, const prk_batch *ref, double *err);
void prk_batch_destroy(prk_batch *b);
/* Tile heights for batched solves, tuned per (pattern, radius, dtype, size
 * class floor(log2 n)).  prk_batch_autotune groups `solves` by that key,
 * takes up to `sample` solves of each group (iterations capped at 10),
 * times each candidate height (16, 32, 64, 128 rows) over `reps` graph
 * replays of a batch of that group alone (repeated until its rounds fill
 * the GPU), and records the fastest if it beats the default 32 rows by at
 * least 5 %; it blocks, and writes the number of keys tuned.  prk_batch_create with
 * rows_per_tile = 0 then takes each grid's height from the table (32 rows
 * for keys not tuned).  Heights never change results.
 * prk_batch_tuning_get reports the height a grid would get. */
int prk_batch_autotune(const prk_solve *solves, int count, int sample, int reps, int *keys_tuned);
int prk_batch_tuning_get(int pattern, int radius, int dtype, int n, int *rows);
/* Sets the tile height of the key of (pattern, radius, dtype, n), as
 * prk_batch_autotune would (a recorded table can be replayed). */
int prk_batch_tuning_set(int pattern, int radius, int dtype, int n, int rows);
int prk_batch_tuning_clear(void);

/* Tiles per round the schedules of later prk_batch_create calls aim for
 * (process-wide tuning knob; never changes results). */
int prk_set_batch_round_tiles(long long tiles);

/* ------------------------------------------------------------------------
 * Continuous-batching server (no reference counterpart).  Solves are
 * submitted at any time; a worker thread admits them first-come first-served
 * into a device memory pool of budget_bytes, keeps their grids resident
 * until they finish and runs every admitted request together in windows of
 * rounds (one stage of each active request per round, one launch per
 * storage-type/pattern/radius class per round, each window's launches
 * captured into a hipGraph while the previous window runs).  Requests that
 * arrive while a window runs join at the next one.
 *   prk_server_submit   queue one solve; PRK_EBUSY if max_pending requests
 *                       are already queued, PRK_ENOMEM if it can never fit
 *                       the budget.  keep_output: hold the request's grids
 *                       after it finishes until prk_server_release (for
 *                       prk_server_fetch)
 *   prk_server_poll     completed results (up to max_results)
 *   prk_server_drain    wait until every accepted request is answered
 *   prk_server_clock    the clock of the result timestamps (seconds)
 * A result's norm is the L1 norm of the solve (as prk_l1norm); times are
 * submission, admission and completion on prk_server_clock.
 * ------------------------------------------------------------------------ */
typedef struct prk_server prk_server;

typedef struct prk_server_config {
    long long budget_bytes; /* device memory for all admitted requests */
    int max_pending;        /* queued (not yet admitted) requests beyond which submit refuses */
    int max_active;         /* requests in flight at most */
    long long round_tiles;  /* admit while the active requests' tiles per round stay below this */
    int shadow_every;       /* > 0: every shadow_every-th FP32/BF16 request of each (dtype, pattern,
                               radius) class also runs in FP64, and its result carries its error against
                               that FP64 solve; 0: none */
} prk_server_config;

typedef struct prk_server_result {
    uint64_t id;
    double norm;
    double t_submit, t_admit, t_done;
    int status;
    /* Shadowed requests (see shadow_every): err_valid 1 and the error of out
     * against the same solve in FP64 over the interior -- max |x-y|, mean
     * |x-y| and the normwise sqrt(sum (x-y)^2 / sum y^2).  t_done is then
     * when both solves had finished. */
    int err_valid;
    double err_max_abs, err_mean_abs, err_normwise;
} prk_server_result;

typedef struct prk_server_stats {
    long long submitted, served, rejected;
    long long windows, launches;
    long long pending, active;
    long long pool_used, budget;
    double build_seconds; /* host time spent building and capturing windows */
    double wait_seconds;  /* host time spent waiting for windows to finish */
    long long graph_updates; /* windows whose graph exec took new arguments in place instead of being rebuilt */
    long long shadow_solves; /* requests that got an FP64 shadow solve */
} prk_server_stats;

int prk_server_create(const prk_server_config *cfg, prk_server **out);
int prk_server_submit(prk_server *s, const prk_solve *solve, int keep_output, uint64_t *id);
int prk_server_poll(prk_server *s, prk_server_result *results, int max_results, int *count);
int prk_server_drain(prk_server *s, double timeout_s);
int prk_server_fetch(prk_server *s, uint64_t id, int which, void *host, size_t bytes);
int prk_server_release(prk_server *s, uint64_t id);
int prk_server_stats_get(prk_server *s, prk_server_stats *out);
double prk_server_clock(void);
void prk_server_destroy(prk_server *s);
int prk_solve_one_by_one(const prk_solve *solves, int count, int dtype_override, double *norms,
                         void *stream);

/* ------------------------------------------------------------------------
 * HBM calibration (nstream triad A += B + s*C, Cxx11/nstream-hip.cc:67-80;
 * bytes 4*len*8 per call, :221) and a plain copy dst = src.
 * ------------------------------------------------------------------------ */
int prk_triad(long long len, double scalar, double *A, const double *B, const double *C,
              void *stream);
int prk_copy(long long bytes, void *dst, const void *src, void *stream);
/* The fused sweep's traffic without its stencil: in_next = in + 1 and
 * out += in over len FP64 values (two streams read, two written, 32 bytes
 * per value), one 16-byte vector per thread.  A plain elementwise kernel
 * with the sweep's mix of streams, for comparison: on MI355X the tiled
 * sweep moves its bytes about 4 % faster than this.  No reference
 * counterpart; a calibration in the manner of the nstream triad above. */
int prk_stream_mix(long long len, const double *in, double *in_next, double *out, void *stream);
/* dst = src (bytes, a multiple of 16) in a chosen shape, to find the
 * fastest streaming pattern on the device: each lane moves `vpl` 16-byte
 * vectors (1, 2, 4 or 8; a workgroup of 256 lanes covers 256*vpl vectors,
 * lanes adjacent within each vector row, every load issued before the first
 * store); `nt` 1 marks loads and stores non-temporal; `blocks` 0 launches
 * one workgroup per 256*vpl vectors, > 0 launches that many workgroups that
 * stride over the array (persistent form).  prk_copy is vpl 1, nt 1,
 * blocks 0.  No reference counterpart. */
int prk_copy_shape(long long bytes, void *dst, const void *src, int vpl, int nt, int blocks, void *stream);

/* Diagnostics.  No counterpart in the reference; for processes that need to
 * know where a host fault happened, such as a serving process whose worker
 * thread dies.  enable != 0 installs handlers for SIGSEGV, SIGBUS, SIGILL,
 * SIGFPE and SIGABRT that write the faulting thread's native frames to
 * stderr and then pass the signal to the handler installed before them (or
 * the default action); enable == 0 puts those handlers back. */
int prk_crash_backtrace(int enable);

/* ------------------------------------------------------------------------
 * Matrix transpose, the next bandwidth-bound kernel of the suite.  One
 * iteration is B += A^T and A += 1 on order x order FP64 row-major matrices
 * (Cxx11/transpose.cc:118-131; the HIP driver's kernels,
 * Cxx11/transpose-hip.cc:64-119, launched once per iteration at :212-219).
 * A and B are separate device matrices of pitch `order`.
 *   prk_transpose        one iteration on the null stream; `tile` is the
 *                        reference's tiling hint and changes no result
 *   prk_transpose_ex     one iteration on `stream`
 *   prk_transpose_init   A = 0, 1, ..., order^2 - 1, B = 0
 *                        (Cxx11/transpose.cc:110-112)
 *   prk_transpose_run    iterations + 1 iterations, the last `iterations`
 *                        timed (the reference's loop, :114-133); blocks
 *   prk_transpose_check  the reference's validation sum
 *                        sum |B[j][i] - ((i*order + j)(K+1) + K(K+1)/2)|
 *                        (:137-146) on the device; blocks; 0 = validates
 * ------------------------------------------------------------------------ */
int prk_transpose(int order, int tile, double *A, double *B);
int prk_transpose_ex(int order, double *A, double *B, void *stream);
int prk_transpose_init(int order, double *A, double *B, void *stream);
int prk_transpose_run(int order, int iterations, double *A, double *B, void *stream, double *avg_seconds);
int prk_transpose_check(int order, int iterations, const double *B, double *abserr, void *stream);
/* Tuning knob (process-wide, results unchanged): tile shape (rows x
 * columns of A, threads) 0 = 64x64, 1 = 128x64, 2 = 64x128 (default),
 * 3 = 32x64, 4 = 128x128 on 1024 threads, 5 = 32x256, 6 = 256x32,
 * 7 = 64x128 on 512 threads, 8 = 128x128 on 1024 threads in two halves
 * of 64 KiB of LDS; tile order 0 = row-major (default), 1 =
 * column-major, 2 = groups of eight tile rows; -1 = default.  Odd orders
 * always use 64x64 tiles with 8-byte accesses. */
int prk_set_transpose_variant(int shape, int mode);

#ifdef __cplusplus
}
#endif

#endif /* PRK_STENCIL_H */
