/*
 * CPU oracle for the Cxx11 stencil path of the Parallel Research Kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it.
 * The product (libprkstencil.so) never links or calls it.
 *
 * It restates, in strict IEEE order (built with -ffp-contract=off and
 * without -ffast-math), what the reference computes:
 *   - the weight table of Cxx11/generate-cxx-stencil.py:209-232 (instance)
 *   - the term order of Cxx11/generate-cxx-stencil.py:8-39 (bodygen):
 *     column offset outer, row offset inner, zero weights skipped, the
 *     products summed left to right and the sum added to out last
 *   - the operators star1..5 / grid1..5 of Cxx11/stencil_seq.hpp, and the
 *     same formulas and order for radii 6..8 (beyond the reference's table)
 *   - the driver loop, the in+=1 refresh and the L1 norm of
 *     Cxx11/stencil.cc:170-216
 *   - the OpenMP structure of Cxx11/stencil-openmp.cc:182-238 with
 *     Cxx11/stencil_openmp.hpp (the CPU baseline)
 * and, for the storage forms the reference does not have (FP32 / BF16
 * storage with FP32 accumulation), the exact arithmetic the GPU kernels
 * define (see prko_run_lp).
 *
 * Every entry point returns 0 on success and a negative value on bad
 * arguments.
 */
#ifndef PRK_ORACLE_H
#define PRK_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PRKO_STAR = 0, PRKO_GRID = 1 };
enum { PRKO_F64 = 0, PRKO_F32 = 1, PRKO_BF16 = 2 };
enum { PRKO_INIT_ANALYTIC = 0, PRKO_INIT_SEEDED = 1 };

/* Weight of tap (row offset di, column offset dj); 0 for taps not in the
 * stencil.  Cxx11/generate-cxx-stencil.py:209-232. */
double prko_weight(int pattern, int radius, int di, int dj);

/* Non-zero taps in generator order (column offset outer, row offset inner).
 * Arrays must hold (2R+1)^2 entries; returns the tap count. */
int prko_taps(int pattern, int radius, int *di, int *dj, double *w);

/* Seeded initial value of element idx: SplitMix64(seed, idx) -> 53-bit
 * uniform u in [0,1) -> 2u-1.  Identical on host and device. */
double prko_seeded_value(uint64_t seed, uint64_t idx);

/* Fill in (n*n, pitch n) and zero out.  init: PRKO_INIT_ANALYTIC gives
 * in[i][j] = i+j (Cxx11/stencil.cc:170-181); PRKO_INIT_SEEDED gives
 * prko_seeded_value(seed, i*n+j). */
int prko_init(int init, uint64_t seed, int n, double *in, double *out);

/* out[i*n+j] += S(in) on the interior [R, n-R)^2, tiled as the generated
 * serial operator (Cxx11/stencil_seq.hpp), strict order. */
int prko_stencil(int pattern, int radius, int n, int t, const double *in, double *out);

/* in += 1 on all n*n points (Cxx11/stencil.cc:189). */
int prko_add(int n, double *in);

/* Sequential L1 norm of the interior divided by (n-2R)^2
 * (Cxx11/stencil.cc:199-206). */
double prko_l1norm(int n, int radius, const double *out);

/* The whole driver: init, iterations+1 sweeps each followed by in+=1.
 * out and in_final (may be NULL) receive n*n values. */
int prko_run(int pattern, int radius, int n, int t, int iterations, int init,
             uint64_t seed, double *out, double *in_final);

/* Output rows [a, b) of prko_run without the full grid: the sweep of an
 * output row depends only on input rows within R of it, and in never
 * depends on out.  out_band receives (b-a)*n values. */
int prko_band(int pattern, int radius, int n, int iterations, int init,
              uint64_t seed, int a, int b, double *out_band);

/* The reduced-precision storage forms, storage dtype PRKO_F32 or PRKO_BF16
 * (raw uint16 bit patterns), FP32 accumulation, weights rounded to float.
 * Every stencil is antisymmetric (W(-d) = -W(d)).  Arithmetic per point,
 * x(a, b) the input at row offset a, column offset b:
 *   star: acc = w1*(x(d1)-x(-d1)); acc = fmaf(wk, x(dk)-x(-dk), acc) ...
 *     over the taps of positive weight, ordered by row-offset magnitude
 *     |di| = 0..R and in generator order within one magnitude;
 *   grid: ring sums.  Ring j carries +-ce_j = +-1/(4j(2j-1)R) on its edges
 *     (columns +-j and rows +-j, corners excluded) and +-cc_j = +-1/(4jR) on
 *     the corners (j, j), (-j, -j):
 *       VS_0(b) = x(0, b), VS_j(b) = VS_{j-1}(b) + (x(j, b) + x(-j, b))
 *       V_j = VS_{j-1}(j) - VS_{j-1}(-j)
 *       H_j = D_j(0) + D_j(-1) + D_j(1) + ... + D_j(-(j-1)) + D_j(j-1),
 *             left to right, D_j(b) = x(j, b) - x(-j, b)
 *       C_j = x(j, j) - x(-j, -j)
 *       acc = ce_1*(H_1 + V_1); acc = fmaf(cc_1, C_1, acc);
 *       acc = fmaf(ce_j, H_j + V_j, acc); acc = fmaf(cc_j, C_j, acc) ...
 *     for j = 2..R;
 * then out = T(float(out) + acc) and in = T(float(in) + 1).  BF16 rounding
 * is round-to-nearest-even.  Initial values: T(float(i+j)) or
 * T(float(seeded double)). */
int prko_run_lp(int dtype, int pattern, int radius, int n, int iterations,
                int init, uint64_t seed, void *out, void *in_final);

/* Row band [a,b) of prko_run_lp (out_band: (b-a)*n elements of dtype). */
int prko_band_lp(int dtype, int pattern, int radius, int n, int iterations,
                 int init, uint64_t seed, int a, int b, void *out_band);

/* One reduced-precision sweep out += S(in) over the interior of given n x n
 * arrays of dtype (no in += 1). */
int prko_sweep_lp(int dtype, int pattern, int radius, int n, const void *in, void *out);

/* L1 norm of a reduced-precision out grid, values widened to double. */
double prko_l1norm_lp(int dtype, int n, int radius, const void *out);

/* float <-> bf16 (round to nearest even) used by the storage forms. */
uint16_t prko_f32_to_bf16(float x);
float prko_bf16_to_f32(uint16_t h);

/* OpenMP baseline in the structure of Cxx11/stencil-openmp.cc: one parallel
 * region, first-touch tiled init, iterations+1 sweeps with the orphaned
 * collapse(2) tile loop and simd inner loop, tiled in+=1.  Runs with
 * `threads` OpenMP threads (<=0: the runtime default).  Returns the
 * average seconds per iteration over sweeps 1..iterations in *avg_s and
 * the L1 norm in *norm; *threads_used receives the team size. */
int prko_omp_run(int pattern, int radius, int n, int t, int iterations,
                 int threads, double *avg_s, double *norm, int *threads_used);

/* Matrix transpose, one iteration of Cxx11/transpose.cc:118-131 with its
 * tiling (tile <= 0 or >= order: untiled): B[i][j] += A[j][i];
 * A[j][i] += 1.0, visiting (i, j) tile by tile.  order x order row-major. */
int prko_transpose(int order, int tile, double *A, double *B);

/* The transpose driver (Cxx11/transpose.cc:110-133): A = 0..order^2-1,
 * B = 0, then iterations+1 transposes.  Writes the final A and B. */
int prko_transpose_run(int order, int tile, int iterations, double *A, double *B);

/* The reference's validation sum (Cxx11/transpose.cc:137-146):
 * sum over (i, j) of |B[j][i] - ((i*order + j)(K+1) + K(K+1)/2)|. */
double prko_transpose_abserr(int order, int iterations, const double *B);

#ifdef __cplusplus
}
#endif

#endif
