// Internal declarations shared by the translation units of libprkstencil.so.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "prk_common.hpp"
#include "prk_stencil.h"

namespace prk {

// Thread-local last-error message behind prk_last_error().
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
int check_hip(hipError_t e, const char *what);

// Geometry of one sweep launch.
struct SweepPlan {
    int dtype, pattern, radius, n;  // n: columns (and rows, for a whole grid)
    RowSpan span;                   // rows read and written (prk_common.hpp)
    long long pitch;
    int rows_per_wave;
    int mode;           // MODE_ELEM / MODE_VEC / MODE_VEC_LD (sweep.hpp)
    int variant;        // pipeline depth / cache policy of the fused vector sweep
    bool pad_writable;  // columns n..pitch-1 may be overwritten
    int order;          // workgroup -> tile order (sweep.hpp tile_of)
    int wgs_per_cu;     // occupancy cap, 0 = none
};

// Number of fused-sweep variants: 0..5 the LDS-ring sweep (pipeline depth
// 1..3 x temporal / non-temporal), 6 and 7 the register-window sweep of the
// FP32/BF16 forms at R <= 3 (depth 2 and 3, non-temporal; elsewhere they run
// as 4 and 5), 8 the LDS-ring sweep at depth 4, non-temporal.
constexpr int kSweepVariants = 9;
inline bool register_window_variant(int variant, int radius, int dtype)
{
    return (variant == 6 || variant == 7) && dtype != F64 && radius <= 3;
}

// The L1 norm's first pass leaves one double-double partial per block, at
// most kNormBlocks of them, in a device buffer of kNormPartialsBytes.
constexpr int kNormBlocks = 1024;
constexpr size_t kNormPartialsBytes = 2 * sizeof(double) * kNormBlocks;

// L1 norm into a device double, enqueued on s, with the caller's partials
// buffer (kNormPartialsBytes of device memory, in use until the norm is done).
int l1norm_async(int dtype, int n, long long pitch, int radius, const void *out, double *d_result, void *partials,
                 hipStream_t s);

// Tiling a sweep of this shape uses: tuned entry, nearest tuned size, or the
// built-in default (autotune.hip).
prk_tuning choose_tuning(int pattern, int radius, int dtype, int n);
prk_tuning heuristic_tuning(int pattern, int radius, int dtype, int n);

// Fills the launch geometry; rows_per_wave <= 0 picks automatically.
int plan_sweep(int pattern, int radius, int dtype, int n, long long pitch, const void *in,
               const void *in_next, const void *out, int rows_per_wave, SweepPlan *plan);

// Enqueue one sweep (in_next == nullptr: unfused).
int launch_sweep(const SweepPlan &p, const void *in, void *in_next, void *out, hipStream_t s);
int launch_sweep_f64(const SweepPlan &p, const void *in, void *in_next, void *out, hipStream_t s);
int launch_sweep_f32(const SweepPlan &p, const void *in, void *in_next, void *out, hipStream_t s);
int launch_sweep_bf16(const SweepPlan &p, const void *in, void *in_next, void *out, hipStream_t s);


}  // namespace prk
