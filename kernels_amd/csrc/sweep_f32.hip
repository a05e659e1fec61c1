// Sweep kernels with F32 storage.
#include "sweep_launch.hpp"

namespace prk {
int launch_sweep_f32(const SweepPlan &p, const void *in, void *in_next, void *out, hipStream_t s)
{
    return launch_dtype<F32>(p, in, in_next, out, s);
}
}  // namespace prk
