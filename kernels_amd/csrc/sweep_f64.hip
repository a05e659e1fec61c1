// Sweep kernels with F64 storage.
#include "sweep_launch.hpp"

namespace prk {
int launch_sweep_f64(const SweepPlan &p, const void *in, void *in_next, void *out, hipStream_t s)
{
    return launch_dtype<F64>(p, in, in_next, out, s);
}
}  // namespace prk
