// Sweep kernels with BF16 storage.
#include "sweep_launch.hpp"

namespace prk {
int launch_sweep_bf16(const SweepPlan &p, const void *in, void *in_next, void *out, hipStream_t s)
{
    return launch_dtype<BF16>(p, in, in_next, out, s);
}
}  // namespace prk
