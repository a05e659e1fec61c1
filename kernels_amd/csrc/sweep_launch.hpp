// Instantiation and launch of the sweep kernel for one storage dtype.
#pragma once

#include "prk_internal.hpp"
#include "sweep.hpp"

namespace prk {

template <int DT, int P, int R, bool FUSE, bool VEC>
int launch_one(const SweepPlan &p, const void *in, void *in_next, void *out, hipStream_t s)
{
    typedef SweepGeom<DT, R> G;
    SweepArgs a;
    a.in = in;
    a.in_next = in_next;
    a.out = out;
    a.pitch = p.pitch;
    a.n = p.n;
    a.rows_per_wave = p.rows_per_wave;
    a.nstrips = (p.n + G::SW - 1) / G::SW;
    a.ngroups = (a.nstrips + kWavesPerBlock - 1) / kWavesPerBlock;
    a.nrowblocks = (p.n + p.rows_per_wave - 1) / p.rows_per_wave;
    a.ntiles = a.ngroups * a.nrowblocks;
    a.chunk = (a.ntiles + 7) / 8;
    const unsigned blocks = 8u * (unsigned)a.chunk;
    hipLaunchKernelGGL((sweep_kernel<DT, P, R, FUSE, VEC>), dim3(blocks), dim3(kThreads), 0, s, a);
    return check_hip(hipGetLastError(), "sweep launch");
}

template <int DT, int P, int R>
int launch_pr(const SweepPlan &p, const void *in, void *in_next, void *out, hipStream_t s)
{
    const bool fuse = in_next != nullptr;
    if (p.vec)
        return fuse ? launch_one<DT, P, R, true, true>(p, in, in_next, out, s)
                    : launch_one<DT, P, R, false, true>(p, in, in_next, out, s);
    return fuse ? launch_one<DT, P, R, true, false>(p, in, in_next, out, s)
                : launch_one<DT, P, R, false, false>(p, in, in_next, out, s);
}

template <int DT>
int launch_dtype(const SweepPlan &p, const void *in, void *in_next, void *out, hipStream_t s)
{
    if (p.pattern == STAR) {
        switch (p.radius) {
        case 1: return launch_pr<DT, STAR, 1>(p, in, in_next, out, s);
        case 2: return launch_pr<DT, STAR, 2>(p, in, in_next, out, s);
        case 3: return launch_pr<DT, STAR, 3>(p, in, in_next, out, s);
        case 4: return launch_pr<DT, STAR, 4>(p, in, in_next, out, s);
        case 5: return launch_pr<DT, STAR, 5>(p, in, in_next, out, s);
        }
    } else {
        switch (p.radius) {
        case 1: return launch_pr<DT, GRID, 1>(p, in, in_next, out, s);
        case 2: return launch_pr<DT, GRID, 2>(p, in, in_next, out, s);
        case 3: return launch_pr<DT, GRID, 3>(p, in, in_next, out, s);
        case 4: return launch_pr<DT, GRID, 4>(p, in, in_next, out, s);
        case 5: return launch_pr<DT, GRID, 5>(p, in, in_next, out, s);
        }
    }
    return set_error(PRK_ENOSTENCIL, "no stencil for pattern %d radius %d", p.pattern, p.radius);
}

}  // namespace prk
