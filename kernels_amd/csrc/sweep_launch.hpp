// Instantiation and launch of the sweep kernel for one storage dtype.
#pragma once

#include <algorithm>

#include "prk_internal.hpp"
#include "sweep.hpp"

namespace prk {

template <int DT, int P, int R, bool FUSE, int MODE, int PF, int NT, bool RW = false>
int launch_one(const SweepPlan &p, const void *in, void *in_next, void *out, hipStream_t s)
{
    typedef SweepGeom<DT, R> G;
    SweepArgs a;
    a.in = in;
    a.in_next = in_next;
    a.out = out;
    a.pitch = p.pitch;
    a.n = p.n;
    a.span = p.span;
    a.rows_per_wave = p.rows_per_wave;
    a.nstrips = (p.n + G::SW - 1) / G::SW;
    a.ngroups = (a.nstrips + kWavesPerBlock - 1) / kWavesPerBlock;
    a.nrowblocks = (p.span.nrows + p.rows_per_wave - 1) / p.rows_per_wave;
    a.ntiles = a.ngroups * a.nrowblocks;
    a.pad_writable = p.pad_writable ? 1 : 0;
    a.order = p.order;
    unsigned blocks;
    if (p.order == 1) {
        a.chunk = 0;
        blocks = (unsigned)a.ntiles;
    } else if (p.order >= 3) {
        a.chunk = (a.ngroups + 7) / 8 * 8;
        blocks = (unsigned)a.chunk * (unsigned)a.nrowblocks;
    } else if (p.order == 2) {
        a.chunk = (a.ngroups + 7) / 8;
        blocks = 8u * (unsigned)a.chunk * (unsigned)a.nrowblocks;
    } else {
        a.chunk = (a.ntiles + 7) / 8;
        blocks = 8u * (unsigned)a.chunk;
    }
    // Extra dynamic LDS caps the workgroups resident per CU (occupancy knob).
    constexpr size_t kLdsPerCu = 160u * 1024u;
    const size_t lds_cap = p.wgs_per_cu > 0 ? kLdsPerCu / (unsigned)p.wgs_per_cu : 0;
    constexpr size_t lds_bytes = RW ? 0 : sweep_lds_bytes<DT, P, R>();
    size_t dyn = lds_cap > lds_bytes ? lds_cap - lds_bytes : 0;
    // The edge stash of the LDS-ring whole-vector sweeps lives at the start
    // of the dynamic LDS: in the occupancy padding when it is large enough,
    // else in kStashBytes added to it where that leaves the workgroups per CU
    // as they were (at most 8: 32 waves per CU), else not at all (edge waves
    // store their straddling vectors at once).  Its flush addresses
    // kStashRows rows of out from one resource, which must stay below the
    // out-of-range offset.
    a.stash = 0;
    if (!RW && MODE == MODE_VEC &&
        (long long)kStashRows * p.pitch * (long long)sizeof(typename G::T) < (long long)kOOB) {
        if (dyn >= kStashBytes) {
            a.stash = 1;
        } else {
            const size_t wgs_before = std::min<size_t>(8, kLdsPerCu / std::max<size_t>(lds_bytes + dyn, 1));
            const size_t wgs_after = kLdsPerCu / (lds_bytes + kStashBytes);
            if (wgs_after >= wgs_before) {
                dyn = kStashBytes;
                a.stash = 1;
            }
        }
    }
    if constexpr (RW)
        hipLaunchKernelGGL((sweep_rw_kernel<DT, P, R, FUSE, MODE, PF, NT>), dim3(blocks), dim3(kThreads), dyn, s, a);
    else
        hipLaunchKernelGGL((sweep_kernel<DT, P, R, FUSE, MODE, PF, NT>), dim3(blocks), dim3(kThreads), dyn, s, a);
    return check_hip(hipGetLastError(), "sweep launch");
}

// The fused, whole-vector sweep is the hot path: it is instantiated for
// every pipeline depth / cache policy the tuner chooses from (radii beyond
// the reference's 5: one, two rows in flight with non-temporal stores).  The
// other forms use the default.  Four rows in flight (variant 8) made FP64
// star R=2 at n=30000 1.3-1.8 % faster than three at 6- and 8-row blocks
// on MI355X (alternating runs on one box), and slower at 12 and 16 rows.
template <int DT, int P, int R>
int launch_pr(const SweepPlan &p, const void *in, void *in_next, void *out, hipStream_t s)
{
    const bool fuse = in_next != nullptr;
    if (fuse && p.mode == MODE_VEC) {
        const int v = p.variant;
        if (v < 0 || v >= kSweepVariants) return set_error(PRK_EINVAL, "unknown sweep variant %d", v);
        if constexpr (R > kRefMaxRadius) {
            return launch_one<DT, P, R, true, MODE_VEC, 2, 1>(p, in, in_next, out, s);
        } else {
            if constexpr (reg_window_ok<DT, P, R>()) {
                if (v == 6) return launch_one<DT, P, R, true, MODE_VEC, 2, 1, true>(p, in, in_next, out, s);
                if (v == 7) return launch_one<DT, P, R, true, MODE_VEC, 3, 1, true>(p, in, in_next, out, s);
            }
            switch (v) {
            case 0: return launch_one<DT, P, R, true, MODE_VEC, 1, 0>(p, in, in_next, out, s);
            case 1: return launch_one<DT, P, R, true, MODE_VEC, 2, 0>(p, in, in_next, out, s);
            case 2: return launch_one<DT, P, R, true, MODE_VEC, 3, 0>(p, in, in_next, out, s);
            case 3: return launch_one<DT, P, R, true, MODE_VEC, 1, 1>(p, in, in_next, out, s);
            case 4: case 6: return launch_one<DT, P, R, true, MODE_VEC, 2, 1>(p, in, in_next, out, s);
            case 8: return launch_one<DT, P, R, true, MODE_VEC, 4, 1>(p, in, in_next, out, s);
            default: return launch_one<DT, P, R, true, MODE_VEC, 3, 1>(p, in, in_next, out, s);
            }
        }
    }
    if (fuse) {
        if (p.mode == MODE_VEC_LD) return launch_one<DT, P, R, true, MODE_VEC_LD, kPrefetch, 0>(p, in, in_next, out, s);
        return launch_one<DT, P, R, true, MODE_ELEM, kPrefetch, 0>(p, in, in_next, out, s);
    }
    if (p.mode == MODE_VEC) return launch_one<DT, P, R, false, MODE_VEC, kPrefetch, 0>(p, in, in_next, out, s);
    if (p.mode == MODE_VEC_LD) return launch_one<DT, P, R, false, MODE_VEC_LD, kPrefetch, 0>(p, in, in_next, out, s);
    return launch_one<DT, P, R, false, MODE_ELEM, kPrefetch, 0>(p, in, in_next, out, s);
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
        case 6: return launch_pr<DT, STAR, 6>(p, in, in_next, out, s);
        case 7: return launch_pr<DT, STAR, 7>(p, in, in_next, out, s);
        case 8: return launch_pr<DT, STAR, 8>(p, in, in_next, out, s);
        }
    } else {
        switch (p.radius) {
        case 1: return launch_pr<DT, GRID, 1>(p, in, in_next, out, s);
        case 2: return launch_pr<DT, GRID, 2>(p, in, in_next, out, s);
        case 3: return launch_pr<DT, GRID, 3>(p, in, in_next, out, s);
        case 4: return launch_pr<DT, GRID, 4>(p, in, in_next, out, s);
        case 5: return launch_pr<DT, GRID, 5>(p, in, in_next, out, s);
        case 6: return launch_pr<DT, GRID, 6>(p, in, in_next, out, s);
        case 7: return launch_pr<DT, GRID, 7>(p, in, in_next, out, s);
        case 8: return launch_pr<DT, GRID, 8>(p, in, in_next, out, s);
        }
    }
    return set_error(PRK_ENOSTENCIL, "no stencil for pattern %d radius %d", p.pattern, p.radius);
}

}  // namespace prk
