// Work fetching of the persistent three-centre kernels (generated j3cq_* kernels).
//
// Each workgroup is one wave.  Lane 0 takes the next entry of a queue part with one device-scope
// atomic add on the part's counter (a vector memory atomic; the value is broadcast to the wave through
// readfirstlane).  Wave b starts on part b % kQueueParts -- the dispatcher deals workgroups to the 8
// XCDs round-robin, so part x is served by the waves of XCD x while it lasts -- and moves on to the next
// parts once its own is exhausted; a wave exits when every part is exhausted, so every wave of the grid
// reaches the exit.  The last wave to leave resets the counters, leaving the control block ready for the
// next launch (or graph replay) without a memset.
#pragma once

#include <hip/hip_runtime.h>

#include "md_args.hpp"

namespace eip {

// Next entry index for this wave (wave-uniform), or -1 when every part is exhausted.  'part' holds the
// part the wave last took an entry from.
__device__ __forceinline__ int queue_next(int* __restrict__ ctl, int& part)
{
    for (int k = 0; k < kQueueParts; ++k) {
        const int p = (part + k) & (kQueueParts - 1);
        const int n = __builtin_amdgcn_readfirstlane(((volatile int*) ctl)[kQueueCount + p]);
        int w = n;
        if (threadIdx.x == 0) w = atomicAdd(ctl + kQueueLine * p, 1);
        w = __builtin_amdgcn_readfirstlane(w);
        if (w < n) {
            part = p;
            return __builtin_amdgcn_readfirstlane(((volatile int*) ctl)[kQueueFirst + p]) + w;
        }
    }
    return -1;
}

// Called by every wave after its last queue_next: the wave that finishes last resets the counters.
// Every other wave has taken its last entry before it counted itself finished (its atomics return
// before the finish count is incremented), so no counter is read after the reset.
__device__ __forceinline__ void queue_finish(int* __restrict__ ctl)
{
    if (threadIdx.x == 0) {
        const int done = atomicAdd(ctl + kQueueDone, 1);
        if (done == (int) gridDim.x - 1) {
            for (int p = 0; p < kQueueParts; ++p) atomicExch(ctl + kQueueLine * p, 0);
            atomicExch(ctl + kQueueDone, 0);
        }
    }
}

}  // namespace eip
