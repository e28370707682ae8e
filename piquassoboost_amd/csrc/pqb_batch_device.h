// Completion protocol of the persistent batch kernels (FP64 and double-double
// sweeps): the per-task sum of the item partials is done by the wave that
// finishes a task's last item, and the grid's counters are restored to zero by
// the last wave to leave, so one kernel launch is the whole batch and a
// captured graph can be replayed.
#pragma once

#include "pqb_device.h"
#include "pqb_kernels.h"

namespace pqb {

// Called by every lane of a wave after lane 0 stored item `idx`'s partial:
// counts the item for its task and, in the wave that brings the count to the
// task's item count, sums the task's partials in double-double in item order
// (lane l takes items l, l+64, ...; then the butterfly of wave_reduce, a fixed
// order, so results do not depend on which wave finishes last), scales by
// 2^-(N-1), writes the 6-double result and resets the count.  The task is
// looked up again from the item here, so that only the item index stays live
// across the sweep.
__device__ __forceinline__ void rep_finish_item(const RepBatchParams& p, unsigned idx, int lane) {
    const cptr<RepTask> T = to_const(p.tasks) + to_const(p.items)[idx].task;
    unsigned* const done_count = T->done;
    // release: this item's partial is visible device-wide before it is counted
    __threadfence();
    const unsigned got = atomicAdd(done_count, lane == 0 ? 1u : 0u);
    const unsigned done = __builtin_amdgcn_readlane(got, 0) + 1u;  // wave-uniform
    const uint32_t n = T->nitems;
    if (done != n) return;
    // acquire: every partial of the task is visible to this wave
    __threadfence();
    const uint32_t first = T->item0;
    ddc s = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t i = first + (uint32_t)lane; i < first + n; i += kWave) {
        const ddc q = p.partials[i];
        ddc_add(s, q);
    }
    s = wave_reduce(s);
    if (lane == 0) {
        s = ddc_scale(s, T->scale);
        double* o = T->out;
        o[0] = s.hr + s.lr;
        o[1] = s.hi + s.li;
        o[2] = s.hr;
        o[3] = s.lr;
        o[4] = s.hi;
        o[5] = s.li;
        *done_count = 0u;  // ready for a replay of the same batch
    }
}

// Called by every lane of every wave after it leaves the item loop: the last
// wave resets the item counter and this counter (every wave's last fetch from
// the item counter precedes its increment here).
__device__ __forceinline__ void rep_finish_wave(const RepBatchParams& p, int lane) {
    __threadfence();
    const unsigned nwaves = gridDim.x * (blockDim.x / kWave);
    unsigned* const waves_done = p.head + kHeadWavesDone;
    const unsigned got = atomicAdd(waves_done, lane == 0 ? 1u : 0u);
    if (__builtin_amdgcn_readlane(got, 0) == nwaves - 1u && lane == 0) {
        *p.head = 0u;
        *waves_done = 0u;
    }
}

}  // namespace pqb
