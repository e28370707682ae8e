// Repeated-row permanent (BBFG with row multiplicities) and the persistent
// batched launch built on it.
//
// Sum restated from BBFGPermanentCalculatorRepeated.hpp:116-307: with an
// anchor row of multiplicity 1 (:74-101) and the other rows i of
// multiplicity m_i,
//
//   perm = 2^-(N-1) sum_{k_i = 0..m_i} prod_i (-1)^{k_i} C(m_i, k_i)
//                    prod_s ( a_anchor,s + sum_i (m_i - 2 k_i) a_i,s ),
//
// where the product runs over column slots s (a column of multiplicity c_j
// occupies c_j slots, i.e. the power colsum_j^{c_j} of :233-237, :271-273).
// Moving one k_i by +/-1 changes every column sum by -/+ 2 a_i (:263-270).
//
// Layout on the GPU:
//
// * The row digits are split into LOW digits, walked inside a chunk in a fixed
//   reflected mixed-radix Gray order that starts at k = 0, and HIGH digits that
//   index the chunks.  The low walk is the same for every lane, so it is a
//   host-built step table shared by all lanes (and by every task with the same
//   low radices): entry t holds the row whose digit moves and the direction
//   (wave-uniform: scalar loads, SGPR operand, +/-2 as an inline constant of
//   the FMA) and the exact weight prod_low (-1)^k C(m, k) of step t (every
//   weight is an integer below 2^53, so it is exact in a double).
// * A lane seeds a chunk from its index: the high digits come from a
//   mixed-radix decomposition of the chunk index, the column sums from the
//   uniform base row (anchor + sum_low m_i a_i) plus (m_h - 2 k_h) a_h for each
//   high digit, and the chunk weight prod_high (-1)^k C(m, k) from a binomial
//   table.
// * Work is a list of items (task, first chunk, chunk count); each wave of a
//   persistent grid pulls items from an atomic counter, each lane takes every
//   64th chunk of the item, and the wave writes one double-double partial per
//   item.  A second kernel sums each task's partials (double-double, fixed
//   order), scales by 2^-(N-1), and resets the item counter, so the pair can be
//   captured once in a hipGraph and replayed for every batch.
// * The kernel is compiled for slot counts 1..kMaxSlots and switches on the
//   item's slot count (wave-uniform), so one launch serves a batch of uneven
//   sizes.

#include "pqb_device.h"
#include "pqb_kernels.h"

namespace pqb {

// p = w * prod_{s < NS-1} c_s, as up to four chains (w folded into the first factor).
template <int NS>
__device__ __forceinline__ void wprod_first(const double (&cr)[NS], const double (&ci)[NS], double w,
                                            double& pr, double& pi) {
    constexpr int M = NS - 1;
    constexpr int NCH = (M >= 12) ? 4 : ((M >= 4) ? 2 : 1);
    double qr[NCH], qi[NCH];
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
        const int lo = q * M / NCH;
        const int hi = (q + 1) * M / NCH;
        if (q == 0) {
            qr[0] = w * cr[0];
            qi[0] = w * ci[0];
        } else {
            qr[q] = cr[lo];
            qi[q] = ci[lo];
        }
#pragma unroll
        for (int j = lo + 1; j < hi; ++j) cmul(qr[q], qi[q], cr[j], ci[j]);
    }
    if constexpr (NCH == 4) {
        cmul(qr[0], qi[0], qr[1], qi[1]);
        cmul(qr[2], qi[2], qr[3], qi[3]);
        cmul(qr[0], qi[0], qr[2], qi[2]);
    } else if constexpr (NCH == 2) {
        cmul(qr[0], qi[0], qr[1], qi[1]);
    }
    pr = qr[0];
    pi = qi[0];
}

// acc += w * prod_s c_s
template <int NS>
__device__ __forceinline__ void add_wterm(const double (&cr)[NS], const double (&ci)[NS], double w,
                                          double& ar, double& ai) {
    if constexpr (NS == 1) {
        ar = fma(w, cr[0], ar);
        ai = fma(w, ci[0], ai);
    } else {
        double pr, pi;
        wprod_first<NS>(cr, ci, w, pr, pi);
        const double lr = cr[NS - 1], li = ci[NS - 1];
        ar = fma(pr, lr, ar);
        ar = fma(-pi, li, ar);
        ai = fma(pr, li, ai);
        ai = fma(pi, lr, ai);
    }
}

// colsum += f * row, f = +/-2.  The row is an SGPR operand (scalar loads); a
// VALU instruction reads at most one SGPR here, so the wave-uniform factor is
// kept in a VGPR (one move per step) rather than branching between two
// inline-constant forms: the two-path loop needed 224 VGPRs at 26 slots
// against 149 for this one (3 waves per SIMD instead of 2).
template <int NS>
__device__ __forceinline__ void rep_add_row(double (&cr)[NS], double (&ci)[NS], cptr<double2> row, double f) {
    asm volatile("" : "+v"(f));
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        const double2 v = row[j];
        cr[j] = fma(f, v.x, cr[j]);
        ci[j] = fma(f, v.y, ci[j]);
        if ((j & 15) == 15) __builtin_amdgcn_sched_barrier(0);
    }
}

// Weighted sum over one chunk: sum_t w_t prod_s colsum_s(t), times the chunk weight.
template <int NS>
__device__ __forceinline__ void rep_chunk(cptr<RepTask> T, const double* __restrict__ binom, uint32_t chunk,
                                          double& outr, double& outi) {
    double cr[NS], ci[NS];
    // The seed row goes straight to VGPRs (vector loads); the pointer is made
    // opaque per chunk so that the invariant loads are not hoisted out of the
    // chunk loop (their registers would then stay live, and spill, across it).
    const double2* base = T->base;
    asm volatile("" : "+s"(base));
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        const double2 v = base[j];
        cr[j] = v.x;
        ci[j] = v.y;
    }
    double wc = 1.0;
    uint32_t c = chunk;
    cptr<double2> A = to_const(T->A);
    asm volatile("" : "+s"(A));
    const int nhigh = T->nhigh;
    for (int h = 0; h < nhigh; ++h) {
        const uint32_t r = T->hrad[h];
        const uint32_t q = c / r;
        const uint32_t k = c - q * r;
        c = q;
        const int m = (int)r - 1;
        const double f = (double)(m - 2 * (int)k);
        const cptr<double2> row = A + (size_t)T->hrow[h] * NS;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const double2 v = row[j];
            cr[j] = fma(f, v.x, cr[j]);
            ci[j] = fma(f, v.y, ci[j]);
            if ((j & 15) == 15) __builtin_amdgcn_sched_barrier(0);
        }
        const double b = binom[m * kBinomStride + k];
        wc *= (k & 1u) ? -b : b;
    }
    double ar = 0.0, ai = 0.0;
    add_wterm<NS>(cr, ci, 1.0, ar, ai);
    const cptr<RepStep> tab = to_const(T->table);
    const int L = T->L;
    if (L > 1) {
        RepStep e = tab[1];
        for (int t = 1; t < L; ++t) {
            const RepStep en = tab[t + 1 < L ? t + 1 : t];  // next entry, loaded ahead
            const cptr<double2> row = A + (size_t)e.row * NS;
            rep_add_row<NS>(cr, ci, row, e.dir ? -2.0 : 2.0);
            add_wterm<NS>(cr, ci, e.w, ar, ai);
            e = en;
        }
    }
    outr = fma(wc, ar, outr);
    outi = fma(wc, ai, outi);
}

template <int NS>
__device__ __forceinline__ void rep_item(cptr<RepTask> T, const double* __restrict__ binom, uint32_t chunk0,
                                         uint32_t count, int lane, double& sr, double& si) {
    for (uint32_t j = lane; j < count; j += kWave) rep_chunk<NS>(T, binom, chunk0 + j, sr, si);
}

// Dispatch on the slot count over the compile-time range [LO, HI].
template <int LO, int HI>
__device__ __forceinline__ void rep_dispatch(int ns, cptr<RepTask> T, const double* __restrict__ binom,
                                             uint32_t chunk0, uint32_t count, int lane, double& sr, double& si) {
    if constexpr (LO <= HI) {
        if (ns == LO) {
            rep_item<LO>(T, binom, chunk0, count, lane, sr, si);
        } else {
            rep_dispatch<LO + 1, HI>(ns, T, binom, chunk0, count, lane, sr, si);
        }
    }
}

// Persistent grid over the work items; slot counts in [LO, HI].  Registers
// are allocated for the largest case, so the host picks the narrowest range
// that covers the batch.  Up to 32 slots the kernel fits 168 VGPRs: three
// waves per SIMD.
template <int LO, int HI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HI <= 32 ? 3 : 1)))
void rep_batch_kernel(RepBatchParams p) {
    const int lane = threadIdx.x & (kWave - 1);
    const cptr<RepItem> items = to_const(p.items);
    const cptr<unsigned> order = to_const(p.order);
    const cptr<RepTask> tasks = to_const(p.tasks);
    const unsigned nitems = to_const(p.counts)[0];
    // A wave's first item is its wave index: no atomic, and the waves past the
    // batch's item count leave at once (a small batch on the full persistent
    // grid otherwise spent ~70 us on one contended counter).  Further items
    // come from the counter, offset by the number of waves.
    const unsigned nwaves = gridDim.x * (blockDim.x / kWave);
    unsigned idx0 = blockIdx.x * (blockDim.x / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    while (idx0 < nitems) {
        const unsigned idx = order[idx0];  // largest items first; partials stay indexed by item
        const RepItem it = items[idx];
        const cptr<RepTask> T = tasks + it.task;
        double sr = 0.0, si = 0.0;
        rep_dispatch<LO, HI>(T->NS, T, p.binom, it.chunk0, it.count, lane, sr, si);
        const ddc v = wave_reduce(ddc_from(sr, si));
        if (lane == 0) p.partials[idx] = v;
        // Every lane executes the atomic (uniform control flow); only lane 0
        // adds 1, and lane 0's old value is the wave's next item.  A
        // lane-0-only branch around the atomic must not be used here: the
        // structurized loop can then re-enter the item fetch for the other
        // lanes without the atomic, and they spin on a stale index.
        const unsigned got = atomicAdd(p.head, lane == 0 ? 1u : 0u);
        idx0 = nwaves + __builtin_amdgcn_readlane(got, 0);  // wave-uniform
    }
}

// One block of 256 threads per task (grid-stride over tasks): sum the task's
// item partials (items of a task are contiguous) in double-double in a fixed
// order -- thread t takes items t, t+256, ... with several loads in flight, then
// the wave butterflies and the four waves in order -- scale, write; block 0
// resets the item counter for the next launch.  (A single 30-photon task has
// ~2000 items, the double-double n = 30 sweep ~16000: one wave per task took
// 11.7 us and 80 us of dependent loads for them.)
__global__ __launch_bounds__(kReduceThreads) void rep_reduce_kernel(RepBatchParams p) {
    __shared__ ddc red[kReduceThreads / kWave];
    const unsigned ntasks = to_const(p.counts)[1];
    for (unsigned task = blockIdx.x; task < ntasks; task += gridDim.x) {
        const RepTask& T = p.tasks[task];
        ddc s = {0.0, 0.0, 0.0, 0.0};
        const uint32_t end = T.item0 + T.nitems;
#pragma unroll 4
        for (uint32_t i = T.item0 + threadIdx.x; i < end; i += kReduceThreads) {
            const ddc q = p.partials[i];
            ddc_add(s, q);
        }
        s = block_reduce(s, red);
        if (threadIdx.x == 0) {
            s = ddc_scale(s, T.scale);
            double* o = p.out + 6 * (size_t)T.out_index;
            o[0] = s.hr + s.lr;
            o[1] = s.hi + s.li;
            o[2] = s.hr;
            o[3] = s.lr;
            o[4] = s.hi;
            o[5] = s.li;
        }
        __syncthreads();  // `red` is reused by the next task
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *p.head = 0u;
}

// Slot ranges of the batch kernel instantiations.
static const void* rep_fn(int range) {
    switch (range) {
        case 0: return reinterpret_cast<const void*>(&rep_batch_kernel<1, 12>);
        case 1: return reinterpret_cast<const void*>(&rep_batch_kernel<1, 20>);
        case 2: return reinterpret_cast<const void*>(&rep_batch_kernel<1, 26>);
        case 3: return reinterpret_cast<const void*>(&rep_batch_kernel<1, 32>);
        default: return reinterpret_cast<const void*>(&rep_batch_kernel<33, 40>);
    }
}

int rep_range_for(int min_ns, int max_ns) {
    if (max_ns <= 12) return 0;
    if (max_ns <= 20) return 1;
    if (max_ns <= 26) return 2;
    if (max_ns <= 32) return 3;
    return min_ns >= 33 ? 4 : -1;  // a batch mixing <= 32 and > 32 slots needs two launches
}

const void* rep_batch_kernel_fn(int range) { return rep_fn(range); }

hipError_t launch_rep_batch(int range, const RepBatchParams& p, int grid_blocks, hipStream_t st) {
    switch (range) {
        case 0: hipLaunchKernelGGL((rep_batch_kernel<1, 12>), dim3(grid_blocks), dim3(256), 0, st, p); break;
        case 1: hipLaunchKernelGGL((rep_batch_kernel<1, 20>), dim3(grid_blocks), dim3(256), 0, st, p); break;
        case 2: hipLaunchKernelGGL((rep_batch_kernel<1, 26>), dim3(grid_blocks), dim3(256), 0, st, p); break;
        case 3: hipLaunchKernelGGL((rep_batch_kernel<1, 32>), dim3(grid_blocks), dim3(256), 0, st, p); break;
        case 4: hipLaunchKernelGGL((rep_batch_kernel<33, 40>), dim3(grid_blocks), dim3(256), 0, st, p); break;
        default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_rep_reduce(p, st);
}

hipError_t launch_rep_reduce(const RepBatchParams& p, hipStream_t st) {
    hipLaunchKernelGGL(rep_reduce_kernel, dim3(kReduceBlocks), dim3(kReduceThreads), 0, st, p);
    return hipGetLastError();
}

}  // namespace pqb
