// Double-double sweep of the repeated-row sum
//
//   perm = 2^-(N-1) sum_{k_i = 0..m_i} prod_i (-1)^{k_i} C(m_i, k_i)
//                    prod_s ( a_anchor,s + sum_i (m_i - 2 k_i) a_i,s )
//
// (BBFGPermanentCalculatorRepeated.hpp:116-307; the plain sum of
// BBFGPermanentCalculator.hpp:87-235 is the case where every multiplicity is
// 1) with 106-bit significands in every column sum, product and partial sum.
// It serves the reference's infinite-precision calculators
// (BBFGPermanentCalculator.cpp:100-107 and BBFGPermanentCalculatorRepeated.cpp:
// 115-122, ComplexInf / FloatInf of InfinitePrecisionComplex.h), which
// evaluate the same sum in exact MPFR arithmetic and round the result once to
// double (mpfr_get_d, round to nearest).  Here:
//
// * chunk seeds are built from the matrix rows with error-free products
//   (integer coefficient x double) and double-double sums; a Gray step adds
//   +-2 a_row (an exact double) with an error-free sum;
// * each complex product and the weighted accumulation are double-double
//   operations with relative error ~2^-104 of the operands' magnitude;
// * partial sums are combined in double-double and rounded once.
// The result is the correctly rounded sum unless the accumulated error
// (~2^-100 of the sum of |terms|) reaches half an ulp of the result, i.e.
// unless the sum cancels by more than ~2^45.
//
// Layout: the tasks, step tables, work items and persistent grid of the FP64
// batched sweep (perm_repeated.hip).  A chunk is handled by a group of G = 1,
// 2 or 4 lanes: lane q of the group holds the column sums of slots
// q*M .. q*M+M-1 (M = ceil(NS/G), four doubles per slot), multiplies them in
// two independent chains, and the group combines its partial products with
// DPP quad permutations (one for a pair, two for a quad).  A wave works on
// 64/G chunks at a time.  G grows with the slot count so that a lane holds at
// most 8 slots, which fits the 168 VGPRs of three waves per SIMD: G = 1 up to
// 8 slots, 2 up to 16, 4 up to 32 (each exchange repeats a product on every
// lane of the group, so smaller groups do less work); 33..40 slots take 10 per
// lane at two waves per SIMD.
//
// The file also holds grad_perm_kernel (the permanent from the finished
// gradient by a double-double Laplace sum).
//
// This translation unit is compiled with -ffp-contract=off (piquassoboost_amd/
// build.py): an error-free transformation is only error-free when every
// product and sum is rounded exactly as written, and a product fused into the
// following TwoSum silently drops the low half.

#include "pqb_device.h"
#include "pqb_kernels.h"

namespace pqb {
namespace {

// Move a double across the lanes of a quad (DPP quad_perm, bound_ctrl).
template <int CTRL>
__device__ __forceinline__ double quad_move(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
constexpr int kXor1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int kXor2 = 0x4E;  // quad_perm [2,3,0,1]

template <int CTRL>
__device__ __forceinline__ ddc quad_move(const ddc& v) {
    ddc r;
    r.hr = quad_move<CTRL>(v.hr);
    r.lr = quad_move<CTRL>(v.lr);
    r.hi = quad_move<CTRL>(v.hi);
    r.li = quad_move<CTRL>(v.li);
    return r;
}

// (h, l) += v for a double v (error-free sum, renormalised).
__device__ __forceinline__ void dd_add_d(double& h, double& l, double v) {
    double s, t;
    two_sum(h, v, s, t);
    t = t + l;
    fast_two_sum(s, t, h, l);
}

// (h, l) += f * a for an integer-valued f with |f| <= 2^11: f a = p + e exactly.
__device__ __forceinline__ void dd_addmul(double& h, double& l, double f, double a) {
    const double p = f * a;
    const double e = fma(f, a, -p);
    double s, t;
    two_sum(h, p, s, t);
    t = t + (l + e);
    fast_two_sum(s, t, h, l);
}

// a *= b in double-double complex arithmetic.  Each component is the sum of
// two exact products (error-free split with FMA), the four cross terms with
// the low parts, and one renormalisation; the lo x lo terms (2^-106 of the
// result's scale) are dropped.
__device__ __forceinline__ void ddc_mul(ddc& a, const ddc& b) {
    // real: ar br - ai bi
    const double p1 = a.hr * b.hr;
    const double e1 = fma(a.hr, b.hr, -p1);
    const double p2 = a.hi * b.hi;
    const double e2 = fma(a.hi, b.hi, -p2);
    double s, t;
    two_sum(p1, -p2, s, t);
    t = t + (e1 - e2);
    t = fma(a.hr, b.lr, t);
    t = fma(a.lr, b.hr, t);
    t = fma(-a.hi, b.li, t);
    t = fma(-a.li, b.hi, t);
    // imaginary: ar bi + ai br
    const double p3 = a.hr * b.hi;
    const double e3 = fma(a.hr, b.hi, -p3);
    const double p4 = a.hi * b.hr;
    const double e4 = fma(a.hi, b.hr, -p4);
    double u, v;
    two_sum(p3, p4, u, v);
    v = v + (e3 + e4);
    v = fma(a.hr, b.li, v);
    v = fma(a.lr, b.hi, v);
    v = fma(a.hi, b.lr, v);
    v = fma(a.li, b.hr, v);
    fast_two_sum(s, t, a.hr, a.lr);
    fast_two_sum(u, v, a.hi, a.li);
}

// acc += w * x for an integer-valued weight w (|w| < 2^53, exact in a double).
__device__ __forceinline__ void ddc_acc_w(ddc& acc, double w, const ddc& x) {
    {
        const double p = w * x.hr;
        double e = fma(w, x.hr, -p);
        e = fma(w, x.lr, e);
        double s, t;
        two_sum(acc.hr, p, s, t);
        t = t + (acc.lr + e);
        fast_two_sum(s, t, acc.hr, acc.lr);
    }
    {
        const double p = w * x.hi;
        double e = fma(w, x.hi, -p);
        e = fma(w, x.li, e);
        double s, t;
        two_sum(acc.hi, p, s, t);
        t = t + (acc.li + e);
        fast_two_sum(s, t, acc.hi, acc.li);
    }
}

// Product of all column sums of the chunk: the lane's M slots as two
// independent chains (same number of products, twice the ILP), then across
// the group of G lanes.
template <int G, int M>
__device__ __forceinline__ ddc group_product(const double (&hr)[M], const double (&lr)[M], const double (&hi)[M],
                                             const double (&li)[M]) {
    constexpr int H = M / 2;  // second chain starts at slot H (M >= 2)
    ddc p = {hr[0], lr[0], hi[0], li[0]};
    if constexpr (M >= 2) {
        ddc u = {hr[H], lr[H], hi[H], li[H]};
#pragma unroll
        for (int k = 1; k < M; ++k) {
            if (k == H) continue;
            const ddc c = {hr[k], lr[k], hi[k], li[k]};
            if (k < H)
                ddc_mul(p, c);
            else
                ddc_mul(u, c);
        }
        ddc_mul(p, u);
    }
    if constexpr (G >= 2) {
        const ddc x = quad_move<kXor1>(p);
        ddc_mul(p, x);  // product of the lane pair
    }
    if constexpr (G == 4) {
        const ddc x = quad_move<kXor2>(p);
        ddc_mul(p, x);  // product of the quad
    }
    return p;
}

// Weighted double-double sum over one chunk (every lane of the group ends
// with the same value; lane q = 0's copy is used).
template <int G, int M>
__device__ __forceinline__ void dd_chunk(cptr<RepTask> T, const double* __restrict__ binom, uint32_t chunk, int q,
                                         ddc& acc) {
    const int NS = T->NS;
    const double2* __restrict__ A = T->A;
    double hr[M], lr[M], hi[M], li[M];
    bool on[M];
    // anchor row (padding slots: factor 1, never updated)
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const int s = q * M + k;
        on[k] = s < NS;
        const double2 a = on[k] ? A[s] : make_double2(0.0, 0.0);
        hr[k] = on[k] ? a.x : 1.0;
        lr[k] = 0.0;
        hi[k] = a.y;
        li[k] = 0.0;
    }
    // low digits at k = 0: coefficient m
    const int nlow = T->nlow;
    for (int dg = 0; dg < nlow; ++dg) {
        const double f = (double)((int)T->lrad[dg] - 1);
        const double2* row = A + (size_t)(1 + dg) * NS;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const double2 a = on[k] ? row[q * M + k] : make_double2(0.0, 0.0);
            dd_addmul(hr[k], lr[k], f, a.x);
            dd_addmul(hi[k], li[k], f, a.y);
        }
    }
    // high digits from the chunk index: coefficient m - 2k, chunk weight (-1)^k C(m, k)
    double wc = 1.0;
    uint32_t c = chunk;
    const int nhigh = T->nhigh;
    for (int h = 0; h < nhigh; ++h) {
        const uint32_t r = T->hrad[h];
        const uint32_t qq = c / r;
        const uint32_t kd = c - qq * r;
        c = qq;
        const int m = (int)r - 1;
        const double f = (double)(m - 2 * (int)kd);
        const double2* row = A + (size_t)T->hrow[h] * NS;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const double2 a = on[k] ? row[q * M + k] : make_double2(0.0, 0.0);
            dd_addmul(hr[k], lr[k], f, a.x);
            dd_addmul(hi[k], li[k], f, a.y);
        }
        const double b = binom[m * kBinomStride + kd];
        wc *= (kd & 1u) ? -b : b;
    }
    ddc_acc_w(acc, wc, group_product<G, M>(hr, lr, hi, li));
    const cptr<RepStep> tab = to_const(T->table);
    const int L = T->L;
    for (int t = 1; t < L; ++t) {
        const RepStep e = tab[t];
        const double f = e.dir ? -2.0 : 2.0;
        const double2* row = A + (size_t)e.row * NS;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const double2 a = on[k] ? row[q * M + k] : make_double2(0.0, 0.0);
            dd_add_d(hr[k], lr[k], f * a.x);
            dd_add_d(hi[k], li[k], f * a.y);
        }
        ddc_acc_w(acc, wc * e.w, group_product<G, M>(hr, lr, hi, li));
    }
}

template <int G, int MLO, int MHI>
__device__ __forceinline__ void dd_dispatch(int m, cptr<RepTask> T, const double* __restrict__ binom, uint32_t chunk0,
                                            uint32_t count, int q, int walker, ddc& acc) {
    if constexpr (MLO <= MHI) {
        if (m == MLO) {
            for (uint32_t j = (uint32_t)walker; j < count; j += kWave / G)
                dd_chunk<G, MLO>(T, binom, chunk0 + j, q, acc);
        } else {
            dd_dispatch<G, MLO + 1, MHI>(m, T, binom, chunk0, count, q, walker, acc);
        }
    }
}

// Persistent grid over the work items (same item protocol as rep_batch_kernel);
// a chunk per group of G lanes, slot counts up to G * MHI.
template <int G, int MHI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MHI <= 8 ? 3 : 1))) void rep_dd_kernel(
    RepBatchParams p) {
    const int lane = threadIdx.x & (kWave - 1);
    const int q = lane & (G - 1);
    const int walker = lane / G;
    const cptr<RepItem> items = to_const(p.items);
    const cptr<unsigned> order = to_const(p.order);
    const cptr<RepTask> tasks = to_const(p.tasks);
    const unsigned nitems = to_const(p.counts)[0];
    // first item = wave index, then the counter offset by the wave count (see rep_batch_kernel)
    const unsigned nwaves = gridDim.x * (blockDim.x / kWave);
    unsigned idx0 = blockIdx.x * (blockDim.x / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    while (idx0 < nitems) {
        const unsigned idx = order[idx0];
        const RepItem it = items[idx];
        const cptr<RepTask> T = tasks + it.task;
        ddc acc = {0.0, 0.0, 0.0, 0.0};
        dd_dispatch<G, 1, MHI>((T->NS + G - 1) / G, T, p.binom, it.chunk0, it.count, q, walker, acc);
        if (q != 0) acc = ddc{0.0, 0.0, 0.0, 0.0};
        const ddc v = wave_reduce(acc);
        if (lane == 0) p.partials[idx] = v;
        // every lane executes the atomic, only lane 0 adds (see rep_batch_kernel)
        const unsigned got = atomicAdd(p.head, lane == 0 ? 1u : 0u);
        idx0 = nwaves + __builtin_amdgcn_readlane(got, 0);
    }
}

// perm(A) = sum_j a_0j G_0j (Laplace expansion along row 0 of the finished
// gradient), products split exactly (FMA) and summed in double-double.  One
// thread per matrix.
__global__ void grad_perm_kernel(const double2* __restrict__ A, const double2* __restrict__ G,
                                 double2* __restrict__ perm, int n, int count) {
    const int task = blockIdx.x * blockDim.x + threadIdx.x;
    if (task >= count) return;
    const double2* a = A + (size_t)task * n * n;  // row 0
    const double2* g = G + (size_t)task * n * n;  // row 0
    double hr = 0.0, lr = 0.0, hi = 0.0, li = 0.0;
    for (int j = 0; j < n; ++j) {
        const double2 x = a[j], y = g[j];
        // re = x.x y.x - x.y y.y, im = x.x y.y + x.y y.x, each product as p + e exactly
        const double p1 = x.x * y.x, e1 = fma(x.x, y.x, -p1);
        const double p2 = -x.y * y.y, e2 = fma(-x.y, y.y, -p2);
        const double p3 = x.x * y.y, e3 = fma(x.x, y.y, -p3);
        const double p4 = x.y * y.x, e4 = fma(x.y, y.x, -p4);
        dd_add(hr, lr, p1, e1);
        dd_add(hr, lr, p2, e2);
        dd_add(hi, li, p3, e3);
        dd_add(hi, li, p4, e4);
    }
    perm[task] = make_double2(hr + lr, hi + li);
}

}  // namespace

hipError_t launch_grad_perm(const double2* A, const double2* G, double2* perm, int n, int count, hipStream_t st) {
    const int bd = 64;
    hipLaunchKernelGGL(grad_perm_kernel, dim3((count + bd - 1) / bd), dim3(bd), 0, st, A, G, perm, n, count);
    return hipGetLastError();
}

// Ranges: lanes per chunk G and slots per lane, by the batch's largest slot count.
int rep_dd_range_for(int max_ns) {
    if (max_ns <= 8) return kDDRange0;       // G = 1, up to 8 slots
    if (max_ns <= 16) return kDDRange0 + 1;  // G = 2, up to 8 per lane
    if (max_ns <= 32) return kDDRange0 + 2;  // G = 4, up to 8 per lane
    return kDDRange0 + 3;                    // G = 4, up to 10 per lane
}

int rep_dd_walkers(int range) {
    switch (range - kDDRange0) {
        case 0: return kWave;
        case 1: return kWave / 2;
        default: return kWave / 4;
    }
}

const void* rep_dd_kernel_fn(int range) {
    switch (range - kDDRange0) {
        case 0: return reinterpret_cast<const void*>(&rep_dd_kernel<1, 8>);
        case 1: return reinterpret_cast<const void*>(&rep_dd_kernel<2, 8>);
        case 2: return reinterpret_cast<const void*>(&rep_dd_kernel<4, 8>);
        case 3: return reinterpret_cast<const void*>(&rep_dd_kernel<4, 10>);
        default: return nullptr;
    }
}

hipError_t launch_rep_dd(int range, const RepBatchParams& p, int grid_blocks, hipStream_t st) {
    switch (range - kDDRange0) {
        case 0: hipLaunchKernelGGL((rep_dd_kernel<1, 8>), dim3(grid_blocks), dim3(256), 0, st, p); break;
        case 1: hipLaunchKernelGGL((rep_dd_kernel<2, 8>), dim3(grid_blocks), dim3(256), 0, st, p); break;
        case 2: hipLaunchKernelGGL((rep_dd_kernel<4, 8>), dim3(grid_blocks), dim3(256), 0, st, p); break;
        case 3: hipLaunchKernelGGL((rep_dd_kernel<4, 10>), dim3(grid_blocks), dim3(256), 0, st, p); break;
        default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_rep_reduce(p, st);
}

}  // namespace pqb
