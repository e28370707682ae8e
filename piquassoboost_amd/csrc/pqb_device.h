// Device-side helpers shared by the permanent kernels (gfx950, wave64).
//
// * complex arithmetic on split (re, im) doubles, written so that hipcc
//   contracts each component into one mul + one fma;
// * double-double ("dd") error-free sums used for every reduction of
//   partial sums across lanes, waves and blocks (the reference accumulates
//   with plain per-thread scalars, BBFGPermanentCalculator.hpp:113,169-176,
//   or with the order-bucketed ComplexM, PicComplexM.hpp:125-201);
// * wave / block reductions of dd complex values.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pqb {

constexpr int kWave = 64;

// Read-only data reached through pointers loaded from memory (task tables,
// step tables, matrices of a batch) is accessed through the constant address
// space, so wave-uniform accesses become scalar loads with SGPR results.
// (The host pass of hipcc parses device templates too; it sees plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using cptr = const __attribute__((address_space(4))) T*;
#else
template <class T>
using cptr = const T*;
#endif

template <class T>
__device__ __forceinline__ cptr<T> to_const(const T* q) {
    return (cptr<T>)q;
}

// ---------------------------------------------------------------- complex --

// (ar + i ai) *= (br + i bi); four FP64 VALU instructions after contraction.
__device__ __forceinline__ void cmul(double& ar, double& ai, double br, double bi) {
    const double t = ar * br - ai * bi;
    ai = ar * bi + ai * br;
    ar = t;
}

// ------------------------------------------------------------ double-double --

struct ddc {  // dd complex: value = (hr + lr) + i (hi + li)
    double hr, lr, hi, li;
};

// Knuth TwoSum: s + e == a + b exactly.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ void fast_two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    e = b - (s - a);
}

// dd + dd (Bailey's sloppy add with renormalisation; relative error ~2^-104
// unless the operands cancel, in which case the absolute error stays at
// 2^-104 of the larger operand, which is what a sum of partials needs).
__device__ __forceinline__ void dd_add(double& h, double& l, double h2, double l2) {
    double s, e;
    two_sum(h, h2, s, e);
    e += l + l2;
    fast_two_sum(s, e, h, l);
}

__device__ __forceinline__ void ddc_add(ddc& a, const ddc& b) {
    dd_add(a.hr, a.lr, b.hr, b.lr);
    dd_add(a.hi, a.li, b.hi, b.li);
}

__device__ __forceinline__ ddc ddc_shfl_xor(const ddc& v, int mask) {
    ddc r;
    r.hr = __shfl_xor(v.hr, mask, kWave);
    r.lr = __shfl_xor(v.lr, mask, kWave);
    r.hi = __shfl_xor(v.hi, mask, kWave);
    r.li = __shfl_xor(v.li, mask, kWave);
    return r;
}

// Butterfly reduction over the 64 lanes of a wave; every lane ends with the sum.
__device__ __forceinline__ ddc wave_reduce(ddc v) {
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1) {
        ddc o = ddc_shfl_xor(v, m);
        ddc_add(v, o);
    }
    return v;
}

// Sum over the block; the result is valid in thread 0.  `sh` needs one ddc
// per wave of the block.  Summation order is fixed (deterministic).
__device__ __forceinline__ ddc block_reduce(ddc v, ddc* sh) {
    v = wave_reduce(v);
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = (blockDim.x + kWave - 1) / kWave;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    ddc r = {0.0, 0.0, 0.0, 0.0};
    if (threadIdx.x == 0) {
        r = sh[0];
        for (int w = 1; w < nw; ++w) ddc_add(r, sh[w]);
    }
    return r;
}

__device__ __forceinline__ ddc ddc_from(double re, double im) {
    ddc r = {re, 0.0, im, 0.0};
    return r;
}

// Exact scaling by a power of two.
__device__ __forceinline__ ddc ddc_scale(ddc v, double s) {
    v.hr *= s; v.lr *= s; v.hi *= s; v.li *= s;
    return v;
}

}  // namespace pqb
