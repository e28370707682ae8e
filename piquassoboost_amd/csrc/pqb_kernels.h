// Kernel parameter blocks and launchers shared between the kernel sources and
// the host engine.  Only plain C types cross this header.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pqb {

struct ddc;

// One plain permanent (R rows, NC columns, R <= NC + 1).
struct PlainParams {
    const double2* M2;   // 2R x NC, packed: rows [0,R) = +2A, rows [R,2R) = -2A
    int R;               // rows
    int K;               // Gray bits walked inside a chunk (rows R-K .. R-1)
    int H;               // chunk-index bits (rows 1 .. R-1-K), H = R-1-K
    int64_t nchunks;     // 2^H; lane g takes chunks g, g+L, ... (L = launched lanes)
    ddc* partials;       // one dd complex per block
    unsigned* counter;   // zero on entry; the last block restores it to zero
    double* out;         // [re, im, hi_re, lo_re, hi_im, lo_im]
    double scale;        // 2^-(R-1)
};

// Exact integer sweep (sum kept modulo 2^128).
struct ExactParams {
    const int* A;        // R x NC integers, packed
    int R, K, H;
    int64_t nlanes;      // = 2^H
    unsigned long long* partials;  // two words per block
    unsigned* counter;
    unsigned long long* out;       // [lo, hi] of sum_delta (...) mod 2^128
};

hipError_t launch_plain(int nc, const PlainParams& p, int nblocks, int bdim, hipStream_t st);
hipError_t launch_plain_prep(const double2* A, double2* M2, int n, hipStream_t st);
hipError_t launch_plain_exact(int nc, const ExactParams& p, int nblocks, int bdim, hipStream_t st);
const void* plain_kernel_fn(int nc);

constexpr int kMaxCols = 40;

}  // namespace pqb
