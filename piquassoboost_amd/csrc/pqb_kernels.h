// Kernel parameter blocks and launchers shared between the kernel sources and
// the host engine.  Only plain C types cross this header.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pqb {

struct ddc;

// Plain permanents (R rows, NC columns, R <= NC + 1), one per grid row
// (blockIdx.y = matrix index; a single matrix is a grid of height 1).
struct PlainParams {
    const double2* M2;   // per matrix 2R x NC, packed: rows [0,R) = +2A, rows [R,2R) = -2A
    int64_t m2_stride;   // elements between consecutive matrices' M2 (2 R NC)
    int R;               // rows
    int K;               // Gray bits walked inside a chunk (rows R-K .. R-1)
    int B;               // lane bits (rows R-K-B .. R-K-1), B <= 6
    int H;               // group bits (rows 1 .. R-1-K-B), K + B + H = R - 1
    ddc* partials;       // one dd complex per block, gridDim.x per matrix
    unsigned* counter;   // one per matrix, zero on entry; the last block restores it to zero
    double* out;         // 6 per matrix: [re, im, hi_re, lo_re, hi_im, lo_im]
    double scale;        // 2^-(R-1)
};

// Exact integer sweep (sum kept modulo 2^128).
struct ExactParams {
    const int* A;        // R x NC integers, packed
    int R, K, H;
    int group;           // column sums per 64-bit partial product (2, 6 or 13), with
                         // (max_j sum_i |a_ij|)^group < 2^63
    int64_t nlanes;      // = 2^H
    unsigned long long* partials;  // two words per block
    unsigned* counter;
    unsigned long long* out;       // [lo, hi] of sum_delta (...) mod 2^128
};

// ---- repeated-row / batched sweep ----
constexpr int kMaxDigits = 40;
constexpr int kMaxSlots = 40;
constexpr int kBinomStride = 41;  // binomial table: binom[m * 41 + k], 0 <= k <= m <= 40

struct RepStep {          // one step of the low-digit Gray walk
    uint32_t row;         // row of A whose digit moves (low digit d lives in row 1 + d)
    uint32_t dir;         // 1: k += 1 (colsum -= 2a), 0: k -= 1 (colsum += 2a)
    double w;             // prod_low (-1)^k C(m, k) after the step (exact integer)
};

struct RepTask {
    const double2* A;      // R x NS: a_i (slots expanded); row 0 is the anchor, rows
                           // 1..nlow the low digits (in walk order), then the high digits
    const double2* base;   // NS: a_anchor + sum_{low i} m_i a_i  (seed of every chunk before high digits)
    const RepStep* table;  // L entries, entry t = step into Gray index t (entry 0 unused)
    int NS;                // column slots
    int L;                 // low-walk length = prod_low (m_i + 1)
    int nhigh;             // high digits
    int nlow;              // low digits (rows 1 .. nlow of A)
    uint32_t item0;        // first work item of this task (items of a task are contiguous)
    uint32_t nitems;
    uint32_t out_index;    // result slot
    uint32_t nchunks;      // prod_high (m_i + 1)
    double scale;          // 2^-(N-1)
    uint32_t hrow[kMaxDigits];  // row (in A) of high digit h
    uint32_t hrad[kMaxDigits];  // radix m_h + 1 of high digit h (digit 0 varies fastest)
    uint32_t lrad[kMaxDigits];  // radix m_d + 1 of low digit d (row 1 + d); read by the
                                // double-double sweep, which seeds chunks from A exactly
};

struct RepItem {
    uint32_t task;
    uint32_t count;   // chunks in the item
    uint32_t chunk0;  // first chunk
    uint32_t pad;
};

struct RepBatchParams {
    const RepTask* tasks;
    const RepItem* items;
    const unsigned* order;  // processing order of the items (largest first)
    const double* binom;    // binomial table
    ddc* partials;          // one per item
    double* out;            // 6 doubles per task result
    unsigned* head;         // item counter, zero on entry; the reduce kernel resets it
    const unsigned* counts; // device: counts[0] = items, counts[1] = tasks (read by the
                            // kernels, so a captured graph serves every batch that fits)
};

constexpr int kReduceBlocks = 512;    // fixed grid of the per-task reduction
constexpr int kReduceThreads = 256;  // one block of four waves per task

// Kernel instantiations cover slot ranges; rep_range_for() picks the narrowest
// one for a batch (-1: the batch needs splitting at 32 slots).
int rep_range_for(int min_ns, int max_ns);
const void* rep_batch_kernel_fn(int range);
hipError_t launch_rep_batch(int range, const RepBatchParams& p, int grid_blocks, hipStream_t st);
// per-task sum of the item partials (shared by the FP64 and double-double sweeps)
hipError_t launch_rep_reduce(const RepBatchParams& p, hipStream_t st);

// Double-double sweep over the same tasks and items (precision PQB_PREC_INF).
// Ranges kDDRange0 .. kDDRange0 + 3 cover up to 8, 16, 32 and 40 slots with a
// chunk per 1, 2, 4 and 4 lanes; rep_dd_walkers() = chunks per wave.
constexpr int kDDRange0 = 5;
constexpr int kRanges = kDDRange0 + 4;
int rep_dd_range_for(int max_ns);
int rep_dd_walkers(int range);
const void* rep_dd_kernel_fn(int range);
hipError_t launch_rep_dd(int range, const RepBatchParams& p, int grid_blocks, hipStream_t st);

// ---- adjoint sweep (permanent + all first derivatives of square matrices) ----
// Adjoint mode: square n x n matrices.  Minors mode: n x (n+1) matrices, the
// n+1 permanents with one column deleted.
struct GradParams {
    const double2* A;    // count x n x cols, packed row-major
    double2* partials;   // count x (wpt / 4) x E, one per block.  Adjoint: E = n*n (G
                         // row-major); minors: E = cols
    int n;               // rows
    int cols;            // columns: n (adjoint) or n + 1 (minors)
    int L;               // delta-lane bits (<= 4)
    int H;               // high bits: K + L + H = n - 1 (K is the kernel's template argument)
    int wpt;             // waves per task, multiple of 4
};

struct GradReduceParams {
    const double2* partials;  // count x nblk x (main + (tail ? 1 : 0))
    double2* out;        // count x main
    double2* tail;       // count, or NULL
    int main, nblk, count;  // nblk: partials (blocks) per task
    double scale;        // 2^-(n-1)
};

const void* grad_kernel_fn(int M, int K, bool minors);
hipError_t launch_grad(int M, int K, bool minors, const GradParams& p, int nblocks, size_t lds_bytes,
                       hipStream_t st);
hipError_t launch_grad_reduce(const GradReduceParams& p, hipStream_t st);
// perm = sum_j a_0j G_0j per matrix (A, G packed count x n x n)
hipError_t launch_grad_perm(const double2* A, const double2* G, double2* perm, int n, int count, hipStream_t st);

hipError_t launch_plain(int nc, const PlainParams& p, int nblocks, int bdim, int ntasks, hipStream_t st);
hipError_t launch_plain_prep(const double2* A, double2* M2, int n, int64_t count, hipStream_t st);
hipError_t launch_plain_exact(int nc, const ExactParams& p, int nblocks, int bdim, hipStream_t st);
const void* plain_kernel_fn(int nc);

constexpr int kMaxCols = 40;

}  // namespace pqb
