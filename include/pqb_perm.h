/*
 * pqb_perm.h -- C ABI of the MI355X permanent engine (libpqb_perm.so).
 *
 * Drop-in boundary for the permanent calculators of piquassoboost.  Every
 * entry point below names the reference interface it replaces
 * (paths relative to the piquassoboost source tree).  Only plain C types
 * cross this boundary: pointers, sizes and status codes.  No C++ exception
 * leaves the library; a non-zero status comes with a message available from
 * pqb_last_error() on the calling thread.
 *
 * Matrices are row-major complex128 (interleaved re, im doubles), `ld` is the
 * row stride in complex elements (ld >= cols), exactly the layout of
 * pic::matrix (piquassoboost/common/source/matrix_base.hpp:44-51) and of a
 * C-contiguous numpy complex128 array.  Occupation vectors are int64, as
 * PicState_int64 (piquassoboost/common/source/PicState.h:33).  A complex
 * result is written as out[0] = real part, out[1] = imaginary part.
 */
#ifndef PQB_PERM_H
#define PQB_PERM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define PQB_OK 0
#define PQB_ERR_ARG 1        /* invalid argument (shape, pointer, code)      */
#define PQB_ERR_SIZE 2       /* problem outside the supported range          */
#define PQB_ERR_STATE 3      /* occupation numbers do not add up (reference   */
                             /* message "Number of input and output states   */
                             /* should be equal")                             */
#define PQB_ERR_DEVICE 4     /* no HIP device / HIP runtime error            */
#define PQB_ERR_EXACT 5      /* exact integer mode requested on a matrix it   */
                             /* cannot represent exactly                      */

/* ---- algorithm codes: which reference calculator's edge cases apply ---- */
/* The arithmetic is the same Gray-code sweep for all three; they differ in
 * the shape rules of the reference:
 *   BBFG  : BBFGPermanentCalculator.hpp:90-95 (empty -> 1, rows >= cols+2 -> 0)
 *   GLYNN : GlynnPermanentCalculator.hpp:50-53 (same rules, recursive order)
 *   SIMPLE: GlynnPermanentCalculatorSimple.hpp:65-70 (empty -> 1,
 *           non-square -> 0)                                                */
#define PQB_ALGO_BBFG 0
#define PQB_ALGO_GLYNN 1
#define PQB_ALGO_SIMPLE 2

/* ---- precision codes ----------------------------------------------------
 * PQB_PREC_DOUBLE   FP64 sweep, double-double reduction of all partial sums
 *                   (reference "double": BBFGPermanentCalculator.cpp:118-121).
 *                   Integer-valued real matrices whose permanent is provably
 *                   representable go through the exact integer sweep instead,
 *                   so the result is the correctly rounded exact permanent.
 * PQB_PREC_EXTENDED Reference "long double" codes (lib 0, 8, 11; repeated
 *                   lib 1, 8).  Same kernels as PQB_PREC_DOUBLE (the error of
 *                   the FP64 sweep with dd reduction is reported per result
 *                   by the tests); kept as a separate code so callers keep
 *                   their intent.
 * PQB_PREC_FP64     Force the FP64 sweep, never the integer route.
 * PQB_PREC_EXACT    Integer sweep modulo 2^128; fails with PQB_ERR_EXACT
 *                   unless every entry is an integer with zero imaginary part
 *                   and pqb_exact_bound() proves 2^(rows-1) |perm| < 2^126.5.
 * PQB_PREC_INF      Double-double sweep: every column sum, product and
 *                   partial sum carries a 106-bit significand and the result
 *                   is rounded once.  Replaces the reference's exact MPFR
 *                   calculators (use_inf: BBFGPermanentCalculator.cpp:100-107,
 *                   BBFGPermanentCalculatorRepeated.cpp:115-122; lib 1 plain,
 *                   lib 9 repeated), whose result is the exact sum rounded to
 *                   nearest: the two agree unless the sum cancels by more than
 *                   ~2^45.  Integer-valued real matrices that PQB_PREC_EXACT
 *                   accepts take the exact integer sweep (plain form).  About
 *                   10x the time of the FP64 sweep. */
#define PQB_PREC_DOUBLE 0
#define PQB_PREC_EXTENDED 1
#define PQB_PREC_FP64 2
#define PQB_PREC_EXACT 3
#define PQB_PREC_INF 4

/* Largest number of columns (distinct column slots for the repeated form)
 * the kernels are instantiated for. */
#define PQB_MAX_COLS 40

/* ---- library / device --------------------------------------------------- */

/* Message of the last failed call on this thread ("" if none). */
const char* pqb_last_error(void);

/* ABI version of this header (bumped on incompatible changes). */
int pqb_abi_version(void);

/* Number of visible HIP devices (0 when there is none; never fails). */
int pqb_device_count(void);

/* Bind the calling thread's engine to a device and create its stream and
 * workspace.  Called implicitly by every compute entry point. */
int pqb_init(int device);

/* ---- plain permanent ------------------------------------------------------
 * Replaces BBFGPermanentCalculator::calculate(matrix&, bool use_extended,
 * bool use_inf)  (piquassoboost/sampling/source/BBFGPermanentCalculator.h:66,
 * .cpp:88-126), GlynnPermanentCalculator<...>::calculate(matrix&)
 * (GlynnPermanentCalculator.h:52, .hpp:47-57) and
 * GlynnPermanentCalculatorSimple<...>::calculate(matrix&)
 * (GlynnPermanentCalculatorSimple.hpp:47-53, 61-168).
 * A: host pointer.  Synchronous.                                             */
int pqb_perm(const double* A, int64_t rows, int64_t cols, int64_t ld, int algo,
             int precision, double* out);

/* Plain permanents of a list of `count` host matrices of any shapes: matrix i
 * at A[i], rows[i] x cols[i], row stride ld[i]; out[2i], out[2i+1] = its
 * permanent.  Replaces the list surface of the plain Python calculator
 * (GlynnPermanentCalculator_Wrapper.hpp:204-293: a list of matrices in, a list
 * of permanents out, one calculator call per matrix).  Same algo and precision
 * codes, edge rules and routes as pqb_perm; the work is regrouped: matrices
 * of one shape on the FP64 route share one launch (a grid row per matrix, each
 * with the launch plan of a single call, so each value is bit-identical to
 * the pqb_perm call on that matrix), the double-double route (PQB_PREC_INF)
 * takes one batched launch (the correctly rounded sum, as per call, whenever
 * the sum cancels by less than ~2^45), exact-route integer matrices one sweep
 * each.  A failing matrix fails the call; the message names its index.
 * Synchronous. */
int pqb_perm_list(const double* const* A, const int64_t* rows, const int64_t* cols, const int64_t* ld,
                  int64_t count, int algo, int precision, double* out);

/* Exact integer permanent of an integer-valued matrix (PQB_PREC_EXACT rules).
 * words[0], words[1] = low and high 64-bit words of the two's complement
 * 128-bit permanent; out = its correctly rounded double value.  No reference
 * counterpart: the reference's exact results come from MPFR (use_inf,
 * BBFGPermanentCalculator.cpp:100-107). */
int pqb_perm_exact(const double* A, int64_t rows, int64_t cols, int64_t ld,
                   int64_t* words, double* out);

/* Host-only eligibility check of the exact integer route (no device needed):
 * PQB_OK with log2_bound = log2 of a proven upper bound on
 * 2^(rows-1) |perm(A)| when the matrix is real and integer valued and the
 * bound is below 2^126.5; PQB_ERR_EXACT otherwise.  Bounds used: for every
 * shape, the sweep's sum is at most 2^(rows-1) prod_j (column j abs sum); for
 * square matrices also |perm(A)| <= prod of row abs sums, and Bregman's bound
 * prod_i (r_i!)^(1/r_i) for entries in {-1, 0, 1}.  No reference
 * counterpart. */
int pqb_exact_bound(const double* A, int64_t rows, int64_t cols, int64_t ld, double* log2_bound);

/* ---- repeated-row permanent -----------------------------------------------
 * Replaces BBFGPermanentCalculatorRepeated::calculate(matrix&,
 * PicState_int64& col_mult, PicState_int64& row_mult, bool use_extended,
 * bool use_inf)  (BBFGPermanentCalculatorRepeated.h:56, .cpp:74-98) and
 * GlynnPermanentCalculatorRepeated<...>::calculate(matrix&,
 * PicState_int64& input_state, PicState_int64& output_state)
 * (GlynnPermanentCalculatorRepeated.h:55-59, .hpp:49-74).
 * Permanent of the matrix whose row i is repeated row_mult[i] times and whose
 * column j is repeated col_mult[j] times (row_mult = output state, col_mult =
 * input state, len(row_mult) = rows, len(col_mult) = cols).
 * algo: PQB_ALGO_BBFG or PQB_ALGO_GLYNN (error message of that calculator). */
int pqb_perm_repeated(const double* A, int64_t rows, int64_t cols, int64_t ld,
                      const int64_t* row_mult, const int64_t* col_mult, int algo,
                      int precision, double* out);

/* ---- batched repeated-row permanents ---------------------------------------
 * One task = the arguments of one pqb_perm_repeated call.  The whole batch
 * runs as one graph-captured launch of a persistent kernel (plus a small
 * reduction) whatever the mix of sizes; out receives 2 doubles per task.
 * Replaces the per-pair loop of BatchednPermanentCalculator::calculate
 * (piquassoboost/sampling/source/BatchedPermanentCalculator.cpp:109-296) and
 * the list-batched surfaces of the Python wrappers
 * (GlynnPermanentCalculator_Wrapper.hpp:204-293,
 * ChinHuhPermanentCalculator_Wrapper.hpp:233-323).  A failing task fails the
 * batch; the message names the task index. */
typedef struct pqb_task {
    const double* A;          /* host, row-major complex128 */
    int64_t rows, cols, ld;
    const int64_t* row_mult;  /* len rows (output state) */
    const int64_t* col_mult;  /* len cols (input state)  */
} pqb_task;

int pqb_perm_batch(const pqb_task* tasks, int64_t ntasks, int algo, int precision, double* out);

/* Permanents of `count` (input state, output state) pairs of ONE interferometer U
 * (rows x cols, row-major complex128, stride ld) in one batched launch: pair i
 * is the repeated-row permanent of U with row multiplicities
 * outputs[i*rows .. i*rows+rows) and column multiplicities
 * inputs[i*cols .. i*cols+cols).  Rows and columns of multiplicity zero drop
 * out during planning, so the caller builds no per-pair matrix.  Replaces
 * BatchednPermanentCalculator::calculate (BatchedPermanentCalculator.cpp:
 * 109-145), which builds the effective matrix of every pair
 * (adaptInterferometer, :412-451) and runs one calculator per pair; same rules,
 * messages and precision codes as pqb_perm_batch.  out: 2 * count doubles. */
int pqb_perm_batch_states(const double* U, int64_t rows, int64_t cols, int64_t ld, const int64_t* inputs,
                          const int64_t* outputs, int64_t count, int algo, int precision, double* out);

/* pqb_perm_batch with an explicit walk length (Gray steps per chunk) instead
 * of the built-in choice; used by the autotuner.  No reference counterpart. */
int pqb_perm_batch_tuned(const pqb_task* tasks, int64_t ntasks, int algo, int precision,
                         int64_t walk_length, double* out);

/* Walk lengths (Gray steps per chunk) chosen by the autotuner, indexed by the
 * task's photon count: walk[n] for n < count, 0 = built-in choice; NULL resets
 * all.  Applies to later pqb_perm_batch / pqb_perm_repeated calls.  No
 * reference counterpart. */
int pqb_set_walk_lengths(const int64_t* walk, int64_t count);

/* Replays the graph of the last batch `iters` times on the engine stream with
 * its inputs already resident and reports the elapsed milliseconds between
 * HIP events around the replays.  No reference counterpart (measurement). */
int pqb_batch_replay(int iters, double* ms);

/* Host-only plan of one repeated task (no device needed): info[0] = column
 * slots, [1] = rows incl. anchor, [2] = low digits, [3] = high digits,
 * [4] = walk length, [5] = chunks.  All zero when no sweep is needed. */
int pqb_repeated_plan(const double* A, int64_t rows, int64_t cols, int64_t ld,
                      const int64_t* row_mult, const int64_t* col_mult,
                      int64_t walk_length, int64_t* info);

/* ---- device-resident plain permanent (inputs already in HBM) -------------
 * dA: device pointer to a packed rows x cols complex128 matrix (ld = cols).
 * d_out: device pointer to 6 doubles: [re, im, hi_re, lo_re, hi_im, lo_im]
 * (the double-double value of the permanent in the last four).
 * d_work: device workspace of at least pqb_perm_device_workspace(rows, cols)
 * bytes whose first 4 bytes are zero before the first call (each call leaves
 * them zero again).  Asynchronous on `stream` (a hipStream_t, NULL = the
 * null stream).  FP64 sweep, BBFG edge rules.  No reference counterpart: it
 * is the form a caller keeping matrices on the GPU (e.g. a PyTorch tensor)
 * uses; it allocates nothing. */
size_t pqb_perm_device_workspace(int64_t rows, int64_t cols);
int pqb_perm_device(const double* dA, int64_t rows, int64_t cols, double* d_out,
                    void* d_work, size_t work_bytes, void* stream);

/* pqb_perm_device without a workspace argument and without any allocation:
 * the workspace (grid counter, per-block partials, +-2A table) is a static
 * device array of the library, pqb_perm_device_noalloc_scratch_bytes() bytes
 * per GPU (about 180 KB, held until the process ends), sized for the largest
 * matrix the kernels take (41 x 40).  Same plan and kernels as
 * pqb_perm_device, so the results are bit-identical to it.  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (two streams,
 * two threads) race and must be ordered by the caller.  Matrices beyond
 * PQB_MAX_COLS columns fail with PQB_ERR_SIZE.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops); the first call on a device looks up the
 * array's address, later calls only launch. */
size_t pqb_perm_device_noalloc_scratch_bytes(void);
int pqb_perm_device_noalloc(const double* dA, int64_t rows, int64_t cols, double* d_out, void* stream);

/* Batch of `count` equal-size plain permanents on device pointers: dA packed
 * count x rows x cols complex128; d_out 6 doubles per matrix (as above);
 * d_work at least pqb_perm_batch_device_workspace bytes, its first
 * 4 * count bytes zero before the first call (each call leaves them zero).
 * One launch for the batch (the matrices share the grid); asynchronous on
 * `stream`; allocates nothing.  Replaces per-matrix calls of the plain
 * calculators over a list (GlynnPermanentCalculator_Wrapper.hpp:204-293) for
 * callers whose matrices are already on the GPU. */
int pqb_perm_batch_device_workspace(int64_t count, int64_t rows, int64_t cols, size_t* bytes);
int pqb_perm_batch_device(const double* dA, int64_t count, int64_t rows, int64_t cols, double* d_out,
                          void* d_work, size_t work_bytes, void* stream);

/* ---- launch plan introspection (host only, no device needed) --------------
 * Writes the sweep decomposition chosen for a rows x cols plain permanent:
 * plan[0] = K (Gray bits walked per group), plan[1] = B (lane bits),
 * plan[2] = H (group bits), plan[3] = launched waves, plan[4] = threads per
 * block, plan[5] = blocks.  `waves_resident` = waves the device keeps
 * resident for this kernel (0 = ask the device). */
int pqb_plain_plan(int64_t rows, int64_t cols, int64_t waves_resident, int64_t* plan);

/* ---- permanent with its gradient (adjoint sweep) ---------------------------
 * For a square n x n complex matrix: perm_out = perm(A) (2 doubles) and
 * grad_out[i*n + j] = d perm(A) / d a_ij = perm(A without row i and column j)
 * (n*n complex, packed row-major), from one Gray-code sweep.  The reference
 * has no derivative; the same numbers come from n*n calls of
 * BBFGPermanentCalculator::calculate on the (n-1) x (n-1) minors
 * (BBFGPermanentCalculator.h:66, .cpp:88-126), which the tests compare with.
 * FP64 sweep, double-double reduction.  n <= PQB_MAX_COLS; n = 0 gives
 * perm 1 and no gradient.  Host pointers, synchronous. */
int pqb_perm_grad(const double* A, int64_t n, int64_t ld, double* perm_out, double* grad_out);

/* Batch of `count` equal-size matrices: matrix b, element (i, j) at
 * A[2 * (b * stride + i * ld + j)]; perm_out: count complex; grad_out:
 * count x n x n complex, packed.  One launch for the whole batch. */
int pqb_perm_grad_batch(const double* A, int64_t count, int64_t n, int64_t ld, int64_t stride,
                        double* perm_out, double* grad_out);

/* Device-resident form (inputs already in HBM, e.g. a PyTorch tensor):
 * dA packed count x n x n complex128; d_perm count complex; d_grad
 * count x n x n complex; d_work at least pqb_perm_grad_workspace bytes (no
 * initial contents needed).  Asynchronous on `stream`; allocates nothing.
 * No reference counterpart (the form a torch.autograd.Function calls). */
int pqb_perm_grad_workspace(int64_t count, int64_t n, size_t* bytes);
int pqb_perm_grad_device(const double* dA, int64_t count, int64_t n, double* d_perm, double* d_grad,
                         void* d_work, size_t work_bytes, void* stream);

/* Column minors of a rows x (rows+1) complex matrix: out[j] (2 doubles each,
 * cols of them) = perm(A without column j), from one sweep over the rows' sign
 * vectors with leave-one-out products of the column sums.  This is what
 * compute_pmf needs per photon step (samplingHelperFunctions.cpp:137-172 calls
 * BBFGPermanentCalculatorRepeated once per occupied input mode, each a minor of
 * the expanded matrix); one call replaces those rows+1 sweeps.  rows = 0 gives
 * out[0] = 1.  Host pointers, synchronous. */
int pqb_perm_minors(const double* A, int64_t rows, int64_t cols, int64_t ld, double* out);

/* The column minors of `count` rows x (rows+1) matrices in one launch: matrix b
 * starts at A + 2 * b * stride doubles (stride >= rows * ld when count > 1) and
 * out[2 * (b * cols + j)] is perm(A_b without column j).  The per-photon-step
 * work of compute_pmf (samplingHelperFunctions.cpp:137-172) for many samples
 * advanced together.  Host pointers, synchronous. */
int pqb_perm_minors_batch(const double* A, int64_t count, int64_t rows, int64_t cols, int64_t ld,
                          int64_t stride, double* out);

/* Launch plan of the adjoint sweep: plan[0] = M (columns per lane), [1] = K
 * (register bits), [2] = L (lane bits), [3] = H (high bits), [4] = waves per
 * matrix, [5] = blocks, [6] = LDS bytes per block.  waves_resident > 0: plan
 * for that many resident waves without a device; 0: ask the device. */
int pqb_grad_plan(int64_t count, int64_t n, int64_t waves_resident, int64_t* plan);

/* ---- timing helper ---------------------------------------------------------
 * Runs `iters` back-to-back pqb_perm_device launches on the engine's stream
 * and reports the elapsed time between HIP events recorded on that stream
 * before the first and after the last launch.  dA, d_out as above. */
int pqb_time_perm_device(const double* dA, int64_t rows, int64_t cols, double* d_out,
                         int iters, double* ms);

/* FP64 issue-rate probe (measurement only): `waves_per_simd` waves on every
 * SIMD each run 16 independent v_fma_f64 chains of 8 x `iters` FMAs
 * (1 <= iters <= 2^30); reports the timed
 * launch's milliseconds (HIP events on the engine stream) and its FP64
 * TFLOP/s (2 flops per FMA).  The ceiling any FP64 VALU kernel can reach on
 * this device at that occupancy.  No reference counterpart. */
int pqb_probe_fp64_fma(int waves_per_simd, int64_t iters, double* ms, double* tflops);

/* Clock stamp (measurement only): a one-wave launch on `stream` writes the
 * shader-clock counter (s_memtime) and the 100 MHz real-time counter
 * (s_memrealtime) to d_dst[0], d_dst[1] (device memory, 16 bytes).  Stamps taken
 * before and after some work on the same stream give its average clock,
 * (c1 - c0) / (r1 - r0) x 100 MHz.  No reference counterpart. */
int pqb_clock_stamp(unsigned long long* d_dst, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PQB_PERM_H */
