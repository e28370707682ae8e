// Adjoint sweep: the permanent and all its first derivatives
//
//     G_ij = d perm(A) / d a_ij = perm(A with row i and column j removed)
//          = 2^-(n-1) sum_delta (prod_k delta_k) delta_i prod_{l != j} c_l(delta),
//     c_l(delta) = sum_k delta_k a_kl,  delta_0 = +1,
//
// in the same Gray-code sweep that gives perm(A) (the Glynn/BBFG sum of
// BBFGPermanentCalculator.hpp:87-235, differentiated term by term; the
// reference has no derivative).  Per step, the leave-one-out products
// P_-j = prod_{l != j} c_l come from prefix and suffix products of the column
// sums, and y_j = s P_-j (s = prod delta) is added to a running sum T_j.  The
// factor delta_i is applied by summation by parts over the walk:
//
//     sum_t delta_i(t) y(t) = delta_i(end) T(end) - sum_{flips of row i at t} 2 delta_i(t) T(t),
//
// T(t) being the sum of y before step t.  A flip touches one row, so each step
// adds one row of G (n complex multiply-adds) instead of n rows.
//
// MI355X layout (one square matrix per task, a batch of equal-size tasks per
// launch):
//
// * A wave holds 16 sign vectors ("delta lanes"), each spread over a quad of
//   lanes: lane q of the quad owns the M = ceil(n/4) columns q*M .. q*M+M-1
//   (column sums, running sums and suffix products for n = 26 would not fit
//   the 256 VGPRs of one lane).  The quad exchanges partial products with two
//   DPP quad permutations: 4 lanes give the full product and, for each lane,
//   the product of the other three quarters.
// * Sign bits (bit b <-> row n-1-b): bits [0, K) are "register rows" walked
//   in time inside a block of 2^K steps (a fully unrolled reflected Gray walk,
//   so the flipped row is a compile-time index and its G accumulators live in
//   registers; its data too up to M = 5, above that it is read from the staged
//   matrix in LDS); bits [K, K+L) are the delta-lane index (constant per lane);
//   bits [K+L, n-1) are "high rows", flipped in time between blocks.  The time
//   index tau = block * 2^K + t runs a reflected Gray code over the K register
//   and H high bits; a wave owns a contiguous run of blocks.
// * A high-row flip is wave-uniform: the 16 delta lanes' T are summed through
//   an LDS transpose and the flipped row's G accumulator (per wave, in LDS) is
//   updated; the flipped row's data comes from the staged matrix in LDS.
// * At the end of its run, a wave closes every row's sum (delta_i(end) T(end)
//   terms, lane rows from the per-lane sign, register rows from their
//   accumulators), sums over its delta lanes and writes one n x n + 1 partial
//   (G and the permanent) in FP64.  A second kernel adds the partials of a
//   task in double-double and scales by 2^-(n-1).
//
// The running sums are kept in the lane's own sign frame (without the lane
// sign s_lane, which is +-1 per lane and constant over the run), so the only
// per-step sign is the alternation (-1)^t, which is static in the unrolled
// block; s_lane is applied where lanes are combined.

#include "pqb_device.h"
#include "pqb_kernels.h"

namespace pqb {

// Move a double across the lanes of a quad (DPP quad_perm, bound_ctrl).
template <int CTRL>
__device__ __forceinline__ double qperm(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
constexpr int kQuadXor1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;  // quad_perm [2,3,0,1]
// (The same permutation through ds_swizzle moves the 8 moves per step to the
// LDS pipe but puts its latency on the step's critical path: measured 4% slower
// on 256 n=20 + 8 n=26 blocks at 2 waves per SIMD.)

// LDS accesses of one wave to locations written by other lanes of the same
// wave: LDS executes a wave's instructions in order, the fences keep the
// compiler from moving loads and stores across this point.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One Gray step's contribution: y_j = (-1)^t * prod_{l != j} c_l for the
// lane's M columns, accumulated into T.  (The permanent itself is not
// accumulated: it is sum_j a_0j G_0j of the result, grad_perm_kernel.)
// PIN: end the step with T in registers.  T is not read again within the step,
// so instruction selection may emit its updates late (after the next flip's
// loads), which keeps the prefix and suffix products of the step live until
// then; an empty asm statement that takes and returns T closes the step.
template <int M, bool ODD, bool PIN>
__device__ __forceinline__ void grad_term(const double (&cr)[M], const double (&ci)[M], double (&tr)[M],
                                          double (&ti)[M]) {
    // suffix products s_k = c_k ... c_{M-1}; s_0 is this quarter's product
    double sr[M], si[M];
    sr[M - 1] = cr[M - 1];
    si[M - 1] = ci[M - 1];
#pragma unroll
    for (int k = M - 2; k >= 0; --k) {
        sr[k] = cr[k] * sr[k + 1] - ci[k] * si[k + 1];
        si[k] = cr[k] * si[k + 1] + ci[k] * sr[k + 1];
    }
    // quad exchange: x1 = partner quarter, g1 = pair product, x2 = other pair
    const double x1r = qperm<kQuadXor1>(sr[0]), x1i = qperm<kQuadXor1>(si[0]);
    const double g1r = sr[0] * x1r - si[0] * x1i;
    const double g1i = sr[0] * x1i + si[0] * x1r;
    const double x2r = qperm<kQuadXor2>(g1r), x2i = qperm<kQuadXor2>(g1i);
    // prefix starts with the other three quarters' product, signed by (-1)^t
    double pr = x1r * x2r - x1i * x2i;
    double pi = x1r * x2i + x1i * x2r;
    if (ODD) {
        pr = -pr;
        pi = -pi;
    }
#pragma unroll
    for (int k = 0; k < M; ++k) {
        if (k + 1 < M) {
            tr[k] = fma(pr, sr[k + 1], tr[k]);
            tr[k] = fma(-pi, si[k + 1], tr[k]);
            ti[k] = fma(pr, si[k + 1], ti[k]);
            ti[k] = fma(pi, sr[k + 1], ti[k]);
            cmul(pr, pi, cr[k], ci[k]);
        } else {
            tr[k] += pr;
            ti[k] += pi;
        }
    }
    if constexpr (PIN) {
#pragma unroll
        for (int k = 0; k < M; ++k) asm volatile("" : "+v"(tr[k]), "+v"(ti[k]));
    }
}

// MINORS = false: the adjoint sweep of square n x n matrices (n = p.n = p.cols).
// MINORS = true: for an R x (R+1) matrix (R = p.n, p.cols = R + 1), the sweep
// over the R rows' sign vectors with the leave-one-out products only: the sum
// of T over all lanes is, per column j, 2^(R-1) perm(A without column j).  No
// row is differentiated, so flips update the column sums only.
template <int M, int K, bool MINORS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void grad_sweep_kernel(GradParams p) {
    extern __shared__ double2 lds[];
    constexpr int NP = 4 * M;   // padded column count
    constexpr int KA = K > 0 ? K : 1;
    // Register budget (two waves per SIMD): up to M = 5 the register rows' data
    // sits in VGPRs; from M = 6 it is read from the staged matrix at each flip,
    // and the scheduler is kept from interleaving steps across blocks (M = 6, 7)
    // or across steps (M >= 8).  Measured on n = 26 (M = 7, K = 3): 4.09 ms for
    // 8 matrices against 4.32 ms with K = 2 and VGPR rows; on n = 20 the VGPR
    // rows are faster (1.59 ms against 1.64-1.70 ms for 256 matrices).  K = 4
    // (M = 5, 7: one high-row flip per 16 steps) fits in 256 VGPRs only with the
    // running sums and accumulators pinned at the end of each step (without, the
    // M = 7 build spills); measured 1.46 / 4.20 ms against 1.53 / 4.27 ms for
    // K = 3.  Pinning at K = 3 costs 3% on n = 26.
    constexpr bool kRowsInLds = M >= 6;
    constexpr int kSchedEvery = M >= 8 ? 1 : (1 << K);  // steps between scheduling barriers
    constexpr bool kPin = K >= 4;
    const int n = p.n, R = p.n, L = p.L, H = p.H, NCOL = p.cols;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = tid / kWave;
    const int q = lane & 3;
    const int d = lane >> 2;
    const int64_t gw = (int64_t)blockIdx.x * 4 + wid;  // global wave
    const int task = (int)(gw / p.wpt);                 // wpt % 4 == 0: one task per block
    const int wi = (int)(gw - (int64_t)task * p.wpt);

    constexpr int kZh = MINORS ? 0 : 1;
    double2* As = lds;  // R x NP, zero-padded columns
    double2* Tx = lds + (size_t)R * NP + (size_t)wid * (16 + kZh * H) * NP;  // 16 x NP transpose scratch
    double2* Zh = Tx + 16 * NP;                                               // H x NP high-row accumulators

    const double2* Ag = p.A + (size_t)task * R * NCOL;
    for (int e = tid; e < R * NP; e += blockDim.x) {
        const int i = e / NP, j = e - i * NP;
        As[e] = j < NCOL ? Ag[(size_t)i * NCOL + j] : make_double2(0.0, 0.0);
    }
    if (!MINORS)
        for (int e = lane; e < H * NP; e += kWave) Zh[e] = make_double2(0.0, 0.0);
    __syncthreads();

    const int64_t NB = (int64_t)1 << H;
    const int64_t g0 = NB * wi / p.wpt;
    const int64_t g1 = NB * (wi + 1) / p.wpt;
    const bool lane_on = d < (1 << L);

    // ---- seed at tau0 = g0 * 2^K: column sums of this quarter, lane sign
    const int64_t tau0 = g0 << K;
    const int64_t gam0 = tau0 ^ (tau0 >> 1);
    double cr[M], ci[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        cr[k] = 0.0;
        ci[k] = 0.0;
    }
    int negs = 0;
    for (int i = 0; i < R; ++i) {
        int neg = 0;
        if (i > 0) {
            const int b = R - 1 - i;
            if (b < K) neg = (int)((gam0 >> b) & 1);
            else if (b < K + L) neg = (d >> (b - K)) & 1;
            else neg = (int)((gam0 >> (b - L)) & 1);
        }
        negs += neg;
        const double s = neg ? -1.0 : 1.0;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const double2 a = As[i * NP + q * M + k];
            cr[k] = fma(s, a.x, cr[k]);
            ci[k] = fma(s, a.y, ci[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < M; ++k) {
        if (q * M + k >= NCOL) {  // padding column: factor 1
            cr[k] = 1.0;
            ci[k] = 0.0;
        }
        if (!lane_on) {  // unused delta lane (n <= 4): every product is zero
            cr[k] = 0.0;
            ci[k] = 0.0;
        }
    }
    const double sl = lane_on ? ((negs & 1) ? -1.0 : 1.0) : 0.0;
    // lane sign = s_time * (-1)^popcount(d): s_time from the time rows (wave-uniform)
    const double s_time = (__popcll((unsigned long long)gam0) & 1) ? -1.0 : 1.0;

    // register rows' data (row n-1-b for bit b < K) when held in VGPRs
    double arr[kRowsInLds ? 1 : KA][M], ari[kRowsInLds ? 1 : KA][M];
    if constexpr (!kRowsInLds) {
#pragma unroll
        for (int b = 0; b < KA; ++b) {
#pragma unroll
            for (int k = 0; k < M; ++k) {
                const double2 a = (b < K) ? As[(R - 1 - b) * NP + q * M + k] : make_double2(0.0, 0.0);
                arr[b][k] = a.x;
                ari[b][k] = a.y;
            }
        }
    }

    double tr[M], ti[M];            // running sum T (lane frame)
    double zr[KA][M], zi[KA][M];    // register rows' accumulators (lane frame)
#pragma unroll
    for (int k = 0; k < M; ++k) {
        tr[k] = 0.0;
        ti[k] = 0.0;
#pragma unroll
        for (int b = 0; b < KA; ++b) {
            zr[b][k] = 0.0;
            zi[b][k] = 0.0;
        }
    }

    // reduction lanes: column `col` summed by one lane (NP > 32) or a lane pair
    constexpr bool kPairs = NP <= 32;
    const int col = kPairs ? (lane >> 1) : lane;
    const int half = kPairs ? (lane & 1) : 0;
    constexpr int kPer = kPairs ? 8 : 16;

    for (int64_t g = g0; g < g1; ++g) {
        // register rows read from the staged matrix at each flip: the offset is
        // made opaque per block so that the loads are not hoisted out of the
        // walk (their values would then occupy VGPRs throughout)
        int rbase = (R - 1) * NP + q * M;
        if constexpr (kRowsInLds) asm volatile("" : "+v"(rbase));
        if (g > g0) {
            // ---- high-row flip: time bit K + h, h = ctz(g); new delta = +1 iff bit h+1 of g
            const int h = __builtin_ctzll((unsigned long long)g);
            const double f = ((g >> (h + 1)) & 1) ? 2.0 : -2.0;  // 2 * new delta
            // sum over the delta lanes of s_lane * T: T is written in the lane
            // frame and the reducer applies (-1)^popcount(d) (static per j and
            // the lane's half) and s_time (wave-uniform, folded into f)
            if constexpr (!MINORS) {
            wave_lds_sync();
#pragma unroll
            for (int k = 0; k < M; ++k) Tx[d * NP + q * M + k] = make_double2(tr[k], ti[k]);
            wave_lds_sync();
            double sr = 0.0, si = 0.0;
            if (col < NP) {
#pragma unroll
                for (int j = 0; j < kPer; ++j) {
                    const double2 v = Tx[(half * kPer + j) * NP + col];
                    if (__builtin_popcount(j) & 1) {
                        sr -= v.x;
                        si -= v.y;
                    } else {
                        sr += v.x;
                        si += v.y;
                    }
                }
            }
            if (kPairs) {  // lanes of the second half hold d = 8 + j: one more lane bit
                if (half) {
                    sr = -sr;
                    si = -si;
                }
                sr += qperm<kQuadXor1>(sr);
                si += qperm<kQuadXor1>(si);
            }
            if (col < NP && half == 0) {
                const double fs = f * s_time;
                double2 z = Zh[h * NP + col];
                z.x = fma(-fs, sr, z.x);
                z.y = fma(-fs, si, z.y);
                Zh[h * NP + col] = z;
            }
            }  // !MINORS
            const int row = R - 1 - (K + L + h);
#pragma unroll
            for (int k = 0; k < M; ++k) {
                const double2 a = As[row * NP + q * M + k];
                cr[k] = fma(f, a.x, cr[k]);
                ci[k] = fma(f, a.y, ci[k]);
            }
        }
        // step sign (-1)^(tau - tau0): blocks of 2^K >= 2 steps start even;
        // with K = 0 every block is one step and the parity alternates
        if (K == 0 && ((g - g0) & 1)) grad_term<M, true, kPin>(cr, ci, tr, ti);
        else grad_term<M, false, kPin>(cr, ci, tr, ti);
        if constexpr (kSchedEvery == 1) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 1; t < (1 << K); ++t) {
            // ---- register-row flip: bit b = ctz(t); new delta = +1 iff bit b+1 of tau
            const int b = __builtin_ctz((unsigned)t);
            const bool up = (b + 1 < K) ? (((t >> (b + 1)) & 1) != 0) : ((g & 1) != 0);
            const double f = up ? 2.0 : -2.0;
#pragma unroll
            for (int k = 0; k < M; ++k) {
                if constexpr (!MINORS) {
                    zr[b][k] = fma(-f, tr[k], zr[b][k]);
                    zi[b][k] = fma(-f, ti[k], zi[b][k]);
                }
                if constexpr (kRowsInLds) {
                    const double2 a = As[rbase - b * NP + k];
                    cr[k] = fma(f, a.x, cr[k]);
                    ci[k] = fma(f, a.y, ci[k]);
                } else {
                    cr[k] = fma(f, arr[b][k], cr[k]);
                    ci[k] = fma(f, ari[b][k], ci[k]);
                }
            }
            if constexpr (kPin && !MINORS) {  // the accumulators too (see grad_term)
#pragma unroll
                for (int k = 0; k < M; ++k) asm volatile("" : "+v"(zr[b][k]), "+v"(zi[b][k]));
            }
            if (t & 1) grad_term<M, true, kPin>(cr, ci, tr, ti);
            else grad_term<M, false, kPin>(cr, ci, tr, ti);
            if constexpr (kRowsInLds) {
                if ((t + 1) % kSchedEvery == 0) __builtin_amdgcn_sched_barrier(0);
            }
        }
    }

    // ---- close the run: per column sums over the 16 delta lanes
    const int64_t tau_end = (g1 << K) - 1;
    const int64_t gam_end = tau_end ^ (tau_end >> 1);
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < M; ++k) Tx[d * NP + q * M + k] = make_double2(sl * tr[k], sl * ti[k]);
    wave_lds_sync();
    double tbr = 0.0, tbi = 0.0;        // sum of T: row 0
    double lr[4] = {0.0, 0.0, 0.0, 0.0}, li[4] = {0.0, 0.0, 0.0, 0.0};  // lane rows
    if (col < NP) {
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int dd = half * kPer + j;
            const double2 v = Tx[dd * NP + col];
            tbr += v.x;
            tbi += v.y;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const double s = ((dd >> b) & 1) ? -1.0 : 1.0;
                lr[b] = fma(s, v.x, lr[b]);
                li[b] = fma(s, v.y, li[b]);
            }
        }
    }
    if (kPairs) {
        tbr += qperm<kQuadXor1>(tbr);
        tbi += qperm<kQuadXor1>(tbi);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            lr[b] += qperm<kQuadXor1>(lr[b]);
            li[b] += qperm<kQuadXor1>(li[b]);
        }
    }
    if constexpr (MINORS) {
        // per column j: sum over the block's lanes of T_j, in wave order
        __syncthreads();
        double2* acc = As;  // NCOL <= R x NP
        for (int w = 0; w < 4; ++w) {
            if (wid == w && col < NCOL && half == 0) {
                double vr = tbr, vi = tbi;
                if (w > 0) {
                    vr += acc[col].x;
                    vi += acc[col].y;
                }
                acc[col] = make_double2(vr, vi);
            }
            __syncthreads();
        }
        double2* out = p.partials + ((size_t)task * (p.wpt / 4) + wi / 4) * NCOL;
        for (int e = tid; e < NCOL; e += blockDim.x) out[e] = acc[e];
        return;
    }
    double zbr[KA], zbi[KA];
#pragma unroll
    for (int b = 0; b < KA; ++b) {
        zbr[b] = 0.0;
        zbi[b] = 0.0;
        if (b < K) {
            wave_lds_sync();
#pragma unroll
            for (int k = 0; k < M; ++k) Tx[d * NP + q * M + k] = make_double2(sl * zr[b][k], sl * zi[b][k]);
            wave_lds_sync();
            if (col < NP) {
#pragma unroll
                for (int j = 0; j < kPer; ++j) {
                    const double2 v = Tx[(half * kPer + j) * NP + col];
                    zbr[b] += v.x;
                    zbi[b] += v.y;
                }
            }
            if (kPairs) {
                zbr[b] += qperm<kQuadXor1>(zbr[b]);
                zbi[b] += qperm<kQuadXor1>(zbi[b]);
            }
        }
    }

    // ---- the block's four waves add their n x n partials in the staged
    // matrix's place (no wave reads it any more), in wave order
    __syncthreads();
    double2* acc = As;  // n x n <= n x NP
    for (int w = 0; w < 4; ++w) {
        if (wid == w && col < n && half == 0) {
            auto put = [&](int row, double vr, double vi) {
                double2* a = acc + row * n + col;
                if (w > 0) {
                    const double2 o = *a;
                    vr += o.x;
                    vi += o.y;
                }
                *a = make_double2(vr, vi);
            };
            put(0, tbr, tbi);  // row 0
#pragma unroll
            for (int b = 0; b < 4; ++b)  // static indices keep lr, li in registers
                if (b < L) put(R - 1 - (K + b), lr[b], li[b]);
#pragma unroll
            for (int b = 0; b < K; ++b) {
                const double s = ((gam_end >> b) & 1) ? -1.0 : 1.0;
                put(R - 1 - b, fma(s, tbr, zbr[b]), fma(s, tbi, zbi[b]));
            }
            for (int h = 0; h < H; ++h) {
                const double s = ((gam_end >> (K + h)) & 1) ? -1.0 : 1.0;
                const double2 z = Zh[h * NP + col];
                put(R - 1 - (K + L + h), fma(s, tbr, z.x), fma(s, tbi, z.y));
            }
        }
        __syncthreads();
    }
    double2* out = p.partials + ((size_t)task * (p.wpt / 4) + wi / 4) * (n * n);
    for (int e = tid; e < n * n; e += blockDim.x) out[e] = acc[e];
}

// (perm(A) = sum_j a_0j G_0j, the Laplace expansion along row 0 of the
// finished gradient, is grad_perm_kernel in perm_dd.hip: its exact product
// splits need a translation unit without contraction.)

// Sum of the per-block partials of each task (double-double, fixed order),
// scaled.  Block = 64 elements x 4 slices of the partial index.
__global__ __launch_bounds__(256) void grad_reduce_kernel(GradReduceParams p) {
    __shared__ ddc red[4][64];
    const int E = p.main + (p.tail ? 1 : 0);
    const int chunks = (E + 63) / 64;
    const int task = blockIdx.x / chunks;
    const int el = threadIdx.x & 63;
    const int slice = threadIdx.x >> 6;
    const int e = (blockIdx.x - task * chunks) * 64 + el;
    ddc s = {0.0, 0.0, 0.0, 0.0};
    if (e < E) {
        const double2* src = p.partials + (size_t)task * p.nblk * E + e;
#pragma unroll 8
        for (int b = slice; b < p.nblk; b += 4) {
            const double2 v = src[(size_t)b * E];
            ddc_add(s, ddc_from(v.x, v.y));
        }
    }
    red[slice][el] = s;
    __syncthreads();
    if (slice == 0 && e < E) {
        for (int k = 1; k < 4; ++k) ddc_add(s, red[k][el]);
        s = ddc_scale(s, p.scale);
        const double2 r = make_double2(s.hr + s.lr, s.hi + s.li);
        if (e < p.main) p.out[(size_t)task * p.main + e] = r;
        else p.tail[task] = r;
    }
}

// ------------------------------------------------------------ dispatch --

#define PQB_GRAD_CASES(X) \
    X(1, 0) X(2, 0) X(2, 1) X(2, 2) X(2, 3) X(3, 3) X(4, 3) X(5, 3) X(6, 3) X(7, 3) X(8, 3) X(9, 2) X(10, 2) \
    X(5, 4) X(7, 4)

const void* grad_kernel_fn(int M, int K, bool minors) {
#define PQB_CASE(MM, KK)                                                                        \
    if (M == MM && K == KK)                                                                     \
        return minors ? reinterpret_cast<const void*>(&grad_sweep_kernel<MM, KK, true>)         \
                      : reinterpret_cast<const void*>(&grad_sweep_kernel<MM, KK, false>);
    PQB_GRAD_CASES(PQB_CASE)
#undef PQB_CASE
    return nullptr;
}

hipError_t launch_grad(int M, int K, bool minors, const GradParams& p, int nblocks, size_t lds_bytes,
                       hipStream_t st) {
#define PQB_CASE(MM, KK)                                                                                     \
    if (M == MM && K == KK) {                                                                                \
        if (minors)                                                                                          \
            hipLaunchKernelGGL((grad_sweep_kernel<MM, KK, true>), dim3(nblocks), dim3(256), lds_bytes, st, p);  \
        else                                                                                                 \
            hipLaunchKernelGGL((grad_sweep_kernel<MM, KK, false>), dim3(nblocks), dim3(256), lds_bytes, st, p); \
        return hipGetLastError();                                                                            \
    }
    PQB_GRAD_CASES(PQB_CASE)
#undef PQB_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_grad_reduce(const GradReduceParams& p, hipStream_t st) {
    const int64_t chunks = (p.main + (p.tail ? 1 : 0) + 63) / 64;
    hipLaunchKernelGGL(grad_reduce_kernel, dim3((unsigned)(chunks * p.count)), dim3(256), 0, st, p);
    return hipGetLastError();
}

}  // namespace pqb
