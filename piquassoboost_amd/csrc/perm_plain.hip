// Plain Glynn / BBFG permanent: the binary reflected Gray-code sweep over the
// 2^(R-1) sign vectors delta in {+1,-1}^R with delta_0 = +1,
//
//     perm(A) = 2^-(R-1) * sum_delta (prod_i delta_i) prod_j sum_i delta_i a_ij,
//
// restated from BBFGPermanentCalculator.hpp:87-235 (sweep :179-221, seed
// :143-176, edge cases :90-95).  As in the reference, Gray bit b is tied to
// row R-1-b (:150-167, :186-191), a flip of row i changes every column sum by
// -/+ 2 a_ij (:205-216) and the running product is taken over all columns.
//
// MI355X layout of the sweep (one permanent per launch):
//
// * The R-1 sign bits are split into K walk bits (rows R-K .. R-1), B <= 6
//   lane bits (rows R-K-B .. R-K-1) and H group bits (rows 1 .. R-1-K-B).
//   Every lane of a wave walks the SAME K-bit reflected Gray sequence from
//   delta = +1, so the flipped row and its direction are wave-uniform: the
//   row of +/-2A is fetched with scalar loads and used as an SGPR operand,
//   and the Gray bookkeeping (ctz, parity) runs on the SALU.
// * Each wave owns a contiguous run of groups (runs differ by at most one
//   group, so any number of resident waves is balanced).  It seeds once, like
//   the reference's per-chunk seed (BBFGPermanentCalculator.hpp:143-167), and
//   moves from group g to g+1 by flipping the rows of the changed group bits,
//   which are again wave-uniform (about three row updates per group).
// * The product of the NC column sums is formed in complex FP64 as four
//   independent chains (for ILP) and the last multiply is fused into the
//   accumulation (4 FMAs), giving 6*NC-4 VALU instructions per Gray step
//   against 8*NC-4 algorithmic flops.
// * Terms of even and odd Gray index go to two accumulators (no sign work).
//   Chunk totals are combined per lane in FP64; everything above the lane
//   (wave, block, grid) is summed in double-double, and the last block to
//   finish reduces the per-block partials and writes the result.
//
// A second kernel evaluates the same sum exactly for integer matrices, with
// integer column sums and products/sums modulo 2^128 (two's complement
// wrap-around); the true sum 2^(R-1) perm is recovered exactly whenever
// |2^(R-1) perm| < 2^127, which the host checks with the row-sum bound.

#include "pqb_device.h"
#include "pqb_kernels.h"

namespace pqb {

// Product of cr[0..M) (complex), as up to four independent chains.
template <int M, int NC>
__device__ __forceinline__ void prod_first(const double (&cr)[NC], const double (&ci)[NC],
                                           double& pr, double& pi) {
    static_assert(M >= 1 && M <= NC, "bad product length");
    constexpr int NCH = (M >= 12) ? 4 : ((M >= 4) ? 2 : 1);
    double qr[NCH], qi[NCH];
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
        const int lo = q * M / NCH;
        const int hi = (q + 1) * M / NCH;
        qr[q] = cr[lo];
        qi[q] = ci[lo];
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

// acc += prod_j c_j  (last factor fused into the accumulation).
template <int NC>
__device__ __forceinline__ void add_term(const double (&cr)[NC], const double (&ci)[NC],
                                         double& ar, double& ai) {
    if constexpr (NC == 1) {
        ar += cr[0];
        ai += ci[0];
    } else {
        double pr, pi;
        prod_first<NC - 1, NC>(cr, ci, pr, pi);
        const double lr = cr[NC - 1], li = ci[NC - 1];
        ar = fma(pr, lr, ar);
        ar = fma(-pi, li, ar);
        ai = fma(pr, li, ai);
        ai = fma(pi, lr, ai);
    }
}

// colsum += row (row is wave-uniform: scalar loads + SGPR operands).
template <int NC>
__device__ __forceinline__ void add_row(double (&cr)[NC], double (&ci)[NC],
                                        const double2* __restrict__ row) {
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const double2 v = row[j];
        cr[j] += v.x;
        ci[j] += v.y;
        // keep at most 16 columns (64 SGPRs) of the row in flight
        if ((j & 15) == 15) __builtin_amdgcn_sched_barrier(0);
    }
}

// colsum += s * row, s = +/-1 per lane.
template <int NC>
__device__ __forceinline__ void add_row_signed(double (&cr)[NC], double (&ci)[NC],
                                               const double2* __restrict__ row, double s) {
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const double2 v = row[j];
        cr[j] = fma(s, v.x, cr[j]);
        ci[j] = fma(s, v.y, ci[j]);
    }
}

template <int NC>
// Up to 32 columns the sweep fits in 168 VGPRs: ask for three waves per SIMD
// so that scalar-load and FP64 latencies of one wave hide behind the others.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NC <= 32 ? 3 : 1)))
void plain_sweep_kernel(PlainParams p) {
    __shared__ ddc red[256 / kWave];
    // batch of equal-size matrices: blockIdx.y is the matrix, gridDim.x blocks each
    const unsigned task = blockIdx.y;
    ddc* const partials = p.partials + (size_t)task * gridDim.x;
    unsigned* const counter = p.counter + task;
    double* const out = p.out + 6 * (size_t)task;

    const int lane = threadIdx.x & (kWave - 1);
    // the wave index as a scalar: the run bounds, walk positions and row offsets
    // derived from it are then uniform for the compiler too (scalar loads, SGPR
    // operands), not only in fact
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x / kWave) +
                         __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;

    const int R = p.R;
    const int K = p.K;  // walk bits:  rows R-K .. R-1
    const int B = p.B;  // lane bits:  rows R-K-B .. R-K-1
    const int H = p.H;  // group bits: rows 1 .. R-1-K-B
    const unsigned rowb = NC * sizeof(double2);
    // byte offsets of the +2A / -2A rows: row r of +2A at r*rowb, of -2A at (R+r)*rowb
    const char* __restrict__ Mb = reinterpret_cast<const char*>(p.M2 + (size_t)task * p.m2_stride);
    const unsigned neg = (unsigned)R * rowb;
    const unsigned low_p = (unsigned)(R - 1) * rowb;  // +2A[R-1]
    const unsigned low_m = neg + low_p;               // -2A[R-1]
    const unsigned W = 1u << K;
    const bool lane_on = lane < (1 << B);

    // contiguous run [tau0, tau1) of the walk index tau = group * 2^K + t for
    // this wave, in units of (even, odd) step pairs, so that every resident
    // wave gets the same work to within one pair and no second round runs at
    // low occupancy; a run may start and end inside a group
    const unsigned step = K >= 1 ? 2u : 1u;
    const int64_t units = ((int64_t)1 << (H + K)) / step;
    const int64_t tau0 = (units * wave / nwaves) * step;
    const int64_t tau1 = (units * (wave + 1) / nwaves) * step;

    double tr = 0.0, ti = 0.0;  // lane total, lane sign applied at the end
    if (tau0 < tau1) {
        const int64_t g0 = tau0 >> K;
        unsigned t0 = (unsigned)(tau0 & (int64_t)(W - 1));  // even when K >= 1
        // ---- seed at (g0, t0): delta_0 = +1, walk row R-1-b is -1 iff bit b of
        // the Gray code t0 ^ (t0 >> 1), lane row R-1-K-b is -1 iff bit b of the
        // lane, group row R-1-K-B-h is -1 iff bit h of g0.  Built from the
        // +/-2A rows and halved (scaling by 2 is exact, so this equals seeding
        // from A as BBFGPermanentCalculator.hpp:143-167).
        double cr[NC], ci[NC];
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            const double2 v = reinterpret_cast<const double2*>(Mb)[j];
            cr[j] = v.x;
            ci[j] = v.y;
        }
        const unsigned gray0 = t0 ^ (t0 >> 1);
        for (int i = R - K; i < R; ++i) {
            const unsigned off = (((gray0 >> (R - 1 - i)) & 1u) ? neg : 0u) + (unsigned)i * rowb;
            add_row<NC>(cr, ci, reinterpret_cast<const double2*>(Mb + off));
        }
        for (int b = 0; b < B; ++b) {
            const int row = R - 1 - K - b;
            const double s = ((lane >> b) & 1) ? -1.0 : 1.0;
            add_row_signed<NC>(cr, ci, reinterpret_cast<const double2*>(Mb + (size_t)row * rowb), s);
        }
        for (int h = 0; h < H; ++h) {
            const int row = R - 1 - K - B - h;
            const unsigned off = (((g0 >> h) & 1) ? neg : 0u) + (unsigned)row * rowb;
            add_row<NC>(cr, ci, reinterpret_cast<const double2*>(Mb + off));
        }
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            cr[j] *= 0.5;
            ci[j] *= 0.5;
        }

        for (int64_t g = g0;;) {
            // this group's part of the run: t in [t0, t_stop)
            const int64_t gbase = g << K;
            // wave-uniform; taken through readfirstlane so that the loop compare is scalar
            const unsigned t_stop = (unsigned)__builtin_amdgcn_readfirstlane(
                (int)((tau1 - gbase >= (int64_t)W) ? W : (unsigned)(tau1 - gbase)));
            double pr = 0.0, pi = 0.0, mr = 0.0, mi = 0.0;
            // t = t0: even Gray index
            add_term<NC>(cr, ci, pr, pi);
            if (K >= 1) {
                // t = t0 + 1: walk bit 0 (row R-1) goes to delta = -1 iff bit 1 of t0 is 0
                add_row<NC>(cr, ci, reinterpret_cast<const double2*>(Mb + (((t0 >> 1) & 1u) ? low_p : low_m)));
                add_term<NC>(cr, ci, mr, mi);
                // one (even, odd) step pair: the even step flips walk row ctz(t),
                // the odd step row R-1
                auto pair = [&](unsigned t) __attribute__((always_inline)) {
                    const unsigned b = __builtin_ctz(t);
                    const unsigned nb = ((t >> b) ^ (t >> (b + 1))) & 1u;
                    const unsigned off = (nb ? neg : 0u) + (unsigned)(R - 1 - (int)b) * rowb;
                    add_row<NC>(cr, ci, reinterpret_cast<const double2*>(Mb + off));
                    add_term<NC>(cr, ci, pr, pi);
                    const unsigned off0 = ((t >> 1) & 1u) ? low_p : low_m;
                    add_row<NC>(cr, ci, reinterpret_cast<const double2*>(Mb + off0));
                    add_term<NC>(cr, ci, mr, mi);
                };
                // Two pairs per iteration: the second pair's even row is then issued
                // during the first pair's odd product instead of after the loop's
                // back-edge (n=30: 2.876 against 2.909 ms, same sums).  At 32 columns
                // the two-pair body does not fit 168 VGPRs.
                constexpr unsigned kPairs = NC == 32 ? 1u : 2u;
                unsigned t = t0 + 2;
                if constexpr (kPairs == 2u) {
                    for (; t + 2 < t_stop; t += 4) {
                        pair(t);
                        pair(t + 2);
                    }
                }
                for (; t < t_stop; t += 2) pair(t);
            }
            // sign of the group's first term relative to the lane: parity of g
            const double gs = (__popcll((unsigned long long)g) & 1) ? -1.0 : 1.0;
            tr = fma(gs, pr - mr, tr);
            ti = fma(gs, pi - mi, ti);
            if (gbase + (int64_t)t_stop >= tau1) break;
            ++g;
            t0 = 0;
            // to the next group's seed: the walk ended at t = 2^K - 1, Gray code
            // 10..0 (row R-K at delta = -1), restore it; then flip the group rows
            // whose bits differ between g-1 and g (wave-uniform).
            if (K >= 1) add_row<NC>(cr, ci, reinterpret_cast<const double2*>(Mb + (size_t)(R - K) * rowb));
            const int64_t diff = (g - 1) ^ g;
            for (int h = 0; h < H; ++h) {
                if ((diff >> h) & 1) {
                    const int row = R - 1 - K - B - h;
                    const unsigned off = (((g >> h) & 1) ? neg : 0u) + (unsigned)row * rowb;
                    add_row<NC>(cr, ci, reinterpret_cast<const double2*>(Mb + off));
                }
            }
        }
    }
    const double ls = (__popc((unsigned)lane) & 1) ? -1.0 : 1.0;
    tr = lane_on ? ls * tr : 0.0;
    ti = lane_on ? ls * ti : 0.0;

    ddc v = block_reduce(ddc_from(tr, ti), red);

    // ---- grid reduction: last block sums the per-block partials ----
    __shared__ bool is_last;
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = v;
        __threadfence();
        const unsigned done = atomicAdd(counter, 1u);
        is_last = (done == gridDim.x - 1);
    }
    __syncthreads();
    if (is_last) {
        __threadfence();
        ddc s = {0.0, 0.0, 0.0, 0.0};
        for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
            const ddc q = partials[b];
            ddc_add(s, q);
        }
        s = block_reduce(s, red);
        if (threadIdx.x == 0) {
            s = ddc_scale(s, p.scale);
            out[0] = s.hr + s.lr;
            out[1] = s.hi + s.li;
            out[2] = s.hr;
            out[3] = s.lr;
            out[4] = s.hi;
            out[5] = s.li;
            *counter = 0u;  // leave the counter ready for the next launch
        }
    }
}

// M2 = [+2A ; -2A] for each of `count` packed matrices of n elements.
__global__ void plain_prep_kernel(const double2* __restrict__ A, double2* __restrict__ M2, int n, int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) {
        const int64_t t = i / n, e = i - t * n;
        const double2 v = A[i];
        M2[2 * t * n + e] = make_double2(2.0 * v.x, 2.0 * v.y);
        M2[2 * t * n + n + e] = make_double2(-2.0 * v.x, -2.0 * v.y);
    }
}

hipError_t launch_plain_prep(const double2* A, double2* M2, int n, int64_t count, hipStream_t st) {
    const int bd = 256;
    const int64_t total = (int64_t)n * count;
    hipLaunchKernelGGL(plain_prep_kernel, dim3((unsigned)((total + bd - 1) / bd)), dim3(bd), 0, st, A, M2, n,
                       total);
    return hipGetLastError();
}

// ------------------------------------------------------------- exact int --

typedef unsigned __int128 u128;

// a * q modulo 2^128 for a signed 64-bit q (two's complement wrap-around):
// one 64 x 64 -> 128 product and one 64-bit low product of |q|, negated for q < 0
__device__ __forceinline__ u128 mul_u128_s64(u128 a, int64_t q) {
    const bool neg = q < 0;
    const unsigned long long m = neg ? 0ull - (unsigned long long)q : (unsigned long long)q;
    const u128 lo = (u128)(unsigned long long)a * m;
    const unsigned long long hi = (unsigned long long)(a >> 64) * m + (unsigned long long)(lo >> 64);
    const u128 r = ((u128)hi << 64) | (unsigned long long)lo;
    return neg ? (u128)0 - r : r;
}

template <int NC, int G>
__global__ __launch_bounds__(256) void plain_exact_kernel(ExactParams p) {
    __shared__ u128 red[256 / kWave];

    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = g < p.nlanes;
    const int64_t gl = active ? g : 0;
    const int R = p.R, K = p.K, H = p.H;
    const int* __restrict__ A = p.A;  // R x NC integers

    int c[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) c[j] = A[j];
    for (int i = R - K; i < R; ++i) {
#pragma unroll
        for (int j = 0; j < NC; ++j) c[j] += A[i * NC + j];
    }
    for (int h = 0; h < H; ++h) {
        const int row = R - 1 - K - h;
        const int s = ((gl >> h) & 1) ? -1 : 1;
#pragma unroll
        for (int j = 0; j < NC; ++j) c[j] += s * A[row * NC + j];
    }
    const bool lane_neg = __popcll((unsigned long long)gl) & 1;

    u128 acc = 0;
    const int64_t W = (int64_t)1 << K;
    for (int64_t t = 0; t < W; ++t) {
        if (t > 0) {
            const int b = __builtin_ctzll((unsigned long long)t);
            const int nb = (int)(((t >> b) ^ (t >> (b + 1))) & 1);
            const int row = R - 1 - b;
            const int s = nb ? -2 : 2;
#pragma unroll
            for (int j = 0; j < NC; ++j) c[j] += s * A[row * NC + j];
        }
        // product of the column sums: G of them at a time exactly in int64 (the
        // host bounds every |c_j| by the largest column abs sum, whose G-th power
        // is below 2^63), each partial product folded into the 128-bit product
        u128 pr = 0;
#pragma unroll
        for (int j0 = 0; j0 < NC; j0 += G) {
            int64_t q = c[j0];
#pragma unroll
            for (int j = j0 + 1; j < (j0 + G < NC ? j0 + G : NC); ++j) q *= (int64_t)c[j];
            pr = (j0 == 0) ? (u128)(__int128)q : mul_u128_s64(pr, q);
        }
        if (t & 1) acc -= pr; else acc += pr;
    }
    if (lane_neg) acc = (u128)0 - acc;
    if (!active) acc = 0;

    // block sum (wrap-around arithmetic is associative: order does not matter)
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    unsigned long long lo = (unsigned long long)acc, hi = (unsigned long long)(acc >> 64);
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1) {
        const unsigned long long olo = __shfl_xor(lo, m, kWave);
        const unsigned long long ohi = __shfl_xor(hi, m, kWave);
        const u128 s = (((u128)hi << 64) | lo) + (((u128)ohi << 64) | olo);
        lo = (unsigned long long)s;
        hi = (unsigned long long)(s >> 64);
    }
    if (lane == 0) red[wid] = ((u128)hi << 64) | lo;
    __syncthreads();
    __shared__ bool is_last;
    if (threadIdx.x == 0) {
        u128 s = 0;
        const int nw = (blockDim.x + kWave - 1) / kWave;
        for (int w = 0; w < nw; ++w) s += red[w];
        p.partials[2 * blockIdx.x] = (unsigned long long)s;
        p.partials[2 * blockIdx.x + 1] = (unsigned long long)(s >> 64);
        __threadfence();
        const unsigned done = atomicAdd(p.counter, 1u);
        is_last = (done == gridDim.x - 1);
    }
    __syncthreads();
    if (is_last && threadIdx.x == 0) {
        __threadfence();
        u128 s = 0;
        for (unsigned b = 0; b < gridDim.x; ++b)
            s += ((u128)p.partials[2 * b + 1] << 64) | p.partials[2 * b];
        p.out[0] = (unsigned long long)s;
        p.out[1] = (unsigned long long)(s >> 64);
        *p.counter = 0u;
    }
}

// ------------------------------------------------------------ dispatch --

template <int NC>
static hipError_t launch_plain_nc(const PlainParams& p, int nblocks, int bdim, int ntasks, hipStream_t st) {
    hipLaunchKernelGGL(plain_sweep_kernel<NC>, dim3(nblocks, ntasks), dim3(bdim), 0, st, p);
    return hipGetLastError();
}

template <int NC>
static hipError_t launch_exact_nc(const ExactParams& p, int nblocks, int bdim, hipStream_t st) {
    switch (p.group) {
        case 13: hipLaunchKernelGGL((plain_exact_kernel<NC, 13>), dim3(nblocks), dim3(bdim), 0, st, p); break;
        case 6: hipLaunchKernelGGL((plain_exact_kernel<NC, 6>), dim3(nblocks), dim3(bdim), 0, st, p); break;
        case 2: hipLaunchKernelGGL((plain_exact_kernel<NC, 2>), dim3(nblocks), dim3(bdim), 0, st, p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int NC>
static const void* plain_fn_nc() {
    return reinterpret_cast<const void*>(&plain_sweep_kernel<NC>);
}

#define PQB_NC_CASES(X) \
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) \
    X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30)    \
    X(31) X(32) X(33) X(34) X(35) X(36) X(37) X(38) X(39) X(40)

hipError_t launch_plain(int nc, const PlainParams& p, int nblocks, int bdim, int ntasks, hipStream_t st) {
    switch (nc) {
#define PQB_CASE(N) \
    case N: return launch_plain_nc<N>(p, nblocks, bdim, ntasks, st);
        PQB_NC_CASES(PQB_CASE)
#undef PQB_CASE
        default: return hipErrorInvalidValue;
    }
}

const void* plain_kernel_fn(int nc) {
    switch (nc) {
#define PQB_CASE(N) \
    case N: return plain_fn_nc<N>();
        PQB_NC_CASES(PQB_CASE)
#undef PQB_CASE
        default: return nullptr;
    }
}

hipError_t launch_plain_exact(int nc, const ExactParams& p, int nblocks, int bdim, hipStream_t st) {
    switch (nc) {
#define PQB_CASE(N) \
    case N: return launch_exact_nc<N>(p, nblocks, bdim, st);
        PQB_NC_CASES(PQB_CASE)
#undef PQB_CASE
        default: return hipErrorInvalidValue;
    }
}

}  // namespace pqb
