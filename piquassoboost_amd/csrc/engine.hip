// Host engine behind the C ABI declared in include/pqb_perm.h.
//
// Responsibilities: argument checks and the reference's edge-case rules,
// choice of the sweep decomposition (chunk bits K, chunk-index bits H, lane
// count), staging host matrices into a per-process device workspace, kernel
// launches on the engine's stream, and error reporting (status code +
// thread-local message; no C++ exception crosses the ABI).
//
// The host-pointer entry points are synchronous (as the reference calculators
// are: the GIL is held and TBB blocks, GlynnPermanentCalculator_Wrapper.hpp:
// 200-377) and thread safe: each call leases one of kEngineContexts execution
// contexts (stream, device workspace, pinned staging, batch buffers and graphs)
// for its duration, so calls from different threads run on different streams
// at the same time.  The device-pointer entry points take the caller's stream
// and workspace and touch no context, only the shared device state (binding,
// occupancy of the kernels).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "pqb_perm.h"
#include "pqb_device.h"
#include "pqb_kernels.h"
#include "pqb_engine.h"

namespace pqb {

thread_local std::string t_err;

int fail(int code, const std::string& msg) {
    t_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(PQB_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

int drain_fail(Engine& E, hipError_t e, const char* what) {
    if (E.stream) (void)hipStreamSynchronize(E.stream);
    return hip_fail(e, what);
}

DeviceState g_dev;
Engine g_ctx[kEngineContexts];

// ----------------------------------------------------------- device state --

// Binding under g_dev.mu (held): device `dev` (< 0: 0) becomes the engine's.
static int set_device_locked(int dev) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return fail(PQB_ERR_DEVICE, "no HIP device available (the permanent engine runs on the GPU only)");
    if (dev < 0) dev = 0;
    if (dev >= n) return fail(PQB_ERR_ARG, "device index out of range");
    e = hipSetDevice(dev);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    g_dev.cus = prop.multiProcessorCount;
    for (int i = 0; i <= PQB_MAX_COLS; ++i) g_dev.blocks_per_cu[i] = 0;
    g_dev.device = dev;
    return PQB_OK;
}

int bind_device(int dev, int* device, int* cus) {
    {
        std::lock_guard<std::mutex> lk(g_dev.mu);
        const bool keep = g_dev.device >= 0 && (dev < 0 || dev == g_dev.device);
        // the first binding: no context has a stream or buffers yet (they are
        // created only after a binding), so nothing needs releasing
        const int st = keep ? PQB_OK : (g_dev.device < 0 ? set_device_locked(dev) : -1);
        if (st >= 0) {
            if (st == PQB_OK) {
                if (device) *device = g_dev.device;
                if (cus) *cus = g_dev.cus;
            }
            return st;
        }
    }
    // a change of device (pqb_init, never called from inside a context's lease):
    // every context is locked -- in index order, before g_dev.mu, the order every
    // caller follows -- and releases its stream and buffers first
    std::unique_lock<std::mutex> held[kEngineContexts];
    for (int k = 0; k < kEngineContexts; ++k) held[k] = std::unique_lock<std::mutex>(g_ctx[k].mu);
    std::lock_guard<std::mutex> lk(g_dev.mu);
    if (!(g_dev.device >= 0 && (dev < 0 || dev == g_dev.device))) {
        for (int k = 0; k < kEngineContexts; ++k) g_ctx[k].release_locked();
        const int st = set_device_locked(dev);
        if (st) return st;
    }
    if (device) *device = g_dev.device;
    if (cus) *cus = g_dev.cus;
    return PQB_OK;
}

int64_t resident_waves(int nc) {
    if (nc < 1 || nc > PQB_MAX_COLS) return 0;
    std::lock_guard<std::mutex> lk(g_dev.mu);
    if (g_dev.blocks_per_cu[nc] == 0) {
        int nb = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, plain_kernel_fn(nc), kBlock, 0);
        if (e != hipSuccess || nb <= 0) nb = 1;
        g_dev.blocks_per_cu[nc] = nb;
    }
    return (int64_t)g_dev.blocks_per_cu[nc] * (kBlock / kWave) * g_dev.cus;
}

// ---------------------------------------------------------------- contexts --

void release_batch_locked(Engine& E);  // batch.hip

void Engine::release_locked() {
    if (stream) (void)hipStreamSynchronize(stream);
    release_batch_locked(*this);
    if (dbuf) (void)hipFree(dbuf);
    if (hbuf) (void)hipHostFree(hbuf);
    dbuf = nullptr;
    hbuf = nullptr;
    dbytes = hbytes = 0;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    device = -1;
}

int Engine::ensure_init_locked() {
    int dev = -1, ncu = 0;
    int st = bind_device(-1, &dev, &ncu);
    if (st) return st;
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if (device == dev && stream) return PQB_OK;
    release_locked();
    e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        stream = nullptr;
        return hip_fail(e, "hipStreamCreate");
    }
    device = dev;
    cus = ncu;
    return PQB_OK;
}

int Engine::ensure_buffers_locked(size_t dneed, size_t hneed) {
    const size_t old_d = dbytes, old_h = hbytes;
    if (dneed > dbytes) {
        if (dbuf) {
            (void)hipStreamSynchronize(stream);
            (void)hipFree(dbuf);
            dbuf = nullptr;
            dbytes = 0;
        }
        // geometric growth: a caller whose sizes creep up (a sampler going photon by
        // photon) reallocates a few times, not at every step
        const size_t sz = std::max<size_t>(std::max<size_t>(dneed, 2 * old_d), (size_t)1 << 20);
        hipError_t e = hipMalloc(&dbuf, sz);
        if (e != hipSuccess) {
            dbuf = nullptr;
            return hip_fail(e, "hipMalloc(workspace)");
        }
        // the grid-reduction counter at offset 0 must start at zero
        e = hipMemsetAsync(dbuf, 0, kCounterBytes, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "hipMemset(workspace)");
        dbytes = sz;
    }
    if (hneed > hbytes) {
        if (hbuf) {
            (void)hipStreamSynchronize(stream);
            (void)hipHostFree(hbuf);
            hbuf = nullptr;
            hbytes = 0;
        }
        const size_t sz = std::max<size_t>(std::max<size_t>(hneed, 2 * old_h), (size_t)1 << 20);
        hipError_t e = hipHostMalloc(&hbuf, sz, hipHostMallocDefault);
        if (e != hipSuccess) {
            hbuf = nullptr;
            return hip_fail(e, "hipHostMalloc(staging)");
        }
        hbytes = sz;
    }
    return PQB_OK;
}

// A thread keeps using the context it was first given when that one is free,
// so a single-threaded caller always runs on one stream and one set of buffers.
static int preferred_context() {
    static std::atomic<unsigned> next{0};
    thread_local int pref = (int)(next.fetch_add(1u) % (unsigned)kEngineContexts);
    return pref;
}

Lease::Lease() {
    const int pref = preferred_context();
    for (int k = 0; k < kEngineContexts && !e_; ++k) {
        Engine& c = g_ctx[(pref + k) % kEngineContexts];
        std::unique_lock<std::mutex> l(c.mu, std::try_to_lock);
        if (l.owns_lock()) {
            lk_ = std::move(l);
            e_ = &c;
        }
    }
    if (!e_) {
        e_ = &g_ctx[pref];
        lk_ = std::unique_lock<std::mutex>(e_->mu);
    }
    st_ = e_->ensure_init_locked();
}

Lease::Lease(Engine& e) : e_(&e), lk_(e.mu) { st_ = e_->ensure_init_locked(); }

// --------------------------------------------------------------- planning --

static int floor_log2(int64_t v) {
    int r = -1;
    while (v > 0) {
        v >>= 1;
        ++r;
    }
    return r;
}

static int ceil_log2(int64_t v) {
    int r = 0;
    while (((int64_t)1 << r) < v) ++r;
    return r;
}

// Split of the R-1 sign bits into walk (K), lane (B) and group (H) bits and
// the number of waves.  The kernel gives every wave one contiguous run of the
// walk index (group * 2^K + t) in step pairs, so runs are equal to within one
// pair and may start and end inside a group: the launch is exactly the
// resident waves, one round at full occupancy.  (Whole groups per wave would
// need a power-of-two wave count: 4096 waves of two groups each at n=30 left a
// quarter of the work to a second round at one wave per SIMD.)  The walk is
// long enough that a run spans a few groups; small sweeps take the shortest
// walk that still amortises the seed.
PlainPlan make_plain_plan(int R, int NC, int64_t waves_resident) {
    (void)NC;
    PlainPlan pl;
    const int T = R - 1;
    const int64_t wr = std::max<int64_t>(waves_resident, 1);
    int K = T - 6 - (ceil_log2(wr) + 1);
    K = std::max(K, std::min(T, 3));
    K = std::min(K, std::min(T, 14));
    pl.K = K;
    pl.B = std::min(6, T - K);
    pl.H = T - K - pl.B;
    const int64_t units = ((int64_t)1 << (pl.H + pl.K)) / (pl.K >= 1 ? 2 : 1);  // step pairs
    int64_t waves = std::min<int64_t>(units, wr);
    if (waves >= 4) {
        waves -= waves % 4;
        pl.block = kBlock;
        pl.blocks = (int)(waves / 4);
    } else {
        pl.block = (int)(64 * waves);
        pl.blocks = 1;
    }
    pl.waves = (int64_t)pl.blocks * pl.block / kWave;
    return pl;
}

// Integer route eligibility.  Fills `ai` (packed R x C int32) when the matrix
// is integer valued and real, and returns log2 of an upper bound on the sweep's
// sum S = sum_delta (prod delta) prod_j (sum_i delta_i a_ij) (or -inf when S is
// provably 0).  For every shape, |S| <= 2^(R-1) prod_j c_j (c_j = column abs
// sums: each of the 2^(R-1) terms is bounded by that product; a zero column
// makes every term 0).  Only for R == C is S = 2^(R-1) perm(A), so only there
// do the row-sum bound perm(|A|) <= prod_i r_i and, for entries in {-1, 0, 1},
// Bregman's bound perm(|A|) <= prod_i (r_i!)^(1/r_i) (also applied to the
// transpose) hold; for R != C the sum is not a permanent of A (R = C + 1 gives
// 2^(R-1) perm(A without row 0), R < C counts maps with odd row multiplicities)
// and a row bound can be arbitrarily wrong.  The sweep sums modulo 2^128, so
// the result is exact when the bound is below 2^127; column sums must also fit
// in int32.
static bool integer_matrix(const double* A, int64_t R, int64_t C, int64_t ld, std::vector<int>& ai,
                           long double* log2_bound) {
    ai.assign((size_t)(R * C), 0);
    std::vector<double> rowabs((size_t)R, 0.0), colabs((size_t)C, 0.0);
    bool unit = true;
    for (int64_t i = 0; i < R; ++i) {
        for (int64_t j = 0; j < C; ++j) {
            const double re = A[2 * (i * ld + j)];
            const double im = A[2 * (i * ld + j) + 1];
            if (im != 0.0 || !(std::fabs(re) < 1073741824.0) || re != std::floor(re)) return false;
            ai[(size_t)(i * C + j)] = (int)re;
            const double a = std::fabs(re);
            if (a > 1.0) unit = false;
            rowabs[(size_t)i] += a;
            colabs[(size_t)j] += a;
        }
    }
    for (int64_t j = 0; j < C; ++j)
        if (colabs[(size_t)j] >= 1073741824.0) return false;  // int32 column sums
    auto lg_prod = [](const std::vector<double>& v, bool bregman) -> long double {
        long double s = 0.0L;
        for (double x : v) {
            if (x == 0.0) return -INFINITY;
            s += bregman ? std::lgamma((long double)x + 1.0L) / (long double)x / std::log(2.0L)
                         : std::log2((long double)x);
        }
        return s;
    };
    long double lb = lg_prod(colabs, false);
    if (R == C) {
        lb = std::min(lb, lg_prod(rowabs, false));
        if (unit) lb = std::min(lb, std::min(lg_prod(rowabs, true), lg_prod(colabs, true)));
    }
    *log2_bound = lb + (long double)(R - 1);
    return true;
}

static bool exact_fits(long double log2_bound) { return log2_bound < 126.5L; }

typedef __int128 i128;

static double i128_to_double(i128 v) { return (double)v; }

// Exact route: sum_delta (prod delta) prod_j c_j  (mod 2^128) on the GPU,
// then divide by 2^(R-1).  The caller holds the lease of E.
static int run_exact_locked(Engine& E, const std::vector<int>& ai, int R, int C, i128* result) {
    const int T = R - 1;
    int64_t lr = E.resident_waves_locked(C) * kWave;
    const int lg = floor_log2(std::max<int64_t>(lr, kWave));
    const int H = std::min(T, std::max(0, lg));
    const int K = T - H;
    if (K > 30) return fail(PQB_ERR_SIZE, "exact sweep too large");
    const int64_t nlanes = (int64_t)1 << H;
    const int block = (int)std::min<int64_t>(kBlock, ((nlanes + kWave - 1) / kWave) * kWave);
    const int64_t blocks = (nlanes + block - 1) / block;

    const size_t a_bytes = ai.size() * sizeof(int);
    const size_t part_off = kCounterBytes;
    const size_t part_bytes = (size_t)blocks * 2 * sizeof(unsigned long long);
    const size_t out_off = part_off + ((part_bytes + 255) / 256) * 256;
    const size_t a_off = out_off + 256;
    int st = E.ensure_buffers_locked(a_off + a_bytes, a_bytes + 256);
    if (st) return st;
    char* d = static_cast<char*>(E.dbuf);
    std::memcpy(E.hbuf, ai.data(), a_bytes);
    hipError_t e = hipMemcpyAsync(d + a_off, E.hbuf, a_bytes, hipMemcpyHostToDevice, E.stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync");
    // every running column sum is bounded by its column's abs sum: 28^13, 1400^6 and
    // (2^30)^2 are below 2^63 (integer_matrix admits column abs sums below 2^30)
    int64_t cmax = 0;
    for (int j = 0; j < C; ++j) {
        int64_t s = 0;
        for (int i = 0; i < R; ++i) s += std::llabs((long long)ai[(size_t)i * C + j]);
        cmax = std::max(cmax, s);
    }
    ExactParams p;
    p.A = reinterpret_cast<const int*>(d + a_off);
    p.R = R;
    p.K = K;
    p.H = H;
    p.group = cmax <= 28 ? 13 : (cmax <= 1400 ? 6 : 2);
    p.nlanes = nlanes;
    p.partials = reinterpret_cast<unsigned long long*>(d + part_off);
    p.counter = reinterpret_cast<unsigned*>(d);
    p.out = reinterpret_cast<unsigned long long*>(d + out_off);
    e = launch_plain_exact(C, p, (int)blocks, block, E.stream);
    if (e != hipSuccess) return drain_fail(E, e, "exact sweep launch");
    unsigned long long* hout = reinterpret_cast<unsigned long long*>(static_cast<char*>(E.hbuf) + a_bytes);
    e = hipMemcpyAsync(hout, d + out_off, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, E.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(E.stream);
    if (e != hipSuccess) return hip_fail(e, "exact sweep");
    const unsigned __int128 u = ((unsigned __int128)hout[1] << 64) | hout[0];
    const i128 sum = (i128)u;
    const i128 mask = (((i128)1) << T) - 1;
    if ((sum & mask) != 0) return fail(PQB_ERR_EXACT, "exact sweep: sum not divisible by 2^(rows-1)");
    *result = sum >> T;  // arithmetic shift: exact division
    return PQB_OK;
}

// FP64 route.  The caller holds the lease of E.
static int run_plain_locked(Engine& E, const double* A, int R, int C, int64_t ld, double* out) {
    const PlainPlan pl = make_plain_plan(R, C, E.resident_waves_locked(C));
    const size_t m2_elems = (size_t)2 * R * C;
    const size_t m2_bytes = m2_elems * sizeof(double2);
    const size_t part_off = kCounterBytes;
    const size_t part_bytes = (size_t)pl.blocks * sizeof(ddc);
    const size_t out_off = part_off + ((part_bytes + 255) / 256) * 256;
    const size_t m2_off = out_off + 256;
    int st = E.ensure_buffers_locked(m2_off + m2_bytes, m2_bytes + 256);
    if (st) return st;
    double* h = static_cast<double*>(E.hbuf);
    for (int i = 0; i < R; ++i) {
        for (int j = 0; j < C; ++j) {
            const double re = A[2 * (i * ld + j)], im = A[2 * (i * ld + j) + 1];
            const size_t k = (size_t)i * C + j;
            h[2 * k] = 2.0 * re;
            h[2 * k + 1] = 2.0 * im;
            h[2 * (k + (size_t)R * C)] = -2.0 * re;
            h[2 * (k + (size_t)R * C) + 1] = -2.0 * im;
        }
    }
    char* d = static_cast<char*>(E.dbuf);
    hipError_t e = hipMemcpyAsync(d + m2_off, h, m2_bytes, hipMemcpyHostToDevice, E.stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync");
    PlainParams p;
    p.M2 = reinterpret_cast<const double2*>(d + m2_off);
    p.m2_stride = (int64_t)m2_elems;
    p.R = R;
    p.K = pl.K;
    p.B = pl.B;
    p.H = pl.H;
    p.partials = reinterpret_cast<ddc*>(d + part_off);
    p.counter = reinterpret_cast<unsigned*>(d);
    p.out = reinterpret_cast<double*>(d + out_off);
    p.scale = std::ldexp(1.0, -(R - 1));
    e = launch_plain(C, p, pl.blocks, pl.block, 1, E.stream);
    if (e != hipSuccess) return drain_fail(E, e, "plain sweep launch");
    double* hout = h + 2 * m2_elems;
    e = hipMemcpyAsync(hout, d + out_off, 6 * sizeof(double), hipMemcpyDeviceToHost, E.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(E.stream);
    if (e != hipSuccess) return hip_fail(e, "plain sweep");
    out[0] = hout[0];
    out[1] = hout[1];
    return PQB_OK;
}

// Shape rules of the reference calculators.  Returns 1 and fills out when
// the result is decided without a sweep.
static int plain_edge(int algo, int64_t rows, int64_t cols, double* out) {
    if (algo == PQB_ALGO_SIMPLE) {
        // GlynnPermanentCalculatorSimple.hpp:65-70
        if (rows == 0 || cols == 0) { out[0] = 1.0; out[1] = 0.0; return 1; }
        if (rows != cols) {
            std::fprintf(stderr, "Matrix is not rectangular! Returning 0.\n");
            out[0] = 0.0; out[1] = 0.0; return 1;
        }
        return 0;
    }
    // BBFGPermanentCalculator.cpp:92-95, .hpp:90-95; GlynnPermanentCalculator.hpp:50-53
    if (rows == 0 || cols == 0) { out[0] = 1.0; out[1] = 0.0; return 1; }
    if (rows >= cols + 2) { out[0] = 0.0; out[1] = 0.0; return 1; }
    return 0;
}

int plain_entry(const double* A, int64_t rows, int64_t cols, int64_t ld, int algo, int precision,
                double* out, int64_t* words) {
    if (!out) return fail(PQB_ERR_ARG, "output pointer is NULL");
    if (rows < 0 || cols < 0) return fail(PQB_ERR_ARG, "negative matrix dimension");
    if (algo < PQB_ALGO_BBFG || algo > PQB_ALGO_SIMPLE) return fail(PQB_ERR_ARG, "unknown algorithm code");
    if (precision < PQB_PREC_DOUBLE || precision > PQB_PREC_INF)
        return fail(PQB_ERR_ARG, "unknown precision code");
    if (plain_edge(algo, rows, cols, out)) {
        if (words) {
            const i128 v = (i128)out[0];
            words[0] = (int64_t)(unsigned long long)v;
            words[1] = (int64_t)(v >> 64);
        }
        return PQB_OK;
    }
    if (!A) return fail(PQB_ERR_ARG, "matrix pointer is NULL");
    if (ld < cols) return fail(PQB_ERR_ARG, "leading dimension smaller than the number of columns");
    if (cols > PQB_MAX_COLS)
        return fail(PQB_ERR_SIZE, "matrix has more than " + std::to_string(PQB_MAX_COLS) + " columns");
    const int R = (int)rows, C = (int)cols;

    std::vector<int> ai;
    long double lb = 0.0L;
    const bool is_int = precision != PQB_PREC_FP64 && integer_matrix(A, rows, cols, ld, ai, &lb);
    const bool exact_ok = is_int && exact_fits(lb);
    if (precision == PQB_PREC_EXACT && !exact_ok)
        return fail(PQB_ERR_EXACT, is_int ? "exact mode: the permanent may exceed the 128-bit range"
                                          : "exact mode needs a real integer-valued matrix");

    Lease L;
    int st = L.status();
    if (st) return st;
    Engine& E = *L;
    if (exact_ok) {
        i128 v = 0;
        st = run_exact_locked(E, ai, R, C, &v);
        if (st) return st;
        out[0] = i128_to_double(v);
        out[1] = 0.0;
        if (words) {
            words[0] = (int64_t)(unsigned long long)v;
            words[1] = (int64_t)(v >> 64);
        }
        return PQB_OK;
    }
    if (precision == PQB_PREC_INF) return run_plain_dd_locked(E, A, R, C, ld, out);
    return run_plain_locked(E, A, R, C, ld, out);
}

static size_t round256(size_t b) { return (b + 255) / 256 * 256; }

// ------------------------------------------------------------- list entry --
//
// Plain permanents of a list of host matrices (any mix of shapes).  Each
// matrix gets the value pqb_perm gives it; the work is regrouped:
//  * exact-route integer matrices: one exact sweep each (as per call);
//  * double-double route: every matrix planned as a single call would plan it,
//    all of them in one batched launch;
//  * FP64 route: one launch per distinct shape, one grid row (blockIdx.y) per
//    matrix, each row with the single-call launch plan -- so every value is
//    bit-identical to its pqb_perm call -- with one upload of all +-2A tables,
//    one download of all results and one synchronisation.

struct ListItem {
    int64_t idx;
    int R, C;
};

static int list_fp64_locked(Engine& E, const double* const* A, const int64_t* ld, std::vector<ListItem>& items,
                            double* out) {
    if (items.empty()) return PQB_OK;
    std::stable_sort(items.begin(), items.end(), [](const ListItem& a, const ListItem& b) {
        return a.R != b.R ? a.R < b.R : a.C < b.C;
    });
    struct Group {
        size_t first, count;
        PlainPlan pl;
        size_t part_off, m2_off;  // device byte offsets of the group's partials and tables
    };
    std::vector<Group> groups;
    for (size_t k = 0; k < items.size();) {
        size_t e = k;
        while (e < items.size() && items[e].R == items[k].R && items[e].C == items[k].C) ++e;
        Group g;
        g.first = k;
        g.count = e - k;
        g.pl = make_plain_plan(items[k].R, items[k].C, E.resident_waves_locked(items[k].C));
        groups.push_back(g);
        k = e;
    }
    const size_t n = items.size();
    const size_t cnt_off = kCounterBytes;  // offset 0 stays the single-call counter
    size_t off = cnt_off + round256(n * sizeof(unsigned));
    const size_t out_off = off;
    off += round256(n * 6 * sizeof(double));
    for (Group& g : groups) {
        g.part_off = off;
        off += round256(g.count * (size_t)g.pl.blocks * sizeof(ddc));
    }
    size_t m2_total = 0;
    for (Group& g : groups) {
        g.m2_off = off + m2_total;
        m2_total += g.count * (size_t)2 * items[g.first].R * items[g.first].C * sizeof(double2);
    }
    const size_t m2_base = off;
    int st = E.ensure_buffers_locked(m2_base + m2_total, m2_total + n * 6 * sizeof(double));
    if (st) return st;
    double* h = static_cast<double*>(E.hbuf);
    for (const Group& g : groups) {
        const int R = items[g.first].R, C = items[g.first].C;
        const size_t rc = (size_t)R * C;
        double* hg = h + (g.m2_off - m2_base) / sizeof(double);
        for (size_t k = 0; k < g.count; ++k) {
            const ListItem& it = items[g.first + k];
            const double* a = A[it.idx];
            const int64_t l = ld[it.idx];
            double* t = hg + 4 * rc * k;  // +2A then -2A, complex interleaved
            for (int i = 0; i < R; ++i)
                for (int j = 0; j < C; ++j) {
                    const double re = a[2 * (i * l + j)], im = a[2 * (i * l + j) + 1];
                    const size_t q = (size_t)i * C + j;
                    t[2 * q] = 2.0 * re;
                    t[2 * q + 1] = 2.0 * im;
                    t[2 * (q + rc)] = -2.0 * re;
                    t[2 * (q + rc) + 1] = -2.0 * im;
                }
        }
    }
    char* d = static_cast<char*>(E.dbuf);
    hipError_t e = hipMemsetAsync(d + cnt_off, 0, n * sizeof(unsigned), E.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d + m2_base, h, m2_total, hipMemcpyHostToDevice, E.stream);
    if (e != hipSuccess) return drain_fail(E, e, "list upload");
    constexpr size_t kMaxGridY = 65535;
    for (const Group& g : groups) {
        const int R = items[g.first].R, C = items[g.first].C;
        const int64_t m2n = (int64_t)2 * R * C;
        for (size_t o = 0; o < g.count; o += kMaxGridY) {
            const size_t nt = std::min(kMaxGridY, g.count - o);
            PlainParams p;
            p.M2 = reinterpret_cast<const double2*>(d + g.m2_off) + (o * (size_t)m2n);
            p.m2_stride = m2n;
            p.R = R;
            p.K = g.pl.K;
            p.B = g.pl.B;
            p.H = g.pl.H;
            p.partials = reinterpret_cast<ddc*>(d + g.part_off) + o * (size_t)g.pl.blocks;
            p.counter = reinterpret_cast<unsigned*>(d + cnt_off) + g.first + o;
            p.out = reinterpret_cast<double*>(d + out_off) + 6 * (g.first + o);
            p.scale = std::ldexp(1.0, -(R - 1));
            e = launch_plain(C, p, g.pl.blocks, g.pl.block, (int)nt, E.stream);
            if (e != hipSuccess) return drain_fail(E, e, "list sweep launch");
        }
    }
    double* hout = h + m2_total / sizeof(double);
    e = hipMemcpyAsync(hout, d + out_off, n * 6 * sizeof(double), hipMemcpyDeviceToHost, E.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(E.stream);
    if (e != hipSuccess) return drain_fail(E, e, "list sweep");
    for (size_t k = 0; k < n; ++k) {
        out[2 * items[k].idx] = hout[6 * k];
        out[2 * items[k].idx + 1] = hout[6 * k + 1];
    }
    return PQB_OK;
}

int list_entry(const double* const* A, const int64_t* rows, const int64_t* cols, const int64_t* ld,
               int64_t count, int algo, int precision, double* out) {
    if (count < 0) return fail(PQB_ERR_ARG, "negative matrix count");
    if (count == 0) return PQB_OK;
    if (!A || !rows || !cols || !ld || !out) return fail(PQB_ERR_ARG, "NULL pointer");
    if (algo < PQB_ALGO_BBFG || algo > PQB_ALGO_SIMPLE) return fail(PQB_ERR_ARG, "unknown algorithm code");
    if (precision < PQB_PREC_DOUBLE || precision > PQB_PREC_INF)
        return fail(PQB_ERR_ARG, "unknown precision code");
    auto bad = [&](int64_t i, int code, const std::string& msg) {
        return fail(code, "matrix " + std::to_string(i) + ": " + msg);
    };
    std::vector<ListItem> fp64, inf;
    std::vector<std::pair<int64_t, std::vector<int>>> exact;
    for (int64_t i = 0; i < count; ++i) {
        if (rows[i] < 0 || cols[i] < 0) return bad(i, PQB_ERR_ARG, "negative matrix dimension");
        if (plain_edge(algo, rows[i], cols[i], out + 2 * i)) continue;
        if (!A[i]) return bad(i, PQB_ERR_ARG, "matrix pointer is NULL");
        if (ld[i] < cols[i]) return bad(i, PQB_ERR_ARG, "leading dimension smaller than the number of columns");
        if (cols[i] > PQB_MAX_COLS)
            return bad(i, PQB_ERR_SIZE, "matrix has more than " + std::to_string(PQB_MAX_COLS) + " columns");
        std::vector<int> ai;
        long double lb = 0.0L;
        const bool is_int = precision != PQB_PREC_FP64 && integer_matrix(A[i], rows[i], cols[i], ld[i], ai, &lb);
        const bool exact_ok = is_int && exact_fits(lb);
        if (precision == PQB_PREC_EXACT && !exact_ok)
            return bad(i, PQB_ERR_EXACT, is_int ? "exact mode: the permanent may exceed the 128-bit range"
                                                : "exact mode needs a real integer-valued matrix");
        const ListItem it{i, (int)rows[i], (int)cols[i]};
        if (exact_ok)
            exact.emplace_back(i, std::move(ai));
        else if (precision == PQB_PREC_INF)
            inf.push_back(it);
        else
            fp64.push_back(it);
    }
    if (exact.empty() && inf.empty() && fp64.empty()) return PQB_OK;
    Lease L;
    int st = L.status();
    if (st) return st;
    Engine& E = *L;
    for (auto& ex : exact) {
        i128 v = 0;
        st = run_exact_locked(E, ex.second, (int)rows[ex.first], (int)cols[ex.first], &v);
        if (st) return bad(ex.first, st, t_err);
        out[2 * ex.first] = i128_to_double(v);
        out[2 * ex.first + 1] = 0.0;
    }
    if (!inf.empty()) {
        std::vector<const double*> pa;
        std::vector<int> pr, pc;
        std::vector<int64_t> pl;
        for (const ListItem& it : inf) {
            pa.push_back(A[it.idx]);
            pr.push_back(it.R);
            pc.push_back(it.C);
            pl.push_back(ld[it.idx]);
        }
        std::vector<double> o(2 * inf.size());
        st = run_plain_dd_list_locked(E, pa.data(), pr.data(), pc.data(), pl.data(), (int64_t)inf.size(),
                                      o.data());
        if (st) return st;
        for (size_t k = 0; k < inf.size(); ++k) {
            out[2 * inf[k].idx] = o[2 * k];
            out[2 * inf[k].idx + 1] = o[2 * k + 1];
        }
    }
    return list_fp64_locked(E, A, ld, fp64, out);
}

// ----------------------------------------------------- device-pointer API --
//
// Workspace of the device entry points: one counter per matrix (rounded up
// to 256 bytes), then `blocks_per_matrix` dd partials per matrix, then the
// +-2A tables of all matrices.  For one matrix the layout is the documented
// pqb_perm_device one (counter slot of 256 bytes, kMaxBlocks partials).

size_t plain_device_bytes(int64_t count, int64_t rows, int64_t cols, int64_t blocks_per_matrix) {
    return round256((size_t)count * sizeof(unsigned)) + (size_t)count * blocks_per_matrix * sizeof(ddc) +
           (size_t)count * 2 * rows * cols * sizeof(double2);
}

int plain_device_check(int64_t count, int64_t rows, int64_t cols) {
    if (count < 1) return fail(PQB_ERR_ARG, "batch count must be positive");
    if (rows < 1 || cols < 1 || rows >= cols + 2)
        return fail(PQB_ERR_ARG, "the device entry points need 1 <= rows <= cols + 1");
    if (cols > PQB_MAX_COLS)
        return fail(PQB_ERR_SIZE, "matrix has more than " + std::to_string(PQB_MAX_COLS) + " columns");
    return PQB_OK;
}

// Launch plan of a batch of `count` equal-size matrices: the single-matrix
// plan for the resident waves divided among the matrices.
int plain_device_blocks(int64_t count, int R, int C, PlainPlan* plan, int64_t* blocks_per_matrix) {
    const int st = bind_device(-1, nullptr, nullptr);
    if (st) return st;
    const int64_t wr = resident_waves(C);
    const PlainPlan pl = make_plain_plan(R, C, std::max<int64_t>(1, wr / count));
    if (pl.blocks > kMaxBlocks) return fail(PQB_ERR_SIZE, "launch plan exceeds the workspace layout");
    if (plan) *plan = pl;
    *blocks_per_matrix = pl.blocks;
    return PQB_OK;
}

int plain_device_run(const double* dA, int64_t count, int64_t rows, int64_t cols, double* d_out, void* d_work,
                     size_t work_bytes, hipStream_t st, bool single) {
    if (!dA || !d_out || !d_work) return fail(PQB_ERR_ARG, "NULL device pointer");
    int s = plain_device_check(count, rows, cols);
    if (s) return s;
    const int R = (int)rows, C = (int)cols;
    PlainPlan pl;
    int64_t bpm = 0;
    s = plain_device_blocks(count, R, C, &pl, &bpm);
    if (s) return s;
    if (single) bpm = kMaxBlocks;  // the documented single-matrix layout
    if (work_bytes < plain_device_bytes(count, rows, cols, bpm)) return fail(PQB_ERR_ARG, "workspace too small");
    char* w = static_cast<char*>(d_work);
    unsigned* counters = reinterpret_cast<unsigned*>(w);
    ddc* partials = reinterpret_cast<ddc*>(w + round256((size_t)count * sizeof(unsigned)));
    double2* m2 = reinterpret_cast<double2*>(reinterpret_cast<char*>(partials) + (size_t)count * bpm * sizeof(ddc));
    const int64_t m2n = (int64_t)2 * R * C;
    hipError_t e = launch_plain_prep(reinterpret_cast<const double2*>(dA), m2, R * C, count, st);
    if (e != hipSuccess) return hip_fail(e, "prep launch");
    constexpr int64_t kMaxGridY = 65535;
    for (int64_t off = 0; off < count; off += kMaxGridY) {
        const int64_t nt = std::min(kMaxGridY, count - off);
        PlainParams p;
        p.M2 = m2 + off * m2n;
        p.m2_stride = m2n;
        p.R = R;
        p.K = pl.K;
        p.B = pl.B;
        p.H = pl.H;
        p.partials = partials + off * pl.blocks;
        p.counter = counters + off;
        p.out = d_out + 6 * off;
        p.scale = std::ldexp(1.0, -(R - 1));
        e = launch_plain(C, p, pl.blocks, pl.block, (int)nt, st);
        if (e != hipSuccess) return hip_fail(e, "plain sweep launch");
    }
    return PQB_OK;
}

// Repeated-row permanent: a batch of one task (batch.hip applies the
// reference's validation and edge rules).
int repeated_entry(const double* A, int64_t rows, int64_t cols, int64_t ld, const int64_t* row_mult,
                   const int64_t* col_mult, int algo, int precision, double* out) {
    if (!out) return fail(PQB_ERR_ARG, "output pointer is NULL");
    if (precision == PQB_PREC_EXACT) return fail(PQB_ERR_ARG, "exact mode is available for the plain permanent only");
    const pqb_task t{A, rows, cols, ld, row_mult, col_mult};
    return batch_entry(&t, 1, algo, precision, out, 0);
}

}  // namespace pqb

using namespace pqb;

extern "C" {

const char* pqb_last_error(void) { return t_err.c_str(); }

int pqb_abi_version(void) { return 1; }

int pqb_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int pqb_init(int device) { return bind_device(device, nullptr, nullptr); }

int pqb_perm(const double* A, int64_t rows, int64_t cols, int64_t ld, int algo, int precision,
             double* out) {
    try {
        return plain_entry(A, rows, cols, ld, algo, precision, out, nullptr);
    } catch (const std::exception& ex) {
        return fail(PQB_ERR_DEVICE, std::string("internal error: ") + ex.what());
    }
}

int pqb_perm_list(const double* const* A, const int64_t* rows, const int64_t* cols, const int64_t* ld,
                  int64_t count, int algo, int precision, double* out) {
    try {
        return list_entry(A, rows, cols, ld, count, algo, precision, out);
    } catch (const std::exception& ex) {
        return fail(PQB_ERR_DEVICE, std::string("internal error: ") + ex.what());
    }
}

int pqb_perm_exact(const double* A, int64_t rows, int64_t cols, int64_t ld, int64_t* words,
                   double* out) {
    if (!words) return fail(PQB_ERR_ARG, "words pointer is NULL");
    try {
        return plain_entry(A, rows, cols, ld, PQB_ALGO_BBFG, PQB_PREC_EXACT, out, words);
    } catch (const std::exception& ex) {
        return fail(PQB_ERR_DEVICE, std::string("internal error: ") + ex.what());
    }
}

int pqb_perm_repeated(const double* A, int64_t rows, int64_t cols, int64_t ld, const int64_t* row_mult,
                      const int64_t* col_mult, int algo, int precision, double* out) {
    try {
        return repeated_entry(A, rows, cols, ld, row_mult, col_mult, algo, precision, out);
    } catch (const std::exception& ex) {
        return fail(PQB_ERR_DEVICE, std::string("internal error: ") + ex.what());
    }
}

int pqb_exact_bound(const double* A, int64_t rows, int64_t cols, int64_t ld, double* log2_bound) {
    if (!A || !log2_bound || rows < 1 || cols < 1 || ld < cols) return fail(PQB_ERR_ARG, "bad arguments");
    std::vector<int> ai;
    long double lb = 0.0L;
    if (!integer_matrix(A, rows, cols, ld, ai, &lb))
        return fail(PQB_ERR_EXACT, "exact mode needs a real integer-valued matrix");
    *log2_bound = (double)lb;
    return exact_fits(lb) ? PQB_OK : fail(PQB_ERR_EXACT, "exact mode: the permanent may exceed the 128-bit range");
}

size_t pqb_perm_device_workspace(int64_t rows, int64_t cols) {
    if (rows < 1 || cols < 1) return kCounterBytes;
    return plain_device_bytes(1, rows, cols, kMaxBlocks);
}

int pqb_perm_device(const double* dA, int64_t rows, int64_t cols, double* d_out, void* d_work,
                    size_t work_bytes, void* stream) {
    return plain_device_run(dA, 1, rows, cols, d_out, d_work, work_bytes, static_cast<hipStream_t>(stream), true);
}

int pqb_perm_batch_device_workspace(int64_t count, int64_t rows, int64_t cols, size_t* bytes) {
    if (!bytes) return fail(PQB_ERR_ARG, "bytes pointer is NULL");
    int st = plain_device_check(count, rows, cols);
    if (st) return st;
    int64_t bpt = 0;
    st = plain_device_blocks(count, (int)rows, (int)cols, nullptr, &bpt);
    if (st) return st;
    *bytes = plain_device_bytes(count, rows, cols, bpt);
    return PQB_OK;
}

int pqb_perm_batch_device(const double* dA, int64_t count, int64_t rows, int64_t cols, double* d_out,
                          void* d_work, size_t work_bytes, void* stream) {
    return plain_device_run(dA, count, rows, cols, d_out, d_work, work_bytes, static_cast<hipStream_t>(stream),
                            false);
}

int pqb_plain_plan(int64_t rows, int64_t cols, int64_t waves_resident, int64_t* plan) {
    if (!plan) return fail(PQB_ERR_ARG, "plan pointer is NULL");
    if (rows < 1 || cols < 1 || rows >= cols + 2 || cols > PQB_MAX_COLS)
        return fail(PQB_ERR_ARG, "no sweep for this shape");
    int64_t wr = waves_resident;
    if (wr <= 0) {
        const int st = bind_device(-1, nullptr, nullptr);
        if (st) return st;
        wr = resident_waves((int)cols);
    }
    const PlainPlan pl = make_plain_plan((int)rows, (int)cols, wr);
    plan[0] = pl.K;
    plan[1] = pl.B;
    plan[2] = pl.H;
    plan[3] = pl.waves;
    plan[4] = pl.block;
    plan[5] = pl.blocks;
    return PQB_OK;
}

int pqb_time_perm_device(const double* dA, int64_t rows, int64_t cols, double* d_out, int iters,
                         double* ms) {
    if (!ms || iters < 1) return fail(PQB_ERR_ARG, "bad timing arguments");
    const size_t wb = pqb_perm_device_workspace(rows, cols);
    void* w = nullptr;
    Lease L;  // the timed launches run on this context's stream, held for the call
    if (L.status()) return L.status();
    hipStream_t st = L->stream;
    hipError_t e = hipMalloc(&w, wb);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(timing workspace)");
    e = hipMemsetAsync(w, 0, kCounterBytes, st);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    int rc = PQB_OK;
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int i = 0; e == hipSuccess && rc == PQB_OK && i < iters; ++i)
        rc = pqb_perm_device(dA, rows, cols, d_out, w, wb, st);
    if (e == hipSuccess && rc == PQB_OK) e = hipEventRecord(e1, st);
    if (e == hipSuccess && rc == PQB_OK) e = hipEventSynchronize(e1);
    float f = 0.f;
    if (e == hipSuccess && rc == PQB_OK) e = hipEventElapsedTime(&f, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(st);
    (void)hipFree(w);
    if (rc != PQB_OK) return rc;
    if (e != hipSuccess) return hip_fail(e, "timing");
    *ms = f;
    return PQB_OK;
}

}  // extern "C"
