// Host side of the repeated-row and batched sweeps: task planning, step
// tables, work items and the graph-captured persistent launch.
//
// Planning a task (one repeated-row permanent), following
// BBFGPermanentCalculatorRepeated.hpp:74-101 and :116-307:
//   * rows / columns with zero multiplicity are dropped (they contribute a
//     factor 1, BatchedPermanentCalculator.cpp:412-451 builds the same compact
//     matrix);
//   * perm(A) = perm(A^T), so the orientation with fewer Gray steps
//     S = m_min prod_{i != min} (m_i + 1) is used (the DFE host code makes the
//     same choice, GlynnPermanentCalculatorRepeatedDFE.cpp:1138);
//   * the row with the smallest multiplicity (first occurrence) becomes the
//     anchor with multiplicity 1 and the rest of its multiplicity stays a digit
//     (BBFGPermanentCalculatorRepeated.hpp:78-94);
//   * columns of multiplicity c occupy c slots;
//   * digits are split into low digits (walked in a chunk, radix product L close
//     to the requested chunk length) and high digits (chunk index).
// The low walk is a reflected mixed-radix Gray code starting at k = 0: each
// digit sweeps up then down while the faster digits complete full sweeps, so
// every step moves exactly one digit by one; tables are shared by all tasks
// whose low radices agree.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <numeric>
#include <string>
#include <system_error>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pqb_perm.h"
#include "pqb_device.h"
#include "pqb_engine.h"
#include "pqb_kernels.h"

namespace pqb {

// ----------------------------------------------------------------- planning --

static uint64_t binom_u64(int n, int k) {
    if (k < 0 || k > n) return 0;
    uint64_t c = 1;
    for (int i = 1; i <= k; ++i) c = c * (uint64_t)(n - k + i) / (uint64_t)i;  // exact at every step
    return c;
}

// Gray steps of the repeated sweep for multiplicities v (nonzero entries).
static long double sweep_steps(const std::vector<int64_t>& v) {
    if (v.empty()) return 0.0L;
    size_t mi = 0;
    for (size_t i = 1; i < v.size(); ++i)
        if (v[i] < v[mi]) mi = i;
    long double s = (long double)v[mi];
    for (size_t i = 0; i < v.size(); ++i)
        if (i != mi) s *= (long double)(v[i] + 1);
    return s;
}

namespace {
struct Digit {
    int64_t src;  // oriented row
    int m;        // multiplicity walked by this digit
};
// Temporaries of plan_task, kept per thread and reused: a batch plans thousands
// of tasks on host threads, and fresh vectors per task made the allocator the
// bottleneck.
struct PlanScratch {
    std::vector<int64_t> ri, rm, ci, cm, slot_col;
    std::vector<Digit> digits, low, high, order_rows;
    std::vector<size_t> ord;
};
}  // namespace

int plan_task(const double* A, int64_t rows, int64_t cols, int64_t ld, const int64_t* rmult,
              const int64_t* cmult, uint64_t target_L, TaskPlan& P, bool plain) {
    P = TaskPlan();
    thread_local PlanScratch sc;
    std::vector<int64_t>& ri = sc.ri;
    std::vector<int64_t>& rm = sc.rm;
    std::vector<int64_t>& ci = sc.ci;
    std::vector<int64_t>& cm = sc.cm;
    ri.clear();
    rm.clear();
    ci.clear();
    cm.clear();
    int64_t N = 0, Nc = 0;
    for (int64_t i = 0; i < rows; ++i)
        if (rmult[i] > 0) {
            ri.push_back(i);
            rm.push_back(rmult[i]);
            N += rmult[i];
        }
    for (int64_t j = 0; j < cols; ++j)
        if (cmult[j] > 0) {
            ci.push_back(j);
            cm.push_back(cmult[j]);
            Nc += cmult[j];
        }
    // plain: a rows x cols matrix with every multiplicity 1 and rows <= cols + 1
    // (the BBFG sum of BBFGPermanentCalculator.hpp:87-235, which is not
    // transpose invariant when rows != cols)
    if (!plain && N != Nc) return fail(PQB_ERR_STATE, "Number of input and output states should be equal");
    if (N == 0) {
        P.trivial = true;
        return PQB_OK;
    }
    if (Nc > kMaxSlots || (!plain && N > kMaxSlots))
        return fail(PQB_ERR_SIZE, "more than " + std::to_string(kMaxSlots) + " photons");
    const bool transpose = !plain && sweep_steps(cm) < sweep_steps(rm);
    // element (r, c) of the oriented compact matrix
    auto elem = [&](int64_t r, int64_t c, double* re, double* im) {
        // oriented row r / column c -> original row i, column j
        const int64_t i = transpose ? ri[(size_t)c] : ri[(size_t)r];
        const int64_t j = transpose ? ci[(size_t)r] : ci[(size_t)c];
        *re = A[2 * (i * ld + j)];
        *im = A[2 * (i * ld + j) + 1];
    };
    const std::vector<int64_t>& row_m = transpose ? cm : rm;
    const std::vector<int64_t>& col_m = transpose ? rm : cm;

    // column slots
    std::vector<int64_t>& slot_col = sc.slot_col;
    slot_col.clear();
    for (size_t c = 0; c < col_m.size(); ++c)
        for (int64_t k = 0; k < col_m[c]; ++k) slot_col.push_back((int64_t)c);
    const int NS = (int)slot_col.size();

    // anchor: smallest multiplicity, first occurrence
    size_t anchor = 0;
    for (size_t r = 1; r < row_m.size(); ++r)
        if (row_m[r] < row_m[anchor]) anchor = r;
    std::vector<Digit>& digits = sc.digits;
    digits.clear();
    for (size_t r = 0; r < row_m.size(); ++r) {
        const int64_t m = (r == anchor) ? row_m[r] - 1 : row_m[r];
        if (m > 0) digits.push_back({(int64_t)r, (int)m});
    }
    // low / high split: largest radices first into the walk while it fits
    std::vector<size_t>& ord = sc.ord;
    ord.resize(digits.size());
    std::iota(ord.begin(), ord.end(), 0);
    // stable, largest first (insertion sort: a few dozen digits, no buffer to allocate)
    for (size_t i = 1; i < ord.size(); ++i) {
        const size_t v = ord[i];
        size_t j = i;
        for (; j > 0 && digits[ord[j - 1]].m < digits[v].m; --j) ord[j] = ord[j - 1];
        ord[j] = v;
    }
    uint64_t L = 1;
    for (;;) {
        std::vector<Digit>& low = sc.low;
        std::vector<Digit>& high = sc.high;
        low.clear();
        high.clear();
        L = 1;
        for (size_t k : ord) {
            const uint64_t rad = (uint64_t)digits[k].m + 1;
            if (L * rad <= target_L) {
                low.push_back(digits[k]);
                L *= rad;
            } else {
                high.push_back(digits[k]);
            }
        }
        long double nch = 1.0L;
        for (const Digit& d : high) nch *= (long double)(d.m + 1);
        if (nch < 4.0e9L) {
            P.nchunks = (uint64_t)nch;
            // rows of A: anchor, low digits in walk order, high digits
            std::vector<Digit>& order_rows = sc.order_rows;
            order_rows.clear();
            order_rows.push_back({(int64_t)anchor, 1});
            for (const Digit& d : low) order_rows.push_back(d);
            for (const Digit& d : high) order_rows.push_back(d);
            P.R = (int)order_rows.size();
            P.NS = NS;
            P.N = (int)N;
            P.nlow = (int)low.size();
            P.nhigh = (int)high.size();
            P.L = L;
            P.A.assign((size_t)P.R * NS * 2, 0.0);
            for (int r = 0; r < P.R; ++r)
                for (int s = 0; s < NS; ++s) {
                    double re, im;
                    elem(order_rows[(size_t)r].src, slot_col[(size_t)s], &re, &im);
                    P.A[2 * ((size_t)r * NS + s)] = re;
                    P.A[2 * ((size_t)r * NS + s) + 1] = im;
                }
            P.base.assign((size_t)NS * 2, 0.0);
            for (int s = 0; s < NS; ++s) {
                // anchor + sum_low m a: in the order of the reference seed
                // (BBFGPermanentCalculatorRepeated.hpp:202-226)
                double br = P.A[2 * (size_t)s], bi = P.A[2 * (size_t)s + 1];
                for (int d = 0; d < P.nlow; ++d) {
                    const double m = (double)low[(size_t)d].m;
                    br += P.A[2 * ((size_t)(1 + d) * NS + s)] * m;
                    bi += P.A[2 * ((size_t)(1 + d) * NS + s) + 1] * m;
                }
                P.base[2 * (size_t)s] = br;
                P.base[2 * (size_t)s + 1] = bi;
            }
            P.low_radix.clear();
            for (const Digit& d : low) P.low_radix.push_back(d.m + 1);
            P.hrad.clear();
            for (const Digit& d : high) P.hrad.push_back((uint32_t)(d.m + 1));
            P.scale = std::ldexp(1.0, -(int)(N - 1));
            return PQB_OK;
        }
        target_L *= 4;  // too many chunks for 32-bit chunk indices: walk more per chunk
    }
}

// Reflected mixed-radix Gray walk from k = 0 over digits with the given radices
// (digit 0 moves fastest).  Entry t describes the move into index t; the table
// has prod(radix) entries.  A step moves one digit by one, so the weight
// prod C(m, k) changes by one exact ratio of binomials (an int64: the digits of a
// walk hold at most 40 photons, prod C <= 2^40) and its sign alternates.
static void build_table(const int* radix, size_t n, RepStep* tab) {
    uint64_t L = 1;
    for (size_t k = 0; k < n; ++k) L *= (uint64_t)radix[k];
    int g[kMaxDigits], o[kMaxDigits];
    int64_t bw[kMaxDigits][kMaxSlots + 1];
    for (size_t k = 0; k < n; ++k) {
        g[k] = 0;
        o[k] = 1;
        for (int v = 0; v < radix[k]; ++v) bw[k][v] = (int64_t)binom_u64(radix[k] - 1, v);
    }
    tab[0] = RepStep{0u, 0u, 1.0};
    int64_t w = 1;
    for (uint64_t t = 1; t < L; ++t) {
        size_t d = 0;
        int nv = 0;
        for (;;) {
            nv = g[d] + o[d];
            if (nv >= 0 && nv <= radix[d] - 1) break;
            o[d] = -o[d];
            ++d;
        }
        const bool up = nv > g[d];
        w = w / bw[d][g[d]] * bw[d][nv];
        g[d] = nv;
        tab[t].row = (uint32_t)(1 + d);
        tab[t].dir = up ? 1u : 0u;
        tab[t].w = (double)((t & 1) ? -w : w);
    }
}

// ------------------------------------------------------------ batch buffers --

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Host worker threads kept for the life of the process (a 16-thread spawn and
// join costs ~0.5 ms, and a batch call runs four parallel phases).  One parallel
// phase at a time uses the pool; a phase that finds it busy (calls from several
// host threads) starts threads of its own.  The pool object is never destroyed,
// so no thread waits on a destroyed condition variable at process exit.
struct HostPool {
    std::mutex call_mu;  // one phase at a time
    std::mutex mu;
    std::condition_variable wake, fin;
    const std::function<void(int64_t, int64_t)>* job = nullptr;
    int64_t n = 0;
    int slices = 0, next = 0, done = 0, workers = 0;

    void worker() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            wake.wait(lk, [&] { return job != nullptr && next < slices; });
            run_one(lk);
        }
    }
    // claims the next slice (mu held, a slice left) and runs it with mu released
    void run_one(std::unique_lock<std::mutex>& lk) {
        const int s = next++;
        const std::function<void(int64_t, int64_t)>* f = job;
        const int64_t lo = n * s / slices, hi = n * (s + 1) / slices;
        lk.unlock();
        (*f)(lo, hi);
        lk.lock();
        if (++done == slices) fin.notify_all();
    }
    void run(int64_t count, int nslices, const std::function<void(int64_t, int64_t)>& f) {
        {
            std::lock_guard<std::mutex> g(mu);
            while (workers < nslices - 1) {  // the calling thread runs slices too
                try {
                    std::thread(&HostPool::worker, this).detach();
                } catch (const std::system_error&) {
                    break;  // fewer workers: the caller claims what is left
                }
                ++workers;
            }
            job = &f;
            n = count;
            slices = nslices;
            next = 0;
            done = 0;
        }
        wake.notify_all();
        std::unique_lock<std::mutex> lk(mu);
        while (next < slices) run_one(lk);
        fin.wait(lk, [&] { return done == slices; });
        job = nullptr;
    }
};

static HostPool& host_pool() {
    static HostPool* pool = new HostPool;  // never destroyed (see above)
    return *pool;
}

// fn(lo, hi) over [0, n) in contiguous slices on up to 16 host threads, at
// least `grain` items per slice; the caller's thread runs slices itself, and all
// of them when no thread can be started.  fn must not throw.
template <class F>
static void host_parallel(int64_t n, int64_t grain, F&& fn) {
    const int64_t hw = std::max<int64_t>(1, (int64_t)std::thread::hardware_concurrency());
    const int64_t nthreads = std::min<int64_t>(hw, std::min<int64_t>(16, (n + grain - 1) / grain));
    if (nthreads <= 1) {
        fn((int64_t)0, n);
        return;
    }
    HostPool& pool = host_pool();
    std::unique_lock<std::mutex> busy(pool.call_mu, std::try_to_lock);
    if (busy.owns_lock()) {
        const std::function<void(int64_t, int64_t)> f = [&fn](int64_t lo, int64_t hi) { fn(lo, hi); };
        pool.run(n, (int)nthreads, f);
        return;
    }
    std::vector<std::thread> spawned;
    spawned.reserve((size_t)nthreads);
    for (int64_t t = 0; t < nthreads; ++t) {
        const int64_t lo = n * t / nthreads, hi = n * (t + 1) / nthreads;
        try {
            spawned.emplace_back(fn, lo, hi);
        } catch (const std::system_error&) {
            fn(lo, hi);
        }
    }
    for (std::thread& th : spawned) th.join();
}

struct BatchBuffers {
    void* dbuf = nullptr;
    void* hbuf = nullptr;
    size_t cap_tasks = 0, cap_items = 0, cap_data = 0;
    size_t off_binom = 0, off_tasks = 0, off_items = 0, off_order = 0, off_data = 0, off_part = 0, off_out = 0,
           total = 0;
    hipGraphExec_t graph[kRanges] = {};  // FP64 slot ranges, then the double-double ones
    int grid[kRanges] = {};
    // last prepared batch (for replays)
    int last_range = -1;
    int last_ntasks = 0;

    void layout(size_t nt, size_t ni, size_t nd) {
        cap_tasks = nt;
        cap_items = ni;
        cap_data = nd;
        off_binom = 256;
        off_tasks = align256(off_binom + sizeof(double) * kBinomStride * kBinomStride);
        off_items = align256(off_tasks + nt * sizeof(RepTask));
        off_order = align256(off_items + ni * sizeof(RepItem));
        off_data = align256(off_order + ni * sizeof(unsigned));
        off_part = align256(off_data + nd);
        off_out = align256(off_part + ni * sizeof(ddc));
        total = align256(off_out + nt * 6 * sizeof(double));
    }
    void drop_graphs() {
        for (auto& g : graph)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
    }
};

// The context's batch buffers, created on first use.
static BatchBuffers& batch_of(Engine& E) {
    if (!E.batch) E.batch = new BatchBuffers();
    return *E.batch;
}

// Releases a context's batch buffers and graphs (its stream is drained); the
// context lock is held.
void release_batch_locked(Engine& E) {
    if (!E.batch) return;
    BatchBuffers& B = *E.batch;
    B.drop_graphs();
    if (B.dbuf) (void)hipFree(B.dbuf);
    if (B.hbuf) (void)hipHostFree(B.hbuf);
    B.dbuf = B.hbuf = nullptr;
    B.cap_tasks = B.cap_items = B.cap_data = 0;
    for (int& g : B.grid) g = 0;
    B.last_range = -1;
    B.last_ntasks = 0;
}

// The context whose batch ran last: pqb_batch_replay replays it.
static std::atomic<Engine*> g_last_batch{nullptr};

static int ensure_batch_buffers_locked(Engine& E, size_t nt, size_t ni, size_t nd) {
    BatchBuffers& B = batch_of(E);
    if (B.dbuf && nt <= B.cap_tasks && ni <= B.cap_items && nd <= B.cap_data) return PQB_OK;
    (void)hipStreamSynchronize(E.stream);
    B.drop_graphs();
    if (B.dbuf) (void)hipFree(B.dbuf);
    if (B.hbuf) (void)hipHostFree(B.hbuf);
    B.dbuf = B.hbuf = nullptr;
    B.layout(std::max<size_t>(nt, std::max<size_t>(2 * B.cap_tasks, 64)),
             std::max<size_t>(ni, std::max<size_t>(2 * B.cap_items, 4096)),
             std::max<size_t>(nd, std::max<size_t>(2 * B.cap_data, (size_t)1 << 20)));
    // on any failure below, leave no buffer behind: the capacity check above
    // must not accept a half-built set on the next call
    auto release = [&B]() {
        if (B.dbuf) (void)hipFree(B.dbuf);
        if (B.hbuf) (void)hipHostFree(B.hbuf);
        B.dbuf = B.hbuf = nullptr;
        B.cap_tasks = B.cap_items = B.cap_data = 0;
    };
    hipError_t e = hipMalloc(&B.dbuf, B.total);
    if (e != hipSuccess) {
        B.dbuf = nullptr;
        release();
        return hip_fail(e, "hipMalloc(batch)");
    }
    e = hipHostMalloc(&B.hbuf, B.off_data + B.cap_data, hipHostMallocDefault);
    if (e != hipSuccess) {
        B.hbuf = nullptr;
        release();
        return hip_fail(e, "hipHostMalloc(batch)");
    }
    std::memset(B.hbuf, 0, B.off_data);
    double* bt = reinterpret_cast<double*>(static_cast<char*>(B.hbuf) + B.off_binom);
    for (int m = 0; m < kBinomStride; ++m)
        for (int k = 0; k < kBinomStride; ++k) bt[m * kBinomStride + k] = (double)binom_u64(m, k);
    // counter + header start at zero on the device
    e = hipMemcpyAsync(B.dbuf, B.hbuf, B.off_tasks, hipMemcpyHostToDevice, E.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(E.stream);
    if (e != hipSuccess) {
        release();
        return hip_fail(e, "batch init");
    }
    return PQB_OK;
}

static RepBatchParams batch_params(Engine& E) {
    BatchBuffers& B = batch_of(E);
    char* d = static_cast<char*>(B.dbuf);
    RepBatchParams p;
    p.tasks = reinterpret_cast<const RepTask*>(d + B.off_tasks);
    p.items = reinterpret_cast<const RepItem*>(d + B.off_items);
    p.order = reinterpret_cast<const unsigned*>(d + B.off_order);
    p.binom = reinterpret_cast<const double*>(d + B.off_binom);
    p.partials = reinterpret_cast<ddc*>(d + B.off_part);
    p.out = reinterpret_cast<double*>(d + B.off_out);
    p.head = reinterpret_cast<unsigned*>(d);
    p.counts = reinterpret_cast<const unsigned*>(d + 64);
    return p;
}

static int batch_grid_locked(Engine& E, int range) {
    BatchBuffers& B = batch_of(E);
    if (B.grid[range] == 0) {
        int nb = 0;
        const void* fn = range >= kDDRange0 ? rep_dd_kernel_fn(range) : rep_batch_kernel_fn(range);
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 256, 0);
        if (e != hipSuccess || nb <= 0) nb = 1;
        B.grid[range] = nb * E.cus;
    }
    return B.grid[range];
}

static int batch_graph_locked(Engine& E, int range, hipGraphExec_t* out) {
    BatchBuffers& B = batch_of(E);
    if (!B.graph[range]) {
        const int grid = batch_grid_locked(E, range);
        const RepBatchParams p = batch_params(E);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
        hipError_t le = range >= kDDRange0 ? launch_rep_dd(range, p, grid, E.stream)
                                           : launch_rep_batch(range, p, grid, E.stream);
        e = hipStreamEndCapture(E.stream, &g);
        if (le != hipSuccess) return hip_fail(le, "batch launch (capture)");
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        hipGraphExec_t ex = nullptr;
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
        B.graph[range] = ex;
    }
    *out = B.graph[range];
    return PQB_OK;
}

// Walk lengths set by the autotuner, indexed by photon count (0 = automatic).
static std::atomic<int64_t> g_walk[PQB_MAX_COLS + 1] = {};

// Target walk length for a task of S steps when the batch holds `total`
// steps: long walks amortise the per-chunk seed, but there must be enough
// chunks to keep every resident lane busy.
static uint64_t choose_L(long double S, long double total, int64_t lanes, bool fp64) {
    // A lane's time is (its chunks) x (walk + per-chunk cost: the item fetch, the
    // mixed-radix decomposition of the chunk index and the seed, tens of FP64 steps
    // with many chunk digits).  With little work per lane one chunk each; in the FP64
    // sweep chunks of up to 128 steps even if up to three quarters of the lanes stay
    // idle (a single 9e6-step task on the whole grid: 0.26 ms at 4 chunks per lane,
    // 0.137 ms at one chunk of ~45 steps per lane, 0.10 ms at 128-step chunks; a
    // 4608-step task keeps its 16-step chunks).  The double-double step costs ~10x an
    // FP64 step, so there the seed is cheap and short walks stay (n = 16: 0.046 ms
    // against 0.19 ms with the floor).  Larger batches take 4 chunks per lane so that
    // uneven items balance.
    const long double per_lane = std::max(total, S) / (long double)std::max<int64_t>(lanes, 1);
    const long double chunks_per_lane = per_lane < 1024.0L ? 1.0L : 4.0L;
    long double want_chunks =
        std::max<long double>(chunks_per_lane * (long double)lanes * S / std::max(total, S), 1.0L);
    long double L = S / want_chunks;
    L = std::min(L, S / 64.0L);  // at least one chunk per lane of a wave
    if (fp64 && per_lane < 1024.0L) L = std::max(L, std::min(128.0L, 4.0L * S / (long double)std::max<int64_t>(lanes, 1)));
    L = std::min<long double>(std::max<long double>(L, 16.0L), 2048.0L);
    return (uint64_t)L;
}

// Run a list of planned tasks; results (6 doubles each) into out6.  dd: the
// double-double sweep (a chunk per quad of lanes) instead of the FP64 one.
static int run_tasks_locked(Engine& E, std::vector<TaskPlan>& plans, std::vector<double>& out6, bool dd) {
    BatchBuffers& B = batch_of(E);
    const size_t nt = plans.size();
    out6.assign(nt * 6, 0.0);
    // sweeps to run
    std::vector<size_t> live;
    int min_ns = 1 << 30, max_ns = 0;
    for (size_t i = 0; i < nt; ++i) {
        if (plans[i].trivial) {
            out6[6 * i] = plans[i].tre;
            out6[6 * i + 1] = plans[i].tim;
            out6[6 * i + 2] = plans[i].tre;
            out6[6 * i + 4] = plans[i].tim;
            continue;
        }
        live.push_back(i);
        min_ns = std::min(min_ns, plans[i].NS);
        max_ns = std::max(max_ns, plans[i].NS);
    }
    if (live.empty()) return PQB_OK;
    const int range = dd ? rep_dd_range_for(max_ns) : rep_range_for(min_ns, max_ns);
    if (range < 0) {
        // split at 32 slots and run the two halves
        std::vector<TaskPlan> lo, hi;
        std::vector<size_t> ilo, ihi;
        for (size_t i : live) {
            if (plans[i].NS <= 32) {
                lo.push_back(std::move(plans[i]));
                ilo.push_back(i);
            } else {
                hi.push_back(std::move(plans[i]));
                ihi.push_back(i);
            }
        }
        std::vector<double> olo, ohi;
        int st = run_tasks_locked(E, lo, olo, dd);
        if (st) return st;
        st = run_tasks_locked(E, hi, ohi, dd);
        if (st) return st;
        for (size_t k = 0; k < ilo.size(); ++k) std::copy(&olo[6 * k], &olo[6 * k] + 6, &out6[6 * ilo[k]]);
        for (size_t k = 0; k < ihi.size(); ++k) std::copy(&ohi[6 * k], &ohi[6 * k] + 6, &out6[6 * ihi[k]]);
        return PQB_OK;
    }

    // ---- work items ----
    std::vector<RepItem> items;
    // cost class of each item: items take few distinct costs (one per task shape
    // and item size), ordered below by a counting pass instead of a sort
    std::vector<unsigned> item_class;
    std::vector<uint64_t> class_cost;
    std::unordered_map<uint64_t, unsigned> class_of;
    class_of.reserve(256);
    std::vector<std::pair<uint32_t, uint32_t>> task_items(live.size());
    // target Gray steps per lane (quad) per item: small items balance the tail of
    // an uneven batch (4096 tasks of 8..26 photons: replay 9.17 / 8.81 / 8.68 / 8.64 ms at 2048 / 1024 /
    // 512 / 256 steps), the cost being one atomic fetch per item
    const uint64_t lane_steps = 512;
    const uint64_t walkers = dd ? (uint64_t)rep_dd_walkers(range) : 64;  // chunks a wave works on at once
    auto per_item_of = [&](const TaskPlan& P) { return walkers * std::max<uint64_t>(1, lane_steps / P.L); };
    // A batch too small to give every resident wave ~4 items at that size
    // (a single mid-size task, say) gets smaller items, down to one chunk per
    // walker, so that the whole grid has work.
    uint64_t cap = ~(uint64_t)0;
    {
        const uint64_t target = 4 * (uint64_t)batch_grid_locked(E, range) * (kBlock / kWave);
        uint64_t n_default = 0, chunks = 0;
        for (size_t k : live) {
            n_default += (plans[k].nchunks + per_item_of(plans[k]) - 1) / per_item_of(plans[k]);
            chunks += plans[k].nchunks;
        }
        if (n_default < target) cap = walkers * std::max<uint64_t>(1, (chunks + target * walkers - 1) / (target * walkers));
    }
    {
        size_t nitems = 0;
        for (size_t k = 0; k < live.size(); ++k) {
            const uint64_t per_item = std::min(per_item_of(plans[live[k]]), cap);
            nitems += (size_t)((plans[live[k]].nchunks + per_item - 1) / per_item);
        }
        items.reserve(nitems);
        item_class.reserve(nitems);
    }
    uint64_t last_cost = ~(uint64_t)0;
    unsigned last_class = 0;
    for (size_t k = 0; k < live.size(); ++k) {
        const TaskPlan& P = plans[live[k]];
        const uint64_t per_item = std::min(per_item_of(P), cap);
        task_items[k].first = (uint32_t)items.size();
        for (uint64_t c0 = 0; c0 < P.nchunks; c0 += per_item) {
            const uint64_t cnt = std::min<uint64_t>(per_item, P.nchunks - c0);
            const uint64_t cost = ((cnt + walkers - 1) / walkers) * P.L * (uint64_t)(P.NS + 4);
            if (cost != last_cost) {
                const auto r = class_of.emplace(cost, (unsigned)class_cost.size());
                if (r.second) class_cost.push_back(cost);
                last_class = r.first->second;
                last_cost = cost;
            }
            item_class.push_back(last_class);
            items.push_back(RepItem{(uint32_t)k, (uint32_t)cnt, (uint32_t)c0, 0u});
        }
        task_items[k].second = (uint32_t)items.size() - task_items[k].first;
    }
    // launch order: largest cost first, ties in item order
    std::vector<unsigned> order(items.size());
    {
        std::vector<unsigned> rank(class_cost.size());
        std::iota(rank.begin(), rank.end(), 0u);
        std::sort(rank.begin(), rank.end(), [&](unsigned a, unsigned b) { return class_cost[a] > class_cost[b]; });
        std::vector<unsigned> start(class_cost.size() + 1, 0u), pos_of(class_cost.size());
        for (size_t r = 0; r < rank.size(); ++r) pos_of[rank[r]] = (unsigned)r;
        for (unsigned c : item_class) ++start[pos_of[c] + 1];
        for (size_t r = 0; r < rank.size(); ++r) start[r + 1] += start[r];
        for (size_t i = 0; i < items.size(); ++i) order[start[pos_of[item_class[i]]]++] = (unsigned)i;
    }

    // ---- data region: A, base per task, deduplicated step tables ----
    // step tables shared by the tasks whose walked radices agree, looked up by the
    // radices as a short string (radices <= 41, up to 40 digits); built below,
    // straight into the staging buffer
    std::unordered_map<std::string, size_t> table_off;
    table_off.reserve(64);
    std::string key;
    std::vector<std::pair<size_t, size_t>> tables;  // (offset, task whose radices it walks)
    std::vector<size_t> a_off(live.size()), b_off(live.size()), t_off(live.size());
    size_t nd = 0;
    for (size_t k = 0; k < live.size(); ++k) {
        const TaskPlan& P = plans[live[k]];
        a_off[k] = nd;
        nd = align256(nd + P.A.size() * sizeof(double));
        b_off[k] = nd;
        nd = align256(nd + P.base.size() * sizeof(double));
        key.assign(P.low_radix.begin(), P.low_radix.end());
        auto it = table_off.find(key);
        if (it == table_off.end()) {
            tables.emplace_back(nd, k);
            table_off.emplace(key, nd);
            t_off[k] = nd;
            nd = align256(nd + (size_t)P.L * sizeof(RepStep));
        } else {
            t_off[k] = it->second;
        }
    }
    int st = ensure_batch_buffers_locked(E, live.size(), items.size(), nd);
    if (st) return st;
    char* h = static_cast<char*>(B.hbuf);
    char* d = static_cast<char*>(B.dbuf);
    // header: head counter (0), counts
    std::memset(h, 0, 256);
    unsigned* counts = reinterpret_cast<unsigned*>(h + 64);
    counts[0] = (unsigned)items.size();
    counts[1] = (unsigned)live.size();
    RepTask* T = reinterpret_cast<RepTask*>(h + B.off_tasks);
    // descriptors and per-task data into the pinned staging buffer (disjoint
    // regions per task; ~20 MB for 4096 tasks of up to 26 photons)
    host_parallel((int64_t)live.size(), 256, [&](int64_t lo, int64_t hi) {
    for (int64_t k = lo; k < hi; ++k) {
        const TaskPlan& P = plans[live[(size_t)k]];
        RepTask t;
        std::memset(&t, 0, sizeof(t));
        t.A = reinterpret_cast<const double2*>(d + B.off_data + a_off[k]);
        t.base = reinterpret_cast<const double2*>(d + B.off_data + b_off[k]);
        t.table = reinterpret_cast<const RepStep*>(d + B.off_data + t_off[k]);
        t.NS = P.NS;
        t.L = (int)P.L;
        t.nhigh = P.nhigh;
        t.nlow = P.nlow;
        for (int ld = 0; ld < P.nlow; ++ld) t.lrad[ld] = (uint32_t)P.low_radix[(size_t)ld];
        t.item0 = task_items[k].first;
        t.nitems = task_items[k].second;
        t.out_index = (uint32_t)k;
        t.nchunks = (uint32_t)P.nchunks;
        t.scale = P.scale;
        for (int hd = 0; hd < P.nhigh; ++hd) {
            t.hrow[hd] = (uint32_t)(1 + P.nlow + hd);
            t.hrad[hd] = P.hrad[(size_t)hd];
        }
        T[k] = t;
        std::memcpy(h + B.off_data + a_off[(size_t)k], P.A.data(), P.A.size() * sizeof(double));
        std::memcpy(h + B.off_data + b_off[(size_t)k], P.base.data(), P.base.size() * sizeof(double));
    }
    });
    host_parallel((int64_t)tables.size(), 4, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            const std::vector<int>& rad = plans[live[tables[(size_t)i].second]].low_radix;
            build_table(rad.data(), rad.size(), reinterpret_cast<RepStep*>(h + B.off_data + tables[(size_t)i].first));
        }
    });
    RepItem* I = reinterpret_cast<RepItem*>(h + B.off_items);
    std::memcpy(I, items.data(), items.size() * sizeof(RepItem));
    unsigned* O = reinterpret_cast<unsigned*>(h + B.off_order);
    std::memcpy(O, order.data(), order.size() * sizeof(unsigned));

    hipError_t e = hipMemcpyAsync(d, h, B.off_data + nd, hipMemcpyHostToDevice, E.stream);
    if (e != hipSuccess) return hip_fail(e, "batch upload");
    hipGraphExec_t gx = nullptr;
    st = batch_graph_locked(E, range, &gx);
    if (st) {
        (void)hipStreamSynchronize(E.stream);  // the upload may still read the staging buffer
        return st;
    }
    e = hipGraphLaunch(gx, E.stream);
    if (e != hipSuccess) return drain_fail(E, e, "hipGraphLaunch");
    std::vector<double> res(live.size() * 6);
    e = hipMemcpyAsync(res.data(), d + B.off_out, res.size() * sizeof(double), hipMemcpyDeviceToHost,
                       E.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(E.stream);
    if (e != hipSuccess) return hip_fail(e, "batch sweep");
    for (size_t k = 0; k < live.size(); ++k) std::copy(&res[6 * k], &res[6 * k] + 6, &out6[6 * live[k]]);
    B.last_range = range;
    B.last_ntasks = (int)live.size();
    g_last_batch.store(&E);
    return PQB_OK;
}

// ------------------------------------------------------------- validation --

// Reference rules for one repeated task (BBFGPermanentCalculatorRepeated.cpp:
// 78-81, .hpp:119-128; GlynnPermanentCalculatorRepeated.hpp:55-63).  Returns
// 1 with *decided set when no sweep is needed.
static int check_task(const pqb_task& t, int algo, bool* decided, double* re) {
    *decided = false;
    if (t.rows < 0 || t.cols < 0) return fail(PQB_ERR_ARG, "negative matrix dimension");
    if ((t.rows > 0 && !t.row_mult) || (t.cols > 0 && !t.col_mult))
        return fail(PQB_ERR_ARG, "occupation pointer is NULL");
    if (algo == PQB_ALGO_BBFG && t.rows == 0) {
        *decided = true;
        *re = 1.0;
        return PQB_OK;
    }
    int64_t sr = 0, sc = 0;
    for (int64_t i = 0; i < t.rows; ++i) {
        if (t.row_mult[i] < 0) return fail(PQB_ERR_ARG, "negative occupation number");
        sr += t.row_mult[i];
    }
    for (int64_t j = 0; j < t.cols; ++j) {
        if (t.col_mult[j] < 0) return fail(PQB_ERR_ARG, "negative occupation number");
        sc += t.col_mult[j];
    }
    if (sr != sc)
        return fail(PQB_ERR_STATE,
                    algo == PQB_ALGO_BBFG
                        ? "BBFGPermanentCalculatorRepeated_Tasks::calculate:  Number of input and output states should be equal"
                        : "GlynnPermanentCalculatorRepeated::calculate:  Number of input and output states should be equal");
    if (t.rows == 0 || t.cols == 0 || sr == 0) {
        *decided = true;
        *re = 1.0;
        return PQB_OK;
    }
    if (!t.A) return fail(PQB_ERR_ARG, "matrix pointer is NULL");
    if (t.ld < t.cols) return fail(PQB_ERR_ARG, "leading dimension smaller than the number of columns");
    if (sr > PQB_MAX_COLS) return fail(PQB_ERR_SIZE, "more than " + std::to_string(PQB_MAX_COLS) + " photons");
    return PQB_OK;
}

int batch_entry(const pqb_task* tasks, int64_t ntasks, int algo, int precision, double* out, int64_t L_override) {
    if (ntasks < 0) return fail(PQB_ERR_ARG, "negative task count");
    if (ntasks > 0 && (!tasks || !out)) return fail(PQB_ERR_ARG, "NULL pointer");
    if (algo != PQB_ALGO_BBFG && algo != PQB_ALGO_GLYNN) return fail(PQB_ERR_ARG, "unknown algorithm code");
    if (precision < PQB_PREC_DOUBLE || precision > PQB_PREC_INF) return fail(PQB_ERR_ARG, "unknown precision code");
    if (precision == PQB_PREC_EXACT) return fail(PQB_ERR_ARG, "exact mode is available for the plain permanent only");
    const bool dd = precision == PQB_PREC_INF;
    std::vector<TaskPlan> plans((size_t)ntasks);
    std::vector<long double> steps((size_t)ntasks, 0.0L);
    std::vector<int> photons((size_t)ntasks, 0);
    // Validation and planning are per task and independent: large batches run
    // them on host threads (4096 tasks are milliseconds of host work either
    // way).  Errors come back per task; the first one by task index is
    // reported, as a serial loop would.
    std::vector<int> pst((size_t)ntasks, PQB_OK);
    std::vector<std::string> perr((size_t)ntasks);
    host_parallel(ntasks, 512, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            try {  // nothing may escape a worker thread
                bool decided = false;
                double v = 1.0;
                const int st = check_task(tasks[i], algo, &decided, &v);
                if (st) {
                    pst[(size_t)i] = st;
                    perr[(size_t)i] = t_err;  // t_err is per thread
                    continue;
                }
                if (decided) {
                    plans[(size_t)i].trivial = true;
                    plans[(size_t)i].tre = v;
                    continue;
                }
                std::vector<int64_t> rm(tasks[i].row_mult, tasks[i].row_mult + tasks[i].rows);
                std::vector<int64_t> cm(tasks[i].col_mult, tasks[i].col_mult + tasks[i].cols);
                rm.erase(std::remove(rm.begin(), rm.end(), 0), rm.end());
                cm.erase(std::remove(cm.begin(), cm.end(), 0), cm.end());
                steps[(size_t)i] = std::min(sweep_steps(rm), sweep_steps(cm));
                photons[(size_t)i] = (int)std::accumulate(rm.begin(), rm.end(), (int64_t)0);
            } catch (const std::exception& ex) {
                pst[(size_t)i] = PQB_ERR_DEVICE;
                perr[(size_t)i] = std::string("internal error: ") + ex.what();
            }
        }
    });
    for (int64_t i = 0; i < ntasks; ++i)
        if (pst[(size_t)i]) {
            t_err = (ntasks > 1 ? "task " + std::to_string(i) + ": " : std::string()) + perr[(size_t)i];
            return pst[(size_t)i];
        }
    long double total = 0.0L;
    int max_photons = 1;  // = the largest slot count of the batch
    for (int64_t i = 0; i < ntasks; ++i) {
        total += steps[(size_t)i];
        max_photons = std::max(max_photons, photons[(size_t)i]);
    }
    bool any_sweep = false;
    for (const TaskPlan& P : plans) any_sweep = any_sweep || !P.trivial;
    if (!any_sweep) {  // every result decided by the edge rules: no device needed
        for (int64_t i = 0; i < ntasks; ++i) {
            out[2 * i] = plans[(size_t)i].tre;
            out[2 * i + 1] = plans[(size_t)i].tim;
        }
        return PQB_OK;
    }
    Lease L;
    int st = L.status();
    if (st) return st;
    Engine& E = *L;
    // chunks in flight: a lane each (FP64), a group of lanes each (double-double)
    const int dd_range = rep_dd_range_for(max_photons);
    const int64_t lanes = dd ? (int64_t)batch_grid_locked(E, dd_range) * (kBlock / kWave) * rep_dd_walkers(dd_range)
                             : (int64_t)batch_grid_locked(E, 2) * kBlock;
    const bool large_batch = total >= 1024.0L * (long double)lanes;  // steps per lane
    auto plan_range = [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            if (plans[(size_t)i].trivial) continue;
            int64_t photons = 0;
            for (int64_t j = 0; j < tasks[i].cols; ++j) photons += tasks[i].col_mult[j];
            // The autotuned walk lengths are measured for the FP64 sweep on a large batch
            // (thousands of steps per lane); a small batch, a single task above all, keeps
            // the automatic rule (a 6e6-step task at the tuned 1024-step walk left all but
            // ~6000 lanes idle: 0.32 ms against 0.10 ms).
            const int64_t tuned =
                (!dd && large_batch && photons >= 0 && photons <= PQB_MAX_COLS) ? g_walk[photons].load() : 0;
            const uint64_t L = L_override > 0 ? (uint64_t)L_override
                               : tuned > 0    ? (uint64_t)tuned
                                              : choose_L(steps[(size_t)i], total, lanes, !dd);
            try {  // nothing may escape a worker thread
                pst[(size_t)i] = plan_task(tasks[i].A, tasks[i].rows, tasks[i].cols, tasks[i].ld,
                                           tasks[i].row_mult, tasks[i].col_mult, L, plans[(size_t)i]);
                if (pst[(size_t)i]) perr[(size_t)i] = t_err;  // t_err is per thread
            } catch (const std::exception& ex) {
                pst[(size_t)i] = PQB_ERR_DEVICE;
                perr[(size_t)i] = std::string("internal error: ") + ex.what();
            }
        }
    };
    host_parallel(ntasks, 256, plan_range);
    for (int64_t i = 0; i < ntasks; ++i)
        if (pst[(size_t)i]) {
            t_err = (ntasks > 1 ? "task " + std::to_string(i) + ": " : std::string()) + perr[(size_t)i];
            return pst[(size_t)i];
        }
    std::vector<double> out6;
    st = run_tasks_locked(E, plans, out6, dd);
    if (st) return st;
    for (int64_t i = 0; i < ntasks; ++i) {
        out[2 * i] = out6[(size_t)(6 * i)];
        out[2 * i + 1] = out6[(size_t)(6 * i + 1)];
    }
    return PQB_OK;
}

int run_plain_dd_list_locked(Engine& E, const double* const* A, const int* R, const int* C, const int64_t* ld,
                             int64_t n, double* out) {
    std::vector<TaskPlan> plans((size_t)n);
    for (int64_t k = 0; k < n; ++k) {
        const std::vector<int64_t> rm((size_t)R[k], 1), cm((size_t)C[k], 1);
        const long double S = std::ldexp(1.0L, R[k] - 1);
        const int range = rep_dd_range_for(C[k]);  // the slots are the columns
        const int64_t walkers = (int64_t)batch_grid_locked(E, range) * (kBlock / kWave) * rep_dd_walkers(range);
        const int st = plan_task(A[k], R[k], C[k], ld[k], rm.data(), cm.data(), choose_L(S, S, walkers, false),
                                 plans[(size_t)k], true);
        if (st) return n > 1 ? fail(st, "matrix " + std::to_string(k) + ": " + t_err) : st;
    }
    std::vector<double> out6;
    const int st = run_tasks_locked(E, plans, out6, true);
    if (st) return st;
    for (int64_t k = 0; k < n; ++k) {
        out[2 * k] = out6[6 * k];
        out[2 * k + 1] = out6[6 * k + 1];
    }
    return PQB_OK;
}

int run_plain_dd_locked(Engine& E, const double* A, int R, int C, int64_t ld, double* out) {
    return run_plain_dd_list_locked(E, &A, &R, &C, &ld, 1, out);
}

int batch_replay(int iters, double* ms) {
    Engine* last = g_last_batch.load();
    if (!last) return fail(PQB_ERR_ARG, "no batch has been run yet");
    Lease L(*last);
    if (L.status()) return L.status();
    Engine& E = *L;
    BatchBuffers& B = batch_of(E);
    if (B.last_range < 0 || !B.graph[B.last_range]) return fail(PQB_ERR_ARG, "no batch has been run yet");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, E.stream);
    for (int i = 0; e == hipSuccess && i < iters; ++i) e = hipGraphLaunch(B.graph[B.last_range], E.stream);
    if (e == hipSuccess) e = hipEventRecord(e1, E.stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float f = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(e, "batch replay");
    *ms = f;
    return PQB_OK;
}

}  // namespace pqb

using namespace pqb;

extern "C" {

int pqb_perm_batch(const pqb_task* tasks, int64_t ntasks, int algo, int precision, double* out) {
    try {
        return batch_entry(tasks, ntasks, algo, precision, out, 0);
    } catch (const std::exception& ex) {
        return fail(PQB_ERR_DEVICE, std::string("internal error: ") + ex.what());
    }
}

int pqb_perm_batch_tuned(const pqb_task* tasks, int64_t ntasks, int algo, int precision, int64_t walk_length,
                         double* out) {
    try {
        return batch_entry(tasks, ntasks, algo, precision, out, walk_length);
    } catch (const std::exception& ex) {
        return fail(PQB_ERR_DEVICE, std::string("internal error: ") + ex.what());
    }
}

int pqb_perm_batch_states(const double* U, int64_t rows, int64_t cols, int64_t ld, const int64_t* inputs,
                          const int64_t* outputs, int64_t count, int algo, int precision, double* out) {
    try {
        if (count < 0 || rows < 0 || cols < 0) return fail(PQB_ERR_ARG, "negative size");
        if (count == 0) return PQB_OK;
        if (!out || (rows > 0 && !outputs) || (cols > 0 && !inputs)) return fail(PQB_ERR_ARG, "NULL pointer");
        if (ld < cols) return fail(PQB_ERR_ARG, "leading dimension smaller than the number of columns");
        // one task per pair, all pointing into U and the two state arrays
        std::vector<pqb_task> t((size_t)count);
        for (int64_t i = 0; i < count; ++i)
            t[(size_t)i] = pqb_task{U, rows, cols, ld, rows > 0 ? outputs + i * rows : nullptr,
                                    cols > 0 ? inputs + i * cols : nullptr};
        return batch_entry(t.data(), count, algo, precision, out, 0);
    } catch (const std::exception& ex) {
        return fail(PQB_ERR_DEVICE, std::string("internal error: ") + ex.what());
    }
}

int pqb_set_walk_lengths(const int64_t* walk, int64_t count) {
    for (int i = 0; i <= PQB_MAX_COLS; ++i) g_walk[i].store((walk && i < count && walk[i] > 0) ? walk[i] : 0);
    return PQB_OK;
}

int pqb_batch_replay(int iters, double* ms) {
    if (!ms || iters < 1) return fail(PQB_ERR_ARG, "bad replay arguments");
    return batch_replay(iters, ms);
}

int pqb_repeated_plan(const double* A, int64_t rows, int64_t cols, int64_t ld, const int64_t* row_mult,
                      const int64_t* col_mult, int64_t walk_length, int64_t* info) {
    if (!info || walk_length < 1) return fail(PQB_ERR_ARG, "bad arguments");
    pqb_task t{A, rows, cols, ld, row_mult, col_mult};
    bool decided = false;
    double v;
    int st = check_task(t, PQB_ALGO_BBFG, &decided, &v);
    if (st) return st;
    if (decided) {
        info[0] = info[1] = info[2] = info[3] = info[4] = info[5] = 0;
        return PQB_OK;
    }
    TaskPlan P;
    st = plan_task(A, rows, cols, ld, row_mult, col_mult, (uint64_t)walk_length, P);
    if (st) return st;
    info[0] = P.NS;
    info[1] = P.R;
    info[2] = P.nlow;
    info[3] = P.nhigh;
    info[4] = (int64_t)P.L;
    info[5] = (int64_t)P.nchunks;
    return PQB_OK;
}

}  // extern "C"
