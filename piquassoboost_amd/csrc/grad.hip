// Host side of the adjoint sweep (perm_grad.hip): launch plan, workspace
// layout and the C-ABI entry points for the permanent together with its
// gradient G_ij = d perm(A) / d a_ij of square complex matrices, one matrix or
// a batch of equal-size matrices per call; and for the column minors of an
// n x (n+1) matrix (the same sweep without the row gradient).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "pqb_perm.h"
#include "pqb_device.h"
#include "pqb_kernels.h"
#include "pqb_engine.h"

namespace pqb {

struct GradPlan {
    bool minors = false;  // column-minors mode (n x (n+1) input)
    int cols = 1;         // columns of the input
    int M = 1;            // columns per lane (4 lanes per sign vector)
    int K = 0;            // register bits (walked inside a block)
    int L = 0;            // delta-lane bits
    int H = 0;            // high bits
    int64_t wpt = 4;      // waves per task
    int64_t blocks = 1;   // launched blocks (4 waves each)
    size_t lds = 0;       // dynamic LDS bytes per block
};

// Register bits per column count: the largest K whose accumulators and row
// data (4 M doubles each) still fit next to the column sums, running sums and
// suffix products in 256 VGPRs (the instantiations of perm_grad.hip); K = 4
// where it was measured faster than K = 3 (n = 17..20 and 25..28).
static int grad_register_bits(int M) {
    if (M == 5 || M == 7) return 4;
    return M <= 8 ? 3 : 2;
}

// Staged matrix (n x 4M) plus, per wave, the 16 x 4M transpose area and (adjoint
// mode) the H x 4M high-row accumulators.
static size_t grad_lds_bytes(int n, int M, int H, bool minors) {
    const size_t NP = 4 * (size_t)M;
    return sizeof(double2) * ((size_t)n * NP + 4 * (16 + (minors ? 0 : (size_t)H)) * NP);
}

static GradPlan grad_shape(int n, bool minors) {
    GradPlan g;
    g.minors = minors;
    g.cols = minors ? n + 1 : n;
    g.M = (g.cols + 3) / 4;
    g.L = std::min(4, n - 1);
    g.K = std::min(grad_register_bits(g.M), n - 1 - g.L);
    g.H = n - 1 - g.L - g.K;
    g.lds = grad_lds_bytes(n, g.M, g.H, minors);
    return g;
}

// Waves per task: about four rounds of the resident waves over the batch,
// at least 4 (one block), at most the number of blocks of 2^K steps.
static GradPlan make_grad_plan(int n, int64_t count, int64_t waves_resident, bool minors = false) {
    GradPlan g = grad_shape(n, minors);
    const int64_t NB = (int64_t)1 << g.H;
    const int64_t target = std::max<int64_t>(4 * std::max<int64_t>(waves_resident, 1), 4);
    int64_t per = (target + count - 1) / count;
    per = ((per + 3) / 4) * 4;
    const int64_t cap = std::max<int64_t>(4, ((NB + 3) / 4) * 4);
    g.wpt = std::max<int64_t>(4, std::min(per, cap));
    g.blocks = count * g.wpt / 4;
    return g;
}

// One partial per block of 4 waves: n x n gradient + permanent, or the n+1 column minors.
static size_t grad_partial_elems(int n, bool minors) { return minors ? (size_t)n + 1 : (size_t)n * n; }

static size_t grad_partials_bytes(int n, int64_t count, int64_t wpt, bool minors = false) {
    return (size_t)count * (size_t)(wpt / 4) * grad_partial_elems(n, minors) * sizeof(double2);
}

// Per-shape LDS limit of the adjoint kernels (set once as a kernel attribute)
// and their occupancy: shared by every caller, guarded by this mutex.
static std::mutex g_grad_mu;

// Resident waves of the adjoint kernel for this shape on the bound device.
static int grad_resident_waves(const GradPlan& g, int device, int cus, int64_t* waves) {
    const void* fn = grad_kernel_fn(g.M, g.K, g.minors);
    if (!fn) return fail(PQB_ERR_SIZE, "no adjoint kernel for this size");
    std::lock_guard<std::mutex> lk(g_grad_mu);
    // dynamic LDS beyond the default: up to the CU's LDS minus the kernel's static LDS
    static size_t dyn_limit[11][4][2] = {};
    size_t& limit = dyn_limit[g.M][g.K][g.minors];
    if (limit == 0) {
        hipFuncAttributes fa;
        hipError_t e = hipFuncGetAttributes(&fa, fn);
        if (e != hipSuccess) return hip_fail(e, "hipFuncGetAttributes(adjoint kernel)");
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, device);
        if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
        const size_t cu_lds = prop.maxSharedMemoryPerMultiProcessor;
        if (cu_lds <= fa.sharedSizeBytes) return fail(PQB_ERR_DEVICE, "no LDS left for the adjoint kernel");
        const size_t lim = cu_lds - fa.sharedSizeBytes;
        e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim);
        if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(adjoint kernel)");
        limit = lim;
    }
    if (g.lds > limit) return fail(PQB_ERR_SIZE, "adjoint sweep needs more LDS than a compute unit has");
    int nb = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 256, g.lds);
    if (e != hipSuccess) return hip_fail(e, "occupancy query (adjoint kernel)");
    if (nb <= 0) return fail(PQB_ERR_SIZE, "adjoint kernel does not fit on a compute unit");
    *waves = (int64_t)nb * 4 * cus;
    return PQB_OK;
}

static int grad_plan(int n, int64_t count, GradPlan* out, bool minors = false) {
    int device = -1, cus = 0;
    int st = bind_device(-1, &device, &cus);
    if (st) return st;
    const GradPlan shape = grad_shape(n, minors);
    int64_t wr = 0;
    st = grad_resident_waves(shape, device, cus, &wr);
    if (st) return st;
    *out = make_grad_plan(n, count, wr, minors);
    return PQB_OK;
}

static int grad_check(int64_t count, int64_t n) {
    if (count < 1) return fail(PQB_ERR_ARG, "batch count must be positive");
    if (n < 1) return fail(PQB_ERR_ARG, "matrix order must be positive");
    if (n > PQB_MAX_COLS)
        return fail(PQB_ERR_SIZE, "matrix order above " + std::to_string(PQB_MAX_COLS));
    return PQB_OK;
}

// Launch on `st`: dA packed count x n x cols, outputs packed.  `work` holds the
// partials (grad_partials_bytes).  Adjoint: d_grad count x n x n and d_perm
// count; minors: d_grad count x (n+1), d_perm unused.
static int grad_launch(const GradPlan& g, const double2* dA, int n, int64_t count, double2* d_perm,
                       double2* d_grad, void* work, hipStream_t st) {
    if (g.blocks > 0x7fffffff) return fail(PQB_ERR_SIZE, "adjoint launch too large");
    GradParams p;
    p.A = dA;
    p.partials = static_cast<double2*>(work);
    p.n = n;
    p.cols = g.cols;
    p.L = g.L;
    p.H = g.H;
    p.wpt = (int)g.wpt;
    hipError_t e = launch_grad(g.M, g.K, g.minors, p, (int)g.blocks, g.lds, st);
    if (e != hipSuccess) return hip_fail(e, "adjoint sweep launch");
    GradReduceParams r;
    r.partials = p.partials;
    r.out = d_grad;
    r.tail = nullptr;
    r.main = g.minors ? n + 1 : n * n;
    r.nblk = (int)(g.wpt / 4);
    r.count = (int)count;
    r.scale = std::ldexp(1.0, -(n - 1));
    e = launch_grad_reduce(r, st);
    if (e != hipSuccess) return hip_fail(e, "adjoint reduction launch");
    if (!g.minors) {  // the permanent from row 0 of the gradient (Laplace expansion)
        e = launch_grad_perm(dA, d_grad, d_perm, n, (int)count, st);
        if (e != hipSuccess) return hip_fail(e, "permanent (Laplace row 0) launch");
    }
    return PQB_OK;
}

static int grad_host_entry(const double* A, int64_t count, int64_t n, int64_t ld, int64_t stride,
                           double* perm_out, double* grad_out) {
    if (!perm_out) return fail(PQB_ERR_ARG, "output pointer is NULL");
    if (count < 1) return fail(PQB_ERR_ARG, "batch count must be positive");
    if (n == 0) {  // perm of the empty matrix is 1; the gradient is empty
        for (int64_t b = 0; b < count; ++b) {
            perm_out[2 * b] = 1.0;
            perm_out[2 * b + 1] = 0.0;
        }
        return PQB_OK;
    }
    int st = grad_check(count, n);
    if (st) return st;
    if (!A || !grad_out) return fail(PQB_ERR_ARG, "matrix or gradient pointer is NULL");
    if (ld < n) return fail(PQB_ERR_ARG, "leading dimension smaller than the matrix order");
    if (count > 1 && stride < n * ld) return fail(PQB_ERR_ARG, "matrix stride smaller than n * ld");
    const int nn = (int)n;
    const size_t mat = (size_t)nn * nn;

    Lease L;
    st = L.status();
    if (st) return st;
    Engine& E = *L;
    GradPlan g;
    st = grad_plan(nn, count, &g);
    if (st) return st;
    const size_t a_bytes = (size_t)count * mat * sizeof(double2);
    const size_t out_bytes = (size_t)count * (mat + 1) * sizeof(double2);
    const size_t part_bytes = grad_partials_bytes(nn, count, g.wpt);
    const size_t a_off = kCounterBytes;  // offset 0 holds the plain sweep's grid counter
    const size_t out_off = a_off + ((a_bytes + 255) / 256) * 256;
    const size_t part_off = out_off + ((out_bytes + 255) / 256) * 256;
    st = E.ensure_buffers_locked(part_off + part_bytes, std::max(a_bytes, out_bytes));
    if (st) return st;
    double* h = static_cast<double*>(E.hbuf);
    for (int64_t b = 0; b < count; ++b)
        for (int i = 0; i < nn; ++i)
            std::memcpy(h + 2 * ((size_t)b * mat + (size_t)i * nn), A + 2 * (b * stride + i * ld),
                        (size_t)nn * 2 * sizeof(double));
    char* d = static_cast<char*>(E.dbuf);
    hipError_t e = hipMemcpyAsync(d + a_off, h, a_bytes, hipMemcpyHostToDevice, E.stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync");
    double2* dperm = reinterpret_cast<double2*>(d + out_off);
    double2* dgrad = dperm + count;
    st = grad_launch(g, reinterpret_cast<const double2*>(d + a_off), nn, count, dperm, dgrad, d + part_off,
                     E.stream);
    if (st) {
        (void)hipStreamSynchronize(E.stream);  // the upload may still read the staging buffer
        return st;
    }
    // the staging buffer is reused for the results once the upload has been consumed
    e = hipMemcpyAsync(h, d + out_off, out_bytes, hipMemcpyDeviceToHost, E.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(E.stream);
    if (e != hipSuccess) return hip_fail(e, "adjoint sweep");
    std::memcpy(perm_out, h, (size_t)count * 2 * sizeof(double));
    std::memcpy(grad_out, h + 2 * count, (size_t)count * mat * 2 * sizeof(double));
    return PQB_OK;
}

// Column minors of `count` rows x (rows+1) matrices (matrix b at A + b * stride
// elements): out[b * cols + j] = perm(A_b without column j).
static int minors_host_entry(const double* A, int64_t count, int64_t rows, int64_t cols, int64_t ld,
                             int64_t stride, double* out) {
    if (!out) return fail(PQB_ERR_ARG, "output pointer is NULL");
    if (count < 1) return fail(PQB_ERR_ARG, "batch count must be positive");
    if (rows < 0 || cols != rows + 1) return fail(PQB_ERR_ARG, "column minors need a rows x (rows + 1) matrix");
    if (rows == 0) {  // the only minor is the empty matrix
        for (int64_t b = 0; b < count; ++b) {
            out[2 * b] = 1.0;
            out[2 * b + 1] = 0.0;
        }
        return PQB_OK;
    }
    if (cols > PQB_MAX_COLS)
        return fail(PQB_ERR_SIZE, "matrix has more than " + std::to_string(PQB_MAX_COLS) + " columns");
    if (!A) return fail(PQB_ERR_ARG, "matrix pointer is NULL");
    if (ld < cols) return fail(PQB_ERR_ARG, "leading dimension smaller than the number of columns");
    if (count > 1 && stride < rows * ld) return fail(PQB_ERR_ARG, "matrix stride smaller than rows * ld");
    const int R = (int)rows, C = (int)cols;

    Lease L;
    int st = L.status();
    if (st) return st;
    Engine& E = *L;
    GradPlan g;
    st = grad_plan(R, count, &g, true);
    if (st) return st;
    const size_t mat = (size_t)R * C;
    const size_t a_bytes = (size_t)count * mat * sizeof(double2);
    const size_t out_bytes = (size_t)count * C * sizeof(double2);
    const size_t part_bytes = grad_partials_bytes(R, count, g.wpt, true);
    const size_t a_off = kCounterBytes;  // offset 0 holds the plain sweep's grid counter
    const size_t out_off = a_off + ((a_bytes + 255) / 256) * 256;
    const size_t part_off = out_off + ((out_bytes + 255) / 256) * 256;
    st = E.ensure_buffers_locked(part_off + part_bytes, std::max(a_bytes, out_bytes));
    if (st) return st;
    double* h = static_cast<double*>(E.hbuf);
    for (int64_t b = 0; b < count; ++b)
        for (int i = 0; i < R; ++i)
            std::memcpy(h + 2 * ((size_t)b * mat + (size_t)i * C), A + 2 * (b * stride + i * ld),
                        (size_t)C * 2 * sizeof(double));
    char* d = static_cast<char*>(E.dbuf);
    hipError_t e = hipMemcpyAsync(d + a_off, h, a_bytes, hipMemcpyHostToDevice, E.stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync");
    st = grad_launch(g, reinterpret_cast<const double2*>(d + a_off), R, count, nullptr,
                     reinterpret_cast<double2*>(d + out_off), d + part_off, E.stream);
    if (st) {
        (void)hipStreamSynchronize(E.stream);  // the upload may still read the staging buffer
        return st;
    }
    e = hipMemcpyAsync(h, d + out_off, out_bytes, hipMemcpyDeviceToHost, E.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(E.stream);
    if (e != hipSuccess) return hip_fail(e, "column-minors sweep");
    std::memcpy(out, h, out_bytes);
    return PQB_OK;
}

}  // namespace pqb

using namespace pqb;

extern "C" {

int pqb_perm_grad(const double* A, int64_t n, int64_t ld, double* perm_out, double* grad_out) {
    try {
        return grad_host_entry(A, 1, n, ld, n * ld, perm_out, grad_out);
    } catch (const std::exception& ex) {
        return fail(PQB_ERR_DEVICE, std::string("internal error: ") + ex.what());
    }
}

int pqb_perm_grad_batch(const double* A, int64_t count, int64_t n, int64_t ld, int64_t stride,
                        double* perm_out, double* grad_out) {
    try {
        return grad_host_entry(A, count, n, ld, stride, perm_out, grad_out);
    } catch (const std::exception& ex) {
        return fail(PQB_ERR_DEVICE, std::string("internal error: ") + ex.what());
    }
}

int pqb_perm_grad_workspace(int64_t count, int64_t n, size_t* bytes) {
    if (!bytes) return fail(PQB_ERR_ARG, "bytes pointer is NULL");
    int st = grad_check(count, n);
    if (st) return st;
    GradPlan g;
    st = grad_plan((int)n, count, &g);
    if (st) return st;
    *bytes = grad_partials_bytes((int)n, count, g.wpt);
    return PQB_OK;
}

int pqb_perm_grad_device(const double* dA, int64_t count, int64_t n, double* d_perm, double* d_grad,
                         void* d_work, size_t work_bytes, void* stream) {
    if (!dA || !d_perm || !d_grad || !d_work) return fail(PQB_ERR_ARG, "NULL device pointer");
    int st = grad_check(count, n);
    if (st) return st;
    GradPlan g;
    st = grad_plan((int)n, count, &g);
    if (st) return st;
    if (work_bytes < grad_partials_bytes((int)n, count, g.wpt)) return fail(PQB_ERR_ARG, "workspace too small");
    return grad_launch(g, reinterpret_cast<const double2*>(dA), (int)n, count, reinterpret_cast<double2*>(d_perm),
                       reinterpret_cast<double2*>(d_grad), d_work, static_cast<hipStream_t>(stream));
}

int pqb_grad_plan(int64_t count, int64_t n, int64_t waves_resident, int64_t* plan) {
    if (!plan) return fail(PQB_ERR_ARG, "plan pointer is NULL");
    int st = grad_check(count, n);
    if (st) return st;
    GradPlan g;
    if (waves_resident > 0) {
        g = make_grad_plan((int)n, count, waves_resident);
    } else {
        st = grad_plan((int)n, count, &g);
        if (st) return st;
    }
    plan[0] = g.M;
    plan[1] = g.K;
    plan[2] = g.L;
    plan[3] = g.H;
    plan[4] = g.wpt;
    plan[5] = g.blocks;
    plan[6] = (int64_t)g.lds;
    return PQB_OK;
}

int pqb_perm_minors(const double* A, int64_t rows, int64_t cols, int64_t ld, double* out) {
    try {
        return minors_host_entry(A, 1, rows, cols, ld, 0, out);
    } catch (const std::exception& ex) {
        return fail(PQB_ERR_DEVICE, std::string("internal error: ") + ex.what());
    }
}

int pqb_perm_minors_batch(const double* A, int64_t count, int64_t rows, int64_t cols, int64_t ld,
                          int64_t stride, double* out) {
    try {
        return minors_host_entry(A, count, rows, cols, ld, stride, out);
    } catch (const std::exception& ex) {
        return fail(PQB_ERR_DEVICE, std::string("internal error: ") + ex.what());
    }
}

}  // extern "C"
