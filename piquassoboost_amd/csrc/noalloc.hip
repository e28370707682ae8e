// Plain permanent on device pointers with no workspace argument and no
// allocation: pqb_perm_device_noalloc (include/pqb_perm.h).
//
// The workspace of pqb_perm_device (grid-reduction counter, one double-double
// partial per block, the +-2A table) lives here in a file-scope device array,
// `scratch`, sized for the largest matrix the kernels take (41 x 40) and the
// largest launch plan (kMaxBlocks blocks).  Each GPU holds its own copy of the
// array for the life of the process; its device address is looked up once per
// device (hipGetSymbolAddress, with that device current) and only that address
// is handed to kernels -- in host code the array's name is not a device
// address.  Every call on one GPU shares the array: calls that overlap on one
// GPU (two streams, two threads) race and must be ordered by the caller.
//
// Nothing here allocates: no hipMalloc / hipFree, no stream or event creation;
// the occupancy query and device attributes are read once per device.  The
// counter is cleared by the first kernel of every call (a plain store, not a
// hipMemsetAsync), so a call captured into a HIP graph replays correctly.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "pqb_device.h"
#include "pqb_engine.h"
#include "pqb_kernels.h"
#include "pqb_perm.h"

namespace pqb {

namespace {

constexpr size_t kScratchCounter = 256;
constexpr size_t kScratchPartials = (size_t)kMaxBlocks * sizeof(ddc);
constexpr size_t kScratchTable = (size_t)2 * (PQB_MAX_COLS + 1) * PQB_MAX_COLS * sizeof(double2);
constexpr size_t kScratchBytes = kScratchCounter + kScratchPartials + kScratchTable;
constexpr int kMaxDevices = 64;

static_assert(kScratchCounter % 256 == 0 && kScratchPartials % 256 == 0, "scratch sections stay 256-byte aligned");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

struct ScratchDevice {
    bool ready = false;
    char* ptr = nullptr;  // device address of this GPU's `scratch`
    int cus = 0;
    int blocks_per_cu[PQB_MAX_COLS + 1] = {0};
};

std::mutex g_scratch_mu;
ScratchDevice g_scratch_dev[kMaxDevices];

// M2 = [+2A ; -2A] of one packed R x C matrix (n = R C elements), and the
// grid-reduction counter set to zero before the sweep reads it.
__global__ void noalloc_prep_kernel(const double2* __restrict__ A, double2* __restrict__ M2, int n,
                                    unsigned* __restrict__ counter) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i == 0) *counter = 0u;
    if (i < n) {
        const double2 v = A[i];
        M2[i] = make_double2(2.0 * v.x, 2.0 * v.y);
        M2[n + i] = make_double2(-2.0 * v.x, -2.0 * v.y);
    }
}

// Set-up of the current device, once: the device address of `scratch`, the CU
// count and (per column count, on first use) the resident blocks of the sweep.
int scratch_device(int nc, ScratchDevice** out, int64_t* waves_resident) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(PQB_ERR_DEVICE, "device index outside the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ScratchDevice& s = g_scratch_dev[dev];
    if (!s.ready) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(PQB_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned NULL");
        int cus = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return hip_fail(e, "hipDeviceGetAttribute");
        s.ptr = static_cast<char*>(p);
        s.cus = cus;
        s.ready = true;
    }
    if (s.blocks_per_cu[nc] == 0) {
        int nb = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, plain_kernel_fn(nc), kBlock, 0);
        if (e != hipSuccess || nb <= 0) nb = 1;
        s.blocks_per_cu[nc] = nb;
    }
    *waves_resident = (int64_t)s.blocks_per_cu[nc] * (kBlock / kWave) * s.cus;
    *out = &s;
    return PQB_OK;
}

}  // namespace

static int noalloc_entry(const double* dA, int64_t rows, int64_t cols, double* d_out, hipStream_t st) {
    if (!dA || !d_out) return fail(PQB_ERR_ARG, "NULL device pointer");
    if (rows < 1 || cols < 1 || rows >= cols + 2)
        return fail(PQB_ERR_ARG, "the device entry points need 1 <= rows <= cols + 1");
    if (cols > PQB_MAX_COLS)
        return fail(PQB_ERR_SIZE, "matrix has more than " + std::to_string(PQB_MAX_COLS) +
                                      " columns: the static scratch buffer holds at most 41 x " +
                                      std::to_string(PQB_MAX_COLS));
    const int R = (int)rows, C = (int)cols;
    ScratchDevice* s = nullptr;
    int64_t wr = 0;
    int rc = scratch_device(C, &s, &wr);
    if (rc) return rc;
    // the launch plan of pqb_perm_device for one matrix: the values agree bit for bit
    const PlainPlan pl = make_plain_plan(R, C, std::max<int64_t>(1, wr));
    if (pl.blocks > kMaxBlocks)
        return fail(PQB_ERR_SIZE, "launch plan exceeds the static scratch buffer (" + std::to_string(kMaxBlocks) +
                                      " partials)");
    const size_t table = (size_t)2 * R * C * sizeof(double2);
    if (table > kScratchTable) return fail(PQB_ERR_SIZE, "matrix exceeds the static scratch buffer");
    unsigned* counter = reinterpret_cast<unsigned*>(s->ptr);
    ddc* partials = reinterpret_cast<ddc*>(s->ptr + kScratchCounter);
    double2* m2 = reinterpret_cast<double2*>(s->ptr + kScratchCounter + kScratchPartials);
    const int n = R * C;
    const int bd = 256;
    hipLaunchKernelGGL(noalloc_prep_kernel, dim3((unsigned)((n + bd - 1) / bd)), dim3(bd), 0, st,
                       reinterpret_cast<const double2*>(dA), m2, n, counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "prep launch");
    PlainParams p;
    p.M2 = m2;
    p.m2_stride = (int64_t)2 * n;
    p.R = R;
    p.K = pl.K;
    p.B = pl.B;
    p.H = pl.H;
    p.partials = partials;
    p.counter = counter;
    p.out = d_out;
    p.scale = std::ldexp(1.0, -(R - 1));
    e = launch_plain(C, p, pl.blocks, pl.block, 1, st);
    if (e != hipSuccess) return hip_fail(e, "plain sweep launch");
    return PQB_OK;
}

}  // namespace pqb

extern "C" {

size_t pqb_perm_device_noalloc_scratch_bytes(void) { return pqb::kScratchBytes; }

int pqb_perm_device_noalloc(const double* dA, int64_t rows, int64_t cols, double* d_out, void* stream) {
    try {
        return pqb::noalloc_entry(dA, rows, cols, d_out, static_cast<hipStream_t>(stream));
    } catch (const std::exception& ex) {
        return pqb::fail(PQB_ERR_DEVICE, std::string("internal error: ") + ex.what());
    }
}

}  // extern "C"
