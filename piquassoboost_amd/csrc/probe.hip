// FP64 issue-rate probe (measurement only): pqb_probe_fp64_fma.
//
// Every lane runs 16 independent chains of v_fma_f64 (no loads in the loop),
// eight dependent FMAs per chain per loop iteration (a loop body of 128 FMAs:
// with one FMA per chain per iteration the loop's branch and counter cost a
// third of a single wave's issue slots), so the only limits are the FP64
// pipe's issue rate and the clock.  Launched
// as exactly `waves_per_simd` waves on every SIMD of the device (blocks of 4
// waves, `waves_per_simd` blocks per CU), timed with HIP events on the engine
// stream.  The achieved rate is what any FP64 VALU kernel can reach on this
// device at that occupancy: the permanent sweeps are compared with it.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "pqb_device.h"
#include "pqb_engine.h"
#include "pqb_perm.h"

namespace pqb {

constexpr int kProbeChains = 16;

constexpr int kProbeDepth = 8;  // dependent FMAs per chain per loop iteration

__global__ __launch_bounds__(256) void fp64_fma_probe_kernel(double a, double b, int iters, double* out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double x[kProbeChains];
#pragma unroll
    for (int c = 0; c < kProbeChains; ++c) x[c] = 1.0 + 1e-3 * (double)(threadIdx.x & 7) + 1e-4 * c;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < kProbeDepth; ++u) {
#pragma unroll
            for (int c = 0; c < kProbeChains; ++c) x[c] = fma(x[c], a, b);
        }
    }
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < kProbeChains; ++c) s += x[c];
    out[g] = s;
}

static int probe_entry(int waves_per_simd, int64_t iters, double* ms, double* tflops) {
    if (waves_per_simd < 1 || waves_per_simd > 8) return fail(PQB_ERR_ARG, "waves_per_simd must be 1..8");
    if (iters < 1 || iters > ((int64_t)1 << 30)) return fail(PQB_ERR_ARG, "iters must be 1..2^30");
    if (!ms || !tflops) return fail(PQB_ERR_ARG, "NULL output pointer");
    Lease L;
    int st = L.status();
    if (st) return st;
    Engine& E = *L;
    const int blocks = E.cus * waves_per_simd;  // 4 waves per block: one per SIMD
    const size_t out_bytes = (size_t)blocks * kBlock * sizeof(double);
    st = E.ensure_buffers_locked(kCounterBytes + out_bytes, 0);
    if (st) return st;
    double* dout = reinterpret_cast<double*>(static_cast<char*>(E.dbuf) + kCounterBytes);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // one untimed launch (code load, clock ramp), then the timed one
    for (int rep = 0; e == hipSuccess && rep < 2; ++rep) {
        if (rep == 1) e = hipEventRecord(e0, E.stream);
        if (e != hipSuccess) break;
        hipLaunchKernelGGL(fp64_fma_probe_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, E.stream,
                           1.0 - 1e-9, 1e-9, (int)iters, dout);
        e = hipGetLastError();
        if (e == hipSuccess && rep == 1) e = hipEventRecord(e1, E.stream);
    }
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float f = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(e, "fp64 probe");
    *ms = f;
    const double flops = (double)blocks * kBlock * (double)iters * kProbeDepth * kProbeChains * 2.0;
    *tflops = flops / ((double)f * 1e-3) / 1e12;
    return PQB_OK;
}

// One lane writes the shader-clock counter (s_memtime) and the 100 MHz
// real-time counter (s_memrealtime).  A stamp before and one after some work on
// the same stream give the average shader clock over that work.  Block 0 of a
// launch goes to the same XCD every time, so both stamps read one counter.
__global__ void clock_stamp_kernel(unsigned long long* dst) {
    if (threadIdx.x == 0) {
        const unsigned long long c = __builtin_amdgcn_s_memtime();
        const unsigned long long r = __builtin_amdgcn_s_memrealtime();
        dst[0] = c;
        dst[1] = r;
    }
}

}  // namespace pqb

extern "C" int pqb_clock_stamp(unsigned long long* d_dst, void* stream) {
    if (!d_dst) return pqb::fail(PQB_ERR_ARG, "NULL device pointer");
    hipLaunchKernelGGL(pqb::clock_stamp_kernel, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), d_dst);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? PQB_OK : pqb::hip_fail(e, "clock stamp launch");
}

extern "C" int pqb_probe_fp64_fma(int waves_per_simd, int64_t iters, double* ms, double* tflops) {
    try {
        return pqb::probe_entry(waves_per_simd, iters, ms, tflops);
    } catch (const std::exception& ex) {
        return pqb::fail(PQB_ERR_DEVICE, std::string("internal error: ") + ex.what());
    }
}
