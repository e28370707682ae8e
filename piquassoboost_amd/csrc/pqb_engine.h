// Host-side engine state shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "pqb_perm.h"

namespace pqb {

constexpr int kBlock = 256;            // threads per block of the sweep kernels
constexpr size_t kCounterBytes = 256;  // grid-reduction counter slot at workspace offset 0
constexpr int kMaxBlocks = 4096;       // partials capacity of the device-API workspace

struct PlainPlan {
    int K = 0;            // walk bits
    int B = 0;            // lane bits
    int H = 0;            // group bits
    int64_t waves = 1;    // launched waves
    int block = 64;       // threads per block
    int blocks = 1;
};

PlainPlan make_plain_plan(int R, int NC, int64_t waves_resident);

// One planned repeated-row permanent (see batch.hip).
struct TaskPlan {
    bool trivial = false;      // decided without a sweep: value (tre, tim)
    double tre = 1.0, tim = 0.0;
    int NS = 0;                // column slots
    int R = 0;                 // rows of A: anchor, low digits, high digits
    int N = 0;                 // photons
    int nlow = 0, nhigh = 0;
    uint64_t L = 1;            // walk length
    uint64_t nchunks = 1;
    double scale = 1.0;        // 2^-(N-1)
    std::vector<double> A;     // R x NS complex (interleaved)
    std::vector<double> base;  // NS complex
    std::vector<int> low_radix;
    std::vector<uint32_t> hrad;
};

int plan_task(const double* A, int64_t rows, int64_t cols, int64_t ld, const int64_t* rmult,
              const int64_t* cmult, uint64_t target_L, TaskPlan& P, bool plain = false);
int batch_entry(const pqb_task* tasks, int64_t ntasks, int algo, int precision, double* out,
                int64_t L_override);
struct Engine;

// Plain permanent (1 <= rows <= cols + 1) by the double-double sweep; context locked.
int run_plain_dd_locked(Engine& E, const double* A, int R, int C, int64_t ld, double* out);
// `n` such permanents (matrix k: A[k], R[k] x C[k], stride ld[k]), each planned as
// run_plain_dd_locked plans it, in one batched launch; out: 2 doubles each.
int run_plain_dd_list_locked(Engine& E, const double* const* A, const int* R, const int* C, const int64_t* ld,
                             int64_t n, double* out);
int list_entry(const double* const* A, const int64_t* rows, const int64_t* cols, const int64_t* ld,
               int64_t count, int algo, int precision, double* out);

// Process-wide device binding: the device every context runs on, its CU count
// and the occupancy of the plain sweep per column count (read-mostly).
struct DeviceState {
    std::mutex mu;
    int device = -1;
    int cus = 0;
    int blocks_per_cu[PQB_MAX_COLS + 1] = {0};
};
extern DeviceState g_dev;

// Binds the process to `dev` (dev < 0: keep the current binding, device 0 if
// none) and reports it; a change of device releases every context's resources.
int bind_device(int dev, int* device, int* cus);
// Resident waves of the plain sweep for `nc` columns on the bound device.
int64_t resident_waves(int nc);

struct BatchBuffers;  // persistent-batch buffers and captured graphs (batch.hip)

// One execution context: a stream, a device workspace, a pinned staging
// buffer and the batch buffers with their captured graphs.  A host entry point
// leases a context for the whole call (class Lease): callers on different
// threads run on different contexts -- their own streams -- and overlap; a
// context is never used by two calls at once.
struct Engine {
    std::mutex mu;
    int device = -1;        // device of the stream and buffers (-1: none yet)
    hipStream_t stream = nullptr;
    int cus = 0;
    void* dbuf = nullptr;   // device workspace (counter slot at offset 0)
    size_t dbytes = 0;
    void* hbuf = nullptr;   // pinned staging
    size_t hbytes = 0;
    BatchBuffers* batch = nullptr;

    // stream and device binding (first use, or after a device change); lock held
    int ensure_init_locked();
    int ensure_buffers_locked(size_t dneed, size_t hneed);
    int64_t resident_waves_locked(int nc) { return resident_waves(nc); }
    void release_locked();  // stream, buffers and graphs (device change)
};

constexpr int kEngineContexts = 4;
extern Engine g_ctx[kEngineContexts];

// The context of a host call: a free one (the calling thread's own first),
// waiting for the thread's own when all are busy.  Bound to the device and
// with its stream created on construction; status() is the result of that.
class Lease {
  public:
    Lease();
    explicit Lease(Engine& e);  // a given context (replay of its last batch)
    Engine& operator*() const { return *e_; }
    Engine* operator->() const { return e_; }
    int status() const { return st_; }

  private:
    Engine* e_ = nullptr;
    std::unique_lock<std::mutex> lk_;
    int st_ = 0;
};

extern thread_local std::string t_err;

int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
// hip_fail after waiting for the context's stream: for failures after an upload
// from the staging buffer was queued, so that no copy still reads it when the
// context is released.
int drain_fail(Engine& E, hipError_t e, const char* what);

}  // namespace pqb
