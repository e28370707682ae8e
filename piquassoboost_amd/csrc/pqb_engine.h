// Host-side engine state shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "pqb_perm.h"

namespace pqb {

constexpr int kBlock = 256;            // threads per block of the sweep kernels
constexpr size_t kCounterBytes = 256;  // grid-reduction counter slot at workspace offset 0
constexpr int kMaxBlocks = 4096;       // partials capacity of the device-API workspace

struct PlainPlan {
    int K = 0;            // Gray bits walked inside a chunk
    int H = 0;            // chunk-index bits
    int64_t nchunks = 1;  // 2^H
    int64_t lanes = 64;   // launched lanes
    int block = 64;       // threads per block
    int blocks = 1;
};

PlainPlan make_plain_plan(int R, int NC, int64_t lanes_resident);

struct Engine {
    std::mutex mu;
    int device = -1;
    hipStream_t stream = nullptr;
    int cus = 0;
    void* dbuf = nullptr;  // device workspace (counter slot at offset 0)
    size_t dbytes = 0;
    void* hbuf = nullptr;  // pinned staging
    size_t hbytes = 0;
    int blocks_per_cu[PQB_MAX_COLS + 1] = {0};

    int ensure_init_locked(int dev);
    int ensure_buffers_locked(size_t dneed, size_t hneed);
    int64_t resident_lanes_locked(int nc);
    void free_buffers();
};

extern Engine g_eng;
extern thread_local std::string t_err;

int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

}  // namespace pqb
