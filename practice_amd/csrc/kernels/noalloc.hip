// Device scratch of the allocation-free execute (cq_db_execute_noalloc).
//
// No reference counterpart: callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc) run an aggregate query with every scratch buffer -- hash
// tables, aggregate slots, materialization and sort buffers -- carved out of
// one file-scope device array of CQ_NOALLOC_SCRATCH_BYTES.  Each GPU holds its
// own copy of the array (at its own address) until the process ends; every
// call on a GPU shares it, so calls that overlap on one GPU race (one
// database serializes its own calls; two databases on one GPU must not call
// it at the same time).  The kernels zero what must start at zero (fill
// kernels, before each use).
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "noalloc.h"

static __device__ __align__(256) unsigned char scratch[CQ_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};  // per device: the address of that device's copy of `scratch`

}  // namespace

namespace db {

void* noAllocScratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("allocation-free scratch: device out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess || cur != device) {
            if (hipSetDevice(device) != hipSuccess) throw std::runtime_error("allocation-free scratch: hipSetDevice failed");
        }
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("allocation-free scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_ptr[device] = p;
    }
    return g_ptr[device];
}

}  // namespace db
