// Fixed device workspace of the allocation-free forward entry point
// (rl_trace_lines_forward_noalloc in rl_api.cpp).
//
// The workspace is a file-scope device array: the runtime places one copy
// on every GPU the library's code object is loaded on, for the life of the
// process, so a call that uses it needs no device allocation.  In host
// code the array's name is only a stand-in for the symbol; its device
// address is taken once per device with hipGetSymbolAddress (from
// rl_create, with that device current) and only that address is passed to
// kernels and copies.  Every context on a GPU shares that GPU's copy:
// calls that overlap on one GPU race, and a call finds the previous call's
// contents (the entry point clears the counters it needs at zero).
#include <hip/hip_runtime.h>

#include <mutex>

#include "rl_engine.h"

namespace rl {

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};
}  // namespace

hipError_t scratch_attach(int device, void **out) {
    *out = nullptr;
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace rl
