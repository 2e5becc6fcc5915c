// mqe_noalloc.hip -- mqe_overlaps for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  The workspace of the transfer-matrix overlaps is a file-scope
// device array of fixed size; each GPU holds its own copy at its own address,
// looked up once per device with hipGetSymbolAddress.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mqe.h"
#include "mqe_internal.h"

namespace {

constexpr long SCRATCH_BYTES = 64L << 20;  // 64 MiB per GPU, held until the process ends
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) char scratch[SCRATCH_BYTES];

std::mutex g_lock;
char *g_ptr[MAX_DEVICES] = {nullptr};

// the device address of this GPU's copy of `scratch` (the array's name is only
// a host-side stand-in and must not be passed to a kernel)
int scratch_ptr(char **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        mqe::set_error("mqe_overlaps_noalloc: no usable current device");
        return MQE_EDEVICE;
    }
    std::lock_guard<std::mutex> g(g_lock);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            mqe::set_error("mqe_overlaps_noalloc: hipGetSymbolAddress of the workspace failed");
            return MQE_EDEVICE;
        }
        g_ptr[dev] = (char *)p;
    }
    *out = g_ptr[dev];
    return MQE_OK;
}

}  // namespace

extern "C" int mqe_overlaps_noalloc(mqe_engine *e, const mqe_mps *other, double *out) {
    long need_min = 0;
    int r = mqe::overlaps_workspace(e, other, nullptr, &need_min);
    if (r != MQE_OK) return r;
    if (need_min > SCRATCH_BYTES) {
        mqe::set_error("mqe_overlaps_noalloc: the other MPS and one walker's transfer matrices exceed the "
                       "64 MiB static workspace; use mqe_overlaps");
        return MQE_ECAPACITY;
    }
    char *ws = nullptr;
    r = scratch_ptr(&ws);
    if (r != MQE_OK) return r;
    return mqe::overlaps_in_workspace(e, other, out, ws, SCRATCH_BYTES);
}
