// nv_pclose_noalloc: the brute close-point search for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  It has no counterpart in the reference
// and no context: the caller passes device-resident inputs and outputs, and
// the search's working memory (tables, work list, segment spheres, chunk
// partials) lives in one static device array of this file, `scratch`.
//
// Each GPU holds its own copy of `scratch` for the life of the process.  In
// host code the array's name is only a host-side handle, so its device
// address is looked up once per device with hipGetSymbolAddress (first call
// on that device) and that address is what the kernels receive.  Every call
// on a GPU shares its array: calls that overlap on one GPU (two streams, two
// threads) race.  Nothing in the array needs to start at zero: each call
// writes its tables and work list first, and every partial the reduction
// reads is written by the same call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "../../include/nuvort.h"
#include "nv_kernels.h"

namespace {

constexpr size_t SCRATCH_BYTES = (size_t)64 << 20;   // NV_NOALLOC_SCRATCH_MIB MiB
static_assert(SCRATCH_BYTES == (size_t)NV_NOALLOC_SCRATCH_MIB << 20, "header and source agree");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
void *g_scratch[MAX_DEVICES] = {};
std::mutex g_mu;

// the device address of this device's `scratch` (looked up once per device)
hipError_t scratch_ptr(int dev, void **out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

}  // namespace

extern "C" int nv_pclose_noalloc(int32_t n, int32_t ld, const double *x, const int32_t *infront,
                                 const int32_t *behind, int32_t *closest, double *closestd, void *stream) {
    if (n < 0 || ld < n) return NV_ERR_ARG;
    if (n == 0) return NV_OK;
    if (!x || !infront || !behind || !closest || !closestd) return NV_ERR_ARG;
    if (n > NV_NOALLOC_MAX_POINTS || nvk_pclose_flat_bytes(n, 1) > SCRATCH_BYTES) return NV_ERR_CAPACITY;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return NV_ERR_HIP;
    int ncu = 256;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1) ncu = 256;
    // chunks per target as nv_pclose chooses them (about 16 work items per
    // CU), fewer when the partials would not fit
    const long ntb = ((long)n + 255) / 256;
    const long per = std::max<long>(4, ntb * ntb / (16L * ncu));
    int S = (int)std::max<long>(1, std::min<long>(ntb, (ntb + per - 1) / per));
    while (S > 1 && nvk_pclose_flat_bytes(n, S) > SCRATCH_BYTES) --S;
    void *work = nullptr;
    if (scratch_ptr(dev, &work) != hipSuccess) return NV_ERR_HIP;
    const hipError_t e = nvk_pclose_flat((hipStream_t)stream, x, ld, infront, behind, closest, closestd, n, S, work);
    return e == hipSuccess ? NV_OK : NV_ERR_HIP;
}
