// noalloc.hip -- ga_wfa_device_noalloc: the device-resident WFA batch for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
// reference.
//
// Every scratch buffer of the call -- the claim order, the plan's histogram
// and class slices, the claim counters and the wavefront stacks -- lives in
// `scratch`, one static device array per GPU (GA_WFA_NOALLOC_SCRATCH_BYTES,
// held by each GPU the library runs on until the process ends).  In host code
// the array's name is only a stand-in for the symbol: its device address is
// looked up with hipGetSymbolAddress once per device, with that device
// current, on the entry point's first call there, and only that address is
// handed to kernels.  The kernels zero what must start at zero (the plan
// kernels zero the histogram and the claim counters), so the array may keep
// the last call's contents.  Calls that overlap on one GPU share its array
// and race.
#include "ga_device.hpp"
#include "../../include/genoaligner/genoaligner_c.h"

#include <mutex>

namespace {

constexpr size_t kScratchBytes = size_t(GA_WFA_NOALLOC_SCRATCH_BYTES);
// claim order (4 bytes a pair) + 256-byte-aligned plan buffers
constexpr size_t kPlanBytes = size_t(GA_WFA_NOALLOC_MAX_PAIRS) * 4 + (size_t(64) << 10);
constexpr size_t kStackBytes = kScratchBytes - kPlanBytes;
static_assert(kPlanBytes < kScratchBytes, "scratch smaller than its plan region");

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t* g_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up on the first call there.
const char* scratch_on_current_device(uint8_t** out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return "hipGetDevice";
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return "hipGetSymbolAddress(scratch)";
        g_ptr[dev] = static_cast<uint8_t*>(p);
    }
    *out = g_ptr[dev];
    return nullptr;
}

}  // namespace

extern "C" int ga_wfa_device_noalloc(const ga_wfa_dev_batch* b, void* stream, const char** error)
{
    if (error) *error = nullptr;
    auto refuse = [&](const char* msg) {
        if (error) *error = msg;
        return GA_INVALID_ARGUMENT;
    };
    if (!b) return refuse("null batch");
    if (b->smax_max < 0 || b->smax_max > GA_MAX_SMAX_EXT) return refuse("smax_max out of range [0, 4000]");
    if (b->n_pairs > GA_WFA_NOALLOC_MAX_PAIRS)
        return refuse("more pairs than ga_wfa_device_noalloc takes (GA_WFA_NOALLOC_MAX_PAIRS)");
    if (b->cigar) {
        const size_t per = size_t(b->smax_max + 1) * size_t(b->smax_max + 1) * (b->max_len > 32766 ? 4 : 2);
        if (per > kStackBytes)
            return refuse("wavefront stack of smax_max larger than ga_wfa_device_noalloc's scratch");
    }
    if (b->n_pairs <= 0) return GA_OK;
    uint8_t* base = nullptr;
    if (const char* e = scratch_on_current_device(&base)) {
        if (error) *error = e;
        return GA_DEVICE_ERROR;
    }
    if (const char* e = ga::wfa_device_in_region(b, static_cast<hipStream_t>(stream), base, kPlanBytes,
                                                 kStackBytes)) {
        if (error) *error = e;
        return GA_DEVICE_ERROR;
    }
    return GA_OK;
}
