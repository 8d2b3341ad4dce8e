// hx_noalloc.hip -- the sorted trace for callers that cannot allocate device
// memory (code under hipGraph capture, serving loops, harnesses that forbid
// hipMalloc).
//
// hx_trace_sorted keeps its ray keys, sorted ray order, radix-sort status words
// and trace counters in context buffers that grow with the largest call.  Here
// they live in one fixed per-device array of kScratchBytes instead, so a call
// never allocates: at most kMaxRays rays per call.  The array's device address
// is looked up once per device when a context is created on it (the device is
// current then); every later use goes through that address.  All calls on one
// GPU share its array, so calls that overlap on one GPU (two streams, two host
// threads) race; each call clears what must start at zero with fill kernels.
#include <mutex>

#include "hx_internal.h"

namespace hx {

namespace {

constexpr int64_t kMaxRays = int64_t(1) << 20;
static_assert(kMaxRays == HX_NOALLOC_MAX_RAYS, "the header states the same maximum");
constexpr size_t kKeyBytes = 8 * (size_t)kMaxRays;     // one u64 key buffer
constexpr size_t kOrderBytes = 4 * (size_t)kMaxRays;   // one int32 order buffer
constexpr size_t kControlBytes = 64 * 1024;             // trace counters (2 int32), padded
constexpr size_t kStatusBytes = 8u << 20;               // radix-sort status words
constexpr size_t kScratchBytes = 2 * kKeyBytes + 2 * kOrderBytes + kControlBytes + kStatusBytes;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device %d: the allocation-free scratch supports devices 0..%d", device, kMaxDevices - 1);
        return HX_ERR_INVALID_ARG;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ptr[device]) return HX_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();  // do not leave the failure for the next launch check
        return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    }
    g_ptr[device] = p;
    return HX_OK;
}

int noalloc_trace_scratch(int device, int64_t n_rays, TraceScratch* s) {
    if (n_rays > kMaxRays) {
        set_error("hx_trace_sorted_noalloc: %lld rays exceed its fixed scratch of %lld rays", (long long)n_rays,
                  (long long)kMaxRays);
        return HX_ERR_INVALID_ARG;
    }
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (device >= 0 && device < kMaxDevices) base = g_ptr[device];
    }
    if (!base) {
        set_error("hx_trace_sorted_noalloc: no scratch address for device %d (context not created on it)", device);
        return HX_ERR_INVALID_ARG;
    }
    unsigned char* p = static_cast<unsigned char*>(base);
    s->keys[0] = p;
    s->keys[1] = p + kKeyBytes;
    s->order[0] = reinterpret_cast<int32_t*>(p + 2 * kKeyBytes);
    s->order[1] = reinterpret_cast<int32_t*>(p + 2 * kKeyBytes + kOrderBytes);
    s->control = reinterpret_cast<int32_t*>(p + 2 * kKeyBytes + 2 * kOrderBytes);
    s->sort.status = reinterpret_cast<uint32_t*>(p + 2 * kKeyBytes + 2 * kOrderBytes + kControlBytes);
    s->sort.status_words = (int64_t)(kStatusBytes / sizeof(uint32_t));
    return HX_OK;
}

}  // namespace hx
