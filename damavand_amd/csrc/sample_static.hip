// dmv_sample_static: dmv_sample without any device allocation, for callers
// that cannot allocate (code under stream capture, serving loops, harnesses
// that forbid hipMalloc).  dmv_sample allocates a device array for the k
// results on every call; here the results are staged in `scratch`, a fixed
// file-scope device array, and the rest of the work uses the state's own
// workspace (block masses and their prefix, allocated with the state).
#include <mutex>

#include "device.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace dmv {

// At most 2^20 samples per call: 8 MiB of int64 results, held on every GPU
// the library samples on until the process ends.
constexpr int64_t kStaticMaxSamples = int64_t(1) << 20;
static __device__ __align__(256) int64_t scratch[kStaticMaxSamples];

// Device address of `scratch` on each device, looked up on the first call made
// with that device current.  In host code the array's name is only a
// stand-in for the symbol; kernels and copies get this address instead.
constexpr int kMaxDevices = 64;
static std::mutex g_static_mu;  // also serialises the calls: they share `scratch`
static int64_t *g_scratch[kMaxDevices] = {};

static int scratch_address(int64_t **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error("hipGetDevice: %s", hipGetErrorString(e));
        return DMV_ERR_HIP;
    }
    DMV_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d devices this entry point tracks", dev,
                  kMaxDevices);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return DMV_ERR_HIP;
        }
        g_scratch[dev] = static_cast<int64_t *>(p);
    }
    *out = g_scratch[dev];
    return DMV_OK;
}

}  // namespace dmv

using namespace dmv;

extern "C" int dmv_sample_static(dmv_state *s, int64_t k, uint64_t seed, int64_t *out)
{
    DMV_CHECK_ARG(s != nullptr && s->amp != nullptr, "null state");
    DMV_CHECK_ARG(out != nullptr || k == 0, "null output");
    DMV_CHECK_ARG(k >= 0 && k <= kStaticMaxSamples,
                  "dmv_sample_static draws at most %lld samples per call, asked for %lld",
                  (long long)kStaticMaxSamples, (long long)k);
    if (k == 0)
        return DMV_OK;
    DMV_TRY(ensure_device());
    std::lock_guard<std::mutex> lk(g_static_mu);
    int64_t *d_out = nullptr;
    DMV_TRY(scratch_address(&d_out));
    // every result slot below k is written by the sampler: nothing to clear
    DMV_TRY(sample_to_device(s, k, seed, nullptr, d_out));
    return copy_d2h(out, d_out, (uint64_t)k * sizeof(int64_t));
}
