// Fixed device scratch block of the allocation-free entry point (mistra_tensor_mult_noalloc): the plan's
// device tables and the staged workspaces of its launches live here instead of in hipMalloc'd memory.
//
// `scratch` is a file-scope device array, so each GPU the library runs on holds its own copy for the
// life of the process.  Host code never uses the array's name as an address (in host code it names a
// host-side shadow, not device memory): the first use on a device, with that device current, asks the
// runtime for that device's address with hipGetSymbolAddress, and every later use goes through the
// pointer kept for the device.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>

#include "driver.hpp"

namespace mistra {
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;  // 64 MiB per device
}  // namespace mistra

static __device__ __align__(256) unsigned char scratch[mistra::kNoallocScratchBytes];

namespace mistra {

namespace {
std::mutex g_scratch_mutex;
std::map<int, void*> g_scratch_ptr;  // device -> its copy of `scratch`
}  // namespace

Status noalloc_scratch(void** ptr, size_t* bytes)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end())
    {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return Status::error(MISTRA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                                     (e != hipSuccess ? hipGetErrorString(e) : "null address"));
        it = g_scratch_ptr.emplace(dev, p).first;
    }
    *ptr = it->second;
    *bytes = kNoallocScratchBytes;
    return Status();
}

}  // namespace mistra
