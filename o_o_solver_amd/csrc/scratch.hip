// The device memory of oosg_solve_refined_device_noalloc: a file-scope array, so that the entry point needs no
// allocation (callers under HIP graph capture, serving loops, harnesses that forbid hipMalloc). Each device that loads
// the library holds its own copy of SCRATCH_BYTES until the process ends; calls that overlap on one device share it and
// race.
//
// In host code `scratch` is a stand-in symbol, not a device address: the address of a device's copy comes from
// hipGetSymbolAddress with that device current, once per device, and only that address is given to kernels and copies.

#include "kernels.h"

#include <map>
#include <mutex>

static __device__ __align__(256) unsigned char scratch[oosg::SCRATCH_BYTES];

namespace oosg
{

hipError_t scratchAddress(int device, void** out)
{
    static std::mutex lock;
    static std::map<int, void*> addresses; // per device, the address of its copy of `scratch`
    std::lock_guard<std::mutex> guard(lock);
    const auto known = addresses.find(device);
    if (known != addresses.end())
    {
        *out = known->second;
        return hipSuccess;
    }
    int current = -1;
    hipError_t e = hipGetDevice(&current);
    if (e != hipSuccess)
        return e;
    if (current != device)
        return hipErrorInvalidDevice;
    void* address = nullptr;
    e = hipGetSymbolAddress(&address, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
        return e;
    addresses[device] = address;
    *out = address;
    return hipSuccess;
}

} // namespace oosg
