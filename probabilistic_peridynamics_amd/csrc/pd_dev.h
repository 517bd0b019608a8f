// libpd_hip internals shared by the set-up entry points (neighbour list,
// stiffness corrections): scoped device buffers and HIP error returns.
#ifndef PD_DEV_H
#define PD_DEV_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/pd_hip.h"

int pd_fail(int code, const char *msg);

namespace pdsetup {

// A device array freed when it leaves scope.
template <typename T>
struct Dev {
    T *p = nullptr;
    Dev() = default;
    Dev(const Dev &) = delete;
    Dev &operator=(const Dev &) = delete;
    ~Dev()
    {
        if (p)
            (void)hipFree(p);
    }
    hipError_t alloc(size_t count)
    {
        return hipMalloc(reinterpret_cast<void **>(&p),
                         std::max<size_t>(count, 1) * sizeof(T));
    }
};

// Makes `device` current after checking that it exists.
inline int use_device(int device)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return pd_fail(PD_E_NODEV, "no HIP device is available");
    if (device < 0 || device >= ndev)
        return pd_fail(PD_E_ARG, "device index out of range");
    if (hipSetDevice(device) != hipSuccess)
        return pd_fail(PD_E_HIP, "hipSetDevice failed");
    return PD_OK;
}

}  // namespace pdsetup

// Returns PD_E_HIP with the call's text from the enclosing function when a
// HIP call fails.
#define PD_SETUP_CHK(call)                                                  \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess)                                               \
            return pd_fail(PD_E_HIP, (std::string(#call) + ": " +           \
                                      hipGetErrorString(e_)).c_str());     \
    } while (0)

#endif
