// device_scope.hpp -- the device an ABI call works on.
#pragma once

#include <hip/hip_runtime.h>

namespace rdamd {

// Makes `device` current for the scope of one ABI call and restores the
// caller's current device afterwards (device < 0: no change).  Handles are
// bound to the device they were created on; without this a call from a thread
// whose current device differs would allocate on the wrong GPU.
//
// It also clears the runtime's last-error slot: every HIP call of the library
// checks its own return code, and the launch checks (hipGetLastError after a
// kernel launch) must report that launch, not an error some earlier, already
// reported or deliberately ignored call left in the slot.
struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int device)
    {
        (void)hipGetLastError();
        if (device < 0) return;
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess || cur == device) return;
        if (hipSetDevice(device) == hipSuccess) prev = cur;
    }
    ~DeviceScope()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
    DeviceScope(const DeviceScope&) = delete;
    DeviceScope& operator=(const DeviceScope&) = delete;
};

}  // namespace rdamd
