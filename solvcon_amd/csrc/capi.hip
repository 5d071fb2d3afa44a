/*
 * Library-wide C entry points: the thread-local last error of the C ABI and
 * device queries.
 */

#include "common.hpp"

#include <cstring>

namespace scamd
{
namespace
{
thread_local int g_code = 0;
thread_local std::string g_msg;
} /* end namespace */

void set_last_error(int code, std::string const & msg)
{
    g_code = code;
    g_msg = msg;
}

} /* end namespace scamd */

extern "C"
{

// Copy the last error message of this thread into buf; return its status
// code and clear it.
int sc_last_error(char * buf, int n)
{
    if (buf != nullptr && n > 0)
    {
        std::strncpy(buf, scamd::g_msg.c_str(), static_cast<size_t>(n) - 1);
        buf[n - 1] = '\0';
    }
    int const c = scamd::g_code;
    scamd::g_code = 0;
    scamd::g_msg.clear();
    return c;
}

// Number of visible GPUs (0 when none), without raising.
int sc_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
    {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

// Wait for all work on the current device.
int sc_device_sync(void)
{
    return scamd::guard([&]() { SC_HIP(hipDeviceSynchronize()); });
}

// Name and architecture of the current device.
int sc_device_name(char * buf, int n)
{
    return scamd::guard(
        [&]()
        {
            int dev = 0;
            SC_HIP(hipGetDevice(&dev));
            hipDeviceProp_t p;
            SC_HIP(hipGetDeviceProperties(&p, dev));
            std::string const s = std::string(p.name) + " " + p.gcnArchName;
            std::strncpy(buf, s.c_str(), static_cast<size_t>(n) - 1);
            buf[n - 1] = '\0';
        });
}

} /* end extern "C" */
