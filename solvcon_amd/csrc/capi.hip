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

namespace
{

// Plain device copy used to measure achievable HBM bandwidth and to
// calibrate the HBM byte counters against a known number of bytes.
__global__ void k_stream_copy(double2 const * __restrict__ src, double2 * __restrict__ dst, int64_t n)
{
    int64_t const stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        dst[i] = src[i];
    }
}

// The same copy with four independent 16-byte transfers in flight per
// thread and non-temporal access (the data is touched once).
__global__ void k_stream_copy4(double2 const * __restrict__ src, double2 * __restrict__ dst, int64_t n)
{
    int64_t const base = static_cast<int64_t>(blockIdx.x) * blockDim.x * 4 + threadIdx.x;
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
    {
        int64_t const i = base + static_cast<int64_t>(u) * blockDim.x;
        if (i < n)
        {
            v[u].x = __builtin_nontemporal_load(&src[i].x);
            v[u].y = __builtin_nontemporal_load(&src[i].y);
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
    {
        int64_t const i = base + static_cast<int64_t>(u) * blockDim.x;
        if (i < n)
        {
            __builtin_nontemporal_store(v[u].x, &dst[i].x);
            __builtin_nontemporal_store(v[u].y, &dst[i].y);
        }
    }
}

} /* end namespace */

} /* end namespace scamd */

extern "C"
{

// Copy nbytes (a multiple of 16) between two device buffers reps times and
// return the average kernel time in milliseconds.
int sc_stream_copy(int64_t nbytes, int reps, double * ms_avg)
{
    return scamd::guard(
        [&]()
        {
            if (nbytes <= 0 || nbytes % 16 != 0 || reps < 1)
            {
                scamd::fail(scamd::SC_EVALUE, "sc_stream_copy: nbytes must be a positive multiple of 16");
            }
            int64_t const n = nbytes / 16;
            scamd::DevBuf<double2> a, b;
            a.alloc(n);
            b.alloc(n);
            SC_HIP(hipMemset(a.ptr, 0, a.bytes()));
            hipEvent_t e0, e1;
            SC_HIP(hipEventCreate(&e0));
            SC_HIP(hipEventCreate(&e1));
            unsigned const grid = 256 * 32;
            scamd::k_stream_copy<<<grid, 256>>>(a.ptr, b.ptr, n);
            SC_HIP(hipGetLastError());
            SC_HIP(hipEventRecord(e0, nullptr));
            for (int r = 0; r < reps; ++r)
            {
                scamd::k_stream_copy<<<grid, 256>>>(a.ptr, b.ptr, n);
            }
            SC_HIP(hipEventRecord(e1, nullptr));
            SC_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            SC_HIP(hipEventElapsedTime(&ms, e0, e1));
            // The unrolled non-temporal copy; the faster of the two counts.
            unsigned const grid4 = static_cast<unsigned>((n + 1023) / 1024);
            scamd::k_stream_copy4<<<grid4, 256>>>(a.ptr, b.ptr, n);
            SC_HIP(hipGetLastError());
            SC_HIP(hipEventRecord(e0, nullptr));
            for (int r = 0; r < reps; ++r)
            {
                scamd::k_stream_copy4<<<grid4, 256>>>(a.ptr, b.ptr, n);
            }
            SC_HIP(hipEventRecord(e1, nullptr));
            SC_HIP(hipEventSynchronize(e1));
            float ms4 = 0.0f;
            SC_HIP(hipEventElapsedTime(&ms4, e0, e1));
            ms = ms4 < ms ? ms4 : ms;
            *ms_avg = ms / reps;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        });
}

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
