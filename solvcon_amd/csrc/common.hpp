#pragma once
/*
 * Shared plumbing of the HIP library: status codes and the thread-local
 * error message of the C ABI, HIP error checks, and launch helpers.
 *
 * Status codes map to the exception types the reference raises through
 * pybind11 (cpp/solvcon/multidim/pymod/wrap_multidim.cpp:42-67):
 *   SC_OK 0, SC_EVALUE 1 (ValueError), SC_EINDEX 2 (IndexError),
 *   SC_ERUNTIME 3 (RuntimeError), SC_EDEVICE 4 (a HIP failure).
 */

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace scamd
{

enum Status : int
{
    SC_OK = 0,
    SC_EVALUE = 1,
    SC_EINDEX = 2,
    SC_ERUNTIME = 3,
    SC_EDEVICE = 4,
};

struct Error : std::runtime_error
{
    int code;
    Error(int c, std::string const & m)
        : std::runtime_error(m)
        , code(c)
    {
    }
};

void set_last_error(int code, std::string const & msg);

[[noreturn]] inline void fail(int code, std::string const & msg) { throw Error(code, msg); }

inline void hip_check(hipError_t e, char const * what)
{
    if (e != hipSuccess)
    {
        fail(SC_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
    }
}

#define SC_HIP(call) ::scamd::hip_check((call), #call)

// Run f and turn exceptions into a status code plus the last-error message.
template <typename F>
int guard(F && f)
{
    try
    {
        f();
        return SC_OK;
    }
    catch (Error const & e)
    {
        set_last_error(e.code, e.what());
        return e.code;
    }
    catch (std::bad_alloc const &)
    {
        set_last_error(SC_ERUNTIME, "host allocation failed");
        return SC_ERUNTIME;
    }
    catch (std::exception const & e)
    {
        set_last_error(SC_ERUNTIME, e.what());
        return SC_ERUNTIME;
    }
}

// The stream that device allocations of the calling thread are ordered on:
// a handle sets it (AllocScope) while it works, so its buffers come from
// the stream-ordered pool (hipMallocAsync / hipFreeAsync) and freeing one
// never waits for the work of other handles' streams, as hipFree would.
inline hipStream_t & alloc_stream()
{
    static thread_local hipStream_t s = nullptr;
    return s;
}

struct AllocScope
{
    hipStream_t prev;
    explicit AllocScope(hipStream_t s)
        : prev(alloc_stream())
    {
        alloc_stream() = s;
    }
    ~AllocScope() { alloc_stream() = prev; }
    AllocScope(AllocScope const &) = delete;
    AllocScope & operator=(AllocScope const &) = delete;
};

// Device buffer owned by a handle; freed on the stream it was allocated on
// (or with hipFree when it was allocated outside any AllocScope).
template <typename T>
struct DevBuf
{
    T * ptr = nullptr;
    int64_t count = 0;
    hipStream_t stream = nullptr;

    DevBuf() = default;
    DevBuf(DevBuf const &) = delete;
    DevBuf & operator=(DevBuf const &) = delete;
    DevBuf(DevBuf && o) noexcept
        : ptr(o.ptr)
        , count(o.count)
        , stream(o.stream)
    {
        o.ptr = nullptr;
        o.count = 0;
        o.stream = nullptr;
    }
    DevBuf & operator=(DevBuf && o) noexcept
    {
        if (this != &o)
        {
            release();
            ptr = o.ptr;
            count = o.count;
            stream = o.stream;
            o.ptr = nullptr;
            o.count = 0;
            o.stream = nullptr;
        }
        return *this;
    }
    ~DevBuf() { release(); }

    void alloc(int64_t n)
    {
        release();
        if (n > 0)
        {
            stream = alloc_stream();
            size_t const nb = sizeof(T) * static_cast<size_t>(n);
            if (stream != nullptr)
            {
                SC_HIP(hipMallocAsync(reinterpret_cast<void **>(&ptr), nb, stream));
            }
            else
            {
                SC_HIP(hipMalloc(&ptr, nb));
            }
        }
        count = n;
    }
    void release()
    {
        if (ptr != nullptr)
        {
            if (stream != nullptr)
            {
                (void)hipFreeAsync(ptr, stream);
            }
            else
            {
                (void)hipFree(ptr);
            }
            ptr = nullptr;
        }
        count = 0;
        stream = nullptr;
    }
    size_t bytes() const { return sizeof(T) * static_cast<size_t>(count); }
};

// A non-blocking stream at the default priority, synchronised and destroyed
// last: a handle declares it before its buffers, so they are freed on it
// while it still exists.
struct StreamHolder
{
    hipStream_t s = nullptr;
    StreamHolder() { SC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~StreamHolder()
    {
        if (s != nullptr)
        {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
    StreamHolder(StreamHolder const &) = delete;
    StreamHolder & operator=(StreamHolder const &) = delete;
};

inline unsigned grid_for(int64_t n, unsigned block)
{
    int64_t const g = (n + block - 1) / block;
    return static_cast<unsigned>(g < 1 ? 1 : g);
}

// gfx950 dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with its own L2.  Remap the hardware block id so that every XCD works on
// one contiguous range of blocks: neighbour gathers of adjacent cells then
// hit the same L2.
__device__ inline unsigned xcd_block(unsigned b, unsigned nblk)
{
    constexpr unsigned NXCD = 8;
    unsigned const full = (nblk / NXCD) * NXCD;
    if (b >= full)
    {
        return b;
    }
    unsigned const per = nblk / NXCD;
    return (b % NXCD) * per + b / NXCD;
}

} /* end namespace scamd */
