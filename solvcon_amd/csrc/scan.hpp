#pragma once
/*
 * Device exclusive prefix sums over int32/int64 used by the mesh builders
 * (face, ghost-node and ghost-face numbering).  Three phases: per-block
 * scans in LDS, a recursive scan of the block totals, and a uniform add.
 */

#include "common.hpp"

namespace scamd
{

constexpr unsigned SCAN_BLOCK = 256;
constexpr unsigned SCAN_ITEMS = 8; // elements per thread
constexpr unsigned SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

template <typename T>
__global__ void k_scan_tile(T const * in, T * out, T * sums, int64_t n)
{
    __shared__ T part[SCAN_BLOCK];
    int64_t const base = static_cast<int64_t>(blockIdx.x) * SCAN_TILE + static_cast<int64_t>(threadIdx.x) * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T acc = 0;
#pragma unroll
    for (unsigned k = 0; k < SCAN_ITEMS; ++k)
    {
        int64_t const i = base + k;
        T const x = i < n ? in[i] : T(0);
        v[k] = acc;
        acc += x;
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    // Hillis-Steele inclusive scan of the per-thread totals.
    for (unsigned off = 1; off < SCAN_BLOCK; off <<= 1)
    {
        T const add = threadIdx.x >= off ? part[threadIdx.x - off] : T(0);
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    T const prefix = threadIdx.x > 0 ? part[threadIdx.x - 1] : T(0);
#pragma unroll
    for (unsigned k = 0; k < SCAN_ITEMS; ++k)
    {
        int64_t const i = base + k;
        if (i < n)
        {
            out[i] = v[k] + prefix;
        }
    }
    if (threadIdx.x == SCAN_BLOCK - 1 && sums != nullptr)
    {
        sums[blockIdx.x] = part[SCAN_BLOCK - 1];
    }
}

template <typename T>
__global__ void k_scan_add(T * out, T const * offs, int64_t n)
{
    int64_t const base = static_cast<int64_t>(blockIdx.x) * SCAN_TILE;
    T const add = offs[blockIdx.x];
    for (unsigned k = threadIdx.x; k < SCAN_TILE; k += SCAN_BLOCK)
    {
        int64_t const i = base + k;
        if (i < n)
        {
            out[i] += add;
        }
    }
}

// out[i] = sum(in[0:i]); returns the total.  in and out may alias.
template <typename T>
T exclusive_scan(T const * in, T * out, int64_t n, hipStream_t s)
{
    if (n <= 0)
    {
        return T(0);
    }
    int64_t const nblk = (n + SCAN_TILE - 1) / SCAN_TILE;
    DevBuf<T> sums;
    sums.alloc(nblk);
    k_scan_tile<T><<<static_cast<unsigned>(nblk), SCAN_BLOCK, 0, s>>>(in, out, sums.ptr, n);
    SC_HIP(hipGetLastError());
    T total = 0;
    if (nblk > 1)
    {
        total = exclusive_scan<T>(sums.ptr, sums.ptr, nblk, s);
        k_scan_add<T><<<static_cast<unsigned>(nblk), SCAN_BLOCK, 0, s>>>(out, sums.ptr, n);
        SC_HIP(hipGetLastError());
    }
    else
    {
        SC_HIP(hipMemcpyAsync(&total, sums.ptr, sizeof(T), hipMemcpyDeviceToHost, s));
        SC_HIP(hipStreamSynchronize(s));
        return total;
    }
    SC_HIP(hipStreamSynchronize(s));
    return total;
}

} /* end namespace scamd */
