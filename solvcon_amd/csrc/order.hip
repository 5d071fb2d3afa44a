/*
 * Key-value radix sort used once per EulerCore to order the body cells
 * (cell type in the high bits, a Hilbert index of the cell centroid below).
 */

#include "common.hpp"

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

namespace scamd
{

void sort_pairs_u64(uint64_t * kin, uint64_t * kout, int32_t * vin, int32_t * vout, int64_t n, hipStream_t s)
{
    if (n <= 0)
    {
        return;
    }
    size_t tmp = 0;
    SC_HIP(rocprim::radix_sort_pairs(nullptr, tmp, kin, kout, vin, vout, static_cast<size_t>(n), 0, 64, s));
    DevBuf<unsigned char> work;
    work.alloc(static_cast<int64_t>(tmp) + 1);
    SC_HIP(rocprim::radix_sort_pairs(work.ptr, tmp, kin, kout, vin, vout, static_cast<size_t>(n), 0, 64, s));
    SC_HIP(hipStreamSynchronize(s));
}

} /* end namespace scamd */
