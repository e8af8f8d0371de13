// Per-row sort of 64-bit keys (the large-k and large-m selection paths):
// rocPRIM's segmented radix sort, ascending over all 64 bits.
#include <cstring>

#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "common.h"

namespace ovx {

size_t segsort_temp_bytes(uint64_t total, uint32_t nseg, const uint64_t* offsets, hipStream_t s) {
    if (total > 0x7fffffffull) throw Error(kInternal, "segmented sort: more than 2^31 keys in one call");
    size_t bytes = 0;
    hipError_t e = rocprim::segmented_radix_sort_keys(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                                      (unsigned int)total, nseg, offsets, offsets + 1, 0, 64, s);
    if (e != hipSuccess) throw Error(kInternal, std::string("segmented sort size query: ") + hipGetErrorString(e));
    return (bytes + 255) & ~(size_t)255;
}

void segsort_keys(void* temp, size_t temp_bytes, const uint64_t* in, uint64_t* out, uint64_t total, uint32_t nseg,
                  const uint64_t* offsets, hipStream_t s) {
    if (total == 0 || nseg == 0) return;
    hipError_t e = rocprim::segmented_radix_sort_keys(temp, temp_bytes, in, out, (unsigned int)total, nseg, offsets,
                                                      offsets + 1, 0, 64, s);
    if (e != hipSuccess) throw Error(kInternal, std::string("segmented sort: ") + hipGetErrorString(e));
}

}  // namespace ovx
