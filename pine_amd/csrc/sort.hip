// Sorting of ray keys for the reordered traversal (PG_LAUNCH_REORDER): the
// (key, ray index) pairs are sorted by rocPRIM's device radix sort (onesweep,
// 8-bit digits), with the ray indices generated by a counting iterator, so the
// only input the caller writes is the key array.  Kept in its own translation
// unit: the sort's templates are large and build in parallel with the rest.
#include <cstring>  // rocPRIM's headers use memset on the host

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <cstdint>

namespace pg {

hipError_t reorder_sort_temp_bytes(int64_t n, int end_bit, size_t& bytes) {
    bytes = 0;
    rocprim::counting_iterator<uint32_t> idx(0u);
    return rocprim::radix_sort_pairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, idx,
                                     (uint32_t*)nullptr, (size_t)n, 0u, (unsigned)end_bit, nullptr);
}

// Sorts keys_in[0, n) (bits [0, end_bit)) and writes, in key order, the
// index of each key's ray to perm; keys_out receives the sorted keys.
hipError_t reorder_sort(void* temp, size_t bytes, const uint32_t* keys_in, uint32_t* keys_out,
                        uint32_t* perm, int64_t n, int end_bit, hipStream_t s) {
    rocprim::counting_iterator<uint32_t> idx(0u);
    return rocprim::radix_sort_pairs(temp, bytes, keys_in, keys_out, idx, perm, (size_t)n, 0u,
                                     (unsigned)end_bit, s);
}

}  // namespace pg
