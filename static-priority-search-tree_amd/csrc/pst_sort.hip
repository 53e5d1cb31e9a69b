// Key/value radix sorts used by tree construction.
//
// The reference sorts size_t index arrays with thrust::sort and an
// indirect comparator (pst/static-pst-gpu.tu:266-267, 317-318). Here the
// comparison keys are first turned into order-preserving unsigned integers
// (pst_common.hpp: ord_key) and sorted with rocPRIM's device radix sort,
// AMD's own library routine for this plain operation; values are 32-bit.
// The sort is stable, which construction relies on for the index
// tie-break.
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include "pst_internal.hpp"

namespace pst {

hipError_t radix_sort_u32(void* tmp, size_t& tmp_bytes, const uint32_t* kin, uint32_t* kout,
                          const uint32_t* vin, uint32_t* vout, size_t n, hipStream_t s) {
  return rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, vin, vout, n, 0, 32, s);
}

hipError_t radix_sort_u64(void* tmp, size_t& tmp_bytes, const uint64_t* kin, uint64_t* kout,
                          const uint32_t* vin, uint32_t* vout, size_t n, hipStream_t s) {
  return rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, vin, vout, n, 0, 64, s);
}

}  // namespace pst
