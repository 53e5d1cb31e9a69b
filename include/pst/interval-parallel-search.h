// Drop-in replacement for the reference's intervalParallelSearch
// (interval-parallel-search/interval-parallel-search.h:23-66): reports the
// points with dim1 <= search_val <= dim2, as IDs or points, into a device
// array of num_elems * sizeof(RetType) bytes allocated by the call.
#ifndef PST_AMD_INTERVAL_PARALLEL_SEARCH_H
#define PST_AMD_INTERVAL_PARALLEL_SEARCH_H

#include <type_traits>

#include "pst_capi.h"
#include "static-pst-gpu.h"

template <typename T, template <typename, typename, size_t> class PointStructTemplate, typename IDType,
          size_t num_IDs, typename RetType>
  requires std::disjunction<std::is_same<RetType, IDType>,
                            std::is_same<RetType, PointStructTemplate<T, IDType, num_IDs>>>::value
void intervalParallelSearch(PointStructTemplate<T, IDType, num_IDs>* pt_arr_d, const size_t num_elems,
                            RetType*& res_arr_d, size_t& num_res_elems, T search_val, const int dev_ind,
                            const int num_devs, const int warp_size, const unsigned num_thread_blocks,
                            const unsigned threads_per_block) {
  constexpr int ret =
      std::is_same<RetType, PointStructTemplate<T, IDType, num_IDs>>::value ? PST_RET_POINT : PST_RET_ID;
  void* p = nullptr;
  pst_amd::check(pst_interval_parallel_search(pt_arr_d, num_elems, pst_amd::dtype_code<T>(),
                                              pst_amd::id_code<IDType, num_IDs>(), ret, &p, &num_res_elems,
                                              &search_val, dev_ind, num_devs, warp_size, num_thread_blocks,
                                              threads_per_block));
  res_arr_d = static_cast<RetType*>(p);
}

#endif
