// Drop-in replacement for the reference's StaticPSTGPU class template
// (pst/static-pst-gpu.h:69-477): same template parameters, constructor,
// search methods, getters, calcGlobalMemNeeded and print, implemented over
// the C ABI of libpst_hip.so (pst_capi.h). Errors throw std::runtime_error
// with the library's message, as gpuErrorCheck does (helpers/gpu-err-chk.h:9-13).
#ifndef PST_AMD_STATIC_PST_GPU_H
#define PST_AMD_STATIC_PST_GPU_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "pst_capi.h"
#include "static-priority-search-tree.h"

namespace pst_amd {

template <typename V>
constexpr int dtype_code() {
  if constexpr (std::is_same<V, float>::value) return PST_F32;
  else if constexpr (std::is_same<V, double>::value) return PST_F64;
  else if constexpr (std::is_integral<V>::value && std::is_signed<V>::value && sizeof(V) == 4) return PST_I32;
  else if constexpr (std::is_integral<V>::value && std::is_signed<V>::value && sizeof(V) == 8) return PST_I64;
  else if constexpr (std::is_integral<V>::value && std::is_unsigned<V>::value && sizeof(V) == 4) return PST_U32;
  else if constexpr (std::is_integral<V>::value && std::is_unsigned<V>::value && sizeof(V) == 8) return PST_U64;
  else return -2;
}

template <typename IDType, size_t num_IDs>
constexpr int id_code() {
  if constexpr (num_IDs == 0 || std::is_void<IDType>::value) return PST_NO_ID;
  else return dtype_code<IDType>();
}

inline void check(int rc) {
  if (rc != PST_OK) throw std::runtime_error(pst_last_error());
}

// Reference-layout tree (dim1, dim2, med, ids, bitcodes) printed the way the
// reference's printRecur draws it (pst/static-pst-gpu.tu:441-483 and
// pst/static-pst-cpu-iter.tpp:821-853): root, then right subtree, then left.
template <typename T, typename IDType>
void print_tree(std::ostream& os, const std::vector<unsigned char>& blob, const size_t off[6], bool has_id,
                size_t s, const std::string& prefix, const std::string& child_prefix) {
  const T* d1 = reinterpret_cast<const T*>(blob.data() + off[0]);
  const T* d2 = reinterpret_cast<const T*>(blob.data() + off[1]);
  const T* md = reinterpret_cast<const T*>(blob.data() + off[2]);
  const unsigned char* bc = blob.data() + off[4];
  os << prefix << '(' << d1[s] << ", " << d2[s] << "; " << md[s];
  if constexpr (!std::is_void<IDType>::value) {
    if (has_id) os << "; " << reinterpret_cast<const IDType*>(blob.data() + off[3])[s];
  }
  os << ')';
  const bool l = bc[s] & 0x2, r = bc[s] & 0x1;
  if (l && r) {
    print_tree<T, IDType>(os, blob, off, has_id, 2 * s + 2, '\n' + child_prefix + "├─(R)─ ", child_prefix + "│      ");
    print_tree<T, IDType>(os, blob, off, has_id, 2 * s + 1, '\n' + child_prefix + "└─(L)─ ", child_prefix + "       ");
  } else if (r) {
    print_tree<T, IDType>(os, blob, off, has_id, 2 * s + 2, '\n' + child_prefix + "└─(R)─ ", child_prefix + "       ");
  } else if (l) {
    print_tree<T, IDType>(os, blob, off, has_id, 2 * s + 1, '\n' + child_prefix + "└─(L)─ ", child_prefix + "       ");
  }
}

}  // namespace pst_amd

template <typename T, template <typename, typename, size_t> class PointStructTemplate, typename IDType = void,
          size_t num_IDs = 0>
class StaticPSTGPU : public StaticPrioritySearchTree<T, PointStructTemplate, IDType, num_IDs> {
 public:
  using Point = PointStructTemplate<T, IDType, num_IDs>;
  static_assert(pst_amd::dtype_code<T>() >= PST_F32 && pst_amd::dtype_code<T>() <= PST_I64,
                "StaticPSTGPU: T must be float, double, int or long");

  StaticPSTGPU(Point* const pt_arr_d, size_t num_elems, const int warp_multiplier = 1, int dev_ind = 0,
               int num_devs = 1, hipDeviceProp_t dev_props = {})
      : num_elems(num_elems), dev_ind(dev_ind), num_devs(num_devs), dev_props(dev_props) {
    static_assert(sizeof(Point) > 0, "");
    if (pst_point_size(pst_amd::dtype_code<T>(), pst_amd::id_code<IDType, num_IDs>()) != sizeof(Point))
      throw std::runtime_error("StaticPSTGPU: PointStructTemplate layout differs from {T, T[, IDType]}");
    pst_amd::check(pst_gpu_create(pt_arr_d, num_elems, pst_amd::dtype_code<T>(), pst_amd::id_code<IDType, num_IDs>(),
                                  warp_multiplier, dev_ind, num_devs, nullptr, &tree));
  }
  virtual ~StaticPSTGPU() {
    if (tree) pst_gpu_destroy(tree);
  }
  StaticPSTGPU(const StaticPSTGPU&) = delete;
  StaticPSTGPU& operator=(const StaticPSTGPU&) = delete;

  virtual void print(std::ostream& os) const {
    if (num_elems == 0) {
      os << "Tree is empty\n";
      return;
    }
    std::vector<unsigned char> blob(pst_gpu_export_bytes(tree));
    size_t off[6];
    pst_amd::check(pst_gpu_export_tree(tree, blob.data(), blob.size(), off));
    pst_amd::print_tree<T, typename std::conditional<num_IDs == 0, void, IDType>::type>(
        os, blob, off, num_IDs != 0, 0, "", "");
  }

  int getDevInd() const { return dev_ind; }
  hipDeviceProp_t getDevProps() const { return dev_props; }
  int getNumDevs() const { return num_devs; }

  template <typename RetType = Point>
    requires std::disjunction<std::is_same<RetType, IDType>, std::is_same<RetType, Point>>::value
  void threeSidedSearch(size_t& num_res_elems, RetType*& res_arr_d, T min_dim1_val, T max_dim1_val,
                        T min_dim2_val, const int warp_multiplier = 1) {
    (void)warp_multiplier;
    void* p = nullptr;
    pst_amd::check(pst_gpu_three_sided_search(tree, ret_code<RetType>(), &num_res_elems, &p, &min_dim1_val,
                                              &max_dim1_val, &min_dim2_val));
    res_arr_d = static_cast<RetType*>(p);
  }

  template <typename RetType = Point>
    requires std::disjunction<std::is_same<RetType, IDType>, std::is_same<RetType, Point>>::value
  void twoSidedLeftSearch(size_t& num_res_elems, RetType*& res_arr_d, T max_dim1_val, T min_dim2_val,
                          const int warp_multiplier = 1) {
    (void)warp_multiplier;
    void* p = nullptr;
    pst_amd::check(pst_gpu_two_sided_left_search(tree, ret_code<RetType>(), &num_res_elems, &p, &max_dim1_val,
                                                 &min_dim2_val));
    res_arr_d = static_cast<RetType*>(p);
  }

  template <typename RetType = Point>
    requires std::disjunction<std::is_same<RetType, IDType>, std::is_same<RetType, Point>>::value
  void twoSidedRightSearch(size_t& num_res_elems, RetType*& res_arr_d, T min_dim1_val, T min_dim2_val,
                           const int warp_multiplier = 1) {
    (void)warp_multiplier;
    void* p = nullptr;
    pst_amd::check(pst_gpu_two_sided_right_search(tree, ret_code<RetType>(), &num_res_elems, &p, &min_dim1_val,
                                                  &min_dim2_val));
    res_arr_d = static_cast<RetType*>(p);
  }

  static size_t calcGlobalMemNeeded(const size_t num_elems) {
    return pst_gpu_calc_global_mem_needed(num_elems, pst_amd::dtype_code<T>(), pst_amd::id_code<IDType, num_IDs>());
  }

  // Beyond the reference: the handle for the batched entry points.
  pst_tree* handle() const { return tree; }

  // Beyond the reference: persistence of the built tree (pst_gpu_save,
  // pst_gpu_load). load() needs neither the input points nor a rebuild.
  void save(const std::string& path) const { pst_amd::check(pst_gpu_save(tree, path.c_str())); }
  // Level index for ID searches (pst_gpu_set_level_index; no counterpart in
  // the reference).
  void setLevelIndex(bool on) { pst_amd::check(pst_gpu_set_level_index(tree, on ? 1 : 0)); }
  static std::unique_ptr<StaticPSTGPU> load(const std::string& path, int dev_ind = 0, int num_devs = 1) {
    pst_tree* h = nullptr;
    pst_amd::check(pst_gpu_load(path.c_str(), dev_ind, num_devs, nullptr, &h));
    if (pst_gpu_slots_per_tree(h) != 0 || pst_gpu_dtype(h) != pst_amd::dtype_code<T>() ||
        pst_gpu_id_dtype(h) != pst_amd::id_code<IDType, num_IDs>()) {
      pst_gpu_destroy(h);
      throw std::runtime_error("StaticPSTGPU::load: " + path + " holds a different tree type");
    }
    return std::unique_ptr<StaticPSTGPU>(new StaticPSTGPU(h, pst_gpu_num_elems(h), dev_ind, num_devs));
  }

 private:
  StaticPSTGPU(pst_tree* h, size_t num_elems, int dev_ind, int num_devs)
      : tree(h), num_elems(num_elems), dev_ind(dev_ind), num_devs(num_devs), dev_props{} {}

  template <typename RetType>
  static constexpr int ret_code() {
    return std::is_same<RetType, Point>::value ? PST_RET_POINT : PST_RET_ID;
  }

  pst_tree* tree = nullptr;
  size_t num_elems;
  int dev_ind;
  int num_devs;
  hipDeviceProp_t dev_props;
};

#endif
