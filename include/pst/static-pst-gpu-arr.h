// Drop-in replacement for the reference's StaticPSTGPUArr class template
// (pst/static-pst-gpu-arr.h:45-366): a forest of shallow trees, one per
// chunk of (smallest power of two above threads_per_block) - 1 consecutive
// input points. Same template parameters, constructor, search methods,
// getters, calcGlobalMemNeeded and print, over the C ABI of libpst_hip.so.
// threeSidedSearch and twoSidedRightSearch, empty in the reference
// (pst/static-pst-gpu-arr.tu:564-578, 645-658), report like the other
// searches.
#ifndef PST_AMD_STATIC_PST_GPU_ARR_H
#define PST_AMD_STATIC_PST_GPU_ARR_H

#include "static-pst-gpu.h"

template <typename T, template <typename, typename, size_t> class PointStructTemplate, typename IDType = void,
          size_t num_IDs = 0>
class StaticPSTGPUArr : public StaticPrioritySearchTree<T, PointStructTemplate, IDType, num_IDs> {
 public:
  using Point = PointStructTemplate<T, IDType, num_IDs>;
  static_assert(pst_amd::dtype_code<T>() >= PST_F32 && pst_amd::dtype_code<T>() <= PST_I64,
                "StaticPSTGPUArr: T must be float, double, int or long");

  StaticPSTGPUArr(Point* const pt_arr_d, const size_t num_elems, const unsigned threads_per_block,
                  const int dev_ind = 0, const int num_devs = 1, const hipDeviceProp_t& dev_props = {})
      : num_elems(num_elems), dev_ind(dev_ind), num_devs(num_devs), dev_props(dev_props) {
    if (pst_point_size(pst_amd::dtype_code<T>(), pst_amd::id_code<IDType, num_IDs>()) != sizeof(Point))
      throw std::runtime_error("StaticPSTGPUArr: PointStructTemplate layout differs from {T, T[, IDType]}");
    pst_amd::check(pst_gpu_arr_create(pt_arr_d, num_elems, pst_amd::dtype_code<T>(),
                                      pst_amd::id_code<IDType, num_IDs>(), threads_per_block, dev_ind, num_devs,
                                      nullptr, &tree));
  }
  virtual ~StaticPSTGPUArr() {
    if (tree) pst_gpu_destroy(tree);
  }
  StaticPSTGPUArr(const StaticPSTGPUArr&) = delete;
  StaticPSTGPUArr& operator=(const StaticPSTGPUArr&) = delete;

  // Each tree printed as StaticPSTGPU::print draws one tree, trees in order.
  virtual void print(std::ostream& os) const {
    if (num_elems == 0) {
      os << "Tree is empty\n";
      return;
    }
    std::vector<unsigned char> blob(pst_gpu_export_bytes(tree));
    size_t off[6];
    pst_amd::check(pst_gpu_export_tree(tree, blob.data(), blob.size(), off));
    const size_t St = pst_gpu_slots_per_tree(tree);
    const size_t sT = sizeof(T), sI = num_IDs ? sizeof(IDType) : 0;
    for (size_t c0 = 0; c0 < off[5]; c0 += St) {
      // re-base the arrays at this tree so that slot 0 is its root
      size_t o[6] = {off[0] + c0 * sT, off[1] + c0 * sT, off[2] + c0 * sT, off[3] + c0 * sI, off[4] + c0, 0};
      pst_amd::print_tree<T, typename std::conditional<num_IDs == 0, void, IDType>::type>(
          os, blob, o, num_IDs != 0, 0, "", "");
      os << '\n';
    }
  }

  int getDevInd() const { return dev_ind; }
  hipDeviceProp_t getDevProps() const { return dev_props; }
  int getNumDevs() const { return num_devs; }

  template <typename RetType = Point>
    requires std::disjunction<std::is_same<RetType, IDType>, std::is_same<RetType, Point>>::value
  void threeSidedSearch(size_t& num_res_elems, RetType*& res_arr_d, const T min_dim1_val, const T max_dim1_val,
                        const T min_dim2_val) {
    void* p = nullptr;
    pst_amd::check(pst_gpu_three_sided_search(tree, ret_code<RetType>(), &num_res_elems, &p, &min_dim1_val,
                                              &max_dim1_val, &min_dim2_val));
    res_arr_d = static_cast<RetType*>(p);
  }

  template <typename RetType = Point>
    requires std::disjunction<std::is_same<RetType, IDType>, std::is_same<RetType, Point>>::value
  void twoSidedLeftSearch(size_t& num_res_elems, RetType*& res_arr_d, const T max_dim1_val, const T min_dim2_val) {
    void* p = nullptr;
    pst_amd::check(pst_gpu_two_sided_left_search(tree, ret_code<RetType>(), &num_res_elems, &p, &max_dim1_val,
                                                 &min_dim2_val));
    res_arr_d = static_cast<RetType*>(p);
  }

  template <typename RetType = Point>
    requires std::disjunction<std::is_same<RetType, IDType>, std::is_same<RetType, Point>>::value
  void twoSidedRightSearch(size_t& num_res_elems, RetType*& res_arr_d, const T min_dim1_val, const T min_dim2_val) {
    void* p = nullptr;
    pst_amd::check(pst_gpu_two_sided_right_search(tree, ret_code<RetType>(), &num_res_elems, &p, &min_dim1_val,
                                                  &min_dim2_val));
    res_arr_d = static_cast<RetType*>(p);
  }

  static size_t calcGlobalMemNeeded(const size_t num_elems, const unsigned threads_per_block) {
    return pst_gpu_arr_calc_global_mem_needed(num_elems, threads_per_block, pst_amd::dtype_code<T>(),
                                              pst_amd::id_code<IDType, num_IDs>());
  }

  // Beyond the reference: the handle for the batched entry points.
  pst_tree* handle() const { return tree; }

  // Beyond the reference: persistence of the built tree (pst_gpu_save,
  // pst_gpu_load). load() needs neither the input points nor a rebuild.
  void save(const std::string& path) const { pst_amd::check(pst_gpu_save(tree, path.c_str())); }
  static std::unique_ptr<StaticPSTGPUArr> load(const std::string& path, int dev_ind = 0, int num_devs = 1) {
    pst_tree* h = nullptr;
    pst_amd::check(pst_gpu_load(path.c_str(), dev_ind, num_devs, nullptr, &h));
    if (pst_gpu_slots_per_tree(h) == 0 || pst_gpu_dtype(h) != pst_amd::dtype_code<T>() ||
        pst_gpu_id_dtype(h) != pst_amd::id_code<IDType, num_IDs>()) {
      pst_gpu_destroy(h);
      throw std::runtime_error("StaticPSTGPUArr::load: " + path + " holds a different tree type");
    }
    return std::unique_ptr<StaticPSTGPUArr>(new StaticPSTGPUArr(h, pst_gpu_num_elems(h), dev_ind, num_devs));
  }

 private:
  StaticPSTGPUArr(pst_tree* h, size_t num_elems, int dev_ind, int num_devs)
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
