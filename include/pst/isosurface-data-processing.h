// Drop-in replacement for the reference's isosurface helpers
// (helpers/isosurface-data-processing.h:21-229, with the Dims indices of
// helpers/linearise-id.h:8 and the exit codes of helpers/exit-status-codes.h):
// calcNumMetacells, formMetacellTags and readInVertices, implemented over
// the C ABI of libpst_hip.so (pst_capi.h).
#ifndef PST_AMD_ISOSURFACE_DATA_PROCESSING_H
#define PST_AMD_ISOSURFACE_DATA_PROCESSING_H

#include <cstddef>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "pst_capi.h"
#include "static-pst-gpu.h"

#ifndef PST_AMD_DIMS_DEFINED
#define PST_AMD_DIMS_DEFINED
// Axis indices of the grid-dimension arrays (x fastest).
enum Dims { X_DIM_IND, Y_DIM_IND, Z_DIM_IND, NUM_DIMS };
// Process exit codes of the reference's tools.
enum ExitStatusCodes { SUCCESS = 0, FILE_NOT_FOUND_ERR, FILE_OPEN_ERR, INPUT_READ_ERR, INSUFFICIENT_NUM_ARGS_ERR,
                       INVALID_ARG_ERR };
#endif

namespace pst_amd {

// The (T, IDType, num_IDs) parameters of a PointStruct-like type.
template <typename P>
struct point_traits;
template <typename T, typename I, size_t k, template <typename, typename, size_t> class PS>
struct point_traits<PS<T, I, k>> {
  using dim_type = T;
  using id_type = I;
  static constexpr size_t num_ids = k;
};

}  // namespace pst_amd

// Metacell grid dimensions ceil((pt_grid_dims[i] - 1) / metacell_dims[i])
// and their product.
template <typename GridDimType>
  requires std::is_integral<GridDimType>::value
void calcNumMetacells(GridDimType pt_grid_dims[Dims::NUM_DIMS], GridDimType metacell_dims[Dims::NUM_DIMS],
                      GridDimType metacell_grid_dims[Dims::NUM_DIMS], size_t& num_metacells) {
  size_t p[3], m[3], g[3] = {0, 0, 0};
  for (int i = 0; i < Dims::NUM_DIMS; i++) {
    p[i] = static_cast<size_t>(pt_grid_dims[i]);
    m[i] = static_cast<size_t>(metacell_dims[i]);
  }
  num_metacells = pst_calc_num_metacells(p, m, g);
  for (int i = 0; i < Dims::NUM_DIMS; i++) metacell_grid_dims[i] = static_cast<GridDimType>(g[i]);
}

// One (min, max, ID) interval per metacell, in a device array allocated by
// the call (free with hipFree or pst_free). dev_ind, num_devs and warp_size
// are accepted for signature parity.
template <class PointStruct, typename T, typename GridDimType>
  requires std::is_integral<GridDimType>::value
PointStruct* formMetacellTags(T* const vertex_arr_d, GridDimType pt_grid_dims[Dims::NUM_DIMS],
                              GridDimType metacell_dims[Dims::NUM_DIMS], GridDimType metacell_grid_dims[NUM_DIMS],
                              const size_t num_metacells, const int dev_ind, const int num_devs,
                              const int warp_size) {
  (void)metacell_grid_dims;
  (void)warp_size;
  using traits = pst_amd::point_traits<PointStruct>;
  static_assert(std::is_same<typename traits::dim_type, T>::value, "formMetacellTags: PointStruct holds T");
  constexpr int dt = pst_amd::dtype_code<T>();
  constexpr int idt = pst_amd::id_code<typename traits::id_type, traits::num_ids>();
  if (pst_point_size(dt, idt) != sizeof(PointStruct))
    throw std::runtime_error("formMetacellTags: PointStruct layout differs from {T, T[, IDType]}");
  size_t p[3], m[3];
  for (int i = 0; i < Dims::NUM_DIMS; i++) {
    p[i] = static_cast<size_t>(pt_grid_dims[i]);
    m[i] = static_cast<size_t>(metacell_dims[i]);
  }
  void* out = nullptr;
  size_t k = 0;
  pst_amd::check(pst_form_metacell_tags(vertex_arr_d, dt, idt, p, m, &out, &k, dev_ind, num_devs));
  if (k != num_metacells) {
    pst_free(out);
    throw std::runtime_error("formMetacellTags: num_metacells differs from calcNumMetacells");
  }
  return static_cast<PointStruct*>(out);
}

// num_vertices values of T from a raw binary file, in a new host array
// (delete[]). On a file that cannot be opened or is too short, prints the
// reference's message and exits with FILE_OPEN_ERR or INPUT_READ_ERR.
template <typename T, typename GridDimType>
  requires std::is_integral<GridDimType>::value
T* readInVertices(std::string input_filename, GridDimType num_vertices) {
  T* vertex_arr = new T[static_cast<size_t>(num_vertices)];
  if (pst_read_vertices(input_filename.c_str(), pst_amd::dtype_code<T>(), static_cast<size_t>(num_vertices),
                        vertex_arr) != PST_OK) {
    const std::string m = pst_last_error();
    std::cerr << m << '\n';
    delete[] vertex_arr;
    std::exit(m.rfind("Failed to open", 0) == 0 ? ExitStatusCodes::FILE_OPEN_ERR : ExitStatusCodes::INPUT_READ_ERR);
  }
  return vertex_arr;
}

#endif
