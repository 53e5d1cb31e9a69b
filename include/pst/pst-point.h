// PointStruct with the reference's layout and orders
// (helpers/point-struct.h:17-199): {T dim1_val; T dim2_val; [IDType id;]};
// compareDim1/compareDim2 order by the dimension value with IEEE `<`/`!=`,
// then by id, then by address. For code that does not already use the
// reference's header.
#ifndef PST_AMD_POINT_H
#define PST_AMD_POINT_H

#include <cstddef>
#include <ostream>
#include <type_traits>

template <typename T, typename IDType, size_t num_IDs>
struct PointStruct;

template <typename T>
struct PointStruct<T, void, 0> {
  static_assert(std::is_arithmetic<T>::value, "T must be arithmetic");
  T dim1_val{};
  T dim2_val{};
  PointStruct() = default;
  PointStruct(T d1, T d2) : dim1_val(d1), dim2_val(d2) {}
  int tie(const PointStruct& o) const { return this == &o ? 0 : (this < &o ? -1 : 1); }
  int compareDim1(const PointStruct& o) const {
    return dim1_val != o.dim1_val ? (dim1_val < o.dim1_val ? -1 : 1) : tie(o);
  }
  int compareDim2(const PointStruct& o) const {
    return dim2_val != o.dim2_val ? (dim2_val < o.dim2_val ? -1 : 1) : tie(o);
  }
  bool operator==(const PointStruct& o) const { return dim1_val == o.dim1_val && dim2_val == o.dim2_val; }
  void print(std::ostream& os) const { os << '(' << dim1_val << ", " << dim2_val << ')'; }
};

template <typename T, typename IDType>
struct PointStruct<T, IDType, 1> {
  static_assert(std::is_arithmetic<T>::value && std::is_arithmetic<IDType>::value, "arithmetic types only");
  T dim1_val{};
  T dim2_val{};
  IDType id{};
  PointStruct() = default;
  PointStruct(T d1, T d2, IDType i) : dim1_val(d1), dim2_val(d2), id(i) {}
  int tie(const PointStruct& o) const {
    if (id < o.id) return -1;
    if (id > o.id) return 1;
    return this == &o ? 0 : (this < &o ? -1 : 1);
  }
  int compareDim1(const PointStruct& o) const {
    return dim1_val != o.dim1_val ? (dim1_val < o.dim1_val ? -1 : 1) : tie(o);
  }
  int compareDim2(const PointStruct& o) const {
    return dim2_val != o.dim2_val ? (dim2_val < o.dim2_val ? -1 : 1) : tie(o);
  }
  bool operator==(const PointStruct& o) const {
    return dim1_val == o.dim1_val && dim2_val == o.dim2_val && id == o.id;
  }
  void print(std::ostream& os) const { os << '(' << dim1_val << ", " << dim2_val << "; " << id << ')'; }
};

template <typename T, typename IDType, size_t num_IDs>
std::ostream& operator<<(std::ostream& os, const PointStruct<T, IDType, num_IDs>& p) {
  p.print(os);
  return os;
}

#endif
