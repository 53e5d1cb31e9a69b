// Driver that runs the reference's own StaticPSTCPUIter, compiled straight
// from its sources under the reference checkout (see oracle/Makefile,
// target `ref`). Output goes to oracle/_ref/ only.
//
// TEST INFRASTRUCTURE ONLY: used to write golden fixtures and to cross-check
// the oracle restatement; never linked into the product.
//
// The reference headers use the CUDA-style qualifiers __host__, __device__,
// __forceinline__ and the built-in blockIdx/threadIdx in helpers/linearise-id.h;
// building this file as HIP (hipcc -x hip) supplies all of them natively, so
// no reference file is edited and nothing is stubbed. <cassert>, <iostream>
// and <typeinfo> are included up front because the reference uses them
// without including them (static-pst-cpu-iter-populate-tree.tpp:58,
// static-pst-cpu-iter.tpp:48).
//
// Input (binary, little endian), file given as argv[1]:
//   char magic[4] = "PSTR"; int32 dtype; int32 id_dtype; uint64 n; uint64 nq;
//   n AoS points as the reference's PointStruct<T, IDType, num_IDs>;
//   nq queries: int32 kind (0 three, 1 left, 2 right); int32 ret_point;
//               T min_dim1, max_dim1, min_dim2.
// Output (text, stdout):
//   "TREE\n" + StaticPSTCPUIter::print() at round-trip precision + "\nENDTREE\n"
//   per query: "Q <index> <count>\n" then one line per result, in the
//   reference's own (LIFO) order: "<id>" or "<dim1> <dim2>[ <id>]".
#include <hip/hip_runtime.h>

#include <cassert>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <typeinfo>
#include <vector>

#include "point-struct.h"
#include "static-pst-cpu-iter.h"

namespace {

template <typename V>
void put(std::ostream& os, V v) {
  if constexpr (std::is_floating_point<V>::value)
    os << std::setprecision(std::numeric_limits<V>::max_digits10) << v;
  else
    os << +v;
}

template <typename T, typename I, size_t NI>
int run(std::ifstream& in, uint64_t n, uint64_t nq) {
  using P = PointStruct<T, I, NI>;
  std::vector<P> pts(n);
  if (n) in.read(reinterpret_cast<char*>(pts.data()), n * sizeof(P));
  StaticPSTCPUIter<T, PointStruct, I, NI> tree(pts.data(), n);
  std::cout << std::setprecision(std::numeric_limits<T>::max_digits10);
  std::cout << "TREE\n";
  tree.print(std::cout);
  std::cout << "\nENDTREE\n";
  for (uint64_t q = 0; q < nq; q++) {
    int32_t kind, ret_point;
    T b[3];
    in.read(reinterpret_cast<char*>(&kind), 4);
    in.read(reinterpret_cast<char*>(&ret_point), 4);
    in.read(reinterpret_cast<char*>(b), sizeof(b));
    size_t k = 0;
    if (ret_point || NI == 0) {
      P* res = nullptr;
      if (kind == 0) tree.template threeSidedSearch<P>(k, res, b[0], b[1], b[2]);
      else if (kind == 1) tree.template twoSidedLeftSearch<P>(k, res, b[1], b[2]);
      else tree.template twoSidedRightSearch<P>(k, res, b[0], b[2]);
      std::cout << "Q " << q << ' ' << k << '\n';
      for (size_t i = 0; i < k; i++) {
        put(std::cout, res[i].dim1_val);
        std::cout << ' ';
        put(std::cout, res[i].dim2_val);
        if constexpr (NI == 1) {
          std::cout << ' ';
          put(std::cout, res[i].id);
        }
        std::cout << '\n';
      }
      delete[] res;
    } else {
      if constexpr (NI == 1) {
        I* res = nullptr;
        if (kind == 0) tree.template threeSidedSearch<I>(k, res, b[0], b[1], b[2]);
        else if (kind == 1) tree.template twoSidedLeftSearch<I>(k, res, b[1], b[2]);
        else tree.template twoSidedRightSearch<I>(k, res, b[0], b[2]);
        std::cout << "Q " << q << ' ' << k << '\n';
        for (size_t i = 0; i < k; i++) {
          put(std::cout, res[i]);
          std::cout << '\n';
        }
        delete[] res;
      }
    }
  }
  return 0;
}

template <typename T>
int by_id(std::ifstream& in, int idt, uint64_t n, uint64_t nq) {
  switch (idt) {
    case -1: return run<T, void, 0>(in, n, nq);
    case 0: return run<T, float, 1>(in, n, nq);
    case 1: return run<T, double, 1>(in, n, nq);
    case 2: return run<T, int, 1>(in, n, nq);
    case 3: return run<T, long, 1>(in, n, nq);
    case 4: return run<T, unsigned, 1>(in, n, nq);
    case 5: return run<T, unsigned long, 1>(in, n, nq);
  }
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s input.bin\n", argv[0]);
    return 1;
  }
  std::ifstream in(argv[1], std::ios::binary);
  char magic[4];
  int32_t dt, idt;
  uint64_t n, nq;
  in.read(magic, 4);
  in.read(reinterpret_cast<char*>(&dt), 4);
  in.read(reinterpret_cast<char*>(&idt), 4);
  in.read(reinterpret_cast<char*>(&n), 8);
  in.read(reinterpret_cast<char*>(&nq), 8);
  if (!in || std::memcmp(magic, "PSTR", 4) != 0) {
    std::fprintf(stderr, "bad input\n");
    return 1;
  }
  switch (dt) {
    case 0: return by_id<float>(in, idt, n, nq);
    case 1: return by_id<double>(in, idt, n, nq);
    case 2: return by_id<int>(in, idt, n, nq);
    case 3: return by_id<long>(in, idt, n, nq);
  }
  return 2;
}
