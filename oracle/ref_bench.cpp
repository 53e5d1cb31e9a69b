// CPU baseline timer over the reference's own StaticPSTCPUIter, compiled
// straight from its sources under the reference checkout (oracle/Makefile,
// target `ref`). Output goes to oracle/_ref/ only.
//
// TEST AND MEASUREMENT INFRASTRUCTURE ONLY: bench.py runs it as the timed CPU
// baseline; nothing in the product links or calls it.
//
// Usage: ref_bench input.bin T1 [T2 ...]
//   input.bin as for ref_driver (magic "PSTR", dtype, id_dtype, n, nq, the
//   points, then per query: int32 kind, int32 ret_point, T min_dim1,
//   max_dim1, min_dim2). Only f32 points with u32 IDs (the float workloads)
//   are timed; results are returned as IDs (RetType = IDType).
//   For each thread count Tk the query list is run once with Tk std::threads
//   sharing the tree read-only, query q on thread q mod Tk, each search
//   allocating and freeing its own result array as the reference does
//   (pst/static-pst-cpu-iter.tpp:197-486).
// Output: one JSON line {"n", "nq", "build_s", "runs": [{"threads", "secs",
// "ids"}]}. Timers are std::chrono::steady_clock, as the reference's testers
// use (testers/pst-tester-functions.tu:466-470, 620-621); construction covers
// the constructor call (index init, two sorts, populate).
#include <hip/hip_runtime.h>

#include <cassert>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <thread>
#include <typeinfo>
#include <vector>

#include "point-struct.h"
#include "static-pst-cpu-iter.h"

namespace {

struct Q {
  int32_t kind;
  float a, b, c;
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s input.bin threads...\n", argv[0]);
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
  if (!in || std::memcmp(magic, "PSTR", 4) != 0 || dt != 0 || idt != 4) {
    std::fprintf(stderr, "bad input (f32 points with u32 IDs expected)\n");
    return 1;
  }
  using P = PointStruct<float, unsigned, 1>;
  std::vector<P> pts(n);
  if (n) in.read(reinterpret_cast<char*>(pts.data()), n * sizeof(P));
  std::vector<Q> qs(nq);
  for (uint64_t q = 0; q < nq; q++) {
    int32_t kind, ret_point;
    float b[3];
    in.read(reinterpret_cast<char*>(&kind), 4);
    in.read(reinterpret_cast<char*>(&ret_point), 4);
    in.read(reinterpret_cast<char*>(b), sizeof(b));
    qs[q] = Q{kind, b[0], b[1], b[2]};
  }
  if (!in) {
    std::fprintf(stderr, "short input\n");
    return 1;
  }
  const auto t0 = std::chrono::steady_clock::now();
  StaticPSTCPUIter<float, PointStruct, unsigned, 1> tree(pts.data(), n);
  const auto t1 = std::chrono::steady_clock::now();
  std::printf("{\"n\": %llu, \"nq\": %llu, \"build_s\": %.6f, \"runs\": [", static_cast<unsigned long long>(n),
              static_cast<unsigned long long>(nq), std::chrono::duration<double>(t1 - t0).count());
  for (int k = 2; k < argc; k++) {
    const int T = std::max(1, std::atoi(argv[k]));
    std::vector<uint64_t> ids(T, 0);
    auto worker = [&](int w) {
      uint64_t cnt = 0;
      for (uint64_t q = w; q < nq; q += T) {
        size_t num = 0;
        unsigned* res = nullptr;
        const Q& x = qs[q];
        if (x.kind == 0) tree.template threeSidedSearch<unsigned>(num, res, x.a, x.b, x.c);
        else if (x.kind == 1) tree.template twoSidedLeftSearch<unsigned>(num, res, x.b, x.c);
        else tree.template twoSidedRightSearch<unsigned>(num, res, x.a, x.c);
        cnt += num;
        delete[] res;
      }
      ids[w] = cnt;
    };
    const auto s0 = std::chrono::steady_clock::now();
    if (T == 1) {
      worker(0);
    } else {
      std::vector<std::thread> th;
      for (int w = 0; w < T; w++) th.emplace_back(worker, w);
      for (auto& x : th) x.join();
    }
    const auto s1 = std::chrono::steady_clock::now();
    uint64_t tot = 0;
    for (auto v : ids) tot += v;
    std::printf("%s{\"threads\": %d, \"secs\": %.6f, \"ids\": %llu}", k > 2 ? ", " : "", T,
                std::chrono::duration<double>(s1 - s0).count(), static_cast<unsigned long long>(tot));
  }
  std::printf("]}\n");
  return 0;
}
