// GPU check of construction's subtree minima and ID arrays: a build with
// them filled during construction (the LDS subtree kernel does the bottom
// levels) must equal, byte for byte, the same arrays recomputed one level at
// a time from the built nodes by build_submin. Sizes around the 2048-point
// LDS subtree, ties in dim2, three point types. Prints one line per case and
// exits non-zero on the first difference. Built by __graft_entry__.build();
// run by tests/test_gpu_submin.py.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "pst_capi.h"
#include "pst_internal.hpp"

#define CHECK(x)                                                                          \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      std::printf("HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return 2;                                                                           \
    }                                                                                     \
  } while (0)

template <typename T, typename I>
struct Pt {
  T d1, d2;
  I id;
};
template <typename T>
struct PtNoId {
  T d1, d2;
};

template <typename P, bool HAS_ID, size_t ID_SIZE>
int run_case(int dt, int idt, size_t n, uint64_t seed) {
  std::mt19937_64 rng(seed);
  std::vector<P> pts(n);
  for (size_t i = 0; i < n; i++) {
    pts[i].d1 = static_cast<decltype(pts[i].d1)>(rng() % 100000);
    pts[i].d2 = static_cast<decltype(pts[i].d2)>(rng() % 5000);  // many ties
    if constexpr (HAS_ID) pts[i].id = static_cast<decltype(pts[i].id)>(i);
  }
  size_t S = 1;
  while (S < n + 1) S <<= 1;
  S -= 1;
  const size_t sn = pst::node_size(dt, idt), sd = pst::dtype_size(dt), si = ID_SIZE;
  void *dp, *nodes, *sub1, *sub2, *ids1 = nullptr, *ids2 = nullptr;
  CHECK(hipMalloc(&dp, n * sizeof(P)));
  CHECK(hipMalloc(&nodes, S * sn));
  CHECK(hipMalloc(&sub1, S * sd));
  CHECK(hipMalloc(&sub2, S * sd));
  CHECK(hipMemset(sub1, 0xAB, S * sd));
  CHECK(hipMemset(sub2, 0xAB, S * sd));
  if (HAS_ID) {
    CHECK(hipMalloc(&ids1, S * si));
    CHECK(hipMalloc(&ids2, S * si));
    CHECK(hipMemset(ids1, 0xAB, S * si));
    CHECK(hipMemset(ids2, 0xAB, S * si));
  }
  CHECK(hipMemcpy(dp, pts.data(), n * sizeof(P), hipMemcpyHostToDevice));
  pst::Status st = pst::build_tree(dt, idt, dp, n, nodes, nullptr, nullptr, sub1, ids1);
  if (st.ok()) st = pst::build_submin(dt, idt, n, 0, nodes, sub2, ids2, nullptr);
  CHECK(hipDeviceSynchronize());
  int bad = st.ok() ? 0 : 3;
  std::vector<unsigned char> a(S * sd), b(S * sd);
  CHECK(hipMemcpy(a.data(), sub1, S * sd, hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(b.data(), sub2, S * sd, hipMemcpyDeviceToHost));
  if (!bad && a != b) bad = 1;
  if (!bad && HAS_ID) {
    std::vector<unsigned char> c(S * si), d(S * si);
    CHECK(hipMemcpy(c.data(), ids1, S * si, hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(d.data(), ids2, S * si, hipMemcpyDeviceToHost));
    if (c != d) bad = 2;
  }
  std::printf("submin dt=%d idt=%d n=%zu: %s\n", dt, idt, n,
              bad == 0 ? "ok" : bad == 1 ? "SUBMIN DIFFERS" : bad == 2 ? "IDS DIFFER" : st.msg.c_str());
  (void)hipFree(dp), (void)hipFree(nodes), (void)hipFree(sub1), (void)hipFree(sub2);
  if (HAS_ID) (void)hipFree(ids1), (void)hipFree(ids2);
  return bad;
}

int main() {
  const size_t sizes[] = {1, 2, 3, 100, 2047, 2048, 2049, 4095, 4097, 10000, 100003, (1u << 20) + 7};
  int rc = 0;
  for (size_t n : sizes) {
    rc |= run_case<Pt<float, uint32_t>, true, 4>(PST_F32, PST_U32, n, n * 7 + 1);
    rc |= run_case<Pt<double, uint64_t>, true, 8>(PST_F64, PST_U64, n, n * 11 + 2);
    rc |= run_case<PtNoId<int32_t>, false, 0>(PST_I32, PST_NO_ID, n, n * 13 + 3);
    if (rc) return rc;
  }
  std::printf("all subtree minima ok\n");
  return 0;
}
