// GPU check of the library's radix sorts (pst::radix_sort_u32 / _u64)
// against std::stable_sort on the host: keys with many duplicates, all-equal
// keys, sizes around the 8192-key tile, values carried along (iota), inputs
// left unchanged. Prints one line per case and exits non-zero on the first
// mismatch. Built by __graft_entry__.build(); run by tests/test_gpu_sort.py.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>

namespace pst {
hipError_t radix_sort_u32(void* tmp, size_t& tmp_bytes, const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
                          uint32_t* vout, size_t n, hipStream_t s, int* fail);
hipError_t radix_sort_u64(void* tmp, size_t& tmp_bytes, const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                          uint32_t* vout, size_t n, hipStream_t s, int* fail);
}  // namespace pst

#define CHECK(x)                                                                      \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      std::printf("HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return 2;                                                                       \
    }                                                                                 \
  } while (0)

template <typename K>
int run_case(size_t n, int dist, uint64_t seed) {
  std::mt19937_64 rng(seed);
  std::vector<K> keys(n);
  for (auto& k : keys) {
    uint64_t r = rng();
    if (dist == 0) k = static_cast<K>(r);                  // spread
    else if (dist == 1) k = static_cast<K>(r % 17) << (sizeof(K) * 8 - 8);  // few values, top digit only
    else if (dist == 2) k = static_cast<K>(r % 1000);      // duplicates, low digits
    else k = static_cast<K>(7);                            // all equal
  }
  std::vector<uint32_t> vals(n);
  std::iota(vals.begin(), vals.end(), 0u);
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
  K *kin, *kout;
  uint32_t *vin, *vout;
  CHECK(hipMalloc(&kin, std::max<size_t>(n, 1) * sizeof(K)));
  CHECK(hipMalloc(&kout, std::max<size_t>(n, 1) * sizeof(K)));
  CHECK(hipMalloc(&vin, std::max<size_t>(n, 1) * 4));
  CHECK(hipMalloc(&vout, std::max<size_t>(n, 1) * 4));
  CHECK(hipMemcpy(kin, keys.data(), n * sizeof(K), hipMemcpyHostToDevice));
  CHECK(hipMemcpy(vin, vals.data(), n * 4, hipMemcpyHostToDevice));
  size_t tb = 0;
  void* tmp = nullptr;
  int* fail;
  CHECK(hipMalloc(&fail, sizeof(int)));
  CHECK(hipMemset(fail, 0, sizeof(int)));
  if constexpr (sizeof(K) == 4) CHECK(pst::radix_sort_u32(nullptr, tb, nullptr, nullptr, nullptr, nullptr, n, 0, nullptr));
  else CHECK(pst::radix_sort_u64(nullptr, tb, nullptr, nullptr, nullptr, nullptr, n, 0, nullptr));
  CHECK(hipMalloc(&tmp, tb));
  // twice with the same scratch: the second sort must not see the first one's state
  for (int rep = 0; rep < 2; rep++) {
    if constexpr (sizeof(K) == 4) CHECK(pst::radix_sort_u32(tmp, tb, kin, kout, vin, vout, n, 0, fail));
    else CHECK(pst::radix_sort_u64(tmp, tb, kin, kout, vin, vout, n, 0, fail));
  }
  CHECK(hipDeviceSynchronize());
  int hfail = 0;
  CHECK(hipMemcpy(&hfail, fail, sizeof(int), hipMemcpyDeviceToHost));
  std::vector<K> gk(n), gin(n);
  std::vector<uint32_t> gv(n);
  CHECK(hipMemcpy(gk.data(), kout, n * sizeof(K), hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(gv.data(), vout, n * 4, hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(gin.data(), kin, n * sizeof(K), hipMemcpyDeviceToHost));
  int bad = 0;
  for (size_t i = 0; i < n && !bad; i++)
    if (gv[i] != order[i] || gk[i] != keys[order[i]]) bad = 1;
  if (gin != keys) bad = 2;
  if (hfail) bad = 3;
  std::printf("sort u%zu n=%zu dist=%d: %s\n", sizeof(K) * 8, n, dist,
              bad == 0 ? "ok" : (bad == 1 ? "MISMATCH" : (bad == 2 ? "INPUT CHANGED" : "WAIT RAN OUT")));
  (void)hipFree(kin), (void)hipFree(kout), (void)hipFree(vin), (void)hipFree(vout), (void)hipFree(tmp), (void)hipFree(fail);
  return bad;
}

int main() {
  const size_t sizes[] = {1, 2, 100, 4095, 4096, 4097, 8191, 8192, 8193, 16385, 100003, (1u << 20) + 3, 1u << 24};
  int rc = 0;
  for (size_t n : sizes)
    for (int dist = 0; dist < 4; dist++) {
      rc |= run_case<uint32_t>(n, dist, n * 31 + dist);
      if (n <= (1u << 20) + 3) rc |= run_case<uint64_t>(n, dist, n * 37 + dist);
      if (rc) return rc;
    }
  std::printf("all sorts ok\n");
  return 0;
}
