// Drop-in check: the reference's class/function names and signatures
// (StaticPSTGPU, intervalParallelSearch) used exactly as its testers use them
// (testers/pst-tester-functions.tu:513-520, 560-575, 685-736), against a
// brute-force scan. Prints "DROPIN OK" on success, exits non-zero otherwise.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <random>
#include <sstream>
#include <vector>

#include "interval-parallel-search.h"
#include "isosurface-data-processing.h"
#include "pst-point.h"
#include "static-pst-gpu-arr.h"
#include "static-pst-gpu.h"

static int failures = 0;
#define EXPECT(c, ...)                \
  do {                                \
    if (!(c)) {                       \
      std::printf("FAIL: " __VA_ARGS__); \
      std::printf("\n");              \
      failures++;                     \
    }                                 \
  } while (0)

template <typename T, typename P>
std::vector<P> to_host(P* d, size_t k) {
  std::vector<P> h(k);
  if (k) (void)hipMemcpy(h.data(), d, k * sizeof(P), hipMemcpyDeviceToHost);
  return h;
}

template <typename T>
void run_with_ids(size_t n, unsigned seed) {
  using Pt = PointStruct<T, unsigned, 1>;
  std::mt19937 rng(seed);
  std::uniform_int_distribution<int> u(0, 50);
  std::vector<Pt> pts(n);
  for (size_t i = 0; i < n; i++) {
    T a = static_cast<T>(u(rng)), w = static_cast<T>(u(rng) % 7);
    pts[i] = Pt(a, a + w, static_cast<unsigned>(rng() % 1000));
  }
  Pt* d = nullptr;
  (void)hipMalloc(&d, n * sizeof(Pt));
  (void)hipMemcpy(d, pts.data(), n * sizeof(Pt), hipMemcpyHostToDevice);
  hipDeviceProp_t props{};
  (void)hipGetDeviceProperties(&props, 0);
  StaticPSTGPU<T, PointStruct, unsigned, 1> tree(d, n, 1, 0, 1, props);
  // the second half of the queries runs with the level index built
  for (int q = 0; q < 40; q++) {
    if (q == 20) tree.setLevelIndex(true);
    T a = static_cast<T>(u(rng)), b = a + static_cast<T>(u(rng) % 20), c = static_cast<T>(u(rng));
    size_t k = 0;
    unsigned* ids = nullptr;
    tree.template threeSidedSearch<unsigned>(k, ids, a, b, c);
    std::vector<unsigned> got = to_host<T>(ids, k), want;
    (void)hipFree(ids);
    for (auto& p : pts)
      if (a <= p.dim1_val && p.dim1_val <= b && c <= p.dim2_val) want.push_back(p.id);
    std::sort(got.begin(), got.end());
    std::sort(want.begin(), want.end());
    EXPECT(got == want, "threeSidedSearch n=%zu q=%d", n, q);

    Pt* res = nullptr;
    tree.twoSidedLeftSearch(k, res, b, c);  // RetType defaults to the point type
    std::vector<Pt> gp = to_host<T>(res, k);
    (void)hipFree(res);
    size_t wl = 0;
    for (auto& p : pts) wl += (p.dim1_val <= b && c <= p.dim2_val);
    EXPECT(gp.size() == wl, "twoSidedLeftSearch count n=%zu", n);
    for (auto& p : gp) EXPECT(p.dim1_val <= b && c <= p.dim2_val, "twoSidedLeftSearch predicate");

    tree.template twoSidedRightSearch<unsigned>(k, ids, a, c);
    got = to_host<T>(ids, k);
    (void)hipFree(ids);
    want.clear();
    for (auto& p : pts)
      if (a <= p.dim1_val && c <= p.dim2_val) want.push_back(p.id);
    std::sort(got.begin(), got.end());
    std::sort(want.begin(), want.end());
    EXPECT(got == want, "twoSidedRightSearch n=%zu", n);

    unsigned* ips = nullptr;
    intervalParallelSearch<T, PointStruct, unsigned, 1, unsigned>(d, n, ips, k, a, 0, 1, 64, 0, 256);
    got = to_host<T>(ips, k);
    (void)hipFree(ips);
    want.clear();
    for (auto& p : pts)
      if (p.dim1_val <= a && a <= p.dim2_val) want.push_back(p.id);
    std::sort(got.begin(), got.end());
    std::sort(want.begin(), want.end());
    EXPECT(got == want, "intervalParallelSearch n=%zu", n);
  }
  EXPECT(tree.calcGlobalMemNeeded(n) > 0, "calcGlobalMemNeeded");
  (void)hipFree(d);
}

void run_no_ids() {
  using Pt = PointStruct<double, void, 0>;
  std::vector<Pt> pts;
  for (int i = 0; i < 9; i++) pts.emplace_back(0.5 * (i % 4), 0.5 * (i % 4) + 1.0);
  Pt* d = nullptr;
  (void)hipMalloc(&d, pts.size() * sizeof(Pt));
  (void)hipMemcpy(d, pts.data(), pts.size() * sizeof(Pt), hipMemcpyHostToDevice);
  StaticPSTGPU<double, PointStruct> tree(d, pts.size());
  size_t k = 0;
  Pt* res = nullptr;
  tree.threeSidedSearch(k, res, 0.5, 1.0, 1.2);
  size_t want = 0;
  for (auto& p : pts) want += (0.5 <= p.dim1_val && p.dim1_val <= 1.0 && 1.2 <= p.dim2_val);
  EXPECT(k == want, "no-ID threeSidedSearch count %zu vs %zu", k, want);
  (void)hipFree(res);
  std::ostringstream os;
  os << tree;
  EXPECT(os.str().find("├─(R)─ ") != std::string::npos, "print draws the tree");
  (void)hipFree(d);
  // empty tree: (0, nullptr) and the reference's message
  StaticPSTGPU<double, PointStruct> empty(nullptr, 0);
  empty.threeSidedSearch(k, res, 0.0, 1.0, 0.0);
  EXPECT(k == 0 && res == nullptr, "empty tree");
}

// StaticPSTGPUArr as the reference's dataset tester uses it
// (testers/pst-tester-functions.tu:172-183): isovalue (q, q) left searches,
// plus the right and three-sided searches the reference leaves empty.
void run_forest(size_t n, unsigned tpb, unsigned seed) {
  using Pt = PointStruct<float, unsigned, 1>;
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> u(0.0f, 1.0f);
  std::vector<Pt> pts(n);
  for (size_t i = 0; i < n; i++) {
    const float a = u(rng);
    pts[i] = Pt(a, a + 0.05f * u(rng), static_cast<unsigned>(i));
  }
  Pt* d = nullptr;
  (void)hipMalloc(&d, n * sizeof(Pt));
  (void)hipMemcpy(d, pts.data(), n * sizeof(Pt), hipMemcpyHostToDevice);
  StaticPSTGPUArr<float, PointStruct, unsigned, 1> forest(d, n, tpb);
  for (int q = 0; q < 10; q++) {
    const float a = u(rng), b = a + 0.1f * u(rng), c = u(rng);
    for (int kind = 0; kind < 3; kind++) {
      size_t k = 0;
      unsigned* ids = nullptr;
      if (kind == 0) forest.template threeSidedSearch<unsigned>(k, ids, a, b, c);
      if (kind == 1) forest.template twoSidedLeftSearch<unsigned>(k, ids, a, a);
      if (kind == 2) forest.template twoSidedRightSearch<unsigned>(k, ids, a, c);
      std::vector<unsigned> got = to_host<float>(ids, k), want;
      (void)hipFree(ids);
      for (auto& p : pts) {
        const bool hit = kind == 0 ? (a <= p.dim1_val && p.dim1_val <= b && c <= p.dim2_val)
                         : kind == 1 ? (p.dim1_val <= a && a <= p.dim2_val)
                                     : (a <= p.dim1_val && c <= p.dim2_val);
        if (hit) want.push_back(p.id);
      }
      std::sort(got.begin(), got.end());
      std::sort(want.begin(), want.end());
      EXPECT(got == want, "StaticPSTGPUArr kind=%d n=%zu tpb=%u q=%d", kind, n, tpb, q);
    }
  }
  EXPECT(forest.calcGlobalMemNeeded(n, tpb) > 0, "StaticPSTGPUArr::calcGlobalMemNeeded");
  std::ostringstream os;
  os << forest;
  EXPECT(os.str().find("└─(L)─ ") != std::string::npos, "StaticPSTGPUArr print draws trees");
  (void)hipFree(d);
}

// StaticPSTGPU::save / load (beyond the reference): the reloaded tree
// answers every search as the original does.
void run_persist(size_t n, unsigned seed) {
  using Pt = PointStruct<float, unsigned, 1>;
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> u(0.0f, 1.0f);
  std::vector<Pt> pts(n);
  for (size_t i = 0; i < n; i++) {
    const float a = u(rng);
    pts[i] = Pt(a, a + 0.05f * u(rng), static_cast<unsigned>(i));
  }
  Pt* d = nullptr;
  (void)hipMalloc(&d, n * sizeof(Pt));
  (void)hipMemcpy(d, pts.data(), n * sizeof(Pt), hipMemcpyHostToDevice);
  const std::string path = "dropin_persist.pst";
  {
    StaticPSTGPU<float, PointStruct, unsigned, 1> tree(d, n);
    tree.save(path);
    auto lo = StaticPSTGPU<float, PointStruct, unsigned, 1>::load(path);
    for (int q = 0; q < 8; q++) {
      const float a = u(rng), c = u(rng);
      size_t k1 = 0, k2 = 0;
      unsigned *r1 = nullptr, *r2 = nullptr;
      tree.template twoSidedRightSearch<unsigned>(k1, r1, a, c);
      lo->template twoSidedRightSearch<unsigned>(k2, r2, a, c);
      std::vector<unsigned> g1 = to_host<float>(r1, k1), g2 = to_host<float>(r2, k2);
      (void)hipFree(r1);
      (void)hipFree(r2);
      std::sort(g1.begin(), g1.end());
      std::sort(g2.begin(), g2.end());
      EXPECT(g1 == g2, "StaticPSTGPU::load q=%d", q);
    }
    bool threw = false;
    try {
      StaticPSTGPUArr<float, PointStruct, unsigned, 1>::load(path);
    } catch (const std::runtime_error&) {
      threw = true;
    }
    EXPECT(threw, "StaticPSTGPUArr::load rejects a StaticPSTGPU file");
  }
  std::remove(path.c_str());
  (void)hipFree(d);
}

// readInVertices + calcNumMetacells + formMetacellTags as the reference's
// dataset tester uses them (testers/pst-tester-functions.tu:97-136), against
// a brute-force min/max per metacell.
void run_isosurface() {
  unsigned pt_dims[NUM_DIMS] = {21, 14, 9}, mc_dims[NUM_DIMS] = {4, 4, 3}, mc_grid[NUM_DIMS];
  const size_t nv = 21 * 14 * 9;
  std::vector<float> vol(nv);
  for (size_t v = 0; v < nv; v++) vol[v] = std::sin(0.37f * v) * 10.0f;
  const std::string path = "dropin_volume.raw";
  {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(vol.data()), nv * sizeof(float));
  }
  float* host = readInVertices<float>(path, nv);
  EXPECT(std::equal(vol.begin(), vol.end(), host), "readInVertices values");
  float* dv = nullptr;
  (void)hipMalloc(&dv, nv * sizeof(float));
  (void)hipMemcpy(dv, host, nv * sizeof(float), hipMemcpyHostToDevice);
  delete[] host;
  size_t num_metacells = 0;
  calcNumMetacells(pt_dims, mc_dims, mc_grid, num_metacells);
  EXPECT(mc_grid[0] == 5 && mc_grid[1] == 4 && mc_grid[2] == 3 && num_metacells == 60, "calcNumMetacells");
  using Pt = PointStruct<float, unsigned, 1>;
  Pt* cells = formMetacellTags<Pt>(dv, pt_dims, mc_dims, mc_grid, num_metacells, 0, 1, 64);
  std::vector<Pt> got = to_host<float>(cells, num_metacells);
  for (unsigned c = 0; c < mc_grid[2]; c++)
    for (unsigned b = 0; b < mc_grid[1]; b++)
      for (unsigned a = 0; a < mc_grid[0]; a++) {
        float mn = INFINITY, mx = -INFINITY;
        for (unsigned z = c * 3; z <= std::min(c * 3 + 3, 8u); z++)
          for (unsigned y = b * 4; y <= std::min(b * 4 + 4, 13u); y++)
            for (unsigned x = a * 4; x <= std::min(a * 4 + 4, 20u); x++) {
              const float v = vol[x + (y + z * 14) * 21];
              mn = std::min(mn, v);
              mx = std::max(mx, v);
            }
        const unsigned id = a + (b + c * mc_grid[1]) * mc_grid[0];
        EXPECT(got[id].dim1_val == mn && got[id].dim2_val == mx && got[id].id == id, "metacell %u", id);
      }
  (void)hipFree(cells);
  (void)hipFree(dv);
  std::remove(path.c_str());
}

int main() {
  run_with_ids<float>(1000, 1);
  run_with_ids<float>(5000, 2);
  run_with_ids<int>(3000, 3);
  run_with_ids<double>(2049, 4);
  run_no_ids();
  run_forest(10000, 64, 5);
  run_forest(3001, 1000, 6);
  run_persist(20000, 7);
  run_isosurface();
  if (failures) {
    std::printf("DROPIN FAILED (%d)\n", failures);
    return 1;
  }
  std::printf("DROPIN OK\n");
  return 0;
}
