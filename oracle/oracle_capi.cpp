// C entry points over the CPU oracle (pst_oracle.hpp) for ctypes.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ and by bench.py's cpu_baseline
// leg, never by the product library.
//
// dtype codes are shared with the product C-ABI (include/pst/pst_capi.h):
//   0 f32, 1 f64, 2 i32, 3 i64, 4 u32, 5 u64, -1 no ID.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "pst_oracle.hpp"

using namespace pst_oracle;

namespace {

template <typename F>
int dispatch_t(int dt, F&& f) {
  switch (dt) {
    case 0: return f(float{});
    case 1: return f(double{});
    case 2: return f(int32_t{});
    case 3: return f(int64_t{});
    default: return -1;
  }
}
struct NoId {};
template <typename F>
int dispatch_id(int idt, F&& f) {
  switch (idt) {
    case -1: return f(NoId{});
    case 0: return f(float{});
    case 1: return f(double{});
    case 2: return f(int32_t{});
    case 3: return f(int64_t{});
    case 4: return f(uint32_t{});
    case 5: return f(uint64_t{});
    default: return -1;
  }
}
template <typename X>
using IdOf = typename std::conditional<std::is_same<X, NoId>::value, void, X>::type;

// Writes one result (ID or AoS point) for slot s.
template <typename T, typename I>
inline void emit_one(const TreeView<T, I>& tv, long long s, int ret_point, unsigned char* out, size_t k) {
  if (ret_point) {
    Pt<T, I> pt{};
    pt.dim1_val = tv.dim1()[s];
    pt.dim2_val = tv.dim2()[s];
    if constexpr (kHasId<T, I>) pt.id = tv.ids()[s];
    std::memcpy(out + k * sizeof(Pt<T, I>), &pt, sizeof(Pt<T, I>));
  } else {
    if constexpr (kHasId<T, I>) {
      I v = tv.ids()[s];
      std::memcpy(out + k * sizeof(I), &v, sizeof(I));
    }
  }
}

}  // namespace

extern "C" {

// Size in bytes of the reference-layout tree blob for n points.
size_t oracle_tree_bytes(int dt, int idt, size_t n) {
  size_t r = 0;
  dispatch_t(dt, [&](auto t) {
    using T = decltype(t);
    return dispatch_id(idt, [&](auto i) {
      using I = IdOf<decltype(i)>;
      r = layout_for<T, I>(n).total;
      return 0;
    });
  });
  return r;
}

// Byte offsets of dim1, dim2, med, ids, bitcodes and the slot count.
int oracle_tree_layout(int dt, int idt, size_t n, size_t* out6) {
  return dispatch_t(dt, [&](auto t) {
    using T = decltype(t);
    return dispatch_id(idt, [&](auto i) {
      using I = IdOf<decltype(i)>;
      Layout L = layout_for<T, I>(n);
      out6[0] = L.off_dim1; out6[1] = L.off_dim2; out6[2] = L.off_med;
      out6[3] = L.off_id; out6[4] = L.off_bits; out6[5] = L.S;
      return 0;
    });
  });
}

// Builds the tree for `n` AoS points into `blob` (oracle_tree_bytes bytes,
// zero-filled here first, as the reference value-initialises its array).
int oracle_build(int dt, int idt, const void* pts, size_t n, void* blob) {
  return dispatch_t(dt, [&](auto t) {
    using T = decltype(t);
    return dispatch_id(idt, [&](auto i) {
      using I = IdOf<decltype(i)>;
      Layout L = layout_for<T, I>(n);
      std::memset(blob, 0, L.total);
      TreeView<T, I> tv{static_cast<unsigned char*>(blob), L};
      Builder<T, I> b(static_cast<const Pt<T, I>*>(pts), n, tv);
      b.run();
      return 0;
    });
  });
}

// One search on a tree blob. kind: 0 three-sided, 1 two-sided left,
// 2 two-sided right. Bounds point at host values of type T; unused bounds
// may be null. Writes up to `cap` results, returns the full count in *num.
int oracle_search(int dt, int idt, void* blob, size_t n, int kind, const void* min1p,
                  const void* max1p, const void* min2p, int ret_point, void* out, size_t cap,
                  size_t* num) {
  return dispatch_t(dt, [&](auto t) {
    using T = decltype(t);
    return dispatch_id(idt, [&](auto i) {
      using I = IdOf<decltype(i)>;
      Layout L = layout_for<T, I>(n);
      TreeView<T, I> tv{static_cast<unsigned char*>(blob), L};
      Searcher<T, I> s{tv, n};
      T min1{}, max1{}, min2{};
      if (min1p) std::memcpy(&min1, min1p, sizeof(T));
      if (max1p) std::memcpy(&max1, max1p, sizeof(T));
      if (min2p) std::memcpy(&min2, min2p, sizeof(T));
      Code c = kind == 0 ? THREE_SEARCH : (kind == 1 ? LEFT_SEARCH : RIGHT_SEARCH);
      size_t k = 0;
      unsigned char* o = static_cast<unsigned char*>(out);
      s.run(c, min1, max1, min2, [&](long long slot) {
        if (k < cap) emit_one(tv, slot, ret_point, o, k);
        k++;
      });
      *num = k;
      return 0;
    });
  });
}

// CPU baseline: runs `nq` searches (kinds[q] plus the three bound arrays of
// type T: min_dim1, max_dim1, min_dim2) over one tree with
// `threads` std::threads, each thread owning its output vector (IDs). The
// walk is the oracle's stack walk; the elapsed wall time is steady_clock.
// Returns total reported count in *total and seconds in *secs.
int oracle_time_queries(int dt, int idt, void* blob, size_t n, size_t nq, const int32_t* kinds,
                        const void* min1s, const void* max1s, const void* min2s, int threads,
                        uint64_t* total, double* secs) {
  return dispatch_t(dt, [&](auto t) {
    using T = decltype(t);
    return dispatch_id(idt, [&](auto i) {
      using I = IdOf<decltype(i)>;
      Layout L = layout_for<T, I>(n);
      TreeView<T, I> tv{static_cast<unsigned char*>(blob), L};
      Searcher<T, I> s{tv, n};
      const T* a = static_cast<const T*>(min1s);
      const T* b = static_cast<const T*>(max1s);
      const T* c = static_cast<const T*>(min2s);
      int P = threads < 1 ? 1 : threads;
      std::vector<uint64_t> counts(P, 0);
      auto worker = [&](int w) {
        std::vector<uint64_t> res;
        res.reserve(1 << 16);
        uint64_t cnt = 0;
        for (size_t q = w; q < nq; q += P) {
          res.clear();
          Code code = kinds[q] == 0 ? THREE_SEARCH : (kinds[q] == 1 ? LEFT_SEARCH : RIGHT_SEARCH);
          s.run(code, a[q], b[q], c[q], [&](long long slot) {
            if constexpr (kHasId<T, I>) res.push_back(static_cast<uint64_t>(tv.ids()[slot]));
            else res.push_back(static_cast<uint64_t>(slot));
          });
          cnt += res.size();
        }
        counts[w] = cnt;
      };
      auto t0 = std::chrono::steady_clock::now();
      if (P == 1) {
        worker(0);
      } else {
        std::vector<std::thread> th;
        for (int w = 0; w < P; w++) th.emplace_back(worker, w);
        for (auto& x : th) x.join();
      }
      auto t1 = std::chrono::steady_clock::now();
      uint64_t tot = 0;
      for (auto v : counts) tot += v;
      *total = tot;
      *secs = std::chrono::duration<double>(t1 - t0).count();
      return 0;
    });
  });
}

// Per-query result counts of the oracle's stack walks for a query list
// (kinds[q], bounds of type T), over `threads` std::threads sharing the
// tree: counts[q] = number of nodes the walk reports.
int oracle_count_queries(int dt, int idt, void* blob, size_t n, size_t nq, const int32_t* kinds,
                         const void* min1s, const void* max1s, const void* min2s, int threads,
                         uint64_t* counts) {
  return dispatch_t(dt, [&](auto t) {
    using T = decltype(t);
    return dispatch_id(idt, [&](auto i) {
      using I = IdOf<decltype(i)>;
      Layout L = layout_for<T, I>(n);
      TreeView<T, I> tv{static_cast<unsigned char*>(blob), L};
      Searcher<T, I> s{tv, n};
      const T* a = static_cast<const T*>(min1s);
      const T* b = static_cast<const T*>(max1s);
      const T* c = static_cast<const T*>(min2s);
      const int P = threads < 1 ? 1 : threads;
      auto worker = [&](int w) {
        for (size_t q = w; q < nq; q += P) {
          uint64_t k = 0;
          Code code = kinds[q] == 0 ? THREE_SEARCH : (kinds[q] == 1 ? LEFT_SEARCH : RIGHT_SEARCH);
          s.run(code, a[q], b[q], c[q], [&](long long) { k++; });
          counts[q] = k;
        }
      };
      std::vector<std::thread> th;
      for (int w = 0; w < P; w++) th.emplace_back(worker, w);
      for (auto& x : th) x.join();
      return 0;
    });
  });
}

// Times the oracle's construction of n points (index init + 2 sorts +
// populate), the span the reference's CONSTR_TIMED blocks cover.
int oracle_time_build(int dt, int idt, const void* pts, size_t n, void* blob, double* secs) {
  auto t0 = std::chrono::steady_clock::now();
  int rc = oracle_build(dt, idt, pts, n, blob);
  auto t1 = std::chrono::steady_clock::now();
  *secs = std::chrono::duration<double>(t1 - t0).count();
  return rc;
}

}  // extern "C"
