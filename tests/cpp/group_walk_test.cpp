// CPU check of the grouped tree walk used by the batched searches: groups of
// 8 queries walk the tree together through node_step (pst_common.hpp), and
// each query's report set must equal the oracle's own walk for that query
// (oracle/pst_oracle.hpp, Searcher::run). Also checks that the closed-form
// node sizes (subtree_size) give the oracle's child bitcodes. Host code only;
// prints "GROUP WALK OK" on success.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <random>
#include <vector>

#include "pst_common.hpp"
#include "pst_oracle.hpp"

namespace {

int failures = 0;

template <typename T>
struct Q {
  T lo, hi, min2;
  uint32_t kind;  // 0 three-sided, 1 left, 2 right
};

template <typename T, typename I>
struct Tree {
  std::vector<unsigned char> blob;
  pst_oracle::TreeView<T, I> tv;
  size_t n;
};

template <typename T, typename I>
Tree<T, I> build(const std::vector<pst_oracle::Pt<T, I>>& pts) {
  Tree<T, I> t;
  t.n = pts.size();
  auto L = pst_oracle::layout_for<T, I>(t.n);
  t.blob.assign(L.total + 8, 0);
  t.tv = pst_oracle::TreeView<T, I>{t.blob.data(), L};
  pst_oracle::Builder<T, I> b(pts.data(), t.n, t.tv);
  b.run();
  return t;
}

template <typename T, typename I>
std::vector<long long> oracle_walk(const Tree<T, I>& t, const Q<T>& q) {
  std::vector<long long> out;
  pst_oracle::Searcher<T, I> s{t.tv, t.n};
  const auto code = q.kind == 0 ? pst_oracle::THREE_SEARCH
                                : (q.kind == 1 ? pst_oracle::LEFT_SEARCH : pst_oracle::RIGHT_SEARCH);
  s.run(code, q.lo, q.hi, q.min2, [&](long long slot) { out.push_back(slot); });
  std::sort(out.begin(), out.end());
  return out;
}

template <typename T>
uint64_t lead_key(const Q<T>& q) {
  return static_cast<uint64_t>(pst::ord_key(q.kind == pst::KIND_LEFT ? q.hi : q.lo));
}

// The device's order: stable by leading bound, then stable by kind.
template <typename T>
std::vector<uint32_t> group_order(const std::vector<Q<T>>& qs) {
  std::vector<uint32_t> p(qs.size());
  for (size_t i = 0; i < p.size(); i++) p[i] = static_cast<uint32_t>(i);
  std::stable_sort(p.begin(), p.end(), [&](uint32_t a, uint32_t b) { return lead_key(qs[a]) < lead_key(qs[b]); });
  std::stable_sort(p.begin(), p.end(), [&](uint32_t a, uint32_t b) {
    return std::min(qs[a].kind, 2u) < std::min(qs[b].kind, 2u);
  });
  return p;
}

// The search kernel's full-subtree shortcut (pst_search.hip): thresholds
// restated here so the walk below takes the same decisions.
constexpr int64_t kCheckMinCount = 16;  // count pass
constexpr int64_t kCopyMin = 256;       // write pass defaults
constexpr int64_t kCopyMax = 2047;

// Subtree minimum of dim2 per slot, from the oracle's tree.
template <typename T, typename I>
std::vector<T> subtree_min(const Tree<T, I>& t) {
  const size_t S = t.tv.L.S;
  std::vector<T> sm(S);
  const uint64_t M = t.n + 1;
  for (size_t s = S; s-- > 0;) {
    const int64_t m = pst::subtree_size(M, s);
    if (m < 1) continue;
    T v = t.tv.dim2()[s];
    if (m >= 2) v = std::min(v, sm[2 * s + 1]);
    if (m >= 3) v = std::min(v, sm[2 * s + 2]);
    sm[s] = v;
  }
  return sm;
}

template <typename T, typename I>
std::vector<std::vector<long long>> group_walk(const Tree<T, I>& t, const std::vector<Q<T>>& qs, size_t& visits,
                                               int shortcut = 0) {
  constexpr int G = 8;
  std::vector<std::vector<long long>> res(qs.size());
  if (t.n == 0) return res;
  const auto order = group_order(qs);
  const uint64_t M = t.n + 1;
  const T* D1 = t.tv.dim1();
  const T* D2 = t.tv.dim2();
  const T* MD = t.tv.med();
  const std::vector<T> sm = shortcut ? subtree_min(t) : std::vector<T>();
  for (size_t g0 = 0; g0 < order.size(); g0 += G) {
    pst::Bounds<T> qb[G] = {};
    uint32_t qid[G], a = 0, l = 0, h = 0;
    for (int j = 0; j < G; j++) {
      qid[j] = ~0u;
      if (g0 + j >= order.size()) continue;
      const Q<T>& q = qs[order[g0 + j]];
      qid[j] = order[g0 + j];
      qb[j] = pst::Bounds<T>{q.lo, q.hi, q.min2};
      a |= 1u << j;
      l |= (q.kind != pst::KIND_LEFT ? 1u : 0u) << j;
      h |= ((q.kind == pst::KIND_THREE || q.kind == pst::KIND_LEFT) ? 1u : 0u) << j;
    }
    struct Item {
      uint64_t slot;
      uint32_t a, l, h, f;
    };
    std::vector<Item> st{{0, a, l, h, 0}};
    while (!st.empty()) {
      Item it = st.back();
      st.pop_back();
      visits++;
      const int64_t m = pst::subtree_size(M, it.slot);
      if (m < 1) {
        std::printf("FAIL: walk reached an empty slot %llu\n", (unsigned long long)it.slot);
        failures++;
        continue;
      }
      if (shortcut) {
        // shortcut 1: count pass (every full subtree of >= 16 points at
        // once); 2: write pass (subtrees of >= kCopyMin points checked, at
        // most kCopyMax copied); 3: write pass with checks from 16 points
        const int64_t check_min = shortcut == 1 || shortcut == 3 ? kCheckMinCount : kCopyMin;
        const uint32_t cand = m >= check_min ? (it.a & ~it.l & ~it.h & ~it.f) : 0u;
        for (int j = 0; j < G; j++)
          if (((cand >> j) & 1u) && qb[j].min2 <= sm[it.slot]) it.f |= 1u << j;
        const uint32_t bulk = (shortcut == 1 || m <= kCopyMax) ? it.f : 0u;
        if (bulk) {
          int64_t copied = 0;
          for (uint32_t k = 0; (1ll << k) <= m; k++)
            for (uint64_t p = 0; p < (1ull << k); p++) {
              const uint64_t ls = ((it.slot + 1) << k) - 1 + p;
              if (pst::subtree_size(M, ls) < 1) continue;
              copied++;
              for (int j = 0; j < G; j++)
                if ((bulk >> j) & 1u) res[qid[j]].push_back(static_cast<long long>(ls));
            }
          if (copied != m) {
            std::printf("FAIL: subtree of slot %llu enumerates %lld points, size %lld\n",
                        (unsigned long long)it.slot, (long long)copied, (long long)m);
            failures++;
          }
        }
        it.a &= ~bulk;
        it.f &= ~bulk;
        if (it.a == 0) continue;
      }
      const pst::StepOut o =
          pst::node_step<T, G>(D1[it.slot], D2[it.slot], MD[it.slot], m, qb, pst::bound_num<T, G>(qb), it.a, it.l,
                               it.h);
      if (it.l == 0 && it.h == 0) {
        // the report-above shortcut must agree with the full transition
        const pst::StepOut f = pst::node_step_report_above<T, G>(D2[it.slot], m, qb, it.a);
        if (f.rep != o.rep || f.al != o.al || f.ar != o.ar || f.ll != o.ll || f.hl != o.hl || f.lr != o.lr ||
            f.hr != o.hr) {
          std::printf("FAIL: report-above shortcut differs at slot %llu\n", (unsigned long long)it.slot);
          failures++;
        }
      }
      for (int j = 0; j < G; j++)
        if ((o.rep >> j) & 1u) res[qid[j]].push_back(static_cast<long long>(it.slot));
      if (o.al) st.push_back({2 * it.slot + 1, o.al, o.ll, o.hl, it.f & o.al});
      if (o.ar) st.push_back({2 * it.slot + 2, o.ar, o.lr, o.hr, it.f & o.ar});
    }
  }
  for (auto& r : res) std::sort(r.begin(), r.end());
  return res;
}

template <typename T, typename I>
void check_shape(const Tree<T, I>& t) {
  const uint64_t M = t.n + 1;
  const unsigned char* B = t.tv.bits();
  for (size_t s = 0; s < t.tv.L.S; s++) {
    const int64_t m = pst::subtree_size(M, s);
    if (m < 1) continue;
    const unsigned char want = static_cast<unsigned char>((m >= 2 ? 2 : 0) | (m >= 3 ? 1 : 0));
    if (B[s] != want) {
      std::printf("FAIL: n=%zu slot %zu bitcode %d, closed form %d\n", t.n, s, B[s], want);
      failures++;
      return;
    }
  }
}

template <typename T>
T draw(std::mt19937_64& rng, int mode) {
  // mode 0: wide range; 1: few distinct values (ties everywhere)
  if constexpr (std::is_floating_point<T>::value) {
    if (mode == 1) return static_cast<T>(std::uniform_int_distribution<int>(-3, 3)(rng)) / T(2);
    return static_cast<T>(std::uniform_real_distribution<double>(0.0, 1.0)(rng));
  } else {
    if (mode == 1) return static_cast<T>(std::uniform_int_distribution<int>(-3, 3)(rng));
    return static_cast<T>(std::uniform_int_distribution<long long>(0, 1000000)(rng));
  }
}

template <typename T>
std::vector<Q<T>> make_queries(std::mt19937_64& rng, size_t nq, int mode) {
  std::vector<Q<T>> qs(nq);
  for (auto& q : qs) {
    q.kind = static_cast<uint32_t>(rng() % 3);
    q.lo = draw<T>(rng, mode);
    q.hi = draw<T>(rng, mode);
    q.min2 = draw<T>(rng, mode);
    const int shape = static_cast<int>(rng() % 8);
    if (shape == 0) std::swap(q.lo, q.hi);  // often an empty range
    if (shape == 1) q.min2 = std::min(q.lo, q.hi);  // dense: every point in range
    if (shape == 2) q.hi = q.lo;  // stabbing-like
    if constexpr (std::is_floating_point<T>::value) {
      if (shape == 3) q.hi = std::numeric_limits<T>::infinity();
      if (shape == 4) q.lo = -std::numeric_limits<T>::infinity();
      if (shape == 5) q.min2 = -T(0);
      if (shape == 6 && (rng() & 7) == 0) q.hi = std::numeric_limits<T>::quiet_NaN();
    }
  }
  return qs;
}

template <typename T, typename I>
void run(const char* name, size_t n, int mode, uint64_t seed, size_t nq) {
  std::mt19937_64 rng(seed);
  std::vector<pst_oracle::Pt<T, I>> pts(n);
  for (size_t i = 0; i < n; i++) {
    T a = draw<T>(rng, mode), b = draw<T>(rng, mode);
    if (mode == 0 && (rng() & 1)) b = a + (b - a) / T(16);
    pts[i].dim1_val = std::min(a, b);
    pts[i].dim2_val = std::max(a, b);
    if (mode == 1 && (rng() & 3) == 0) pts[i].dim1_val = -T(0);
    if constexpr (pst_oracle::kHasId<T, I>) pts[i].id = static_cast<I>(rng() % (n / 2 + 1));
  }
  const auto t = build(pts);
  check_shape(t);
  const auto qs = make_queries<T>(rng, nq, mode);
  size_t visits = 0, v1 = 0, v2 = 0;
  const auto got = group_walk(t, qs, visits);
  const auto got_count = group_walk(t, qs, v1, 1);
  const auto got_write = group_walk(t, qs, v2, 2);
  size_t v3 = 0;
  const auto got_write16 = group_walk(t, qs, v3, 3);
  size_t reported = 0, single_visits = 0;
  for (size_t i = 0; i < qs.size(); i++) {
    const auto want = oracle_walk(t, qs[i]);
    reported += want.size();
    if (got_count[i] != want || got_write[i] != want || got_write16[i] != want) {
      std::printf("FAIL: %s n=%zu mode=%d query %zu kind %u: full-subtree walk differs (%zu / %zu results, oracle %zu)\n",
                  name, n, mode, i, qs[i].kind, got_count[i].size(), got_write[i].size(), want.size());
      failures++;
      return;
    }
    if (got[i] != want) {
      std::printf("FAIL: %s n=%zu mode=%d query %zu kind %u: %zu results, oracle %zu\n", name, n, mode, i,
                  qs[i].kind, got[i].size(), want.size());
      failures++;
      return;
    }
  }
  (void)single_visits;
  std::printf("ok %s n=%zu mode=%d queries=%zu reported=%zu group-walk visits=%zu (full-subtree: %zu count, %zu write)\n",
              name, n, mode, nq, reported, visits, v1, v2);
}

}  // namespace

int main() {
  for (size_t n : {0, 1, 2, 3, 7, 8, 100, 1000, 4097, 30000}) {
    for (int mode : {0, 1}) {
      run<float, uint32_t>("f32/u32", n, mode, 1000 + n + mode, 200);
      run<int32_t, uint32_t>("i32/u32", n, mode, 2000 + n + mode, 200);
      run<double, uint64_t>("f64/u64", n, mode, 3000 + n + mode, 100);
      run<int64_t, void>("i64/none", n, mode, 4000 + n + mode, 100);
    }
  }
  // subtrees above the copy limit: the write pass flags and walks them on
  run<float, uint32_t>("f32/u32", 200000, 0, 77, 60);
  if (failures) {
    std::printf("%d failures\n", failures);
    return 1;
  }
  std::printf("GROUP WALK OK\n");
  return 0;
}
