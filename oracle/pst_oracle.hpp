// CPU oracle for the static priority search tree (PST) hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is a clean-room restatement of the
// reference's single-threaded iterative tree (StaticPSTCPUIter) and is used
// by the test-suite as the checker and by bench.py's cpu_baseline leg. The
// product library (libpst_hip.so) never links, loads or calls it.
//
// What is restated, and from where (paths relative to the reference root):
//   * point orders  ............ helpers/point-struct.h:54-92, 146-188
//       key1 = (dim1, id, address) ascending, key2 = (dim2, id, address);
//       values compared with IEEE `!=` and `<` in the input type T, so
//       -0.0 and +0.0 tie and fall through to the tie-break.
//   * construction ............. pst/static-pst-cpu-iter.tpp:16-177 (ctor),
//       pst/static-pst-cpu-iter-populate-tree.tpp:9-114 (stack build and the
//       level-parity swap of the two dim2 index arrays),
//       pst/static-pst-cpu-iter.tpp:488-598 (constructNode: median rule,
//       dim1 shift, stable dim2 partition), :791-818 (binarySearch)
//   * searches ................. pst/static-pst-cpu-iter.tpp:197-486 and the
//       delegation helpers :600-735 (LIFO stack of (node, code) pairs)
//   * memory layout ............ pst/static-pst-cpu-iter.h:132-155 and
//       pst/static-pst-cpu-iter.tpp:737-789: one blob holding
//       dim1[S] dim2[S] med[S] (T), ids[S] (IDType, aligned up), bitcodes[S]
//   * slot count ............... S(n) = 2^ceil(log2(n+1)) - 1
//       (helpers/power-of-2-functions.h:95-106, static-pst-cpu-iter.tpp:21)
//   * bitcodes ................. HAS_LEFT = 0x2, HAS_RIGHT = 0x1
//       (pst/cpu-iter-tree-node.h:42-46)
//
// Parity of this restatement is pinned by tests/test_oracle.py against
//   (1) golden fixtures written by the reference's own StaticPSTCPUIter
//       (compiled from its sources by oracle/Makefile, target `ref`),
//   (2) a brute-force scan for every search,
//   (3) the closed-form construction (node = max key2, median = L-th of the
//       rest with L = floor(m/2)) written independently in the tests,
//   (4) structural invariants (heap order, BST split, permutation of ids).
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stack>
#include <type_traits>
#include <vector>

namespace pst_oracle {

// AoS point record; identical layout to the reference's PointStruct
// (helpers/point-struct.h:25-33, 107-117).
template <typename T, typename I>
struct Pt {
  T dim1_val;
  T dim2_val;
  I id;
};
template <typename T>
struct Pt<T, void> {
  T dim1_val;
  T dim2_val;
};

template <typename T, typename I>
constexpr bool kHasId = !std::is_void<I>::value;

// Three-way orders. `a` and `b` are indices into the same point array, so
// "address" order is index order.
template <typename T, typename I>
inline int tiebreak(const Pt<T, I>* p, size_t a, size_t b) {
  if constexpr (kHasId<T, I>) {
    if (p[a].id < p[b].id) return -1;
    if (p[a].id > p[b].id) return 1;
  }
  return a == b ? 0 : (a < b ? -1 : 1);
}
template <typename T, typename I>
inline int cmp1(const Pt<T, I>* p, size_t a, size_t b) {
  if (p[a].dim1_val != p[b].dim1_val) return p[a].dim1_val < p[b].dim1_val ? -1 : 1;
  return tiebreak(p, a, b);
}
template <typename T, typename I>
inline int cmp2(const Pt<T, I>* p, size_t a, size_t b) {
  if (p[a].dim2_val != p[b].dim2_val) return p[a].dim2_val < p[b].dim2_val ? -1 : 1;
  return tiebreak(p, a, b);
}

inline size_t slots_for(size_t n) {
  if (n == 0) return 0;
  size_t p = 1;
  while (p <= n) p <<= 1;
  return p - 1;
}

// Byte offsets of the five sub-arrays inside one tree blob; follows the
// reference's getIDsRoot / getBitcodesRoot / calcTotArrSize* arithmetic.
struct Layout {
  size_t S = 0;
  size_t off_dim1 = 0, off_dim2 = 0, off_med = 0, off_id = 0, off_bits = 0;
  size_t total = 0;  // bytes
};
template <typename T, typename I>
inline Layout layout_for(size_t n) {
  Layout L;
  L.S = slots_for(n);
  const size_t sT = sizeof(T);
  L.off_dim1 = 0;
  L.off_dim2 = L.S * sT;
  L.off_med = 2 * L.S * sT;
  size_t after_vals = 3 * L.S * sT;
  if constexpr (kHasId<T, I>) {
    const size_t sI = sizeof(I);
    L.off_id = ((after_vals + sI - 1) / sI) * sI;
    L.off_bits = L.off_id + L.S * sI;
  } else {
    L.off_id = after_vals;
    L.off_bits = after_vals;
  }
  // Every sub-array fits; rounded to 8 bytes. (The reference sizes its blob
  // in units of T, which can fall a few bytes short of the bitcode tail when
  // sizeof(IDType) > sizeof(T); the sub-array offsets above are what matter.)
  L.total = ((L.off_bits + L.S + 7) / 8) * 8;
  return L;
}

template <typename T, typename I>
struct TreeView {
  unsigned char* base;
  Layout L;
  T* dim1() const { return reinterpret_cast<T*>(base + L.off_dim1); }
  T* dim2() const { return reinterpret_cast<T*>(base + L.off_dim2); }
  T* med() const { return reinterpret_cast<T*>(base + L.off_med); }
  unsigned char* bits() const { return base + L.off_bits; }
  template <typename J = I>
  J* ids() const { return reinterpret_cast<J*>(base + L.off_id); }
};

enum : unsigned char { HAS_RIGHT = 0x1, HAS_LEFT = 0x2 };
enum Code : unsigned char { REPORT_ABOVE = 0, LEFT_SEARCH = 1, RIGHT_SEARCH = 2, THREE_SEARCH = 3 };

// ---------------------------------------------------------------------------
// Construction (restates the iterative CPU tree).
// ---------------------------------------------------------------------------
template <typename T, typename I>
class Builder {
 public:
  Builder(const Pt<T, I>* pts, size_t n, TreeView<T, I> tv) : p_(pts), n_(n), tv_(tv) {}

  void run() {
    if (n_ == 0) return;
    std::vector<size_t> d1(n_), d2a(n_), d2b(n_);
    for (size_t i = 0; i < n_; i++) d1[i] = d2a[i] = i;
    const Pt<T, I>* p = p_;
    std::sort(d1.begin(), d1.end(), [p](size_t a, size_t b) { return cmp1(p, a, b) < 0; });
    std::sort(d2a.begin(), d2a.end(), [p](size_t a, size_t b) { return cmp2(p, a, b) > 0; });
    populate(d1.data(), d2a.data(), d2b.data());
  }

 private:
  // Position of element `e` inside the dim1-sorted run [lo, lo+cnt).
  long long locate(const size_t* d1, size_t e, size_t lo, size_t cnt) const {
    size_t a = lo, b = lo + cnt;
    while (a < b) {
      size_t mid = (a + b) / 2;
      int c = cmp1(p_, e, d1[mid]);
      if (c == 0) return static_cast<long long>(mid);
      if (c < 0) b = mid; else a = mid + 1;
    }
    return -1;
  }

  static unsigned level_parity_primary(size_t slot) {
    // number of bits in (slot+1); odd levels (1-based) keep the original
    // primary/secondary roles of the two dim2 arrays.
    size_t v = slot + 1;
    unsigned bits = 0;
    while (v) { bits++; v >>= 1; }
    return bits & 1u;
  }

  void populate(size_t* d1, size_t* d2a, size_t* d2b) {
    struct Job { size_t start, cnt, slot; };
    std::stack<Job> st;
    st.push({0, n_, 0});
    while (!st.empty()) {
      Job j = st.top();
      st.pop();
      size_t* cur = level_parity_primary(j.slot) ? d2a : d2b;
      size_t* nxt = level_parity_primary(j.slot) ? d2b : d2a;
      const size_t node_pt = cur[j.start];
      const long long q = locate(d1, node_pt, j.start, j.cnt);
      const size_t qi = static_cast<size_t>(q);

      size_t med_pos, left_cnt, right_cnt;
      if (j.cnt == 1) {
        med_pos = j.start; left_cnt = 0; right_cnt = 0;
      } else if (qi < j.start + j.cnt / 2) {
        med_pos = j.start + j.cnt / 2;
        left_cnt = med_pos - j.start;
        right_cnt = j.start + j.cnt - med_pos - 1;
      } else {
        med_pos = j.start + j.cnt / 2 - 1;
        left_cnt = med_pos - j.start + 1;
        right_cnt = j.start + j.cnt - med_pos - 2;
      }
      // node record: the max-dim2 point, and the median's dim1 read before
      // the removal shift (the shift never touches med_pos itself).
      tv_.dim1()[j.slot] = p_[node_pt].dim1_val;
      tv_.dim2()[j.slot] = p_[node_pt].dim2_val;
      tv_.med()[j.slot] = p_[d1[med_pos]].dim1_val;
      if constexpr (kHasId<T, I>) tv_.ids()[j.slot] = p_[node_pt].id;

      unsigned char bc = 0;
      if (left_cnt > 0) {
        bc |= HAS_LEFT;
        if (qi < med_pos)
          for (size_t i = qi; i < med_pos; i++) d1[i] = d1[i + 1];
      }
      if (right_cnt > 0) {
        bc |= HAS_RIGHT;
        if (qi > med_pos)
          for (size_t i = qi; i < j.start + j.cnt - 1; i++) d1[i] = d1[i + 1];
      }
      tv_.bits()[j.slot] = bc;
      const size_t right_start = med_pos + 1;
      if (bc) {
        size_t li = j.start, ri = right_start;
        const size_t med_pt = d1[med_pos];
        for (size_t i = j.start + 1; i < j.start + j.cnt; i++) {
          if (cmp1(p_, cur[i], med_pt) <= 0) nxt[li++] = cur[i];
          else nxt[ri++] = cur[i];
        }
      }
      if (right_cnt > 0) st.push({right_start, right_cnt, 2 * j.slot + 2});
      if (left_cnt > 0) st.push({j.start, left_cnt, 2 * j.slot + 1});
    }
  }

  const Pt<T, I>* p_;
  size_t n_;
  TreeView<T, I> tv_;
};

// ---------------------------------------------------------------------------
// Searches (restates the stack walks). Results are appended to `out`; the
// caller decides whether an entry is an ID or a whole point.
// ---------------------------------------------------------------------------
template <typename T, typename I>
struct Searcher {
  TreeView<T, I> tv;
  size_t n;

  template <typename Emit>
  void run(Code start, T min1, T max1, T min2, Emit&& emit) const {
    if (n == 0) return;
    // (node, code) pairs, LIFO as in the reference.
    std::vector<std::pair<long long, unsigned char>> st;
    st.reserve(256);
    st.push_back({0, start});
    const T* D1 = tv.dim1();
    const T* D2 = tv.dim2();
    const T* MD = tv.med();
    const unsigned char* B = tv.bits();
    while (!st.empty()) {
      auto [s, code] = st.back();
      st.pop_back();
      const T d1 = D1[s], d2 = D2[s], md = MD[s];
      const unsigned char bc = B[s];
      if (!(min2 <= d2)) continue;
      bool rep;
      switch (start) {
        case THREE_SEARCH: rep = (min1 <= d1 && d1 <= max1); break;
        case LEFT_SEARCH: rep = (d1 <= max1); break;
        default: rep = (d1 >= min1); break;
      }
      if (rep) emit(s);
      if (!(bc & (HAS_LEFT | HAS_RIGHT))) continue;
      const long long lc = 2 * s + 1, rc = 2 * s + 2;
      switch (code) {
        case THREE_SEARCH:
          if (min1 <= md && md <= max1) {
            if (bc & HAS_LEFT) st.push_back({lc, RIGHT_SEARCH});
            if (bc & HAS_RIGHT) st.push_back({rc, LEFT_SEARCH});
          } else if (max1 < md && (bc & HAS_LEFT)) {
            st.push_back({lc, THREE_SEARCH});
          } else if (md < min1 && (bc & HAS_RIGHT)) {
            st.push_back({rc, THREE_SEARCH});
          }
          break;
        case LEFT_SEARCH:
          if (md <= max1) {
            if (bc & HAS_LEFT) st.push_back({lc, REPORT_ABOVE});
            if (bc & HAS_RIGHT) st.push_back({rc, LEFT_SEARCH});
          } else if (bc & HAS_LEFT) {
            st.push_back({lc, LEFT_SEARCH});
          }
          break;
        case RIGHT_SEARCH:
          if (md >= min1) {
            if (bc & HAS_LEFT) st.push_back({lc, RIGHT_SEARCH});
            if (bc & HAS_RIGHT) st.push_back({rc, REPORT_ABOVE});
          } else if (bc & HAS_RIGHT) {
            st.push_back({rc, RIGHT_SEARCH});
          }
          break;
        default:
          if (bc & HAS_LEFT) st.push_back({lc, REPORT_ABOVE});
          if (bc & HAS_RIGHT) st.push_back({rc, REPORT_ABOVE});
          break;
      }
    }
  }
};

}  // namespace pst_oracle
