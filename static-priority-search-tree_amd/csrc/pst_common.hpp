// Shared definitions for the MI355X (gfx950) static priority search tree.
//
// Device-side tree layout (one tree, S = 2^ceil(log2(n+1)) - 1 slots):
//   nodes[S] : { T dim1; T dim2; T med; IDType id; }
//           one record per slot, read with one vector load per visit: a
//           walk needs dim2 (prune key), dim1 and med (bounds and split) and
//           the ID (payload) of every node it reports.
// Children of slot s sit at 2s+1 and 2s+2 (pst/gpu-tree-node.h:11-13 of the
// reference). No bitcode array is stored: the tree shape depends on n only
// (a node holding m points gives floor(m/2) to its left child and
// ceil(m/2)-1 to its right child, pst/static-pst-cpu-iter.tpp:512-545), so a
// node's point count, and hence which children exist, is computed in closed
// form (subtree_size below). The reference-layout blob with bitcodes is
// produced on request by the export entry point.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>

namespace pst {

// dtype codes of the C-ABI (include/pst/pst_capi.h)
enum DType : int { F32 = 0, F64 = 1, I32 = 2, I64 = 3, U32 = 4, U64 = 5, NO_ID = -1 };

// Query kinds of the C-ABI.
enum Kind : uint32_t { KIND_THREE = 0, KIND_LEFT = 1, KIND_RIGHT = 2 };

struct NoId {};

template <typename T, typename I>
struct Node {
  T dim1;
  T dim2;
  T med;
  I id;
};
template <typename T>
struct Node<T, NoId> {
  T dim1;
  T dim2;
  T med;
};

// AoS input/output record with the reference's PointStruct layout
// (helpers/point-struct.h:25-33, 107-117).
template <typename T, typename I>
struct Point {
  T dim1_val;
  T dim2_val;
  I id;
};
template <typename T>
struct Point<T, NoId> {
  T dim1_val;
  T dim2_val;
};

template <typename I>
constexpr bool has_id_v = !std::is_same<I, NoId>::value;

// Query record for batched searches: bounds in T plus the kind.
template <typename T>
struct Query {
  T min_dim1;
  T max_dim1;
  T min_dim2;
  uint32_t kind;
};

// ---------------------------------------------------------------------------
// Order-preserving unsigned keys. IEEE compares treat -0.0 == +0.0, so both
// map to the key of +0.0 and the tie-break decides, as in the reference's
// compareDim1/compareDim2 (helpers/point-struct.h:62-68, 154-161).
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t ord_key(float v) {
  uint32_t u = __builtin_bit_cast(uint32_t, v);
  if ((u & 0x7fffffffu) == 0) u = 0;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline uint64_t ord_key(double v) {
  uint64_t u = __builtin_bit_cast(uint64_t, v);
  if ((u & 0x7fffffffffffffffull) == 0) u = 0;
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline uint32_t ord_key(int32_t v) { return static_cast<uint32_t>(v) ^ 0x80000000u; }
__host__ __device__ inline uint64_t ord_key(int64_t v) {
  return static_cast<uint64_t>(v) ^ 0x8000000000000000ull;
}
__host__ __device__ inline uint32_t ord_key(uint32_t v) { return v; }
__host__ __device__ inline uint64_t ord_key(uint64_t v) { return v; }

template <typename V>
__host__ __device__ inline bool is_nan(V v) {
  if constexpr (std::is_floating_point<V>::value) return v != v;
  else return false;
}

// ---------------------------------------------------------------------------
// Tree shape (data independent). With M = n + 1, the node at depth t and
// position j (slot 2^t - 1 + j) holds
//     floor(M / 2^t) - 1 + [brev_t(j) < (M mod 2^t)]
// points: the recursion m -> (floor(m/2), ceil(m/2)-1) splits M = m+1 into
// (ceil(M/2), floor(M/2)), i.e. repeated halving with the larger half on the
// left, whose parts are floor(M/2^t) or that plus one, the larger ones at the
// bit-reversed positions below M mod 2^t.
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t brev_bits(uint32_t x, uint32_t t) {
  if (t == 0) return 0;
#if defined(__HIP_DEVICE_COMPILE__)
  return __brev(x) >> (32u - t);
#else
  uint32_t r = 0;
  for (uint32_t b = 0; b < t; b++)
    if ((x >> b) & 1u) r |= 1u << (t - 1 - b);
  return r;
#endif
}

__host__ __device__ inline uint32_t depth_of(uint64_t slot) {
  // floor(log2(slot + 1))
#if defined(__HIP_DEVICE_COMPILE__)
  return 63u - static_cast<uint32_t>(__clzll(static_cast<long long>(slot + 1)));
#else
  return 63u - static_cast<uint32_t>(__builtin_clzll(slot + 1));
#endif
}

// Point count of level-t position j in a subtree whose root holds M-1
// points (negative when the position is past the last level).
__host__ __device__ inline int64_t level_size(uint64_t M, uint32_t t, uint32_t j) {
  const int64_t base = static_cast<int64_t>(M >> t);
  const uint64_t r = M & ((1ull << t) - 1ull);
  return base - 1 + (static_cast<uint64_t>(brev_bits(j, t)) < r ? 1 : 0);
}

__host__ __device__ inline int64_t subtree_size(uint64_t M, uint64_t slot) {
  const uint32_t t = depth_of(slot);
  const uint32_t j = static_cast<uint32_t>(slot + 1 - (1ull << t));
  return level_size(M, t, j);
}

// subtree_size for slot < 2^31 and M < 2^31, in 32-bit arithmetic (the
// walks' per-visit form).
__host__ __device__ inline int32_t subtree_size32(uint32_t M, uint32_t slot) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint32_t t = 31u - static_cast<uint32_t>(__clz(static_cast<int>(slot + 1)));
#else
  const uint32_t t = 31u - static_cast<uint32_t>(__builtin_clz(slot + 1));
#endif
  const uint32_t j = slot + 1 - (1u << t);
  const uint32_t r = M & ((1u << t) - 1u);
  return static_cast<int32_t>(M >> t) - 1 + (brev_bits(j, t) < r ? 1 : 0);
}

// Number of points held by level-t positions 0..j-1 (their segments are
// consecutive in key1 order), in O(t): position j' < j is grouped by the
// highest bit where it differs from j; each group is a set of positions
// sharing a top prefix, whose bit reversals are an arithmetic progression.
// Meaningful while all positions before j exist (level_size >= 0).
__host__ __device__ inline int64_t level_start(uint64_t M, uint32_t t, uint32_t j) {
  const int64_t base = static_cast<int64_t>(M >> t) - 1;
  const uint64_t r = M & ((1ull << t) - 1ull);
  int64_t extra = 0;
  for (uint32_t b = 0; b < t; b++) {
    if ((j >> b) & 1u) {
      const uint32_t prefix = (j >> b) ^ 1u;
      const uint64_t c = brev_bits(prefix, t - b);
      if (r > c) {
        const uint64_t step = 1ull << (t - b);
        const uint64_t cnt = (r - c + step - 1) / step;
        const uint64_t cap = 1ull << b;
        extra += static_cast<int64_t>(cnt < cap ? cnt : cap);
      }
    }
  }
  return static_cast<int64_t>(j) * base + extra;
}

// ---------------------------------------------------------------------------
// Search state of a query at a node, as bits j of three masks for the G
// queries of a group: `a` the query is walking this subtree, `l` its lower
// dim1 bound still cuts the subtree, `h` its upper bound still does. These
// are the reference's search codes (pst/static-pst-gpu.h:408-414):
// REPORT_ABOVE = a only, RIGHT_SEARCH = a+l, LEFT_SEARCH = a+h,
// THREE_SEARCH = a+l+h. A closed bound holds for every point of the subtree,
// because a left subtree's dim1 values are <= the parent's med and a right
// subtree's are >= it.
// ---------------------------------------------------------------------------
template <typename T>
struct Bounds {
  T lo, hi, min2;
};

// Initial state bits of a query of the given kind (codes other than 0..2
// walk as right searches).
__host__ __device__ inline uint32_t lo_open(uint32_t kind) { return kind != KIND_LEFT ? 1u : 0u; }
__host__ __device__ inline uint32_t hi_open(uint32_t kind) {
  return (kind == KIND_THREE || kind == KIND_LEFT) ? 1u : 0u;
}

struct StepOut {
  uint32_t rep;         // queries that report the node
  uint32_t al, ll, hl;  // state of the left child (al == 0: not visited)
  uint32_t ar, lr, hr;  // state of the right child
};

// One node visit for G queries. d1, d2, md: the node's dim1, dim2 and med;
// m: its point count (left child exists iff m >= 2, right iff m >= 3).
// Per query this is the reference's delegation, with its comparisons
// (pst/static-pst-cpu-iter.tpp:600-735): a node with dim2 < min_dim2 ends
// the query's walk below it; otherwise it is reported if its dim1 meets the
// open bounds, and its children inherit
//   REPORT_ABOVE -> both children REPORT_ABOVE;
//   LEFT_SEARCH  -> med <= hi ? (left REPORT_ABOVE, right LEFT) : left LEFT;
//   RIGHT_SEARCH -> med >= lo ? (left RIGHT, right REPORT_ABOVE) : right RIGHT;
//   THREE_SEARCH -> lo <= med <= hi ? (left RIGHT, right LEFT)
//                   : hi < med ? left THREE : med < lo ? right THREE : none.
// Bit j set when query j's lower (upper) dim1 bound is a number. A node's
// med is never NaN (construction rejects NaN points), so for such a bound
// md < lo is the complement of md >= lo and hi < md that of md <= hi; a NaN
// bound fails every comparison.
struct BoundNum {
  uint32_t lo, hi;
};
template <typename T, int G>
__host__ __device__ inline BoundNum bound_num(const Bounds<T>* qb) {
  BoundNum b{0, 0};
#pragma unroll
  for (int j = 0; j < G; j++) {
    b.lo |= (qb[j].lo == qb[j].lo) ? (1u << j) : 0u;
    b.hi |= (qb[j].hi == qb[j].hi) ? (1u << j) : 0u;
  }
  return b;
}

template <typename T, int G>
__host__ __device__ inline StepOut node_step(T d1, T d2, T md, int32_t m, const Bounds<T>* qb, BoundNum bn,
                                             uint32_t a, uint32_t l, uint32_t h) {
  uint32_t ge2 = 0, d1lo = 0, d1hi = 0, mdhi = 0, mdlo = 0;
#pragma unroll
  for (int j = 0; j < G; j++) {
    const uint32_t bit = 1u << j;
    ge2 |= (qb[j].min2 <= d2) ? bit : 0u;
    d1lo |= (qb[j].lo <= d1) ? bit : 0u;
    d1hi |= (d1 <= qb[j].hi) ? bit : 0u;
    mdhi |= (md <= qb[j].hi) ? bit : 0u;
    mdlo |= (md >= qb[j].lo) ? bit : 0u;
  }
  const uint32_t himd = ~mdhi & bn.hi, mdltlo = ~mdlo & bn.lo;
  const uint32_t act = a & ge2;
  const uint32_t ra = act & ~l & ~h, lf = act & h & ~l, rt = act & l & ~h, th = act & l & h;
  const uint32_t split = th & mdlo & mdhi;
  const uint32_t tl = th & ~split & himd;
  const uint32_t tr = th & ~split & ~himd & mdltlo;
  StepOut o;
  o.rep = act & (~l | d1lo) & (~h | d1hi);
  o.al = ra | lf | (rt & mdlo) | split | tl;
  o.ll = (rt & mdlo) | split | tl;
  o.hl = (lf & ~mdhi) | tl;
  o.ar = ra | (lf & mdhi) | rt | split | tr;
  o.lr = (rt & ~mdlo) | tr;
  o.hr = (lf & mdhi) | split | tr;
  if (m < 2) o.al = o.ll = o.hl = 0;
  if (m < 3) o.ar = o.lr = o.hr = 0;
  return o;
}

// node_step for items with no open bound (every query REPORT_ABOVE): only
// the dim2 prune applies, and both children inherit the surviving queries.
template <typename T, int G>
__host__ __device__ inline StepOut node_step_report_above(T d2, int32_t m, const Bounds<T>* qb, uint32_t a) {
  uint32_t ge2 = 0;
#pragma unroll
  for (int j = 0; j < G; j++) ge2 |= (qb[j].min2 <= d2) ? (1u << j) : 0u;
  StepOut o;
  o.rep = a & ge2;
  o.al = m >= 2 ? o.rep : 0u;
  o.ar = m >= 3 ? o.rep : 0u;
  o.ll = o.hl = o.lr = o.hr = 0;
  return o;
}

// ---------------------------------------------------------------------------
// Wave64 helpers (gfx950 wavefront = 64 lanes).
// ---------------------------------------------------------------------------
constexpr int kWave = 64;

__device__ inline uint32_t lane_id() { return __lane_id(); }

// Exclusive count of set bits of `mask` below this lane.
__device__ inline uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

template <typename V>
__device__ inline V shfl(V v, int src) {
  return __shfl(v, src, kWave);
}

}  // namespace pst
