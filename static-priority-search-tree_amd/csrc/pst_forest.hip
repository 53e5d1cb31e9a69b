// Batched searches of a forest of shallow trees (StaticPSTGPUArr) with the
// trees held stationary: each wavefront owns a segment of consecutive trees
// and evaluates every query of the batch against one tree at a time, so the
// forest is read from HBM once per pass instead of once per query walk.
//
// The reference searches a forest with one thread block per tree and query
// (pst/static-pst-gpu-arr-search-functions.tu:10-196). Here a wave stages
// one tree of at most 2^L - 1 <= 511 points at a time: its slots that hold
// a point are read once (coalesced) and compacted, in slot order, into LDS
// as (dim1, dim2) pairs and IDs. A lane then owns a query (up to 512 per
// batch, 64 per pass), and per pass:
//  - the tree is skipped when none of the pass's queries may hit it: the
//    query's dim2 bound is above the root's dim2 (the tree's maximum, by
//    heap order), or its dim1 bounds miss the tree's dim1 range;
//  - otherwise every lane tests every staged record (LDS broadcast reads)
//    against its query's predicate, which is exactly the set a PST walk of
//    that tree reports (a walk reports every node whose dim1 meets the
//    query's bounds and whose dim2 reaches min_dim2: heap order keeps such
//    a node's ancestors unpruned, and the median splits steer the walk to
//    it).
// Pass 1 counts the hits of every (query, segment) in registers; a scan per
// query turns them into segment offsets and the query's total; pass 2 has
// each lane write its query's hits at its offset + segment offset + running
// count, so output positions are fixed (no atomics) and results are in
// tree, then slot order. Predicates follow the search kinds of the C ABI:
// THREE min_dim1 <= dim1 <= max_dim1, LEFT dim1 <= max_dim1, any other kind
// dim1 >= min_dim1 (a RIGHT search), each with dim2 >= min_dim2, compared
// in T.
#include <algorithm>
#include <limits>

#include "pst_internal.hpp"

namespace pst {

namespace {

constexpr int kFBlock = 256;                   // four waves per workgroup, one segment each
constexpr int kFWaves = kFBlock / kWave;
constexpr uint32_t kScanBlock = 1024;

template <typename T>
__device__ inline T wave_min(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const T o = __shfl_xor(v, off, kWave);
    v = o < v ? o : v;
  }
  return v;
}
template <typename T>
__device__ inline T wave_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const T o = __shfl_xor(v, off, kWave);
    v = o > v ? o : v;
  }
  return v;
}

// A query as three closed bounds: lo <= dim1 <= hi, dim2 >= min2. A bound
// the kind leaves open is the type's lowest or highest value (infinities
// for floats), which every stored dim1 meets; a NaN bound stays NaN and
// fails every comparison, as it does in the walk.
template <typename T>
struct Box {
  T lo, hi, min2, pad;
};

template <typename T>
__device__ inline Box<T> box_of(const Query<T>& q) {
  constexpr T lowest = std::numeric_limits<T>::has_infinity ? -std::numeric_limits<T>::infinity()
                                                              : std::numeric_limits<T>::lowest();
  constexpr T highest = std::numeric_limits<T>::has_infinity ? std::numeric_limits<T>::infinity()
                                                               : std::numeric_limits<T>::max();
  Box<T> b;
  b.lo = q.kind == KIND_LEFT ? lowest : q.min_dim1;
  b.hi = (q.kind == KIND_THREE || q.kind == KIND_LEFT) ? q.max_dim1 : highest;
  b.min2 = q.min_dim2;
  b.pad = T(0);
  return b;
}

template <typename T>
__device__ inline bool hit(const Box<T>& q, T d1, T d2) {
  return (q.min2 <= d2) & (q.lo <= d1) & (d1 <= q.hi);
}

// Whether the query can hit a tree whose dim1 values lie in [lo1, hi1] and
// whose largest dim2 is max2 (a necessary condition).
template <typename T>
__device__ inline bool may_hit(const Box<T>& q, T lo1, T hi1, T max2) {
  return (q.min2 <= max2) & (q.lo <= hi1) & (lo1 <= q.hi);
}

struct ForestArgs {
  uint32_t ntrees;      // trees of the forest
  uint32_t shift;       // record slots per tree = 2^shift
  uint32_t full_M;      // points + 1 of a full tree
  uint32_t last_M;      // points + 1 of the last tree
  uint32_t nseg;        // segments (one wave each)
  uint32_t seg_trees;   // trees per segment
  uint32_t nq;
};

// Which bounds a 64-query pass compares: FORM 0 all three, FORM 1 no
// lower dim1 bound (every live query's lo is the lowest value, which every
// non-NaN dim1 meets and a NaN dim1 fails through d1 <= hi anyway), FORM 2
// no upper dim1 bound (likewise).
template <int FORM, typename T>
__device__ inline bool hit_f(const Box<T>& q, T d1, T d2) {
  if constexpr (FORM == 1) return (q.min2 <= d2) & (d1 <= q.hi);
  if constexpr (FORM == 2) return (q.min2 <= d2) & (q.lo <= d1);
  return hit(q, d1, d2);
}

template <typename T>
__device__ inline int pass_form(bool mh, const Box<T>& q) {
  constexpr T lowest = std::numeric_limits<T>::has_infinity ? -std::numeric_limits<T>::infinity()
                                                              : std::numeric_limits<T>::lowest();
  constexpr T highest = std::numeric_limits<T>::has_infinity ? std::numeric_limits<T>::infinity()
                                                               : std::numeric_limits<T>::max();
  if (!__ballot(mh && !(q.lo == lowest))) return 1;
  if (!__ballot(mh && !(q.hi == highest))) return 2;
  return 0;
}

template <typename T>
struct Pair {
  T d1, d2;
};

// A tree staged for the wave: the records of its slots that hold a point,
// compacted in slot order, as (dim1, dim2) pairs and IDs in LDS.
template <typename T, typename IS, int R>
struct TreeStage {
  Pair<T> pr[kWave * R];
  typename std::conditional<has_id_v<IS>, IS, uint32_t>::type id[kWave * R];
};

// Stages tree `tree` (R slots per lane: slot lane + 64 r; the last slot of
// the 2^shift is never used) and returns its point count, its dim1 range
// and its largest dim2 (the root's, by heap order). Slots holding no point
// are not read.
template <typename T, typename IS, int R>
__device__ inline uint32_t stage_tree(const Node<T, IS>* __restrict__ nodes, const ForestArgs& F, uint32_t tree,
                                      TreeStage<T, IS, R>& st, T& lo1, T& hi1, T& max2) {
  const uint32_t M = tree + 1 == F.ntrees ? F.last_M : F.full_M;
  const uint64_t base = static_cast<uint64_t>(tree) << F.shift;
  const uint32_t lane = lane_id();
  const uint32_t slots = 1u << F.shift;
  Node<T, IS> rec[R];
  bool v[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t s = lane + r * kWave;
    v[r] = s + 1 < slots && subtree_size32(M, s) > 0;
    if (v[r]) rec[r] = nodes[base + s];
  }
  T mn = __shfl(rec[0].dim1, 0, kWave);  // slot 0 always holds a point (the root)
  T mx = mn;
  max2 = __shfl(rec[0].dim2, 0, kWave);
  __builtin_amdgcn_wave_barrier();  // the previous tree's records are read
  uint32_t nv = 0;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint64_t m = __ballot(v[r]);
    if (v[r]) {
      mn = rec[r].dim1 < mn ? rec[r].dim1 : mn;
      mx = rec[r].dim1 > mx ? rec[r].dim1 : mx;
      const uint32_t k = nv + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
      st.pr[k] = Pair<T>{rec[r].dim1, rec[r].dim2};
      if constexpr (has_id_v<IS>) st.id[k] = rec[r].id;
    }
    nv += static_cast<uint32_t>(__builtin_popcountll(m));
  }
  lo1 = wave_min(mn);
  hi1 = wave_max(mx);
  __builtin_amdgcn_wave_barrier();
  return nv;
}

// Queries per lane pass: lane l of pass p holds query 64 p + l.
constexpr int kQPass = static_cast<int>(kForestDenseMax / kWave);

template <int FORM, typename T>
__device__ inline uint32_t count_run(const Box<T>& q, const Pair<T>* __restrict__ pr, uint32_t nv) {
  uint32_t c = 0;
#pragma unroll 8
  for (uint32_t k = 0; k < nv; k++) {
    const Pair<T> v = pr[k];
    c += hit_f<FORM>(q, v.d1, v.d2) ? 1u : 0u;
  }
  return c;
}

// Pass 1: segcnt[q * nseg + seg] = hits of query q in segment seg. One
// wave per segment, a lane per query: per tree, a 64-query pass runs only
// when one of its queries may hit the tree; it then tests every staged
// record (an LDS broadcast) against each lane's own query.
template <typename T, typename IS, int R>
__global__ __launch_bounds__(kFBlock) void k_forest_count(const Node<T, IS>* __restrict__ nodes, ForestArgs F,
                                                          const Query<T>* __restrict__ queries,
                                                          uint32_t* __restrict__ segcnt) {
  __shared__ TreeStage<T, IS, R> stage_all[kFWaves];
  const uint32_t wave = threadIdx.x / kWave, lane = lane_id();
  const uint32_t seg = blockIdx.x * kFWaves + wave;
  if (seg >= F.nseg) return;
  TreeStage<T, IS, R>& st = stage_all[wave];
  Box<T> qb[kQPass];
  bool live[kQPass];
  uint32_t acc[kQPass];
#pragma unroll
  for (int p = 0; p < kQPass; p++) {
    const uint32_t q = p * kWave + lane;
    live[p] = q < F.nq;
    qb[p] = box_of(queries[live[p] ? q : 0]);
    acc[p] = 0;
  }
  const uint32_t npass = (F.nq + kWave - 1) / kWave;
  const uint32_t t0 = seg * F.seg_trees;
  const uint32_t t1 = std::min(F.ntrees, t0 + F.seg_trees);
  for (uint32_t tree = t0; tree < t1; tree++) {
    T lo1, hi1, max2;
    const uint32_t nv = stage_tree<T, IS, R>(nodes, F, tree, st, lo1, hi1, max2);
#pragma unroll
    for (int p = 0; p < kQPass; p++) {
      if (static_cast<uint32_t>(p) >= npass) break;
      const bool mh = live[p] && may_hit(qb[p], lo1, hi1, max2);
      if (!__ballot(mh)) continue;
      const int form = pass_form(mh, qb[p]);
      uint32_t c;
      if (form == 1) c = count_run<1>(qb[p], st.pr, nv);
      else if (form == 2) c = count_run<2>(qb[p], st.pr, nv);
      else c = count_run<0>(qb[p], st.pr, nv);
      acc[p] += mh ? c : 0u;
    }
  }
#pragma unroll
  for (int p = 0; p < kQPass; p++)
    if (live[p]) segcnt[static_cast<uint64_t>(p * kWave + lane) * F.nseg + seg] = acc[p];
}

// Per query (one workgroup each): exclusive scan of its segment counts in
// place; totals[q] = the query's hit count.
__global__ __launch_bounds__(kScanBlock) void k_forest_seg_scan(uint32_t* __restrict__ segcnt, uint32_t nseg,
                                                                unsigned long long* __restrict__ totals) {
  __shared__ unsigned long long part[kScanBlock / kWave];
  __shared__ unsigned long long carry_s;
  const uint32_t q = blockIdx.x;
  uint32_t* row = segcnt + static_cast<uint64_t>(q) * nseg;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  if (tid == 0) carry_s = 0;
  __syncthreads();
  for (uint32_t base = 0; base < nseg; base += kScanBlock) {
    const uint32_t i = base + tid;
    const unsigned long long v = i < nseg ? row[i] : 0ull;
    unsigned long long inc = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const unsigned long long o = __shfl_up(inc, off, kWave);
      if (static_cast<int>(lane) >= off) inc += o;
    }
    if (lane == kWave - 1) part[wave] = inc;
    __syncthreads();
    unsigned long long before = carry_s;
    for (uint32_t w = 0; w < wave; w++) before += part[w];
    if (i < nseg) row[i] = static_cast<uint32_t>(before + inc - v);
    __syncthreads();
    if (tid == kScanBlock - 1) carry_s = before + inc;
    __syncthreads();
  }
  if (tid == 0) totals[q] = carry_s;
}

// Writes the lane's hits among the staged records from position w on,
// those at positions >= lim dropped; returns the new position.
template <int FORM, bool RP, typename T, typename IS, int R, typename Rec>
__device__ inline uint32_t write_run(const Box<T>& q, const TreeStage<T, IS, R>& st, uint32_t nv, Rec* __restrict__ o,
                                     uint32_t w, uint32_t lim) {
#pragma unroll 4
  for (uint32_t k = 0; k < nv; k++) {
    const Pair<T> v = st.pr[k];
    if (hit_f<FORM>(q, v.d1, v.d2)) {
      if (w < lim) {
        if constexpr (RP) {
          Point<T, IS> pt;
          pt.dim1_val = v.d1;
          pt.dim2_val = v.d2;
          if constexpr (has_id_v<IS>) pt.id = st.id[k];
          o[w] = pt;
        } else {
          o[w] = st.id[k];
        }
      }
      w++;
    }
  }
  return w;
}

// Pass 2: query q's hits of segment seg at out + offsets[q] + segoff[q][seg]
// + running count, in (tree, slot) order; positions at or past the query's
// limit (its capacity, and the output's end) are dropped. A lane per query
// as in pass 1: each lane writes its own query's run.
template <typename T, typename IS, bool RP, int R>
__global__ __launch_bounds__(kFBlock) void k_forest_write(const Node<T, IS>* __restrict__ nodes, ForestArgs F,
                                                          const Query<T>* __restrict__ queries,
                                                          const uint32_t* __restrict__ segoff,
                                                          const unsigned long long* __restrict__ offsets,
                                                          const unsigned long long* __restrict__ caps,
                                                          unsigned long long cap_all, unsigned long long out_total,
                                                          void* __restrict__ out) {
  using Rec = typename std::conditional<RP, Point<T, IS>, IS>::type;
  __shared__ TreeStage<T, IS, R> stage_all[kFWaves];
  const uint32_t wave = threadIdx.x / kWave, lane = lane_id();
  const uint32_t seg = blockIdx.x * kFWaves + wave;
  if (seg >= F.nseg) return;
  TreeStage<T, IS, R>& st = stage_all[wave];
  Box<T> qb[kQPass];
  bool live[kQPass];
  Rec* qo[kQPass];                   // the query's first output record
  uint32_t at[kQPass], lim[kQPass];  // position within the query; positions below lim are written
#pragma unroll
  for (int p = 0; p < kQPass; p++) {
    const uint32_t q = p * kWave + lane;
    live[p] = q < F.nq;
    const uint32_t qq = live[p] ? q : 0;
    qb[p] = box_of(queries[qq]);
    const unsigned long long off = offsets ? offsets[qq] : 0ull;
    const unsigned long long cap = caps ? caps[qq] : cap_all;
    const unsigned long long room = off < out_total ? out_total - off : 0ull;
    const unsigned long long limit = cap < room ? cap : room;
    qo[p] = static_cast<Rec*>(out) + off;
    at[p] = segoff[static_cast<uint64_t>(qq) * F.nseg + seg];
    lim[p] = live[p] ? static_cast<uint32_t>(limit < 0xffffffffull ? limit : 0xffffffffull) : 0u;
  }
  const uint32_t npass = (F.nq + kWave - 1) / kWave;
  const uint32_t t0 = seg * F.seg_trees;
  const uint32_t t1 = std::min(F.ntrees, t0 + F.seg_trees);
  for (uint32_t tree = t0; tree < t1; tree++) {
    T lo1, hi1, max2;
    const uint32_t nv = stage_tree<T, IS, R>(nodes, F, tree, st, lo1, hi1, max2);
#pragma unroll
    for (int p = 0; p < kQPass; p++) {
      if (static_cast<uint32_t>(p) >= npass) break;
      const bool mh = live[p] && may_hit(qb[p], lo1, hi1, max2);
      if (!__ballot(mh)) continue;
      const int form = pass_form(mh, qb[p]);
      if (mh) {
        if (form == 1) at[p] = write_run<1, RP>(qb[p], st, nv, qo[p], at[p], lim[p]);
        else if (form == 2) at[p] = write_run<2, RP>(qb[p], st, nv, qo[p], at[p], lim[p]);
        else at[p] = write_run<0, RP>(qb[p], st, nv, qo[p], at[p], lim[p]);
      }
    }
  }
}

template <typename F>
Status fwith_t(int dt, F&& f) {
  switch (dt) {
    case F32: return f(float{});
    case F64: return f(double{});
    case I32: return f(int32_t{});
    case I64: return f(int64_t{});
  }
  return Status{1, "Error: unsupported data type code " + std::to_string(dt)};
}
template <typename F>
Status fwith_is(int idt, F&& f) {
  switch (idt) {
    case NO_ID: return f(NoId{});
    case F32: case I32: case U32: return f(uint32_t{});
    case F64: case I64: case U64: return f(uint64_t{});
  }
  return Status{1, "Error: unsupported ID type code " + std::to_string(idt)};
}

ForestArgs forest_args(const SearchArgs& a, int cus) {
  ForestArgs F{};
  const uint64_t St = a.slots_per_tree;
  F.shift = 0;
  while ((1ull << F.shift) < St + 1) F.shift++;
  F.ntrees = static_cast<uint32_t>((a.n + St - 1) / St);
  F.full_M = static_cast<uint32_t>(St + 1);
  F.last_M = static_cast<uint32_t>(a.n - (static_cast<uint64_t>(F.ntrees) - 1) * St + 1);
  // about 64 segments per CU (8 waves per SIMD resident, launched in
  // several turns), each a run of consecutive trees
  const uint32_t want = static_cast<uint32_t>(std::max(1, cus)) * 64u;
  F.seg_trees = std::max(1u, (F.ntrees + want - 1) / want);
  F.nseg = (F.ntrees + F.seg_trees - 1) / F.seg_trees;
  F.nq = static_cast<uint32_t>(a.nq);
  return F;
}

}  // namespace

size_t forest_dense_scratch_bytes(size_t n, size_t slots_per_tree, size_t nq, int cus) {
  SearchArgs a{};
  a.n = n;
  a.slots_per_tree = slots_per_tree;
  a.nq = nq;
  const ForestArgs F = forest_args(a, cus);
  return static_cast<size_t>(F.nseg) * nq * 4;
}

Status forest_dense_search(const SearchArgs& a, void* segbuf, hipStream_t s, int plan) {
  if (a.nq == 0 || a.n == 0) return Status{};
  if (a.nq > kForestDenseMax) return Status{1, "Error: too many queries for a stationary forest pass"};
  int dev = 0, cus = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const ForestArgs F = forest_args(a, cus);
  uint32_t* segcnt = static_cast<uint32_t*>(segbuf);
  return fwith_t(a.dtype, [&](auto tv) -> Status {
    using T = decltype(tv);
    return fwith_is(a.id_dtype, [&](auto iv) -> Status {
      using IS = decltype(iv);
      const auto* nodes = static_cast<const Node<T, IS>*>(a.nodes);
      const auto* qs = static_cast<const Query<T>*>(a.queries);
      const unsigned grid = (F.nseg + kFWaves - 1) / kFWaves;
      auto go = [&](auto rtag) -> Status {
        constexpr int R = decltype(rtag)::value;
        const bool count = a.mode == 0 || plan != 2;
        if (count) {
          // plan 2 writes from the segment offsets a plan 1 count pass kept
          hipLaunchKernelGGL((k_forest_count<T, IS, R>), dim3(grid), dim3(kFBlock), 0, s, nodes, F, qs, segcnt);
          // write mode without kept offsets: the per-query totals go to the cursor
          hipLaunchKernelGGL(k_forest_seg_scan, dim3(F.nq), dim3(kScanBlock), 0, s, segcnt, F.nseg, a.counts);
        }
        if (a.mode == 0) return hip_status(hipGetLastError(), "Error in launching forest search");
        if (a.ret_point) {
          hipLaunchKernelGGL((k_forest_write<T, IS, true, R>), dim3(grid), dim3(kFBlock), 0, s, nodes, F, qs,
                             segcnt, a.offsets, a.caps, a.cap_all, a.out_total, a.out);
        } else {
          if constexpr (has_id_v<IS>) {
            hipLaunchKernelGGL((k_forest_write<T, IS, false, R>), dim3(grid), dim3(kFBlock), 0, s, nodes,
                               F, qs, segcnt, a.offsets, a.caps, a.cap_all, a.out_total, a.out);
          } else {
            return Status{1, "Error: IDs requested from a tree built without IDs"};
          }
        }
        return hip_status(hipGetLastError(), "Error in launching forest search");
      };
      switch (F.shift) {
        case 0: case 1: case 2: case 3: case 4: case 5: case 6:
          return go(std::integral_constant<int, 1>{});
        case 7: return go(std::integral_constant<int, 2>{});
        case 8: return go(std::integral_constant<int, 4>{});
        case 9: return go(std::integral_constant<int, 8>{});
      }
      return Status{1, "Error: stationary forest passes hold trees of at most 511 points"};
    });
  });
}

}  // namespace pst
