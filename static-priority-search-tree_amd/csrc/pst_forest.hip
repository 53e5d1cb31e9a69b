// Batched searches of a forest of shallow trees (StaticPSTGPUArr) with the
// trees held stationary: each wavefront owns a segment of consecutive trees
// and evaluates every query of the batch against one tree at a time, so the
// forest is read from HBM once per pass instead of once per query walk.
//
// The reference searches a forest with one thread block per tree and query
// (pst/static-pst-gpu-arr-search-functions.tu:10-196). A tree of at most
// 2^L - 1 <= 2047 points is loaded by one wave as 2^L / 64 records per lane
// (one coalesced 16-byte load each for f32/u32), and per query:
//  - the tree is skipped when the query's dim2 bound is above the root's
//    dim2 (the tree's maximum, by heap order), or its dim1 bounds miss the
//    tree's dim1 range (a wave min/max reduction per tree);
//  - otherwise each lane tests its records against the query predicate,
//    which is exactly the set a PST walk of that tree reports (a walk
//    reports every node whose dim1 meets the query's bounds and whose dim2
//    reaches min_dim2: heap order keeps such a node's ancestors unpruned,
//    and the median splits steer the walk to it), and a ballot per record
//    position gives the wave's hits.
// Pass 1 counts the hits of every (query, segment); a scan per query turns
// them into segment offsets and the query's total; pass 2 writes each
// query's hits at its offset + segment offset + running count, compacted by
// ballot/mbcnt, so output positions are fixed (no atomics) and results are
// in tree order. Predicates follow the search kinds of the C ABI: THREE
// min_dim1 <= dim1 <= max_dim1, LEFT dim1 <= max_dim1, any other kind
// dim1 >= min_dim1 (a RIGHT search), each with dim2 >= min_dim2, compared
// in T.
#include <algorithm>

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

template <typename T>
__device__ inline bool hit(const Query<T>& q, T d1, T d2) {
  if (!(q.min_dim2 <= d2)) return false;
  if (q.kind == KIND_THREE) return q.min_dim1 <= d1 && d1 <= q.max_dim1;
  if (q.kind == KIND_LEFT) return d1 <= q.max_dim1;
  return q.min_dim1 <= d1;
}

// Whether query q can hit a tree whose dim1 values lie in [lo1, hi1] and
// whose largest dim2 is max2 (a necessary condition; NaN bounds fail it).
template <typename T>
__device__ inline bool may_hit(const Query<T>& q, T lo1, T hi1, T max2) {
  if (!(q.min_dim2 <= max2)) return false;
  if (q.kind == KIND_THREE) return q.min_dim1 <= hi1 && lo1 <= q.max_dim1;
  if (q.kind == KIND_LEFT) return lo1 <= q.max_dim1;
  return q.min_dim1 <= hi1;
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

// R records per lane: slot lane + 64 r of the tree.
template <typename T, typename IS, int R>
__device__ inline void load_tree(const Node<T, IS>* __restrict__ nodes, const ForestArgs& F, uint32_t tree,
                                 Node<T, IS> (&rec)[R], bool (&valid)[R], T& lo1, T& hi1, T& max2) {
  const uint32_t M = tree + 1 == F.ntrees ? F.last_M : F.full_M;
  const uint64_t base = static_cast<uint64_t>(tree) << F.shift;
  const uint32_t lane = lane_id();
  const uint32_t slots = 1u << F.shift;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t s = lane + r * kWave;
    if (s < slots) rec[r] = nodes[base + s];
  }
  T mn = rec[0].dim1, mx = rec[0].dim1;  // slot 0 always holds a point (the root)
  mn = __shfl(mn, 0, kWave);
  mx = mn;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t s = lane + r * kWave;
    valid[r] = s + 1 < slots && subtree_size32(M, s) > 0;
    if (valid[r]) {
      mn = rec[r].dim1 < mn ? rec[r].dim1 : mn;
      mx = rec[r].dim1 > mx ? rec[r].dim1 : mx;
    }
  }
  lo1 = wave_min(mn);
  hi1 = wave_max(mx);
  max2 = __shfl(rec[0].dim2, 0, kWave);
}

// Pass 1: segcnt[q * nseg + seg] = hits of query q in segment seg.
template <typename T, typename IS, int R>
__global__ __launch_bounds__(kFBlock) void k_forest_count(const Node<T, IS>* __restrict__ nodes, ForestArgs F,
                                                          const Query<T>* __restrict__ queries,
                                                          uint32_t* __restrict__ segcnt) {
  extern __shared__ unsigned char smem[];
  Query<T>* sq = reinterpret_cast<Query<T>*>(smem);
  uint32_t* cnt_all = reinterpret_cast<uint32_t*>(sq + F.nq);
  for (uint32_t i = threadIdx.x; i < F.nq; i += kFBlock) sq[i] = queries[i];
  const uint32_t wave = threadIdx.x / kWave, lane = lane_id();
  uint32_t* cnt = cnt_all + wave * F.nq;
  for (uint32_t i = lane; i < F.nq; i += kWave) cnt[i] = 0;
  __syncthreads();
  const uint32_t seg = blockIdx.x * kFWaves + wave;
  if (seg >= F.nseg) return;
  const uint32_t t0 = seg * F.seg_trees;
  const uint32_t t1 = std::min(F.ntrees, t0 + F.seg_trees);
  for (uint32_t tree = t0; tree < t1; tree++) {
    Node<T, IS> rec[R];
    bool valid[R];
    T lo1, hi1, max2;
    load_tree<T, IS, R>(nodes, F, tree, rec, valid, lo1, hi1, max2);
    for (uint32_t q = 0; q < F.nq; q++) {
      const Query<T> Q = sq[q];
      if (!may_hit(Q, lo1, hi1, max2)) continue;
      uint32_t c = 0;
#pragma unroll
      for (int r = 0; r < R; r++) c += static_cast<uint32_t>(__builtin_popcountll(__ballot(valid[r] && hit(Q, rec[r].dim1, rec[r].dim2))));
      if (lane == 0) cnt[q] += c;
    }
  }
  __builtin_amdgcn_wave_barrier();
  for (uint32_t q = lane; q < F.nq; q += kWave) segcnt[static_cast<uint64_t>(q) * F.nseg + seg] = cnt[q];
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

// Pass 2: query q's hits of segment seg at out + offsets[q] + segoff[q][seg]
// + running count, those at query positions >= the query's limit dropped.
template <typename T, typename IS, bool RP, int R>
__global__ __launch_bounds__(kFBlock) void k_forest_write(const Node<T, IS>* __restrict__ nodes, ForestArgs F,
                                                          const Query<T>* __restrict__ queries,
                                                          const uint32_t* __restrict__ segoff,
                                                          const unsigned long long* __restrict__ offsets,
                                                          const unsigned long long* __restrict__ caps,
                                                          unsigned long long cap_all, unsigned long long out_total,
                                                          void* __restrict__ out) {
  using Rec = typename std::conditional<RP, Point<T, IS>, IS>::type;
  extern __shared__ unsigned char smem[];
  Query<T>* sq = reinterpret_cast<Query<T>*>(smem);
  unsigned long long* qoff = reinterpret_cast<unsigned long long*>(sq + F.nq);  // the query's output offset
  uint32_t* pos_all = reinterpret_cast<uint32_t*>(qoff + F.nq);  // running position within the query
  uint32_t* lim_all = pos_all + kFWaves * F.nq;                  // positions below this are written
  for (uint32_t i = threadIdx.x; i < F.nq; i += kFBlock) {
    sq[i] = queries[i];
    qoff[i] = offsets ? offsets[i] : 0ull;
  }
  const uint32_t wave = threadIdx.x / kWave, lane = lane_id();
  const uint32_t seg = blockIdx.x * kFWaves + wave;
  uint32_t* pos = pos_all + wave * F.nq;
  uint32_t* lim = lim_all + wave * F.nq;
  Rec* o = static_cast<Rec*>(out);
  if (seg < F.nseg) {
    for (uint32_t q = lane; q < F.nq; q += kWave) {
      // positions of this query that fit its capacity and the output
      const unsigned long long off = offsets ? offsets[q] : 0ull;
      const unsigned long long cap = caps ? caps[q] : cap_all;
      const unsigned long long room = off < out_total ? out_total - off : 0ull;
      const unsigned long long limit = cap < room ? cap : room;
      pos[q] = segoff[static_cast<uint64_t>(q) * F.nseg + seg];
      lim[q] = static_cast<uint32_t>(limit < 0xffffffffull ? limit : 0xffffffffull);
    }
  }
  __syncthreads();
  if (seg >= F.nseg) return;
  const uint32_t t0 = seg * F.seg_trees;
  const uint32_t t1 = std::min(F.ntrees, t0 + F.seg_trees);
  for (uint32_t tree = t0; tree < t1; tree++) {
    Node<T, IS> rec[R];
    bool valid[R];
    T lo1, hi1, max2;
    load_tree<T, IS, R>(nodes, F, tree, rec, valid, lo1, hi1, max2);
    for (uint32_t q = 0; q < F.nq; q++) {
      const Query<T> Q = sq[q];
      if (!may_hit(Q, lo1, hi1, max2)) continue;
      uint32_t at = pos[q];
      const uint32_t limit = lim[q];
      const unsigned long long base = qoff[q];
#pragma unroll
      for (int r = 0; r < R; r++) {
        const bool h = valid[r] && hit(Q, rec[r].dim1, rec[r].dim2);
        const uint64_t b = __ballot(h);
        if (h) {
          const uint32_t p = at + lanes_below(b);
          if (p < limit) {
            if constexpr (RP) {
              Point<T, IS> pt;
              pt.dim1_val = rec[r].dim1;
              pt.dim2_val = rec[r].dim2;
              if constexpr (has_id_v<IS>) pt.id = rec[r].id;
              o[base + p] = pt;
            } else {
              o[base + p] = rec[r].id;
            }
          }
        }
        at += static_cast<uint32_t>(__builtin_popcountll(b));
      }
      if (lane == 0) pos[q] = at;
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
  // about 16 waves per CU, each with a run of consecutive trees
  const uint32_t want = static_cast<uint32_t>(std::max(1, cus)) * 16u;
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
        const size_t lds_count = static_cast<size_t>(F.nq) * (sizeof(Query<T>) + kFWaves * 4);
        const size_t lds_write = static_cast<size_t>(F.nq) * (sizeof(Query<T>) + 8 + kFWaves * 8);
        const bool count = a.mode == 0 || plan != 2;
        if (count) {
          // plan 2 writes from the segment offsets a plan 1 count pass kept
          hipLaunchKernelGGL((k_forest_count<T, IS, R>), dim3(grid), dim3(kFBlock), lds_count, s, nodes, F, qs, segcnt);
          // write mode without kept offsets: the per-query totals go to the cursor
          hipLaunchKernelGGL(k_forest_seg_scan, dim3(F.nq), dim3(kScanBlock), 0, s, segcnt, F.nseg, a.counts);
        }
        if (a.mode == 0) return hip_status(hipGetLastError(), "Error in launching forest search");
        if (a.ret_point) {
          hipLaunchKernelGGL((k_forest_write<T, IS, true, R>), dim3(grid), dim3(kFBlock), lds_write, s, nodes, F, qs,
                             segcnt, a.offsets, a.caps, a.cap_all, a.out_total, a.out);
        } else {
          if constexpr (has_id_v<IS>) {
            hipLaunchKernelGGL((k_forest_write<T, IS, false, R>), dim3(grid), dim3(kFBlock), lds_write, s, nodes,
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
