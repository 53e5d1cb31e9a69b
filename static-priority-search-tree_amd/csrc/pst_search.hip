// 2-sided and 3-sided searches on one deep tree (StaticPSTGPU), for one
// query or a batch of them, on gfx950.
//
// The reference walks the tree with a single thread block, two block
// barriers per node step, an atomicCAS hand-off of split work to idle
// threads and dynamic-parallelism child grids for overflow
// (pst/static-pst-gpu-search-functions.tu:12-722, pst/static-pst-gpu.tu:
// 812-1148). Here every wavefront is an independent worker:
//
//   * a batch's queries are sorted by their leading dim1 bound and cut into
//     groups of G (8) neighbours; one walk serves a whole group. A work
//     item is (group, slot, per-query state masks, node_step in
//     pst_common.hpp): a node is loaded once for all queries of the group
//     that still walk its subtree, and each query's report set is exactly
//     the one its own walk would give. Neighbouring queries share most of
//     their paths and reported subtrees, so one node load feeds several
//     queries' outputs. Single searches use G = 1 and 32-bit items;
//   * a wave owns one group at a time and walks it depth first with a stack
//     of items in LDS, popping up to 64 items per step (one per lane);
//   * results are staged per query in LDS and flushed with one atomic per
//     flush on the query's output cursor (the reference's warp scan +
//     atomicAdd, helpers/calc-alloc-report-ind-offset.h:17-44, becomes
//     ballot/mbcnt within the wave and one atomic per 64..256 results);
//   * load balance: the launch is repeated for a few rounds; in all but the
//     last, a wave that spent its step budget on one batch pushes its
//     remaining stack to a shared list that the next round spreads over all
//     waves. The last round runs every stack to empty;
//   * whole subtrees that a query reports are counted at once (count pass)
//     or recorded as copy tasks that a streaming kernel writes after the
//     walk (write pass), so the walk stops at their roots.
// Whether a child exists follows from n alone (pst_common.hpp), so no
// bitcode is read.
#include <algorithm>

#include "pst_internal.hpp"

namespace pst {

namespace {

constexpr int kMaxRounds = 8;
constexpr uint64_t kEmpty = ~0ull;
// Per-wave LDS stack (64-bit items); deeper stacks move their oldest
// kMoveBlock items at a time to a per-wave overflow stack in device memory
// (the walk pops the newest 64 items per step, so the stack can grow by 64
// per tree level). One-query walks are latency bound (a refill from device
// memory stalls the only wave on a subtree) and get a deeper LDS stack;
// group walks keep a small one for occupancy.
template <int G>
constexpr int lds_stack() {
  return G == 1 ? 1024 : 512;
}
constexpr int kMoveBlock = 256;

// Full-subtree shortcut: a REPORT_ABOVE item of at least check_min points
// loads its subtree's minimum dim2; queries with min_dim2 <= it report the
// whole subtree. The count pass (check_min = kCheckMinCount) adds the
// subtree's size; the write pass checks subtrees of at least copy_min
// points, records those of at most copy_max points as copy tasks (copied in
// breadth-first order by k_copy_tasks after the last round), and walks
// larger ones on with the query flagged in the children, so that the check
// is not repeated (copy_max = 0: no copies). A task costs a list
// reservation in the walk and an output reservation in the copy, so small
// subtrees are cheaper to walk.
constexpr int32_t kCheckMinCount = 16;

// Spilled work item: the group it belongs to (kEmpty: no item) and the item.
struct Entry {
  uint64_t item;
  uint64_t group;
};

// A whole subtree that some queries of a group report, found by the write
// pass's walk and copied to their outputs by k_copy_tasks after the last
// round (qmask 0: no task).
struct CopyTask {
  uint32_t root;   // slot of the subtree's root
  uint32_t qmask;  // bit j: query j of the group
  uint64_t group;
};
// Copy tasks one write pass can hold; subtrees found beyond it are copied
// by the walk itself.
constexpr size_t kCopyTaskCap = size_t{1} << 22;
constexpr int kCopyWavesPerCu = 32;
// Loads in flight per lane in a copy pass (4-byte records): 4096 positions
// per memory round trip in the copy kernel (195 VGPRs, two waves per SIMD,
// 1% faster on a 2^26-point batch than 32 at four waves); the walk's own
// copies (task list full) use narrower passes to keep its register budget.
constexpr int kCopyPass = 64;
constexpr int kCopyPassWalk = 8;

// Work items: slot | a << 32 | l << 40 | h << 48 | f << 56. Bit j of a, l
// and h is query j's search state at the slot (node_step); bit j of f (a
// subset of a) says query j reports every node of this subtree.
struct Items {
  using type = uint64_t;
  __device__ static type make(uint32_t slot, uint32_t a, uint32_t l, uint32_t h, uint32_t f) {
    return static_cast<uint64_t>(slot) | (static_cast<uint64_t>(a | (l << 8) | (h << 16) | (f << 24)) << 32);
  }
  __device__ static uint32_t slot(type it) { return static_cast<uint32_t>(it); }
  __device__ static uint32_t a(type it) { return static_cast<uint32_t>(it >> 32) & 0xffu; }
  __device__ static uint32_t l(type it) { return static_cast<uint32_t>(it >> 40) & 0xffu; }
  __device__ static uint32_t h(type it) { return static_cast<uint32_t>(it >> 48) & 0xffu; }
  __device__ static uint32_t f(type it) { return static_cast<uint32_t>(it >> 56) & 0xffu; }
};

template <typename T, typename IS>
struct Params {
  const Node<T, IS>* nodes;
  const T* submin;  // subtree minimum of dim2 per slot
  const IS* ids;    // ID per slot (whole-subtree copies of IDs read these)
  const Query<T>* queries;
  const uint32_t* gq;  // G query indices per group (~0u: none); null when G == 1
  Entry* list[2];
  unsigned long long list_cap;
  unsigned long long* ctr;  // [2r] items in list r, [2r+1] grab counter of round r
  unsigned long long* counts;  // count mode: totals; write mode: cursors
  const unsigned long long* offsets;
  const unsigned long long* caps;
  unsigned long long cap_all;
  unsigned long long out_total;  // records the output buffer holds
  void* out;
  unsigned long long* err;  // bit 0: stack would overflow; bit 1: invalid work item
  unsigned long long* dbg;  // optional per-round counters (nullable)
  Query<T> single;          // the query when `queries` is null (one-query searches)
  uint32_t S;               // slot count (< 2^30): items outside [0, S) are rejected
  // Forest of trees: slot = tree << tree_shift | slot within the tree; tree
  // c holds full_M - 1 points, the last one last_M - 1. One deep tree:
  // tree_shift 31, ntrees 1, last_M = n + 1.
  uint32_t tree_shift;
  uint32_t ntrees, full_M, last_M;
  int32_t check_min, copy_max;  // full-subtree shortcut thresholds of this pass
  uint32_t nq;              // query count
  uint32_t ngroups;         // groups: items of other groups are rejected
  void* gstack;              // per-wave overflow stacks, gcap items each
  uint32_t gcap;
  CopyTask* tasks;           // copy tasks of the write pass (null: copies in the walk)
  unsigned long long task_cap;
  unsigned long long* task_ctr;
  uint32_t total_waves;
  uint32_t budget_small, budget_large;
};

__device__ inline uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

template <typename V>
__device__ inline V read_lane(V v, int j) {
  if constexpr (sizeof(V) == 4) {
    return __builtin_bit_cast(V, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), j));
  } else {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(u), j);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(u >> 32), j);
    return __builtin_bit_cast(V, (static_cast<uint64_t>(hi) << 32) | lo);
  }
}


template <typename T, int G>
__global__ void k_search_init(const Query<T>* __restrict__ queries, Query<T> single, const uint32_t* gq,
                              uint32_t nq, uint32_t ngroups, uint64_t ntrees, uint32_t tree_shift,
                              Entry* __restrict__ roots, unsigned long long* __restrict__ ctr, uint32_t nctr,
                              unsigned long long* __restrict__ counts, unsigned long long* err) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  // root entry k: group k / ntrees at the root of tree k % ntrees
  if (k < ngroups * ntrees) {
    const uint32_t i = static_cast<uint32_t>(k / ntrees);
    const uint64_t tree = k - static_cast<uint64_t>(i) * ntrees;
    uint64_t item;
    if constexpr (G == 1) {
      const uint32_t kind = queries ? queries[i].kind : single.kind;
      item = (1ull << 32) | (static_cast<uint64_t>(lo_open(kind)) << 40) | (static_cast<uint64_t>(hi_open(kind)) << 48);
    } else {
      uint32_t a = 0, l = 0, h = 0;
      for (int j = 0; j < G; j++) {
        const uint32_t q = gq[static_cast<uint64_t>(i) * G + j];
        if (q < nq) {
          const uint32_t kind = queries[q].kind;
          a |= 1u << j;
          l |= lo_open(kind) << j;
          h |= hi_open(kind) << j;
        }
      }
      item = (static_cast<uint64_t>(a) << 32) | (static_cast<uint64_t>(l) << 40) | (static_cast<uint64_t>(h) << 48);
    }
    item |= tree << tree_shift;
    roots[k] = Entry{item, i};
  }
  if (k < nq) counts[k] = 0;
  if (k < nctr) ctr[k] = (k == 0) ? ngroups * ntrees : 0;
  if (k == 0) {
    err[0] = 0;  // error flags
    err[1] = 0;  // copy tasks recorded
  }
}

template <typename T, typename IS, bool RET_POINT>
struct OutRec {
  using type = typename std::conditional<RET_POINT, Point<T, IS>, IS>::type;
};

// Staged results per query before a flush (a step adds at most 64).
template <int G, size_t REC>
constexpr int stage_per_query() {
  return G == 1 ? 256 : (REC <= 8 ? 128 : 64);
}

// Copies the msz records of the subtree rooted at slot `root` (tree at
// slot tb, shape parameter Mt) to the outputs of the queries in qmask (lane
// j holds query j's index, output offset and capacity): one reservation per
// query (lanes j together), then the subtree's positions in breadth-first
// order as one flat range (position x is level k = floor(log2(x + 1)), local
// slot (r + 1) 2^k - 1 + x + 1 - 2^k), U x 64 positions per pass with every
// load of the pass issued before its stores (one memory round trip per
// pass); positions past the tree's last points are skipped. UW is the
// pass width for 4-byte records (scaled down for larger ones).
template <typename T, typename IS, bool RET_POINT, int G, int UW>
__device__ __attribute__((always_inline)) inline void copy_subtree(const Params<T, IS>& P, uint32_t lane,
                                                                   uint32_t qid, unsigned long long qoff,
                                                                   unsigned long long qcap, uint32_t root,
                                                                   uint32_t tb, uint32_t Mt, uint32_t msz,
                                                                   uint32_t qmask) {
  using Rec = typename OutRec<T, IS, RET_POINT>::type;
  constexpr int U = sizeof(Rec) <= 4 ? UW : (sizeof(Rec) <= 8 ? UW / 2 : UW / 4);
  unsigned long long b = 0;
  uint32_t fit = 0;
  if (lane < G && ((qmask >> lane) & 1u)) {
    b = atomicAdd(&P.counts[qid], static_cast<unsigned long long>(msz));
    const unsigned long long room_q = b < qcap ? qcap - b : 0;
    const unsigned long long at = qoff + b;
    const unsigned long long room_o = at < P.out_total ? P.out_total - at : 0;
    const unsigned long long r = room_q < room_o ? room_q : room_o;
    fit = static_cast<uint32_t>(r < msz ? r : msz);
    b = at;
  }
  const uint32_t rl = root - tb;
  const uint32_t levels = 32u - static_cast<uint32_t>(__clz(static_cast<int>(msz)));
  const uint32_t npos = (1u << levels) - 1u;
  // a subtree with every position filled (all but those on the tree's
  // right edge) needs no per-position check
  const bool complete = msz == npos;
  auto local_slot = [&](uint32_t x) {
    const uint32_t k = 31u - static_cast<uint32_t>(__clz(static_cast<int>(x + 1)));
    return ((rl + 1) << k) + x - (1u << k);
  };
  auto load_rec = [&](uint32_t ls, Rec& r) {
    if constexpr (RET_POINT) {
      const Node<T, IS> nd2 = P.nodes[tb + ls];
      r.dim1_val = nd2.dim1;
      r.dim2_val = nd2.dim2;
      if constexpr (has_id_v<IS>) r.id = nd2.id;
    } else {
      r = P.ids[tb + ls];
    }
  };
  if (complete) {
    // position x goes to output index x; positions past the subtree load
    // the root again (discarded), so every load of a pass is issued
    for (uint32_t x0 = 0; x0 < npos; x0 += kWave * U) {
      Rec rec[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const uint32_t x = x0 + u * kWave + lane;
        load_rec(x < npos ? local_slot(x) : rl, rec[u]);
      }
      uint32_t qm = qmask;
      while (qm) {
        const int j = __builtin_ctz(qm);
        qm &= qm - 1;
        Rec* out = static_cast<Rec*>(P.out) + read_lane(b, j);
        const uint32_t lim = read_lane(fit, j);  // at most msz = npos
#pragma unroll
        for (int u = 0; u < U; u++) {
          const uint32_t x = x0 + u * kWave + lane;
          if (x < lim) out[x] = rec[u];
        }
      }
    }
    return;
  }
  // subtrees on the tree's right edge: narrower passes (rare)
  constexpr int UE = U >= 8 ? U / 4 : U;
  auto exists = [&](uint32_t x) { return x < npos && subtree_size32(Mt, local_slot(x)) >= 1; };
  uint32_t written = 0;
  for (uint32_t x0 = 0; x0 < npos; x0 += kWave * UE) {
    Rec rec[UE];
#pragma unroll
    for (int u = 0; u < UE; u++) {
      const uint32_t x = x0 + u * kWave + lane;
      if (exists(x)) load_rec(local_slot(x), rec[u]);
    }
#pragma unroll
    for (int u = 0; u < UE; u++) {
      const uint32_t x = x0 + u * kWave + lane;
      const bool ex = exists(x);
      const uint64_t em = __ballot(ex);
      const uint32_t idx = written + lanes_below(em);
#pragma unroll
      for (int j = 0; j < G; j++) {
        if (!((qmask >> j) & 1u)) continue;
        Rec* out = static_cast<Rec*>(P.out) + read_lane(b, j);
        if (ex && idx < read_lane(fit, j)) out[idx] = rec[u];
      }
      written += __popcll(em);
    }
  }
}

// Copies the recorded tasks' subtrees: one task per wave at a time. The
// walk's tasks were written by the previous launch and are read with
// device-coherent loads.
template <typename T, typename IS, bool RET_POINT, int G>
__global__ __launch_bounds__(kWave) void k_copy_tasks(Params<T, IS> P) {
  const uint32_t lane = threadIdx.x;
  const unsigned long long n =
      min(__hip_atomic_load(P.task_ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), P.task_cap);
  for (unsigned long long t = blockIdx.x; t < n; t += gridDim.x) {
    const uint64_t group = __hip_atomic_load(&P.tasks[t].group, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t root = __hip_atomic_load(&P.tasks[t].root, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t qmask = __hip_atomic_load(&P.tasks[t].qmask, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (qmask == 0) continue;
    if (group >= P.ngroups || root >= P.S || qmask >= (1u << G)) {
      if (lane == 0) atomicOr(P.err, 2ull);
      continue;
    }
    const uint32_t tree = root >> P.tree_shift;
    const uint32_t tb = tree << P.tree_shift;
    const uint32_t Mt = tree + 1 < P.ntrees ? P.full_M : P.last_M;
    const int32_t msz = subtree_size32(Mt, root - tb);
    if (msz < 1) {
      if (lane == 0) atomicOr(P.err, 2ull);
      continue;
    }
    uint32_t qid = ~0u;
    unsigned long long qoff = 0, qcap = 0;
    if (lane < G) {
      qid = (G == 1) ? static_cast<uint32_t>(group) : P.gq[group * G + lane];
      if (qid < P.nq) {
        qoff = P.offsets ? P.offsets[qid] : 0;
        qcap = P.caps ? P.caps[qid] : read_lane(P.cap_all, 0);
      }
    }
    qmask &= static_cast<uint32_t>(__ballot(lane < G && qid < P.nq));
    copy_subtree<T, IS, RET_POINT, G, kCopyPass>(P, lane, qid, qoff, qcap, root, tb, Mt, static_cast<uint32_t>(msz),
                                                 qmask);
  }
}

template <typename T, typename IS, bool RET_POINT, bool WRITE, int G>
__global__ __launch_bounds__(kWave) void k_search_round(Params<T, IS> P, uint32_t round, uint32_t last) {
  using Rec = typename OutRec<T, IS, RET_POINT>::type;
  using It = Items;
  using Item = typename It::type;
  constexpr int kStage = stage_per_query<G, sizeof(Rec)>();
  static_assert(kStage % kWave == 0, "stages are flushed in whole wave-widths");
  constexpr int kLdsStack = lds_stack<G>();
  __shared__ Item stack[kLdsStack];
  __shared__ Entry pend[kWave];
  Item* gstack = static_cast<Item*>(P.gstack) + static_cast<uint64_t>(blockIdx.x) * P.gcap;
  __shared__ Rec stage[WRITE ? G * kStage : 1];

  const uint32_t lane = threadIdx.x;
  // Counters and lists written by the previous launch are read with
  // device-coherent loads (never through the scalar or L1 caches).
  const unsigned long long count_in =
      min(__hip_atomic_load(&P.ctr[2 * round], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), P.list_cap);
  if (P.dbg && lane == 0) {
    atomicMax(&P.dbg[round * 8 + 0], count_in);
  }
  unsigned long long* grab = &P.ctr[2 * round + 1];
  unsigned long long* count_out = &P.ctr[2 * round + 2];
  const Entry* list_in = P.list[round & 1];
  Entry* list_out = P.list[(round + 1) & 1];

  unsigned long long chunk = count_in / P.total_waves;
  chunk = chunk < 1 ? 1 : (chunk > kWave ? kWave : chunk);
  const uint32_t budget =
      last ? 0xffffffffu : (count_in < P.total_waves ? P.budget_small : P.budget_large);

  // The first chunk of each wave is fixed by its index; later chunks are
  // grabbed from a shared counter, which is read first so that waves with
  // nothing left to take leave without an atomic update.
  const unsigned long long dyn0 = static_cast<unsigned long long>(P.total_waves) * chunk;
  bool fixed_done = false;
  uint32_t npend = 0;  // uniform: valid pending entries are pend[0..npend) with group != kEmpty
  while (true) {
    // ---------- pick the next batch: the pending items of one group ----------
    const uint64_t mine_g = lane < npend ? pend[lane].group : kEmpty;
    const uint64_t have = __ballot(mine_g != kEmpty);
    if (have == 0) {
      unsigned long long base = 0;
      if (!fixed_done) {
        fixed_done = true;
        base = static_cast<unsigned long long>(blockIdx.x) * chunk;
        if (base >= count_in) continue;
      } else {
        if (dyn0 >= count_in ||
            dyn0 + __hip_atomic_load(grab, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= count_in)
          break;
        if (lane == 0) base = dyn0 + atomicAdd(grab, chunk);
        base = __shfl(base, 0, kWave);
        if (base >= count_in) break;
      }
      const unsigned long long cnt = min(chunk, count_in - base);
      Entry e{0, kEmpty};
      if (lane < cnt) {
        e.group = __hip_atomic_load(&list_in[base + lane].group, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        e.item = __hip_atomic_load(&list_in[base + lane].item, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      // reject items that cannot come from this search
      if (e.group != kEmpty && (e.group >= P.ngroups || It::slot(static_cast<Item>(e.item)) >= P.S ||
                                It::a(static_cast<Item>(e.item)) == 0)) {
        atomicOr(P.err, 2ull);
        if (P.dbg) atomicAdd(&P.dbg[round * 8 + 3], 1ull);
        e.group = kEmpty;
      }
      pend[lane] = e;
      npend = static_cast<uint32_t>(cnt);
      if (P.dbg && lane == 0) atomicAdd(&P.dbg[round * 8 + 2], cnt);
      __builtin_amdgcn_wave_barrier();
      continue;
    }
    const int first = __builtin_ctzll(have);
    const uint32_t g = uniform(static_cast<uint32_t>(__shfl(mine_g, first, kWave)));
    const bool take = mine_g == g;
    const uint64_t tmask = __ballot(take);
    if (take) {
      stack[lanes_below(tmask)] = static_cast<Item>(pend[lane].item);
      pend[lane].group = kEmpty;
    }
    uint32_t sp = __popcll(tmask);
    __builtin_amdgcn_wave_barrier();

    // ---------- the group's queries: lane j holds query j ----------
    uint32_t qid = ~0u;
    Query<T> qv{};
    if (lane < G) {
      qid = (G == 1) ? g : P.gq[static_cast<uint64_t>(g) * G + lane];
      if (qid < P.nq) {
        if (P.queries) {
          qv = P.queries[qid];
        } else {
          qv.min_dim1 = read_lane(P.single.min_dim1, 0);
          qv.max_dim1 = read_lane(P.single.max_dim1, 0);
          qv.min_dim2 = read_lane(P.single.min_dim2, 0);
          qv.kind = read_lane(P.single.kind, 0);
        }
      }
    }
    Bounds<T> qb[G];
#pragma unroll
    for (int j = 0; j < G; j++) {
      qb[j].lo = read_lane(qv.min_dim1, j);
      qb[j].hi = read_lane(qv.max_dim1, j);
      qb[j].min2 = read_lane(qv.min_dim2, j);
    }
    BoundNum bn;  // queries whose dim1 bounds are numbers (lane j holds query j)
    bn.lo = static_cast<uint32_t>(__ballot(lane < G && qv.min_dim1 == qv.min_dim1));
    bn.hi = static_cast<uint32_t>(__ballot(lane < G && qv.max_dim1 == qv.max_dim1));
    uint32_t found[G];  // results of this batch (count mode; a batch reports < 2^32 per query)
    uint32_t nst[G];    // staged results (write mode)
#pragma unroll
    for (int j = 0; j < G; j++) {
      found[j] = 0;
      nst[j] = 0;
    }
    // lane j keeps query j's output offset and capacity (write mode)
    unsigned long long qoff = 0, qcap = 0;
    if constexpr (WRITE) {
      if (lane < G && qid < P.nq) {
        qoff = P.offsets ? P.offsets[qid] : 0;
        qcap = P.caps ? P.caps[qid] : read_lane(P.cap_all, 0);
      }
    }
    // Flushes every query whose stage holds at least `thr` results: lanes
    // 0..G-1 reserve their query's output range together (one atomic
    // round trip for the group), then the wave copies the stages out. The
    // stages are read from LDS while the reservation is in flight.
    auto flush = [&](uint32_t thr) {
      if constexpr (WRITE) {
        uint32_t fl = 0, mine = 0;
#pragma unroll
        for (int j = 0; j < G; j++) {
          if (nst[j] >= thr && nst[j] > 0) fl |= 1u << j;
          if (lane == static_cast<uint32_t>(j)) mine = nst[j];
        }
        if (fl == 0) return;
        constexpr int kPer = kStage / kWave;
        Rec v[G][kPer];
#pragma unroll
        for (int j = 0; j < G; j++)
#pragma unroll
          for (int k = 0; k < kPer; k++) v[j][k] = stage[j * kStage + k * kWave + lane];
        // lane j: first free index of query j's range after the reservation,
        // and how many of its staged records still fit
        unsigned long long b = 0;
        uint32_t fit = 0;
        if (lane < G && ((fl >> lane) & 1u)) {
          b = atomicAdd(&P.counts[qid], static_cast<unsigned long long>(mine));
          const unsigned long long room_q = b < qcap ? qcap - b : 0;
          const unsigned long long at = qoff + b;
          const unsigned long long room_o = at < P.out_total ? P.out_total - at : 0;
          const unsigned long long r = room_q < room_o ? room_q : room_o;
          fit = static_cast<uint32_t>(r < mine ? r : mine);
          b = at;
        }
#pragma unroll
        for (int j = 0; j < G; j++) {
          if (!((fl >> j) & 1u)) continue;
          Rec* out = static_cast<Rec*>(P.out) + read_lane(b, j);
          const uint32_t nfit = read_lane(fit, j);
#pragma unroll
          for (int k = 0; k < kPer; k++) {
            const uint32_t x = k * kWave + lane;
            if (x < nfit) out[x] = v[j][k];
          }
          nst[j] = 0;
        }
        __builtin_amdgcn_wave_barrier();
      }
    };
    // a step adds at most 64 results per query; a full stage triggers a
    // flush of every stage at least this full
    constexpr uint32_t kFlushAll = kStage > kWave ? (kStage - kWave) / 2 + 1 : 1;

    // One batch of popped items: prep() decodes the items and issues their
    // loads; finish() steps the nodes, reports, and pushes the children.
    struct Half {
      Node<T, IS> nd;
      T sm;  // subtree minimum of dim2 (read for candidates only)
      uint32_t slot, tbase, tM, ia, il, ih, ifull, cand;
      int32_t m;
      bool ok, bad;
    };
    // tree shape in scalar registers (a per-lane choice between two kernel
    // arguments would otherwise become a load from the argument block)
    const uint32_t tshift = uniform(P.tree_shift), ntrees = uniform(P.ntrees);
    const uint32_t full_M = uniform(P.full_M), last_M = uniform(P.last_M);
    auto prep = [&](const Item item, const bool act, Half& h) {
      h.slot = It::slot(item);
      h.m = 0;
      h.ok = act;
      h.bad = false;
      h.tbase = 0;
      if (act) {
        const uint32_t tree = h.slot >> tshift;
        h.tbase = tree << tshift;
        h.m = subtree_size32(tree + 1 < ntrees ? full_M : last_M, h.slot - h.tbase);
        if (h.slot >= P.S || h.m < 1) {
          h.bad = true;
          h.ok = false;
        }
      }
      h.ia = It::a(item);
      h.il = It::l(item);
      h.ih = It::h(item);
      h.ifull = It::f(item);
      h.tM = (h.slot >> tshift) + 1 < ntrees ? full_M : last_M;
      // REPORT_ABOVE queries not yet known to report the whole subtree
      h.cand = (h.ok && h.m >= P.check_min) ? (h.ia & ~h.il & ~h.ih & ~h.ifull) : 0u;
      h.sm = P.submin[h.cand ? h.slot : 0u];
      h.nd = P.nodes[h.ok ? h.slot : 0u];
    };
    bool saw_bad = false;  // an invalid item was popped (reported after the walk)
    auto finish = [&](Half& h) {
      saw_bad |= h.bad;
      const uint32_t slot = h.slot, tbase = h.tbase, tM = h.tM, il = h.il, ih = h.ih;
      const int32_t m = h.m;
      uint32_t ia = h.ia, ifull = h.ifull;
      bool ok = h.ok;
      const Node<T, IS>& nd = h.nd;
      StepOut so{};
      if (h.cand) {
#pragma unroll
        for (int j = 0; j < G; j++)
          if (((h.cand >> j) & 1u) && qb[j].min2 <= h.sm) ifull |= 1u << j;
      }
      // whole subtrees: counted at once, or copied when small enough
      const uint32_t bulk = (!WRITE || m <= P.copy_max) ? ifull : 0u;
      if constexpr (!WRITE) {
        if (__ballot(bulk != 0)) {
#pragma unroll
          for (int j = 0; j < G; j++) {
            uint32_t v = ((bulk >> j) & 1u) ? static_cast<uint32_t>(m) : 0u;
            for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
            found[j] += uniform(v);
          }
        }
      } else {
        uint64_t bm = __ballot(bulk != 0);
        if (bm) {
          // record the subtrees as copy tasks (one list reservation for the
          // step); a reservation that does not fit marks its in-capacity
          // part empty and the wave copies the subtrees itself
          const uint32_t ntask = __popcll(bm);
          bool listed = false;
          if (P.tasks) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(P.task_ctr, static_cast<unsigned long long>(ntask));
            base = __shfl(base, 0, kWave);
            listed = base < P.task_cap && ntask <= P.task_cap - base;
            if (listed) {
              if (bulk) P.tasks[base + lanes_below(bm)] = CopyTask{slot, bulk, g};
            } else if (lane < ntask && base + lane < P.task_cap) {
              P.tasks[base + lane] = CopyTask{0, 0, kEmpty};
            }
          }
          while (!listed && bm) {
            const int src = __builtin_ctzll(bm);
            bm &= bm - 1;
            copy_subtree<T, IS, RET_POINT, G, kCopyPassWalk>(
                P, lane, qid, qoff, qcap, static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(slot), src)),
                static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(tbase), src)),
                static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(tM), src)),
                static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(m), src)),
                static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(bulk), src)));
          }
        }
      }
      ia &= ~bulk;
      ifull &= ~bulk;
      if (ia == 0) ok = false;
      // deep in reported subtrees every item is REPORT_ABOVE for all its
      // queries: then only the dim2 prune is evaluated
      const bool open = ok && (il | ih) != 0;
      if (__ballot(open) == 0) {
        if (ok) so = node_step_report_above<T, G>(nd.dim2, m, qb, ia);
      } else if (ok) {
        so = node_step<T, G>(nd.dim1, nd.dim2, nd.med, m, qb, bn, ia, il, ih);
      }
      // ---- results, per query of the group
#pragma unroll
      for (int j = 0; j < G; j++) {
        const bool r = (so.rep >> j) & 1u;
        const uint64_t rmask = __ballot(r);
        if (rmask == 0) continue;
        if constexpr (WRITE) {
          if (r) {
            Rec rec;
            if constexpr (RET_POINT) {
              rec.dim1_val = nd.dim1;
              rec.dim2_val = nd.dim2;
              if constexpr (has_id_v<IS>) rec.id = nd.id;
            } else {
              rec = nd.id;
            }
            stage[j * kStage + nst[j] + lanes_below(rmask)] = rec;
          }
          nst[j] += __popcll(rmask);
        } else {
          found[j] += __popcll(rmask);
        }
      }
      if constexpr (WRITE) {
        __builtin_amdgcn_wave_barrier();
        bool full = false;
#pragma unroll
        for (int j = 0; j < G; j++) full |= nst[j] > static_cast<uint32_t>(kStage - kWave);
        if (full) flush(kFlushAll);
      }
      // ---- children, pushed as (left, right) pairs in lane order: the next
      // pop then takes whole sibling pairs of neighbouring parents, which sit
      // side by side in the level below (a few cache lines per pop)
      const bool cl = so.al != 0, cr = so.ar != 0;
      const uint64_t lm = __ballot(cl), rm = __ballot(cr);
      const uint32_t pos = sp + lanes_below(lm) + lanes_below(rm);
      const uint32_t lchild = tbase + 2 * (slot - tbase) + 1;
      if (cl) stack[pos] = It::make(lchild, so.al, so.ll, so.hl, ifull & so.al);
      if (cr) stack[pos + (cl ? 1u : 0u)] = It::make(lchild + 1, so.ar, so.lr, so.hr, ifull & so.ar);
      sp += __popcll(lm) + __popcll(rm);
      __builtin_amdgcn_wave_barrier();
    };
    uint32_t steps = 0;
    unsigned long long visits = 0;  // items popped (debug counters only)
    uint32_t gsp = 0;  // items in the wave's global overflow stack
    while (sp > 0 || gsp > 0) {
      if (sp == 0) {
        // refill from the overflow stack (blocks of kMoveBlock, each lane
        // reads back the entries it wrote)
        gsp -= kMoveBlock;
#pragma unroll
        for (int k = 0; k < kMoveBlock / kWave; k++) stack[lane + k * kWave] = gstack[gsp + lane + k * kWave];
        sp = kMoveBlock;
        __builtin_amdgcn_wave_barrier();
      }
      if (sp > kLdsStack - 2 * kWave) {
        // a step (two batches) grows the stack by at most 128: move the
        // oldest kMoveBlock items to the overflow stack and shift the rest
        // down
        if (gsp + kMoveBlock > P.gcap) {
          if (lane == 0) atomicOr(P.err, 1ull);
          sp = 0;
          gsp = 0;
          break;
        }
#pragma unroll
        for (int k = 0; k < kMoveBlock / kWave; k++) gstack[gsp + lane + k * kWave] = stack[lane + k * kWave];
        gsp += kMoveBlock;
        // shift in passes of kMoveBlock: each pass reads a block above the
        // moved one before writing it kMoveBlock lower
        for (uint32_t b0 = kMoveBlock; b0 < sp; b0 += kMoveBlock) {
          Item keep[kMoveBlock / kWave];
#pragma unroll
          for (int k = 0; k < kMoveBlock / kWave; k++) {
            const uint32_t x = b0 + lane + k * kWave;
            keep[k] = x < sp ? stack[x] : Item(0);
          }
          __builtin_amdgcn_wave_barrier();
#pragma unroll
          for (int k = 0; k < kMoveBlock / kWave; k++) {
            const uint32_t x = b0 + lane + k * kWave;
            if (x < sp) stack[x - kMoveBlock] = keep[k];
          }
          __builtin_amdgcn_wave_barrier();
        }
        sp -= kMoveBlock;
        __builtin_amdgcn_wave_barrier();
      }
      // ---- pop two batches of up to 64 items, A (the newest) and B below
      // it, and load both batches' nodes before A is processed, so that B's
      // loads are in flight while A is stepped. Every lane issues both loads
      // of each batch (lanes without a valid item read slot 0, discarded),
      // which keeps the number of loads after A's fixed.
      const uint32_t na = sp < kWave ? sp : kWave;
      sp -= na;
      const uint32_t nb = sp < kWave ? sp : kWave;
      sp -= nb;
      visits += na + nb;
      const Item item_a = lane < na ? stack[sp + nb + lane] : Item(0);
      const Item item_b = lane < nb ? stack[sp + lane] : Item(0);
      __builtin_amdgcn_wave_barrier();
      Half A, B;
      prep(item_a, lane < na, A);
      prep(item_b, lane < nb, B);
      finish(A);
      if (nb) finish(B);
      // ---- hand remaining work to the next round when over budget
      steps += nb ? 2 : 1;  // batches stepped
      if (steps >= budget && sp + gsp > 0) {
        // Reservations only grow the counter; one that ends past the list
        // capacity fails, marks its in-capacity part empty (the next round
        // skips empty entries and reads at most list_cap of them) and the
        // wave keeps walking its batch instead.
        unsigned long long base = 0;
        const uint32_t tot = sp + gsp;
        if (lane == 0) base = atomicAdd(count_out, static_cast<unsigned long long>(tot));
        base = __shfl(base, 0, kWave);
        if (base + tot <= P.list_cap) {
          for (uint32_t x = lane; x < gsp; x += kWave) list_out[base + x] = Entry{gstack[x], g};
          for (uint32_t x = lane; x < sp; x += kWave) list_out[base + gsp + x] = Entry{stack[x], g};
          if (P.dbg && lane == 0) {
            atomicAdd(&P.dbg[round * 8 + 4], 1ull);
            atomicAdd(&P.dbg[round * 8 + 5], static_cast<unsigned long long>(tot));
          }
          sp = 0;
          gsp = 0;
        } else {
          for (unsigned long long x = base + lane; x < P.list_cap && x < base + tot; x += kWave)
            list_out[x].group = kEmpty;
          steps = 0;
        }
      }
    }
    if (saw_bad) atomicOr(P.err, 2ull);
    if (P.dbg && lane == 0) {
      atomicAdd(&P.dbg[round * 8 + 6], static_cast<unsigned long long>(steps));
      atomicMax(&P.dbg[round * 8 + 7], static_cast<unsigned long long>(steps));
      atomicAdd(&P.dbg[round * 8 + 1], visits);
    }
    if constexpr (WRITE) {
      flush(1);
    } else {
      uint32_t mine = 0;
#pragma unroll
      for (int j = 0; j < G; j++)
        if (lane == static_cast<uint32_t>(j)) mine = found[j];
      if (lane < G && mine) atomicAdd(&P.counts[qid], static_cast<unsigned long long>(mine));
    }
  }
}

// ---------------------------------------------------------------------------
// Grouping: queries are ordered by kind, then by their leading dim1 bound
// (max_dim1 for left searches, min_dim1 otherwise), with two stable radix
// sorts; group k holds sorted positions [kG, kG + G).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_group_keys(const Query<T>* __restrict__ q, uint32_t nq, uint64_t* __restrict__ key,
                             uint32_t* __restrict__ idx) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const Query<T> r = q[i];
  key[i] = static_cast<uint64_t>(ord_key(r.kind == KIND_LEFT ? r.max_dim1 : r.min_dim1));
  idx[i] = i;
}

template <typename T>
__global__ void k_kind_keys(const Query<T>* __restrict__ q, const uint32_t* __restrict__ perm, uint32_t nq,
                            uint32_t* __restrict__ kk) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const uint32_t k = q[perm[i]].kind;
  kk[i] = k > KIND_RIGHT ? KIND_RIGHT : k;
}

// Batches of up to kSmallSort queries are ordered by one workgroup with a
// bitonic sort in LDS, on (kind, leading-bound key, index): one kernel and
// no library calls, so the whole batched search can be captured in a HIP
// graph. Larger batches use the two radix sorts below.
constexpr uint32_t kSmallSort = 8192;
constexpr int kSortBlock = 1024;

template <typename T>
__global__ __launch_bounds__(kSortBlock) void k_group_sort_small(const Query<T>* __restrict__ q, uint32_t nq,
                                                                 uint32_t N, uint32_t total,
                                                                 uint32_t* __restrict__ gq) {
  extern __shared__ __align__(16) unsigned char lds[];
  uint64_t* key = reinterpret_cast<uint64_t*>(lds);
  uint32_t* kind = reinterpret_cast<uint32_t*>(key + N);
  uint16_t* perm = reinterpret_cast<uint16_t*>(kind + N);
  for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) {
    perm[i] = static_cast<uint16_t>(i);
    if (i < nq) {
      const Query<T> r = q[i];
      key[i] = static_cast<uint64_t>(ord_key(r.kind == KIND_LEFT ? r.max_dim1 : r.min_dim1));
      kind[i] = r.kind > KIND_RIGHT ? KIND_RIGHT : r.kind;
    }
  }
  __syncthreads();
  auto before = [&](uint32_t a, uint32_t b) {
    if (a >= nq) return false;
    if (b >= nq) return true;
    if (kind[a] != kind[b]) return kind[a] < kind[b];
    if (key[a] != key[b]) return key[a] < key[b];
    return a < b;
  };
  for (uint32_t k = 2; k <= N; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t t = threadIdx.x; t < N / 2; t += blockDim.x) {
        const uint32_t i = (t / j) * 2 * j + (t % j);
        const uint32_t l = i + j;
        const uint16_t a = perm[i], b = perm[l];
        const bool up = (i & k) == 0;
        if (up ? before(b, a) : before(a, b)) {
          perm[i] = b;
          perm[l] = a;
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) gq[i] = i < nq ? perm[i] : ~0u;
}

__global__ void k_group_pad(uint32_t* __restrict__ gq, uint32_t nq, uint32_t total) {
  const uint32_t i = nq + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) gq[i] = ~0u;
}

// Exclusive prefix sum of per-query counts: offsets[q] = sum of counts
// before q, offsets[nq] = total. One workgroup of 1024 threads walks the
// array in chunks of 4096 (batches hold fewer than 2^24 queries).
constexpr int kScanBlock = 1024;
__global__ __launch_bounds__(kScanBlock) void k_exclusive_scan(const unsigned long long* __restrict__ in,
                                                               unsigned long long* __restrict__ out,
                                                               uint32_t n) {
  __shared__ unsigned long long wsum[kScanBlock / kWave];
  __shared__ unsigned long long carry;
  const uint32_t tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < n; base += 4 * kScanBlock) {
    unsigned long long v[4], loc = 0;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const uint32_t i = base + tid * 4 + e;
      v[e] = i < n ? in[i] : 0;
      loc += v[e];
    }
    unsigned long long incl = loc;
    for (int off = 1; off < kWave; off <<= 1) {
      const unsigned long long o = __shfl_up(incl, off, kWave);
      if (static_cast<int>(lane) >= off) incl += o;
    }
    if (lane == kWave - 1) wsum[wave] = incl;
    __syncthreads();
    unsigned long long wbase = 0, total = 0;
    for (uint32_t w = 0; w < kScanBlock / kWave; w++) {
      wbase += w < wave ? wsum[w] : 0;
      total += wsum[w];
    }
    unsigned long long run = carry + wbase + incl - loc;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const uint32_t i = base + tid * 4 + e;
      if (i < n) out[i] = run;
      run += v[e];
    }
    __syncthreads();
    if (tid == 0) carry += total;
    __syncthreads();
  }
  if (tid == 0) out[n] = carry;
}

inline uint32_t group_size(const SearchArgs& a) {
  return (a.single_query || a.nq == 1 || a.group == 1) ? 1u : static_cast<uint32_t>(kGroup);
}

struct TreeShape {
  uint64_t S, ntrees, full_M, last_M;
  uint32_t shift, levels;
};
TreeShape tree_shape(const SearchArgs& a) {
  TreeShape t{};
  if (a.slots_per_tree == 0) {
    t.S = (a.n == 0) ? 0 : ((1ull << (64 - __builtin_clzll(a.n))) - 1);
    t.ntrees = 1;
    t.full_M = t.last_M = a.n + 1;
    t.shift = 31;
  } else {
    const uint64_t St = a.slots_per_tree;
    t.shift = 0;
    while ((1ull << t.shift) < St + 1) t.shift++;
    t.ntrees = (a.n + St - 1) / St;
    t.full_M = St + 1;
    t.last_M = a.n - (t.ntrees - 1) * St + 1;
    t.S = t.ntrees << t.shift;
  }
  t.levels = 0;
  const uint64_t biggest = t.ntrees > 1 ? t.full_M - 1 : t.last_M - 1;
  while ((1ull << t.levels) <= biggest) t.levels++;
  return t;
}

template <typename T, typename IS, bool RP, bool W, int G>
Status launch_rounds(const SearchArgs& a, SearchScratch& sc, hipStream_t s) {
  Params<T, IS> P{};
  P.nodes = static_cast<const Node<T, IS>*>(a.nodes);
  P.submin = static_cast<const T*>(a.submin);
  P.ids = static_cast<const IS*>(a.ids);
  if (W && !RP && a.copy_max > 0 && !a.ids) return Status{1, "Error: the tree has no ID array"};
  if (!W) {
    P.check_min = kCheckMinCount;
    P.copy_max = 0;
  } else if (a.copy_max > 0) {
    P.check_min = a.copy_min > 1 ? a.copy_min : 1;
    P.copy_max = a.copy_max;
  } else {
    P.check_min = 0x7fffffff;
    P.copy_max = 0;
  }
  P.queries = static_cast<const Query<T>*>(a.queries);
  if (a.single_query) P.single = *static_cast<const Query<T>*>(a.single_query);
  P.gq = G == 1 ? nullptr : sc.groups;
  P.list[0] = static_cast<Entry*>(sc.list[0]);
  P.list[1] = static_cast<Entry*>(sc.list[1]);
  P.list_cap = sc.list_cap;
  P.ctr = sc.ctr;
  P.counts = a.counts;
  P.offsets = a.offsets;
  P.caps = a.caps;
  P.cap_all = a.cap_all;
  P.out_total = a.out_total;
  P.out = a.out;
  P.err = sc.ctr + 2 * kMaxRounds + 2;
  P.dbg = a.debug;
  const TreeShape sh = tree_shape(a);
  P.S = static_cast<uint32_t>(sh.S);
  P.tree_shift = sh.shift;
  P.ntrees = static_cast<uint32_t>(sh.ntrees);
  P.full_M = static_cast<uint32_t>(sh.full_M);
  P.last_M = static_cast<uint32_t>(sh.last_M);
  P.nq = static_cast<uint32_t>(a.nq);
  P.ngroups = static_cast<uint32_t>((a.nq + G - 1) / G);
  const uint32_t H = sh.levels;
  // deepest stack: 128 items (a step's two batches) per level below the
  // first popped batch, plus a step's growth, in whole move blocks
  const uint32_t bound = 2 * kWave * (H + 2) + 4 * kWave;
  P.gcap = ((bound + kMoveBlock - 1) / kMoveBlock) * kMoveBlock;
  P.budget_small = a.budget_small > 0 ? a.budget_small : 6;
  // a write step costs more than a count step (staging, flushes, copy
  // tasks), so the write pass hands work on after fewer steps
  P.budget_large = a.budget_large > 0 ? a.budget_large : (W ? 192 : 384);
  auto kern = k_search_round<T, IS, RP, W, G>;
  int dev = 0;
  hipGetDevice(&dev);
  int cus = 0;
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  // Waves per CU (one wave per workgroup), measured on the 2^26-point batch:
  // both walk kinds gain up to about 28-32 (launching more workgroups than
  // fit at once also overlaps the round's tail with late starters).
  int occ = G == 1 ? 28 : 32;
  if (a.waves_per_cu > 0) occ = std::min(a.waves_per_cu, 64);
  hipError_t e = hipSuccess;
  const uint32_t grid = static_cast<uint32_t>(cus * occ);
  P.total_waves = grid;
  const size_t gbytes = static_cast<size_t>(grid) * P.gcap * sizeof(Items::type);
  if (sc.gstack_bytes < gbytes) {
    if (sc.pinned) return Status{1, kPinnedScratchMsg};
    if (sc.gstack) hipFree(sc.gstack);
    sc.gstack = nullptr;
    sc.gstack_bytes = 0;
    e = hipMalloc(&sc.gstack, gbytes);
    if (e != hipSuccess) return hip_status(e, "Error in allocating search overflow stacks");
    sc.gstack_bytes = gbytes;
  }
  P.gstack = sc.gstack;
  if (W && P.copy_max > 0) {
    const size_t want = a.copy_tasks > 0 ? std::min(a.copy_tasks, kCopyTaskCap) : kCopyTaskCap;
    if (sc.task_cap < want) {
      if (sc.pinned) return Status{1, kPinnedScratchMsg};
      if (sc.tasks) hipFree(sc.tasks);
      sc.tasks = nullptr;
      sc.task_cap = 0;
      e = hipMalloc(&sc.tasks, want * sizeof(CopyTask));
      if (e != hipSuccess) return hip_status(e, "Error in allocating the copy task list");
      sc.task_cap = want;
    }
    P.tasks = static_cast<CopyTask*>(sc.tasks);
    P.task_cap = want;
    P.task_ctr = sc.ctr + 2 * kMaxRounds + 3;
  }
  const int R = std::max(1, std::min(a.rounds, kMaxRounds));
  for (int r = 0; r < R; r++) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kWave), 0, s, P, static_cast<uint32_t>(r),
                       static_cast<uint32_t>(r == R - 1));
  }
  if constexpr (W) {
    if (P.tasks) {
      // the copy kernel is a streaming loop with no LDS: fill every CU
      hipLaunchKernelGGL((k_copy_tasks<T, IS, RP, G>), dim3(static_cast<uint32_t>(cus * kCopyWavesPerCu)),
                         dim3(kWave), 0, s, P);
    }
  }
  return hip_status(hipGetLastError(), "Error in launching search kernel");
}

template <typename F>
Status with_t(int dt, F&& f) {
  switch (dt) {
    case F32: return f(float{});
    case F64: return f(double{});
    case I32: return f(int32_t{});
    case I64: return f(int64_t{});
  }
  return Status{1, "Error: unsupported data type code " + std::to_string(dt)};
}
template <typename F>
Status with_is(int idt, F&& f) {
  switch (idt) {
    case NO_ID: return f(NoId{});
    case F32: case I32: case U32: return f(uint32_t{});
    case F64: case I64: case U64: return f(uint64_t{});
  }
  return Status{1, "Error: unsupported ID type code " + std::to_string(idt)};
}

template <typename T>
Status make_groups(const SearchArgs& a, SearchScratch& sc, hipStream_t s) {
  const uint32_t nq = static_cast<uint32_t>(a.nq);
  const uint32_t total = ((nq + kGroup - 1) / kGroup) * kGroup;
  if (nq <= kSmallSort) {
    uint32_t N = 2;
    while (N < nq) N <<= 1;
    const size_t lds = static_cast<size_t>(N) * (8 + 4 + 2);
    auto kern = k_group_sort_small<T>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return hip_status(e, "Error in sizing the query sort");
    hipLaunchKernelGGL(kern, dim3(1), dim3(kSortBlock), lds, s, static_cast<const Query<T>*>(a.queries), nq, N,
                       total, sc.groups);
    return hip_status(hipGetLastError(), "Error in launching query grouping");
  }
  const unsigned blocks = (nq + 255) / 256;
  auto* keys = static_cast<uint64_t*>(sc.gkeys);
  auto* idx = reinterpret_cast<uint32_t*>(keys + 2 * sc.gcap);
  uint32_t* perm1 = idx + sc.gcap;
  uint32_t* kk = perm1 + sc.gcap;
  k_group_keys<T><<<blocks, 256, 0, s>>>(static_cast<const Query<T>*>(a.queries), nq, keys, idx);
  size_t tb = sc.gtmp_bytes;
  hipError_t e = radix_sort_u64(sc.gtmp, tb, keys, keys + sc.gcap, idx, perm1, nq, s);
  if (e != hipSuccess) return hip_status(e, "Error in sorting a query batch");
  k_kind_keys<T><<<blocks, 256, 0, s>>>(static_cast<const Query<T>*>(a.queries), perm1, nq, kk);
  tb = sc.gtmp_bytes;
  e = radix_sort_u32(sc.gtmp, tb, kk, kk + sc.gcap, perm1, sc.groups, nq, s);
  if (e != hipSuccess) return hip_status(e, "Error in sorting a query batch");
  if (total > nq) k_group_pad<<<1, 256, 0, s>>>(sc.groups, nq, total);
  return hip_status(hipGetLastError(), "Error in launching query grouping");
}

}  // namespace

size_t search_scratch_bytes(size_t n, int dtype, int cus) {
  SearchArgs a{};
  a.n = n;
  const uint32_t H = tree_shape(a).levels;
  const uint32_t bound = 2 * kWave * (H + 2) + 4 * kWave;
  const size_t gcap = ((bound + kMoveBlock - 1) / kMoveBlock) * kMoveBlock;
  const size_t waves = static_cast<size_t>(cus > 0 ? cus : 1) * 32;  // the default grid of group walks
  const size_t lists = 2 * (size_t{1} << 22) * sizeof(Entry);
  const size_t tasks = kCopyTaskCap * sizeof(CopyTask);
  const size_t ctr = (2 * kMaxRounds + 4) * 8 + 8;
  (void)dtype;
  return lists + tasks + waves * gcap * sizeof(Items::type) + ctr;
}

const char* const kPinnedScratchMsg =
    "Error: a batch graph of this tree uses its search buffers, which this call would reallocate; "
    "destroy the graph first";

Status scratch_reserve(SearchScratch& sc, size_t nq, size_t roots, hipStream_t) {
  const size_t spill_cap = 1ull << 22;  // 4M items per round list
  const size_t need = std::max(std::max(nq, roots), spill_cap);
  const bool grow = sc.list_cap < need || !sc.ctr || sc.cursor_cap < std::max<size_t>(nq, 1) ||
                    (nq > 1 && sc.gcap < nq);
  if (grow && sc.pinned) return Status{1, kPinnedScratchMsg};
  if (sc.list_cap < need) {
    if (sc.list[0]) hipFree(sc.list[0]);
    if (sc.list[1]) hipFree(sc.list[1]);
    sc.list[0] = sc.list[1] = nullptr;
    sc.list_cap = 0;
    hipError_t e = hipMalloc(&sc.list[0], need * sizeof(Entry));
    if (e == hipSuccess) e = hipMalloc(&sc.list[1], need * sizeof(Entry));
    if (e != hipSuccess) return hip_status(e, "Error in allocating search work lists");
    sc.list_cap = need;
  }
  if (!sc.ctr) {
    hipError_t e = hipMalloc(&sc.ctr, (2 * kMaxRounds + 4) * sizeof(unsigned long long));
    if (e != hipSuccess) return hip_status(e, "Error in allocating search counters");
  }
  if (sc.cursor_cap < std::max<size_t>(nq, 1)) {
    if (sc.cursor) hipFree(sc.cursor);
    hipError_t e = hipMalloc(&sc.cursor, std::max<size_t>(nq, 1) * 8);
    if (e != hipSuccess) return hip_status(e, "Error in allocating result cursors");
    sc.cursor_cap = std::max<size_t>(nq, 1);
  }
  if (nq > 1 && sc.gcap < nq) {
    // group table (padded to whole groups) and the sort's keys and values
    const size_t cap = ((nq + kGroup - 1) / kGroup) * kGroup;
    if (sc.groups) hipFree(sc.groups);
    if (sc.gkeys) hipFree(sc.gkeys);
    if (sc.gtmp) hipFree(sc.gtmp);
    sc.groups = nullptr;
    sc.gkeys = sc.gtmp = nullptr;
    sc.gcap = 0;
    size_t t64 = 0, t32 = 0;
    hipError_t e = radix_sort_u64(nullptr, t64, nullptr, nullptr, nullptr, nullptr, cap, nullptr);
    if (e == hipSuccess) e = radix_sort_u32(nullptr, t32, nullptr, nullptr, nullptr, nullptr, cap, nullptr);
    if (e == hipSuccess) e = hipMalloc(&sc.groups, cap * 4);
    if (e == hipSuccess) e = hipMalloc(&sc.gkeys, cap * (2 * 8 + 4 * 4));
    sc.gtmp_bytes = std::max<size_t>(std::max(t64, t32), 16);
    if (e == hipSuccess) e = hipMalloc(&sc.gtmp, sc.gtmp_bytes);
    if (e != hipSuccess) return hip_status(e, "Error in allocating query grouping storage");
    sc.gcap = cap;
  }
  return Status{};
}

void scratch_release(SearchScratch& sc) {
  if (sc.list[0]) hipFree(sc.list[0]);
  if (sc.list[1]) hipFree(sc.list[1]);
  if (sc.ctr) hipFree(sc.ctr);
  if (sc.cursor) hipFree(sc.cursor);
  if (sc.groups) hipFree(sc.groups);
  if (sc.gkeys) hipFree(sc.gkeys);
  if (sc.gtmp) hipFree(sc.gtmp);
  if (sc.gstack) hipFree(sc.gstack);
  if (sc.tasks) hipFree(sc.tasks);
  if (sc.fseg) hipFree(sc.fseg);
  sc = SearchScratch{};
}

template <typename T, int G>
Status init_typed(const SearchArgs& a, SearchScratch& sc, hipStream_t s) {
  Query<T> one{};
  if (a.single_query) one = *static_cast<const Query<T>*>(a.single_query);
  const uint32_t nctr = 2 * kMaxRounds + 2;
  const uint32_t ngroups = static_cast<uint32_t>((a.nq + G - 1) / G);
  const TreeShape sh = tree_shape(a);
  const uint64_t work = std::max<uint64_t>(std::max<uint64_t>(a.nq, nctr), ngroups * sh.ntrees);
  k_search_init<T, G><<<static_cast<unsigned>((work + 255) / 256), 256, 0, s>>>(
      a.single_query ? nullptr : static_cast<const Query<T>*>(a.queries), one, G == 1 ? nullptr : sc.groups,
      static_cast<uint32_t>(a.nq), ngroups, sh.ntrees, sh.shift, static_cast<Entry*>(sc.list[0]), sc.ctr, nctr,
      a.counts, sc.ctr + 2 * kMaxRounds + 2);
  return hip_status(hipGetLastError(), "Error in launching search initialisation");
}

// Enqueues the grouping, initialisation and search rounds on `s`
// (asynchronous). A single query is passed by value in a.single_query (host
// Query<T>) and reaches the device as a kernel argument.
Status run_search(const SearchArgs& a_in, SearchScratch& sc, hipStream_t s) {
  SearchArgs a = a_in;
  if (a.nq == 0) return Status{};
  if (a.single_query) a.queries = nullptr;
  const uint32_t G = group_size(a);
  const size_t roots = ((a.nq + G - 1) / G) * tree_shape(a).ntrees;
  Status st = scratch_reserve(sc, a.nq, roots, s);
  if (!st.ok()) return st;
  return with_t(a.dtype, [&](auto t) -> Status {
    using T = decltype(t);
    return with_is(a.id_dtype, [&](auto i) -> Status {
      using IS = decltype(i);
      if (!a.ret_point && a.mode == 1) {
        if constexpr (!has_id_v<IS>) return Status{1, "Error: IDs requested from a tree built without IDs"};
      }
      auto go = [&](auto gtag) -> Status {
        constexpr int GG = decltype(gtag)::value;
        if constexpr (GG > 1) {
          if (!a.groups_ready) {
            Status sg = make_groups<T>(a, sc, s);
            if (!sg.ok()) return sg;
          }
        }
        Status s0 = init_typed<T, GG>(a, sc, s);
        if (!s0.ok()) return s0;
        if (a.mode == 0) return launch_rounds<T, IS, false, false, GG>(a, sc, s);
        if (a.ret_point) return launch_rounds<T, IS, true, true, GG>(a, sc, s);
        if constexpr (has_id_v<IS>) return launch_rounds<T, IS, false, true, GG>(a, sc, s);
        return Status{1, "Error: IDs requested from a tree built without IDs"};
      };
      if (G == 1) return go(std::integral_constant<int, 1>{});
      return go(std::integral_constant<int, kGroup>{});
    });
  });
}

unsigned long long* search_error_flag(SearchScratch& sc) { return sc.ctr + 2 * kMaxRounds + 2; }

Status exclusive_scan(const unsigned long long* in, unsigned long long* out, size_t n, hipStream_t s) {
  k_exclusive_scan<<<1, kScanBlock, 0, s>>>(in, out, static_cast<uint32_t>(n));
  return hip_status(hipGetLastError(), "Error in launching offset scan");
}

}  // namespace pst
