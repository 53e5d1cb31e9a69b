// Batched searches of a forest of shallow trees (StaticPSTGPUArr) with the
// trees held stationary: each wavefront owns a segment of trees and
// evaluates the batch's queries against one tree at a time, so the forest
// is read from HBM once per pass instead of once per query walk.
//
// The reference searches a forest with one thread block per tree and query
// (pst/static-pst-gpu-arr-search-functions.tu:10-196). Here, per tree of at
// most 2^L - 1 <= 511 points:
//  - lanes first take a query each (64 per step, up to 512 per batch) to
//    find the candidates, the queries that may hit the tree: those whose
//    dim2 bound does not exceed the root's dim2 (the tree's largest, by
//    heap order) and whose dim1 bounds meet the tree's dim1 range (kept per
//    tree from construction);
//  - lanes then take the tree's slots (slot 64 r + l in lane l, coalesced
//    loads), and each candidate, its bounds read from its lane's registers,
//    is compared with all of them at once. The hits are exactly what a PST
//    walk of that tree reports (a walk reports every node whose dim1 meets
//    the query's bounds and whose dim2 reaches min_dim2: heap order keeps
//    such a node's ancestors unpruned, and the median splits steer the walk
//    to it).
// Pass 1 adds each candidate's hit count to its (query, segment) counter; a
// scan per query turns the counters into segment offsets and the query's
// total; pass 2 writes each candidate's hits in the tree at its offset +
// segment offset + running count, the hit lanes storing their records at
// consecutive positions. Output positions are fixed (no atomics); within a
// query and segment, results come in (tree, slot) order. Segment s holds
// trees s, s + nseg, s + 2 nseg, ... so that segments get alike work even
// when hits cluster in one region of the input. Predicates follow the
// search kinds of the C ABI: THREE min_dim1 <= dim1 <= max_dim1, LEFT
// dim1 <= max_dim1, any other kind dim1 >= min_dim1 (a RIGHT search), each
// with dim2 >= min_dim2, compared in T.
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

// Whether the query can hit a tree whose dim1 values lie in [lo1, hi1] and
// whose largest dim2 is max2 (a necessary condition).
template <typename T>
__device__ inline bool may_hit(const Box<T>& q, T lo1, T hi1, T max2) {
  return (q.min2 <= max2) & (q.lo <= hi1) & (lo1 <= q.hi);
}

// A tree's dim1 range over its points (NaN values ignored).
template <typename T>
struct Range1 {
  T lo, hi;
};

struct ForestArgs {
  uint32_t ntrees;      // trees of the forest
  uint32_t shift;       // record slots per tree = 2^shift
  uint32_t full_M;      // points + 1 of a full tree
  uint32_t last_M;      // points + 1 of the last tree
  uint32_t nseg;        // segments (one wave each); segment s holds trees s + k nseg
  uint32_t nq;
};

// Lane masks of a <= b (ordered: false where either is NaN), one vector
// compare each.
__device__ inline uint64_t le_mask(float a, float b) { return __builtin_amdgcn_fcmpf(a, b, 5); }
__device__ inline uint64_t le_mask(double a, double b) { return __builtin_amdgcn_fcmp(a, b, 5); }
__device__ inline uint64_t le_mask(int32_t a, int32_t b) { return __builtin_amdgcn_sicmp(a, b, 41); }
__device__ inline uint64_t le_mask(int64_t a, int64_t b) {
  return __builtin_amdgcn_sicmpl(static_cast<long long>(a), static_cast<long long>(b), 41);
}

// The box of a query slot past the batch's end (such slots are never
// candidates; the box keeps their LDS entries defined).
template <typename T>
__device__ inline Box<T> dead_box() {
  Box<T> b;
  b.lo = std::numeric_limits<T>::max();
  b.hi = std::numeric_limits<T>::lowest();
  b.min2 = std::numeric_limits<T>::has_quiet_NaN ? std::numeric_limits<T>::quiet_NaN() : std::numeric_limits<T>::max();
  b.pad = T(0);
  return b;
}

// The lanes whose slot of the tree chunk (d1, d2 in the lane) the query box
// hits (a bound the query's kind leaves open is the lowest or highest value,
// which every dim1 but NaN meets).
template <typename T>
__device__ inline uint64_t hit_lanes(const Box<T>& q, T d1, T d2) {
  return le_mask(q.min2, d2) & le_mask(q.lo, d1) & le_mask(d1, q.hi);
}

// Lane j's value, as a uniform (SGPR) value.
template <typename T>
__device__ inline T bcast(T v, uint32_t j) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int32_t, v), j));
  } else {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int32_t>(u), j);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int32_t>(u >> 32), j);
    return __builtin_bit_cast(T, (static_cast<uint64_t>(hi) << 32) | lo);
  }
}

// Lane ql's box, as uniform values.
template <typename T>
__device__ inline Box<T> lane_box(const Box<T>& b, uint32_t ql) {
  Box<T> r;
  r.lo = bcast(b.lo, ql);
  r.hi = bcast(b.hi, ql);
  r.min2 = bcast(b.min2, ql);
  r.pad = T(0);
  return r;
}

// Queries per pass over the batch: lane l of pass p tests query 64 p + l.
constexpr uint32_t kQSlots = static_cast<uint32_t>(kForestDenseMax);

// Stages the batch's query boxes in LDS (the whole block), query
// slots past the batch's end up to a multiple of 64 getting dead boxes.
template <typename T>
__device__ inline void stage_boxes(const Query<T>* __restrict__ queries, uint32_t nq, Box<T>* sbox) {
  const uint32_t padded = (nq + kWave - 1) / kWave * kWave;
  for (uint32_t i = threadIdx.x; i < padded; i += kFBlock) sbox[i] = i < nq ? box_of(queries[i]) : dead_box<T>();
}

// A tree staged for the wave: slot 64 r + lane of the tree in lane `lane`
// (coalesced vector loads), and per chunk r the slots that hold a point.
template <typename T, typename IS, int R>
__device__ inline void load_tree(const Node<T, IS>* __restrict__ tn, const ForestArgs& F, uint32_t M, T (&d1)[R],
                                 T (&d2)[R], uint64_t (&vm)[R]) {
  const uint32_t lane = lane_id();
  const uint32_t end = (1u << F.shift) - 1;  // the last slot of a tree's storage is never used
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t s = r * kWave + lane;
    const bool v = s < end && (M == F.full_M || subtree_size32(M, s) > 0);
    d1[r] = T(0);
    d2[r] = T(0);
    if (v) {
      const Node<T, IS> nd = tn[s];
      d1[r] = nd.dim1;
      d2[r] = nd.dim2;
    }
    vm[r] = __ballot(v);
  }
}


// The records of the lanes in m, in lane order, at oq[w], oq[w + 1], ...
// (positions at or past lim dropped); returns w + |m|. The mask is the
// lanes' predicate as it stands; the per-lane limit test only runs when
// the run crosses the limit.
template <typename Rec>
__device__ inline uint32_t emit_run(Rec* oq, uint32_t w, uint32_t lim, uint64_t m, const Rec& rec) {
  const uint32_t at = w + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
  const uint32_t pop = static_cast<uint32_t>(__builtin_popcountll(m));
  if (w + pop <= lim) {
    if (__builtin_amdgcn_inverse_ballot_w64(m)) oq[at] = rec;
  } else {
    if (__builtin_amdgcn_inverse_ballot_w64(m) && at < lim) oq[at] = rec;
  }
  return w + pop;
}

// Two chunks' runs, chunk a's records then chunk b's, from w on: one store
// takes every lane with a hit (its chunk-a record where it has one, else
// its chunk-b record), a second only the lanes that hit in both chunks.
// Returns w + |ma| + |mb|.
template <typename Rec>
__device__ inline uint32_t emit_two(Rec* oq, uint32_t w, uint32_t lim, uint64_t ma, uint64_t mb, const Rec& ra,
                                    const Rec& rb) {
  const uint32_t pa = static_cast<uint32_t>(__builtin_popcountll(ma));
  const uint32_t pb = static_cast<uint32_t>(__builtin_popcountll(mb));
  if (w + pa + pb > lim) return emit_run(oq, emit_run(oq, w, lim, ma, ra), lim, mb, rb);
  const uint32_t at_a = w + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(ma >> 32),
                                                      __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(ma), 0u));
  const uint32_t at_b = w + pa + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mb >> 32),
                                                           __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mb), 0u));
  if (__builtin_amdgcn_inverse_ballot_w64(ma | mb)) {
    const bool in_a = __builtin_amdgcn_inverse_ballot_w64(ma);
    oq[in_a ? at_a : at_b] = in_a ? ra : rb;
  }
  const uint64_t both = ma & mb;
  if (both && __builtin_amdgcn_inverse_ballot_w64(both)) oq[at_b] = rb;
  return w + pa + pb;
}

// Pass 1: segcnt[q * nseg + seg] = hits of query q in segment seg. One
// wave per segment. Per tree, lanes first take a query each (64 per pass)
// to find the candidates: the queries that may hit the tree (its dim1 range
// and its root's dim2, the tree's largest by heap order). Lanes then take
// the tree's slots, and each candidate is compared with all of them at
// once; its hit count goes to its LDS counter.
template <typename T, typename IS, int R>
__global__ __launch_bounds__(kFBlock) void k_forest_count(const Node<T, IS>* __restrict__ nodes,
                                                          const Range1<T>* __restrict__ bounds, ForestArgs F,
                                                          const Query<T>* __restrict__ queries,
                                                          const uint32_t* __restrict__ stab_flag,
                                                          uint32_t* __restrict__ segcnt) {
  if (*stab_flag) return;  // a stabbing batch: k_stab_count
  __shared__ Box<T> sbox[kQSlots];
  __shared__ uint32_t cnt_all[kFWaves][kQSlots];
  stage_boxes(queries, F.nq, sbox);
  const uint32_t lane = lane_id();
  uint32_t* cnt = cnt_all[threadIdx.x / kWave];
  for (uint32_t i = lane; i < kQSlots; i += kWave) cnt[i] = 0;
  __syncthreads();
  // the wave's segment, known to the compiler as uniform
  const uint32_t seg = __builtin_amdgcn_readfirstlane(blockIdx.x * kFWaves + threadIdx.x / kWave);
  if (seg >= F.nseg) return;
  const uint32_t npass = (F.nq + kWave - 1) / kWave;
  // segment seg holds trees seg, seg + nseg, seg + 2 nseg, ...: trees spread
  // over the whole forest, so that the segments' work is alike even where
  // hits cluster in some region of the input
  for (uint32_t tree = seg; tree < F.ntrees; tree += F.nseg) {
    const Node<T, IS>* tn = nodes + (static_cast<uint64_t>(tree) << F.shift);
    const Range1<T> b = bounds[tree];
    const T max2 = tn[0].dim2;
    const uint32_t M = tree + 1 == F.ntrees ? F.last_M : F.full_M;
    T d1[R], d2[R];
    uint64_t vm[R];
    bool loaded = false;
    for (uint32_t p = 0; p < npass; p++) {
      const Box<T> lb = sbox[p * kWave + lane];  // the lane's query of this pass
      uint64_t cm = __ballot(p * kWave + lane < F.nq && may_hit(lb, b.lo, b.hi, max2));
      if (!cm) continue;
      if (!loaded) {
        load_tree<T, IS, R>(tn, F, M, d1, d2, vm);
        loaded = true;
      }
      // two candidates per step: their compare chains are independent
      while (cm) {
        const uint32_t qa = static_cast<uint32_t>(__builtin_ctzll(cm));
        cm &= cm - 1;
        const bool two = cm != 0;
        const uint32_t qb = two ? static_cast<uint32_t>(__builtin_ctzll(cm)) : qa;
        cm &= cm - (two ? 1 : 0);
        const Box<T> ba = lane_box(lb, qa), bb = lane_box(lb, qb);  // from the candidates' lanes
        uint32_t ca = 0, cb = 0;
#pragma unroll
        for (int r = 0; r < R; r++) {
          ca += static_cast<uint32_t>(__builtin_popcountll(hit_lanes(ba, d1[r], d2[r]) & vm[r]));
          cb += static_cast<uint32_t>(__builtin_popcountll(hit_lanes(bb, d1[r], d2[r]) & vm[r]));
        }
        if (lane == 0) {
          if (ca) atomicAdd(&cnt[p * kWave + qa], ca);  // LDS adds, no return
          if (two && cb) atomicAdd(&cnt[p * kWave + qb], cb);
        }
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  for (uint32_t q = lane; q < F.nq; q += kWave) segcnt[static_cast<uint64_t>(q) * F.nseg + seg] = cnt[q];
}

// Per query (one workgroup each): exclusive scan of its segment counts in
// place; totals[q] = the query's hit count. Wave w takes a contiguous
// chunk of the row: it sums the chunk (coalesced loads), the waves' sums
// give each wave's prefix, and the wave then scans its chunk 64 counts at
// a time from that prefix (the row stays in L2 between the two reads).
__global__ __launch_bounds__(kScanBlock) void k_forest_seg_scan(uint32_t* __restrict__ segcnt, uint32_t nseg,
                                                                unsigned long long* __restrict__ totals) {
  constexpr uint32_t kWaves = kScanBlock / kWave;
  __shared__ unsigned long long part[kWaves];
  const uint32_t q = blockIdx.x;
  uint32_t* row = segcnt + static_cast<uint64_t>(q) * nseg;
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  const uint32_t chunk = (nseg + kWaves * kWave - 1) / (kWaves * kWave) * kWave;  // a multiple of 64
  const uint32_t lo = std::min(nseg, wave * chunk), hi = std::min(nseg, lo + chunk);
  // kU rows of 64 entries per step, their loads issued together
  constexpr int kU = 8;
  auto load_rows = [&](uint32_t i0, uint32_t (&v)[kU]) {
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const uint32_t i = i0 + u * kWave + lane;
      v[u] = i < hi ? row[i] : 0u;
    }
  };
  unsigned long long sum = 0;
  for (uint32_t i0 = lo; i0 < hi; i0 += kU * kWave) {
    uint32_t v[kU];
    load_rows(i0, v);
#pragma unroll
    for (int u = 0; u < kU; u++) sum += v[u];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, kWave);
  if (lane == 0) part[wave] = sum;
  __syncthreads();
  unsigned long long run = 0;
  for (uint32_t w = 0; w < wave; w++) run += part[w];
  for (uint32_t i0 = lo; i0 < hi; i0 += kU * kWave) {
    uint32_t v[kU];
    load_rows(i0, v);
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const uint32_t i = i0 + u * kWave + lane;
      unsigned long long inc = v[u];
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        const unsigned long long o = __shfl_up(inc, off, kWave);
        if (static_cast<int>(lane) >= off) inc += o;
      }
      if (i < hi) row[i] = static_cast<uint32_t>(run + inc - v[u]);
      run += __shfl(inc, kWave - 1, kWave);
    }
  }
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (uint32_t w = 0; w < kWaves; w++) t += part[w];
    totals[q] = t;
  }
}

// Pass 2: query q's hits of segment seg at out + offsets[q] + segoff[q][seg]
// + running count, in (tree, slot) order; positions at or past the query's
// limit (its capacity, and the output's end) are dropped. Candidates are
// found as in pass 1; per candidate and chunk of 64 slots, the lanes whose
// slot it hits write their record at consecutive positions (one contiguous
// store per candidate and chunk).
template <typename T, typename IS, bool RP, int R>
__global__ __launch_bounds__(kFBlock) void k_forest_write(const Node<T, IS>* __restrict__ nodes,
                                                          const void* __restrict__ ids_v,
                                                          const Range1<T>* __restrict__ bounds, ForestArgs F,
                                                          const Query<T>* __restrict__ queries,
                                                          const uint32_t* __restrict__ stab_flag,
                                                          const uint32_t* __restrict__ segoff,
                                                          const unsigned long long* __restrict__ offsets,
                                                          const unsigned long long* __restrict__ caps,
                                                          unsigned long long cap_all, unsigned long long out_total,
                                                          void* __restrict__ out) {
  if (*stab_flag) return;  // a stabbing batch: k_stab_write
  using Rec = typename std::conditional<RP, Point<T, IS>, IS>::type;
  using ISt = typename std::conditional<has_id_v<IS>, IS, uint32_t>::type;
  const ISt* __restrict__ ids = static_cast<const ISt*>(ids_v);
  __shared__ Box<T> sbox[kQSlots];
  __shared__ uint32_t pos_all[kFWaves][kQSlots];  // the query's next position (within the query's results)
  __shared__ unsigned long long soff[kQSlots];     // the query's first output record
  __shared__ uint32_t slim[kQSlots];               // positions below this are written
  stage_boxes(queries, F.nq, sbox);
  for (uint32_t i = threadIdx.x; i < F.nq; i += kFBlock) {
    const unsigned long long off = offsets ? offsets[i] : 0ull;
    const unsigned long long cap = caps ? caps[i] : cap_all;
    const unsigned long long room = off < out_total ? out_total - off : 0ull;
    const unsigned long long limit = cap < room ? cap : room;
    soff[i] = off;
    slim[i] = static_cast<uint32_t>(limit < 0xffffffffull ? limit : 0xffffffffull);
  }
  const uint32_t lane = lane_id();
  uint32_t* pos = pos_all[threadIdx.x / kWave];
  const uint32_t seg = __builtin_amdgcn_readfirstlane(blockIdx.x * kFWaves + threadIdx.x / kWave);
  if (seg < F.nseg)
    for (uint32_t i = lane; i < F.nq; i += kWave) pos[i] = segoff[static_cast<uint64_t>(i) * F.nseg + seg];
  __syncthreads();
  if (seg >= F.nseg) return;
  Rec* const o = static_cast<Rec*>(out);
  const uint32_t npass = (F.nq + kWave - 1) / kWave;
  // segment seg holds trees seg, seg + nseg, seg + 2 nseg, ...: trees spread
  // over the whole forest, so that the segments' work is alike even where
  // hits cluster in some region of the input
  for (uint32_t tree = seg; tree < F.ntrees; tree += F.nseg) {
    const uint64_t base = static_cast<uint64_t>(tree) << F.shift;
    const Node<T, IS>* tn = nodes + base;
    const Range1<T> b = bounds[tree];
    const T max2 = tn[0].dim2;
    const uint32_t M = tree + 1 == F.ntrees ? F.last_M : F.full_M;
    T d1[R], d2[R];
    uint64_t vm[R];
    Rec rv[R];  // lane l's record of chunk r: slot 64 r + l
    bool vl[R];  // and whether that slot holds a point
    bool loaded = false;
    for (uint32_t p = 0; p < npass; p++) {
      const Box<T> lb = sbox[p * kWave + lane];  // the lane's query of this pass
      uint64_t cm = __ballot(p * kWave + lane < F.nq && may_hit(lb, b.lo, b.hi, max2));
      if (!cm) continue;
      uint32_t lpos = pos[p * kWave + lane];  // the lane's query's next position, offset and limit
      const unsigned long long loff = soff[p * kWave + lane];
      const uint32_t llim = slim[p * kWave + lane];
      bool moved = false;
      if (!loaded) {
        load_tree<T, IS, R>(tn, F, M, d1, d2, vm);
#pragma unroll
        for (int r = 0; r < R; r++) {
          rv[r] = Rec{};
          vl[r] = (vm[r] >> lane) & 1ull;
          if (vl[r]) {
            if constexpr (RP) {
              rv[r].dim1_val = d1[r];
              rv[r].dim2_val = d2[r];
              if constexpr (has_id_v<IS>) rv[r].id = tn[r * kWave + lane].id;
            } else {
              rv[r] = ids[base + r * kWave + lane];
            }
          }
        }
        loaded = true;
      }
      // one candidate's hits in the tree: its lanes' records, in slot order,
      // from its next position on
      auto emit = [&](uint32_t ql, const uint64_t (&S)[R]) {
        Rec* const oq = o + bcast(loff, ql);
        const uint32_t lim = __builtin_amdgcn_readlane(llim, ql);
        uint32_t w = __builtin_amdgcn_readlane(lpos, ql);
#pragma unroll
        for (int r = 0; r + 1 < R; r += 2) w = emit_two(oq, w, lim, S[r], S[r + 1], rv[r], rv[r + 1]);
        if constexpr (R % 2 == 1) {
          if (S[R - 1]) w = emit_run(oq, w, lim, S[R - 1], rv[R - 1]);
        }
        lpos = lane == ql ? w : lpos;
      };
      // two candidates per step: their compare chains are independent
      while (cm) {
        const uint32_t qa = static_cast<uint32_t>(__builtin_ctzll(cm));
        cm &= cm - 1;
        const bool two = cm != 0;
        const uint32_t qb = two ? static_cast<uint32_t>(__builtin_ctzll(cm)) : qa;
        cm &= cm - (two ? 1 : 0);
        const Box<T> ba = lane_box(lb, qa), bb = lane_box(lb, qb);
        bool ma[R], mb[R];  // the candidate hits the lane's slot of chunk r
        uint64_t Sa[R], Sb[R];
        uint64_t anya = 0, anyb = 0;
#pragma unroll
        for (int r = 0; r < R; r++) {
          ma[r] = vl[r] & (ba.min2 <= d2[r]) & (ba.lo <= d1[r]) & (d1[r] <= ba.hi);
          mb[r] = vl[r] & two & (bb.min2 <= d2[r]) & (bb.lo <= d1[r]) & (d1[r] <= bb.hi);
          Sa[r] = __ballot(ma[r]);
          Sb[r] = __ballot(mb[r]);
          anya |= Sa[r];
          anyb |= Sb[r];
        }
        if (anya) emit(qa, Sa);
        if (anyb) emit(qb, Sb);
        moved = true;
      }
      if (moved) pos[p * kWave + lane] = lpos;
    }
  }
}

// ---------------------------------------------------------------------------
// Stabbing batches: every query a LEFT search whose dim1 bound equals its
// dim2 bound v (an isovalue query: dim1 <= v <= dim2). With the batch's
// values sorted, the queries that hit a point form one run of sorted
// positions, [a, b) with a = #values < dim1 and b = #values <= dim2, so a
// pass counts hits with two binary searches per point and a difference
// array per segment, and stores every slot's run; the write pass tests a
// candidate against a slot's stored run with two integer compares. NaN
// values hit nothing and sort last.
// ---------------------------------------------------------------------------
template <typename T>
struct StabPrep {
  uint32_t all_stab;      // 1: every query of the batch is a stabbing query
  uint32_t ns;            // non-NaN values
  uint32_t pad[2];
  T sv[kQSlots];          // the values in ascending order (ties by query index)
  uint32_t sidx[kQSlots];  // sorted position -> query index (NaN values after ns)
};

template <typename T>
__device__ inline bool is_stab(const Query<T>& q) {
  return q.kind == KIND_LEFT && (q.max_dim1 == q.min_dim2 || (is_nan(q.max_dim1) && is_nan(q.min_dim2)));
}

// One workgroup of kQSlots threads: the flag, then the sort (rank of each
// value among the batch's, ties by index).
template <typename T>
__global__ __launch_bounds__(kQSlots) void k_stab_prep(const Query<T>* __restrict__ queries, uint32_t nq,
                                                       StabPrep<T>* __restrict__ prep) {
  __shared__ T v[kQSlots];
  __shared__ uint32_t ok[kQSlots];
  const uint32_t i = threadIdx.x;
  bool stab = true, num = false;
  T x = T(0);
  if (i < nq) {
    const Query<T> q = queries[i];
    stab = is_stab(q);
    x = q.max_dim1;
    num = !is_nan(x);
  }
  v[i] = x;
  ok[i] = num ? 1u : 0u;
  const int all = __syncthreads_and(stab ? 1 : 0);
  const uint32_t ns = static_cast<uint32_t>(__syncthreads_count(num ? 1 : 0));
  if (i < nq) {
    uint32_t r = 0;
    if (num) {
      for (uint32_t j = 0; j < nq; j++) r += ok[j] & ((v[j] < x) | (v[j] == x && j < i) ? 1u : 0u);
    } else {
      r = ns;
      for (uint32_t j = 0; j < i; j++) r += ok[j] ^ 1u;
    }
    if (num) prep->sv[r] = x;
    prep->sidx[r] = i;
  }
  if (i == 0) {
    prep->all_stab = all ? 1u : 0u;
    prep->ns = ns;
  }
}

__device__ inline uint32_t top_step(uint32_t ns) { return ns ? 1u << (31 - __builtin_clz(ns)) : 0u; }

// Pass 1 of a stabbing batch: segcnt[q * nseg + seg] = hits of query q in
// segment seg, from each point's run [a, b) of sorted queries (+1 at a,
// -1 at b in the wave's difference array, then a prefix sum); each slot's
// run is stored (a | b << 16, 0 for an empty slot) for pass 2.
template <typename T, typename IS, int R>
__global__ __launch_bounds__(kFBlock) void k_stab_count(const Node<T, IS>* __restrict__ nodes, ForestArgs F,
                                                        const StabPrep<T>* __restrict__ prep,
                                                        uint32_t* __restrict__ segcnt, uint32_t* __restrict__ runs) {
  if (!prep->all_stab) return;
  __shared__ T ssv[kQSlots];
  __shared__ int32_t diff_all[kFWaves][kQSlots + 1];
  const uint32_t ns = prep->ns;
  for (uint32_t i = threadIdx.x; i < ns; i += kFBlock) ssv[i] = prep->sv[i];
  const uint32_t lane = lane_id();
  int32_t* diff = diff_all[threadIdx.x / kWave];
  for (uint32_t i = lane; i <= kQSlots; i += kWave) diff[i] = 0;
  __syncthreads();
  const uint32_t seg = __builtin_amdgcn_readfirstlane(blockIdx.x * kFWaves + threadIdx.x / kWave);
  if (seg >= F.nseg) return;
  const uint32_t top = top_step(ns);
  // the next tree's slots are loaded while the current one is worked on
  T n1[R], n2[R];
  uint64_t nvm[R];
  auto load = [&](uint32_t tree) {
    load_tree<T, IS, R>(nodes + (static_cast<uint64_t>(tree) << F.shift), F,
                        tree + 1 == F.ntrees ? F.last_M : F.full_M, n1, n2, nvm);
  };
  if (seg < F.ntrees) load(seg);
  for (uint32_t tree = seg; tree < F.ntrees; tree += F.nseg) {
    T d1[R], d2[R];
    uint64_t vm[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      d1[r] = n1[r];
      d2[r] = n2[r];
      vm[r] = nvm[r];
    }
    if (tree + F.nseg < F.ntrees) load(tree + F.nseg);
    uint32_t* const tr = runs + (static_cast<uint64_t>(tree) << F.shift);
    // the 2R searches of the lane in lockstep: independent LDS reads per step
    uint32_t pa[R], pb[R];
#pragma unroll
    for (int r = 0; r < R; r++) pa[r] = pb[r] = 0;
    for (uint32_t step = top; step > 0; step >>= 1) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const uint32_t ia = pa[r] + step, ib = pb[r] + step;
        if (ia <= ns && ssv[ia - 1] < d1[r]) pa[r] = ia;   // #values < dim1
        if (ib <= ns && ssv[ib - 1] <= d2[r]) pb[r] = ib;  // #values <= dim2
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      uint32_t a = 0, b = 0;
      if ((vm[r] >> lane) & 1ull) {
        a = pa[r];
        b = pb[r];
        if (a < b) {
          atomicAdd(&diff[a], 1);
          atomicAdd(&diff[b], -1);
        }
      }
      const uint32_t s = r * kWave + lane;
      if (s < (1u << F.shift)) tr[s] = a | (b << 16);  // the write pass's runs (empty: 0)
    }
  }
  __builtin_amdgcn_wave_barrier();
  // prefix sum over sorted positions, 64 at a time
  int32_t run = 0;
  const uint32_t nq = F.nq;
  for (uint32_t k0 = 0; k0 < nq; k0 += kWave) {
    const uint32_t k = k0 + lane;
    const int32_t d = k < ns ? diff[k] : 0;
    int32_t inc = d;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int32_t o = __shfl_up(inc, off, kWave);
      if (static_cast<int>(lane) >= off) inc += o;
    }
    if (k < nq) segcnt[static_cast<uint64_t>(prep->sidx[k]) * F.nseg + seg] = k < ns ? static_cast<uint32_t>(run + inc) : 0u;
    run += __shfl(inc, kWave - 1, kWave);
  }
}

// Pass 2 of a stabbing batch, with the output positions of k_forest_write:
// per tree, each slot's run [a, b) as pass 1 stored it; the candidates are
// the sorted positions [kmin, kmax) that some run covers, taken 64 at a
// time (lane l holding position w0 + l's next position, offset and limit);
// a candidate hits the lanes whose run holds it.
template <typename T, typename IS, bool RP, int R>
__global__ __launch_bounds__(kFBlock) void k_stab_write(const Node<T, IS>* __restrict__ nodes,
                                                        const void* __restrict__ ids_v, ForestArgs F,
                                                        const StabPrep<T>* __restrict__ prep,
                                                        const uint32_t* __restrict__ runs,
                                                        const uint32_t* __restrict__ segoff,
                                                        const unsigned long long* __restrict__ offsets,
                                                        const unsigned long long* __restrict__ caps,
                                                        unsigned long long cap_all, unsigned long long out_total,
                                                        void* __restrict__ out) {
  if (!prep->all_stab) return;
  using Rec = typename std::conditional<RP, Point<T, IS>, IS>::type;
  using ISt = typename std::conditional<has_id_v<IS>, IS, uint32_t>::type;
  const ISt* __restrict__ ids = static_cast<const ISt*>(ids_v);
  __shared__ uint32_t spos_all[kFWaves][kQSlots];  // by sorted position: the query's next position
  __shared__ unsigned long long soff[kQSlots];      // by sorted position: the query's first output record
  __shared__ uint32_t slim[kQSlots];                // and the positions below which records are written
  const uint32_t ns = prep->ns;
  for (uint32_t k = threadIdx.x; k < ns; k += kFBlock) {
    const uint32_t q = prep->sidx[k];
    const unsigned long long off = offsets ? offsets[q] : 0ull;
    const unsigned long long cap = caps ? caps[q] : cap_all;
    const unsigned long long room = off < out_total ? out_total - off : 0ull;
    const unsigned long long limit = cap < room ? cap : room;
    soff[k] = off;
    slim[k] = static_cast<uint32_t>(limit < 0xffffffffull ? limit : 0xffffffffull);
  }
  const uint32_t lane = lane_id();
  uint32_t* spos = spos_all[threadIdx.x / kWave];
  const uint32_t seg = __builtin_amdgcn_readfirstlane(blockIdx.x * kFWaves + threadIdx.x / kWave);
  if (seg < F.nseg)
    for (uint32_t k = lane; k < ns; k += kWave) spos[k] = segoff[static_cast<uint64_t>(prep->sidx[k]) * F.nseg + seg];
  __syncthreads();
  if (seg >= F.nseg) return;
  Rec* const o = static_cast<Rec*>(out);
  // the next tree's runs and records are loaded while the current one is
  // written (every slot of a tree's storage can be read)
  uint32_t nab[R];
  Rec nrv[R];
  auto load = [&](uint32_t tree) {
    const uint64_t base = static_cast<uint64_t>(tree) << F.shift;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const uint32_t s = r * kWave + lane;
      nab[r] = 0u;
      nrv[r] = Rec{};
      if (s < (1u << F.shift)) {
        nab[r] = runs[base + s];
        if constexpr (RP) {
          const Node<T, IS> nd = nodes[base + s];
          nrv[r].dim1_val = nd.dim1;
          nrv[r].dim2_val = nd.dim2;
          if constexpr (has_id_v<IS>) nrv[r].id = nd.id;
        } else {
          nrv[r] = ids[base + s];
        }
      }
    }
  };
  if (seg < F.ntrees) load(seg);
  for (uint32_t tree = seg; tree < F.ntrees; tree += F.nseg) {
    uint32_t a[R], b[R];  // the lane's slot of chunk r is hit by sorted positions [a, b) (empty: 0, 0)
    Rec rv[R];            // and its record
    uint32_t kmin = ns, kmax = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      a[r] = nab[r] & 0xffffu;
      b[r] = nab[r] >> 16;
      rv[r] = nrv[r];
      if (a[r] < b[r]) {
        kmin = a[r] < kmin ? a[r] : kmin;
        kmax = b[r] > kmax ? b[r] : kmax;
      }
    }
    if (tree + F.nseg < F.ntrees) load(tree + F.nseg);
    kmin = wave_min(kmin);
    kmax = wave_max(kmax);
    kmin = __builtin_amdgcn_readfirstlane(kmin);
    kmax = __builtin_amdgcn_readfirstlane(kmax);
    if (kmin >= kmax) continue;
    for (uint32_t w0 = kmin & ~(kWave - 1u); w0 < kmax; w0 += kWave) {
      const uint32_t kl = w0 + lane;  // the lane's sorted position of this window
      uint32_t lpos = 0, llim = 0;
      unsigned long long loff = 0;
      if (kl < ns) {
        lpos = spos[kl];
        loff = soff[kl];
        llim = slim[kl];
      }
      const uint32_t k_lo = kmin > w0 ? kmin : w0;
      const uint32_t k_hi = kmax < w0 + kWave ? kmax : w0 + kWave;
      for (uint32_t k = k_lo; k < k_hi; k++) {
        uint64_t S[R];
        uint64_t any = 0;
#pragma unroll
        for (int r = 0; r < R; r++) {
          S[r] = __builtin_amdgcn_uicmp(a[r], k, 37) & __builtin_amdgcn_uicmp(b[r], k, 34);  // a <= k < b
          any |= S[r];
        }
        if (!any) continue;
        const uint32_t ql = k - w0;
        Rec* const oq = o + bcast(loff, ql);
        const uint32_t lim = __builtin_amdgcn_readlane(llim, ql);
        uint32_t w = __builtin_amdgcn_readlane(lpos, ql);
#pragma unroll
        for (int r = 0; r + 1 < R; r += 2) w = emit_two(oq, w, lim, S[r], S[r + 1], rv[r], rv[r + 1]);
        if constexpr (R % 2 == 1) {
          if (S[R - 1]) w = emit_run(oq, w, lim, S[R - 1], rv[R - 1]);
        }
        lpos = lane == ql ? w : lpos;
      }
      if (kl < ns && kl >= k_lo && kl < k_hi) spos[kl] = lpos;
    }
  }
}

// bounds[tree] = the tree's dim1 range (one wave per tree); NaN values
// compare false and so never widen it.
template <typename T, typename IS>
__global__ __launch_bounds__(kFBlock) void k_forest_bounds(const Node<T, IS>* __restrict__ nodes, ForestArgs F,
                                                           Range1<T>* __restrict__ bounds) {
  const uint32_t tree = blockIdx.x * kFWaves + threadIdx.x / kWave, lane = lane_id();
  if (tree >= F.ntrees) return;
  constexpr T lowest = std::numeric_limits<T>::has_infinity ? -std::numeric_limits<T>::infinity()
                                                              : std::numeric_limits<T>::lowest();
  constexpr T highest = std::numeric_limits<T>::has_infinity ? std::numeric_limits<T>::infinity()
                                                               : std::numeric_limits<T>::max();
  const uint32_t M = tree + 1 == F.ntrees ? F.last_M : F.full_M;
  const Node<T, IS>* tn = nodes + (static_cast<uint64_t>(tree) << F.shift);
  const uint32_t end = (1u << F.shift) - 1;
  T mn = highest, mx = lowest;
  for (uint32_t s = lane; s < end; s += kWave) {
    if (subtree_size32(M, s) > 0) {
      const T d1 = tn[s].dim1;
      mn = d1 < mn ? d1 : mn;
      mx = d1 > mx ? d1 : mx;
    }
  }
  mn = wave_min(mn);
  mx = wave_max(mx);
  if (lane == 0) bounds[tree] = Range1<T>{mn, mx};
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
  // 112 segments per CU: about four turns of resident waves, so that the
  // hardware's dispatch evens out what the interleaving leaves uneven
  const uint32_t want = static_cast<uint32_t>(std::max(1, cus)) * 112u;
  F.nseg = std::min(F.ntrees, want);
  F.nq = static_cast<uint32_t>(a.nq);
  return F;
}

// The per (query, segment) counters, then the stabbing-batch table, then
// per slot of the forest its run of sorted positions (two 16-bit values).
size_t stab_prep_at(const ForestArgs& F) { return (static_cast<size_t>(F.nseg) * F.nq * 4 + 255) & ~size_t{255}; }
size_t stab_runs_at(const ForestArgs& F) { return stab_prep_at(F) + ((sizeof(StabPrep<double>) + 255) & ~size_t{255}); }

}  // namespace

size_t forest_dense_scratch_bytes(size_t n, size_t slots_per_tree, size_t nq, int cus) {
  SearchArgs a{};
  a.n = n;
  a.slots_per_tree = slots_per_tree;
  a.nq = nq;
  const ForestArgs F = forest_args(a, cus);
  return stab_runs_at(F) + (static_cast<size_t>(F.ntrees) << F.shift) * 4;
}

Status forest_dense_search(const SearchArgs& a, void* segbuf, hipStream_t s, int plan) {
  if (a.nq == 0 || a.n == 0) return Status{};
  if (a.nq > kForestDenseMax) return Status{1, "Error: too many queries for a stationary forest pass"};
  if (!a.tree_bounds) return Status{1, "Error: the forest has no per-tree dim1 ranges"};
  int dev = 0, cus = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const ForestArgs F = forest_args(a, cus);
  if (F.shift > 9) return Status{1, "Error: stationary forest passes hold trees of at most 511 points"};
  uint32_t* segcnt = static_cast<uint32_t*>(segbuf);
  return fwith_t(a.dtype, [&](auto tv) -> Status {
    using T = decltype(tv);
    return fwith_is(a.id_dtype, [&](auto iv) -> Status {
      using IS = decltype(iv);
      const auto* nodes = static_cast<const Node<T, IS>*>(a.nodes);
      const auto* qs = static_cast<const Query<T>*>(a.queries);
      const auto* bounds = static_cast<const Range1<T>*>(a.tree_bounds);
      auto* prep = reinterpret_cast<StabPrep<T>*>(static_cast<unsigned char*>(segbuf) + stab_prep_at(F));
      const uint32_t* stab = &prep->all_stab;
      auto* runs = reinterpret_cast<uint32_t*>(static_cast<unsigned char*>(segbuf) + stab_runs_at(F));
      const unsigned grid = (F.nseg + kFWaves - 1) / kFWaves;
      auto go = [&](auto rtag) -> Status {
        constexpr int R = decltype(rtag)::value;
        // whether the batch is all stabbing queries, and their sorted values;
        // of each pair of passes below, the one for the other kind of batch
        // returns at once. Plan 2 writes the batch a plan 1 count pass has
        // just counted, from its kept segment offsets, and its prep.
        const bool kept = a.mode != 0 && plan == 2;
        if (!kept) {
          hipLaunchKernelGGL((k_stab_prep<T>), dim3(1), dim3(kQSlots), 0, s, qs, F.nq, prep);
          // the stabbing pass first: while it runs, the host queues the rest
          hipLaunchKernelGGL((k_stab_count<T, IS, R>), dim3(grid), dim3(kFBlock), 0, s, nodes, F, prep, segcnt, runs);
          hipLaunchKernelGGL((k_forest_count<T, IS, R>), dim3(grid), dim3(kFBlock), 0, s, nodes, bounds, F, qs,
                             stab, segcnt);
          // write mode without kept offsets: the per-query totals go to the cursor
          hipLaunchKernelGGL(k_forest_seg_scan, dim3(F.nq), dim3(kScanBlock), 0, s, segcnt, F.nseg, a.counts);
        }
        if (a.mode == 0) return hip_status(hipGetLastError(), "Error in launching forest search");
        if (a.ret_point) {
          hipLaunchKernelGGL((k_forest_write<T, IS, true, R>), dim3(grid), dim3(kFBlock), 0, s, nodes, a.ids, bounds,
                             F, qs, stab, segcnt, a.offsets, a.caps, a.cap_all, a.out_total, a.out);
          hipLaunchKernelGGL((k_stab_write<T, IS, true, R>), dim3(grid), dim3(kFBlock), 0, s, nodes, a.ids, F, prep,
                             runs, segcnt, a.offsets, a.caps, a.cap_all, a.out_total, a.out);
        } else {
          if constexpr (has_id_v<IS>) {
            hipLaunchKernelGGL((k_forest_write<T, IS, false, R>), dim3(grid), dim3(kFBlock), 0, s, nodes, a.ids,
                               bounds, F, qs, stab, segcnt, a.offsets, a.caps, a.cap_all, a.out_total, a.out);
            hipLaunchKernelGGL((k_stab_write<T, IS, false, R>), dim3(grid), dim3(kFBlock), 0, s, nodes, a.ids, F,
                               prep, runs, segcnt, a.offsets, a.caps, a.cap_all, a.out_total, a.out);
          }
        }
        return hip_status(hipGetLastError(), "Error in launching forest search");
      };
      if (!a.ret_point && !has_id_v<IS> && a.mode != 0)
        return Status{1, "Error: IDs requested from a tree built without IDs"};
      switch (F.shift) {
        case 7: return go(std::integral_constant<int, 2>{});
        case 8: return go(std::integral_constant<int, 4>{});
        case 9: return go(std::integral_constant<int, 8>{});
        default: return go(std::integral_constant<int, 1>{});
      }
    });
  });
}

size_t forest_bounds_bytes(int dtype, size_t n, size_t slots_per_tree) {
  if (n == 0 || slots_per_tree == 0) return 0;
  return (n + slots_per_tree - 1) / slots_per_tree * 2 * dtype_size(dtype);
}

Status build_forest_bounds(int dtype, int id_dtype, size_t n, size_t slots_per_tree, const void* nodes_d,
                           void* bounds_d, hipStream_t s) {
  if (n == 0) return Status{};
  SearchArgs a{};
  a.n = n;
  a.slots_per_tree = slots_per_tree;
  a.nq = 1;
  const ForestArgs F = forest_args(a, 1);
  return fwith_t(dtype, [&](auto tv) -> Status {
    using T = decltype(tv);
    return fwith_is(id_dtype, [&](auto iv) -> Status {
      using IS = decltype(iv);
      const unsigned grid = (F.ntrees + kFWaves - 1) / kFWaves;
      hipLaunchKernelGGL((k_forest_bounds<T, IS>), dim3(grid), dim3(kFBlock), 0, s,
                         static_cast<const Node<T, IS>*>(nodes_d), F, static_cast<Range1<T>*>(bounds_d));
      return hip_status(hipGetLastError(), "Error in launching the forest's per-tree ranges");
    });
  });
}

}  // namespace pst
