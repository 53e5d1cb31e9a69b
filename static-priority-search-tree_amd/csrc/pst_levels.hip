// Level index of one deep tree, and the batched searches that use it.
//
// For every depth t of the tree (t = 0 .. H-1, H = bit length of n) the
// index holds the points of all subtrees rooted at depth t, subtree by
// subtree in position order, each subtree's points sorted by dim2
// descending:
//   key[t * ns + k] = ~ord_key(dim2)   (ascending along a subtree's run)
//   id [t * ns + k] = the point's ID
// (level stride ns = n rounded up to 64, so every level starts aligned).
// The run of level-t position j starts at level_start(M, t, j) and holds
// level_size(M, t, j) points (M = n + 1, the closed-form tree shape).
//
// Why: a walk that reaches a subtree in REPORT_ABOVE state (both dim1
// bounds closed for it) reports exactly the subtree's points with
// dim2 >= min_dim2 (heap order: a node's dim2 is at least its
// descendants'), i.e. a prefix of the subtree's run, found by one binary
// search and copied as one contiguous range. What remains of the walk are
// the at most two paths on which a dim1 bound is still open, at most two
// nodes per level. A batch then costs a few dozen node visits and binary
// searches per query and one streaming copy of 4 bytes read + 4 bytes
// written per reported 4-byte ID, instead of a node visit per ID.
//
// Equal dim2 values keep breadth-first slot order inside a run, so a
// subtree root (the smallest slot of its subtree, with the largest dim2)
// is the first element of its run: a path node's own report is the
// one-element range at the front of its run.
//
// Construction: level 0 is one stable radix sort of all points by key (the
// input in slot order, last-level slots in bit-reversed position order);
// level t+1 is the stable partition of every level-t run into the runs of
// the root's left and right children (the root itself dropped). Each
// point's destination is its child run's closed-form start plus its rank
// among the points of its run going the same way, from a global prefix
// count of left- and right-going points (per 4096-point tile, then
// scanned) less the closed-form count of same-way points in earlier runs.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pst_internal.hpp"

namespace pst {
namespace {

constexpr int kLvBlock = 256;
constexpr int kLvPer = 16;
constexpr uint32_t kLvTile = kLvBlock * kLvPer;  // points per partition tile
constexpr uint32_t kTaskCap = 128;               // ranges per query: <= 4 per level, <= 31 levels
constexpr uint32_t kSearchBit = 0x80000000u;     // range needs a binary search (a REPORT_ABOVE subtree)
constexpr uint64_t kMinChunk = 32768;            // IDs per copy chunk (at least)
constexpr uint64_t kChunkCap = 1u << 20;         // chunk table entries
constexpr int kCopyBlock = 256;
constexpr int kCopyBytes = 256;  // bytes of loads in flight per lane in the copy
constexpr uint32_t kFlat = 4096;  // positions per flat pass of the copy (16 per lane)

template <typename T>
using key_of = decltype(ord_key(T{}));

__host__ __device__ inline uint64_t level_stride(uint64_t n) { return (n + 63) & ~63ull; }

// The kLvPer consecutive elements from a[base] (zeros past nt): aligned
// 16-byte loads when all are in range (base is a multiple of kLvPer and
// every level starts 256-byte aligned).
template <typename V>
__device__ inline void load_run(const V* __restrict__ a, uint32_t base, uint32_t nt, V (&v)[kLvPer]) {
  if (base + kLvPer <= nt) {
    const uint4* q = reinterpret_cast<const uint4*>(a + base);
    constexpr int kVec = sizeof(V) * kLvPer / 16;
    constexpr int kPer = 16 / sizeof(V);
#pragma unroll
    for (int w = 0; w < kVec; w++) {
      const uint4 x = q[w];
      __builtin_memcpy(&v[w * kPer], &x, 16);
    }
  } else {
#pragma unroll
    for (int u = 0; u < kLvPer; u++) v[u] = base + u < nt ? a[base + u] : V{};
  }
}

// #{ j' < j : brev_t(j') < r } (positions are t-bit numbers; r may exceed
// 2^t - 1, counting every position). Positions j' < j are grouped by the
// highest bit b where they differ from j (bit b set in j, clear in j'):
// the group's bit reversals are c + k 2^(t-b), k < 2^b.
__host__ __device__ inline uint64_t brev_below(uint32_t t, uint32_t j, uint64_t r) {
  uint64_t cnt = 0;
  for (uint32_t b = 0; b < t; b++) {
    if ((j >> b) & 1u) {
      const uint64_t c = brev_bits((j >> b) ^ 1u, t - b);
      if (r > c) {
        const uint64_t step = 1ull << (t - b);
        const uint64_t k = (r - c + step - 1) / step;
        const uint64_t cap = 1ull << b;
        cnt += k < cap ? k : cap;
      }
    }
  }
  return cnt;
}

// Points of level-(t+1) positions 2j'' + side, j'' < j: the same-way points
// of the level-t runs before run j.
__host__ __device__ inline uint64_t side_before(uint64_t M, uint32_t t, uint32_t j, uint32_t side) {
  const uint64_t base = (M >> (t + 1)) - 1;
  const uint64_t r1 = M & ((2ull << t) - 1ull);
  const uint64_t off = static_cast<uint64_t>(side) << t;
  const uint64_t r = r1 > off ? r1 - off : 0;
  return static_cast<uint64_t>(j) * base + brev_below(t, j, r);
}

// Slot of input element e of the level-0 sort: every slot above the last
// level in order, then the last level's occupied positions (those whose
// bit reversal is below the count) in bit-reversed order.
__host__ __device__ inline uint32_t input_slot(uint32_t e, uint32_t H) {
  const uint32_t F = (1u << (H - 1)) - 1u;
  return e < F ? e : F + brev_bits(e - F, H - 1);
}

template <typename T, typename IS>
__global__ __launch_bounds__(kLvBlock) void k_lv_keys(const Node<T, IS>* __restrict__ nodes, uint32_t n, uint32_t H,
                                                      key_of<T>* __restrict__ keys, uint32_t* __restrict__ slots) {
  const uint32_t e = blockIdx.x * kLvBlock + threadIdx.x;
  if (e >= n) return;
  const uint32_t s = input_slot(e, H);
  keys[e] = ~ord_key(nodes[s].dim2);
  slots[e] = s;
}

template <typename IS>
__global__ __launch_bounds__(kLvBlock) void k_lv_ids0(const IS* __restrict__ ids, const uint32_t* __restrict__ slots,
                                                      uint32_t n, IS* __restrict__ out) {
  const uint32_t e = blockIdx.x * kLvBlock + threadIdx.x;
  if (e >= n) return;
  out[e] = ids[slots[e]];
}

// Direction of the level-t point in slot s: 0 left, 1 right, 2 the root of
// its run (dropped).
__device__ inline uint32_t lv_side(uint32_t s, uint32_t t) {
  const uint32_t d = depth_of(s);
  if (d == t) return 2u;
  return ((s + 1u) >> (d - t - 1u)) & 1u;
}

// Per tile: left-going count (low 32 bits) and right-going count (high).
__global__ __launch_bounds__(kLvBlock) void k_lv_tile_counts(const uint32_t* __restrict__ slots, uint32_t nt,
                                                             uint32_t t, unsigned long long* __restrict__ agg) {
  __shared__ unsigned long long wsum[kLvBlock / kWave];
  const uint32_t base = blockIdx.x * kLvTile + threadIdx.x * kLvPer;
  uint32_t sl[kLvPer];
  load_run(slots, base, nt, sl);
  unsigned long long c = 0;
#pragma unroll
  for (int u = 0; u < kLvPer; u++) {
    if (base + u < nt) {
      const uint32_t sd = lv_side(sl[u], t);
      c += sd == 0 ? 1ull : (sd == 1 ? (1ull << 32) : 0ull);
    }
  }
  for (int o = kWave / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    for (int w = 0; w < kLvBlock / kWave; w++) s += wsum[w];
    agg[blockIdx.x] = s;
  }
}

// Level t -> t+1: every point but its run's root moves to its child run.
// Phase 1 (a thread per 16 consecutive points): each point's destination,
// into LDS. Phase 2 (a thread per point, 256 consecutive points per pass):
// the moves, so that a wave's stores fall in at most a few contiguous
// stretches of the child runs and leave as whole lines.
constexpr uint32_t kDstPad = kLvPer + 1;  // LDS row of a thread's 16 destinations, padded against bank conflicts
template <typename K, typename IS>
__global__ __launch_bounds__(kLvBlock) void k_lv_scatter(const uint32_t* __restrict__ slots_in,
                                                         uint32_t* __restrict__ slots_out,
                                                         const K* __restrict__ key_in, K* __restrict__ key_out,
                                                         const IS* __restrict__ id_in, IS* __restrict__ id_out,
                                                         const unsigned long long* __restrict__ tile_pre,
                                                         uint32_t nt, uint32_t t, uint64_t M) {
  __shared__ unsigned long long wsum[kLvBlock / kWave];
  __shared__ uint32_t s_dst[kLvBlock * kDstPad];
  const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint32_t tile = blockIdx.x * kLvTile;
  const uint32_t base = tile + tid * kLvPer;
  uint32_t s[kLvPer];
  load_run(slots_in, base, nt, s);
  unsigned long long loc = 0;
#pragma unroll
  for (int u = 0; u < kLvPer; u++) {
    if (base + u < nt) {
      const uint32_t sd = lv_side(s[u], t);
      loc += sd == 0 ? 1ull : (sd == 1 ? (1ull << 32) : 0ull);
    }
  }
  unsigned long long incl = loc;
  for (int o = 1; o < kWave; o <<= 1) {
    const unsigned long long y = __shfl_up(incl, o, kWave);
    if (static_cast<int>(lane) >= o) incl += y;
  }
  if (lane == kWave - 1) wsum[wave] = incl;
  __syncthreads();
  unsigned long long run = tile_pre[blockIdx.x] + incl - loc;
  for (uint32_t w = 0; w < wave; w++) run += wsum[w];
  // A thread's points lie in runs cj, cj + 1, ... (every level-t run
  // holds at least its root), so the closed-form counts of the first run
  // are advanced by the child run sizes from then on.
  uint32_t cj = 0;
  uint64_t cl = 0, cr = 0;
  bool have = false;
#pragma unroll
  for (int u = 0; u < kLvPer; u++) {
    uint32_t dst = ~0u;  // the run's root (dropped) or past the level
    if (base + u < nt) {
      const uint32_t d = depth_of(s[u]);
      const uint32_t j = ((s[u] + 1u) >> (d - t)) - (1u << t);
      if (!have) {
        cj = j;
        cl = side_before(M, t, j, 0);
        cr = side_before(M, t, j, 1);
        have = true;
      }
      while (cj < j) {
        cl += static_cast<uint64_t>(level_size(M, t + 1, 2 * cj));
        cr += static_cast<uint64_t>(level_size(M, t + 1, 2 * cj + 1));
        cj++;
      }
      if (d != t) {
        const uint32_t sd = ((s[u] + 1u) >> (d - t - 1u)) & 1u;
        const uint64_t start = sd == 0 ? cl + cr : cl + cr + static_cast<uint64_t>(level_size(M, t + 1, 2 * cj));
        const uint64_t rank = sd == 0 ? (run & 0xffffffffull) - cl : (run >> 32) - cr;
        dst = static_cast<uint32_t>(start + rank);
        run += sd == 0 ? 1ull : (1ull << 32);
      }
    }
    s_dst[tid * kDstPad + u] = dst;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kLvPer; u++) {
    const uint32_t r = u * kLvBlock + tid;  // point within the tile
    const uint32_t p = tile + r;
    const uint32_t dst = s_dst[(r / kLvPer) * kDstPad + r % kLvPer];
    if (p < nt && dst != ~0u) {
      key_out[dst] = key_in[p];
      id_out[dst] = id_in[p];
      slots_out[dst] = slots_in[p];
    }
  }
}

__host__ __device__ inline uint64_t chunk_len(uint64_t W) {
  uint64_t c = (W + kChunkCap - 1) / kChunkCap;
  c = (c + 255) & ~255ull;
  return c > kMinChunk ? c : kMinChunk;
}

template <typename T, typename IS>
struct LvParams {
  const Node<T, IS>* nodes;
  const key_of<T>* key;
  const IS* id;
  uint64_t n;
  uint64_t ns;  // level stride
  uint32_t M, H;
  const Query<T>* queries;
  Query<T> single;
  uint32_t use_single;
  uint32_t keep;  // count mode: also keep the ranges (unclipped) for a later write from them
  uint32_t nq;
  unsigned long long* counts;  // count mode: per-query counts; write mode: per-query totals
  const unsigned long long* offsets;
  const unsigned long long* caps;
  unsigned long long cap_all;
  unsigned long long out_total;
  uint64_t* tsrc;   // per query, kTaskCap ranges: first index in the level arrays
  uint32_t* tpre;   // per query, kTaskCap: output position of each range (then the query's total)
  uint32_t* ntask;  // ranges per query
  unsigned long long* wq;  // IDs written per query
  unsigned long long* woff;  // one query (use_single, write mode): its work offsets, written by the plan
  uint32_t* chunk_task;      // and its chunk table
};

// One workgroup (one wave) per query: the open paths, level by level, then
// the REPORT_ABOVE subtrees' binary searches, one lane per subtree.
template <typename T, typename IS, bool W>
__global__ __launch_bounds__(kWave) void k_lv_plan(LvParams<T, IS> P) {
  using K = key_of<T>;
  __shared__ uint64_t s_src[kTaskCap];
  __shared__ uint32_t s_len[kTaskCap];
  const uint32_t lane = threadIdx.x;
  const uint32_t q = blockIdx.x;
  const Query<T> Q = P.use_single ? P.single : P.queries[q];
  const Bounds<T> qb{Q.min_dim1, Q.max_dim1, Q.min_dim2};
  const BoundNum bn = bound_num<T, 1>(&qb);
  const uint64_t ns = P.ns;
  // a path node: slot | l << 32 | h << 33 | valid << 34
  constexpr uint64_t kValid = 1ull << 34;
  uint64_t mine = lane == 0 ? (kValid | (static_cast<uint64_t>(lo_open(Q.kind)) << 32) |
                               (static_cast<uint64_t>(hi_open(Q.kind)) << 33))
                            : 0ull;
  uint32_t ntask = 0;
  for (uint32_t t = 0; t < P.H; t++) {
    if (__ballot((mine & kValid) != 0) == 0) break;
    uint32_t emit = 0;  // bit 0: own point, bit 1: a REPORT_ABOVE child
    uint64_t own_src = 0, ra_src = 0;
    uint32_t ra_size = 0;
    uint64_t nx0 = 0, nx1 = 0;
    uint32_t nnx = 0;
    if (mine & kValid) {
      const uint32_t slot = static_cast<uint32_t>(mine);
      const uint32_t l = static_cast<uint32_t>(mine >> 32) & 1u, h = static_cast<uint32_t>(mine >> 33) & 1u;
      const Node<T, IS> nd = P.nodes[slot];
      const int32_t m = subtree_size32(P.M, slot);
      const StepOut o = node_step<T, 1>(nd.dim1, nd.dim2, nd.med, m, &qb, bn, 1u, l, h);
      const uint32_t j = slot + 1u - (1u << t);
      if (o.rep) {
        emit |= 1u;
        own_src = static_cast<uint64_t>(t) * ns + static_cast<uint64_t>(level_start(P.M, t, j));
      }
      auto child = [&](uint32_t a, uint32_t cl, uint32_t ch, uint32_t side) {
        if (!a) return;
        const uint32_t cs = 2u * slot + 1u + side;
        if (!cl && !ch) {
          emit |= 2u;
          const uint32_t cjj = 2u * j + side;
          ra_src = static_cast<uint64_t>(t + 1) * ns + static_cast<uint64_t>(level_start(P.M, t + 1, cjj));
          ra_size = static_cast<uint32_t>(level_size(P.M, t + 1, cjj));
        } else {
          const uint64_t e = kValid | cs | (static_cast<uint64_t>(cl) << 32) | (static_cast<uint64_t>(ch) << 33);
          if (nnx == 0) nx0 = e;
          else nx1 = e;
          nnx++;
        }
      };
      child(o.al, o.ll, o.hl, 0u);
      child(o.ar, o.lr, o.hr, 1u);
    }
    // ranges of this level: lane 0's (own, subtree), then lane 1's
    const uint32_t c0 = __shfl(__builtin_popcount(emit), 0, kWave);
    if (lane < 2 && emit) {
      uint32_t at = ntask + (lane == 0 ? 0u : c0);
      if (emit & 1u) {
        s_src[at] = own_src;
        s_len[at] = 1u;
        at++;
      }
      if (emit & 2u) {
        s_src[at] = ra_src;
        s_len[at] = ra_size | kSearchBit;
      }
    }
    ntask += c0 + __shfl(__builtin_popcount(emit), 1, kWave);
    // next level's paths: lane 0's (at most two), then lane 1's (at most one)
    const uint64_t a0 = __shfl(nx0, 0, kWave), a1 = __shfl(nx1, 0, kWave), b0 = __shfl(nx0, 1, kWave);
    const uint32_t k0 = __shfl(nnx, 0, kWave);
    const uint64_t p0 = k0 >= 1 ? a0 : b0;
    const uint64_t p1 = k0 >= 2 ? a1 : (k0 == 1 ? b0 : 0ull);
    mine = lane == 0 ? p0 : (lane == 1 ? p1 : 0ull);
  }
  __syncthreads();
  // binary searches: points of the subtree's run with key <= ~ord_key(min2)
  const K qk = ~ord_key(qb.min2);
  uint32_t cnt[2] = {0, 0};
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const uint32_t i = lane + r * kWave;
    if (i < ntask) {
      const uint32_t len = s_len[i];
      if (len & kSearchBit) {
        const K* kk = P.key + s_src[i];
        uint32_t lo = 0, hi = len & ~kSearchBit;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (kk[mid] <= qk) lo = mid + 1;
          else hi = mid;
        }
        cnt[r] = lo;
      } else {
        cnt[r] = len;
      }
    }
  }
  // inclusive prefix over the ranges in order
  uint32_t inc0 = cnt[0], inc1 = cnt[1];
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y0 = __shfl_up(inc0, o, kWave), y1 = __shfl_up(inc1, o, kWave);
    if (static_cast<int>(lane) >= o) {
      inc0 += y0;
      inc1 += y1;
    }
  }
  const uint32_t tot0 = __shfl(inc0, kWave - 1, kWave);
  inc1 += tot0;
  const uint32_t total = __shfl(inc1, kWave - 1, kWave);
  if (!W && !P.keep) {
    if (lane == 0) P.counts[q] = total;
  } else {
    unsigned long long clip = total;
    if constexpr (W) {
      const unsigned long long off = P.offsets ? P.offsets[q] : 0ull;
      const unsigned long long cap = P.caps ? P.caps[q] : P.cap_all;
      const unsigned long long room = off < P.out_total ? P.out_total - off : 0ull;
      if (cap < clip) clip = cap;
      if (room < clip) clip = room;
    }
    const uint32_t cw = static_cast<uint32_t>(clip);
    // keep the ranges that start before the clip point and hold IDs
    const uint32_t pre0 = inc0 - cnt[0], pre1 = inc1 - cnt[1];
    const bool k0 = lane < ntask && cnt[0] > 0 && pre0 < cw;
    const bool k1 = lane + kWave < ntask && cnt[1] > 0 && pre1 < cw;
    const uint64_t b0 = __ballot(k0), b1 = __ballot(k1);
    const uint32_t at0 = lanes_below(b0), at1 = static_cast<uint32_t>(__builtin_popcountll(b0)) + lanes_below(b1);
    const uint64_t qb0 = static_cast<uint64_t>(q) * kTaskCap;
    if (k0) {
      P.tsrc[qb0 + at0] = s_src[lane];
      P.tpre[qb0 + at0] = pre0;
    }
    if (k1) {
      P.tsrc[qb0 + at1] = s_src[lane + kWave];
      P.tpre[qb0 + at1] = pre1;
    }
    if (lane == 0) {
      const uint32_t nt = static_cast<uint32_t>(__builtin_popcountll(b0) + __builtin_popcountll(b1));
      P.tpre[qb0 + nt] = cw;
      P.ntask[q] = nt;
      if constexpr (W) P.wq[q] = clip;
      P.counts[q] = total;
    }
    if constexpr (W) {
      if (P.use_single) {
        // one query: its work offsets and chunk table directly (no scan or
        // marking launch); a kept range ends where the next one starts
        if (lane == 0) {
          P.woff[0] = 0;
          P.woff[1] = clip;
        }
        const uint64_t L = chunk_len(clip);
        auto mark = [&](uint32_t at, uint32_t pre, uint32_t c) {
          const uint64_t a = pre, b = pre + c < cw ? pre + c : cw;
          for (uint64_t ch = (a + L - 1) / L; ch * L < b; ch++) P.chunk_task[ch] = at;
        };
        if (k0) mark(at0, pre0, cnt[0]);
        if (k1) mark(at1, pre1, cnt[1]);
      }
    }
  }
}

// Write step from ranges kept by a count pass: clip each query's ranges to
// its capacity and the output's size (a thread per query).
__global__ __launch_bounds__(kLvBlock) void k_lv_clip(uint32_t* __restrict__ tpre, uint32_t* __restrict__ ntask,
                                                      unsigned long long* __restrict__ wq,
                                                      unsigned long long* __restrict__ totals,
                                                      const unsigned long long* __restrict__ offsets,
                                                      const unsigned long long* __restrict__ caps,
                                                      unsigned long long cap_all, unsigned long long out_total,
                                                      uint32_t nq) {
  const uint32_t q = blockIdx.x * kLvBlock + threadIdx.x;
  if (q >= nq) return;
  const uint64_t qb = static_cast<uint64_t>(q) * kTaskCap;
  uint32_t nt = ntask[q];
  const uint32_t total = tpre[qb + nt];
  const unsigned long long off = offsets ? offsets[q] : 0ull;
  const unsigned long long cap = caps ? caps[q] : cap_all;
  const unsigned long long room = off < out_total ? out_total - off : 0ull;
  unsigned long long clip = total;
  if (cap < clip) clip = cap;
  if (room < clip) clip = room;
  if (clip < total) {
    while (nt > 0 && tpre[qb + nt - 1] >= clip) nt--;
    ntask[q] = nt;
    tpre[qb + nt] = static_cast<uint32_t>(clip);
  }
  wq[q] = clip;
  totals[q] = total;
}

// Chunk table: chunk c (IDs [c L, (c+1) L) of the batch's concatenated
// writes) starts in range i of query q: entry q << 7 | i.
__global__ __launch_bounds__(kLvBlock) void k_lv_mark(const unsigned long long* __restrict__ woff,
                                                      const uint32_t* __restrict__ tpre,
                                                      const uint32_t* __restrict__ ntask, uint32_t nq,
                                                      uint32_t* __restrict__ chunk_task) {
  // one thread per (query, range)
  const uint64_t g = static_cast<uint64_t>(blockIdx.x) * kLvBlock + threadIdx.x;
  const uint32_t q = static_cast<uint32_t>(g / kTaskCap), i = static_cast<uint32_t>(g % kTaskCap);
  if (q >= nq || i >= ntask[q]) return;
  const uint64_t L = chunk_len(woff[nq]);
  const uint64_t base = woff[q];
  const uint64_t qb = static_cast<uint64_t>(q) * kTaskCap;
  const uint64_t a = base + tpre[qb + i], b = base + tpre[qb + i + 1];
  for (uint64_t c = (a + L - 1) / L; c * L < b; c++) chunk_task[c] = (q << 7) | i;
}

// All lanes of the workgroup copy len IDs, a multiple of the pass size,
// with kCopyBytes of loads in flight per lane (every load of a pass issued
// before its stores).
template <typename IS>
__device__ inline void copy_passes(const IS* __restrict__ src, IS* __restrict__ dst, uint64_t len, uint32_t tid) {
  constexpr int U = kCopyBytes / sizeof(IS);
  constexpr uint64_t kPass = static_cast<uint64_t>(kCopyBlock) * U;
  for (uint64_t k0 = 0; k0 < len; k0 += kPass) {
    IS v[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = src[k0 + u * kCopyBlock + tid];
#pragma unroll
    for (int u = 0; u < U; u++) dst[k0 + u * kCopyBlock + tid] = v[u];
  }
}

// Streams the batch's ranges to the outputs, a chunk of IDs per workgroup
// pass. The current query's range table (output positions and sources) is
// staged in LDS. A range's whole passes are copied with all lanes on it;
// what is left of it, and the ranges after it in the same query and chunk,
// go in flat passes of kFlat positions where each lane finds its
// position's range in the table, so short ranges share a memory round
// trip.
template <typename IS>
__global__ __launch_bounds__(kCopyBlock) void k_lv_copy(const IS* __restrict__ ids, const uint64_t* __restrict__ tsrc,
                                                        const uint32_t* __restrict__ tpre,
                                                        const uint32_t* __restrict__ ntask,
                                                        const unsigned long long* __restrict__ woff,
                                                        const unsigned long long* __restrict__ offsets,
                                                        const uint32_t* __restrict__ chunk_task, uint32_t nq,
                                                        IS* __restrict__ out) {
  __shared__ uint32_t s_pre[kTaskCap + 1];
  __shared__ uint64_t s_src[kTaskCap];
  const uint64_t W = woff[nq];
  const uint64_t L = chunk_len(W);
  const uint64_t nch = (W + L - 1) / L;
  const uint32_t tid = threadIdx.x;
  for (uint64_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const uint32_t e = chunk_task[c];
    uint32_t q = e >> 7, i = e & 127u;
    uint64_t x = c * L;
    const uint64_t x1 = x + L < W ? x + L : W;
    uint32_t staged = ~0u, nt = 0;
    uint64_t wbase = 0, obase = 0;
    while (x < x1) {
      if (q != staged) {
        __syncthreads();  // every lane is done with the previous table
        nt = ntask[q];
        wbase = woff[q];
        obase = offsets ? offsets[q] : 0ull;
        const uint64_t qb = static_cast<uint64_t>(q) * kTaskCap;
        if (tid <= nt) s_pre[tid] = tpre[qb + tid];
        if (tid < nt) s_src[tid] = tsrc[qb + tid];
        __syncthreads();
        staged = q;
      }
      if (i >= nt) {
        // the next query with IDs: the last q' with woff[q'] <= x
        uint32_t lo = q + 1, hi = nq - 1;
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) >> 1;
          if (woff[mid] <= x) lo = mid;
          else hi = mid - 1;
        }
        if (lo > hi) break;
        q = lo;
        i = 0;
        continue;
      }
      const uint64_t a = wbase + s_pre[i], b = wbase + s_pre[i + 1];
      if (x < a || x >= b) break;  // not a range of this batch's plan
      const uint64_t hi = b < x1 ? b : x1;
      constexpr uint64_t kPass = static_cast<uint64_t>(kCopyBlock) * (kCopyBytes / sizeof(IS));
      const uint64_t bulk = ((hi - x) / kPass) * kPass;
      if (bulk) {
        copy_passes(ids + s_src[i] + (x - a), out + obase + (x - wbase), bulk, tid);
        x += bulk;
        if (x == b) i++;
        continue;
      }
      // a flat pass from x to the query's, the chunk's or the pass's end
      const uint64_t qend = wbase + s_pre[nt];
      uint64_t fend = x + kFlat;
      if (fend > qend) fend = qend;
      if (fend > x1) fend = x1;
      {
        constexpr int V = kFlat / kCopyBlock;
        IS v[V];
        uint32_t loc[V];
        uint32_t ii = i;
#pragma unroll
        for (int u = 0; u < V; u++) {
          const uint64_t pos = x + u * kCopyBlock + tid;
          loc[u] = static_cast<uint32_t>(pos - wbase);
          if (pos < fend) {
            while (s_pre[ii + 1] <= loc[u]) ii++;
            v[u] = ids[s_src[ii] + (loc[u] - s_pre[ii])];
          }
        }
#pragma unroll
        for (int u = 0; u < V; u++)
          if (x + u * kCopyBlock + tid < fend) out[obase + loc[u]] = v[u];
      }
      x = fend;
      while (i < nt && wbase + s_pre[i + 1] <= x) i++;
    }
  }
}

template <typename F>
Status lv_with_t(int dt, F&& f) {
  switch (dt) {
    case F32: return f(float{});
    case F64: return f(double{});
    case I32: return f(int32_t{});
    case I64: return f(int64_t{});
  }
  return Status{1, "Error: unsupported data type code " + std::to_string(dt)};
}
template <typename F>
Status lv_with_is(int idt, F&& f) {
  switch (idt) {
    case F32: case I32: case U32: return f(uint32_t{});
    case F64: case I64: case U64: return f(uint64_t{});
  }
  return Status{1, "Error: the level index needs a tree with IDs"};
}

uint32_t bit_length(uint64_t n) { return n ? 64u - static_cast<uint32_t>(__builtin_clzll(n)) : 0u; }

}  // namespace

size_t level_index_bytes(int dtype, int id_dtype, size_t n) {
  if (n == 0) return 0;
  const size_t per = dtype_size(dtype) + (id_dtype == F64 || id_dtype == I64 || id_dtype == U64 ? 8 : 4);
  return static_cast<size_t>(bit_length(n)) * level_stride(n) * per;
}

Status build_level_index(int dtype, int id_dtype, size_t n, const void* nodes_d, const void* ids_d, void* key_d,
                         void* id_d, hipStream_t s) {
  if (n == 0) return Status{};
  if (n >= (1ull << 31)) return Status{1, "Error: the level index holds trees of fewer than 2^31 points"};
  if (!ids_d) return Status{1, "Error: the level index needs a tree with IDs"};
  const uint32_t H = bit_length(n);
  const uint64_t M = n + 1;
  const size_t ns = level_stride(n);
  return lv_with_t(dtype, [&](auto tv) -> Status {
    using T = decltype(tv);
    using K = key_of<T>;
    return lv_with_is(id_dtype, [&](auto iv) -> Status {
      using IS = decltype(iv);
      K* key = static_cast<K*>(key_d);
      IS* id = static_cast<IS*>(id_d);
      const uint32_t n32 = static_cast<uint32_t>(n);
      const uint32_t ntiles = (n32 + kLvTile - 1) / kLvTile;
      K* kin = nullptr;
      uint32_t *sa = nullptr, *sb = nullptr;
      unsigned long long *agg = nullptr, *pre = nullptr;
      void* tmp = nullptr;
      size_t tmp_bytes = 0;
      hipError_t e = sizeof(K) == 4
                         ? radix_sort_u32(nullptr, tmp_bytes, nullptr, nullptr, nullptr, nullptr, n, s)
                         : radix_sort_u64(nullptr, tmp_bytes, nullptr, nullptr, nullptr, nullptr, n, s);
      if (e == hipSuccess) e = hipMalloc(&kin, n * sizeof(K));
      if (e == hipSuccess) e = hipMalloc(&sa, n * 4);
      if (e == hipSuccess) e = hipMalloc(&sb, n * 4);
      if (e == hipSuccess) e = hipMalloc(&agg, (ntiles + 1) * 8ull);
      if (e == hipSuccess) e = hipMalloc(&pre, (ntiles + 1) * 8ull);
      if (e == hipSuccess) e = hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 16));
      Status st = hip_status(e, "Error in allocating level index construction buffers");
      if (st.ok()) {
        const unsigned blocks = (n32 + kLvBlock - 1) / kLvBlock;
        const Node<T, IS>* nodes = static_cast<const Node<T, IS>*>(nodes_d);
        hipLaunchKernelGGL((k_lv_keys<T, IS>), dim3(blocks), dim3(kLvBlock), 0, s, nodes, n32, H, kin, sb);
        size_t tb = std::max<size_t>(tmp_bytes, 16);
        if constexpr (sizeof(K) == 4)
          e = radix_sort_u32(tmp, tb, kin, key, sb, sa, n, s);
        else
          e = radix_sort_u64(tmp, tb, kin, key, sb, sa, n, s);
        st = hip_status(e, "Error in sorting the level index");
      }
      if (st.ok()) {
        const unsigned blocks = (n32 + kLvBlock - 1) / kLvBlock;
        hipLaunchKernelGGL((k_lv_ids0<IS>), dim3(blocks), dim3(kLvBlock), 0, s, static_cast<const IS*>(ids_d), sa,
                           n32, id);
        for (uint32_t t = 0; t + 1 < H && st.ok(); t++) {
          const uint32_t nt = n32 - ((1u << t) - 1u);
          const uint32_t tiles = (nt + kLvTile - 1) / kLvTile;
          hipLaunchKernelGGL(k_lv_tile_counts, dim3(tiles), dim3(kLvBlock), 0, s, sa, nt, t, agg);
          st = exclusive_scan(agg, pre, tiles, s);
          if (!st.ok()) break;
          hipLaunchKernelGGL((k_lv_scatter<K, IS>), dim3(tiles), dim3(kLvBlock), 0, s, sa, sb,
                             key + static_cast<size_t>(t) * ns, key + static_cast<size_t>(t + 1) * ns,
                             id + static_cast<size_t>(t) * ns, id + static_cast<size_t>(t + 1) * ns, pre, nt, t, M);
          std::swap(sa, sb);
        }
        if (st.ok()) st = hip_status(hipGetLastError(), "Error in launching level index construction");
        if (st.ok()) st = hip_status(hipStreamSynchronize(s), "Error in building the level index");
      }
      if (kin) hipFree(kin);
      if (sa) hipFree(sa);
      if (sb) hipFree(sb);
      if (agg) hipFree(agg);
      if (pre) hipFree(pre);
      if (tmp) hipFree(tmp);
      return st;
    });
  });
}

Status level_scratch_reserve(LevelScratch& sc, size_t nq) {
  if (sc.cap >= nq && sc.chunk_task) return Status{};
  if (sc.external) return Status{1, "Error: the batch holds more queries than the search buffers"};
  if (sc.pinned) return Status{1, kPinnedScratchMsg};
  level_scratch_release(sc);
  const size_t cap = std::max<size_t>(nq, 64);
  hipError_t e = hipMalloc(&sc.tsrc, cap * kTaskCap * 8);
  if (e == hipSuccess) e = hipMalloc(&sc.tpre, cap * kTaskCap * 4);
  if (e == hipSuccess) e = hipMalloc(&sc.ntask, cap * 4);
  if (e == hipSuccess) e = hipMalloc(&sc.wq, cap * 8);
  if (e == hipSuccess) e = hipMalloc(&sc.woff, (cap + 1) * 8);
  if (e == hipSuccess) e = hipMalloc(&sc.chunk_task, kChunkCap * 4);
  if (e != hipSuccess) {
    level_scratch_release(sc);
    return hip_status(e, "Error in allocating level search buffers");
  }
  sc.cap = cap;
  return Status{};
}

void level_scratch_release(LevelScratch& sc) {
  if (sc.external) {
    sc = LevelScratch{};
    return;
  }
  if (sc.tsrc) hipFree(sc.tsrc);
  if (sc.tpre) hipFree(sc.tpre);
  if (sc.ntask) hipFree(sc.ntask);
  if (sc.wq) hipFree(sc.wq);
  if (sc.woff) hipFree(sc.woff);
  if (sc.chunk_task) hipFree(sc.chunk_task);
  sc = LevelScratch{};
}

Status level_search(const SearchArgs& a, const void* key_d, const void* id_d, LevelScratch& sc, hipStream_t s,
                    int plan) {
  if (a.nq == 0 || a.n == 0) return Status{};
  if (a.mode == 1 && a.ret_point) return Status{1, "Error: the level index returns IDs only"};
  if (plan == 2 && (a.mode != 1 || sc.cap < a.nq || !sc.tsrc))
    return Status{1, "Error: no kept ranges for this batch"};
  if (a.mode == 1 || plan == 1) {
    Status st = level_scratch_reserve(sc, a.nq);
    if (!st.ok()) return st;
  }
  return lv_with_t(a.dtype, [&](auto tv) -> Status {
    using T = decltype(tv);
    using K = key_of<T>;
    return lv_with_is(a.id_dtype, [&](auto iv) -> Status {
      using IS = decltype(iv);
      LvParams<T, IS> P{};
      P.nodes = static_cast<const Node<T, IS>*>(a.nodes);
      P.key = static_cast<const K*>(key_d);
      P.id = static_cast<const IS*>(id_d);
      P.n = a.n;
      P.ns = level_stride(a.n);
      P.M = static_cast<uint32_t>(a.n + 1);
      P.H = bit_length(a.n);
      P.queries = static_cast<const Query<T>*>(a.queries);
      if (a.single_query) {
        P.single = *static_cast<const Query<T>*>(a.single_query);
        P.use_single = 1;
      }
      P.nq = static_cast<uint32_t>(a.nq);
      P.counts = a.counts;
      P.offsets = a.offsets;
      P.caps = a.caps;
      P.cap_all = a.cap_all;
      P.out_total = a.out_total;
      P.tsrc = sc.tsrc;
      P.tpre = sc.tpre;
      P.ntask = sc.ntask;
      P.wq = sc.wq;
      P.woff = sc.woff;
      P.chunk_task = sc.chunk_task;
      if (a.mode == 0) {
        P.keep = plan == 1;
        hipLaunchKernelGGL((k_lv_plan<T, IS, false>), dim3(P.nq), dim3(kWave), 0, s, P);
        return hip_status(hipGetLastError(), "Error in launching level search");
      }
      if (plan == 2) {
        hipLaunchKernelGGL(k_lv_clip, dim3((P.nq + kLvBlock - 1) / kLvBlock), dim3(kLvBlock), 0, s, sc.tpre, sc.ntask,
                           sc.wq, a.counts, a.offsets, a.caps, a.cap_all, P.out_total, P.nq);
      } else {
        hipLaunchKernelGGL((k_lv_plan<T, IS, true>), dim3(P.nq), dim3(kWave), 0, s, P);
      }
      if (!(P.use_single && plan != 2)) {  // a single query's plan wrote these itself
        Status st = exclusive_scan(sc.wq, sc.woff, a.nq, s);
        if (!st.ok()) return st;
        const uint64_t mark_threads = static_cast<uint64_t>(P.nq) * kTaskCap;
        hipLaunchKernelGGL(k_lv_mark, dim3(static_cast<unsigned>((mark_threads + kLvBlock - 1) / kLvBlock)),
                           dim3(kLvBlock), 0, s, sc.woff, sc.tpre, sc.ntask, P.nq, sc.chunk_task);
      }
      int dev = 0, cus = 0;
      hipGetDevice(&dev);
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      // three workgroups per CU: measured on the 2^26 batch, 2-4 beat 8
      // (fewer waves, each with 256 B of loads in flight)
      hipLaunchKernelGGL((k_lv_copy<IS>), dim3(static_cast<unsigned>(cus * 3)), dim3(kCopyBlock), 0, s, P.id,
                         sc.tsrc, sc.tpre, sc.ntask, sc.woff, a.offsets, sc.chunk_task, P.nq,
                         static_cast<IS*>(a.out));
      return hip_status(hipGetLastError(), "Error in launching level search");
    });
  });
}

}  // namespace pst
