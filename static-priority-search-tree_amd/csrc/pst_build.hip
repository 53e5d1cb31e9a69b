// Tree construction for StaticPSTGPU on gfx950.
//
// Replaces the reference's construction (pst/static-pst-gpu.tu:19-437:
// index init, two thrust::sorts, then populateTree, a one-block kernel that
// builds nodes thread-per-node and relaunches itself through dynamic
// parallelism, pst/static-pst-gpu-populate-tree.tu:10-158) with rank arrays
// and level-synchronous passes:
//
//   1. key2 ranks: r2[i] = position of point i in the key2-descending order
//      (key2 = (dim2, id, index), the reference's compareDim2 order).
//   2. A = r2 listed in key1-ascending order (key1 = (dim1, id, index)).
//      From here on construction only compares 32-bit ranks.
//   3. Level t holds 2^t segments of A, one per slot, in slot order; their
//      sizes and offsets follow from n alone (pst_common.hpp). Per segment,
//      the node is the minimum rank (max key2); med is the dim1 of the
//      floor(m/2)-th remaining point; removing the node from A (a coalesced
//      compaction) yields level t+1, whose left/right segments are the
//      first floor(m/2) / the remaining points.
//   4. Once every segment fits in LDS (<= kFinishTile points), one
//      workgroup finishes its whole subtree in LDS.
// The node chosen, the median rule and the child split are those of the
// reference's constructNode (pst/static-pst-gpu.tu:706-806 and
// pst/static-pst-cpu-iter.tpp:488-598); the tree is bit-identical to the
// reference's CPU iterative tree on the same input.
#include <algorithm>
#include <vector>

#include "pst_internal.hpp"

namespace pst {

namespace {

constexpr int kLevelBlock = 256;
constexpr int kLevelElems = 32;
constexpr uint32_t kLevelTile = kLevelBlock * kLevelElems;  // 8192 points
constexpr int kFinishBlock = 512;      // threads per LDS subtree of the deep tree
constexpr uint32_t kFinishTile = 2048;  // max points per LDS subtree
constexpr int kForestBlock = 256;       // threads per forest tree

template <typename I>
struct IdStore {
  using type = typename std::conditional<sizeof(I) == 8, uint64_t, uint32_t>::type;
};
template <>
struct IdStore<NoId> {
  using type = NoId;
};

// ---------------------------------------------------------------------------
// Key kernels. Keys are order-preserving unsigned integers; for the
// descending key2 order the keys are complemented and the input visited
// back to front, so that a stable ascending sort also orders equal keys by
// descending index, as the reference's comparisonTiebreaker does for
// compareDim2(...) > 0.
// ---------------------------------------------------------------------------
// ORDER_CHECK: also sets flags[1] when some point's ID is above the next
// point's (IDs not in index order), and flags[0] for a NaN ID. With key1
// (32-bit keys only), the same read of point i also stores its ascending
// dim1 key at key1[i], the keys of the dim1 sort.
template <typename T, typename I, typename K, bool DIM2, bool DESC, bool WITH_ID, bool ORDER_CHECK = false>
__global__ void k_keys_single(const Point<T, I>* __restrict__ pts, uint32_t n, K* __restrict__ keys,
                              uint32_t* __restrict__ vals, const uint32_t* __restrict__ r2,
                              int* __restrict__ flags, uint32_t* __restrict__ key1 = nullptr) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t i = DESC ? n - 1 - j : j;
  const Point<T, I> p = pts[i];
  const T d = DIM2 ? p.dim2_val : p.dim1_val;
  bool bad = is_nan(d);
  K k;
  if constexpr (WITH_ID) {
    bad |= is_nan(p.id);
    k = (static_cast<K>(ord_key(d)) << 32) | static_cast<K>(ord_key(p.id));
  } else {
    k = static_cast<K>(ord_key(d));
  }
  if constexpr (ORDER_CHECK) {
    bad |= is_nan(p.id);
    if (i + 1 < n && ord_key(p.id) > ord_key(pts[i + 1].id)) flags[1] = 1;
  }
  if constexpr (sizeof(K) == 4) {
    if (key1) {
      bad |= is_nan(p.dim1_val);
      key1[i] = static_cast<uint32_t>(ord_key(p.dim1_val));
    }
  }
  if (bad) flags[0] = 1;
  keys[j] = DESC ? ~k : k;
  vals[j] = r2 ? r2[i] : i;
}

// Two-pass keys (when dim or id is 8 bytes): pass A orders by id.
template <typename T, typename I, typename K, bool DESC>
__global__ void k_keys_id(const Point<T, I>* __restrict__ pts, uint32_t n, K* __restrict__ keys,
                          uint32_t* __restrict__ vals, int* __restrict__ nan_flag) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t i = DESC ? n - 1 - j : j;
  const I id = pts[i].id;
  if (is_nan(id)) *nan_flag = 1;
  const K k = static_cast<K>(ord_key(id));
  keys[j] = DESC ? static_cast<K>(~k) : k;
  vals[j] = i;
}

// Pass B: orders by dim, visiting points in pass-A order.
template <typename T, typename I, typename K, bool DIM2, bool DESC>
__global__ void k_keys_dim_perm(const Point<T, I>* __restrict__ pts, uint32_t n,
                                const uint32_t* __restrict__ perm, K* __restrict__ keys,
                                uint32_t* __restrict__ vals, const uint32_t* __restrict__ r2,
                                int* __restrict__ nan_flag) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t i = perm[j];
  const T d = DIM2 ? pts[i].dim2_val : pts[i].dim1_val;
  if (is_nan(d)) *nan_flag = 1;
  const K k = static_cast<K>(ord_key(d));
  keys[j] = DESC ? static_cast<K>(~k) : k;
  vals[j] = r2 ? r2[i] : i;
}

__global__ void k_scatter_rank(const uint32_t* __restrict__ perm2, uint32_t n, uint32_t* __restrict__ r2) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) r2[perm2[p]] = p;
}

// Segment starts of levels 0..t0, slot-indexed (level t at offset 2^t - 1):
// the closed form of pst_common.hpp, clamped at 0.
__global__ void k_level_starts(uint64_t M, uint32_t nstart, uint32_t* __restrict__ starts) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nstart) return;
  const uint32_t t = depth_of(idx);
  const uint32_t j = idx + 1u - (1u << t);
  const int64_t v = level_start(M, t, j);
  starts[idx] = static_cast<uint32_t>(v > 0 ? v : 0);
}

// Per level t < t0 and tile b of kLevelTile points: the segment holding the
// tile's first point, i.e. the last j with starts[j] <= b * kLevelTile.
__global__ void k_tile_segments(const uint32_t* __restrict__ starts, uint32_t n, uint32_t t0, uint32_t stride,
                                uint32_t* __restrict__ tileseg) {
  const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= static_cast<uint64_t>(t0) * stride) return;
  const uint32_t t = static_cast<uint32_t>(g / stride), b = static_cast<uint32_t>(g % stride);
  const uint32_t n_level = n - ((1u << t) - 1u);
  const uint32_t tiles = (n_level + kLevelTile - 1) / kLevelTile;
  uint32_t seg = 0;
  if (b < tiles) {
    const uint32_t* lv = starts + (1u << t) - 1u;
    const uint32_t pos = b * kLevelTile;
    uint32_t lo = 0, hi = 1u << t;  // lv[lo] <= pos (lv[0] = 0); answer in [lo, hi)
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (lv[mid] <= pos) lo = mid;
      else hi = mid;
    }
    seg = lo;
  }
  tileseg[g] = seg;
}

__global__ void k_fill_u64(unsigned long long* p, size_t count, unsigned long long v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

// ---------------------------------------------------------------------------
// Level passes over A (all segments of level t are non-empty here).
// ---------------------------------------------------------------------------

// Segments met by one tile of kLevelTile points: every segment of a level
// walked by these passes holds at least kFinishTile points.
constexpr uint32_t kSegSlots = kLevelTile / kFinishTile + 2;

// The tile's elements, each thread's kLevelElems loaded before any is used
// (position tile + threadIdx.x + e * kLevelBlock).
__device__ inline void load_tile(const uint32_t* __restrict__ A, uint32_t n_level, uint32_t tile,
                                 uint32_t (&av)[kLevelElems]) {
  const uint32_t* p = A + tile + threadIdx.x;
  if (tile + kLevelTile <= n_level) {
#pragma unroll
    for (int e = 0; e < kLevelElems; e++) av[e] = p[e * kLevelBlock];
  } else {
#pragma unroll
    for (int e = 0; e < kLevelElems; e++)
      if (tile + threadIdx.x + e * kLevelBlock < n_level) av[e] = p[e * kLevelBlock];
  }
}

// The starts of the tile's segments seg0 + k, k <= kSegSlots (n_level past
// the level's last segment), into LDS. Call before a barrier.
__device__ inline void load_seg_starts(const uint32_t* __restrict__ starts, uint32_t seg0, uint32_t nseg,
                                       uint32_t n_level, uint32_t* sst) {
  if (threadIdx.x <= kSegSlots) {
    const uint32_t j = seg0 + threadIdx.x;
    sst[threadIdx.x] = j < nseg ? starts[j] : n_level;
  }
}

// Per tile and segment met: min over (rank << 32 | local position), i.e.
// the max-key2 point and where it sits in the segment, into tmin (per tile
// kSegSlots entries, combined per segment by k_tile_min_reduce).
__global__ __launch_bounds__(kLevelBlock) void k_level_argmin(
    const uint32_t* __restrict__ A, uint32_t n_level, uint32_t nseg,
    const uint32_t* __restrict__ starts, const uint32_t* __restrict__ tileseg,
    unsigned long long* __restrict__ tmin) {
  __shared__ unsigned long long smin[kSegSlots];
  __shared__ uint32_t sst[kSegSlots + 1];
  const uint32_t tile = blockIdx.x * kLevelTile;
  const uint32_t seg0 = tileseg[blockIdx.x];  // the segment holding the tile's first point
  uint32_t av[kLevelElems];
  load_tile(A, n_level, tile, av);
  if (threadIdx.x < kSegSlots) smin[threadIdx.x] = ~0ull;
  load_seg_starts(starts, seg0, nseg, n_level, sst);
  __syncthreads();
  const uint32_t first = tile + threadIdx.x;
  uint32_t k = 0;
  uint32_t st = sst[0], next = sst[1];
  unsigned long long best = ~0ull;
#pragma unroll
  for (int e = 0; e < kLevelElems; e++) {
    const uint32_t P = first + e * kLevelBlock;
    if (P < n_level) {
      while (P >= next) {
        if (best != ~0ull) atomicMin(&smin[k], best);
        best = ~0ull;
        k++;
        st = next;
        next = sst[k + 1];
      }
      const unsigned long long key = (static_cast<unsigned long long>(av[e]) << 32) | (P - st);
      best = key < best ? key : best;
    }
  }
  if (best != ~0ull) atomicMin(&smin[k], best);
  __syncthreads();
  if (threadIdx.x < kSegSlots) tmin[static_cast<size_t>(blockIdx.x) * kSegSlots + threadIdx.x] = smin[threadIdx.x];
}

// out[i] for i < gridDim.x (one workgroup each): the minimum of slot
// i - (tileseg[b] << sh) of the per-tile minima of every tile b that meets
// segment i >> sh of the level (slots: entries per tile). For the levels
// with few segments, where the tiles' global atomics would queue on a few
// addresses.
__global__ __launch_bounds__(kLevelBlock) void k_tile_min_reduce(
    const unsigned long long* __restrict__ tmin, uint32_t slots, const uint32_t* __restrict__ tileseg,
    const uint32_t* __restrict__ starts, uint32_t nseg, uint32_t n_level, uint32_t sh,
    unsigned long long* __restrict__ out) {
  __shared__ unsigned long long wbest[kLevelBlock / kWave];
  const uint32_t i = blockIdx.x;
  const uint32_t j = i >> sh;
  const uint32_t lo = starts[j];
  const uint32_t hi = j + 1 < nseg ? starts[j + 1] : n_level;
  unsigned long long best = ~0ull;
  if (hi > lo) {
    const uint32_t b0 = lo / kLevelTile, b1 = (hi - 1) / kLevelTile;
    for (uint32_t b = b0 + threadIdx.x; b <= b1; b += kLevelBlock) {
      const uint32_t slot = i - (tileseg[b] << sh);
      if (slot < slots) {
        const unsigned long long v = tmin[static_cast<size_t>(b) * slots + slot];
        best = v < best ? v : best;
      }
    }
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(best, off, kWave);
    best = o < best ? o : best;
  }
  if (lane_id() == 0) wbest[threadIdx.x / kWave] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = wbest[0];
    for (int w = 1; w < kLevelBlock / kWave; w++) m = wbest[w] < m ? wbest[w] : m;
    out[i] = m;
  }
}

template <typename T, typename IS>
__device__ inline void write_node(Node<T, IS>* nodes, uint64_t slot, const Point<T, IS>& p, T med) {
  Node<T, IS> r;
  r.dim1 = p.dim1_val;
  r.dim2 = p.dim2_val;
  r.med = med;
  if constexpr (has_id_v<IS>) r.id = p.id;
  nodes[slot] = r;
}

// Emits the level's nodes and compacts A into the next level: the point at
// local position x of segment j (x != q_j) moves to start_j - j + x', with
// x' = x - [x > q_j]. Unless level t + 1 is built in LDS, the same pass
// finds the next level's minima: x' < floor(m_j / 2) lands in child 2j at
// position x', the rest in child 2j + 1 at x' - floor(m_j / 2), and the
// minimum of (rank << 32 | child position) per child goes to nextmin.
// The tile's elements are loaded first and its segments' starts, minima and
// halves staged in LDS, so that the pass streams A at full occupancy.
template <typename T, typename IS>
__global__ __launch_bounds__(kLevelBlock) void k_level_emit(
    const uint32_t* __restrict__ A, uint32_t* __restrict__ Anext, uint32_t n_level, uint32_t t,
    uint64_t M, const uint32_t* __restrict__ starts, const uint32_t* __restrict__ tileseg,
    const unsigned long long* __restrict__ segmin, unsigned long long* __restrict__ nextmin,
    unsigned long long* __restrict__ tmin, const uint32_t* __restrict__ perm2, const Point<T, IS>* __restrict__ pts,
    Node<T, IS>* __restrict__ nodes) {
  constexpr uint32_t kChildSlots = 2 * kSegSlots;  // children per tile
  const uint32_t nseg = 1u << t;
  __shared__ unsigned long long cmin[kChildSlots];
  __shared__ unsigned long long smn[kSegSlots];
  __shared__ uint32_t sst[kSegSlots + 1];
  __shared__ uint32_t shalf[kSegSlots];
  const uint32_t tile = blockIdx.x * kLevelTile;
  const uint32_t seg0 = tileseg[blockIdx.x];  // the segment holding the tile's first point
  uint32_t av[kLevelElems];
  load_tile(A, n_level, tile, av);
  if (threadIdx.x < kChildSlots) cmin[threadIdx.x] = ~0ull;
  load_seg_starts(starts, seg0, nseg, n_level, sst);
  if (threadIdx.x < kSegSlots) {
    const uint32_t j = seg0 + threadIdx.x;
    smn[threadIdx.x] = j < nseg ? segmin[j] : ~0ull;
    shalf[threadIdx.x] = j < nseg ? static_cast<uint32_t>(level_size(M, t, j)) >> 1 : 0u;
  }
  __syncthreads();
  const uint32_t c0 = 2 * seg0;  // the tile's first child segment
  const uint32_t first = tile + threadIdx.x;
  uint32_t k = 0;
  uint32_t st = sst[0], next = sst[1];
  uint32_t q = static_cast<uint32_t>(smn[0]);
  uint32_t half = shalf[0];
  unsigned long long best = ~0ull;  // running minimum of child bc
  uint32_t bc = 0;
#pragma unroll
  for (int e = 0; e < kLevelElems; e++) {
    const uint32_t P = first + e * kLevelBlock;
    if (P < n_level) {
      while (P >= next) {
        k++;
        st = next;
        next = sst[k + 1];
        q = static_cast<uint32_t>(smn[k]);
        half = shalf[k];
      }
      const uint32_t x = P - st;
      const uint32_t a = av[e];
      const uint32_t j = seg0 + k;
      if (x == q) {
        const uint32_t L = half;
        const Point<T, IS> p = pts[perm2[a]];
        T med = p.dim1_val;
        if (L >= 1) {
          const uint32_t pm = q < L ? L : L - 1;
          med = pts[perm2[A[st + pm]]].dim1_val;
        }
        write_node(nodes, (1ull << t) - 1 + j, p, med);
      } else {
        const uint32_t xr = x - (x > q ? 1u : 0u);
        Anext[st - j + xr] = a;
        if (nextmin) {
          const bool right = xr >= half;
          const uint32_t c = 2 * k + (right ? 1u : 0u);
          const unsigned long long key = (static_cast<unsigned long long>(a) << 32) | (right ? xr - half : xr);
          if (c != bc) {
            if (best != ~0ull) atomicMin(&cmin[bc], best);
            best = ~0ull;
            bc = c;
          }
          best = key < best ? key : best;
        }
      }
    }
  }
  if (!nextmin) return;
  if (best != ~0ull) atomicMin(&cmin[bc], best);
  __syncthreads();
  if (tmin) {  // per-tile minima, reduced by k_tile_min_reduce
    if (threadIdx.x < kChildSlots) tmin[static_cast<size_t>(blockIdx.x) * kChildSlots + threadIdx.x] = cmin[threadIdx.x];
  } else if (threadIdx.x < kChildSlots && cmin[threadIdx.x] != ~0ull) {
    atomicMin(&nextmin[c0 + threadIdx.x], cmin[threadIdx.x]);
  }
}

// ---------------------------------------------------------------------------
// Whole subtree in LDS: A-segment of m <= kFinishTile points in bufA.
// ---------------------------------------------------------------------------
template <typename T, typename IS, typename PermF>
__device__ void lds_build(uint32_t* bufA, uint32_t* bufB, unsigned long long* sres,
                          unsigned long long* wred, uint32_t m, uint64_t root_slot,
                          const Point<T, IS>* __restrict__ pts, PermF perm, Node<T, IS>* __restrict__ nodes) {
  const uint32_t tid = threadIdx.x;
  const uint32_t nthr = blockDim.x;
  const uint64_t Mp = static_cast<uint64_t>(m) + 1;
  for (uint32_t u = 0; (Mp >> u) != 0; u++) {
    const uint32_t nsub = 1u << u;
    if (nsub <= nthr) {
      const uint32_t G = nthr / nsub;
      const uint32_t k = tid / G, g = tid % G;
      const int64_t sz = level_size(Mp, u, k);
      const uint32_t st = static_cast<uint32_t>(sz > 0 ? level_start(Mp, u, k) : 0);
      unsigned long long best = ~0ull;
      for (int64_t x = g; x < sz; x += G) {
        const unsigned long long key = (static_cast<unsigned long long>(bufA[st + x]) << 32) |
                                       static_cast<unsigned long long>(x);
        best = key < best ? key : best;
      }
      const uint32_t wg = G < kWave ? G : kWave;
      for (uint32_t off = wg >> 1; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(best, off, kWave);
        best = o < best ? o : best;
      }
      if (G > kWave) {  // uniform across the block
        if ((tid & (kWave - 1)) == 0) wred[tid / kWave] = best;
        __syncthreads();
        if (g == 0) {
          const uint32_t w0 = tid / kWave, nw = G / kWave;
          for (uint32_t w = 0; w < nw; w++) best = wred[w0 + w] < best ? wred[w0 + w] : best;
        }
      }
      if (g == 0) sres[k] = best;
      __syncthreads();
      if (sz > 0) {
        const unsigned long long mn = sres[k];
        const uint32_t q = static_cast<uint32_t>(mn);
        const uint32_t L = static_cast<uint32_t>(sz) >> 1;
        if (g == 0) {
          const Point<T, IS> p = pts[perm(static_cast<uint32_t>(mn >> 32))];
          T med = p.dim1_val;
          if (L >= 1) med = pts[perm(bufA[st + (q < L ? L : L - 1)])].dim1_val;
          write_node(nodes, (root_slot + 1) * nsub - 1 + k, p, med);
        }
        const uint32_t dst0 = static_cast<uint32_t>(sz > 1 ? level_start(Mp, u + 1, 2 * k) : 0);
        for (int64_t x = g; x < sz; x += G) {
          if (static_cast<uint32_t>(x) == q) continue;
          bufB[dst0 + x - (static_cast<uint32_t>(x) > q ? 1 : 0)] = bufA[st + x];
        }
      }
    } else {
      for (uint32_t k = tid; k < nsub; k += nthr) {
        const int64_t sz = level_size(Mp, u, k);
        if (sz <= 0) continue;
        const uint32_t st = static_cast<uint32_t>(level_start(Mp, u, k));
        unsigned long long best = ~0ull;
        for (int64_t x = 0; x < sz; x++) {
          const unsigned long long key = (static_cast<unsigned long long>(bufA[st + x]) << 32) |
                                         static_cast<unsigned long long>(x);
          best = key < best ? key : best;
        }
        const uint32_t q = static_cast<uint32_t>(best);
        const uint32_t L = static_cast<uint32_t>(sz) >> 1;
        const Point<T, IS> p = pts[perm(static_cast<uint32_t>(best >> 32))];
        T med = p.dim1_val;
        if (L >= 1) med = pts[perm(bufA[st + (q < L ? L : L - 1)])].dim1_val;
        write_node(nodes, (root_slot + 1) * nsub - 1 + k, p, med);
        if (sz > 1) {
          const uint32_t dst0 = static_cast<uint32_t>(level_start(Mp, u + 1, 2 * k));
          for (int64_t x = 0; x < sz; x++) {
            if (static_cast<uint32_t>(x) == q) continue;
            bufB[dst0 + x - (static_cast<uint32_t>(x) > q ? 1 : 0)] = bufA[st + x];
          }
        }
      }
    }
    __syncthreads();
    uint32_t* tmp = bufA;
    bufA = bufB;
    bufB = tmp;
  }
}

// Orders a wave's LDS accesses around it (the compiler keeps them on their
// side; the wave's LDS instructions execute in order).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Threads of the subtree passes: from level kWaveLevel of a subtree on, the
// segment of position k and every segment below it belong to wave
// k >> (u - kWaveLevel), so those levels need no workgroup barrier.
constexpr uint32_t kFinishLog = 9;
static_assert((1u << kFinishLog) == kFinishBlock, "block size is a power of two");
constexpr uint32_t kWaveLevel = kFinishLog - 6;
static_assert(kWave == 64, "64-lane waves");

// The subtree build of the deep tree's last levels. Every point's rank is
// read once up front, all reads in flight together; the levels then work
// on 16-bit local positions in LDS: the node of a
// segment is the position of least rank, and the level only notes, per
// point, the subtree slot it fills and its med point. The node records are
// read from the input and written after the last level, all reads in
// flight together, so that no level waits on global memory. The levels
// above kWaveLevel share segments between waves and end in a workgroup
// barrier; from it on each wave works on its own part of the subtree.
template <typename T, typename IS>
__device__ void lds_build_local(uint16_t* bufA, uint16_t* bufB, const uint32_t* rank, uint16_t* slot_of,
                                uint16_t* med_of, uint16_t* sst, unsigned long long* sres, unsigned long long* wred,
                                uint32_t m) {
  const uint32_t tid = threadIdx.x;
  constexpr uint32_t nthr = kFinishBlock;  // the launch's block size
  const uint64_t Mp = static_cast<uint64_t>(m) + 1;
  // node of segment k of level u: point bufA[st + q]; its subtree slot
  // h = 2^u - 1 + k; med: the point at position L or L - 1 (L = floor(sz /
  // 2)), or the node itself when L = 0. Segment starts are kept per slot
  // (sst[h], sst[0] = 0 set by the caller): with the node removed, the
  // children start at st and st + L of the next level's array. Children stay
  // within their parent's positions (the next array has gaps where nodes
  // left), so that a wave's part of both arrays never moves.
  auto note = [&](uint32_t u, uint32_t k, uint32_t st, uint32_t sz, uint32_t q) {
    const uint32_t L = sz >> 1;
    const uint16_t x = bufA[st + q];
    const uint32_t h = (1u << u) - 1 + k;
    slot_of[x] = static_cast<uint16_t>(h);
    med_of[x] = L >= 1 ? bufA[st + (q < L ? L : L - 1)] : x;
    if (sz > 1) {
      sst[2 * h + 1] = static_cast<uint16_t>(st);
      sst[2 * h + 2] = static_cast<uint16_t>(st + L);
    }
  };
  // level sync: a workgroup barrier while segments span waves, else the
  // wave's own
  auto level_sync = [&](uint32_t u) {
    if (u < kWaveLevel) __syncthreads();
    else wave_sync();
  };
  for (uint32_t u = 0; (Mp >> u) != 0; u++) {
    const uint32_t nsub = 1u << u;
    if (nsub <= nthr) {
      // segment k on threads [k G, (k + 1) G)
      const uint32_t G = nthr / nsub;
      const uint32_t k = tid / G, g = tid % G;
      const int64_t sz = level_size(Mp, u, k);
      const uint32_t st = sz > 0 ? sst[(1u << u) - 1 + k] : 0u;
      unsigned long long best = ~0ull;
      for (int64_t x = g; x < sz; x += G) {
        const unsigned long long key =
            (static_cast<unsigned long long>(rank[bufA[st + x]]) << 32) | static_cast<unsigned long long>(x);
        best = key < best ? key : best;
      }
      const uint32_t wg = G < kWave ? G : kWave;
      for (uint32_t off = wg >> 1; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(best, off, kWave);
        best = o < best ? o : best;
      }
      if (G > kWave) {  // uniform across the block
        if ((tid & (kWave - 1)) == 0) wred[tid / kWave] = best;
        __syncthreads();
        if (g == 0) {
          const uint32_t w0 = tid / kWave, nw = G / kWave;
          for (uint32_t w = 0; w < nw; w++) best = wred[w0 + w] < best ? wred[w0 + w] : best;
        }
      }
      // the segment's result sits at its first thread's entry, which no
      // other wave uses at any level
      if (g == 0) sres[k * G] = best;
      level_sync(u);
      if (sz > 0) {
        const uint32_t q = static_cast<uint32_t>(sres[k * G]);
        if (g == 0) note(u, k, st, static_cast<uint32_t>(sz), q);
        const uint32_t dst0 = st;
        for (int64_t x = g; x < sz; x += G) {
          if (static_cast<uint32_t>(x) == q) continue;
          bufB[dst0 + x - (static_cast<uint32_t>(x) > q ? 1 : 0)] = bufA[st + x];
        }
      }
    } else {
      // thread t takes the 2^(u - kFinishLog) segments below its segment of
      // level kFinishLog
      const uint32_t per = nsub >> kFinishLog;
      for (uint32_t k = tid * per; k < (tid + 1) * per; k++) {
        const int64_t sz = level_size(Mp, u, k);
        if (sz <= 0) continue;
        const uint32_t st = sst[(1u << u) - 1 + k];
        unsigned long long best = ~0ull;
        for (int64_t x = 0; x < sz; x++) {
          const unsigned long long key =
              (static_cast<unsigned long long>(rank[bufA[st + x]]) << 32) | static_cast<unsigned long long>(x);
          best = key < best ? key : best;
        }
        const uint32_t q = static_cast<uint32_t>(best);
        note(u, k, st, static_cast<uint32_t>(sz), q);
        if (sz > 1) {
          const uint32_t dst0 = st;
          for (int64_t x = 0; x < sz; x++) {
            if (static_cast<uint32_t>(x) == q) continue;
            bufB[dst0 + x - (static_cast<uint32_t>(x) > q ? 1 : 0)] = bufA[st + x];
          }
        }
      }
    }
    level_sync(u);
    uint16_t* tmp = bufA;
    bufA = bufB;
    bufB = tmp;
  }
  __syncthreads();
}

// One workgroup per level-t0 segment: its whole subtree in LDS, then every
// node record; with `submin` also the subtree minima of dim2 and the IDs of
// the subtree's slots (the bottom levels of what k_submin_level computes).
template <typename T, typename IS>
__global__ __launch_bounds__(kFinishBlock) void k_finish(
    const uint32_t* __restrict__ A, uint32_t t0, uint64_t M, const uint32_t* __restrict__ starts,
    const uint32_t* __restrict__ perm2, const Point<T, IS>* __restrict__ pts, Node<T, IS>* __restrict__ nodes,
    T* __restrict__ submin, IS* __restrict__ ids) {
  union FinishLds {
    struct {
      uint16_t bufA[kFinishTile];
      uint16_t bufB[kFinishTile];
      uint16_t sst[2 * kFinishTile];  // segment start per subtree slot
    } b;
    T d2[2 * kFinishTile];  // after the build: dim2 per subtree slot, then its subtree minimum
  };
  __shared__ FinishLds U;
  union RankLds {
    uint32_t rank[kFinishTile];
    T d1[kFinishTile];  // after the lookups: dim1 of every point, for the meds
  };
  __shared__ RankLds R;
  uint32_t* rank = R.rank;
  __shared__ uint16_t slot_of[kFinishTile];  // the subtree slot (heap order) each point fills
  __shared__ uint16_t med_of[kFinishTile];   // and the point whose dim1 is its med
  __shared__ unsigned long long sres[kFinishBlock];  // a segment's result, at its first thread's entry
  __shared__ unsigned long long wred[kFinishBlock / kWave];
  const uint32_t j = blockIdx.x;
  const int64_t m = level_size(M, t0, j);
  if (m <= 0) return;
  const uint32_t st = starts[j];
  const uint64_t root = (1ull << t0) - 1 + j;
  constexpr int kPer = kFinishTile / kFinishBlock;
#pragma unroll
  for (int e = 0; e < kPer; e++) {
    const uint32_t x = threadIdx.x + e * kFinishBlock;
    if (x < m) {
      rank[x] = A[st + x];
      U.b.bufA[x] = static_cast<uint16_t>(x);
    }
  }
  if (threadIdx.x == 0) U.b.sst[0] = 0;
  __syncthreads();
  lds_build_local<T, IS>(U.b.bufA, U.b.bufB, rank, slot_of, med_of, U.b.sst, sres, wred, static_cast<uint32_t>(m));
  // every point's node record (rank -> input index -> record; all loads in
  // flight together): subtree slot h (level u, position k) is slot
  // ((root + 1) << u) - 1 + k of the tree. A med is the dim1 of a point of
  // the same subtree, exchanged through LDS.
  uint32_t ix[kPer];
#pragma unroll
  for (int e = 0; e < kPer; e++) {
    const uint32_t x = threadIdx.x + e * kFinishBlock;
    if (x < m) ix[e] = perm2[rank[x]];
  }
  Point<T, IS> p[kPer];
#pragma unroll
  for (int e = 0; e < kPer; e++) {
    const uint32_t x = threadIdx.x + e * kFinishBlock;
    if (x < m) p[e] = pts[ix[e]];
  }
  __syncthreads();  // the build's LDS arrays and the ranks are free from here
#pragma unroll
  for (int e = 0; e < kPer; e++) {
    const uint32_t x = threadIdx.x + e * kFinishBlock;
    if (x < m) R.d1[x] = p[e].dim1_val;
  }
  __syncthreads();
  T med[kPer];
#pragma unroll
  for (int e = 0; e < kPer; e++) {
    const uint32_t x = threadIdx.x + e * kFinishBlock;
    if (x < m) med[e] = R.d1[med_of[x]];
  }
  const uint32_t mdepth = 32u - static_cast<uint32_t>(__builtin_clz(static_cast<uint32_t>(m)));
#pragma unroll
  for (int e = 0; e < kPer; e++) {
    const uint32_t x = threadIdx.x + e * kFinishBlock;
    if (x < m) {
      const uint32_t h = slot_of[x] + 1u;
      if ((h >> mdepth) != 0) continue;  // never: the subtree has mdepth levels
      const uint32_t u = 31u - static_cast<uint32_t>(__builtin_clz(h));
      const uint64_t slot = ((root + 1) << u) - 1 + (h - (1u << u));
      write_node(nodes, slot, p[e], med[e]);
      if (submin) {
        U.d2[h - 1] = p[e].dim2_val;
        if constexpr (has_id_v<IS>) {
          if (ids) ids[slot] = p[e].id;
        }
      }
    }
  }
  if (!submin) return;
  // subtree minima, bottom level first: min(dim2(s), minima of the children
  // that exist); a slot of size ms has a left child if ms >= 2, a right one
  // if ms >= 3. Levels from kWaveLevel down: wave w takes the slots below
  // slot w of level kWaveLevel, with no workgroup barrier between levels;
  // the levels above them are one thread's.
  __syncthreads();
  const uint64_t Mp = static_cast<uint64_t>(m) + 1;
  const int depth = 32 - __builtin_clz(static_cast<uint32_t>(m));
  auto min_slot = [&](uint32_t u, uint32_t k) {
    const int64_t ms = level_size(Mp, u, k);
    if (ms < 1) return;
    const uint32_t sl = (1u << u) - 1 + k;
    T v = U.d2[sl];
    if (ms >= 2) {
      const T c = U.d2[2 * sl + 1];
      v = c < v ? c : v;
    }
    if (ms >= 3) {
      const T c = U.d2[2 * sl + 2];
      v = c < v ? c : v;
    }
    U.d2[sl] = v;
    submin[((root + 1) << u) - 1 + k] = v;
  };
  const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (int u = depth - 1; u >= static_cast<int>(kWaveLevel); u--) {
    const uint32_t per = 1u << (u - kWaveLevel);
    for (uint32_t i = lane; i < per; i += kWave) min_slot(static_cast<uint32_t>(u), wave * per + i);
    wave_sync();
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int top = depth - 1 < static_cast<int>(kWaveLevel) - 1 ? depth - 1 : static_cast<int>(kWaveLevel) - 1;
    for (int u = top; u >= 0; u--)
      for (uint32_t k = 0; k < (1u << u); k++) min_slot(static_cast<uint32_t>(u), k);
  }
}

// ---------------------------------------------------------------------------
// Subtree minimum of dim2, one level at a time from the bottom: submin(s) =
// min(dim2(s), submin(2s+1), submin(2s+2)) over the children that exist.
// A walk that meets a REPORT_ABOVE subtree with min_dim2 <= submin knows
// that every node in it is reported. Level t of every tree of a forest is
// done by one launch (ntrees x 2^t positions). The same pass copies each
// slot's ID into a dense ID array (null: none), which whole-subtree copies
// read instead of the full records.
// ---------------------------------------------------------------------------
template <typename T, typename IS>
__global__ void k_submin_level(const Node<T, IS>* __restrict__ nodes, T* __restrict__ submin, IS* __restrict__ ids,
                               uint32_t t, uint32_t shift, uint32_t ntrees, uint32_t full_M, uint32_t last_M) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t per_tree = 1ull << t;
  if (k >= per_tree * ntrees) return;
  const uint32_t tree = static_cast<uint32_t>(k >> t);
  const uint32_t local = static_cast<uint32_t>((1ull << t) - 1 + (k & (per_tree - 1)));
  const uint32_t M = tree + 1 < ntrees ? full_M : last_M;
  const int32_t m = subtree_size32(M, local);
  if (m < 1) return;
  const uint64_t base = static_cast<uint64_t>(tree) << shift;
  const Node<T, IS> nd = nodes[base + local];
  if constexpr (has_id_v<IS>) {
    if (ids) ids[base + local] = nd.id;
  }
  T v = nd.dim2;
  if (m >= 2) {
    const T c = submin[base + 2 * local + 1];
    v = c < v ? c : v;
  }
  if (m >= 3) {
    const T c = submin[base + 2 * local + 2];
    v = c < v ? c : v;
  }
  submin[base + local] = v;
}

template <typename T, typename IS>
Status submin_typed(const void* nodes, void* submin, void* ids, uint32_t levels, uint32_t shift, uint32_t ntrees,
                    uint32_t full_M, uint32_t last_M, hipStream_t s) {
  for (int t = static_cast<int>(levels) - 1; t >= 0; t--) {
    const uint64_t work = (1ull << t) * ntrees;
    k_submin_level<T, IS><<<static_cast<unsigned>((work + 255) / 256), 256, 0, s>>>(
        static_cast<const Node<T, IS>*>(nodes), static_cast<T*>(submin), static_cast<IS*>(ids),
        static_cast<uint32_t>(t), shift, ntrees, full_M, last_M);
  }
  return hip_status(hipGetLastError(), "Error in launching subtree minima");
}

// ---------------------------------------------------------------------------
// Forest of shallow trees (StaticPSTGPUArr). Tree c holds the input points
// [c*St, min(n, (c+1)*St)), the reference's consecutive chunks
// (pst/static-pst-gpu-arr-populate-trees.tu:25-40, pst/static-pst-gpu-arr.tu:
// 1037-1060), and is the tree the deep construction would build for that
// chunk alone. One workgroup per tree: the chunk's key1 and key2 orders come
// from bitonic sorts in LDS, then lds_build fills the tree level by level.
// Tree c's records start at c << L, with St = 2^L - 1.
// ---------------------------------------------------------------------------
template <typename Before>
__device__ void bitonic_sort(uint16_t* idx, uint32_t N, Before before) {
  for (uint32_t k = 2; k <= N; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t t = threadIdx.x; t < N / 2; t += blockDim.x) {
        const uint32_t i = (t / j) * 2 * j + (t % j);
        const uint32_t l = i + j;
        const uint16_t a = idx[i], b = idx[l];
        const bool up = (i & k) == 0;
        if (up ? before(b, a) : before(a, b)) {
          idx[i] = b;
          idx[l] = a;
        }
      }
      __syncthreads();
    }
  }
}

struct LocalPerm {
  const uint16_t* ord2;
  uint64_t first;
  __device__ uint32_t operator()(uint32_t r) const { return static_cast<uint32_t>(first + ord2[r]); }
};

// Dynamic LDS for a tree of up to N (a power of two) points.
__host__ __device__ constexpr size_t forest_lds_bytes(uint32_t N) {
  return static_cast<size_t>(N) * (3 * 8 + 2 * 2 + 2 * 4) + kForestBlock * 8 + (kForestBlock / kWave) * 8;
}

template <typename T, typename I>
__global__ __launch_bounds__(kForestBlock) void k_forest_build(const Point<T, I>* __restrict__ pts, uint64_t n,
                                                               uint32_t St, uint32_t L, uint32_t N,
                                                               Node<T, typename IdStore<I>::type>* __restrict__ nodes,
                                                               int* __restrict__ nan_flag) {
  using IS = typename IdStore<I>::type;
  extern __shared__ __align__(16) unsigned char lds[];
  uint64_t* k1 = reinterpret_cast<uint64_t*>(lds);
  uint64_t* k2 = k1 + N;
  uint64_t* kid = k2 + N;
  unsigned long long* sres = reinterpret_cast<unsigned long long*>(kid + N);
  unsigned long long* wred = sres + kForestBlock;
  uint32_t* bufA = reinterpret_cast<uint32_t*>(wred + kForestBlock / kWave);
  uint32_t* bufB = bufA + N;
  uint16_t* o1 = reinterpret_cast<uint16_t*>(bufB + N);
  uint16_t* o2 = o1 + N;

  const uint64_t c = blockIdx.x;
  const uint64_t first = c * St;
  const uint32_t m = static_cast<uint32_t>(n - first < St ? n - first : St);
  for (uint32_t x = threadIdx.x; x < N; x += blockDim.x) {
    o1[x] = static_cast<uint16_t>(x);
    o2[x] = static_cast<uint16_t>(x);
    if (x < m) {
      const Point<T, I> p = pts[first + x];
      bool bad = is_nan(p.dim1_val) || is_nan(p.dim2_val);
      k1[x] = static_cast<uint64_t>(ord_key(p.dim1_val));
      k2[x] = static_cast<uint64_t>(ord_key(p.dim2_val));
      if constexpr (has_id_v<I>) {
        bad |= is_nan(p.id);
        kid[x] = static_cast<uint64_t>(ord_key(p.id));
      } else {
        kid[x] = 0;
      }
      if (bad) *nan_flag = 1;
    }
  }
  __syncthreads();
  // key1 = (dim1, id, index) ascending; key2 = (dim2, id, index) descending
  // (rank 0 is the node point), the reference's compareDim1/compareDim2 with
  // comparisonTiebreaker (helpers/point-struct.h:54-92, 146-188)
  bitonic_sort(o1, N, [&](uint16_t a, uint16_t b) {
    if (a >= m) return false;
    if (b >= m) return true;
    if (k1[a] != k1[b]) return k1[a] < k1[b];
    if (kid[a] != kid[b]) return kid[a] < kid[b];
    return a < b;
  });
  bitonic_sort(o2, N, [&](uint16_t a, uint16_t b) {
    if (a >= m) return false;
    if (b >= m) return true;
    if (k2[a] != k2[b]) return k2[a] > k2[b];
    if (kid[a] != kid[b]) return kid[a] > kid[b];
    return a > b;
  });
  for (uint32_t r = threadIdx.x; r < m; r += blockDim.x) bufB[o2[r]] = r;
  __syncthreads();
  for (uint32_t x = threadIdx.x; x < m; x += blockDim.x) bufA[x] = bufB[o1[x]];
  __syncthreads();
  lds_build<T, IS>(bufA, bufB, sres, wred, m, 0, reinterpret_cast<const Point<T, IS>*>(pts), LocalPerm{o2, first},
                   nodes + (c << L));
}

// Forest export: per tree the reference's layout for its own slot count,
// trees concatenated (tree c at position c*St of every array).
template <typename T, typename IS>
__global__ void k_forest_export(const Node<T, IS>* __restrict__ nodes, uint64_t n, uint64_t St, uint32_t L,
                                uint64_t ntrees, uint64_t total, unsigned char* __restrict__ blob, size_t o1,
                                size_t o2, size_t om, size_t oi, size_t ob) {
  const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= total) return;
  uint64_t c = p / St;
  if (c >= ntrees) c = ntrees - 1;
  const uint64_t s = p - c * St;
  const uint64_t mc = (c + 1 < ntrees) ? St : n - c * St;
  const int64_t m = subtree_size(mc + 1, s);
  T d1 = T(0), d2 = T(0), md = T(0);
  unsigned char bc = 0;
  const Node<T, IS>* nd = nodes + (c << L) + s;
  if (m >= 1) {
    d1 = nd->dim1;
    md = nd->med;
    d2 = nd->dim2;
    bc = static_cast<unsigned char>((m >= 2 ? 2 : 0) | (m >= 3 ? 1 : 0));
  }
  reinterpret_cast<T*>(blob + o1)[p] = d1;
  reinterpret_cast<T*>(blob + o2)[p] = d2;
  reinterpret_cast<T*>(blob + om)[p] = md;
  if constexpr (has_id_v<IS>) {
    IS id = 0;
    if (m >= 1) id = nd->id;
    reinterpret_cast<IS*>(blob + oi)[p] = id;
  }
  blob[ob + p] = bc;
}

// Reference-layout export: dim1[S] dim2[S] med[S] ids[S] bitcodes[S].
template <typename T, typename IS>
__global__ void k_export(const Node<T, IS>* __restrict__ nodes, uint64_t S,
                         uint64_t M, unsigned char* __restrict__ blob, size_t o1, size_t o2, size_t om,
                         size_t oi, size_t ob) {
  const uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const int64_t m = subtree_size(M, s);
  T d1 = T(0), d2 = T(0), md = T(0);
  unsigned char bc = 0;
  if (m >= 1) {
    d1 = nodes[s].dim1;
    md = nodes[s].med;
    d2 = nodes[s].dim2;
    bc = static_cast<unsigned char>((m >= 2 ? 2 : 0) | (m >= 3 ? 1 : 0));
  }
  reinterpret_cast<T*>(blob + o1)[s] = d1;
  reinterpret_cast<T*>(blob + o2)[s] = d2;
  reinterpret_cast<T*>(blob + om)[s] = md;
  if constexpr (has_id_v<IS>) {
    IS id = 0;
    if (m >= 1) id = nodes[s].id;
    reinterpret_cast<IS*>(blob + oi)[s] = id;
  }
  blob[ob + s] = bc;
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() { release(); }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 1); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  template <typename U>
  U* as() const {
    return static_cast<U*>(p);
  }
};

inline unsigned grid_for(size_t n, unsigned block) { return static_cast<unsigned>((n + block - 1) / block); }

// Construction buffers of n points, carved from one allocation: two key
// arrays, four 32-bit arrays (sort values, the key2 order, the level arrays
// and the key2 ranks), flags, radix sort storage, level tables.
struct BuildLayout {
  size_t key0, key1, v0, v1, v2, r2, flags, tmp, starts, segmin, tileseg, tmin;
  size_t total_a, total_b, total;  // part A, part B, both
  size_t tile_stride;  // tileseg entries per level
  size_t tmp_bytes;
  uint32_t t0;
};

inline size_t al256(size_t b) { return (b + 255) & ~size_t{255}; }

// Levels walked by global passes: every segment of level t0 fits in LDS.
inline uint32_t finish_level(uint64_t M) {
  uint32_t t0 = 0;
  while (true) {
    const uint64_t base = M >> t0;
    const uint64_t r = M & ((1ull << t0) - 1ull);
    const uint64_t mx = base - 1 + (r ? 1 : 0);
    if (mx <= kFinishTile) break;
    t0++;
  }
  return t0;
}

template <typename K>
BuildLayout build_layout(size_t n) {
  BuildLayout L{};
  size_t tb = 0;
  if constexpr (sizeof(K) == 4) radix_sort_u32(nullptr, tb, nullptr, nullptr, nullptr, nullptr, n, nullptr);
  else radix_sort_u64(nullptr, tb, nullptr, nullptr, nullptr, nullptr, n, nullptr);
  L.tmp_bytes = tb;
  L.t0 = finish_level(static_cast<uint64_t>(n) + 1);
  const size_t nstart = (1ull << (L.t0 + 1)) - 1;
  const size_t nmin = L.t0 ? (1ull << L.t0) - 1 : 1;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += al256(bytes ? bytes : 1);
    return at;
  };
  // part A: what the first key kernel writes
  L.key0 = take(n * sizeof(K));
  L.v0 = take(n * 4);
  L.v2 = take(n * 4);
  L.flags = take(16);
  L.total_a = o;
  // part B (offsets from its own start): the rest
  o = 0;
  L.key1 = take(n * sizeof(K));
  L.v1 = take(n * 4);
  L.r2 = take(n * 4);
  L.tmp = take(tb);
  L.starts = take(nstart * 4);
  L.segmin = take(nmin * 8);
  L.tile_stride = (n + kLevelTile - 1) / kLevelTile;
  L.tileseg = take(static_cast<size_t>(L.t0) * L.tile_stride * 4);
  L.tmin = take(L.tile_stride * 2 * kSegSlots * 8);
  L.total_b = o;
  L.total = L.total_a + L.total_b;
  return L;
}

// Sort keys wide enough for a dim or an ID of the point type.
template <typename T, typename I>
using BuildKey = typename std::conditional<(sizeof(T) == 8 || (has_id_v<I> && sizeof(I) == 8)), uint64_t,
                                           uint32_t>::type;

template <typename T, typename I>
Status build_typed(const void* pts_v, size_t n64, BuildOut& out, hipStream_t s, BuildTimes* times) {
  using IS = typename IdStore<I>::type;
  using K = BuildKey<T, I>;
  Status st;
  if (n64 == 0) return st;
  if (n64 >= (1ull << 32) - 1) {
    st.code = 1;
    st.msg = "Error: StaticPSTGPU supports fewer than 2^32 - 1 points";
    return st;
  }
  const uint32_t n = static_cast<uint32_t>(n64);
  const uint64_t M = n64 + 1;
  const auto* pts = static_cast<const Point<T, I>*>(pts_v);
  const auto* ptsS = static_cast<const Point<T, IS>*>(pts_v);
  constexpr bool kHasId = has_id_v<I>;

  hipEvent_t ev[5] = {};
  if (times)
    for (auto& e : ev) hipEventCreate(&e);
  auto mark = [&](int k) {
    if (times) hipEventRecord(ev[k], s);
  };

  hipError_t e;
#define PST_TRY(expr, what)                         \
  do {                                              \
    e = (expr);                                     \
    if (e != hipSuccess) return hip_status(e, what); \
  } while (0)
  const BuildLayout BL = build_layout<K>(n64);
  // scratch in two parts: what the first key kernel writes, then the rest,
  // allocated while that kernel runs
  DevBuf arena_a, arena_b;
  PST_TRY(arena_a.alloc(BL.total_a), "Error in allocating construction arrays");
  unsigned char* base_a = arena_a.as<unsigned char>();
  K* key0 = reinterpret_cast<K*>(base_a + BL.key0);
  uint32_t* v0 = reinterpret_cast<uint32_t*>(base_a + BL.v0);
  uint32_t* v2 = reinterpret_cast<uint32_t*>(base_a + BL.v2);
  int* flags = reinterpret_cast<int*>(base_a + BL.flags);
  // flags[0]: NaN seen; flags[1]: IDs not in index order; flags[2]: a sort's wait ran out
  PST_TRY(hipMemsetAsync(flags, 0, 3 * sizeof(int), s), "Error in clearing construction flags");
  const unsigned B = 256;
  const unsigned G = grid_for(n, B);
  int* nan_flag = flags;
  mark(0);
  // ---- key2 descending -> perm2 (v1), r2. One sort by dim2 when the IDs
  // are in index order (the sort's index tie-break is then the (id, index)
  // tie-break), else a sort by ID and then a stable one by dim2.
  // 32-bit keys: the dim1 keys are stored by the same kernel, into v2
  constexpr bool kKey1First = sizeof(K) == 4;
  k_keys_single<T, I, K, true, true, false, kHasId>
      <<<G, B, 0, s>>>(pts, n, key0, v0, nullptr, flags, kKey1First ? v2 : nullptr);
  PST_TRY(arena_b.alloc(BL.total_b), "Error in allocating construction arrays");
  unsigned char* base_b = arena_b.as<unsigned char>();
  K* key1 = reinterpret_cast<K*>(base_b + BL.key1);
  uint32_t* v1 = reinterpret_cast<uint32_t*>(base_b + BL.v1);
  uint32_t* r2 = reinterpret_cast<uint32_t*>(base_b + BL.r2);
  void* tmp = base_b + BL.tmp;
  uint32_t* starts_d = reinterpret_cast<uint32_t*>(base_b + BL.starts);
  uint32_t* tileseg_d = reinterpret_cast<uint32_t*>(base_b + BL.tileseg);
  // level tables, which depend on n alone: segment starts of levels 0..t0
  // and, per level t < t0 and tile of kLevelTile points, the segment
  // holding the tile's first point (so that no workgroup searches the
  // starts itself)
  {
    const uint32_t nstart = (2u << BL.t0) - 1u;
    k_level_starts<<<grid_for(nstart, 256), 256, 0, s>>>(n64 + 1, nstart, starts_d);
    if (BL.t0 > 0) {
      const uint64_t cells = static_cast<uint64_t>(BL.t0) * BL.tile_stride;
      k_tile_segments<<<grid_for(cells, 256), 256, 0, s>>>(starts_d, static_cast<uint32_t>(n64), BL.t0,
                                                           static_cast<uint32_t>(BL.tile_stride), tileseg_d);
    }
  }
  // stable radix sort of (key, value) pairs: kin/vin to kout/vout
  auto sortK = [&](const K* kin, K* kout, const uint32_t* vin, uint32_t* vout) -> hipError_t {
    size_t tb = BL.tmp_bytes;
    if constexpr (sizeof(K) == 4) return radix_sort_u32(tmp, tb, kin, kout, vin, vout, n, s, flags + 2);
    else return radix_sort_u64(tmp, tb, kin, kout, vin, vout, n, s, flags + 2);
  };
  bool ordered = true;
  if constexpr (kHasId) {
    int hflags[2] = {0, 0};
    PST_TRY(hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, s), "Error in reading construction flags");
    PST_TRY(hipStreamSynchronize(s), "Error in synchronising construction");
    ordered = hflags[1] == 0;
  }
  if (ordered) {
    PST_TRY(sortK(key0, key1, v0, v1), "Error in sorting by dimension 2");
  } else {
    if constexpr (kHasId) {
      k_keys_id<T, I, K, true><<<G, B, 0, s>>>(pts, n, key0, v0, nan_flag);
      PST_TRY(sortK(key0, key1, v0, v2), "Error in sorting by ID");
      k_keys_dim_perm<T, I, K, true, true><<<G, B, 0, s>>>(pts, n, v2, key0, v0, nullptr, nan_flag);
      PST_TRY(sortK(key0, key1, v0, v1), "Error in sorting by dimension 2");
    }
  }
  const uint32_t* perm2 = v1;
  k_scatter_rank<<<G, B, 0, s>>>(perm2, n, r2);
  // ---- key1 ascending, carrying r2 -> A (v2; v0 when the dim1 keys came
  // with the dim2 keys, whose sort reads them from v2 and r2 in place)
  uint32_t* Acur = v2;
  uint32_t* Anxt = v0;
  if (ordered && kKey1First) {
    PST_TRY(sortK(reinterpret_cast<const K*>(v2), key0, r2, v0), "Error in sorting by dimension 1");
    std::swap(Acur, Anxt);
  } else if (ordered) {
    k_keys_single<T, I, K, false, false, false><<<G, B, 0, s>>>(pts, n, key0, v0, r2, flags);
    PST_TRY(sortK(key0, key1, v0, v2), "Error in sorting by dimension 1");
  } else {
    if constexpr (kHasId) {
      k_keys_id<T, I, K, false><<<G, B, 0, s>>>(pts, n, key0, v0, nan_flag);
      PST_TRY(sortK(key0, key1, v0, v2), "Error in sorting by ID");
      k_keys_dim_perm<T, I, K, false, false><<<G, B, 0, s>>>(pts, n, v2, key0, v0, r2, nan_flag);
      PST_TRY(sortK(key0, key1, v0, v2), "Error in sorting by dimension 1");
    }
  }
  mark(1);
  // the output arrays: the caller's hook allocates them while the sorts run
  if (out.alloc) {
    Status sa = out.alloc(out);
    if (!sa.ok()) {
      (void)hipStreamSynchronize(s);  // the queued passes still use the scratch
      return sa;
    }
  }
  auto* nodes = static_cast<Node<T, IS>*>(out.nodes);
  void* submin_v = out.submin;
  void* ids_v = out.ids;

  // ---- level schedule (tables computed above)
  const uint32_t t0 = BL.t0;
  const size_t nmin = t0 ? (1ull << t0) - 1 : 1;
  unsigned long long* segmin = reinterpret_cast<unsigned long long*>(base_b + BL.segmin);
  k_fill_u64<<<grid_for(nmin, B), B, 0, s>>>(segmin, nmin, ~0ull);
  const uint32_t* starts = starts_d;
  // level 0's minimum by its own pass; each emit pass then finds the next
  // level's minima while compacting. Segment sizes within a level differ by
  // at most one, and some segment of every level t < t0 holds more than
  // kFinishTile points, so all hold at least kFinishTile: a tile of
  // kLevelTile points meets at most kLevelTile / kFinishTile + 2 segments,
  // twice as many children.
  // Levels with at most kReduceSegs segments combine the tiles' minima by a
  // reduction instead of global atomics.
  constexpr uint32_t kReduceSegs = 32;
  unsigned long long* tmin = reinterpret_cast<unsigned long long*>(base_b + BL.tmin);
  if (t0 > 0) {
    k_level_argmin<<<grid_for(n, kLevelTile), kLevelBlock, 0, s>>>(Acur, n, 1u, starts, tileseg_d, tmin);
    k_tile_min_reduce<<<1, kLevelBlock, 0, s>>>(tmin, kSegSlots, tileseg_d, starts, 1u, n, 0u, segmin);
  }
  for (uint32_t t = 0; t < t0; t++) {
    const uint32_t n_level = n - ((1u << t) - 1);
    const unsigned tiles = grid_for(n_level, kLevelTile);
    unsigned long long* nextmin = t + 1 < t0 ? segmin + (2u << t) - 1 : nullptr;
    const bool reduce = nextmin && (1u << t) <= kReduceSegs;
    const uint32_t* st_t = starts + (1u << t) - 1;
    const uint32_t* ts_t = tileseg_d + static_cast<size_t>(t) * BL.tile_stride;
    k_level_emit<T, IS><<<tiles, kLevelBlock, 0, s>>>(Acur, Anxt, n_level, t, M, st_t, ts_t, segmin + (1u << t) - 1,
                                                      nextmin, reduce ? tmin : nullptr, perm2, ptsS, nodes);
    if (reduce)
      k_tile_min_reduce<<<2u << t, kLevelBlock, 0, s>>>(tmin, 2 * kSegSlots, ts_t, st_t, 1u << t, n_level, 1u,
                                                        nextmin);
    std::swap(Acur, Anxt);
  }
  mark(2);
  k_finish<T, IS><<<1u << t0, kFinishBlock, 0, s>>>(Acur, t0, M, starts + (1u << t0) - 1, perm2, ptsS, nodes,
                                                    static_cast<T*>(submin_v), static_cast<IS*>(ids_v));
  // subtree minima and IDs of the slots above the LDS subtrees
  if (submin_v && t0 > 0) {
    const uint32_t M32 = static_cast<uint32_t>(M);
    Status st2 = submin_typed<T, IS>(nodes, submin_v, ids_v, t0, 31u, 1u, M32, M32, s);
    if (!st2.ok()) return st2;
  }
  mark(3);
  PST_TRY(hipGetLastError(), "Error in launching construction kernels");
  int hflag[3] = {0, 0, 0};
  PST_TRY(hipMemcpyAsync(hflag, flags, sizeof(hflag), hipMemcpyDeviceToHost, s),
          "Error in reading construction flag");
  PST_TRY(hipStreamSynchronize(s), "Error in synchronising after construction");
  arena_b.release();
  arena_a.release();
  if (times) {
    hipEventElapsedTime(&times->sort_ms, ev[0], ev[1]);
    hipEventElapsedTime(&times->levels_ms, ev[1], ev[2]);
    hipEventElapsedTime(&times->finish_ms, ev[2], ev[3]);
    for (auto& x : ev) hipEventDestroy(x);
  }
  if (hflag[2]) {
    st.code = 5;
    st.msg = "Error: a construction sort pass waited too long for earlier tiles";
  } else if (hflag[0]) {
    st.code = 3;
    st.msg = "Error: input contains NaN values, which have no order";
  }
#undef PST_TRY
  return st;
}

template <typename T, typename IS>
Status export_typed(size_t n, const void* nodes, void* blob, size_t o1, size_t o2, size_t om,
                    size_t oi, size_t ob, hipStream_t s) {
  if (n == 0) return Status{};
  uint64_t S = 1;
  while (S <= n) S <<= 1;
  S -= 1;
  k_export<T, IS><<<grid_for(S, 256), 256, 0, s>>>(static_cast<const Node<T, IS>*>(nodes), S, n + 1,
                                                    static_cast<unsigned char*>(blob), o1, o2, om, oi, ob);
  return hip_status(hipGetLastError(), "Error in launching tree export");
}

template <typename T, typename I>
Status forest_build_typed(const void* pts_v, size_t n, size_t St, void* nodes_v, hipStream_t s) {
  using IS = typename IdStore<I>::type;
  if (n == 0) return Status{};
  uint32_t L = 0;
  while ((1ull << L) < St + 1) L++;
  if ((1ull << L) != St + 1 || St > kFinishTile - 1)
    return Status{1, "Error: forest trees hold 2^k - 1 <= 2047 points"};
  const uint64_t ntrees = (n + St - 1) / St;
  const uint32_t N = static_cast<uint32_t>(St + 1);
  const size_t lds = forest_lds_bytes(N);
  auto kern = k_forest_build<T, I>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
  if (e != hipSuccess) return hip_status(e, "Error in sizing forest construction LDS");
  DevBuf flag;
  e = flag.alloc(sizeof(int));
  if (e == hipSuccess) e = hipMemsetAsync(flag.p, 0, sizeof(int), s);
  if (e != hipSuccess) return hip_status(e, "Error in allocating construction flag");
  hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(ntrees)), dim3(kForestBlock), lds, s,
                     static_cast<const Point<T, I>*>(pts_v), static_cast<uint64_t>(n), static_cast<uint32_t>(St), L,
                     N, static_cast<Node<T, IS>*>(nodes_v), flag.as<int>());
  e = hipGetLastError();
  if (e != hipSuccess) return hip_status(e, "Error in launching forest construction");
  int hflag = 0;
  e = hipMemcpyAsync(&hflag, flag.p, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_status(e, "Error in synchronising after forest construction");
  if (hflag) return Status{3, "Error: input contains NaN values, which have no order"};
  return Status{};
}

template <typename T, typename IS>
Status forest_export_typed(size_t n, size_t St, const void* nodes, void* blob, size_t total, size_t o1, size_t o2,
                           size_t om, size_t oi, size_t ob, hipStream_t s) {
  if (n == 0) return Status{};
  uint32_t L = 0;
  while ((1ull << L) < St + 1) L++;
  const uint64_t ntrees = (n + St - 1) / St;
  k_forest_export<T, IS><<<grid_for(total, 256), 256, 0, s>>>(static_cast<const Node<T, IS>*>(nodes), n, St, L,
                                                              ntrees, total, static_cast<unsigned char*>(blob), o1,
                                                              o2, om, oi, ob);
  return hip_status(hipGetLastError(), "Error in launching forest export");
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
Status with_id(int idt, F&& f) {
  switch (idt) {
    case NO_ID: return f(NoId{});
    case F32: return f(float{});
    case F64: return f(double{});
    case I32: return f(int32_t{});
    case I64: return f(int64_t{});
    case U32: return f(uint32_t{});
    case U64: return f(uint64_t{});
  }
  return Status{1, "Error: unsupported ID type code " + std::to_string(idt)};
}

}  // namespace

Status build_tree(int dtype, int id_dtype, const void* pts_d, size_t n, BuildOut& out, hipStream_t stream,
                  BuildTimes* times) {
  return with_t(dtype, [&](auto t) {
    return with_id(id_dtype, [&](auto i) { return build_typed<decltype(t), decltype(i)>(pts_d, n, out, stream, times); });
  });
}

Status build_tree(int dtype, int id_dtype, const void* pts_d, size_t n, void* nodes_d, hipStream_t stream,
                  BuildTimes* times, void* submin_d, void* ids_d) {
  BuildOut out;
  out.nodes = nodes_d;
  out.submin = submin_d;
  out.ids = ids_d;
  return build_tree(dtype, id_dtype, pts_d, n, out, stream, times);
}

size_t build_scratch_bytes(int dtype, int id_dtype, size_t n) {
  size_t bytes = 0;
  Status st = with_t(dtype, [&](auto t) {
    return with_id(id_dtype, [&](auto i) {
      bytes = build_layout<BuildKey<decltype(t), decltype(i)>>(n).total;
      return Status{};
    });
  });
  return st.ok() ? bytes : 0;
}

Status export_tree(int dtype, int id_dtype, size_t n, const void* nodes_d, void* blob_d,
                   size_t o1, size_t o2, size_t om, size_t oi, size_t ob, hipStream_t stream) {
  return with_t(dtype, [&](auto t) {
    return with_id(id_dtype, [&](auto i) {
      using IS = typename IdStore<decltype(i)>::type;
      return export_typed<decltype(t), IS>(n, nodes_d, blob_d, o1, o2, om, oi, ob, stream);
    });
  });
}

}  // namespace pst

namespace pst {

Status build_forest(int dtype, int id_dtype, const void* pts_d, size_t n, size_t slots_per_tree, void* nodes_d,
                    hipStream_t stream) {
  return with_t(dtype, [&](auto t) {
    return with_id(id_dtype, [&](auto i) {
      return forest_build_typed<decltype(t), decltype(i)>(pts_d, n, slots_per_tree, nodes_d, stream);
    });
  });
}

Status export_forest(int dtype, int id_dtype, size_t n, size_t slots_per_tree, const void* nodes_d, void* blob_d,
                     size_t total_slots, size_t o1, size_t o2, size_t om, size_t oi, size_t ob,
                     hipStream_t stream) {
  return with_t(dtype, [&](auto t) {
    return with_id(id_dtype, [&](auto i) {
      using IS = typename IdStore<decltype(i)>::type;
      return forest_export_typed<decltype(t), IS>(n, slots_per_tree, nodes_d, blob_d, total_slots, o1, o2, om, oi,
                                                  ob, stream);
    });
  });
}

}  // namespace pst

namespace pst {

Status build_submin(int dtype, int id_dtype, size_t n, size_t slots_per_tree, const void* nodes_d, void* submin_d,
                    void* ids_d, hipStream_t stream) {
  if (n == 0) return Status{};
  uint32_t shift, ntrees, full_M, last_M, levels = 0;
  if (slots_per_tree == 0) {
    shift = 31;
    ntrees = 1;
    full_M = last_M = static_cast<uint32_t>(n + 1);
    while ((1ull << levels) <= n) levels++;
  } else {
    shift = 0;
    while ((1ull << shift) < slots_per_tree + 1) shift++;
    ntrees = static_cast<uint32_t>((n + slots_per_tree - 1) / slots_per_tree);
    full_M = static_cast<uint32_t>(slots_per_tree + 1);
    last_M = static_cast<uint32_t>(n - (ntrees - 1ull) * slots_per_tree + 1);
    levels = shift;
  }
  return with_t(dtype, [&](auto t) {
    return with_id(id_dtype, [&](auto i) {
      using IS = typename IdStore<decltype(i)>::type;
      return submin_typed<decltype(t), IS>(nodes_d, submin_d, ids_d, levels, shift, ntrees, full_M, last_M, stream);
    });
  });
}

}  // namespace pst
