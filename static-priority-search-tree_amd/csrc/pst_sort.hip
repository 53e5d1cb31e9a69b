// Stable LSD radix sort of (unsigned key, 32-bit value) pairs for gfx950,
// used by tree construction, level index construction and query grouping.
//
// The reference sorts size_t index arrays with thrust::sort and an
// indirect comparator (pst/static-pst-gpu.tu:266-267, 317-318). Here the
// comparison keys are order-preserving unsigned integers (pst_common.hpp:
// ord_key), sorted 8 bits per pass, least significant digit first. Every
// pass is stable, which construction relies on for its index tie-breaks.
//
// A sort is one histogram pass over the keys and then one kernel per digit:
//   k_sort_hist     every digit's count for every pass, from one read of the
//                   keys (LDS counters, then one global atomic per counter
//                   and workgroup);
//   k_digit_starts  per pass, each digit's first output position;
//   k_sort_pass     per tile of 8192 keys (eight waves of 1024, lane l of
//                   round r holding key r * 64 + l of its wave):
//     1. takes the next tile number from a counter, so that tiles start in
//        tile order;
//     2. ranks every key stably within the tile: within a wave by ballot
//        matching of the 8 digit bits (lanes with the same digit in a
//        round), across rounds by per-wave digit counters in LDS, across
//        waves by their digit prefix;
//     3. publishes its digit counts, then adds up the counts of the tiles
//        before it by reading their published words backwards until one
//        carries an inclusive prefix (decoupled look-back), and publishes
//        its own inclusive prefix;
//     4. reorders the tile in LDS by (digit, rank) and writes the keys, then
//        the values through the same LDS buffer, so that each digit's keys
//        leave as one contiguous run.
// A tile waits only for tiles with smaller numbers, which took their
// numbers earlier and are therefore running, and every tile publishes its
// counts before it waits: every wait ends. The wait is still bounded; a
// tile that exceeds the bound sets the caller's failure word and goes on.
// Passes alternate between the output and a scratch buffer, ending in the
// output; the inputs are not modified.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "pst_internal.hpp"

namespace pst {

namespace {

constexpr int kDigitBits = 8;
constexpr int kDigits = 1 << kDigitBits;
constexpr int kPassBlock = 512;
constexpr int kPassWaves = kPassBlock / kWave;            // 8
constexpr int kPassPer = 16;                              // keys per lane
constexpr uint32_t kWaveKeys = kWave * kPassPer;          // 1024
constexpr uint32_t kSortTile = kPassBlock * kPassPer;     // 8192
constexpr int kHistBlock = 512;
constexpr int kHistPer = 16;
constexpr uint32_t kHistChunk = kHistBlock * kHistPer;
constexpr unsigned kHistMaxBlocks = 1024;
constexpr uint32_t kLookbackSpins = 1u << 24;             // about a second
constexpr uint32_t kLookAhead = 4;                        // look-back words read at once

template <typename K>
__device__ inline uint32_t digit_of(K k, uint32_t shift) {
  return static_cast<uint32_t>(k >> shift) & (kDigits - 1);
}

// LDS digit counters of the histogram are kept in kSub copies, lane l
// adding to copy l % kSub: lanes that share a digit (frequent in the high
// digits of clustered keys) collide on one counter kSub times less often.
constexpr int kSub = 4;

// Published per (tile, digit): the count in the low 32 bits, a tag in the
// high 32 bits: 0 nothing yet, 2 * pass + 2 the tile's own count, 2 * pass
// + 3 the inclusive prefix over tiles 0..tile. Tags differ between the
// passes of one sort, so the words are cleared once per sort.
__device__ inline uint32_t tag_count(uint32_t pass) { return 2 * pass + 2; }
__device__ inline uint32_t tag_prefix(uint32_t pass) { return 2 * pass + 3; }
__device__ inline unsigned long long pack_status(uint32_t tag, uint32_t v) {
  return (static_cast<unsigned long long>(tag) << 32) | v;
}

__global__ void k_fill_zero(uint32_t* __restrict__ p, size_t count) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0u;
}

// hist[p * 256 + d] += keys of digit d at pass p.
template <typename K>
__global__ __launch_bounds__(kHistBlock) void k_sort_hist(const K* __restrict__ keys, uint32_t n,
                                                          uint32_t* __restrict__ hist) {
  constexpr int P = static_cast<int>(sizeof(K));
  __shared__ uint32_t h[kSub][P][kDigits];
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < kSub * P * kDigits; i += kHistBlock) (&h[0][0][0])[i] = 0u;
  __syncthreads();
  uint32_t(*mine)[kDigits] = h[tid % kSub];
  for (uint64_t b = static_cast<uint64_t>(blockIdx.x) * kHistChunk; b < n;
       b += static_cast<uint64_t>(gridDim.x) * kHistChunk) {
    K k[kHistPer];
#pragma unroll
    for (int e = 0; e < kHistPer; e++) {
      const uint64_t i = b + static_cast<uint64_t>(e) * kHistBlock + tid;
      k[e] = i < n ? keys[i] : K(0);
    }
#pragma unroll
    for (int e = 0; e < kHistPer; e++) {
      const bool ok = b + static_cast<uint64_t>(e) * kHistBlock + tid < n;
      if (ok) {
#pragma unroll
        for (int p = 0; p < P - 1; p++) atomicAdd(&mine[p][digit_of(k[e], p * kDigitBits)], 1u);
      }
      // the top digit, often shared by a whole wave of clustered keys: such
      // a wave adds once
      const uint32_t d = digit_of(k[e], (P - 1) * kDigitBits);
      const uint64_t act = __ballot(ok);
      const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
      if (act == ~0ull && __ballot(d != d0) == 0ull) {
        if (tid % kWave == 0) atomicAdd(&h[0][P - 1][d0], static_cast<uint32_t>(kWave));
      } else if (ok) {
        atomicAdd(&mine[P - 1][d], 1u);
      }
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < P * kDigits; i += kHistBlock) {
    uint32_t c = 0;
#pragma unroll
    for (int u = 0; u < kSub; u++) c += (&h[u][0][0])[i];
    if (c) atomicAdd(&hist[i], c);
  }
}

// Exclusive scan of each of `passes` rows of 256 digit totals, in place (one block).
__global__ __launch_bounds__(kDigits) void k_digit_starts(uint32_t* __restrict__ hist, uint32_t passes) {
  __shared__ uint32_t s[kDigits];
  for (uint32_t p = 0; p < passes; p++) {
    const uint32_t v = hist[p * kDigits + threadIdx.x];
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t off = 1; off < kDigits; off <<= 1) {
      const uint32_t o = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
      __syncthreads();
      s[threadIdx.x] += o;
      __syncthreads();
    }
    hist[p * kDigits + threadIdx.x] = s[threadIdx.x] - v;
    __syncthreads();
  }
}

template <typename K>
struct PassLds {
  union {
    unsigned long long peer[kPassWaves][kDigits];  // while ranking: per-wave lane masks of a round's digits
    K sbuf[kSortTile];                             // then keys, then values, in (digit, rank) order
  };
  uint32_t wcnt[kPassWaves][kDigits];   // per-wave digit counters, then their wave prefix
  uint32_t dcnt[kDigits];               // the tile's count of each digit
  uint32_t dstart[kDigits];             // first tile position of each digit
  uint32_t goff[kDigits];               // global position of each digit's first key of the tile
  uint32_t wtot[kDigits / kWave];
  uint32_t tile;
};

// One tile of a digit pass. FULL: the tile holds kSortTile keys, so no
// element needs a bounds check (and each lane's loads share one address).
template <typename K, bool FULL>
__device__ __forceinline__ void sort_tile(PassLds<K>& L, const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                          K* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t tile,
                                          uint32_t valid, uint32_t pass, unsigned long long* __restrict__ status,
                                          const uint32_t* __restrict__ gstart, int* __restrict__ fail) {
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const uint32_t shift = pass * kDigitBits;
  const uint64_t base = static_cast<uint64_t>(tile) * kSortTile;
  const uint32_t wbase = wave * kWaveKeys;
  auto in_tile = [&](uint32_t i) { return FULL || i < valid; };
  const K* kp = kin + base + wbase + lane;
  K k[kPassPer];
#pragma unroll
  for (int r = 0; r < kPassPer; r++)
    if (in_tile(wbase + r * kWave + lane)) k[r] = kp[r * kWave];
  // stable rank within the wave: rounds in element order, lanes in order.
  // The lanes of a round that share a digit find each other through the
  // wave's peer word of that digit: each ORs in its lane bit, all read the
  // word back (one wave's LDS instructions complete in order), and the
  // lowest of them advances the digit's counter and clears the word.
  uint32_t rank[kPassPer];
  uint32_t* cnt = L.wcnt[wave];
  unsigned long long* pw = L.peer[wave];
  const uint64_t me = 1ull << lane;
  const uint64_t lt = me - 1ull;
#pragma unroll
  for (int r = 0; r < kPassPer; r++) {
    const bool ok = in_tile(wbase + r * kWave + lane);
    const uint32_t d = digit_of(k[r], shift);
    uint64_t peers = 0;
    if (ok) atomicOr(&pw[d], me);
    __builtin_amdgcn_wave_barrier();
    if (ok) peers = pw[d];
    const uint32_t below = static_cast<uint32_t>(__builtin_popcountll(peers & lt));
    uint32_t before = 0;
    if (ok) before = cnt[d];
    __builtin_amdgcn_wave_barrier();
    if (ok && below == 0) {
      cnt[d] = before + static_cast<uint32_t>(__builtin_popcountll(peers));
      pw[d] = 0ull;
    }
    __builtin_amdgcn_wave_barrier();
    rank[r] = before + below;
    // keep the rank computed here: sunk to its use, every round's peer
    // word would stay live until the reorder
    asm volatile("" : "+v"(rank[r]));
  }
  __syncthreads();
  // per digit (thread tid = digit): wave prefixes and the tile's count,
  // published at once, and the in-tile exclusive scan of the counts
  if (tid < kDigits) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kPassWaves; w++) {
      const uint32_t c = L.wcnt[w][tid];
      L.wcnt[w][tid] = run;
      run += c;
    }
    L.dcnt[tid] = run;
    __hip_atomic_store(&status[static_cast<size_t>(tile) * kDigits + tid],
                       pack_status(tile == 0 ? tag_prefix(pass) : tag_count(pass), run), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    uint32_t inc = run;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t o = __shfl_up(inc, off, kWave);
      if (static_cast<int>(lane) >= off) inc += o;
    }
    if (lane == kWave - 1) L.wtot[wave] = inc;
    L.dstart[tid] = inc - run;  // exclusive within the wave for now
  }
  __syncthreads();
  if (tid < kDigits) {
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; w++) before += L.wtot[w];
    L.dstart[tid] += before;
  }
  __syncthreads();
  // keys into LDS by (digit, rank); rank[] becomes the LDS position
#pragma unroll
  for (int r = 0; r < kPassPer; r++) {
    if (in_tile(wbase + r * kWave + lane)) {
      const uint32_t d = digit_of(k[r], shift);
      const uint32_t p = L.dstart[d] + L.wcnt[wave][d] + rank[r];
      L.sbuf[p] = k[r];
      rank[r] = p;
    }
  }
  const uint32_t* vp = vin + base + wbase + lane;
  uint32_t v[kPassPer];
#pragma unroll
  for (int r = 0; r < kPassPer; r++)
    if (in_tile(wbase + r * kWave + lane)) v[r] = vp[r * kWave];
  // look-back: the digit's count over the tiles before this one
  if (tid < kDigits) {
    const uint32_t run = L.dcnt[tid];
    uint32_t excl = 0;
    if (tile > 0) {
      const uint32_t ta = tag_count(pass), tp = tag_prefix(pass);
      uint32_t j = tile - 1;  // the next tile to read; tiles above j are summed
      uint32_t spins = 0;
      bool done = false;
      while (!done) {
        // kLookAhead words at a time (independent loads), then taken in
        // order: a prefix ends the walk, a count is added, a word not yet
        // published stops this step
        unsigned long long w[kLookAhead];
#pragma unroll
        for (uint32_t i = 0; i < kLookAhead; i++)
          w[i] = i <= j ? __hip_atomic_load(&status[static_cast<size_t>(j - i) * kDigits + tid], __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT)
                        : 0ull;
        bool stalled = false;
#pragma unroll
        for (uint32_t i = 0; i < kLookAhead; i++) {
          if (done || stalled) continue;
          const uint32_t tag = static_cast<uint32_t>(w[i] >> 32);
          if (tag == tp) {
            excl += static_cast<uint32_t>(w[i]);
            done = true;
          } else if (tag == ta && j > 0) {  // tile 0 publishes a prefix
            excl += static_cast<uint32_t>(w[i]);
            j--;
          } else {
            stalled = true;
          }
        }
        if (done || !stalled) continue;
        if (++spins > kLookbackSpins) {
          *fail = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      __hip_atomic_store(&status[static_cast<size_t>(tile) * kDigits + tid], pack_status(tp, excl + run),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    L.goff[tid] = gstart[tid] + excl;
  }
  __syncthreads();
  // write out: each digit's keys form a contiguous run of the output
  uint32_t g[kPassPer];
#pragma unroll
  for (int r = 0; r < kPassPer; r++) {
    const uint32_t i = r * kPassBlock + tid;
    if (in_tile(i)) {
      const K key = L.sbuf[i];
      const uint32_t d = digit_of(key, shift);
      g[r] = L.goff[d] + (i - L.dstart[d]);
      kout[g[r]] = key;
    }
  }
  __syncthreads();
  uint32_t* sval = reinterpret_cast<uint32_t*>(L.sbuf);
#pragma unroll
  for (int r = 0; r < kPassPer; r++)
    if (in_tile(wbase + r * kWave + lane)) sval[rank[r]] = v[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kPassPer; r++) {
    const uint32_t i = r * kPassBlock + tid;
    if (in_tile(i)) vout[g[r]] = sval[i];
  }
}

// One digit pass: tiles in counter order, look-back over the published
// per-digit counts of earlier tiles, then a stable scatter.
// Six waves per SIMD: three workgroups per CU for 32-bit keys (LDS bounds
// 64-bit keys to two).
template <typename K>
__global__ __launch_bounds__(kPassBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_sort_pass(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                          K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                          uint32_t n, uint32_t pass, uint32_t* __restrict__ tile_ctr,
                                                          unsigned long long* __restrict__ status,
                                                          const uint32_t* __restrict__ gstart,
                                                          int* __restrict__ fail) {
  __shared__ PassLds<K> L;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) L.tile = atomicAdd(&tile_ctr[pass], 1u);
  for (uint32_t i = tid; i < kPassWaves * kDigits; i += kPassBlock) {
    (&L.wcnt[0][0])[i] = 0u;
    (&L.peer[0][0])[i] = 0ull;
  }
  __syncthreads();
  const uint32_t tile = L.tile;
  if (tile >= n / kSortTile + (n % kSortTile != 0 ? 1u : 0u)) return;  // one tile per workgroup
  // 32-bit arithmetic: tile * kSortTile < n < 2^32. (A 64-bit minimum here
  // was lowered to a select on a stale SCC, giving full tiles past n.)
  const uint32_t rem = n - tile * kSortTile;
  const uint32_t valid = rem < kSortTile ? rem : kSortTile;
  if (valid == kSortTile)
    sort_tile<K, true>(L, kin, vin, kout, vout, tile, valid, pass, status, gstart, fail);
  else
    sort_tile<K, false>(L, kin, vin, kout, vout, tile, valid, pass, status, gstart, fail);
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t{255}; }

template <typename K>
hipError_t radix_sort(void* tmp, size_t& tmp_bytes, const K* kin, K* kout, const uint32_t* vin, uint32_t* vout,
                      size_t n, hipStream_t s, int* fail) {
  constexpr uint32_t passes = sizeof(K) * 8 / kDigitBits;
  const size_t ntiles = (n + kSortTile - 1) / kSortTile;
  const size_t need_keys = align256(n * sizeof(K));
  const size_t need_vals = align256(n * 4);
  // cleared once per sort: the status words, the tile counters and the
  // failure word, the digit totals
  const size_t status_bytes = std::max<size_t>(ntiles, 1) * kDigits * 8;
  const size_t ctr_bytes = 256;
  const size_t hist_bytes = passes * kDigits * 4;
  const size_t need_clear = align256(status_bytes + ctr_bytes + hist_bytes);
  const size_t need = need_keys + need_vals + need_clear;
  if (!tmp) {
    tmp_bytes = std::max<size_t>(need, 256);
    return hipSuccess;
  }
  if (tmp_bytes < need) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  if (n >= (1ull << 32)) return hipErrorInvalidValue;
  unsigned char* p = static_cast<unsigned char*>(tmp);
  K* kt = reinterpret_cast<K*>(p);
  uint32_t* vt = reinterpret_cast<uint32_t*>(p + need_keys);
  unsigned char* clr = p + need_keys + need_vals;
  auto* status = reinterpret_cast<unsigned long long*>(clr);
  uint32_t* ctr = reinterpret_cast<uint32_t*>(clr + status_bytes);
  uint32_t* hist = reinterpret_cast<uint32_t*>(clr + status_bytes + ctr_bytes);
  if (!fail) fail = reinterpret_cast<int*>(ctr + passes);
  const uint32_t n32 = static_cast<uint32_t>(n);
  const size_t clear_words = (status_bytes + ctr_bytes + hist_bytes) / 4;
  hipLaunchKernelGGL(k_fill_zero, dim3(static_cast<unsigned>((clear_words + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<uint32_t*>(clr), clear_words);
  const unsigned hblocks =
      static_cast<unsigned>(std::min<size_t>((n + kHistChunk - 1) / kHistChunk, kHistMaxBlocks));
  hipLaunchKernelGGL((k_sort_hist<K>), dim3(hblocks), dim3(kHistBlock), 0, s, kin, n32, hist);
  hipLaunchKernelGGL(k_digit_starts, dim3(1), dim3(kDigits), 0, s, hist, passes);
  // passes alternate so that the last one writes kout: an even number of
  // passes starts into the scratch, an odd number into kout
  const K* ksrc = kin;
  const uint32_t* vsrc = vin;
  for (uint32_t pass = 0; pass < passes; pass++) {
    const bool to_out = ((passes - pass) % 2) == 1;
    K* kdst = to_out ? kout : kt;
    uint32_t* vdst = to_out ? vout : vt;
    hipLaunchKernelGGL((k_sort_pass<K>), dim3(static_cast<unsigned>(ntiles)), dim3(kPassBlock), 0, s, ksrc, vsrc,
                       kdst, vdst, n32, pass, ctr, status, hist + pass * kDigits, fail);
    ksrc = kdst;
    vsrc = vdst;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t radix_sort_u32(void* tmp, size_t& tmp_bytes, const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
                          uint32_t* vout, size_t n, hipStream_t s, int* fail) {
  return radix_sort<uint32_t>(tmp, tmp_bytes, kin, kout, vin, vout, n, s, fail);
}

hipError_t radix_sort_u64(void* tmp, size_t& tmp_bytes, const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                          uint32_t* vout, size_t n, hipStream_t s, int* fail) {
  return radix_sort<uint64_t>(tmp, tmp_bytes, kin, kout, vin, vout, n, s, fail);
}

}  // namespace pst
