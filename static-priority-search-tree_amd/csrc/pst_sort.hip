// Stable LSD radix sort of (unsigned key, 32-bit value) pairs for gfx950,
// used by tree construction, level index construction and query grouping.
//
// The reference sorts size_t index arrays with thrust::sort and an
// indirect comparator (pst/static-pst-gpu.tu:266-267, 317-318). Here the
// comparison keys are order-preserving unsigned integers (pst_common.hpp:
// ord_key), sorted 8 bits per pass, least significant digit first. Every
// pass is stable, which construction relies on for its index tie-breaks.
//
// One pass over tiles of 8192 keys (four waves of 2048, each wave's keys
// striped so that lane l of round r holds key r * 64 + l of the wave):
//   k_upsweep    per tile, the count of every digit (LDS atomics);
//   k_scan_rows  per digit, an exclusive scan of its tile counts in tile
//                order, and the digit's total;
//   k_digit_starts  the digits' starts: an exclusive scan of the totals;
//   k_downsweep  per tile, a stable rank for every key: within a wave by
//                ballot matching of the 8 digit bits (lanes with the same
//                digit in a round), across rounds by per-wave digit
//                counters in LDS, across waves by their digit prefix; the
//                tile is then reordered in LDS by (digit, rank) and written
//                out in that order, so each digit's keys leave as one
//                contiguous run.
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
constexpr int kSortBlock = 256;
constexpr int kSortWaves = kSortBlock / kWave;
constexpr int kSortPer = 32;                                    // keys per lane
constexpr uint32_t kWaveKeys = kWave * kSortPer;                // 2048
constexpr uint32_t kSortTile = kSortBlock * kSortPer;           // 8192
constexpr int kScanThreads = 1024;

template <typename K>
__device__ inline uint32_t digit_of(K k, uint32_t shift) {
  return static_cast<uint32_t>(k >> shift) & (kDigits - 1);
}

// LDS digit counters are kept in kSub copies, lane l adding to copy l %
// kSub: lanes of a wave that share a digit (frequent in the high digits of
// clustered keys) then collide on one counter kSub times less often.
constexpr int kSub = 4;

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

// counts[d * ntiles + tile] = keys of digit d in the tile.
template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_upsweep(const K* __restrict__ keys, uint32_t n, uint32_t shift,
                                                        uint32_t ntiles, uint32_t* __restrict__ counts) {
  __shared__ uint32_t hs[kSub][kDigits];
#pragma unroll
  for (int u = 0; u < kSub; u++) hs[u][threadIdx.x] = 0;
  __syncthreads();
  uint32_t* h = hs[threadIdx.x % kSub];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
  const uint32_t valid = static_cast<uint32_t>(std::min<uint64_t>(kSortTile, n - base));
  K k[kSortPer];
#pragma unroll
  for (int e = 0; e < kSortPer; e++) {
    const uint32_t i = e * kSortBlock + threadIdx.x;
    if (i < valid) k[e] = keys[base + i];
  }
#pragma unroll
  for (int e = 0; e < kSortPer; e++) {
    const uint32_t i = e * kSortBlock + threadIdx.x;
    if (i < valid) atomicAdd(&h[digit_of(k[e], shift)], 1u);
  }
  __syncthreads();
  uint32_t c = 0;
#pragma unroll
  for (int u = 0; u < kSub; u++) c += hs[u][threadIdx.x];
  counts[static_cast<uint64_t>(threadIdx.x) * ntiles + blockIdx.x] = c;
}

// Per digit (one block each): exclusive scan of its tile counts in place;
// totals[d] = the digit's count over all tiles.
__global__ __launch_bounds__(kScanThreads) void k_scan_rows(uint32_t* __restrict__ counts, uint32_t ntiles,
                                                            uint32_t* __restrict__ totals) {
  __shared__ uint32_t part[kScanThreads / kWave];
  __shared__ uint32_t carry;
  const uint32_t d = blockIdx.x;
  uint32_t* row = counts + static_cast<uint64_t>(d) * ntiles;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (uint32_t b = 0; b < ntiles; b += kScanThreads) {
    const uint32_t i = b + tid;
    const uint32_t v = i < ntiles ? row[i] : 0u;
    uint32_t inc = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t o = __shfl_up(inc, off, kWave);
      if (static_cast<int>(lane) >= off) inc += o;
    }
    if (lane == kWave - 1) part[wave] = inc;
    __syncthreads();
    uint32_t before = carry;
    for (uint32_t w = 0; w < wave; w++) before += part[w];
    if (i < ntiles) row[i] = before + inc - v;
    __syncthreads();
    if (tid == kScanThreads - 1) carry = before + inc;
    __syncthreads();
  }
  if (tid == 0) totals[d] = carry;
}

template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_downsweep(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                          K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                          uint32_t n, uint32_t shift, uint32_t ntiles,
                                                          const uint32_t* __restrict__ offs,
                                                          const uint32_t* __restrict__ starts) {
  __shared__ K skey[kSortTile];
  __shared__ uint32_t sval[kSortTile];
  __shared__ uint32_t wcnt[kSortWaves][kDigits];  // per-wave digit counters, then their wave prefix
  __shared__ uint32_t dstart[kDigits];             // first tile position of each digit
  __shared__ uint32_t goff[kDigits];               // global position of each digit's first key of the tile
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const uint32_t tile = blockIdx.x;
  const uint64_t base = static_cast<uint64_t>(tile) * kSortTile;
  const uint32_t valid = static_cast<uint32_t>(std::min<uint64_t>(kSortTile, n - base));
#pragma unroll
  for (int w = 0; w < kSortWaves; w++) wcnt[w][tid] = 0;
  goff[tid] = starts[tid] + offs[static_cast<uint64_t>(tid) * ntiles + tile];
  K k[kSortPer];
  uint32_t v[kSortPer];
  const uint32_t wbase = wave * kWaveKeys;
#pragma unroll
  for (int r = 0; r < kSortPer; r++) {
    const uint32_t i = wbase + r * kWave + lane;
    if (i < valid) {
      k[r] = kin[base + i];
      v[r] = vin[base + i];
    }
  }
  __syncthreads();
  // stable rank within the wave: rounds in element order, lanes in order
  uint32_t rank[kSortPer];
  uint32_t* cnt = wcnt[wave];
  const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int r = 0; r < kSortPer; r++) {
    const bool ok = wbase + r * kWave + lane < valid;
    const uint32_t d = ok ? digit_of(k[r], shift) : 0u;
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < kDigitBits; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t below = static_cast<uint32_t>(__builtin_popcountll(peers & lt));
    uint32_t before = 0;
    if (ok) before = cnt[d];
    __builtin_amdgcn_wave_barrier();
    if (ok && below == 0) cnt[d] = before + static_cast<uint32_t>(__builtin_popcountll(peers));
    __builtin_amdgcn_wave_barrier();
    rank[r] = before + below;
  }
  __syncthreads();
  // per digit (thread tid = digit): wave prefixes and the tile's digit starts
  {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; w++) {
      const uint32_t c = wcnt[w][tid];
      wcnt[w][tid] = run;
      run += c;
    }
    // block-wide exclusive scan of the digit totals
    uint32_t inc = run;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t o = __shfl_up(inc, off, kWave);
      if (static_cast<int>(lane) >= off) inc += o;
    }
    dstart[tid] = inc;  // inclusive within the wave for now
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; w++) before += dstart[w * kWave + kWave - 1];
    __syncthreads();
    dstart[tid] = before + inc - run;
  }
  __syncthreads();
  // reorder the tile in LDS by (digit, rank)
#pragma unroll
  for (int r = 0; r < kSortPer; r++) {
    const uint32_t i = wbase + r * kWave + lane;
    if (i < valid) {
      const uint32_t d = digit_of(k[r], shift);
      const uint32_t p = dstart[d] + wcnt[wave][d] + rank[r];
      skey[p] = k[r];
      sval[p] = v[r];
    }
  }
  __syncthreads();
  // write out: each digit's keys form a contiguous run of the output
#pragma unroll
  for (int r = 0; r < kSortPer; r++) {
    const uint32_t i = r * kSortBlock + tid;
    if (i < valid) {
      const K key = skey[i];
      const uint32_t d = digit_of(key, shift);
      const uint32_t g = goff[d] + (i - dstart[d]);
      kout[g] = key;
      vout[g] = sval[i];
    }
  }
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t{255}; }

template <typename K>
hipError_t radix_sort(void* tmp, size_t& tmp_bytes, const K* kin, K* kout, const uint32_t* vin, uint32_t* vout,
                      size_t n, hipStream_t s) {
  constexpr uint32_t passes = sizeof(K) * 8 / kDigitBits;
  const uint32_t ntiles = static_cast<uint32_t>((n + kSortTile - 1) / kSortTile);
  const size_t need_keys = align256(n * sizeof(K));
  const size_t need_vals = align256(n * 4);
  const size_t need_counts = align256(static_cast<size_t>(ntiles) * kDigits * 4);
  const size_t need_hist = align256(kDigits * 4);
  const size_t need = need_keys + need_vals + need_counts + need_hist;
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
  uint32_t* counts = reinterpret_cast<uint32_t*>(p + need_keys + need_vals);
  uint32_t* hist = reinterpret_cast<uint32_t*>(p + need_keys + need_vals + need_counts);
  const uint32_t n32 = static_cast<uint32_t>(n);
  // passes alternate so that the last one writes kout: an even number of
  // passes starts into the scratch, an odd number into kout
  const K* ksrc = kin;
  const uint32_t* vsrc = vin;
  for (uint32_t pass = 0; pass < passes; pass++) {
    const bool to_out = ((passes - pass) % 2) == 1;
    K* kdst = to_out ? kout : kt;
    uint32_t* vdst = to_out ? vout : vt;
    const uint32_t shift = pass * kDigitBits;
    hipLaunchKernelGGL((k_upsweep<K>), dim3(ntiles), dim3(kSortBlock), 0, s, ksrc, n32, shift, ntiles, counts);
    // digit totals from the tile counts, then the digits' starts
    hipLaunchKernelGGL(k_scan_rows, dim3(kDigits), dim3(kScanThreads), 0, s, counts, ntiles, hist);
    hipLaunchKernelGGL(k_digit_starts, dim3(1), dim3(kDigits), 0, s, hist, 1u);
    hipLaunchKernelGGL((k_downsweep<K>), dim3(ntiles), dim3(kSortBlock), 0, s, ksrc, vsrc, kdst, vdst, n32, shift,
                       ntiles, counts, hist);
    ksrc = kdst;
    vsrc = vdst;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t radix_sort_u32(void* tmp, size_t& tmp_bytes, const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
                          uint32_t* vout, size_t n, hipStream_t s) {
  return radix_sort<uint32_t>(tmp, tmp_bytes, kin, kout, vin, vout, n, s);
}

hipError_t radix_sort_u64(void* tmp, size_t& tmp_bytes, const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                          uint32_t* vout, size_t n, hipStream_t s) {
  return radix_sort<uint64_t>(tmp, tmp_bytes, kin, kout, vin, vout, n, s);
}

}  // namespace pst
