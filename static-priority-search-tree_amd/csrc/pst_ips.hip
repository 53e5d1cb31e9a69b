// Interval parallel search (Liu et al. 2016), the brute-force baseline the
// PST is compared with: report every point with dim1 <= q <= dim2
// (interval-parallel-search/interval-parallel-search.h:94-96 of the
// reference), as IDs or whole points, in unspecified order.
//
// The reference launches a grid-stride loop that, per tile, block-scans
// the hit flags through cooperative groups and takes one atomicAdd per
// tile (interval-parallel-search.h:70-117 with
// helpers/calc-alloc-report-ind-offset.h:46-133). Here every wavefront is
// an independent streaming filter with no block barriers:
//  - a wave takes chunks of 64 x 16 consecutive points (grid-stride over
//    chunks) and issues all 16 record loads of a chunk before testing any,
//    each load one lane-contiguous run of 64 records;
//  - per record position, a ballot gives the wave's hits; hits go to a
//    per-wave stage in LDS at their ballot/mbcnt prefix;
//  - a full stage is flushed with one atomic on the per-call counter
//    followed by coalesced stores; a chunk with many hits (dense queries)
//    reserves its own range and is written directly.
// So the counter sees about one atomic per 512 results (sparse) or per
// chunk (dense), not one per tile.
#include <algorithm>

#include "pst_internal.hpp"

namespace pst {

namespace {

constexpr int kIpsBlock = 256;              // four independent waves per workgroup
constexpr int kIpsElems = 16;               // records per lane per chunk
constexpr uint32_t kIpsChunk = kWave * kIpsElems;  // 1024 records per wave chunk
constexpr uint32_t kIpsStage = 512;         // staged results per wave
constexpr uint32_t kIpsDirect = 256;        // a chunk with at least this many hits is written directly

template <typename T, typename IS, bool RET_POINT>
__global__ __launch_bounds__(kIpsBlock) void k_ips(const Point<T, IS>* __restrict__ pts, uint64_t n, T q,
                                                   void* __restrict__ out, unsigned long long cap,
                                                   unsigned long long* __restrict__ counter) {
  using Rec = typename std::conditional<RET_POINT, Point<T, IS>, IS>::type;
  __shared__ Rec stage_all[kIpsBlock / kWave][kIpsStage];
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  Rec* stage = stage_all[wave];
  Rec* o = static_cast<Rec*>(out);
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (kIpsBlock / kWave);
  const uint64_t nchunks = (n + kIpsChunk - 1) / kIpsChunk;
  uint32_t staged = 0;

  auto rec_of = [](const Point<T, IS>& p) -> Rec {
    if constexpr (RET_POINT) return p;
    else return p.id;
  };
  auto flush = [&]() {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(counter, static_cast<unsigned long long>(staged));
    base = __shfl(base, 0, kWave);
    __builtin_amdgcn_wave_barrier();
    for (uint32_t i = lane; i < staged; i += kWave)
      if (base + i < cap) o[base + i] = stage[i];
    __builtin_amdgcn_wave_barrier();
    staged = 0;
  };

  for (uint64_t c = static_cast<uint64_t>(blockIdx.x) * (kIpsBlock / kWave) + wave; c < nchunks; c += waves) {
    const uint64_t first = c * kIpsChunk;
    Point<T, IS> p[kIpsElems];
    uint64_t hb[kIpsElems];
    if (first + kIpsChunk <= n) {
      // full chunk: every load issued before any test
#pragma unroll
      for (int e = 0; e < kIpsElems; e++) p[e] = pts[first + static_cast<uint64_t>(e) * kWave + lane];
#pragma unroll
      for (int e = 0; e < kIpsElems; e++) hb[e] = __ballot(p[e].dim1_val <= q && q <= p[e].dim2_val);
    } else {
#pragma unroll
      for (int e = 0; e < kIpsElems; e++) {
        const uint64_t i = first + static_cast<uint64_t>(e) * kWave + lane;
        bool h = false;
        if (i < n) {
          p[e] = pts[i];
          h = p[e].dim1_val <= q && q <= p[e].dim2_val;
        }
        hb[e] = __ballot(h);
      }
    }
    uint32_t hits = 0;
#pragma unroll
    for (int e = 0; e < kIpsElems; e++) hits += static_cast<uint32_t>(__builtin_popcountll(hb[e]));
    if (hits == 0) continue;
    if (hits >= kIpsDirect) {
      // dense chunk: its own range, written straight from the registers
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(counter, static_cast<unsigned long long>(hits));
      base = __shfl(base, 0, kWave);
      uint32_t run = 0;
#pragma unroll
      for (int e = 0; e < kIpsElems; e++) {
        if ((hb[e] >> lane) & 1ull) {
          const unsigned long long at = base + run + lanes_below(hb[e]);
          if (at < cap) o[at] = rec_of(p[e]);
        }
        run += static_cast<uint32_t>(__builtin_popcountll(hb[e]));
      }
      continue;
    }
    if (staged + hits > kIpsStage) flush();
    uint32_t run = staged;
#pragma unroll
    for (int e = 0; e < kIpsElems; e++) {
      if ((hb[e] >> lane) & 1ull) stage[run + lanes_below(hb[e])] = rec_of(p[e]);
      run += static_cast<uint32_t>(__builtin_popcountll(hb[e]));
    }
    staged = run;
  }
  if (staged) flush();
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

}  // namespace

// counter_d must be zero on entry; afterwards it holds the number of hits
// (results beyond `cap` are counted but not written).
Status ips_search(int dtype, int id_dtype, const void* pts_d, size_t n, int ret_point, const void* search_val,
                  void* out_d, unsigned long long cap, unsigned long long* counter_d, unsigned num_blocks,
                  hipStream_t stream) {
  if (n == 0) return Status{};
  return with_t(dtype, [&](auto t) -> Status {
    using T = decltype(t);
    return with_is(id_dtype, [&](auto i) -> Status {
      using IS = decltype(i);
      T q;
      __builtin_memcpy(&q, search_val, sizeof(T));
      unsigned grid = num_blocks;
      if (grid == 0) {
        // four workgroups (16 waves) per CU, fewer when there are fewer chunks
        int dev = 0, cus = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        const uint64_t chunks = (n + kIpsChunk - 1) / kIpsChunk;
        const uint64_t blocks = (chunks + kIpsBlock / kWave - 1) / (kIpsBlock / kWave);
        grid = static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(blocks, static_cast<uint64_t>(cus) * 4)));
      }
      const auto* p = static_cast<const Point<T, IS>*>(pts_d);
      if (ret_point) {
        k_ips<T, IS, true><<<grid, kIpsBlock, 0, stream>>>(p, n, q, out_d, cap, counter_d);
      } else {
        if constexpr (has_id_v<IS>)
          k_ips<T, IS, false><<<grid, kIpsBlock, 0, stream>>>(p, n, q, out_d, cap, counter_d);
        else
          return Status{1, "Error: IDs requested from points without IDs"};
      }
      return hip_status(hipGetLastError(), "Error in launching interval parallel search");
    });
  });
}

}  // namespace pst
