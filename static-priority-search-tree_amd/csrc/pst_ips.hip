// Interval parallel search (Liu et al. 2016), the brute-force baseline the
// PST is compared with: report every point with dim1 <= q <= dim2
// (interval-parallel-search/interval-parallel-search.h:94-96 of the
// reference), as IDs or whole points, in unspecified order.
//
// The reference launches a grid-stride loop that, per tile, block-scans
// the hit flags through cooperative groups and takes one atomicAdd per
// tile (interval-parallel-search.h:70-117 with
// helpers/calc-alloc-report-ind-offset.h:46-133). Here: a streaming filter
// over the AoS array with 16 points per thread per tile (coalesced 12-byte
// records for f32/u32), wave64 ballot/mbcnt counts, one LDS scan across the
// block's four waves and one global atomic per 4096-point tile. The
// output counter is a per-call device word, not a global symbol.
#include <algorithm>

#include "pst_internal.hpp"

namespace pst {

namespace {

constexpr int kIpsBlock = 256;
constexpr int kIpsElems = 16;
constexpr uint32_t kIpsTile = kIpsBlock * kIpsElems;

template <typename T, typename IS, bool RET_POINT>
__global__ __launch_bounds__(kIpsBlock) void k_ips(const Point<T, IS>* __restrict__ pts, uint64_t n, T q,
                                                   void* __restrict__ out, unsigned long long cap,
                                                   unsigned long long* __restrict__ counter) {
  using Rec = typename std::conditional<RET_POINT, Point<T, IS>, IS>::type;
  __shared__ uint32_t wcount[kIpsBlock / kWave];
  __shared__ unsigned long long tile_base;
  const uint32_t tid = threadIdx.x;
  const uint32_t wave = tid / kWave;
  Rec* o = static_cast<Rec*>(out);
  for (uint64_t tile = static_cast<uint64_t>(blockIdx.x) * kIpsTile; tile < n;
       tile += static_cast<uint64_t>(gridDim.x) * kIpsTile) {
    Point<T, IS> p[kIpsElems];
    uint32_t hit = 0;
#pragma unroll
    for (int e = 0; e < kIpsElems; e++) {
      const uint64_t i = tile + static_cast<uint64_t>(e) * kIpsBlock + tid;
      if (i < n) {
        p[e] = pts[i];
        if (p[e].dim1_val <= q && q <= p[e].dim2_val) hit |= 1u << e;
      }
    }
    const uint32_t mine = __popc(hit);
    // exclusive scan of per-thread counts: within the wave by prefix sums
    // over lanes, across the block's waves through LDS
    uint32_t incl = mine;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t v = __shfl_up(incl, off, kWave);
      if (static_cast<int>(lane_id()) >= off) incl += v;
    }
    if (lane_id() == kWave - 1) wcount[wave] = incl;
    __syncthreads();
    uint32_t wbase = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kIpsBlock / kWave; w++) {
      const uint32_t c = wcount[w];
      wbase += w < static_cast<int>(wave) ? c : 0;
      total += c;
    }
    if (tid == 0) tile_base = total ? atomicAdd(counter, static_cast<unsigned long long>(total)) : 0;
    __syncthreads();
    unsigned long long pos = tile_base + wbase + (incl - mine);
#pragma unroll
    for (int e = 0; e < kIpsElems; e++) {
      if (hit & (1u << e)) {
        if (pos < cap) {
          if constexpr (RET_POINT) o[pos] = p[e];
          else o[pos] = p[e].id;
        }
        pos++;
      }
    }
    __syncthreads();  // wcount / tile_base reuse
  }
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
        int dev = 0, cus = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        const uint64_t tiles = (n + kIpsTile - 1) / kIpsTile;
        grid = static_cast<unsigned>(std::min<uint64_t>(tiles, static_cast<uint64_t>(cus) * 8));
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
