// 2-sided and 3-sided searches on one deep tree (StaticPSTGPU), for one
// query or a batch of them, on gfx950.
//
// The reference walks the tree with a single thread block, two block
// barriers per node step, an atomicCAS hand-off of split work to idle
// threads and dynamic-parallelism child grids for overflow
// (pst/static-pst-gpu-search-functions.tu:12-722, pst/static-pst-gpu.tu:
// 812-1148). Here every wavefront is an independent worker:
//
//   * work items are (query, slot, search code); a wave owns one query at a
//     time and walks it depth first with a stack of (slot, code) words in
//     LDS, popping up to 64 items per step (one per lane);
//   * the delegation rules per code are the reference's
//     (pst/static-pst-cpu-iter.tpp:600-735): REPORT_ABOVE reports every node
//     with dim2 >= min_dim2 below it, LEFT/RIGHT split into REPORT_ABOVE
//     plus one continuing child at the median, THREE splits into RIGHT +
//     LEFT when the median falls in [min_dim1, max_dim1];
//   * results are staged per wave in LDS and flushed with one atomic per
//     flush on the query's output cursor (the reference's warp scan +
//     atomicAdd, helpers/calc-alloc-report-ind-offset.h:17-44, becomes
//     ballot/mbcnt within the wave and one atomic per ~200 results);
//   * load balance: the launch is repeated for a few rounds; in all but the
//     last, a wave that spent its step budget on one batch pushes its
//     remaining stack to a shared list that the next round spreads over all
//     waves. The last round runs every stack to empty.
// Whether a child exists follows from n alone (pst_common.hpp), so no
// bitcode is read, and each node record carries its children's dim2, so
// pruned children are never pushed or loaded: every visit is a live node. A
// REPORT_ABOVE node needs no dim1/med load because every node below a
// REPORT_ABOVE split already satisfies the dim1 bounds.
#include <algorithm>

#include "pst_internal.hpp"

namespace pst {

namespace {

constexpr int kStageIds = 256;  // staged results per wave before a flush
constexpr int kFlushAt = kStageIds - 2 * kWave;
constexpr int kMaxRounds = 8;
constexpr uint64_t kEmpty = ~0ull;

template <typename T, typename IS>
struct Params {
  const KV<T, IS>* kv;
  const DM<T>* dm;
  uint64_t M;  // n + 1
  const Query<T>* queries;
  uint64_t* list[2];
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
  uint64_t S;               // slot count: items outside [0, S) are rejected
  uint32_t nq;              // query count: items of other queries are rejected
  uint32_t stack_cap;
  uint32_t total_waves;
  uint32_t budget_small, budget_large;
};

__device__ inline uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ inline uint32_t start_code(uint32_t kind) {
  return kind == KIND_THREE ? THREE_SEARCH : (kind == KIND_LEFT ? LEFT_SEARCH : RIGHT_SEARCH);
}

template <typename T>
__global__ void k_search_init(const Query<T>* __restrict__ queries, Query<T> single, uint32_t nq,
                              uint64_t* __restrict__ roots, unsigned long long* __restrict__ ctr,
                              uint32_t nctr, unsigned long long* __restrict__ counts,
                              unsigned long long* err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nq) {
    const uint32_t kind = queries ? queries[i].kind : single.kind;
    roots[i] = (static_cast<uint64_t>(i) << 40) | start_code(kind);
    counts[i] = 0;
  }
  if (i < nctr) ctr[i] = (i == 0) ? nq : 0;
  if (i == 0) *err = 0;
}

template <typename T, typename IS, bool RET_POINT>
struct OutRec {
  using type = typename std::conditional<RET_POINT, Point<T, IS>, IS>::type;
};

template <typename T, typename IS, bool RET_POINT, bool WRITE>
__global__ __launch_bounds__(kWave) void k_search_round(Params<T, IS> P, uint32_t round, uint32_t last) {
  using Rec = typename OutRec<T, IS, RET_POINT>::type;
  extern __shared__ uint32_t stack[];
  __shared__ uint64_t pend[kWave];
  __shared__ Rec stage[WRITE ? kStageIds : 1];

  const uint32_t lane = threadIdx.x;
  // Counters and lists written by the previous launch are read with
  // device-coherent loads (never through the scalar or L1 caches).
  const unsigned long long count_in =
      min(__hip_atomic_load(&P.ctr[2 * round], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), P.list_cap);
  if (P.dbg && lane == 0) {
    atomicMax(&P.dbg[round * 8 + 0], count_in);
    atomicMin(&P.dbg[round * 8 + 1], count_in);
  }
  unsigned long long* grab = &P.ctr[2 * round + 1];
  unsigned long long* count_out = &P.ctr[2 * round + 2];
  const uint64_t* list_in = P.list[round & 1];
  uint64_t* list_out = P.list[(round + 1) & 1];

  unsigned long long chunk = count_in / P.total_waves;
  chunk = chunk < 1 ? 1 : (chunk > kWave ? kWave : chunk);
  const uint32_t budget =
      last ? 0xffffffffu : (count_in < P.total_waves ? P.budget_small : P.budget_large);

  uint32_t npend = 0;  // uniform: valid pending entries are pend[0..npend) != kEmpty
  while (true) {
    // ---------- pick the next batch: the pending items of one query ----------
    const uint64_t mine = lane < npend ? pend[lane] : kEmpty;
    const uint64_t have = __ballot(mine != kEmpty);
    if (have == 0) {
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(grab, chunk);
      base = __shfl(base, 0, kWave);
      if (base >= count_in) break;
      const unsigned long long cnt = min(chunk, count_in - base);
      uint64_t it = kEmpty;
      if (lane < cnt)
        it = __hip_atomic_load(&list_in[base + lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // reject items that cannot come from this search
      if (it != kEmpty && (static_cast<uint32_t>(it >> 40) >= P.nq || ((it & 0xffffffffull) >> 2) >= P.S)) {
        atomicOr(P.err, 2ull);
        if (P.dbg) atomicAdd(&P.dbg[round * 8 + 3], 1ull);
        it = kEmpty;
      }
      pend[lane] = it;
      npend = static_cast<uint32_t>(cnt);
      if (P.dbg && lane == 0) atomicAdd(&P.dbg[round * 8 + 2], cnt);
      __builtin_amdgcn_wave_barrier();
      continue;
    }
    const int first = __builtin_ctzll(have);
    const uint64_t head = __shfl(mine, first, kWave);
    const uint32_t q = uniform(static_cast<uint32_t>(head >> 40));
    const bool take = mine != kEmpty && static_cast<uint32_t>(mine >> 40) == q;
    const uint64_t tmask = __ballot(take);
    if (take) {
      stack[lanes_below(tmask)] = static_cast<uint32_t>(mine);
      pend[lane] = kEmpty;
    }
    uint32_t sp = __popcll(tmask);
    __builtin_amdgcn_wave_barrier();

    const Query<T> Q = P.queries ? P.queries[q] : P.single;
    const uint32_t kind = uniform(Q.kind);
    unsigned long long found = 0;  // results of this batch (count mode)
    uint32_t nstage = 0;           // staged results (write mode)
    unsigned long long off_q = 0, cap_q = 0;
    if constexpr (WRITE) {
      off_q = P.offsets ? P.offsets[q] : 0;
      cap_q = P.caps ? P.caps[q] : P.cap_all;
    }
    auto flush = [&]() {
      if constexpr (WRITE) {
        if (nstage == 0) return;
        unsigned long long b = 0;
        if (lane == 0) b = atomicAdd(&P.counts[q], static_cast<unsigned long long>(nstage));
        b = __shfl(b, 0, kWave);
        Rec* out = static_cast<Rec*>(P.out) + off_q;
        for (uint32_t x = lane; x < nstage; x += kWave)
          if (b + x < cap_q && off_q + b + x < P.out_total) out[b + x] = stage[x];
        nstage = 0;
        __builtin_amdgcn_wave_barrier();
      }
    };

    uint32_t steps = 0;
    while (sp > 0) {
      // children (<= 2 per popped item) must fit: sp - take + 2 take <= cap
      const uint32_t room = P.stack_cap - sp;
      uint32_t takecnt = sp < kWave ? sp : kWave;
      if (takecnt > room) takecnt = room;
      if (takecnt == 0) {
        if (lane == 0) atomicOr(P.err, 1ull);
        sp = 0;
        break;
      }
      sp -= takecnt;
      const bool act = lane < takecnt;
      const uint32_t item = act ? stack[sp + lane] : 0;
      __builtin_amdgcn_wave_barrier();
      const uint64_t slot = item >> 2;
      const uint32_t code = item & 3u;
      bool rep = false, cl = false, cr = false;
      uint32_t codel = REPORT_ABOVE, coder = REPORT_ABOVE;
      KV<T, IS> a{};
      DM<T> b{};
      bool ok = act;
      if (act && (slot >= P.S || subtree_size(P.M, slot) < 1)) {
        atomicOr(P.err, 2ull);
        ok = false;
      }
      if (ok) {
        a = P.kv[slot];
        if (RET_POINT || code != REPORT_ABOVE) b = P.dm[slot];
        if (Q.min_dim2 <= a.dim2) {
          // a child is pushed only if it exists and survives the dim2 bound,
          // which its parent's record tells without loading the child
          const int64_t m = subtree_size(P.M, slot);
          const bool hl = m >= 2 && Q.min_dim2 <= a.cdim2[0];
          const bool hr = m >= 3 && Q.min_dim2 <= a.cdim2[1];
          if (code == REPORT_ABOVE) {
            rep = true;
            cl = hl;
            cr = hr;
          } else {
            const T d1 = b.dim1, md = b.med;
            if (kind == KIND_THREE) rep = Q.min_dim1 <= d1 && d1 <= Q.max_dim1;
            else if (kind == KIND_LEFT) rep = d1 <= Q.max_dim1;
            else rep = d1 >= Q.min_dim1;
            if (code == LEFT_SEARCH) {
              if (md <= Q.max_dim1) {
                cl = hl; codel = REPORT_ABOVE;
                cr = hr; coder = LEFT_SEARCH;
              } else {
                cl = hl; codel = LEFT_SEARCH;
              }
            } else if (code == RIGHT_SEARCH) {
              if (md >= Q.min_dim1) {
                cl = hl; codel = RIGHT_SEARCH;
                cr = hr; coder = REPORT_ABOVE;
              } else {
                cr = hr; coder = RIGHT_SEARCH;
              }
            } else {  // THREE_SEARCH
              if (Q.min_dim1 <= md && md <= Q.max_dim1) {
                cl = hl; codel = RIGHT_SEARCH;
                cr = hr; coder = LEFT_SEARCH;
              } else if (Q.max_dim1 < md) {
                cl = hl; codel = THREE_SEARCH;
              } else if (md < Q.min_dim1) {
                cr = hr; coder = THREE_SEARCH;
              }
            }
          }
        }
      }
      // ---- results
      const uint64_t rmask = __ballot(rep);
      if constexpr (WRITE) {
        if (rep) {
          Rec r;
          if constexpr (RET_POINT) {
            r.dim1_val = b.dim1;
            r.dim2_val = a.dim2;
            if constexpr (has_id_v<IS>) r.id = a.id;
          } else {
            r = a.id;
          }
          stage[nstage + lanes_below(rmask)] = r;
        }
        nstage += __popcll(rmask);
        __builtin_amdgcn_wave_barrier();
        if (nstage >= kFlushAt) flush();
      } else {
        found += __popcll(rmask);
      }
      // ---- children
      const uint64_t lm = __ballot(cl), rm = __ballot(cr);
      const uint32_t nl = __popcll(lm);
      if (cl) stack[sp + lanes_below(lm)] = static_cast<uint32_t>(((2 * slot + 1) << 2) | codel);
      if (cr) stack[sp + nl + lanes_below(rm)] = static_cast<uint32_t>(((2 * slot + 2) << 2) | coder);
      sp += nl + __popcll(rm);
      __builtin_amdgcn_wave_barrier();
      // ---- hand remaining work to the next round when over budget
      if (++steps >= budget && sp > 0) {
        // Reservations only grow the counter; one that ends past the list
        // capacity fails, marks its in-capacity part empty (the next round
        // skips empty entries and reads at most list_cap of them) and the
        // wave keeps walking its batch instead.
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(count_out, static_cast<unsigned long long>(sp));
        base = __shfl(base, 0, kWave);
        if (base + sp <= P.list_cap) {
          for (uint32_t x = lane; x < sp; x += kWave)
            list_out[base + x] = (static_cast<uint64_t>(q) << 40) | stack[x];
          if (P.dbg && lane == 0) {
            atomicAdd(&P.dbg[round * 8 + 4], 1ull);
            atomicAdd(&P.dbg[round * 8 + 5], static_cast<unsigned long long>(sp));
          }
          sp = 0;
        } else {
          for (unsigned long long x = base + lane; x < P.list_cap && x < base + sp; x += kWave)
            list_out[x] = kEmpty;
          steps = 0;
        }
      }
    }
    if constexpr (WRITE) {
      flush();
    } else {
      if (lane == 0 && found) atomicAdd(&P.counts[q], found);
    }
  }
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

template <typename T, typename IS, bool RP, bool W>
Status launch_rounds(const SearchArgs& a, SearchScratch& sc, hipStream_t s) {
  Params<T, IS> P{};
  P.kv = static_cast<const KV<T, IS>*>(a.kv);
  P.dm = static_cast<const DM<T>*>(a.dm);
  P.M = a.n + 1;
  P.queries = static_cast<const Query<T>*>(a.queries);
  if (a.single_query) P.single = *static_cast<const Query<T>*>(a.single_query);
  P.list[0] = static_cast<uint64_t*>(sc.list[0]);
  P.list[1] = static_cast<uint64_t*>(sc.list[1]);
  P.list_cap = sc.list_cap;
  P.ctr = sc.ctr;
  P.counts = a.counts;
  P.offsets = a.offsets;
  P.caps = a.caps;
  P.cap_all = a.cap_all;
  P.out_total = a.out_total ? a.out_total : ~0ull;
  P.out = a.out;
  P.err = sc.ctr + 2 * kMaxRounds + 2;
  P.dbg = a.debug;
  P.S = (a.n == 0) ? 0 : ((1ull << (64 - __builtin_clzll(a.n))) - 1);
  P.nq = static_cast<uint32_t>(a.nq);
  uint32_t H = 0;
  while ((1ull << H) <= a.n) H++;  // levels
  P.stack_cap = static_cast<uint32_t>(kWave * (H + 2) + 2 * kWave);
  P.budget_small = a.budget_small > 0 ? a.budget_small : 6;
  P.budget_large = a.budget_large > 0 ? a.budget_large : 384;
  const size_t lds = P.stack_cap * sizeof(uint32_t);
  auto kern = k_search_round<T, IS, RP, W>;
  int dev = 0;
  hipGetDevice(&dev);
  int cus = 0;
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  int occ = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kWave, lds);
  if (e != hipSuccess || occ <= 0) occ = 8;
  if (a.waves_per_cu > 0 && a.waves_per_cu < occ) occ = a.waves_per_cu;
  const uint32_t grid = static_cast<uint32_t>(cus * occ);
  P.total_waves = grid;
  const int R = std::max(1, std::min(a.rounds, kMaxRounds));
  for (int r = 0; r < R; r++) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kWave), lds, s, P, static_cast<uint32_t>(r),
                       static_cast<uint32_t>(r == R - 1));
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

}  // namespace

Status scratch_reserve(SearchScratch& sc, size_t nq, int dtype, hipStream_t) {
  const size_t spill_cap = 1ull << 22;  // 4M items per round list
  const size_t need = std::max(nq, spill_cap);
  if (sc.list_cap < need) {
    if (sc.list[0]) hipFree(sc.list[0]);
    if (sc.list[1]) hipFree(sc.list[1]);
    sc.list[0] = sc.list[1] = nullptr;
    sc.list_cap = 0;
    hipError_t e = hipMalloc(&sc.list[0], need * 8);
    if (e == hipSuccess) e = hipMalloc(&sc.list[1], need * 8);
    if (e != hipSuccess) return hip_status(e, "Error in allocating search work lists");
    sc.list_cap = need;
  }
  if (!sc.ctr) {
    hipError_t e = hipMalloc(&sc.ctr, (2 * kMaxRounds + 4) * sizeof(unsigned long long));
    if (e != hipSuccess) return hip_status(e, "Error in allocating search counters");
  }
  const size_t qb = 32;
  if (sc.qbuf_bytes < qb) {
    if (sc.qbuf) hipFree(sc.qbuf);
    hipError_t e = hipMalloc(&sc.qbuf, qb);
    if (e != hipSuccess) return hip_status(e, "Error in allocating query record");
    sc.qbuf_bytes = qb;
  }
  if (sc.cursor_cap < std::max<size_t>(nq, 1)) {
    if (sc.cursor) hipFree(sc.cursor);
    hipError_t e = hipMalloc(&sc.cursor, std::max<size_t>(nq, 1) * 8);
    if (e != hipSuccess) return hip_status(e, "Error in allocating result cursors");
    sc.cursor_cap = std::max<size_t>(nq, 1);
  }
  (void)dtype;
  return Status{};
}

void scratch_release(SearchScratch& sc) {
  if (sc.list[0]) hipFree(sc.list[0]);
  if (sc.list[1]) hipFree(sc.list[1]);
  if (sc.ctr) hipFree(sc.ctr);
  if (sc.qbuf) hipFree(sc.qbuf);
  if (sc.cursor) hipFree(sc.cursor);
  sc = SearchScratch{};
}

template <typename T>
Status init_typed(const SearchArgs& a, SearchScratch& sc, hipStream_t s) {
  Query<T> one{};
  if (a.single_query) one = *static_cast<const Query<T>*>(a.single_query);
  const uint32_t nctr = 2 * kMaxRounds + 2;
  const uint32_t work = static_cast<uint32_t>(std::max<size_t>(a.nq, nctr));
  k_search_init<T><<<(work + 255) / 256, 256, 0, s>>>(
      a.single_query ? nullptr : static_cast<const Query<T>*>(a.queries), one, static_cast<uint32_t>(a.nq),
      static_cast<uint64_t*>(sc.list[0]), sc.ctr, nctr, a.counts, sc.ctr + 2 * kMaxRounds + 2);
  return hip_status(hipGetLastError(), "Error in launching search initialisation");
}

// Enqueues the initialisation and the search rounds on `s` (asynchronous).
// A single query is passed by value in a.single_query (host Query<T>) and
// copied to the device by the initialisation kernel.
Status run_search(const SearchArgs& a_in, SearchScratch& sc, hipStream_t s) {
  SearchArgs a = a_in;
  if (a.nq == 0) return Status{};
  Status st = scratch_reserve(sc, a.nq, a.dtype, s);
  if (!st.ok()) return st;
  if (a.single_query) a.queries = nullptr;
  return with_t(a.dtype, [&](auto t) -> Status {
    using T = decltype(t);
    return with_is(a.id_dtype, [&](auto i) -> Status {
      using IS = decltype(i);
      if (!a.ret_point && a.mode == 1) {
        if constexpr (!has_id_v<IS>) return Status{1, "Error: IDs requested from a tree built without IDs"};
      }
      Status s0 = init_typed<T>(a, sc, s);
      if (!s0.ok()) return s0;
      if (a.mode == 0) return launch_rounds<T, IS, false, false>(a, sc, s);
      if (a.ret_point) return launch_rounds<T, IS, true, true>(a, sc, s);
      if constexpr (has_id_v<IS>) return launch_rounds<T, IS, false, true>(a, sc, s);
      return Status{1, "Error: IDs requested from a tree built without IDs"};
    });
  });
}

unsigned long long* search_error_flag(SearchScratch& sc) { return sc.ctr + 2 * kMaxRounds + 2; }

Status exclusive_scan(const unsigned long long* in, unsigned long long* out, size_t n, hipStream_t s) {
  k_exclusive_scan<<<1, kScanBlock, 0, s>>>(in, out, static_cast<uint32_t>(n));
  return hip_status(hipGetLastError(), "Error in launching offset scan");
}

}  // namespace pst
