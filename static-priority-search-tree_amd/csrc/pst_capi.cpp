// extern "C" entry points of libpst_hip.so (declared in include/pst/pst_capi.h).
#include "pst_capi.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "pst_internal.hpp"

using pst::Status;

struct pst_tree {
  int dtype = 0, id_dtype = -1;
  size_t n = 0, S = 0;
  int dev_ind = 0, num_devs = 1, device = 0;
  void* nodes = nullptr;
  void* submin = nullptr;  // subtree minimum of dim2 per record slot
  void* ids = nullptr;     // ID per record slot (whole-subtree copies read these)
  size_t slots_per_tree = 0;  // forest (StaticPSTGPUArr): slots of a full tree; 0 for one deep tree
  void* tree_bounds = nullptr;  // forest: each tree's dim1 range (stationary batch passes)
  pst::SearchScratch sc;
  int rounds = 4, budget_small = 0, budget_large = 0, waves_per_cu = 0;
  int group = pst::kGroup;  // queries per walk in batched searches
  int copy_min = 256, copy_max = 65535;  // write pass: full subtrees copied in this size range
  size_t copy_tasks = 0;                 // copy task list entries (0: library default)
  void* lv_key = nullptr;  // level index (deep trees with IDs, on request): sorted dim2 keys per level
  void* lv_id = nullptr;   // and the IDs in the same order
  pst::LevelScratch lsc;
  hipStream_t stream = nullptr;
  unsigned long long* debug = nullptr;  // per-round search counters (debug builds)
  int graphs = 0;  // live batch graphs: they hold the scratch and index addresses
};

namespace {

thread_local std::string g_err;

int fail(const Status& s) {
  g_err = s.msg;
  return s.code ? s.code : PST_ERR_INTERNAL;
}
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

std::string dev_suffix(int dev_ind, int num_devs) {
  return " on device " + std::to_string(dev_ind + 1) + " (1-indexed) of " + std::to_string(num_devs) + ": ";
}

bool valid_t(int d) { return d >= PST_F32 && d <= PST_I64; }
bool valid_id(int d) { return d == PST_NO_ID || (d >= PST_F32 && d <= PST_U64); }

size_t id_size(int idt) {
  switch (idt) {
    case PST_NO_ID: return 0;
    case PST_F32: case PST_I32: case PST_U32: return 4;
    default: return 8;
  }
}

size_t slots_for(size_t n) {
  if (n == 0) return 0;
  size_t p = 1;
  while (p <= n) p <<= 1;
  return p - 1;
}

// Slots of a full forest tree: the smallest power of two above
// threads_per_block, less one (the reference's calcNumElemSlotsPerTree with
// minPowerOf2GreaterThan, pst/static-pst-gpu-arr.h:256-261).
size_t forest_slots_per_tree(unsigned tpb) {
  size_t p = 1;
  while (p <= tpb) p <<= 1;
  return p - 1;
}

// Per-device counter used by the blocking interval search.
std::mutex g_ctr_mu;
std::map<int, unsigned long long*> g_ctr;
unsigned long long* device_counter(int dev) {
  std::lock_guard<std::mutex> lk(g_ctr_mu);
  auto it = g_ctr.find(dev);
  if (it != g_ctr.end()) return it->second;
  unsigned long long* p = nullptr;
  if (hipMalloc(&p, sizeof(unsigned long long)) != hipSuccess) return nullptr;
  g_ctr[dev] = p;
  return p;
}

template <typename T>
void make_query(void* dst, uint32_t kind, const void* a, const void* b, const void* c) {
  pst::Query<T> q{};
  if (a) std::memcpy(&q.min_dim1, a, sizeof(T));
  if (b) std::memcpy(&q.max_dim1, b, sizeof(T));
  if (c) std::memcpy(&q.min_dim2, c, sizeof(T));
  q.kind = kind;
  std::memcpy(dst, &q, sizeof(q));
}

void fill_query(int dt, void* dst, uint32_t kind, const void* a, const void* b, const void* c) {
  switch (dt) {
    case PST_F32: make_query<float>(dst, kind, a, b, c); break;
    case PST_F64: make_query<double>(dst, kind, a, b, c); break;
    case PST_I32: make_query<int32_t>(dst, kind, a, b, c); break;
    default: make_query<int64_t>(dst, kind, a, b, c); break;
  }
}

pst::SearchArgs base_args(pst_tree* t) {
  pst::SearchArgs a{};
  a.dtype = t->dtype;
  a.id_dtype = t->id_dtype;
  a.n = t->n;
  a.nodes = t->nodes;
  a.submin = t->submin;
  a.ids = t->ids;
  a.slots_per_tree = t->slots_per_tree;
  a.tree_bounds = t->tree_bounds;
  a.rounds = t->rounds;
  a.budget_small = t->budget_small;
  a.budget_large = t->budget_large;
  a.waves_per_cu = t->waves_per_cu;
  a.group = t->group;
  a.copy_min = t->copy_min;
  a.copy_max = t->copy_max;
  a.copy_tasks = t->copy_tasks;
  a.debug = t->debug;
  return a;
}

// One query, results written to out (capacity cap), blocking; *num = count.
int single_search(pst_tree* t, int kind, int ret, const void* a, const void* b, const void* c, void* out,
                  size_t cap, size_t* num, hipStream_t s) {
  alignas(16) unsigned char qrec[32];
  fill_query(t->dtype, qrec, static_cast<uint32_t>(kind), a, b, c);
  pst::SearchArgs args = base_args(t);
  args.single_query = qrec;
  args.nq = 1;
  args.mode = 1;
  args.ret_point = ret;
  Status st = pst::scratch_reserve(t->sc, 1, 0, s);
  if (!st.ok()) return fail(st);
  args.counts = t->sc.cursor;
  args.cap_all = cap;
  args.out = out;
  const bool via_index = t->lv_key && ret != PST_RET_POINT;
  st = via_index ? pst::level_search(args, t->lv_key, t->lv_id, t->lsc, s) : pst::run_search(args, t->sc, s);
  if (!st.ok()) return fail(st);
  unsigned long long host[2] = {0, 0};
  hipError_t e = hipMemcpyAsync(&host[0], t->sc.cursor, 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && !via_index)
    e = hipMemcpyAsync(&host[1], pst::search_error_flag(t->sc), 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess)
    return fail(PST_ERR_HIP, "Error in copying global result array final index from device" +
                                 dev_suffix(t->dev_ind, t->num_devs) + hipGetErrorString(e));
  if (host[1] & 1ull) return fail(PST_ERR_INTERNAL, "Error: search stack overflow");
  if (host[1] & 2ull) return fail(PST_ERR_INTERNAL, "Error: search read an invalid work item");
  *num = host[0];
  if (host[0] > cap) return fail(PST_ERR_CAPACITY, "Error: search results exceed the output capacity");
  return PST_OK;
}

int blocking_search(pst_tree* t, int kind, int ret, size_t* num, void** res, const void* a, const void* b,
                    const void* c) {
  if (!t || !num || !res) return fail(PST_ERR_INVALID, "Error: null argument");
  if (ret != PST_RET_POINT && t->id_dtype == PST_NO_ID)
    return fail(PST_ERR_INVALID, "Error: IDs requested from a tree built without IDs");
  if (t->n == 0) {
    std::printf("Tree is empty; nothing to search\n");
    std::fflush(stdout);
    *num = 0;
    *res = nullptr;
    return PST_OK;
  }
  hipSetDevice(t->device);
  const size_t rec = ret == PST_RET_POINT ? pst_point_size(t->dtype, t->id_dtype) : id_size(t->id_dtype);
  void* out = nullptr;
  hipError_t e = hipMalloc(&out, t->n * rec);
  if (e != hipSuccess)
    return fail(PST_ERR_HIP, "Error in allocating array to store PointStruct search result" +
                                 dev_suffix(t->dev_ind, t->num_devs) + hipGetErrorString(e));
  int rc = single_search(t, kind, ret, a, b, c, out, t->n, num, t->stream);
  if (rc != PST_OK) {
    hipFree(out);
    *res = nullptr;
    return rc;
  }
  *res = out;
  return PST_OK;
}

// Forest batches small enough for the stationary two-pass search
// (pst_forest.hip) take it instead of per-tree walks.
bool forest_dense(const pst_tree* t, size_t nq) {
  return t->slots_per_tree && t->slots_per_tree <= pst::kForestDenseMaxSlots && nq <= pst::kForestDenseMax;
}

// Grows the handle's per-(query, segment) buffer of forest batches.
Status forest_seg_reserve(pst_tree* t, size_t nq) {
  int dev = 0, cus = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const size_t need = std::max<size_t>(pst::forest_dense_scratch_bytes(t->n, t->slots_per_tree, nq, cus), 16);
  if (t->sc.fseg_bytes >= need) return Status{};
  if (t->sc.pinned) return Status{1, pst::kPinnedScratchMsg};
  if (t->sc.fseg) hipFree(t->sc.fseg);
  t->sc.fseg = nullptr;
  t->sc.fseg_bytes = 0;
  Status st = pst::hip_status(hipMalloc(&t->sc.fseg, need), "Error in allocating forest search buffers");
  if (st.ok()) t->sc.fseg_bytes = need;
  return st;
}

// ---- the allocation-free batched search (pst_gpu_batch_search_noalloc) ----
// Its search buffers live in one file-scope device array per GPU, sized for
// kNoAllocMaxQueries queries through the level index: 128 ranges per query
// (8-byte source + 4-byte output position), the range count, IDs written,
// their prefix and the write cursor per query, and the copy chunk table.
constexpr size_t kNoAllocMaxQueries = 8192;
constexpr size_t kNoAllocRanges = 128;      // ranges per query of a level search
constexpr size_t kNoAllocChunks = 1u << 20;  // chunk table entries of a level search
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t{255}; }
constexpr size_t kNoAllocTsrc = 0;
constexpr size_t kNoAllocTpre = kNoAllocTsrc + align256(kNoAllocMaxQueries * kNoAllocRanges * 8);
constexpr size_t kNoAllocNtask = kNoAllocTpre + align256(kNoAllocMaxQueries * kNoAllocRanges * 4);
constexpr size_t kNoAllocWq = kNoAllocNtask + align256(kNoAllocMaxQueries * 4);
constexpr size_t kNoAllocWoff = kNoAllocWq + align256(kNoAllocMaxQueries * 8);
constexpr size_t kNoAllocCursor = kNoAllocWoff + align256((kNoAllocMaxQueries + 1) * 8);
constexpr size_t kNoAllocChunk = kNoAllocCursor + align256(kNoAllocMaxQueries * 8);
constexpr size_t kNoAllocBytes = kNoAllocChunk + align256(kNoAllocChunks * 4);  // 16.25 MiB

// In host code this name is only a stand-in for the symbol: its device
// address comes from hipGetSymbolAddress (noalloc_scratch below).
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch;  // device ordinal -> that device's copy of `scratch`

// The device address of `scratch` on `device` (looked up once per device,
// with that device current); nullptr and a message in g_err on failure.
unsigned char* noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(device);
  if (it != g_scratch.end()) return it->second;
  int prev = 0;
  hipGetDevice(&prev);
  if (prev != device && hipSetDevice(device) != hipSuccess) {
    g_err = "Error in selecting device " + std::to_string(device) + " for the search scratch";
    return nullptr;
  }
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (prev != device) hipSetDevice(prev);
  if (e != hipSuccess || !ptr) {
    g_err = std::string("Error in locating the search scratch: ") + hipGetErrorString(e);
    return nullptr;
  }
  unsigned char* p = static_cast<unsigned char*>(ptr);
  g_scratch[device] = p;
  return p;
}

__global__ void k_zero_u64(unsigned long long* p, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

namespace pst {

Status hip_status(hipError_t e, const std::string& what) {
  Status s;
  if (e != hipSuccess) {
    int dev = 0;
    hipGetDevice(&dev);
    s.code = PST_ERR_HIP;
    s.msg = what + dev_suffix(dev, 1) + hipGetErrorString(e);
  }
  return s;
}

size_t dtype_size(int dt) { return (dt == PST_F64 || dt == PST_I64 || dt == PST_U64) ? 8 : 4; }

// sizeof(Node<T, IdStore>): { T dim1; T dim2; T med; [IDType id]; }
size_t node_size(int dt, int idt) {
  const size_t a = dtype_size(dt), b = id_size(idt);
  if (b == 0) return 3 * a;
  const size_t al = a > b ? a : b;
  const size_t off = ((3 * a + b - 1) / b) * b;
  return ((off + b + al - 1) / al) * al;
}

size_t point_size(int dt, int idt) {
  const size_t a = dtype_size(dt), b = id_size(idt);
  if (b == 0) return 2 * a;
  const size_t al = a > b ? a : b;
  const size_t off = ((2 * a + b - 1) / b) * b;
  return ((off + b + al - 1) / al) * al;
}

}  // namespace pst

extern "C" {

const char* pst_last_error(void) { return g_err.c_str(); }

size_t pst_point_size(int dtype, int id_dtype) {
  if (!valid_t(dtype) || !valid_id(id_dtype)) return 0;
  return pst::point_size(dtype, id_dtype);
}

int pst_free(void* p) {
  if (!p) return PST_OK;
  hipError_t e = hipFree(p);
  if (e != hipSuccess) return fail(pst::hip_status(e, "Error in freeing search result array"));
  return PST_OK;
}

int pst_gpu_create(const void* pt_arr_d, size_t n, int dtype, int id_dtype, int warp_multiplier, int dev_ind,
                   int num_devs, pst_stream_t stream, pst_tree** out) {
  (void)warp_multiplier;
  if (!out) return fail(PST_ERR_INVALID, "Error: null output handle");
  *out = nullptr;
  if (!valid_t(dtype) || !valid_id(id_dtype))
    return fail(PST_ERR_INVALID, "Error: unsupported data/ID type codes " + std::to_string(dtype) + "/" +
                                     std::to_string(id_dtype));
  if (n > 0 && !pt_arr_d) return fail(PST_ERR_INVALID, "Error: null point array");
  // search work items pack a slot into 30 bits: S = 2^ceil(log2(n+1)) - 1 < 2^30
  if (n >= (1ull << 29)) return fail(PST_ERR_INVALID, "Error: StaticPSTGPU holds fewer than 2^29 points");
  pst_tree* t = new pst_tree();
  t->dtype = dtype;
  t->id_dtype = id_dtype;
  t->n = n;
  t->S = slots_for(n);
  t->dev_ind = dev_ind;
  t->num_devs = num_devs;
  t->stream = reinterpret_cast<hipStream_t>(stream);
  hipGetDevice(&t->device);
  if (n == 0) {
    *out = t;
    return PST_OK;
  }
  // the tree's arrays are allocated by the build once its sorts are queued,
  // so that the allocation overlaps them
  pst::BuildOut bo;
  bo.alloc = [&](pst::BuildOut& o) {
    hipError_t e = hipMalloc(&t->nodes, t->S * pst::node_size(dtype, id_dtype));
    if (e == hipSuccess) e = hipMalloc(&t->submin, t->S * pst::dtype_size(dtype));
    if (e == hipSuccess && id_size(id_dtype) > 0) e = hipMalloc(&t->ids, t->S * id_size(id_dtype));
    if (e != hipSuccess)
      return Status{PST_ERR_HIP, "Error in allocating priority search tree storage array" +
                                     dev_suffix(dev_ind, num_devs) + hipGetErrorString(e)};
    o.nodes = t->nodes;
    o.submin = t->submin;
    o.ids = t->ids;
    return Status{};
  };
  Status st = pst::build_tree(dtype, id_dtype, pt_arr_d, n, bo, t->stream);
  if (!st.ok()) {
    pst_gpu_destroy(t);
    return fail(st);
  }
  *out = t;
  return PST_OK;
}

int pst_gpu_destroy(pst_tree* t) {
  if (!t) return PST_OK;
  if (t->graphs > 0)
    return fail(PST_ERR_INVALID, "Error: the tree still has batch graphs; destroy them before the tree");
  hipError_t e = hipSuccess;
  if (t->nodes) e = hipFree(t->nodes);
  if (t->submin) {
    hipError_t e2 = hipFree(t->submin);
    if (e == hipSuccess) e = e2;
  }
  if (t->ids) {
    hipError_t e2 = hipFree(t->ids);
    if (e == hipSuccess) e = e2;
  }
  if (t->tree_bounds) {
    hipError_t e2 = hipFree(t->tree_bounds);
    if (e == hipSuccess) e = e2;
  }
  pst::scratch_release(t->sc);
  if (t->lv_key) hipFree(t->lv_key);
  if (t->lv_id) hipFree(t->lv_id);
  pst::level_scratch_release(t->lsc);
  const int di = t->dev_ind, nd = t->num_devs;
  delete t;
  if (e != hipSuccess)
    return fail(PST_ERR_HIP, "Error in freeing array storing PST" + dev_suffix(di, nd) + hipGetErrorString(e));
  return PST_OK;
}

int pst_gpu_arr_create(const void* pt_arr_d, size_t n, int dtype, int id_dtype, unsigned threads_per_block,
                       int dev_ind, int num_devs, pst_stream_t stream, pst_tree** out) {
  if (!out) return fail(PST_ERR_INVALID, "Error: null output handle");
  *out = nullptr;
  if (!valid_t(dtype) || !valid_id(id_dtype))
    return fail(PST_ERR_INVALID, "Error: unsupported data/ID type codes " + std::to_string(dtype) + "/" +
                                     std::to_string(id_dtype));
  if (threads_per_block == 0 || threads_per_block > 1024)
    return fail(PST_ERR_INVALID, "Error: threads_per_block must be in [1, 1024]");
  if (n > 0 && !pt_arr_d) return fail(PST_ERR_INVALID, "Error: null point array");
  if (n >= (1ull << 29)) return fail(PST_ERR_INVALID, "Error: StaticPSTGPUArr holds fewer than 2^29 points");
  pst_tree* t = new pst_tree();
  t->dtype = dtype;
  t->id_dtype = id_dtype;
  t->n = n;
  t->slots_per_tree = forest_slots_per_tree(threads_per_block);
  const size_t ntrees = n ? (n + t->slots_per_tree - 1) / t->slots_per_tree : 0;
  t->S = ntrees * (t->slots_per_tree + 1);  // records: trees at a stride of slots_per_tree + 1
  t->dev_ind = dev_ind;
  t->num_devs = num_devs;
  t->stream = reinterpret_cast<hipStream_t>(stream);
  hipGetDevice(&t->device);
  if (n == 0) {
    *out = t;
    return PST_OK;
  }
  hipError_t e = hipMalloc(&t->nodes, t->S * pst::node_size(dtype, id_dtype));
  if (e == hipSuccess) e = hipMalloc(&t->submin, t->S * pst::dtype_size(dtype));
  if (e == hipSuccess && id_size(id_dtype) > 0) e = hipMalloc(&t->ids, t->S * id_size(id_dtype));
  if (e == hipSuccess) e = hipMalloc(&t->tree_bounds, pst::forest_bounds_bytes(dtype, n, t->slots_per_tree));
  if (e != hipSuccess) {
    std::string m = "Error in allocating array of PSTs on device " + std::to_string(dev_ind + 1) +
                    " (1-indexed) of " + std::to_string(num_devs) + ": " + hipGetErrorString(e);
    pst_gpu_destroy(t);
    return fail(PST_ERR_HIP, m);
  }
  Status st = pst::build_forest(dtype, id_dtype, pt_arr_d, n, t->slots_per_tree, t->nodes, t->stream);
  if (st.ok())
    st = pst::build_submin(dtype, id_dtype, n, t->slots_per_tree, t->nodes, t->submin, t->ids, t->stream);
  if (st.ok())
    st = pst::build_forest_bounds(dtype, id_dtype, n, t->slots_per_tree, t->nodes, t->tree_bounds, t->stream);
  if (st.ok()) st = pst::hip_status(hipStreamSynchronize(t->stream), "Error in synchronising after construction");
  if (!st.ok()) {
    pst_gpu_destroy(t);
    return fail(st);
  }
  *out = t;
  return PST_OK;
}

size_t pst_gpu_slots_per_tree(const pst_tree* t) { return t ? t->slots_per_tree : 0; }

int pst_gpu_save(const pst_tree* t, const char* path) {
  if (!t || !path) return fail(PST_ERR_INVALID, "Error: null tree or path");
  Status st = pst::save_tree_file(path, t->slots_per_tree ? 1 : 0, t->dtype, t->id_dtype, t->n, t->S,
                                  t->slots_per_tree, t->nodes, t->stream);
  if (!st.ok()) return fail(st);
  return PST_OK;
}

int pst_gpu_load(const char* path, int dev_ind, int num_devs, pst_stream_t stream, pst_tree** out) {
  if (!out || !path) return fail(PST_ERR_INVALID, "Error: null output handle or path");
  *out = nullptr;
  pst::TreeFileInfo info{};
  Status st = pst::read_tree_header(path, &info);
  if (!st.ok()) return fail(st);
  pst_tree* t = new pst_tree();
  t->dtype = info.dtype;
  t->id_dtype = info.id_dtype;
  t->n = info.n;
  t->S = info.S;
  t->slots_per_tree = info.slots_per_tree;
  t->dev_ind = dev_ind;
  t->num_devs = num_devs;
  t->stream = reinterpret_cast<hipStream_t>(stream);
  hipGetDevice(&t->device);
  if (t->n == 0) {
    *out = t;
    return PST_OK;
  }
  hipError_t e = hipMalloc(&t->nodes, t->S * pst::node_size(t->dtype, t->id_dtype));
  if (e == hipSuccess) e = hipMalloc(&t->submin, t->S * pst::dtype_size(t->dtype));
  if (e == hipSuccess && id_size(t->id_dtype) > 0) e = hipMalloc(&t->ids, t->S * id_size(t->id_dtype));
  if (e == hipSuccess && t->slots_per_tree)
    e = hipMalloc(&t->tree_bounds, pst::forest_bounds_bytes(t->dtype, t->n, t->slots_per_tree));
  if (e != hipSuccess) {
    std::string m = "Error in allocating priority search tree storage array" + dev_suffix(dev_ind, num_devs) +
                    hipGetErrorString(e);
    pst_gpu_destroy(t);
    return fail(PST_ERR_HIP, m);
  }
  st = pst::load_tree_file(path, info, t->nodes, t->stream);
  if (st.ok())
    st = pst::build_submin(t->dtype, t->id_dtype, t->n, t->slots_per_tree, t->nodes, t->submin, t->ids, t->stream);
  if (st.ok() && t->slots_per_tree)
    st = pst::build_forest_bounds(t->dtype, t->id_dtype, t->n, t->slots_per_tree, t->nodes, t->tree_bounds,
                                  t->stream);
  if (st.ok()) st = pst::hip_status(hipStreamSynchronize(t->stream), "Error in synchronising after loading");
  if (!st.ok()) {
    pst_gpu_destroy(t);
    return fail(st);
  }
  *out = t;
  return PST_OK;
}

int pst_read_vertices(const char* path, int dtype, size_t num_vertices, void* vertex_arr) {
  if (!path || (!vertex_arr && num_vertices)) return fail(PST_ERR_INVALID, "Error: null path or output");
  if (!valid_t(dtype)) return fail(PST_ERR_INVALID, "Error: unsupported data type code " + std::to_string(dtype));
  Status st = pst::read_raw(path, num_vertices * pst::dtype_size(dtype), vertex_arr, nullptr, nullptr);
  if (!st.ok()) return fail(st);
  return PST_OK;
}

int pst_read_vertices_to_device(const char* path, int dtype, size_t num_vertices, void* vertex_arr_d,
                                pst_stream_t stream) {
  if (!path || (!vertex_arr_d && num_vertices)) return fail(PST_ERR_INVALID, "Error: null path or output");
  if (!valid_t(dtype)) return fail(PST_ERR_INVALID, "Error: unsupported data type code " + std::to_string(dtype));
  Status st = pst::read_raw(path, num_vertices * pst::dtype_size(dtype), nullptr, vertex_arr_d,
                            reinterpret_cast<hipStream_t>(stream));
  if (!st.ok()) return fail(st);
  return PST_OK;
}
int pst_gpu_dtype(const pst_tree* t) { return t ? t->dtype : -2; }
int pst_gpu_id_dtype(const pst_tree* t) { return t ? t->id_dtype : -2; }

size_t pst_gpu_num_elems(const pst_tree* t) { return t ? t->n : 0; }
int pst_gpu_dev_ind(const pst_tree* t) { return t ? t->dev_ind : 0; }
int pst_gpu_num_devs(const pst_tree* t) { return t ? t->num_devs : 0; }

int pst_gpu_three_sided_search(pst_tree* t, int ret, size_t* num, void** res, const void* min1, const void* max1,
                               const void* min2) {
  return blocking_search(t, PST_THREE_SIDED, ret, num, res, min1, max1, min2);
}
int pst_gpu_two_sided_left_search(pst_tree* t, int ret, size_t* num, void** res, const void* max1,
                                  const void* min2) {
  return blocking_search(t, PST_TWO_SIDED_LEFT, ret, num, res, nullptr, max1, min2);
}
int pst_gpu_two_sided_right_search(pst_tree* t, int ret, size_t* num, void** res, const void* min1,
                                   const void* min2) {
  return blocking_search(t, PST_TWO_SIDED_RIGHT, ret, num, res, min1, nullptr, min2);
}

int pst_gpu_search_into(pst_tree* t, int kind, int ret, const void* min1, const void* max1, const void* min2,
                        void* out_d, size_t capacity, size_t* num, pst_stream_t stream) {
  if (!t || !num) return fail(PST_ERR_INVALID, "Error: null argument");
  if (kind < 0 || kind > 2) return fail(PST_ERR_INVALID, "Error: unknown query kind");
  if (ret != PST_RET_POINT && t->id_dtype == PST_NO_ID)
    return fail(PST_ERR_INVALID, "Error: IDs requested from a tree built without IDs");
  if (t->n == 0) {
    *num = 0;
    return PST_OK;
  }
  return single_search(t, kind, ret, min1, max1, min2, out_d, capacity, num,
                       reinterpret_cast<hipStream_t>(stream));
}

size_t pst_gpu_calc_global_mem_needed(size_t n, int dtype, int id_dtype) {
  if (!valid_t(dtype) || !valid_id(id_dtype)) return 0;
  if (n == 0) return 0;
  const size_t S = slots_for(n);
  // resident: node records, subtree minima and the per-slot ID array
  const size_t tree = S * (pst::node_size(dtype, id_dtype) + pst::dtype_size(dtype) + id_size(id_dtype));
  // construction's buffers, freed when it ends
  const size_t construct = pst::build_scratch_bytes(dtype, id_dtype, n);
  // one blocking search: the handle's search buffers (kept) and a result
  // array of n points (the reference's allocation, pst/static-pst-gpu.tu:563)
  int dev = 0, cus = 0;
  hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  const size_t search = pst::search_scratch_bytes(n, dtype, cus) + n * pst::point_size(dtype, id_dtype);
  return tree + (construct > search ? construct : search);
}

size_t pst_gpu_arr_calc_global_mem_needed(size_t n, unsigned threads_per_block, int dtype, int id_dtype) {
  if (!valid_t(dtype) || !valid_id(id_dtype) || threads_per_block == 0 || threads_per_block > 1024) return 0;
  const size_t St = forest_slots_per_tree(threads_per_block);
  const size_t ntrees = n ? (n + St - 1) / St : 0;
  const size_t tree =
      ntrees * (St + 1) * (pst::node_size(dtype, id_dtype) + pst::dtype_size(dtype) + id_size(id_dtype)) +
      pst::forest_bounds_bytes(dtype, n, St);
  // construction works in LDS; a blocking search needs a result array of n
  // points and the handle's walk buffers (work lists, copy tasks and the
  // default grid's overflow stacks, sized as for a deep tree of n points,
  // which is at least as deep as a forest tree); a stationary batch of up
  // to 512 queries its per-(query, segment) counters, the stabbing table
  // and 4 B per slot; all for the current device
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  const size_t search = n ? n * pst::point_size(dtype, id_dtype) + pst::search_scratch_bytes(n, dtype, cus) +
                                pst::forest_dense_scratch_bytes(n, St, pst::kForestDenseMax, cus)
                          : 0;
  return tree + search;
}

// Reference-layout slot count of the export: one tree's S, or for a forest
// the full trees' slots plus the last tree's own S.
static size_t export_slots(const pst_tree* t) {
  if (t->slots_per_tree == 0 || t->n == 0) return t->S;
  const size_t St = t->slots_per_tree;
  const size_t ntrees = (t->n + St - 1) / St;
  return (ntrees - 1) * St + slots_for(t->n - (ntrees - 1) * St);
}

size_t pst_gpu_export_bytes(const pst_tree* t) {
  if (!t) return 0;
  const size_t sT = pst::dtype_size(t->dtype), sI = id_size(t->id_dtype);
  const size_t S = export_slots(t);
  size_t after = 3 * S * sT;
  size_t off_id = sI ? ((after + sI - 1) / sI) * sI : after;
  size_t off_bits = off_id + S * sI;
  return ((off_bits + S + 7) / 8) * 8;
}

int pst_gpu_export_tree(pst_tree* t, void* host_blob, size_t blob_bytes, size_t* off6) {
  if (!t || !off6) return fail(PST_ERR_INVALID, "Error: null argument");
  const size_t sT = pst::dtype_size(t->dtype), sI = id_size(t->id_dtype);
  const size_t S = export_slots(t);
  off6[0] = 0;
  off6[1] = S * sT;
  off6[2] = 2 * S * sT;
  const size_t after = 3 * S * sT;
  off6[3] = sI ? ((after + sI - 1) / sI) * sI : after;
  off6[4] = off6[3] + S * sI;
  off6[5] = S;
  const size_t need = pst_gpu_export_bytes(t);
  if (blob_bytes < need) return fail(PST_ERR_INVALID, "Error: export buffer too small");
  if (t->n == 0) return PST_OK;
  void* d = nullptr;
  hipError_t e = hipMalloc(&d, need);
  if (e != hipSuccess) return fail(pst::hip_status(e, "Error in allocating tree export buffer"));
  Status st = t->slots_per_tree
                  ? pst::export_forest(t->dtype, t->id_dtype, t->n, t->slots_per_tree, t->nodes, d, S, off6[0],
                                       off6[1], off6[2], off6[3], off6[4], t->stream)
                  : pst::export_tree(t->dtype, t->id_dtype, t->n, t->nodes, d, off6[0], off6[1], off6[2], off6[3],
                                     off6[4], t->stream);
  if (st.ok()) st = pst::hip_status(hipMemcpyAsync(host_blob, d, need, hipMemcpyDeviceToHost, t->stream),
                                    "Error in copying tree export");
  if (st.ok()) st = pst::hip_status(hipStreamSynchronize(t->stream), "Error in synchronising tree export");
  hipFree(d);
  if (!st.ok()) return fail(st);
  return PST_OK;
}

int pst_gpu_batch_count(pst_tree* t, const void* queries_d, size_t nq, unsigned long long* counts_d,
                        pst_stream_t stream) {
  if (!t) return fail(PST_ERR_INVALID, "Error: null tree");
  if (nq == 0) return PST_OK;
  if (nq >= (1ull << 24)) return fail(PST_ERR_INVALID, "Error: batches hold fewer than 2^24 queries");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (t->n == 0) {
    Status st = pst::hip_status(hipMemsetAsync(counts_d, 0, nq * 8, s), "Error in clearing counts");
    return st.ok() ? PST_OK : fail(st);
  }
  pst::SearchArgs a = base_args(t);
  a.queries = queries_d;
  a.nq = nq;
  a.mode = 0;
  a.ret_point = 0;
  a.counts = counts_d;
  Status st;
  if (forest_dense(t, nq)) {
    st = forest_seg_reserve(t, nq);
    if (st.ok()) st = pst::forest_dense_search(a, t->sc.fseg, s, 0);
  } else {
    st = t->lv_key ? pst::level_search(a, t->lv_key, t->lv_id, t->lsc, s) : pst::run_search(a, t->sc, s);
  }
  return st.ok() ? PST_OK : fail(st);
}

int pst_gpu_batch_write(pst_tree* t, const void* queries_d, size_t nq, int ret, const unsigned long long* offsets_d,
                        const unsigned long long* caps_d, void* out_d, pst_stream_t stream) {
  if (!t) return fail(PST_ERR_INVALID, "Error: null tree");
  if (nq == 0 || t->n == 0) return PST_OK;
  if (nq >= (1ull << 24)) return fail(PST_ERR_INVALID, "Error: batches hold fewer than 2^24 queries");
  if (ret != PST_RET_POINT && t->id_dtype == PST_NO_ID)
    return fail(PST_ERR_INVALID, "Error: IDs requested from a tree built without IDs");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  Status st = pst::scratch_reserve(t->sc, nq, 0, s);
  if (!st.ok()) return fail(st);
  pst::SearchArgs a = base_args(t);
  a.queries = queries_d;
  a.nq = nq;
  a.mode = 1;
  a.ret_point = ret;
  a.counts = t->sc.cursor;
  a.offsets = offsets_d;
  a.caps = caps_d;
  a.cap_all = ~0ull;
  a.out = out_d;
  if (forest_dense(t, nq)) {
    st = forest_seg_reserve(t, nq);
    if (st.ok()) st = pst::forest_dense_search(a, t->sc.fseg, s, 0);
  } else {
    st = (t->lv_key && ret != PST_RET_POINT) ? pst::level_search(a, t->lv_key, t->lv_id, t->lsc, s)
                                             : pst::run_search(a, t->sc, s);
  }
  return st.ok() ? PST_OK : fail(st);
}

int pst_gpu_batch_search(pst_tree* t, const void* queries_d, size_t nq, int ret, unsigned long long* counts_d,
                         unsigned long long* offsets_d, void* out_d, size_t out_capacity, pst_stream_t stream) {
  if (!t || !counts_d || !offsets_d) return fail(PST_ERR_INVALID, "Error: null argument");
  if (nq == 0) return PST_OK;
  if (nq >= (1ull << 24)) return fail(PST_ERR_INVALID, "Error: batches hold fewer than 2^24 queries");
  if (ret != PST_RET_POINT && t->id_dtype == PST_NO_ID)
    return fail(PST_ERR_INVALID, "Error: IDs requested from a tree built without IDs");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (forest_dense(t, nq) && t->n > 0) {
    // stationary forest passes: the count keeps its segment offsets and the
    // write step places every hit from them
    Status st = forest_seg_reserve(t, nq);
    pst::SearchArgs a = base_args(t);
    a.queries = queries_d;
    a.nq = nq;
    a.mode = 0;
    a.counts = counts_d;
    if (st.ok()) st = pst::forest_dense_search(a, t->sc.fseg, s, 1);
    if (st.ok()) st = pst::exclusive_scan(counts_d, offsets_d, nq, s);
    if (!st.ok()) return fail(st);
    if (out_capacity == 0) return PST_OK;
    a.mode = 1;
    a.ret_point = ret;
    a.counts = nullptr;  // the kept segment offsets already give every position
    a.offsets = offsets_d;
    a.caps = counts_d;
    a.cap_all = ~0ull;
    a.out = out_d;
    a.out_total = out_capacity;
    st = pst::forest_dense_search(a, t->sc.fseg, s, 2);
    return st.ok() ? PST_OK : fail(st);
  }
  if (t->lv_key && ret != PST_RET_POINT && t->n > 0) {
    // through the level index: the count pass keeps its ranges and the
    // write step only clips them
    Status st = pst::scratch_reserve(t->sc, nq, 0, s);
    pst::SearchArgs a = base_args(t);
    a.queries = queries_d;
    a.nq = nq;
    a.mode = 0;
    a.counts = counts_d;
    if (st.ok()) st = pst::level_search(a, t->lv_key, t->lv_id, t->lsc, s, 1);
    if (st.ok()) st = pst::exclusive_scan(counts_d, offsets_d, nq, s);
    if (!st.ok()) return fail(st);
    if (out_capacity == 0) return PST_OK;  // a sizing call: counts and offsets only
    a.mode = 1;
    a.ret_point = ret;
    a.counts = t->sc.cursor;
    a.offsets = offsets_d;
    a.caps = counts_d;
    a.cap_all = ~0ull;
    a.out = out_d;
    a.out_total = out_capacity;
    if (st.ok()) st = pst::level_search(a, t->lv_key, t->lv_id, t->lsc, s, 2);
    return st.ok() ? PST_OK : fail(st);
  }
  int rc = pst_gpu_batch_count(t, queries_d, nq, counts_d, stream);
  if (rc != PST_OK) return rc;
  Status st = pst::exclusive_scan(counts_d, offsets_d, nq, s);
  if (!st.ok()) return fail(st);
  if (t->n == 0 || out_capacity == 0) return PST_OK;  // nothing to write (a sizing call: counts and offsets only)
  st = pst::scratch_reserve(t->sc, nq, 0, s);
  if (!st.ok()) return fail(st);
  pst::SearchArgs a = base_args(t);
  a.queries = queries_d;
  a.nq = nq;
  a.mode = 1;
  a.ret_point = ret;
  a.counts = t->sc.cursor;
  a.offsets = offsets_d;
  a.caps = counts_d;
  a.cap_all = ~0ull;
  a.out = out_d;
  a.out_total = out_capacity;
  a.groups_ready = !t->lv_key;  // the count pass above grouped this batch (unless the index counted it)
  st = pst::run_search(a, t->sc, s);
  return st.ok() ? PST_OK : fail(st);
}

int pst_gpu_set_search_tuning(pst_tree* t, int rounds, int budget_small, int budget_large, int waves_per_cu) {
  if (!t) return fail(PST_ERR_INVALID, "Error: null tree");
  if (rounds > 0) t->rounds = rounds;
  if (budget_small > 0) t->budget_small = budget_small;
  if (budget_large > 0) t->budget_large = budget_large;
  if (waves_per_cu >= 0) t->waves_per_cu = waves_per_cu;
  return PST_OK;
}

int pst_gpu_autotune(pst_tree* t, const void* queries_d, size_t nq, int ret, pst_stream_t stream, int* chosen6) {
  if (!t || !queries_d || nq == 0) return fail(PST_ERR_INVALID, "Error: autotuning needs a tree and queries");
  if (t->graphs > 0) return fail(PST_ERR_INVALID, pst::kPinnedScratchMsg);
  if (ret != PST_RET_POINT && t->id_dtype == PST_NO_ID)
    return fail(PST_ERR_INVALID, "Error: IDs requested from a tree built without IDs");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  unsigned long long *counts = nullptr, *offs = nullptr;
  void* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    if (counts) hipFree(counts);
    if (offs) hipFree(offs);
    if (out) hipFree(out);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
  };
  hipError_t e = hipMalloc(&counts, nq * 8);
  if (e == hipSuccess) e = hipMalloc(&offs, (nq + 1) * 8);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    cleanup();
    return fail(pst::hip_status(e, "Error in allocating autotuning buffers"));
  }
  int rc = pst_gpu_batch_count(t, queries_d, nq, counts, stream);
  if (rc != PST_OK) {
    cleanup();
    return rc;
  }
  Status st = pst::exclusive_scan(counts, offs, nq, s);
  unsigned long long total = 0;
  if (st.ok()) st = pst::hip_status(hipMemcpyAsync(&total, offs + nq, 8, hipMemcpyDeviceToHost, s), "Error in reading the batch total");
  if (st.ok()) st = pst::hip_status(hipStreamSynchronize(s), "Error in synchronising");
  const size_t rec = ret == PST_RET_POINT ? pst::point_size(t->dtype, t->id_dtype) : id_size(t->id_dtype);
  if (st.ok()) st = pst::hip_status(hipMalloc(&out, std::max<size_t>(total * rec, 1)), "Error in allocating autotuning output");
  if (!st.ok()) {
    cleanup();
    return fail(st);
  }
  // candidates: waves per CU x queries per walk x step budget of the spill rounds
  const int wpcs[] = {16, 24, 32, 40};
  const int groups[] = {pst::kGroup, 1};
  const int budgets[] = {128, 384, 1024};
  const int keep[5] = {t->rounds, t->budget_small, t->budget_large, t->waves_per_cu, t->group};
  float best = 1e30f;
  int best_cfg[5] = {keep[0], keep[1], keep[2], keep[3], keep[4]};
  for (int g : groups)
    for (int w : wpcs)
      for (int b : budgets) {
        t->group = g;
        t->waves_per_cu = w;
        t->budget_large = b;
        float ms = 0;
        for (int rep = 0; rep < 3 && rc == PST_OK; rep++) {
          if (rep == 1) hipEventRecord(e0, s);
          rc = pst_gpu_batch_search(t, queries_d, nq, ret, counts, offs, out, total, stream);
        }
        if (rc != PST_OK) break;
        hipEventRecord(e1, s);
        hipEventSynchronize(e1);
        hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
          best = ms;
          best_cfg[0] = t->rounds;
          best_cfg[1] = t->budget_small;
          best_cfg[2] = b;
          best_cfg[3] = w;
          best_cfg[4] = g;
        }
      }
  cleanup();
  if (rc != PST_OK) {
    t->rounds = keep[0], t->budget_small = keep[1], t->budget_large = keep[2], t->waves_per_cu = keep[3], t->group = keep[4];
    return rc;
  }
  t->rounds = best_cfg[0];
  t->budget_small = best_cfg[1];
  t->budget_large = best_cfg[2];
  t->waves_per_cu = best_cfg[3];
  t->group = best_cfg[4];
  if (chosen6) {
    for (int k = 0; k < 5; k++) chosen6[k] = best_cfg[k];
    chosen6[5] = static_cast<int>(best * 500.0f);  // microseconds per batch (two timed repetitions)
  }
  return PST_OK;
}

struct pst_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  pst_tree* tree = nullptr;
};

int pst_gpu_batch_graph_create(pst_tree* t, const void* queries_d, size_t nq, int ret, unsigned long long* counts_d,
                               unsigned long long* offsets_d, void* out_d, size_t out_capacity, pst_stream_t stream,
                               pst_graph** out) {
  if (!out) return fail(PST_ERR_INVALID, "Error: null output handle");
  *out = nullptr;
  if (!t) return fail(PST_ERR_INVALID, "Error: null tree");
  if (!stream) return fail(PST_ERR_INVALID, "Error: graph capture needs a stream other than the null stream");
  if (nq > 8192) return fail(PST_ERR_INVALID, "Error: captured batches hold at most 8192 queries");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // eager run first: every allocation the batch needs happens here
  int rc = pst_gpu_batch_search(t, queries_d, nq, ret, counts_d, offsets_d, out_d, out_capacity, stream);
  if (rc != PST_OK) return rc;
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(pst::hip_status(e, "Error in synchronising before capture"));
  e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) return fail(pst::hip_status(e, "Error in beginning graph capture"));
  rc = pst_gpu_batch_search(t, queries_d, nq, ret, counts_d, offsets_d, out_d, out_capacity, stream);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(s, &g);
  if (rc != PST_OK || e != hipSuccess) {
    if (g) hipGraphDestroy(g);
    return rc != PST_OK ? rc : fail(pst::hip_status(e, "Error in ending graph capture"));
  }
  pst_graph* pg = new pst_graph();
  pg->graph = g;
  e = hipGraphInstantiate(&pg->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(g);
    delete pg;
    return fail(pst::hip_status(e, "Error in instantiating the batch graph"));
  }
  pg->tree = t;
  // the graph holds the addresses of the handle's buffers: from now on a
  // call that would reallocate them fails instead (until the last graph goes)
  t->graphs++;
  t->sc.pinned = t->lsc.pinned = 1;
  *out = pg;
  return PST_OK;
}

int pst_gpu_batch_graph_launch(pst_graph* g, pst_stream_t stream) {
  if (!g) return fail(PST_ERR_INVALID, "Error: null graph");
  hipError_t e = hipGraphLaunch(g->exec, reinterpret_cast<hipStream_t>(stream));
  return e == hipSuccess ? PST_OK : fail(pst::hip_status(e, "Error in launching the batch graph"));
}

int pst_gpu_batch_graph_destroy(pst_graph* g) {
  if (!g) return PST_OK;
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  if (g->tree && --g->tree->graphs == 0) g->tree->sc.pinned = g->tree->lsc.pinned = 0;
  delete g;
  return PST_OK;
}

int pst_gpu_set_subtree_copy(pst_tree* t, int copy_min, int copy_max, size_t max_tasks) {
  if (!t) return fail(PST_ERR_INVALID, "Error: null tree");
  if (copy_max < 0 || copy_min < 0 || (copy_max > 0 && copy_min > copy_max))
    return fail(PST_ERR_INVALID, "Error: need 0 <= copy_min <= copy_max (copy_max 0: off)");
  t->copy_min = copy_min;
  t->copy_max = copy_max;
  t->copy_tasks = max_tasks;
  return PST_OK;
}

size_t pst_gpu_level_index_bytes(size_t n, int dtype, int id_dtype) {
  if (!valid_t(dtype) || id_dtype == PST_NO_ID || !valid_id(id_dtype)) return 0;
  return pst::level_index_bytes(dtype, id_dtype, n);
}

int pst_gpu_set_level_index(pst_tree* t, int on) {
  if (!t) return fail(PST_ERR_INVALID, "Error: null tree");
  if (!on) {
    if (t->graphs > 0 && t->lv_key)
      return fail(PST_ERR_INVALID, "Error: a batch graph of this tree uses its level index; destroy the graph first");
    hipSetDevice(t->device);
    if (t->lv_key) hipFree(t->lv_key);
    if (t->lv_id) hipFree(t->lv_id);
    t->lv_key = t->lv_id = nullptr;
    pst::level_scratch_release(t->lsc);
    return PST_OK;
  }
  if (t->slots_per_tree) return fail(PST_ERR_INVALID, "Error: the level index is for one deep tree, not a forest");
  if (t->id_dtype == PST_NO_ID) return fail(PST_ERR_INVALID, "Error: the level index needs a tree with IDs");
  if (t->lv_key || t->n == 0) return PST_OK;
  if (t->n >= (1ull << 31)) return fail(PST_ERR_INVALID, "Error: the level index holds trees of fewer than 2^31 points");
  hipSetDevice(t->device);
  const size_t elems = pst::level_index_bytes(t->dtype, t->id_dtype, t->n) /
                       (pst::dtype_size(t->dtype) + id_size(t->id_dtype));
  void *k = nullptr, *i = nullptr;
  hipError_t e = hipMalloc(&k, elems * pst::dtype_size(t->dtype));
  if (e == hipSuccess) e = hipMalloc(&i, elems * id_size(t->id_dtype));
  Status st = pst::hip_status(e, "Error in allocating the level index" + dev_suffix(t->dev_ind, t->num_devs));
  if (st.ok()) st = pst::build_level_index(t->dtype, t->id_dtype, t->n, t->nodes, t->ids, k, i, t->stream);
  if (!st.ok()) {
    if (k) hipFree(k);
    if (i) hipFree(i);
    return fail(st);
  }
  t->lv_key = k;
  t->lv_id = i;
  // set-up of the allocation-free batched search on this device
  if (!noalloc_scratch(t->device)) return PST_ERR_HIP;
  return PST_OK;
}

int pst_gpu_batch_search_noalloc(pst_tree* t, const void* queries_d, size_t nq, unsigned long long* counts_d,
                                 unsigned long long* offsets_d, void* out_d, size_t out_capacity,
                                 pst_stream_t stream) {
  if (!t || !counts_d || !offsets_d || (nq && !queries_d)) return fail(PST_ERR_INVALID, "Error: null argument");
  if (nq == 0) return PST_OK;
  if (nq > kNoAllocMaxQueries)
    return fail(PST_ERR_INVALID, "Error: the allocation-free batched search holds at most " +
                                     std::to_string(kNoAllocMaxQueries) + " queries per call");
  if (t->slots_per_tree || t->id_dtype == PST_NO_ID)
    return fail(PST_ERR_INVALID, "Error: the allocation-free batched search needs one deep tree with IDs");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (t->n == 0) {  // nothing to report: zero counts and offsets, with a kernel (capturable)
    hipLaunchKernelGGL(k_zero_u64, dim3(static_cast<unsigned>((nq + 256) / 256)), dim3(256), 0, s, offsets_d,
                       nq + 1);
    hipLaunchKernelGGL(k_zero_u64, dim3(static_cast<unsigned>((nq + 255) / 256)), dim3(256), 0, s, counts_d, nq);
    Status st = pst::hip_status(hipGetLastError(), "Error in launching the count reset");
    return st.ok() ? PST_OK : fail(st);
  }
  if (!t->lv_key)
    return fail(PST_ERR_INVALID, "Error: the allocation-free batched search runs through the level index; "
                                 "build it first (pst_gpu_set_level_index)");
  unsigned char* base = noalloc_scratch(t->device);
  if (!base) return PST_ERR_HIP;
  pst::LevelScratch ls{};
  ls.tsrc = reinterpret_cast<uint64_t*>(base + kNoAllocTsrc);
  ls.tpre = reinterpret_cast<uint32_t*>(base + kNoAllocTpre);
  ls.ntask = reinterpret_cast<uint32_t*>(base + kNoAllocNtask);
  ls.wq = reinterpret_cast<unsigned long long*>(base + kNoAllocWq);
  ls.woff = reinterpret_cast<unsigned long long*>(base + kNoAllocWoff);
  ls.chunk_task = reinterpret_cast<uint32_t*>(base + kNoAllocChunk);
  ls.cap = kNoAllocMaxQueries;
  ls.external = 1;
  unsigned long long* cursor = reinterpret_cast<unsigned long long*>(base + kNoAllocCursor);
  // count pass keeping each query's ranges, offsets, then the write step
  // clips the kept ranges: every scratch entry a step reads was written by
  // an earlier step of this call, so nothing needs clearing
  pst::SearchArgs a = base_args(t);
  a.queries = queries_d;
  a.nq = nq;
  a.mode = 0;
  a.counts = counts_d;
  Status st = pst::level_search(a, t->lv_key, t->lv_id, ls, s, 1);
  if (st.ok()) st = pst::exclusive_scan(counts_d, offsets_d, nq, s);
  if (!st.ok()) return fail(st);
  if (out_capacity == 0) return PST_OK;
  a.mode = 1;
  a.ret_point = 0;
  a.counts = cursor;
  a.offsets = offsets_d;
  a.caps = counts_d;
  a.cap_all = ~0ull;
  a.out = out_d;
  a.out_total = out_capacity;
  st = pst::level_search(a, t->lv_key, t->lv_id, ls, s, 2);
  return st.ok() ? PST_OK : fail(st);
}

int pst_gpu_set_query_grouping(pst_tree* t, int queries_per_walk) {
  if (!t) return fail(PST_ERR_INVALID, "Error: null tree");
  if (queries_per_walk != 1 && queries_per_walk != pst::kGroup)
    return fail(PST_ERR_INVALID, "Error: queries per walk must be 1 or " + std::to_string(pst::kGroup));
  t->group = queries_per_walk;
  return PST_OK;
}

size_t pst_calc_num_metacells(const size_t pt_grid_dims[3], const size_t metacell_dims[3],
                              size_t metacell_grid_dims[3]) {
  if (!pt_grid_dims || !metacell_dims) return 0;
  size_t total = 1;
  for (int d = 0; d < 3; d++) {
    if (pt_grid_dims[d] < 1 || metacell_dims[d] == 0) return 0;
    const size_t g = (pt_grid_dims[d] - 1 + metacell_dims[d] - 1) / metacell_dims[d];
    if (metacell_grid_dims) metacell_grid_dims[d] = g;
    total *= g;
  }
  return total;
}

static int metacell_dims_u32(const size_t pt[3], const size_t mc[3], uint32_t p32[3], uint32_t m32[3]) {
  for (int d = 0; d < 3; d++) {
    if (pt[d] >= (1ull << 31) || mc[d] >= (1ull << 31))
      return fail(PST_ERR_INVALID, "Error: grid dimensions must be below 2^31");
    p32[d] = static_cast<uint32_t>(pt[d]);
    m32[d] = static_cast<uint32_t>(mc[d]);
  }
  return PST_OK;
}

int pst_form_metacells_into(const void* vertex_arr_d, int dtype, int id_dtype, const size_t pt_grid_dims[3],
                            const size_t metacell_dims[3], void* metacell_arr_d, pst_stream_t stream) {
  if (!vertex_arr_d || !metacell_arr_d || !pt_grid_dims || !metacell_dims)
    return fail(PST_ERR_INVALID, "Error: null argument");
  if (!valid_t(dtype) || !valid_id(id_dtype)) return fail(PST_ERR_INVALID, "Error: unsupported type codes");
  uint32_t p32[3], m32[3];
  int rc = metacell_dims_u32(pt_grid_dims, metacell_dims, p32, m32);
  if (rc != PST_OK) return rc;
  Status st = pst::form_metacells(dtype, id_dtype, vertex_arr_d, p32, m32, metacell_arr_d,
                                  reinterpret_cast<hipStream_t>(stream));
  return st.ok() ? PST_OK : fail(st);
}

namespace {
int iso_args(const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3], const size_t metacell_dims[3],
             const void* ids_d, int id_dtype, size_t num_ids, const void* isovalue, uint32_t p32[3],
             uint32_t m32[3]) {
  if (!vertex_arr_d || !pt_grid_dims || !metacell_dims || !isovalue || (num_ids && !ids_d))
    return fail(PST_ERR_INVALID, "Error: null argument");
  if (!valid_t(dtype) || id_dtype == PST_NO_ID || !valid_id(id_dtype))
    return fail(PST_ERR_INVALID, "Error: unsupported type codes");
  return metacell_dims_u32(pt_grid_dims, metacell_dims, p32, m32);
}
}  // namespace

namespace {
using IsoCountFn = pst::Status (*)(int, int, const void*, const uint32_t*, const uint32_t*, const void*, size_t,
                                   const void*, unsigned long long*, unsigned long long*, hipStream_t);
using IsoWriteFn = pst::Status (*)(int, int, const void*, const uint32_t*, const uint32_t*, const void*, size_t,
                                   const void*, const unsigned long long*, size_t, float*, hipStream_t);

int surface_count(IsoCountFn fn, const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3],
                  const size_t metacell_dims[3], const void* metacell_ids_d, int id_dtype, size_t num_ids,
                  const void* isovalue, unsigned long long* counts_d, pst_stream_t stream) {
  uint32_t p32[3], m32[3];
  int rc = iso_args(vertex_arr_d, dtype, pt_grid_dims, metacell_dims, metacell_ids_d, id_dtype, num_ids, isovalue,
                    p32, m32);
  if (rc != PST_OK) return rc;
  if (num_ids && !counts_d) return fail(PST_ERR_INVALID, "Error: null argument");
  int dev = 0;
  hipGetDevice(&dev);
  unsigned long long* err = device_counter(dev);
  if (!err) return fail(PST_ERR_HIP, "Error in allocating a device counter");
  Status st = fn(dtype, id_dtype, vertex_arr_d, p32, m32, metacell_ids_d, num_ids, isovalue, counts_d, err,
                 reinterpret_cast<hipStream_t>(stream));
  return st.ok() ? PST_OK : fail(st);
}

int surface_write(IsoWriteFn fn, const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3],
                  const size_t metacell_dims[3], const void* metacell_ids_d, int id_dtype, size_t num_ids,
                  const void* isovalue, const unsigned long long* offsets_d, float* triangles_d, size_t capacity,
                  pst_stream_t stream) {
  uint32_t p32[3], m32[3];
  int rc = iso_args(vertex_arr_d, dtype, pt_grid_dims, metacell_dims, metacell_ids_d, id_dtype, num_ids, isovalue,
                    p32, m32);
  if (rc != PST_OK) return rc;
  if (num_ids && (!offsets_d || (capacity && !triangles_d))) return fail(PST_ERR_INVALID, "Error: null argument");
  Status st = fn(dtype, id_dtype, vertex_arr_d, p32, m32, metacell_ids_d, num_ids, isovalue, offsets_d, capacity,
                 triangles_d, reinterpret_cast<hipStream_t>(stream));
  return st.ok() ? PST_OK : fail(st);
}

int surface_blocking(IsoCountFn cfn, IsoWriteFn wfn, const void* vertex_arr_d, int dtype,
                     const size_t pt_grid_dims[3], const size_t metacell_dims[3], const void* metacell_ids_d,
                     int id_dtype, size_t num_ids, const void* isovalue, float** triangles_d, size_t* num_triangles) {
  if (!triangles_d || !num_triangles) return fail(PST_ERR_INVALID, "Error: null argument");
  *triangles_d = nullptr;
  *num_triangles = 0;
  uint32_t p32[3], m32[3];
  int rc = iso_args(vertex_arr_d, dtype, pt_grid_dims, metacell_dims, metacell_ids_d, id_dtype, num_ids, isovalue,
                    p32, m32);
  if (rc != PST_OK || num_ids == 0) return rc;
  int dev = 0;
  hipGetDevice(&dev);
  unsigned long long* err = device_counter(dev);
  if (!err) return fail(PST_ERR_HIP, "Error in allocating a device counter");
  unsigned long long* counts = nullptr;  // counts, then offsets (num_ids + 1)
  hipError_t e = hipMalloc(&counts, (2 * num_ids + 1) * sizeof(unsigned long long));
  if (e != hipSuccess) return fail(PST_ERR_HIP, std::string("Error in allocating isosurface counts: ") + hipGetErrorString(e));
  unsigned long long* offsets = counts + num_ids;
  Status st = pst::hip_status(hipMemsetAsync(err, 0, 8, nullptr), "Error in clearing the error flag");
  if (st.ok()) st = cfn(dtype, id_dtype, vertex_arr_d, p32, m32, metacell_ids_d, num_ids, isovalue, counts, err, nullptr);
  if (st.ok()) st = pst::exclusive_scan(counts, offsets, num_ids, nullptr);
  unsigned long long total = 0, bad = 0;
  if (st.ok())
    st = pst::hip_status(hipMemcpy(&total, offsets + num_ids, 8, hipMemcpyDeviceToHost),
                         "Error in reading the triangle count");
  if (st.ok()) st = pst::hip_status(hipMemcpy(&bad, err, 8, hipMemcpyDeviceToHost), "Error in reading the error flag");
  if (st.ok() && bad) st = Status{PST_ERR_INVALID, "Error: a metacell ID lies past the metacell grid"};
  float* tri = nullptr;
  if (st.ok() && total)
    st = pst::hip_status(hipMalloc(&tri, total * 9 * sizeof(float)), "Error in allocating the triangle array");
  if (st.ok() && total)
    st = wfn(dtype, id_dtype, vertex_arr_d, p32, m32, metacell_ids_d, num_ids, isovalue, offsets, total, tri, nullptr);
  if (st.ok()) st = pst::hip_status(hipStreamSynchronize(nullptr), "Error in synchronising after isosurface extraction");
  hipFree(counts);
  if (!st.ok()) {
    if (tri) hipFree(tri);
    return fail(st);
  }
  *triangles_d = tri;
  *num_triangles = total;
  return PST_OK;
}
}  // namespace

int pst_isosurface_count(const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3],
                         const size_t metacell_dims[3], const void* metacell_ids_d, int id_dtype, size_t num_ids,
                         const void* isovalue, unsigned long long* counts_d, pst_stream_t stream) {
  return surface_count(pst::iso_count, vertex_arr_d, dtype, pt_grid_dims, metacell_dims, metacell_ids_d, id_dtype,
                       num_ids, isovalue, counts_d, stream);
}

int pst_isosurface_write(const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3],
                         const size_t metacell_dims[3], const void* metacell_ids_d, int id_dtype, size_t num_ids,
                         const void* isovalue, const unsigned long long* offsets_d, float* triangles_d,
                         size_t capacity, pst_stream_t stream) {
  return surface_write(pst::iso_write, vertex_arr_d, dtype, pt_grid_dims, metacell_dims, metacell_ids_d, id_dtype,
                       num_ids, isovalue, offsets_d, triangles_d, capacity, stream);
}

int pst_isosurface(const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3], const size_t metacell_dims[3],
                   const void* metacell_ids_d, int id_dtype, size_t num_ids, const void* isovalue,
                   float** triangles_d, size_t* num_triangles) {
  return surface_blocking(pst::iso_count, pst::iso_write, vertex_arr_d, dtype, pt_grid_dims, metacell_dims,
                          metacell_ids_d, id_dtype, num_ids, isovalue, triangles_d, num_triangles);
}

int pst_marching_cubes_count(const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3],
                             const size_t metacell_dims[3], const void* metacell_ids_d, int id_dtype, size_t num_ids,
                             const void* isovalue, unsigned long long* counts_d, pst_stream_t stream) {
  return surface_count(pst::mc_count, vertex_arr_d, dtype, pt_grid_dims, metacell_dims, metacell_ids_d, id_dtype,
                       num_ids, isovalue, counts_d, stream);
}

int pst_marching_cubes_write(const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3],
                             const size_t metacell_dims[3], const void* metacell_ids_d, int id_dtype, size_t num_ids,
                             const void* isovalue, const unsigned long long* offsets_d, float* triangles_d,
                             size_t capacity, pst_stream_t stream) {
  return surface_write(pst::mc_write, vertex_arr_d, dtype, pt_grid_dims, metacell_dims, metacell_ids_d, id_dtype,
                       num_ids, isovalue, offsets_d, triangles_d, capacity, stream);
}

int pst_marching_cubes(const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3],
                       const size_t metacell_dims[3], const void* metacell_ids_d, int id_dtype, size_t num_ids,
                       const void* isovalue, float** triangles_d, size_t* num_triangles) {
  return surface_blocking(pst::mc_count, pst::mc_write, vertex_arr_d, dtype, pt_grid_dims, metacell_dims,
                          metacell_ids_d, id_dtype, num_ids, isovalue, triangles_d, num_triangles);
}

int pst_form_metacell_tags(const void* vertex_arr_d, int dtype, int id_dtype, const size_t pt_grid_dims[3],
                           const size_t metacell_dims[3], void** metacell_arr_d, size_t* num_metacells, int dev_ind,
                           int num_devs) {
  if (!metacell_arr_d || !num_metacells) return fail(PST_ERR_INVALID, "Error: null argument");
  *metacell_arr_d = nullptr;
  const size_t nm = pst_calc_num_metacells(pt_grid_dims, metacell_dims, nullptr);
  if (nm == 0) return fail(PST_ERR_INVALID, "Error: empty metacell grid");
  void* out = nullptr;
  hipError_t e = hipMalloc(&out, nm * pst::point_size(dtype, id_dtype));
  if (e != hipSuccess)
    return fail(PST_ERR_HIP, "Error in allocating metacell tag storage array" + dev_suffix(dev_ind, num_devs) +
                                 hipGetErrorString(e));
  int rc = pst_form_metacells_into(vertex_arr_d, dtype, id_dtype, pt_grid_dims, metacell_dims, out, nullptr);
  if (rc == PST_OK && hipStreamSynchronize(nullptr) != hipSuccess)
    rc = fail(PST_ERR_HIP, "Error in synchronising after metacell formation");
  if (rc != PST_OK) {
    hipFree(out);
    return rc;
  }
  *metacell_arr_d = out;
  *num_metacells = nm;
  return PST_OK;
}

int pst_synth_gaussian_volume(float* out_d, const size_t dims[3], const float* gaussians_d, unsigned num_gaussians,
                              float noise_std, unsigned long long seed, pst_stream_t stream) {
  if (!out_d || !dims || (num_gaussians && !gaussians_d)) return fail(PST_ERR_INVALID, "Error: null argument");
  uint32_t d32[3];
  for (int d = 0; d < 3; d++) {
    if (dims[d] == 0 || dims[d] >= (1ull << 31)) return fail(PST_ERR_INVALID, "Error: bad volume dimensions");
    d32[d] = static_cast<uint32_t>(dims[d]);
  }
  Status st = pst::gauss_volume(out_d, d32, gaussians_d, num_gaussians, noise_std, seed,
                                reinterpret_cast<hipStream_t>(stream));
  return st.ok() ? PST_OK : fail(st);
}

int pst_interval_parallel_search(const void* pt_arr_d, size_t n, int dtype, int id_dtype, int ret, void** res,
                                 size_t* num, const void* search_val, int dev_ind, int num_devs, int warp_size,
                                 unsigned num_thread_blocks, unsigned threads_per_block) {
  (void)warp_size;
  (void)threads_per_block;
  if (!res || !num || !search_val) return fail(PST_ERR_INVALID, "Error: null argument");
  if (!valid_t(dtype) || !valid_id(id_dtype)) return fail(PST_ERR_INVALID, "Error: unsupported type codes");
  if (ret != PST_RET_POINT && id_dtype == PST_NO_ID)
    return fail(PST_ERR_INVALID, "Error: IDs requested from points without IDs");
  const size_t rec = ret == PST_RET_POINT ? pst::point_size(dtype, id_dtype) : id_size(id_dtype);
  void* out = nullptr;
  hipError_t e = hipMalloc(&out, (n ? n : 1) * rec);
  if (e != hipSuccess)
    return fail(PST_ERR_HIP, "Error in allocating array to store PointStruct search result" +
                                 dev_suffix(dev_ind, num_devs) + hipGetErrorString(e));
  int dev = 0;
  hipGetDevice(&dev);
  unsigned long long* ctr = device_counter(dev);
  if (!ctr) {
    hipFree(out);
    return fail(PST_ERR_HIP, "Error in allocating global result array index" + dev_suffix(dev_ind, num_devs));
  }
  e = hipMemsetAsync(ctr, 0, 8, nullptr);
  if (e != hipSuccess) {
    hipFree(out);
    return fail(PST_ERR_HIP, "Error in initialising global result array index to 0" +
                                 dev_suffix(dev_ind, num_devs) + hipGetErrorString(e));
  }
  Status st = pst::ips_search(dtype, id_dtype, pt_arr_d, n, ret, search_val, out, n, ctr, num_thread_blocks, nullptr);
  unsigned long long h = 0;
  if (st.ok()) st = pst::hip_status(hipMemcpy(&h, ctr, 8, hipMemcpyDeviceToHost),
                                    "Error in copying global result array final index from device");
  if (!st.ok()) {
    hipFree(out);
    return fail(st);
  }
  *num = h;
  *res = out;
  return PST_OK;
}

int pst_interval_parallel_search_into(const void* pt_arr_d, size_t n, int dtype, int id_dtype, int ret,
                                      const void* search_val, void* out_d, size_t capacity,
                                      unsigned long long* counter_d, pst_stream_t stream) {
  if (!valid_t(dtype) || !valid_id(id_dtype) || !counter_d || !search_val)
    return fail(PST_ERR_INVALID, "Error: invalid argument");
  if (ret != PST_RET_POINT && id_dtype == PST_NO_ID)
    return fail(PST_ERR_INVALID, "Error: IDs requested from points without IDs");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  Status st = pst::hip_status(hipMemsetAsync(counter_d, 0, 8, s), "Error in clearing the result counter");
  if (st.ok()) st = pst::ips_search(dtype, id_dtype, pt_arr_d, n, ret, search_val, out_d, capacity, counter_d, 0, s);
  return st.ok() ? PST_OK : fail(st);
}

#ifdef PST_DEBUG_EXPORTS
// Debug builds only: per-round counters of the search launches.
int pst_debug_set_counters(pst_tree* t, unsigned long long* dbg) {
  t->debug = dbg;
  return PST_OK;
}
int pst_debug_get_round_counters(pst_tree* t, unsigned long long* host20) {
  if (!t->sc.ctr) return PST_ERR_INVALID;
  return hipMemcpy(host20, t->sc.ctr, 20 * sizeof(unsigned long long), hipMemcpyDeviceToHost) == hipSuccess
             ? PST_OK
             : PST_ERR_HIP;
}
#endif

}  // extern "C"
