// Query server over resident trees (no counterpart in the reference):
// continuous batching of 2-/3-sided queries against several StaticPSTGPU /
// StaticPSTGPUArr handles, with admission by output memory.
//
// Per tree, on that tree's own stream (trees overlap on the device):
//   pending   --(count batch: H2D queries, batch_count, D2H counts)-->
//   counted   --(admission: the longest prefix whose exact outputs fit one
//                contiguous range of the output pool)-->
//   writing   --(H2D queries/offsets/counts, batch_write)-->  complete.
// pst_server_step advances every stage without blocking (events are
// polled), so requests submitted while earlier batches run are admitted as
// soon as memory and a staging slot free up. A result stays in the pool
// until its ticket is released; a range returns to the pool when all its
// tickets are released.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "pst_capi.h"
#include "pst_internal.hpp"

namespace {

constexpr int kSlots = 6;  // staging slots per tree (in-flight write batches + one count)

struct Slot {
  void* q_dev = nullptr;                      // query records
  unsigned long long* cnt_dev = nullptr;      // counts
  unsigned long long* off_dev = nullptr;      // offsets within the batch's range
  unsigned char* q_host = nullptr;            // pinned staging
  unsigned long long* cnt_host = nullptr;     // pinned
  unsigned long long* off_host = nullptr;     // pinned
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<unsigned long long> tickets;
  size_t range = 0;  // pool range (write batches)
};

struct Ticket {
  int tree = 0;
  int ret = PST_RET_ID;
  unsigned char rec[32];
  unsigned long long count = 0;
  size_t range = 0, byte_off = 0;
  bool complete = false;
  bool failed = false;  // output larger than the whole pool
};

struct TreeQ {
  pst_tree* tree = nullptr;
  hipStream_t stream = nullptr;
  size_t qrec = 0;           // bytes per query record
  size_t rec_bytes[2] = {};  // bytes per result record: IDs, points
  std::deque<unsigned long long> pending[2];  // per RetType: submitted, not counted
  std::deque<unsigned long long> counted[2];  // counted, waiting for memory
  int counting = -1;                          // slot of the count batch in flight
  int count_ret = 0;
  std::deque<int> writing;                    // slots of write batches in flight
  Slot slots[kSlots];
};

// First-fit allocator over [0, size) with coalescing; offsets are aligned.
class Pool {
 public:
  explicit Pool(size_t size) { free_[0] = size; }
  bool alloc(size_t bytes, size_t& off) {
    bytes = (bytes + 255) & ~size_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        off = it->first;
        const size_t rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[off + bytes] = rest;
        used_[off] = bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t off) {
    auto u = used_.find(off);
    if (u == used_.end()) return;
    size_t o = off, sz = u->second;
    used_.erase(u);
    auto next = free_.lower_bound(o);
    if (next != free_.end() && o + sz == next->first) {
      sz += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == o) {
        prev->second += sz;
        return;
      }
    }
    free_[o] = sz;
  }

 private:
  std::map<size_t, size_t> free_, used_;
};

}  // namespace

struct pst_server {
  std::vector<TreeQ> trees;
  void* pool_dev = nullptr;
  size_t pool_bytes = 0;
  Pool* pool = nullptr;
  int max_batch = 0;
  unsigned long long next_ticket = 1;
  std::unordered_map<unsigned long long, Ticket> tickets;
  std::map<size_t, size_t> range_refs;  // pool range -> unreleased tickets
  std::vector<unsigned long long> done;  // completed since the last poll
  std::string err;
};

namespace {

thread_local std::string g_srv_err;

int sfail(int code, const std::string& m) {
  g_srv_err = m;
  return code;
}

size_t query_record_bytes(int dtype) { return (dtype == PST_F64 || dtype == PST_I64) ? 32 : 16; }

int start_count(pst_server* s, TreeQ& tq) {
  if (tq.counting >= 0) return PST_OK;
  int ret = -1;
  for (int r = 0; r < 2; r++)
    if (!tq.pending[r].empty()) {
      ret = r;
      break;
    }
  if (ret < 0) return PST_OK;
  int slot = -1;
  for (int k = 0; k < kSlots; k++)
    if (!tq.slots[k].busy) {
      slot = k;
      break;
    }
  if (slot < 0) return PST_OK;
  Slot& sl = tq.slots[slot];
  sl.tickets.clear();
  while (!tq.pending[ret].empty() && static_cast<int>(sl.tickets.size()) < s->max_batch) {
    const unsigned long long id = tq.pending[ret].front();
    tq.pending[ret].pop_front();
    std::memcpy(sl.q_host + sl.tickets.size() * tq.qrec, s->tickets[id].rec, tq.qrec);
    sl.tickets.push_back(id);
  }
  const size_t nq = sl.tickets.size();
  hipError_t e = hipMemcpyAsync(sl.q_dev, sl.q_host, nq * tq.qrec, hipMemcpyHostToDevice, tq.stream);
  if (e != hipSuccess) return sfail(PST_ERR_HIP, std::string("Error in staging queries: ") + hipGetErrorString(e));
  int rc = pst_gpu_batch_count(tq.tree, sl.q_dev, nq, sl.cnt_dev, tq.stream);
  if (rc != PST_OK) return sfail(rc, pst_last_error());
  e = hipMemcpyAsync(sl.cnt_host, sl.cnt_dev, nq * 8, hipMemcpyDeviceToHost, tq.stream);
  if (e == hipSuccess) e = hipEventRecord(sl.done, tq.stream);
  if (e != hipSuccess) return sfail(PST_ERR_HIP, std::string("Error in queueing counts: ") + hipGetErrorString(e));
  sl.busy = true;
  tq.counting = slot;
  tq.count_ret = ret;
  return PST_OK;
}

int finish_count(pst_server* s, TreeQ& tq) {
  if (tq.counting < 0) return PST_OK;
  Slot& sl = tq.slots[tq.counting];
  const hipError_t q = hipEventQuery(sl.done);
  if (q == hipErrorNotReady) return PST_OK;
  if (q != hipSuccess) return sfail(PST_ERR_HIP, std::string("Error in count batch: ") + hipGetErrorString(q));
  for (size_t i = 0; i < sl.tickets.size(); i++) {
    s->tickets[sl.tickets[i]].count = sl.cnt_host[i];
    tq.counted[tq.count_ret].push_back(sl.tickets[i]);
  }
  sl.tickets.clear();
  sl.busy = false;
  tq.counting = -1;
  return PST_OK;
}

int admit(pst_server* s, TreeQ& tq) {
  for (int ret = 0; ret < 2; ret++) {
    while (!tq.counted[ret].empty()) {
      int slot = -1;
      for (int k = 0; k < kSlots; k++)
        if (!tq.slots[k].busy) {
          slot = k;
          break;
        }
      if (slot < 0) return PST_OK;
      const size_t rs = tq.rec_bytes[ret];
      // longest prefix (<= max_batch) whose outputs fit one pool range
      size_t take = 0, total = 0;
      while (take < tq.counted[ret].size() && static_cast<int>(take) < s->max_batch) {
        total += s->tickets[tq.counted[ret][take]].count;
        take++;
      }
      size_t off = 0;
      while (take > 0 && !s->pool->alloc(std::max<size_t>(total * rs, 1), off)) {
        take--;
        total -= s->tickets[tq.counted[ret][take]].count;
      }
      if (take == 0) {
        // a query whose output exceeds the whole pool can never be admitted
        const unsigned long long id = tq.counted[ret].front();
        Ticket& tk = s->tickets[id];
        if (tk.count * rs > s->pool_bytes) {
          tq.counted[ret].pop_front();
          tk.failed = tk.complete = true;
          s->done.push_back(id);
          continue;
        }
        return PST_OK;  // wait for memory
      }
      Slot& sl = tq.slots[slot];
      sl.tickets.clear();
      unsigned long long run = 0;
      for (size_t i = 0; i < take; i++) {
        const unsigned long long id = tq.counted[ret].front();
        tq.counted[ret].pop_front();
        Ticket& tk = s->tickets[id];
        std::memcpy(sl.q_host + i * tq.qrec, tk.rec, tq.qrec);
        sl.cnt_host[i] = tk.count;
        sl.off_host[i] = run;
        tk.range = off;
        tk.byte_off = off + run * rs;
        run += tk.count;
        sl.tickets.push_back(id);
      }
      s->range_refs[off] = take;
      sl.range = off;
      hipError_t e = hipMemcpyAsync(sl.q_dev, sl.q_host, take * tq.qrec, hipMemcpyHostToDevice, tq.stream);
      if (e == hipSuccess) e = hipMemcpyAsync(sl.cnt_dev, sl.cnt_host, take * 8, hipMemcpyHostToDevice, tq.stream);
      if (e == hipSuccess) e = hipMemcpyAsync(sl.off_dev, sl.off_host, take * 8, hipMemcpyHostToDevice, tq.stream);
      if (e != hipSuccess) return sfail(PST_ERR_HIP, std::string("Error in staging a write batch: ") + hipGetErrorString(e));
      int rc = pst_gpu_batch_write(tq.tree, sl.q_dev, take, ret, sl.off_dev, sl.cnt_dev,
                                   static_cast<unsigned char*>(s->pool_dev) + off, tq.stream);
      if (rc != PST_OK) return sfail(rc, pst_last_error());
      e = hipEventRecord(sl.done, tq.stream);
      if (e != hipSuccess) return sfail(PST_ERR_HIP, std::string("Error in recording a write batch: ") + hipGetErrorString(e));
      sl.busy = true;
      tq.writing.push_back(slot);
    }
  }
  return PST_OK;
}

int retire(pst_server* s, TreeQ& tq) {
  while (!tq.writing.empty()) {
    Slot& sl = tq.slots[tq.writing.front()];
    const hipError_t q = hipEventQuery(sl.done);
    if (q == hipErrorNotReady) return PST_OK;
    if (q != hipSuccess) return sfail(PST_ERR_HIP, std::string("Error in write batch: ") + hipGetErrorString(q));
    for (auto id : sl.tickets) {
      s->tickets[id].complete = true;
      s->done.push_back(id);
    }
    sl.tickets.clear();
    sl.busy = false;
    tq.writing.pop_front();
  }
  return PST_OK;
}

}  // namespace

extern "C" {

const char* pst_server_last_error(void) { return g_srv_err.c_str(); }

int pst_server_create(pst_tree* const* trees, int num_trees, size_t out_budget_bytes, int max_batch,
                      pst_server** out) {
  if (!out || !trees || num_trees <= 0 || max_batch <= 0 || out_budget_bytes == 0)
    return sfail(PST_ERR_INVALID, "Error: invalid server arguments");
  for (int i = 0; i < num_trees; i++)
    if (!trees[i]) return sfail(PST_ERR_INVALID, "Error: null tree");
  *out = nullptr;
  pst_server* s = new pst_server();
  s->max_batch = max_batch;
  s->pool_bytes = out_budget_bytes;
  hipError_t e = hipMalloc(&s->pool_dev, out_budget_bytes);
  if (e != hipSuccess) {
    delete s;
    return sfail(PST_ERR_HIP, std::string("Error in allocating the output pool: ") + hipGetErrorString(e));
  }
  s->pool = new Pool(out_budget_bytes);
  s->trees.resize(num_trees);
  for (int i = 0; i < num_trees; i++) {
    TreeQ& tq = s->trees[i];
    tq.tree = trees[i];
    const int dt = pst_gpu_dtype(trees[i]), idt = pst_gpu_id_dtype(trees[i]);
    tq.qrec = query_record_bytes(dt);
    const size_t pt = pst_point_size(dt, idt);
    tq.rec_bytes[PST_RET_POINT] = pt;
    tq.rec_bytes[PST_RET_ID] = idt == PST_NO_ID ? pt : ((idt == PST_F64 || idt == PST_I64 || idt == PST_U64) ? 8 : 4);
    e = hipStreamCreateWithFlags(&tq.stream, hipStreamNonBlocking);
    for (int k = 0; k < kSlots && e == hipSuccess; k++) {
      Slot& sl = tq.slots[k];
      e = hipMalloc(&sl.q_dev, max_batch * tq.qrec);
      if (e == hipSuccess) e = hipMalloc(&sl.cnt_dev, (max_batch + 1) * 8);
      if (e == hipSuccess) e = hipMalloc(&sl.off_dev, (max_batch + 1) * 8);
      if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&sl.q_host), max_batch * tq.qrec);
      if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&sl.cnt_host), (max_batch + 1) * 8);
      if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&sl.off_host), (max_batch + 1) * 8);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
      pst_server_destroy(s);
      return sfail(PST_ERR_HIP, std::string("Error in allocating server staging: ") + hipGetErrorString(e));
    }
    // size the tree's search scratch for a full batch now, so that serving
    // never allocates: max_batch three-sided queries with min_dim1 = 1 >
    // max_dim1 = 0, which report nothing, and zero offsets
    {
      Slot& s0 = tq.slots[0];
      std::memset(s0.q_host, 0, max_batch * tq.qrec);
      for (int q = 0; q < max_batch; q++) {
        unsigned char* r = s0.q_host + q * tq.qrec;
        if (dt == PST_F32) { const float one = 1.0f; std::memcpy(r, &one, 4); }
        else if (dt == PST_I32) { const int32_t one = 1; std::memcpy(r, &one, 4); }
        else if (dt == PST_F64) { const double one = 1.0; std::memcpy(r, &one, 8); }
        else { const int64_t one = 1; std::memcpy(r, &one, 8); }
      }
      std::memset(s0.off_host, 0, (max_batch + 1) * 8);
      e = hipMemcpyAsync(s0.q_dev, s0.q_host, max_batch * tq.qrec, hipMemcpyHostToDevice, tq.stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(s0.off_dev, s0.off_host, (max_batch + 1) * 8, hipMemcpyHostToDevice, tq.stream);
    }
    int rc = e == hipSuccess ? pst_gpu_batch_count(tq.tree, tq.slots[0].q_dev, max_batch, tq.slots[0].cnt_dev,
                                                   tq.stream)
                             : PST_ERR_HIP;
    if (rc == PST_OK) rc = pst_gpu_batch_write(tq.tree, tq.slots[0].q_dev, max_batch, PST_RET_POINT,
                                               tq.slots[0].off_dev, tq.slots[0].cnt_dev, s->pool_dev, tq.stream);
    // ID writes of a tree with a level index use that index's own search
    // buffers: size them too
    if (rc == PST_OK && idt != PST_NO_ID)
      rc = pst_gpu_batch_write(tq.tree, tq.slots[0].q_dev, max_batch, PST_RET_ID, tq.slots[0].off_dev,
                               tq.slots[0].cnt_dev, s->pool_dev, tq.stream);
    if (rc == PST_OK && hipStreamSynchronize(tq.stream) != hipSuccess) rc = PST_ERR_HIP;
    if (rc != PST_OK) {
      const std::string m = std::string("Error in warming up the server: ") + pst_last_error();
      pst_server_destroy(s);
      return sfail(rc, m);
    }
  }
  *out = s;
  return PST_OK;
}

int pst_server_destroy(pst_server* s) {
  if (!s) return PST_OK;
  for (auto& tq : s->trees) {
    if (tq.stream) hipStreamSynchronize(tq.stream);
    for (auto& sl : tq.slots) {
      if (sl.q_dev) hipFree(sl.q_dev);
      if (sl.cnt_dev) hipFree(sl.cnt_dev);
      if (sl.off_dev) hipFree(sl.off_dev);
      if (sl.q_host) hipHostFree(sl.q_host);
      if (sl.cnt_host) hipHostFree(sl.cnt_host);
      if (sl.off_host) hipHostFree(sl.off_host);
      if (sl.done) hipEventDestroy(sl.done);
    }
    if (tq.stream) hipStreamDestroy(tq.stream);
  }
  if (s->pool_dev) hipFree(s->pool_dev);
  delete s->pool;
  delete s;
  return PST_OK;
}

int pst_server_submit(pst_server* s, int tree, const void* query, int ret, unsigned long long* ticket) {
  if (!s || !query || !ticket || tree < 0 || tree >= static_cast<int>(s->trees.size()) ||
      (ret != PST_RET_ID && ret != PST_RET_POINT))
    return sfail(PST_ERR_INVALID, "Error: invalid submission");
  TreeQ& tq = s->trees[tree];
  if (ret == PST_RET_ID && pst_gpu_id_dtype(tq.tree) == PST_NO_ID)
    return sfail(PST_ERR_INVALID, "Error: IDs requested from a tree built without IDs");
  const unsigned long long id = s->next_ticket++;
  Ticket tk;
  tk.tree = tree;
  tk.ret = ret;
  std::memcpy(tk.rec, query, tq.qrec);
  s->tickets[id] = tk;
  tq.pending[ret].push_back(id);
  *ticket = id;
  return PST_OK;
}

int pst_server_submit_many(pst_server* s, int tree, const void* queries, size_t num, int ret,
                           unsigned long long* tickets) {
  if (!s || (num && (!queries || !tickets)) || tree < 0 || tree >= static_cast<int>(s->trees.size()))
    return sfail(PST_ERR_INVALID, "Error: invalid submission");
  const size_t rec = s->trees[tree].qrec;
  for (size_t i = 0; i < num; i++) {
    int rc = pst_server_submit(s, tree, static_cast<const unsigned char*>(queries) + i * rec, ret, &tickets[i]);
    if (rc != PST_OK) return rc;
  }
  return PST_OK;
}

int pst_server_step(pst_server* s) {
  if (!s) return sfail(PST_ERR_INVALID, "Error: null server");
  for (auto& tq : s->trees) {
    int rc = retire(s, tq);
    if (rc == PST_OK) rc = finish_count(s, tq);
    if (rc == PST_OK) rc = admit(s, tq);
    if (rc == PST_OK) rc = start_count(s, tq);
    if (rc != PST_OK) return rc;
  }
  return PST_OK;
}

int pst_server_poll(pst_server* s, unsigned long long* tickets, size_t capacity, size_t* num) {
  if (!s || !num) return sfail(PST_ERR_INVALID, "Error: null argument");
  const size_t k = std::min(capacity, s->done.size());
  for (size_t i = 0; i < k; i++) tickets[i] = s->done[i];
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *num = k;
  return PST_OK;
}

int pst_server_poll_counts(pst_server* s, unsigned long long* tickets, unsigned long long* counts, size_t capacity,
                           size_t* num) {
  if (!s || !num || (capacity && (!tickets || !counts))) return sfail(PST_ERR_INVALID, "Error: null argument");
  const size_t k = std::min(capacity, s->done.size());
  for (size_t i = 0; i < k; i++) {
    const Ticket& tk = s->tickets[s->done[i]];
    tickets[i] = s->done[i];
    counts[i] = tk.failed ? ~0ull : tk.count;
  }
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *num = k;
  return PST_OK;
}

int pst_server_result(pst_server* s, unsigned long long ticket, void** res_d, size_t* num) {
  if (!s || !res_d || !num) return sfail(PST_ERR_INVALID, "Error: null argument");
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end() || !it->second.complete) return sfail(PST_ERR_INVALID, "Error: ticket not complete");
  if (it->second.failed) {
    *res_d = nullptr;
    *num = it->second.count;
    return sfail(PST_ERR_CAPACITY, "Error: the query's output exceeds the server's output budget");
  }
  *res_d = static_cast<unsigned char*>(s->pool_dev) + it->second.byte_off;
  *num = it->second.count;
  return PST_OK;
}

int pst_server_release(pst_server* s, unsigned long long ticket) {
  if (!s) return sfail(PST_ERR_INVALID, "Error: null server");
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end() || !it->second.complete) return sfail(PST_ERR_INVALID, "Error: ticket not complete");
  const size_t range = it->second.range;
  const bool failed = it->second.failed;
  s->tickets.erase(it);
  if (failed) return PST_OK;
  auto r = s->range_refs.find(range);
  if (r != s->range_refs.end() && --r->second == 0) {
    s->pool->release(range);
    s->range_refs.erase(r);
  }
  return PST_OK;
}

int pst_server_release_many(pst_server* s, const unsigned long long* tickets, size_t num) {
  if (!s || (num && !tickets)) return sfail(PST_ERR_INVALID, "Error: null argument");
  for (size_t i = 0; i < num; i++) {
    int rc = pst_server_release(s, tickets[i]);
    if (rc != PST_OK) return rc;
  }
  return PST_OK;
}

size_t pst_server_outstanding(const pst_server* s) { return s ? s->tickets.size() : 0; }

}  // extern "C"
