// Internal interfaces between the translation units of libpst_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>

#include "pst_common.hpp"

namespace pst {

// ---- sorting (pst_sort.hip) ----
// `fail` (device int, optional) is set to 1 if a pass's bounded wait for
// earlier tiles ran out, which leaves the output undefined.
hipError_t radix_sort_u32(void* tmp, size_t& tmp_bytes, const uint32_t* kin, uint32_t* kout,
                          const uint32_t* vin, uint32_t* vout, size_t n, hipStream_t s, int* fail = nullptr);
hipError_t radix_sort_u64(void* tmp, size_t& tmp_bytes, const uint64_t* kin, uint64_t* kout,
                          const uint32_t* vin, uint32_t* vout, size_t n, hipStream_t s, int* fail = nullptr);

// Error carrier: a HIP status plus the reference-style context message
// ("Error in ... on device X (1-indexed) of Y: " + hipGetErrorString).
struct Status {
  int code = 0;  // 0 ok, else a PST_ERR_* value
  std::string msg;
  bool ok() const { return code == 0; }
};

// Sizes of one tree's device arrays.
size_t node_size(int dtype, int id_dtype);
size_t point_size(int dtype, int id_dtype);
size_t dtype_size(int dtype);

// ---- construction (pst_build.hip) ----
struct BuildTimes {
  float keys_ms = 0, sort_ms = 0, levels_ms = 0, finish_ms = 0;
};
// Builds the tree for n AoS points (device pointer) into `nodes`, which
// holds S = 2^ceil(log2(n+1)) - 1 records. Returns PST_ERR_NAN if a NaN dim
// (or floating ID) is found. With `submin_d` (and `ids_d`, null for trees
// without IDs) also fills what build_submin(.., slots_per_tree = 0, ..) does.
Status build_tree(int dtype, int id_dtype, const void* pts_d, size_t n, void* nodes_d, hipStream_t stream,
                  BuildTimes* times, void* submin_d = nullptr, void* ids_d = nullptr);
// The arrays a build writes: node records and, optionally, per-slot subtree
// minima and IDs. With `alloc` set, the build calls it once its sorts are
// queued (the arrays are first written after them), so that allocating
// them overlaps the sorts; alloc sets nodes, submin and ids.
struct BuildOut {
  void* nodes = nullptr;
  void* submin = nullptr;
  void* ids = nullptr;
  std::function<Status(BuildOut&)> alloc;
};
Status build_tree(int dtype, int id_dtype, const void* pts_d, size_t n, BuildOut& out, hipStream_t stream,
                  BuildTimes* times = nullptr);

// Device bytes one construction of n points allocates (freed when it ends).
size_t build_scratch_bytes(int dtype, int id_dtype, size_t n);
// Device bytes a handle keeps for one search of a deep tree of n points
// (work lists, copy tasks, the overflow stacks of the default grid on a GPU
// with `cus` compute units), allocated at its first search.
size_t search_scratch_bytes(size_t n, int dtype, int cus);

// Reference-layout blob (dim1, dim2, med, ids, bitcodes) written to a
// device buffer, for export / parity checks.
Status export_tree(int dtype, int id_dtype, size_t n, const void* nodes_d, void* blob_d, size_t off_dim1,
                   size_t off_dim2, size_t off_med, size_t off_id, size_t off_bits, hipStream_t stream);

// ---- forest of shallow trees (pst_build.hip) ----
// Tree c holds points [c*St, min(n, (c+1)*St)), St = 2^L - 1 <= 2047; its
// records start at slot c << L of `nodes` (ceil(n/St) << L records).
Status build_forest(int dtype, int id_dtype, const void* pts_d, size_t n, size_t slots_per_tree, void* nodes_d,
                    hipStream_t stream);
// Per-tree reference layouts, trees concatenated: tree c at position c*St
// of each array; total_slots = (trees - 1)*St + S(last tree's points).
Status export_forest(int dtype, int id_dtype, size_t n, size_t slots_per_tree, const void* nodes_d, void* blob_d,
                     size_t total_slots, size_t off_dim1, size_t off_dim2, size_t off_med, size_t off_id,
                     size_t off_bits, hipStream_t stream);

// Subtree minimum of dim2 per slot (T per record slot) and, when ids_d is
// not null, each slot's ID (stored ID type per slot), for a tree or a forest
// (slots_per_tree > 0), after construction.
Status build_submin(int dtype, int id_dtype, size_t n, size_t slots_per_tree, const void* nodes_d, void* submin_d,
                    void* ids_d, hipStream_t stream);

// ---- searches on one deep tree (pst_search.hip) ----
constexpr int kGroup = 8;  // queries per group walk in batched searches
struct SearchScratch {
  void* list[2] = {nullptr, nullptr};  // work lists (16-byte entries)
  size_t list_cap = 0;                 // entries per list
  unsigned long long* ctr = nullptr;   // round counters
  unsigned long long* cursor = nullptr;  // per-query output cursors
  size_t cursor_cap = 0;
  uint32_t* groups = nullptr;  // group table: kGroup query indices per group
  void* gkeys = nullptr;       // grouping sort keys and values
  void* gtmp = nullptr;        // grouping sort temporary storage
  size_t gtmp_bytes = 0;
  size_t gcap = 0;  // queries the grouping storage holds (a multiple of kGroup)
  void* gstack = nullptr;  // per-wave overflow stacks of the walks
  size_t gstack_bytes = 0;
  void* tasks = nullptr;  // copy tasks of write passes (16-byte entries)
  size_t task_cap = 0;
  void* fseg = nullptr;  // forest batches: per (query, segment) hit counts, then offsets
  size_t fseg_bytes = 0;
  // batch graphs of the handle hold these buffers' addresses: while any
  // exists, a call that would reallocate them fails instead
  int pinned = 0;
};
// Message of the error a pinned scratch gives for a call that needs more.
extern const char* const kPinnedScratchMsg;
// Work lists for nq queries whose walks start from `roots` root items.
Status scratch_reserve(SearchScratch& sc, size_t nq, size_t roots, hipStream_t stream);
void scratch_release(SearchScratch& sc);

// mode 0: count into counts_d[q]; mode 1: write results for query q at
// out + offsets_d[q] (offsets_d may be null: all offsets 0), at most
// caps_d[q] of them (caps_d null: cap_all), counting into cursor[q].
struct SearchArgs {
  int dtype, id_dtype;
  size_t n;
  const void* nodes;
  const void* submin;     // subtree minimum of dim2 per slot
  const void* ids;        // ID per slot (null for trees without IDs)
  size_t slots_per_tree;  // 0: one deep tree; else a forest of trees of this many slots
  const void* queries;  // device Query<T>[nq]
  size_t nq;
  int mode;             // 0 count, 1 write
  int ret_point;
  unsigned long long* counts;        // mode 0 output / mode 1 cursors
  const unsigned long long* offsets;  // mode 1
  const unsigned long long* caps;     // mode 1 (nullable)
  unsigned long long cap_all;
  void* out;
  int rounds;           // kernel launches enqueued (>= 1)
  int budget_small, budget_large;
  int waves_per_cu;
  int group;            // 1: one query per walk; otherwise kGroup
  int copy_min, copy_max;  // write pass: full subtrees of this many points are copied (copy_max 0: never)
  size_t copy_tasks;       // copy task list entries (0: default; beyond it the walk copies)
  int groups_ready;     // the group table in the scratch already holds this batch
  const void* single_query;  // host Query<T> for one query (queries ignored)
  unsigned long long* debug;  // optional per-round counters (8 per round)
  unsigned long long out_total = ~0ull;  // records `out` holds (~0: unlimited; 0: nothing is written)
  const void* tree_bounds = nullptr;     // forest: each tree's dim1 range (T lo, hi)
};
Status run_search(const SearchArgs& a, SearchScratch& sc, hipStream_t stream);
unsigned long long* search_error_flag(SearchScratch& sc);
// offsets[q] = sum of in[0..q), offsets[n] = total (one workgroup).
Status exclusive_scan(const unsigned long long* in, unsigned long long* offsets, size_t n, hipStream_t s);

// ---- forest batches with the trees held stationary (pst_forest.hip) ----
// Batches of at most this many queries on a forest of trees of at most 511
// points run as two passes over the forest instead of one walk per tree.
constexpr size_t kForestDenseMax = 512;
constexpr size_t kForestDenseMaxSlots = 511;
size_t forest_dense_scratch_bytes(size_t n, size_t slots_per_tree, size_t nq, int cus);
// mode 0: counts[q]; plan 1 also keeps the per-segment offsets in segbuf.
// mode 1: writes query q's results at out + offsets[q] (at most caps[q] or
// cap_all, and within out_total); plan 2 uses the offsets kept by a plan 1
// count of the same batch, otherwise the pass counts first (and stores each
// query's total in counts). Asynchronous.
Status forest_dense_search(const SearchArgs& a, void* segbuf, hipStream_t stream, int plan);
// Each forest tree's dim1 range (two T per tree), read by the stationary passes.
size_t forest_bounds_bytes(int dtype, size_t n, size_t slots_per_tree);
Status build_forest_bounds(int dtype, int id_dtype, size_t n, size_t slots_per_tree, const void* nodes_d,
                           void* bounds_d, hipStream_t stream);

// ---- level index of a deep tree (pst_levels.hip) ----
// Device bytes of the index: bit_length(n) levels of n (key, ID) pairs,
// each level padded to a multiple of 64 pairs.
size_t level_index_bytes(int dtype, int id_dtype, size_t n);
// Builds the index of a deep tree with IDs (ids_d: ID per slot) into key_d
// (unsigned keys of the tree's dtype width) and id_d (blocking).
Status build_level_index(int dtype, int id_dtype, size_t n, const void* nodes_d, const void* ids_d, void* key_d,
                         void* id_d, hipStream_t stream);
struct LevelScratch {
  uint64_t* tsrc = nullptr;   // per query, 128 ranges: first index in the level arrays
  uint32_t* tpre = nullptr;   // per query, 128: output position of each range, then the total
  uint32_t* ntask = nullptr;  // ranges per query
  unsigned long long* wq = nullptr;    // IDs written per query
  unsigned long long* woff = nullptr;  // their exclusive prefix (nq + 1)
  uint32_t* chunk_task = nullptr;      // copy chunk -> first range
  size_t cap = 0;                      // queries the buffers hold
  int pinned = 0;                      // as SearchScratch::pinned
  int external = 0;                    // the buffers belong to the caller (never freed or grown here)
};
Status level_scratch_reserve(LevelScratch& sc, size_t nq);
void level_scratch_release(LevelScratch& sc);
// Batched count (a.mode 0) or ID write (a.mode 1) with the semantics of
// run_search, through the level index (asynchronous). plan 1 (count): also
// keep the batch's ranges in the scratch; plan 2 (write): write from the
// ranges kept by the count pass of the same batch instead of planning again.
Status level_search(const SearchArgs& a, const void* key_d, const void* id_d, LevelScratch& sc, hipStream_t stream,
                    int plan = 0);

// ---- interval parallel search (pst_ips.hip) ----
// counter_d must be zero on entry; results past `cap` are counted only.
Status ips_search(int dtype, int id_dtype, const void* pts_d, size_t n, int ret_point,
                  const void* search_val, void* out_d, unsigned long long cap,
                  unsigned long long* counter_d, unsigned num_blocks, hipStream_t stream);

// ---- metacells (pst_metacell.hip) ----
// One PointStruct per metacell, at index = its ID (asynchronous).
Status form_metacells(int dtype, int id_dtype, const void* vertices_d, const uint32_t pt_dims[3],
                      const uint32_t mc_dims[3], void* out_pts_d, hipStream_t stream);
// Sum of Gaussians plus normal noise on a dims[0] x dims[1] x dims[2] grid
// (x fastest); params_d: per Gaussian (cx, cy, cz, -1/(2 s^2), amp, 0, 0, 0).
Status gauss_volume(float* out_d, const uint32_t dims[3], const float* params_d, uint32_t ngauss, float noise_std,
                    uint64_t seed, hipStream_t stream);

// helpers
Status hip_status(hipError_t e, const std::string& what);

// ---- isosurface of active metacells (pst_isosurface.hip) ----
// counts_d[i]: triangles of metacell ids_d[i]; bit 0 of *err_d is set for an
// ID past the metacell grid (counted as 0).
Status iso_count(int dtype, int id_dtype, const void* vertices_d, const uint32_t pt_dims[3],
                 const uint32_t mc_dims[3], const void* ids_d, size_t nids, const void* isovalue,
                 unsigned long long* counts_d, unsigned long long* err_d, hipStream_t stream);
// Triangles of metacell ids_d[i] at tri_d + 9 offsets_d[i] (floats);
// triangles at or past index `capacity` are not written.
Status iso_write(int dtype, int id_dtype, const void* vertices_d, const uint32_t pt_dims[3],
                 const uint32_t mc_dims[3], const void* ids_d, size_t nids, const void* isovalue,
                 const unsigned long long* offsets_d, size_t capacity, float* tri_d, hipStream_t stream);

// Marching cubes over the listed metacells: the same contract as iso_count
// and iso_write, one surface patch per voxel from the 256-case table.
Status mc_count(int dtype, int id_dtype, const void* vertices_d, const uint32_t pt_dims[3],
                const uint32_t mc_dims[3], const void* ids_d, size_t nids, const void* isovalue,
                unsigned long long* counts_d, unsigned long long* err_d, hipStream_t stream);
Status mc_write(int dtype, int id_dtype, const void* vertices_d, const uint32_t pt_dims[3],
                const uint32_t mc_dims[3], const void* ids_d, size_t nids, const void* isovalue,
                const unsigned long long* offsets_d, size_t capacity, float* tri_d, hipStream_t stream);

// ---- tree files and raw volumes (pst_persist.hip) ----
struct TreeFileInfo {
  int kind;  // 0: one deep tree, 1: forest
  int dtype, id_dtype;
  size_t n, S, slots_per_tree;
  uint64_t checksum;
};
// Writes the node records of a tree (S slots) to `path`.
Status save_tree_file(const char* path, int kind, int dtype, int id_dtype, size_t n, size_t S,
                      size_t slots_per_tree, const void* nodes_d, hipStream_t stream);
// Reads and validates a tree file's header (no device work).
Status read_tree_header(const char* path, TreeFileInfo* info);
// Reads the node records into nodes_d (info.S slots) and checks the checksum.
Status load_tree_file(const char* path, const TreeFileInfo& info, void* nodes_d, hipStream_t stream);
// Reads `bytes` bytes of a raw file into host_out, or else into dev_out.
Status read_raw(const char* path, size_t bytes, void* host_out, void* dev_out, hipStream_t stream);

}  // namespace pst
