/*
 * C ABI of libpst_hip.so: the MI355X (gfx950) static priority search tree.
 *
 * The reference (arraylistnumstring/static-priority-search-tree) exposes its
 * hot path as header-only C++ class templates; there is no FFI. This header
 * is the flat boundary those templates are re-implemented over (see
 * include/pst/static-pst-gpu.h, static-pst-gpu-arr.h and
 * interval-parallel-search.h for the drop-in C++ classes). Each entry point
 * names the reference member it replaces (paths relative to the reference).
 *
 * Conventions
 *  - All pointers named *_d are device pointers on the current HIP device.
 *  - Point arrays use the reference's AoS PointStruct<T, IDType, num_IDs>
 *    layout {T dim1_val; T dim2_val; [IDType id;]} with C struct padding
 *    (helpers/point-struct.h:25-33, 107-117); pst_point_size() gives its size.
 *  - Bounds and search values are passed as pointers to ONE host value of
 *    the tree's data type.
 *  - Every function returns PST_OK (0) or an error code; pst_last_error()
 *    then returns a message in the reference's style, "Error in <what> on
 *    device <i> (1-indexed) of <n>: <HIP error string>"
 *    (helpers/gpu-err-chk.h:9-13).
 *  - Results of the blocking searches are returned as a device array the
 *    library allocates with hipMalloc, of num_elems * sizeof(RetType) bytes
 *    as the reference does (pst/static-pst-gpu.tu:563); release it with
 *    pst_free() or hipFree(). Only the first *num_res_elems entries are
 *    valid; their order is unspecified.
 *  - An empty tree prints "Tree is empty; nothing to search" and returns
 *    *num_res_elems = 0, *res_arr_d = NULL (pst/static-pst-gpu.tu:504-510).
 *  - Handles are not shared between threads without external locking; each
 *    handle owns its own device counters (the reference's single global
 *    counter, helpers/dev-symbols.h:6, made searches non re-entrant).
 */
#ifndef PST_CAPI_H
#define PST_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* hipStream_t-compatible opaque stream handle (NULL = default stream). */
typedef struct ihipStream_t* pst_stream_t;

enum pst_status {
  PST_OK = 0,
  PST_ERR_INVALID = 1,  /* bad argument or unsupported type combination */
  PST_ERR_HIP = 2,      /* a HIP runtime call failed */
  PST_ERR_NAN = 3,      /* a dimension (or floating ID) is NaN: no order */
  PST_ERR_CAPACITY = 4, /* results exceeded the caller's buffer (counted only) */
  PST_ERR_INTERNAL = 5,
  PST_ERR_IO = 6        /* a file could not be opened, read or written, or is not a valid tree file */
};

/* Data type codes: T in {F32, F64, I32, I64} (the reference's float,
 * double, int, long: testers/pst-test-info-struct.h:72-119); IDType in all
 * six or PST_NO_ID (num_IDs = 0) (pst-test-info-struct.h:234-284). */
enum pst_dtype {
  PST_F32 = 0,
  PST_F64 = 1,
  PST_I32 = 2,
  PST_I64 = 3,
  PST_U32 = 4,
  PST_U64 = 5,
  PST_NO_ID = -1
};

/* RetType: IDType or the whole PointStruct (pst/static-pst-gpu.h:103-128). */
enum pst_ret { PST_RET_ID = 0, PST_RET_POINT = 1 };

/* Query kinds; bounds are inclusive.
 *  THREE: min_dim1 <= dim1 <= max_dim1 && dim2 >= min_dim2
 *  LEFT : dim1 <= max_dim1 && dim2 >= min_dim2   (min_dim1 ignored)
 *  RIGHT: dim1 >= min_dim1 && dim2 >= min_dim2   (max_dim1 ignored) */
enum pst_query_kind { PST_THREE_SIDED = 0, PST_TWO_SIDED_LEFT = 1, PST_TWO_SIDED_RIGHT = 2 };

/* Batched query records (device arrays), one struct per data type. */
typedef struct { float min_dim1, max_dim1, min_dim2; uint32_t kind; } pst_query_f32;
typedef struct { double min_dim1, max_dim1, min_dim2; uint32_t kind; } pst_query_f64;
typedef struct { int32_t min_dim1, max_dim1, min_dim2; uint32_t kind; } pst_query_i32;
typedef struct { int64_t min_dim1, max_dim1, min_dim2; uint32_t kind; } pst_query_i64;

/* One tree (StaticPSTGPU) or a forest of shallow trees (StaticPSTGPUArr);
 * both are searched, exported and destroyed through the same entry points. */
typedef struct pst_tree pst_tree;

/* Message of the last failing call on this thread ("" if none). */
const char* pst_last_error(void);
/* sizeof(PointStruct<T, IDType, num_IDs>) for the given codes (0 if invalid). */
size_t pst_point_size(int dtype, int id_dtype);
/* Releases a result array returned by a blocking search. */
int pst_free(void* res_arr_d);

/* ------------------------------------------------------------------ */
/* StaticPSTGPU (pst/static-pst-gpu.h:72-477)                          */
/* ------------------------------------------------------------------ */

/* Replaces StaticPSTGPU::StaticPSTGPU(pt_arr_d, num_elems, warp_multiplier,
 * dev_ind, num_devs, dev_props) (pst/static-pst-gpu.h:76-78,
 * pst/static-pst-gpu.tu:19-437). pt_arr_d is read during construction only
 * and not retained. warp_multiplier is accepted for signature parity; the
 * launch geometry is the library's own (wave64). Blocking. */
int pst_gpu_create(const void* pt_arr_d, size_t num_elems, int dtype, int id_dtype,
                   int warp_multiplier, int dev_ind, int num_devs, pst_stream_t stream,
                   pst_tree** out);
/* Replaces ~StaticPSTGPU (pst/static-pst-gpu.h:80-88). */
int pst_gpu_destroy(pst_tree* tree);
/* Getters (pst/static-pst-gpu.h:96-98) and sizes; the dtype getters have
 * no counterpart (the reference's types are template parameters). */
size_t pst_gpu_num_elems(const pst_tree* tree);
int pst_gpu_dtype(const pst_tree* tree);
int pst_gpu_id_dtype(const pst_tree* tree);
int pst_gpu_dev_ind(const pst_tree* tree);
int pst_gpu_num_devs(const pst_tree* tree);

/* Replaces StaticPSTGPU::threeSidedSearch(num_res_elems, res_arr_d,
 * min_dim1_val, max_dim1_val, min_dim2_val, warp_multiplier)
 * (pst/static-pst-gpu.h:108-110, pst/static-pst-gpu.tu:489-539). Blocking. */
int pst_gpu_three_sided_search(pst_tree* tree, int ret, size_t* num_res_elems, void** res_arr_d,
                               const void* min_dim1_val, const void* max_dim1_val,
                               const void* min_dim2_val);
/* Replaces StaticPSTGPU::twoSidedLeftSearch(num_res_elems, res_arr_d,
 * max_dim1_val, min_dim2_val, warp_multiplier)
 * (pst/static-pst-gpu.h:117-119, pst/static-pst-gpu.tu:541-591). Blocking. */
int pst_gpu_two_sided_left_search(pst_tree* tree, int ret, size_t* num_res_elems, void** res_arr_d,
                                  const void* max_dim1_val, const void* min_dim2_val);
/* Replaces StaticPSTGPU::twoSidedRightSearch(num_res_elems, res_arr_d,
 * min_dim1_val, min_dim2_val, warp_multiplier)
 * (pst/static-pst-gpu.h:126-128, pst/static-pst-gpu.tu:593-642). Blocking. */
int pst_gpu_two_sided_right_search(pst_tree* tree, int ret, size_t* num_res_elems, void** res_arr_d,
                                   const void* min_dim1_val, const void* min_dim2_val);

/* One search into a caller-owned device buffer of `capacity` records
 * (no allocation). Blocking; *num_res_elems is the full result count, and
 * PST_ERR_CAPACITY is returned (with the first `capacity` results written)
 * when it exceeds capacity. No counterpart in the reference: the
 * preallocated-output form of the three searches above. */
int pst_gpu_search_into(pst_tree* tree, int kind, int ret, const void* min_dim1_val,
                        const void* max_dim1_val, const void* min_dim2_val, void* out_d,
                        size_t capacity, size_t* num_res_elems, pst_stream_t stream);

/* Replaces StaticPSTGPU::calcGlobalMemNeeded(num_elems)
 * (pst/static-pst-gpu.h:132, pst/static-pst-gpu.tu:646-671): bytes of
 * device memory a tree needs: its resident arrays (node records, subtree
 * minima, IDs) plus the larger of construction's buffers (freed when it
 * ends) and one blocking search (the handle's search buffers, kept, and a
 * result array of num_elems points). A level index, built on request, adds
 * pst_gpu_level_index_bytes on top. */
size_t pst_gpu_calc_global_mem_needed(size_t num_elems, int dtype, int id_dtype);

/* The tree in the reference's memory layout, copied to host memory:
 * dim1[S] dim2[S] med[S] (T), ids[S] (IDType, offset rounded up to its
 * alignment), bitcodes[S] (HAS_LEFT 0x2 | HAS_RIGHT 0x1), S = 2^ceil(log2(n+1))-1
 * (pst/static-pst-gpu.h:316-365). offsets6 receives the byte offsets of the
 * five arrays and S. Used by StaticPSTGPU::print and parity tests. */
int pst_gpu_export_tree(pst_tree* tree, void* host_blob, size_t blob_bytes, size_t* offsets6);
size_t pst_gpu_export_bytes(const pst_tree* tree);

/* ------------------------------------------------------------------ */
/* StaticPSTGPUArr (pst/static-pst-gpu-arr.h:45-366)                   */
/* ------------------------------------------------------------------ */

/* Replaces StaticPSTGPUArr::StaticPSTGPUArr(pt_arr_d, num_elems,
 * threads_per_block, dev_ind, num_devs, dev_props)
 * (pst/static-pst-gpu-arr.h:51-54, pst/static-pst-gpu-arr.tu:17-487): the
 * input is cut into consecutive chunks of St = (smallest power of two above
 * threads_per_block) - 1 points (pst/static-pst-gpu-arr.h:256-261) and each
 * chunk gets its own tree, the one StaticPSTGPU would build for that chunk
 * alone. threads_per_block in [1, 1024]. The returned handle is used with
 * the searches, export, batch entry points and pst_gpu_destroy above and
 * below; searches report from every tree. Right and three-sided searches,
 * empty stubs in the reference (pst/static-pst-gpu-arr.tu:564-578,
 * 645-658), are defined like the other searches. Blocking. */
int pst_gpu_arr_create(const void* pt_arr_d, size_t num_elems, int dtype, int id_dtype,
                       unsigned threads_per_block, int dev_ind, int num_devs, pst_stream_t stream,
                       pst_tree** out);
/* Replaces StaticPSTGPUArr::calcGlobalMemNeeded(num_elems, threads_per_block)
 * (pst/static-pst-gpu-arr.h:103, pst/static-pst-gpu-arr.tu:662-686). */
size_t pst_gpu_arr_calc_global_mem_needed(size_t num_elems, unsigned threads_per_block, int dtype,
                                          int id_dtype);
/* St of a forest (0 for a StaticPSTGPU handle). The forest export holds
 * each tree in the reference's per-tree layout, trees concatenated: tree c
 * at position c*St of every array, the last tree with its own S. */
size_t pst_gpu_slots_per_tree(const pst_tree* tree);

/* ------------------------------------------------------------------ */
/* Tree persistence (no counterpart in the reference, which only       */
/* prints its trees: pst/static-pst-gpu.tu:441-483)                    */
/* ------------------------------------------------------------------ */

/* Writes a built tree or forest to `path`: a 64-byte header (magic
 * "PSTAMD", version 1, kind 0 = StaticPSTGPU / 1 = StaticPSTGPUArr, dtype,
 * id_dtype, n, S, slots per forest tree, checksum) followed by the node
 * records as SoA sections dim1[S], dim2[S], med[S] and, with IDs, id[S], in
 * slot order (slots without a point are zero). Blocking. */
int pst_gpu_save(const pst_tree* tree, const char* path);
/* Creates a handle from a file written by pst_gpu_save, without the input
 * points and without rebuilding: a StaticPSTGPU or StaticPSTGPUArr handle
 * as the file says (pst_gpu_slots_per_tree tells which). The header is
 * checked against the file size and the sections against the checksum
 * (PST_ERR_IO on any mismatch). Blocking; searches then run on `stream`. */
int pst_gpu_load(const char* path, int dev_ind, int num_devs, pst_stream_t stream, pst_tree** out);

/* ------------------------------------------------------------------ */
/* Metacell formation (helpers/isosurface-data-processing.h:21-137)    */
/* ------------------------------------------------------------------ */

/* Replaces calcNumMetacells(pt_grid_dims, metacell_dims,
 * metacell_grid_dims, num_metacells) (isosurface-data-processing.h:21-36):
 * ceil((pt_grid_dims[d] - 1) / metacell_dims[d]) metacells along axis d;
 * returns their product (0 on invalid dimensions). */
size_t pst_calc_num_metacells(const size_t pt_grid_dims[3], const size_t metacell_dims[3],
                              size_t metacell_grid_dims[3]);
/* Replaces formMetacellTags<PointStruct, T, GridDimType>(vertex_arr_d,
 * pt_grid_dims, metacell_dims, metacell_grid_dims, num_metacells, dev_ind,
 * num_devs, warp_size) (isosurface-data-processing.h:66-137,
 * isosurface-data-processing.tu:7-127): vertex_arr_d is a T grid, x fastest;
 * metacell (a, b, c) gets (min, max) over the vertices
 * [a*mx, min((a+1)*mx, px-1)] x ... and ID a + (b + c*gy)*gx, stored at
 * index ID of a newly allocated PointStruct array (free with pst_free).
 * Blocking. For NaN-free volumes. */
int pst_form_metacell_tags(const void* vertex_arr_d, int dtype, int id_dtype, const size_t pt_grid_dims[3],
                           const size_t metacell_dims[3], void** metacell_arr_d, size_t* num_metacells,
                           int dev_ind, int num_devs);
/* The same into a caller-owned array, asynchronous on `stream`. No
 * counterpart in the reference. */
int pst_form_metacells_into(const void* vertex_arr_d, int dtype, int id_dtype, const size_t pt_grid_dims[3],
                            const size_t metacell_dims[3], void* metacell_arr_d, pst_stream_t stream);
/* Replaces readInVertices<T>(input_filename, num_vertices)
 * (helpers/isosurface-data-processing.h:205-229): reads num_vertices
 * values of dtype (native byte order, x fastest) from a raw binary file into
 * the caller's host array. A file that cannot be opened gives "Failed to
 * open <file> for reading", one that is too short "Failure: read <have>
 * bytes, expected <want> bytes" (PST_ERR_IO; the reference prints these and
 * exits with FILE_OPEN_ERR / INPUT_READ_ERR). */
int pst_read_vertices(const char* path, int dtype, size_t num_vertices, void* vertex_arr);
/* The same straight into device memory (vertex_arr_d holds num_vertices
 * values), through two pinned staging buffers so that reading overlaps the
 * copy. No counterpart in the reference. Blocking. */
int pst_read_vertices_to_device(const char* path, int dtype, size_t num_vertices, void* vertex_arr_d,
                                pst_stream_t stream);

/* Isosurface of the active metacells (no counterpart in the reference,
 * which stops at the metacell IDs of an isovalue search and copies them back
 * "for on-device marching cubes", testers/pst-tester-functions.tu:230-252).
 * Marching tetrahedra: each voxel of each listed metacell is cut into the
 * 6 tetrahedra {0, e_a, e_a + e_b, 7}; a vertex is below when its value is
 * < isovalue; surface vertices interpolate linearly along grid edges from
 * the endpoint of lower linear index (shared edges give identical vertices,
 * so the mesh has no cracks); triangles are wound with their normal toward
 * values >= isovalue. Triangles are 9 floats (x, y, z of 3 vertices) in grid
 * coordinates. The metacell grid is that of pst_calc_num_metacells;
 * metacell_ids_d holds IDs of an ID type of 4 or 8 bytes (the search
 * results), isovalue points to one value of dtype.
 *   pst_isosurface_count: counts_d[i] = triangles of metacell ids[i]
 *                          (an ID past the grid counts 0), asynchronous;
 *   pst_isosurface_write: metacell ids[i]'s triangles at triangles_d +
 *                          9 * offsets_d[i] (in the order: 64-voxel chunk,
 *                          tetrahedron, voxel), triangles at or past index
 *                          `capacity` not written, asynchronous;
 *   pst_isosurface:       both, blocking, into an array allocated by the
 *                          call (free with pst_free); PST_ERR_INVALID for an
 *                          ID past the grid. */
int pst_isosurface_count(const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3],
                         const size_t metacell_dims[3], const void* metacell_ids_d, int id_dtype, size_t num_ids,
                         const void* isovalue, unsigned long long* counts_d, pst_stream_t stream);
int pst_isosurface_write(const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3],
                         const size_t metacell_dims[3], const void* metacell_ids_d, int id_dtype, size_t num_ids,
                         const void* isovalue, const unsigned long long* offsets_d, float* triangles_d,
                         size_t capacity, pst_stream_t stream);
int pst_isosurface(const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3], const size_t metacell_dims[3],
                   const void* metacell_ids_d, int id_dtype, size_t num_ids, const void* isovalue,
                   float** triangles_d, size_t* num_triangles);

/* Marching cubes over the active metacells: the consumer the reference
 * anticipates for the IDs of an isovalue search (its sample
 * .marching-cubes/marchingCubes_kernel.cu:155-434 classifies voxels, scans
 * their triangle counts and generates the triangles of a 256-case table).
 * Same arguments, ID rules, output layout (9 floats per triangle, grid
 * coordinates) and blocking/asynchronous forms as pst_isosurface_* above,
 * with one patch per voxel instead of six tetrahedra: corner k of a voxel
 * (x = k & 1, y = k >> 1 & 1, z = k >> 2) is below when its value is <
 * isovalue; on each cube face every run of below corners is cut off by its
 * own segment (two diagonal below corners stay apart), so neighbouring
 * voxels cut shared faces alike and the mesh has no cracks; the segments
 * close into loops that are fanned into triangles wound with their normal
 * toward values >= isovalue; surface vertices interpolate along grid edges
 * from the endpoint of lower linear index. Triangles of a metacell come in
 * voxel order (x fastest). At most 5 triangles per voxel. */
int pst_marching_cubes_count(const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3],
                             const size_t metacell_dims[3], const void* metacell_ids_d, int id_dtype, size_t num_ids,
                             const void* isovalue, unsigned long long* counts_d, pst_stream_t stream);
int pst_marching_cubes_write(const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3],
                             const size_t metacell_dims[3], const void* metacell_ids_d, int id_dtype, size_t num_ids,
                             const void* isovalue, const unsigned long long* offsets_d, float* triangles_d,
                             size_t capacity, pst_stream_t stream);
int pst_marching_cubes(const void* vertex_arr_d, int dtype, const size_t pt_grid_dims[3],
                       const size_t metacell_dims[3], const void* metacell_ids_d, int id_dtype, size_t num_ids,
                       const void* isovalue, float** triangles_d, size_t* num_triangles);

/* Synthetic f32 volume (no counterpart in the reference; a seeded input of
 * the isosurface workload): value = sum_g amp_g * exp(-|p - c_g|^2 /
 * (2 s_g^2)) + noise_std * N(0, 1) at p = (i, j, k) / (dims - 1), x fastest.
 * gaussians_d: device floats, 8 per Gaussian: cx, cy, cz, -1/(2 s^2), amp,
 * 0, 0, 0. The normal draw of vertex v comes from SplitMix64 over (seed, v),
 * so the volume depends only on the inputs. At most 256 Gaussians. */
int pst_synth_gaussian_volume(float* out_d, const size_t dims[3], const float* gaussians_d, unsigned num_gaussians,
                              float noise_std, unsigned long long seed, pst_stream_t stream);

/* ---- Batched searches (no counterpart in the reference) ----
 * queries_d: device array of pst_query_<dtype>. Count pass then write pass:
 *   pst_gpu_batch_count   counts_d[q] = number of results of query q
 *   pst_gpu_batch_write   writes query q's results at out_d + offsets_d[q]
 *                          (records of RetType), at most caps_d[q] of them
 *                          (caps_d may be NULL: no limit beyond offsets).
 * Both are asynchronous on `stream` and allocate nothing after the first
 * call with a given batch size (scratch is kept by the handle). */
int pst_gpu_batch_count(pst_tree* tree, const void* queries_d, size_t num_queries,
                        unsigned long long* counts_d, pst_stream_t stream);
int pst_gpu_batch_write(pst_tree* tree, const void* queries_d, size_t num_queries, int ret,
                        const unsigned long long* offsets_d, const unsigned long long* caps_d,
                        void* out_d, pst_stream_t stream);
/* Count, scan and write in one asynchronous call: counts_d[q] (nq
 * entries) and offsets_d (nq + 1 entries, offsets_d[nq] = total) are
 * written on the device; results go to out_d, of which only the first
 * out_capacity records are written (compare offsets_d[nq] with it).
 * out_capacity = 0 is a sizing call: counts and offsets only, out_d is not
 * touched (it may be NULL). No host synchronisation and no allocation after
 * the first call for a batch size, so the call can be captured into a HIP
 * graph. */
int pst_gpu_batch_search(pst_tree* tree, const void* queries_d, size_t num_queries, int ret,
                         unsigned long long* counts_d, unsigned long long* offsets_d, void* out_d,
                         size_t out_capacity, pst_stream_t stream);
/* pst_gpu_batch_search through the level index with no device allocation
 * at all: no counterpart in the reference; it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Same arguments and results as
 * pst_gpu_batch_search with ret = PST_RET_ID (IDs only; out_capacity = 0 is
 * the sizing call), for a deep tree whose level index is built
 * (pst_gpu_set_level_index; PST_ERR_INVALID otherwise, and for forests or
 * trees without IDs). Its search buffers are one file-scope device array
 * per GPU, 16.25 MiB, held from the library's load until the process ends,
 * which limits a call to at most 8192 queries (PST_ERR_INVALID beyond).
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (other streams, other threads) race and must be serialised by the caller.
 * Asynchronous on `stream`; nothing is cleared or synchronised. */
int pst_gpu_batch_search_noalloc(pst_tree* tree, const void* queries_d, size_t num_queries,
                                 unsigned long long* counts_d, unsigned long long* offsets_d, void* out_d,
                                 size_t out_capacity, pst_stream_t stream);
/* Tuning knobs of the search rounds (0 keeps the default). */
int pst_gpu_set_search_tuning(pst_tree* tree, int rounds, int budget_small, int budget_large,
                              int waves_per_cu);
/* Autotuner (no counterpart in the reference): times pst_gpu_batch_search
 * on this batch over waves per CU {16, 24, 32, 40} x queries per walk
 * {8, 1} x spill-round step budget {128, 384, 1024}, keeps the fastest
 * setting on the tree and reports it in chosen6 (nullable): rounds,
 * budget_small, budget_large, waves_per_cu, queries per walk, microseconds
 * per batch. Blocking; allocates its own count/offset/output buffers. */
int pst_gpu_autotune(pst_tree* tree, const void* queries_d, size_t num_queries, int ret, pst_stream_t stream,
                     int* chosen6);
/* A batched search (count, scan, write; pst_gpu_batch_search with these
 * exact arguments) captured as a HIP graph (no counterpart in the
 * reference). Runs the batch once eagerly (all allocation happens then),
 * then captures it on `stream` (not the null stream). At most 8192 queries:
 * such batches are grouped by one kernel, with no library call. Replays read
 * the current contents of queries_d and rewrite counts_d, offsets_d, out_d.
 * The graph holds the addresses of the tree's search buffers and level
 * index. While any graph of a tree exists, the calls on that tree that would
 * free or reallocate them fail with PST_ERR_INVALID and change nothing: a
 * batch larger than the buffers hold, pst_gpu_autotune,
 * pst_gpu_set_level_index(tree, 0) and pst_gpu_destroy. Destroy the graphs
 * first. */
typedef struct pst_graph pst_graph;
int pst_gpu_batch_graph_create(pst_tree* tree, const void* queries_d, size_t num_queries, int ret,
                               unsigned long long* counts_d, unsigned long long* offsets_d, void* out_d,
                               size_t out_capacity, pst_stream_t stream, pst_graph** out);
int pst_gpu_batch_graph_launch(pst_graph* graph, pst_stream_t stream);
int pst_gpu_batch_graph_destroy(pst_graph* graph);
/* Full-subtree copies of the write pass (no counterpart in the reference):
 * a subtree whose every point a query reports, with copy_min..copy_max
 * points, is not walked but recorded as a copy task; after the walk a copy
 * kernel writes each task's subtree as one block (one reservation per
 * query, breadth-first loads from a per-slot ID array). copy_max = 0 turns
 * copies off. Defaults 256..65535. max_tasks bounds the task list (0: the
 * default of 4M entries, also the most); subtrees found once it is full are
 * copied by the walk itself. Results are the same either way. */
int pst_gpu_set_subtree_copy(pst_tree* tree, int copy_min, int copy_max, size_t max_tasks);
/* Queries per tree walk in batched searches: 8 (default; a batch is sorted
 * by leading dim1 bound and each walk serves 8 neighbouring queries, with
 * each query's results exactly those of its own walk) or 1 (one walk per
 * query). No counterpart in the reference. */
int pst_gpu_set_query_grouping(pst_tree* tree, int queries_per_walk);
/* Level index of a deep tree with IDs (no counterpart in the reference):
 * on = 1 builds it (blocking; a no-op when built), on = 0 frees it. For
 * every depth t it holds the points of each subtree rooted at depth t
 * sorted by dim2 descending, so a subtree whose dim1 range lies inside a
 * query's bounds reports a prefix of its sorted run: one binary search and
 * one contiguous copy. While it exists, pst_gpu_batch_count, and the ID
 * returning pst_gpu_batch_write / pst_gpu_batch_search (and the batch
 * graphs made from them), single searches and pst_gpu_search_into walk
 * only the query's boundary paths and stream the IDs; results are the same
 * sets as without it (order within a query's results differs). Trees of
 * fewer than 2^31 points; not for a forest or a tree without IDs
 * (PST_ERR_INVALID). Device memory: pst_gpu_level_index_bytes, plus
 * per-handle batch scratch of about 1.5 KB per query of the largest batch
 * written (128 ranges of 12 bytes). */
int pst_gpu_set_level_index(pst_tree* tree, int on);
/* Bytes of the level index for n points (bit_length(n) levels of n key/ID
 * pairs, each level padded to a multiple of 64); 0 for n = 0 or a tree
 * without IDs. */
size_t pst_gpu_level_index_bytes(size_t n, int dtype, int id_dtype);

/* ------------------------------------------------------------------ */
/* Query server (no counterpart in the reference)                      */
/* ------------------------------------------------------------------ */

/* Continuous batching over resident trees with admission by output memory.
 * Queries (host pst_query_<dtype> records) are submitted per tree; each
 * pst_server_step, without blocking: retires finished write batches,
 * collects finished count passes, admits counted queries in arrival order
 * while their exact outputs fit one range of the out_budget_bytes pool,
 * launches their write pass, and starts a count pass (up to max_batch
 * queries) for newly submitted ones. Each tree works on its own stream, so
 * batches of different trees overlap on the device. A completed ticket's
 * results (pst_server_result: device pointer, count) stay valid until
 * pst_server_release. Each result equals the same query searched alone.
 * Not thread-safe: one host thread drives a server. */
typedef struct pst_server pst_server;
int pst_server_create(pst_tree* const* trees, int num_trees, size_t out_budget_bytes, int max_batch,
                      pst_server** out);
int pst_server_destroy(pst_server* server);
int pst_server_submit(pst_server* server, int tree, const void* query, int ret, unsigned long long* ticket);
int pst_server_step(pst_server* server);
int pst_server_poll(pst_server* server, unsigned long long* tickets, size_t capacity, size_t* num);
int pst_server_result(pst_server* server, unsigned long long ticket, void** res_d, size_t* num_res_elems);
int pst_server_release(pst_server* server, unsigned long long ticket);
/* Batched forms of submit, poll and release for drivers that handle many
 * queries per loop: num records of the tree's query type submitted in
 * order (tickets[i] for records[i]; stops at the first invalid record),
 * completed tickets with their result counts (~0 for a ticket whose output
 * exceeds the server's budget), and releases of num tickets. */
int pst_server_submit_many(pst_server* server, int tree, const void* queries, size_t num, int ret,
                           unsigned long long* tickets);
int pst_server_poll_counts(pst_server* server, unsigned long long* tickets, unsigned long long* counts,
                           size_t capacity, size_t* num);
int pst_server_release_many(pst_server* server, const unsigned long long* tickets, size_t num);
/* Create the server after building the trees' level indexes: creation
 * sizes every tree's search buffers for max_batch queries (so serving does
 * not allocate), and an index built later sizes its own on first use. */
/* Tickets submitted and not yet released. */
size_t pst_server_outstanding(const pst_server* server);
const char* pst_server_last_error(void);

/* ------------------------------------------------------------------ */
/* Interval parallel search (interval-parallel-search/                 */
/* interval-parallel-search.h:23-66)                                   */
/* ------------------------------------------------------------------ */

/* Replaces intervalParallelSearch<T, PointStructTemplate, IDType, num_IDs,
 * RetType>(pt_arr_d, num_elems, res_arr_d, num_res_elems, search_val,
 * dev_ind, num_devs, warp_size, num_thread_blocks, threads_per_block):
 * reports points with dim1 <= search_val <= dim2. warp_size and
 * threads_per_block are accepted for signature parity; num_thread_blocks =
 * 0 lets the library size the grid. Blocking; allocates
 * num_elems * sizeof(RetType) for the result like the reference. */
int pst_interval_parallel_search(const void* pt_arr_d, size_t num_elems, int dtype, int id_dtype,
                                 int ret, void** res_arr_d, size_t* num_res_elems,
                                 const void* search_val, int dev_ind, int num_devs, int warp_size,
                                 unsigned num_thread_blocks, unsigned threads_per_block);
/* Same filter into a caller buffer; asynchronous. counter_d (one
 * unsigned long long, zeroed by the call) receives the hit count. */
int pst_interval_parallel_search_into(const void* pt_arr_d, size_t num_elems, int dtype,
                                      int id_dtype, int ret, const void* search_val, void* out_d,
                                      size_t capacity, unsigned long long* counter_d,
                                      pst_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* PST_CAPI_H */
