"""Python mirror of the drop-in interface, over libpst_hip.so (ctypes).

The classes keep the reference's names and argument order
(pst/static-pst-gpu.h:76-132, interval-parallel-search/
interval-parallel-search.h:29-32): StaticPSTGPU(pt_arr_d, num_elems, ...)
with threeSidedSearch(min_dim1, max_dim1, min_dim2), twoSidedLeftSearch(
max_dim1, min_dim2), twoSidedRightSearch(min_dim1, min_dim2),
calcGlobalMemNeeded(num_elems), and intervalParallelSearch(...). Device
memory is handled with torch tensors (plumbing only); every search runs in
the HIP library. The module refuses to work without the built library.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpst_hip.so")

DT = {"f32": 0, "f64": 1, "i32": 2, "i64": 3, "u32": 4, "u64": 5, None: -1}
NP = {"f32": np.float32, "f64": np.float64, "i32": np.int32, "i64": np.int64,
      "u32": np.uint32, "u64": np.uint64}
THREE, LEFT, RIGHT = 0, 1, 2
RET_ID, RET_POINT = 0, 1
ERR_CAPACITY = 4

EXPORTED = [
    "pst_last_error", "pst_point_size", "pst_free",
    "pst_gpu_create", "pst_gpu_destroy", "pst_gpu_num_elems", "pst_gpu_dev_ind",
    "pst_gpu_num_devs", "pst_gpu_three_sided_search", "pst_gpu_two_sided_left_search",
    "pst_gpu_two_sided_right_search", "pst_gpu_search_into", "pst_gpu_calc_global_mem_needed",
    "pst_gpu_export_tree", "pst_gpu_export_bytes", "pst_gpu_batch_count", "pst_gpu_batch_write",
    "pst_gpu_batch_search", "pst_gpu_set_search_tuning", "pst_gpu_set_query_grouping",
    "pst_gpu_arr_create", "pst_gpu_arr_calc_global_mem_needed", "pst_gpu_slots_per_tree",
    "pst_calc_num_metacells", "pst_form_metacell_tags", "pst_form_metacells_into",
    "pst_synth_gaussian_volume", "pst_gpu_dtype", "pst_gpu_id_dtype",
    "pst_server_create", "pst_server_destroy", "pst_server_submit", "pst_server_step", "pst_server_poll",
    "pst_server_result", "pst_server_release", "pst_server_outstanding", "pst_server_last_error",
    "pst_server_submit_many", "pst_server_poll_counts", "pst_server_release_many",
    "pst_gpu_autotune", "pst_gpu_set_subtree_copy", "pst_gpu_batch_graph_create", "pst_gpu_batch_graph_launch", "pst_gpu_batch_graph_destroy",
    "pst_interval_parallel_search",
    "pst_interval_parallel_search_into",
    "pst_gpu_save", "pst_gpu_load", "pst_read_vertices", "pst_read_vertices_to_device",
    "pst_gpu_set_level_index", "pst_gpu_level_index_bytes", "pst_gpu_batch_search_noalloc",
    "pst_isosurface_count", "pst_isosurface_write", "pst_isosurface",
    "pst_marching_cubes_count", "pst_marching_cubes_write", "pst_marching_cubes",
]
ERR_IO = 6
DT_NAME = {v: k for k, v in DT.items()}


def point_dtype(t: str, i: Optional[str]) -> np.dtype:
    """numpy view of PointStruct<T, IDType, num_IDs> (C layout)."""
    f = [("dim1", NP[t]), ("dim2", NP[t])]
    if i is not None:
        f.append(("id", NP[i]))
    return np.dtype(f, align=True)


def query_dtype(t: str) -> np.dtype:
    """numpy view of pst_query_<t>."""
    return np.dtype([("min_dim1", NP[t]), ("max_dim1", NP[t]), ("min_dim2", NP[t]),
                     ("kind", np.uint32)], align=True)


_lib = None
_hip = None


class PSTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run "
                               "`python static-priority-search-tree_amd/build.py`")
        # One HIP runtime per process: when torch is present, load it first so
        # that the library binds to the runtime torch already carries.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        c = ctypes
        vp, sz, i32 = c.c_void_p, c.c_size_t, c.c_int
        L.pst_last_error.restype = c.c_char_p
        L.pst_point_size.restype = sz
        L.pst_point_size.argtypes = [i32, i32]
        L.pst_free.argtypes = [vp]
        L.pst_gpu_create.argtypes = [vp, sz, i32, i32, i32, i32, i32, vp, c.POINTER(vp)]
        L.pst_gpu_destroy.argtypes = [vp]
        L.pst_gpu_arr_create.argtypes = [vp, sz, i32, i32, c.c_uint, i32, i32, vp, c.POINTER(vp)]
        L.pst_gpu_arr_calc_global_mem_needed.restype = sz
        L.pst_gpu_arr_calc_global_mem_needed.argtypes = [sz, c.c_uint, i32, i32]
        L.pst_gpu_slots_per_tree.restype = sz
        L.pst_gpu_slots_per_tree.argtypes = [vp]
        L.pst_gpu_autotune.argtypes = [vp, vp, sz, i32, vp, c.POINTER(c.c_int)]
        L.pst_gpu_set_subtree_copy.argtypes = [vp, i32, i32, sz]
        L.pst_gpu_save.argtypes = [vp, c.c_char_p]
        L.pst_gpu_load.argtypes = [c.c_char_p, i32, i32, vp, c.POINTER(vp)]
        L.pst_read_vertices.argtypes = [c.c_char_p, i32, sz, vp]
        L.pst_read_vertices_to_device.argtypes = [c.c_char_p, i32, sz, vp, vp]
        L.pst_gpu_batch_graph_create.argtypes = [vp, vp, sz, i32, vp, vp, vp, sz, vp, c.POINTER(vp)]
        L.pst_gpu_batch_graph_launch.argtypes = [vp, vp]
        L.pst_gpu_batch_graph_destroy.argtypes = [vp]
        L.pst_gpu_dtype.argtypes = [vp]
        L.pst_gpu_id_dtype.argtypes = [vp]
        L.pst_server_create.argtypes = [c.POINTER(vp), i32, sz, i32, c.POINTER(vp)]
        L.pst_server_destroy.argtypes = [vp]
        L.pst_server_submit.argtypes = [vp, i32, vp, i32, c.POINTER(c.c_ulonglong)]
        L.pst_server_step.argtypes = [vp]
        L.pst_server_poll.argtypes = [vp, vp, sz, c.POINTER(sz)]
        L.pst_server_result.argtypes = [vp, c.c_ulonglong, c.POINTER(vp), c.POINTER(sz)]
        L.pst_server_release.argtypes = [vp, c.c_ulonglong]
        L.pst_server_submit_many.argtypes = [vp, i32, vp, sz, i32, vp]
        L.pst_server_poll_counts.argtypes = [vp, vp, vp, sz, c.POINTER(sz)]
        L.pst_server_release_many.argtypes = [vp, vp, sz]
        L.pst_server_outstanding.restype = sz
        L.pst_server_outstanding.argtypes = [vp]
        L.pst_server_last_error.restype = c.c_char_p
        dims3 = c.c_size_t * 3
        L.pst_isosurface_count.argtypes = [vp, i32, dims3, dims3, vp, i32, sz, vp, vp, vp]
        L.pst_isosurface_write.argtypes = [vp, i32, dims3, dims3, vp, i32, sz, vp, vp, vp, sz, vp]
        L.pst_isosurface.argtypes = [vp, i32, dims3, dims3, vp, i32, sz, vp, c.POINTER(vp), c.POINTER(sz)]
        L.pst_marching_cubes_count.argtypes = L.pst_isosurface_count.argtypes
        L.pst_marching_cubes_write.argtypes = L.pst_isosurface_write.argtypes
        L.pst_marching_cubes.argtypes = L.pst_isosurface.argtypes
        L.pst_calc_num_metacells.restype = sz
        L.pst_calc_num_metacells.argtypes = [dims3, dims3, dims3]
        L.pst_form_metacell_tags.argtypes = [vp, i32, i32, dims3, dims3, c.POINTER(vp), c.POINTER(sz), i32, i32]
        L.pst_form_metacells_into.argtypes = [vp, i32, i32, dims3, dims3, vp, vp]
        L.pst_synth_gaussian_volume.argtypes = [vp, dims3, vp, c.c_uint, c.c_float, c.c_ulonglong, vp]
        L.pst_gpu_num_elems.restype = sz
        L.pst_gpu_num_elems.argtypes = [vp]
        L.pst_gpu_three_sided_search.argtypes = [vp, i32, c.POINTER(sz), c.POINTER(vp), vp, vp, vp]
        L.pst_gpu_two_sided_left_search.argtypes = [vp, i32, c.POINTER(sz), c.POINTER(vp), vp, vp]
        L.pst_gpu_two_sided_right_search.argtypes = [vp, i32, c.POINTER(sz), c.POINTER(vp), vp, vp]
        L.pst_gpu_search_into.argtypes = [vp, i32, i32, vp, vp, vp, vp, sz, c.POINTER(sz), vp]
        L.pst_gpu_calc_global_mem_needed.restype = sz
        L.pst_gpu_calc_global_mem_needed.argtypes = [sz, i32, i32]
        L.pst_gpu_export_bytes.restype = sz
        L.pst_gpu_export_bytes.argtypes = [vp]
        L.pst_gpu_export_tree.argtypes = [vp, vp, sz, c.POINTER(sz)]
        L.pst_gpu_batch_count.argtypes = [vp, vp, sz, vp, vp]
        L.pst_gpu_batch_write.argtypes = [vp, vp, sz, i32, vp, vp, vp, vp]
        L.pst_gpu_batch_search.argtypes = [vp, vp, sz, i32, vp, vp, vp, sz, vp]
        L.pst_gpu_set_search_tuning.argtypes = [vp, i32, i32, i32, i32]
        L.pst_gpu_set_query_grouping.argtypes = [vp, i32]
        L.pst_gpu_set_level_index.argtypes = [vp, i32]
        L.pst_gpu_level_index_bytes.restype = sz
        L.pst_gpu_level_index_bytes.argtypes = [sz, i32, i32]
        L.pst_interval_parallel_search.argtypes = [vp, sz, i32, i32, i32, c.POINTER(vp),
                                                   c.POINTER(sz), vp, i32, i32, i32,
                                                   c.c_uint, c.c_uint]
        L.pst_interval_parallel_search_into.argtypes = [vp, sz, i32, i32, i32, vp, vp, sz, vp, vp]
        _lib = L
    return _lib


def hip():
    """libamdhip64 for plain memcpy of library-allocated results."""
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def _check(rc):
    if rc != 0:
        raise PSTError(rc, lib().pst_last_error().decode())


def _scalar(v, t):
    return np.array([v], dtype=NP[t])


def _d2h(ptr, nbytes) -> np.ndarray:
    out = np.empty(max(nbytes, 1), dtype=np.uint8)
    if nbytes:
        rc = hip().hipMemcpy(out.ctypes.data, ptr, nbytes, 2)
        if rc != 0:
            raise PSTError(2, f"hipMemcpy failed ({rc})")
    return out[:nbytes]


def _stream_ptr(stream):
    if stream is None:
        return None
    return ctypes.c_void_p(stream.cuda_stream)


def to_device(arr: np.ndarray, device="cuda"):
    import torch
    b = np.ascontiguousarray(arr).view(np.uint8).reshape(-1)
    return torch.from_numpy(b.copy()).to(device)


class StaticPSTGPU:
    """Drop-in for the reference's StaticPSTGPU (pst/static-pst-gpu.h:72)."""

    def __init__(self, pt_arr_d, num_elems: int, dtype: str = "f32", id_dtype: Optional[str] = "u32",
                 warp_multiplier: int = 1, dev_ind: int = 0, num_devs: int = 1, stream=None):
        self.t, self.i, self.n = dtype, id_dtype, int(num_elems)
        h = ctypes.c_void_p()
        ptr = pt_arr_d.data_ptr() if hasattr(pt_arr_d, "data_ptr") else pt_arr_d
        _check(lib().pst_gpu_create(ptr, self.n, DT[dtype], DT[id_dtype], warp_multiplier,
                                    dev_ind, num_devs, _stream_ptr(stream), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            _check(lib().pst_gpu_destroy(self.h))
            self.h = None

    # ---- persistence (no counterpart in the reference) ----
    def save(self, path: str):
        """Write the built tree (or forest) to `path` (pst_gpu_save)."""
        _check(lib().pst_gpu_save(self.h, os.fsencode(path)))

    @staticmethod
    def load(path: str, dev_ind: int = 0, num_devs: int = 1, stream=None):
        """A tree or forest from a file written by save(), without its input
        points and without rebuilding (pst_gpu_load)."""
        h = ctypes.c_void_p()
        _check(lib().pst_gpu_load(os.fsencode(path), dev_ind, num_devs, _stream_ptr(stream), ctypes.byref(h)))
        L = lib()
        spt = L.pst_gpu_slots_per_tree(h)
        obj = object.__new__(StaticPSTGPUArr if spt else StaticPSTGPU)
        obj.h = h
        obj.t = DT_NAME[L.pst_gpu_dtype(h)]
        obj.i = DT_NAME[L.pst_gpu_id_dtype(h)]
        obj.n = L.pst_gpu_num_elems(h)
        if spt:
            obj.slots_per_tree = spt
        return obj

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- reference-named blocking searches (results copied to host) ----
    def _rec(self, ret):
        if ret == RET_POINT or self.i is None:
            return point_dtype(self.t, self.i), RET_POINT
        return np.dtype(NP[self.i]), RET_ID

    def _finish(self, num, ptr, ret):
        dt, rp = self._rec(ret)
        if not ptr.value:
            return np.zeros(0, dtype=dt)
        raw = _d2h(ptr.value, num.value * dt.itemsize)
        _check(lib().pst_free(ptr.value))
        return raw.view(dt)

    def threeSidedSearch(self, min_dim1_val, max_dim1_val, min_dim2_val, ret=RET_ID):
        num, ptr = ctypes.c_size_t(), ctypes.c_void_p()
        _, rp = self._rec(ret)
        a, b, c = (_scalar(v, self.t) for v in (min_dim1_val, max_dim1_val, min_dim2_val))
        _check(lib().pst_gpu_three_sided_search(self.h, rp, ctypes.byref(num), ctypes.byref(ptr),
                                                a.ctypes.data, b.ctypes.data, c.ctypes.data))
        return self._finish(num, ptr, rp)

    def twoSidedLeftSearch(self, max_dim1_val, min_dim2_val, ret=RET_ID):
        num, ptr = ctypes.c_size_t(), ctypes.c_void_p()
        _, rp = self._rec(ret)
        b, c = _scalar(max_dim1_val, self.t), _scalar(min_dim2_val, self.t)
        _check(lib().pst_gpu_two_sided_left_search(self.h, rp, ctypes.byref(num), ctypes.byref(ptr),
                                                   b.ctypes.data, c.ctypes.data))
        return self._finish(num, ptr, rp)

    def twoSidedRightSearch(self, min_dim1_val, min_dim2_val, ret=RET_ID):
        num, ptr = ctypes.c_size_t(), ctypes.c_void_p()
        _, rp = self._rec(ret)
        a, c = _scalar(min_dim1_val, self.t), _scalar(min_dim2_val, self.t)
        _check(lib().pst_gpu_two_sided_right_search(self.h, rp, ctypes.byref(num), ctypes.byref(ptr),
                                                    a.ctypes.data, c.ctypes.data))
        return self._finish(num, ptr, rp)

    def search(self, kind, min1=0, max1=0, min2=0, ret=RET_ID):
        if kind == THREE:
            return self.threeSidedSearch(min1, max1, min2, ret)
        if kind == LEFT:
            return self.twoSidedLeftSearch(max1, min2, ret)
        return self.twoSidedRightSearch(min1, min2, ret)

    # ---- preallocated output ----
    def search_into(self, kind, min1, max1, min2, out, capacity, ret=RET_ID, stream=None):
        """out: device tensor (or pointer) of `capacity` records. Returns count."""
        num = ctypes.c_size_t()
        _, rp = self._rec(ret)
        a, b, c = (_scalar(v, self.t) for v in (min1, max1, min2))
        ptr = out.data_ptr() if hasattr(out, "data_ptr") else out
        rc = lib().pst_gpu_search_into(self.h, kind, rp, a.ctypes.data, b.ctypes.data, c.ctypes.data,
                                       ptr, capacity, ctypes.byref(num), _stream_ptr(stream))
        if rc not in (0, ERR_CAPACITY):
            _check(rc)
        return num.value

    # ---- batched ----
    def batch_count(self, queries_d, nq, counts_d, stream=None):
        _check(lib().pst_gpu_batch_count(self.h, queries_d.data_ptr(), nq, counts_d.data_ptr(),
                                         _stream_ptr(stream)))

    def batch_write(self, queries_d, nq, offsets_d, out_d, ret=RET_ID, caps_d=None, stream=None):
        _, rp = self._rec(ret)
        _check(lib().pst_gpu_batch_write(self.h, queries_d.data_ptr(), nq, rp, offsets_d.data_ptr(),
                                         caps_d.data_ptr() if caps_d is not None else None,
                                         out_d.data_ptr(), _stream_ptr(stream)))

    def batch_search(self, queries_d, nq, counts_d, offsets_d, out_d, out_capacity, ret=RET_ID,
                     stream=None):
        """Count + scan + write, asynchronous; offsets_d has nq + 1 entries."""
        _, rp = self._rec(ret)
        _check(lib().pst_gpu_batch_search(self.h, queries_d.data_ptr(), nq, rp, counts_d.data_ptr(),
                                          offsets_d.data_ptr(), out_d.data_ptr(), out_capacity,
                                          _stream_ptr(stream)))

    def set_tuning(self, rounds=0, budget_small=0, budget_large=0, waves_per_cu=-1):
        _check(lib().pst_gpu_set_search_tuning(self.h, rounds, budget_small, budget_large, waves_per_cu))

    def autotune(self, queries_d, nq, ret=RET_ID, stream=None):
        """Pick rounds/budgets/waves/grouping for this batch; returns the choice."""
        _, rp = self._rec(ret)
        out = (ctypes.c_int * 6)()
        _check(lib().pst_gpu_autotune(self.h, queries_d.data_ptr(), nq, rp, _stream_ptr(stream), out))
        return dict(zip(("rounds", "budget_small", "budget_large", "waves_per_cu", "queries_per_walk",
                         "us_per_batch"), list(out)))

    def batch_graph(self, queries_d, nq, counts_d, offsets_d, out_d, out_capacity, stream, ret=RET_ID):
        """Capture count + scan + write as a HIP graph on `stream` (not the
        null stream); returns a BatchGraph whose launch(stream) replays it."""
        _, rp = self._rec(ret)
        g = ctypes.c_void_p()
        _check(lib().pst_gpu_batch_graph_create(self.h, queries_d.data_ptr(), nq, rp, counts_d.data_ptr(),
                                                offsets_d.data_ptr(), out_d.data_ptr(), out_capacity,
                                                _stream_ptr(stream), ctypes.byref(g)))
        return BatchGraph(g, self)

    def set_subtree_copy(self, copy_min: int, copy_max: int, max_tasks: int = 0):
        """Write pass: copy fully reported subtrees of copy_min..copy_max points
        (copy_max 0: off) through a task list of max_tasks entries (0: default);
        subtrees found beyond it are copied by the walk."""
        _check(lib().pst_gpu_set_subtree_copy(self.h, copy_min, copy_max, max_tasks))

    def set_level_index(self, on: bool = True):
        """Build (on) or free (off) the level index: per depth, each subtree's
        points sorted by dim2 descending. While built, batched counts and ID
        writes walk only the boundary paths and copy sorted runs."""
        _check(lib().pst_gpu_set_level_index(self.h, 1 if on else 0))

    @staticmethod
    def level_index_bytes(num_elems, dtype="f32", id_dtype="u32"):
        return lib().pst_gpu_level_index_bytes(num_elems, DT[dtype], DT[id_dtype])

    def set_grouping(self, queries_per_walk: int):
        """Queries per tree walk in batched searches: 8 (default) or 1."""
        _check(lib().pst_gpu_set_query_grouping(self.h, queries_per_walk))

    # ---- layout ----
    def export_tree(self):
        """Reference-layout arrays dim1, dim2, med, ids, bits (host numpy)."""
        nb = lib().pst_gpu_export_bytes(self.h)
        blob = np.zeros(max(nb, 8), dtype=np.uint8)
        off = (ctypes.c_size_t * 6)()
        _check(lib().pst_gpu_export_tree(self.h, blob.ctypes.data, nb, off))
        S = off[5]
        tb = np.dtype(NP[self.t]).itemsize
        d = {"dim1": blob[off[0]:off[0] + S * tb].view(NP[self.t]),
             "dim2": blob[off[1]:off[1] + S * tb].view(NP[self.t]),
             "med": blob[off[2]:off[2] + S * tb].view(NP[self.t]),
             "bits": blob[off[4]:off[4] + S]}
        if self.i is not None:
            ib = np.dtype(NP[self.i]).itemsize
            d["ids"] = blob[off[3]:off[3] + S * ib].view(NP[self.i])
        return d

    @staticmethod
    def calcGlobalMemNeeded(num_elems, dtype="f32", id_dtype="u32"):
        return lib().pst_gpu_calc_global_mem_needed(num_elems, DT[dtype], DT[id_dtype])


class BatchGraph:
    """A captured batch. It keeps its tree alive: the tree's search buffers
    and level index cannot be freed or reallocated while a graph exists."""

    def __init__(self, g, tree=None):
        self.g = g
        self.tree = tree

    def launch(self, stream):
        _check(lib().pst_gpu_batch_graph_launch(self.g, _stream_ptr(stream)))

    def close(self):
        if getattr(self, "g", None):
            _check(lib().pst_gpu_batch_graph_destroy(self.g))
            self.g = None
        self.tree = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StaticPSTGPUArr(StaticPSTGPU):
    """Drop-in for the reference's StaticPSTGPUArr (pst/static-pst-gpu-arr.h:45):
    a forest of shallow trees over consecutive chunks of the input. Searches,
    batches and export are those of StaticPSTGPU; the export concatenates the
    trees' reference layouts (see forest_sizes)."""

    def __init__(self, pt_arr_d, num_elems: int, threads_per_block: int, dtype: str = "f32",
                 id_dtype: Optional[str] = "u32", dev_ind: int = 0, num_devs: int = 1, stream=None):
        self.t, self.i, self.n = dtype, id_dtype, int(num_elems)
        h = ctypes.c_void_p()
        ptr = pt_arr_d.data_ptr() if hasattr(pt_arr_d, "data_ptr") else pt_arr_d
        _check(lib().pst_gpu_arr_create(ptr, self.n, DT[dtype], DT[id_dtype], threads_per_block,
                                        dev_ind, num_devs, _stream_ptr(stream), ctypes.byref(h)))
        self.h = h
        self.slots_per_tree = lib().pst_gpu_slots_per_tree(self.h)

    def forest_sizes(self):
        """Slot count of each tree in the export, in order."""
        if self.n == 0:
            return []
        st = self.slots_per_tree
        nt = (self.n + st - 1) // st
        last = self.n - (nt - 1) * st
        return [st] * (nt - 1) + [(1 << int(last).bit_length()) - 1]

    @staticmethod
    def calcGlobalMemNeeded(num_elems, threads_per_block, dtype="f32", id_dtype="u32"):
        return lib().pst_gpu_arr_calc_global_mem_needed(num_elems, threads_per_block, DT[dtype], DT[id_dtype])


def readInVertices(path: str, num_vertices: int, dtype: str = "f32") -> np.ndarray:
    """Drop-in for the reference's readInVertices<T> (helpers/isosurface-data-
    processing.h:205-229): num_vertices values of dtype from a raw file."""
    out = np.empty(num_vertices, dtype=NP[dtype])
    _check(lib().pst_read_vertices(os.fsencode(path), DT[dtype], num_vertices, out.ctypes.data))
    return out


def read_vertices_to_device(path: str, num_vertices: int, out_d, dtype: str = "f32", stream=None):
    """Raw volume file straight into device memory (pst_read_vertices_to_device)."""
    ptr = out_d.data_ptr() if hasattr(out_d, "data_ptr") else out_d
    _check(lib().pst_read_vertices_to_device(os.fsencode(path), DT[dtype], num_vertices, ptr,
                                             _stream_ptr(stream)))


def isosurface(vertex_arr_d, pt_grid_dims, metacell_dims, ids_d, num_ids, isovalue, dtype="f32",
               id_dtype="u32", method="tetrahedra") -> np.ndarray:
    """Triangles (n, 3, 3) of the isosurface in the listed metacells, grid
    coordinates: method "tetrahedra" (pst_isosurface) or "cubes"
    (pst_marching_cubes)."""
    vptr = vertex_arr_d.data_ptr() if hasattr(vertex_arr_d, "data_ptr") else vertex_arr_d
    iptr = ids_d.data_ptr() if hasattr(ids_d, "data_ptr") else ids_d
    dims3 = ctypes.c_size_t * 3
    q = _scalar(isovalue, dtype)
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    fn = lib().pst_marching_cubes if method == "cubes" else lib().pst_isosurface
    _check(fn(vptr, DT[dtype], dims3(*pt_grid_dims), dims3(*metacell_dims), iptr, DT[id_dtype],
              num_ids, q.ctypes.data, ctypes.byref(p), ctypes.byref(n)))
    if not p.value:
        return np.zeros((0, 3, 3), np.float32)
    raw = _d2h(p.value, n.value * 36)
    _check(lib().pst_free(p.value))
    return raw.view(np.float32).reshape(-1, 3, 3)


def isosurface_count(vertex_arr_d, pt_grid_dims, metacell_dims, ids_d, num_ids, isovalue, counts_d,
                     dtype="f32", id_dtype="u32", stream=None, method="tetrahedra"):
    """Triangles per listed metacell (pst_isosurface_count, or
    pst_marching_cubes_count for method "cubes"), asynchronous."""
    dims3 = ctypes.c_size_t * 3
    q = _scalar(isovalue, dtype)
    fn = lib().pst_marching_cubes_count if method == "cubes" else lib().pst_isosurface_count
    _check(fn(vertex_arr_d.data_ptr(), DT[dtype], dims3(*pt_grid_dims), dims3(*metacell_dims),
              ids_d.data_ptr(), DT[id_dtype], num_ids, q.ctypes.data, counts_d.data_ptr(), _stream_ptr(stream)))


def isosurface_write(vertex_arr_d, pt_grid_dims, metacell_dims, ids_d, num_ids, isovalue, offsets_d, tri_d,
                     capacity, dtype="f32", id_dtype="u32", stream=None, method="tetrahedra"):
    """Triangles of the listed metacells at their offsets (pst_isosurface_write,
    or pst_marching_cubes_write for method "cubes")."""
    dims3 = ctypes.c_size_t * 3
    q = _scalar(isovalue, dtype)
    fn = lib().pst_marching_cubes_write if method == "cubes" else lib().pst_isosurface_write
    _check(fn(vertex_arr_d.data_ptr(), DT[dtype], dims3(*pt_grid_dims), dims3(*metacell_dims),
              ids_d.data_ptr(), DT[id_dtype], num_ids, q.ctypes.data, offsets_d.data_ptr(),
              tri_d.data_ptr(), capacity, _stream_ptr(stream)))


def intervalParallelSearch(pt_arr_d, num_elems, search_val, dtype="f32", id_dtype="u32",
                           ret=RET_ID, dev_ind=0, num_devs=1, warp_size=64,
                           num_thread_blocks=0, threads_per_block=256):
    """Drop-in for the reference's intervalParallelSearch (results to host)."""
    if id_dtype is None:
        ret = RET_POINT
    num, ptr = ctypes.c_size_t(), ctypes.c_void_p()
    v = _scalar(search_val, dtype)
    p = pt_arr_d.data_ptr() if hasattr(pt_arr_d, "data_ptr") else pt_arr_d
    _check(lib().pst_interval_parallel_search(p, num_elems, DT[dtype], DT[id_dtype], ret,
                                              ctypes.byref(ptr), ctypes.byref(num), v.ctypes.data,
                                              dev_ind, num_devs, warp_size, num_thread_blocks,
                                              threads_per_block))
    dt = point_dtype(dtype, id_dtype) if ret == RET_POINT else np.dtype(NP[id_dtype])
    raw = _d2h(ptr.value, num.value * dt.itemsize) if ptr.value else np.zeros(0, np.uint8)
    if ptr.value:
        _check(lib().pst_free(ptr.value))
    return raw.view(dt)


def ips_into(pt_arr_d, num_elems, search_val, out_d, capacity, counter_d, dtype="f32",
             id_dtype="u32", ret=RET_ID, stream=None):
    v = _scalar(search_val, dtype)
    _check(lib().pst_interval_parallel_search_into(pt_arr_d.data_ptr(), num_elems, DT[dtype],
                                                   DT[id_dtype], ret, v.ctypes.data,
                                                   out_d.data_ptr(), capacity,
                                                   counter_d.data_ptr(), _stream_ptr(stream)))


# ---- metacells (helpers/isosurface-data-processing.h) ----
def calcNumMetacells(pt_grid_dims, metacell_dims):
    """(number of metacells, metacell grid dims), as calcNumMetacells."""
    d3 = ctypes.c_size_t * 3
    g = d3()
    n = lib().pst_calc_num_metacells(d3(*pt_grid_dims), d3(*metacell_dims), g)
    return n, tuple(g)


def form_metacells_into(vertices_d, pt_grid_dims, metacell_dims, out_d, dtype="f32", id_dtype="u32", stream=None):
    """Metacell intervals (min, max, id) into out_d (one PointStruct per metacell)."""
    d3 = ctypes.c_size_t * 3
    _check(lib().pst_form_metacells_into(vertices_d.data_ptr(), DT[dtype], DT[id_dtype], d3(*pt_grid_dims),
                                         d3(*metacell_dims), out_d.data_ptr(), _stream_ptr(stream)))


def gaussian_volume_params(n_gauss, seed):
    """Host parameters of the synthetic isosurface field: centres U[0,1]^3,
    sigma ~ U[0.02, 0.15], amplitude ~ U[-1, 1] (SplitMix64 streams 10-14
    of pst_data), packed 8 floats per Gaussian as the volume kernel reads them."""
    import pst_data as D
    c = np.stack([D.u01_f64(seed, 10 + a, 0, n_gauss) for a in range(3)], axis=1)
    s = 0.02 + 0.13 * D.u01_f64(seed, 13, 0, n_gauss)
    amp = -1.0 + 2.0 * D.u01_f64(seed, 14, 0, n_gauss)
    p = np.zeros((n_gauss, 8), np.float32)
    p[:, 0:3] = c
    p[:, 3] = -1.0 / (2.0 * s * s)
    p[:, 4] = amp
    return p


def synth_gaussian_volume(out_d, dims, params_d, n_gauss, noise_std=0.01, seed=1, stream=None):
    d3 = ctypes.c_size_t * 3
    _check(lib().pst_synth_gaussian_volume(out_d.data_ptr(), d3(*dims), params_d.data_ptr(), n_gauss,
                                           noise_std, seed, _stream_ptr(stream)))


# ---- query server (continuous batching over resident trees) ----
class Server:
    """Drives the library's query server: submit(tree index, query record),
    step(), poll() -> completed tickets, result(ticket) -> device pointer
    and count, release(ticket)."""

    def __init__(self, trees, out_budget_bytes, max_batch=4096):
        self.trees = list(trees)
        hs = [t.h.value if isinstance(t.h, ctypes.c_void_p) else t.h for t in self.trees]
        arr = (ctypes.c_void_p * len(hs))(*hs)
        h = ctypes.c_void_p()
        rc = lib().pst_server_create(arr, len(self.trees), out_budget_bytes, max_batch, ctypes.byref(h))
        if rc != 0:
            raise PSTError(rc, lib().pst_server_last_error().decode())
        self.h = h
        self._buf = (ctypes.c_ulonglong * 65536)()

    def _chk(self, rc):
        if rc != 0:
            raise PSTError(rc, lib().pst_server_last_error().decode())

    def submit(self, tree, query, ret=RET_ID):
        """query: one record of query_dtype(tree's dtype) (numpy void or 1-element array)."""
        q = np.ascontiguousarray(np.asarray(query).reshape(1))
        t = ctypes.c_ulonglong()
        self._chk(lib().pst_server_submit(self.h, tree, q.ctypes.data, ret, ctypes.byref(t)))
        return t.value

    def step(self):
        self._chk(lib().pst_server_step(self.h))

    def poll(self):
        n = ctypes.c_size_t()
        self._chk(lib().pst_server_poll(self.h, self._buf, len(self._buf), ctypes.byref(n)))
        return list(self._buf[:n.value])

    def result(self, ticket):
        p, n = ctypes.c_void_p(), ctypes.c_size_t()
        self._chk(lib().pst_server_result(self.h, ticket, ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value

    def release(self, ticket):
        self._chk(lib().pst_server_release(self.h, ticket))

    def submit_many(self, tree, queries, ret=RET_ID):
        """queries: contiguous records of query_dtype(tree's dtype); returns their tickets (uint64)."""
        q = np.ascontiguousarray(queries)
        t = np.zeros(len(q), np.uint64)
        self._chk(lib().pst_server_submit_many(self.h, tree, q.ctypes.data, len(q), ret, t.ctypes.data))
        return t

    def poll_counts(self, capacity=1 << 16):
        """Completed tickets and their result counts (~0 for a failed ticket)."""
        t = np.zeros(capacity, np.uint64)
        c = np.zeros(capacity, np.uint64)
        n = ctypes.c_size_t()
        self._chk(lib().pst_server_poll_counts(self.h, t.ctypes.data, c.ctypes.data, capacity, ctypes.byref(n)))
        return t[:n.value], c[:n.value]

    def release_many(self, tickets):
        t = np.ascontiguousarray(tickets, dtype=np.uint64)
        self._chk(lib().pst_server_release_many(self.h, t.ctypes.data, len(t)))

    def outstanding(self):
        return lib().pst_server_outstanding(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().pst_server_destroy(self.h)
            self.h = None
