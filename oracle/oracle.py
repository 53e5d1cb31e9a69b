"""CPU oracle for the PST hot path (Python side).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker, never as the thing
measured for the GPU path or shipped.

Pieces
  * ctypes bindings to oracle/_build/liboracle.so, the C++ restatement of the
    reference's StaticPSTCPUIter (construction + 3 stack searches), see
    oracle/pst_oracle.hpp for the file:line map;
  * brute_force(): the closed-form predicates of the three searches and of
    interval-parallel search (pst/static-pst-cpu-iter.tpp:355-358, 446-449,
    262-264; interval-parallel-search/interval-parallel-search.h:94-96);
  * closed_form_tree(): the construction written independently from the
    rule "node = max key2 of the set; the other m-1 points sorted by key1;
    the first floor(m/2) go left; med = dim1 of the floor(m/2)-th of them"
    (derived from pst/static-pst-cpu-iter.tpp:512-545, 579-597);
  * run_reference(): runs oracle/_ref/pst_ref (the reference's own code,
    compiled by `make -C oracle ref`) and parses its printed tree/results.
"""
from __future__ import annotations

import ctypes
import os
import struct
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "pst_ref")
REF_BENCH = os.path.join(HERE, "_ref", "ref_bench")

# dtype codes shared with include/pst/pst_capi.h
DT = {"f32": 0, "f64": 1, "i32": 2, "i64": 3, "u32": 4, "u64": 5, None: -1}
NP = {"f32": np.float32, "f64": np.float64, "i32": np.int32, "i64": np.int64,
      "u32": np.uint32, "u64": np.uint64}
THREE, LEFT, RIGHT = 0, 1, 2


def point_dtype(t: str, i: Optional[str]) -> np.dtype:
    """numpy view of the reference's AoS PointStruct<T, IDType, num_IDs>."""
    fields = [("dim1", NP[t]), ("dim2", NP[t])]
    if i is not None:
        fields.append(("id", NP[i]))
    return np.dtype(fields, align=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-C", HERE, "oracle"], check=True,
                           stdout=subprocess.DEVNULL)
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_tree_bytes.restype = ctypes.c_size_t
        L.oracle_tree_bytes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t]
        L.oracle_tree_layout.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t,
                                         ctypes.POINTER(ctypes.c_size_t)]
        L.oracle_build.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_size_t, ctypes.c_void_p]
        L.oracle_search.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                    ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.oracle_time_queries.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                          ctypes.POINTER(ctypes.c_double)]
        L.oracle_time_build.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_size_t, ctypes.c_void_p,
                                        ctypes.POINTER(ctypes.c_double)]
        L.oracle_count_queries.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                           ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        _lib = L
    return _lib


class Tree:
    """A reference-layout tree blob with named views of its sub-arrays."""

    def __init__(self, t: str, i: Optional[str], n: int, blob: np.ndarray):
        self.t, self.i, self.n, self.blob = t, i, n, blob
        off = (ctypes.c_size_t * 6)()
        lib().oracle_tree_layout(DT[t], DT[i], n, off)
        self.S = off[5]
        tb = np.dtype(NP[t]).itemsize
        S = self.S
        self.dim1 = blob[off[0]:off[0] + S * tb].view(NP[t])
        self.dim2 = blob[off[1]:off[1] + S * tb].view(NP[t])
        self.med = blob[off[2]:off[2] + S * tb].view(NP[t])
        if i is not None:
            ib = np.dtype(NP[i]).itemsize
            self.ids = blob[off[3]:off[3] + S * ib].view(NP[i])
        else:
            self.ids = None
        self.bits = blob[off[4]:off[4] + S]

    def arrays(self):
        d = {"dim1": self.dim1, "dim2": self.dim2, "med": self.med, "bits": self.bits}
        if self.ids is not None:
            d["ids"] = self.ids
        return d


def empty_tree(t: str, i: Optional[str], n: int) -> Tree:
    nbytes = lib().oracle_tree_bytes(DT[t], DT[i], n)
    return Tree(t, i, n, np.zeros(max(nbytes, 8), dtype=np.uint8))


def build(pts: np.ndarray, t: str, i: Optional[str]) -> Tree:
    pts = np.ascontiguousarray(pts)
    assert pts.dtype == point_dtype(t, i)
    tr = empty_tree(t, i, len(pts))
    lib().oracle_build(DT[t], DT[i], pts.ctypes.data, len(pts), tr.blob.ctypes.data)
    return tr


def search(tr: Tree, kind: int, min1=0, max1=0, min2=0, ret_point: bool = False):
    """Runs one oracle search; returns IDs (or points) in the walk's order."""
    T = NP[tr.t]
    b = [np.array([v], dtype=T) for v in (min1, max1, min2)]
    if ret_point or tr.i is None:
        odt = point_dtype(tr.t, tr.i)
        rp = 1
    else:
        odt = np.dtype(NP[tr.i])
        rp = 0
    num = ctypes.c_size_t(0)
    cap = max(tr.n, 1)
    out = np.zeros(cap, dtype=odt)
    lib().oracle_search(DT[tr.t], DT[tr.i], tr.blob.ctypes.data, tr.n, kind,
                        b[0].ctypes.data, b[1].ctypes.data, b[2].ctypes.data, rp,
                        out.ctypes.data, cap, ctypes.byref(num))
    return out[:num.value]


def time_queries(tr: Tree, kinds, min1s, max1s, min2s, threads: int = 1):
    """Times the oracle's stack walks over a query list (CPU baseline)."""
    T = NP[tr.t]
    kinds = np.ascontiguousarray(kinds, dtype=np.int32)
    a = np.ascontiguousarray(min1s, dtype=T)
    b = np.ascontiguousarray(max1s, dtype=T)
    c = np.ascontiguousarray(min2s, dtype=T)
    tot = ctypes.c_uint64(0)
    secs = ctypes.c_double(0)
    lib().oracle_time_queries(DT[tr.t], DT[tr.i], tr.blob.ctypes.data, tr.n, len(kinds),
                              kinds.ctypes.data, a.ctypes.data, b.ctypes.data, c.ctypes.data,
                              threads, ctypes.byref(tot), ctypes.byref(secs))
    return tot.value, secs.value


def count_queries(tr: Tree, kinds, min1s, max1s, min2s, threads: int = 1) -> np.ndarray:
    """Per-query result counts of the oracle's walks (kinds 0 three-sided,
    1 left, 2 right), over `threads` threads."""
    T = NP[tr.t]
    kinds = np.ascontiguousarray(kinds, dtype=np.int32)
    a = np.ascontiguousarray(min1s, dtype=T)
    b = np.ascontiguousarray(max1s, dtype=T)
    c = np.ascontiguousarray(min2s, dtype=T)
    out = np.zeros(len(kinds), np.uint64)
    lib().oracle_count_queries(DT[tr.t], DT[tr.i], tr.blob.ctypes.data, tr.n, len(kinds), kinds.ctypes.data,
                               a.ctypes.data, b.ctypes.data, c.ctypes.data, threads, out.ctypes.data)
    return out.astype(np.int64)


def time_build(pts: np.ndarray, t: str, i: Optional[str]):
    tr = empty_tree(t, i, len(pts))
    secs = ctypes.c_double(0)
    lib().oracle_time_build(DT[t], DT[i], np.ascontiguousarray(pts).ctypes.data, len(pts),
                            tr.blob.ctypes.data, ctypes.byref(secs))
    return tr, secs.value


# ---------------------------------------------------------------------------
# Brute force and closed form
# ---------------------------------------------------------------------------
def brute_force(pts: np.ndarray, kind: int, min1=0, max1=0, min2=0) -> np.ndarray:
    """Indices of points satisfying the query predicate (inclusive bounds)."""
    d1, d2 = pts["dim1"], pts["dim2"]
    T = d1.dtype.type
    min1, max1, min2 = T(min1), T(max1), T(min2)
    if kind == THREE:
        m = (min1 <= d1) & (d1 <= max1) & (min2 <= d2)
    elif kind == LEFT:
        m = (d1 <= max1) & (min2 <= d2)
    else:
        m = (d1 >= min1) & (min2 <= d2)
    return np.nonzero(m)[0]


def brute_force_ips(pts: np.ndarray, q) -> np.ndarray:
    """interval-parallel-search.h:94-96: dim1 <= q <= dim2."""
    d1, d2 = pts["dim1"], pts["dim2"]
    q = d1.dtype.type(q)
    return np.nonzero((d1 <= q) & (q <= d2))[0]


def _key_order(vals: np.ndarray, ids: Optional[np.ndarray]) -> np.ndarray:
    """argsort by (value, id, index) ascending with IEEE semantics
    (-0.0 == +0.0); used as the key1/key2 total orders."""
    idx = np.arange(len(vals))
    v = vals.astype(vals.dtype, copy=True)
    if v.dtype.kind == "f":
        v = np.where(v == 0, v.dtype.type(0), v)  # -0.0 ties +0.0
    keys = [idx]
    if ids is not None:
        iv = ids.copy()
        if iv.dtype.kind == "f":
            iv = np.where(iv == 0, iv.dtype.type(0), iv)
        keys.append(iv)
    keys.append(v)
    return np.lexsort(keys)


def shape_sizes(n: int) -> np.ndarray:
    """Subtree point count for every slot (data-independent tree shape):
    a node holding m points gives floor(m/2) to the left child and
    ceil(m/2)-1 to the right child."""
    S = (1 << n.bit_length()) - 1 if n else 0
    m = np.zeros(S, dtype=np.int64)
    if n:
        m[0] = n
        for s in range(S):
            if 2 * s + 2 < S + 1 and m[s] > 0:
                if 2 * s + 1 < S:
                    m[2 * s + 1] = m[s] // 2
                if 2 * s + 2 < S:
                    m[2 * s + 2] = (m[s] + 1) // 2 - 1
    return m


def closed_form_tree(pts: np.ndarray, has_id: bool):
    """Independent construction from the closed-form rule (small n only).
    Returns dict of arrays dim1, dim2, med, ids (or None), bits."""
    n = len(pts)
    S = (1 << n.bit_length()) - 1 if n else 0
    d1, d2 = pts["dim1"], pts["dim2"]
    ids = pts["id"] if has_id else None
    ord1 = _key_order(d1, ids)
    rank1 = np.empty(n, dtype=np.int64)
    rank1[ord1] = np.arange(n)
    ord2 = _key_order(d2, ids)
    rank2 = np.empty(n, dtype=np.int64)
    rank2[ord2] = np.arange(n)
    out = {"dim1": np.zeros(S, d1.dtype), "dim2": np.zeros(S, d1.dtype),
           "med": np.zeros(S, d1.dtype), "bits": np.zeros(S, np.uint8)}
    out["ids"] = np.zeros(S, ids.dtype) if has_id else None
    stack = [(0, np.arange(n))] if n else []
    while stack:
        slot, members = stack.pop()
        node = members[np.argmax(rank2[members])]
        rest = members[members != node]
        rest = rest[np.argsort(rank1[rest], kind="stable")]
        m = len(members)
        L = m // 2
        out["dim1"][slot] = d1[node]
        out["dim2"][slot] = d2[node]
        if has_id:
            out["ids"][slot] = ids[node]
        out["med"][slot] = d1[rest[L - 1]] if L >= 1 else d1[node]
        left, right = rest[:L], rest[L:]
        bc = (2 if len(left) else 0) | (1 if len(right) else 0)
        out["bits"][slot] = bc
        if len(left):
            stack.append((2 * slot + 1, left))
        if len(right):
            stack.append((2 * slot + 2, right))
    return out


# ---------------------------------------------------------------------------
# The reference's own CPU tree (oracle/_ref/pst_ref)
# ---------------------------------------------------------------------------
def ref_available() -> bool:
    return os.path.exists(REF_BIN)


def _parse_val(s: str, T):
    if np.dtype(T).kind == "f":
        return T(float(s))
    return T(int(s))


def run_reference(pts: np.ndarray, t: str, i: Optional[str], queries, workdir: str):
    """queries: list of (kind, ret_point, min1, max1, min2).
    Returns (tree arrays dict, list of result arrays in the reference order)."""
    T = NP[t]
    n = len(pts)
    path = os.path.join(workdir, "ref_in.bin")
    with open(path, "wb") as f:
        f.write(b"PSTR" + struct.pack("<iiQQ", DT[t], DT[i], n, len(queries)))
        f.write(np.ascontiguousarray(pts).tobytes())
        for (kind, rp, a, b, c) in queries:
            f.write(struct.pack("<ii", kind, rp))
            f.write(np.array([a, b, c], dtype=T).tobytes())
    txt = subprocess.run([REF_BIN, path], check=True, capture_output=True, text=True).stdout
    lines = txt.split("\n")
    assert lines[0] == "TREE"
    end = lines.index("ENDTREE")
    tree_lines = [ln for ln in lines[1:end] if ln.strip()]
    S = (1 << n.bit_length()) - 1 if n else 0
    IT = NP[i] if i is not None else None
    arrs = {"dim1": np.zeros(S, T), "dim2": np.zeros(S, T), "med": np.zeros(S, T),
            "bits": np.zeros(S, np.uint8)}
    arrs["ids"] = np.zeros(S, IT) if i is not None else None
    if n:
        _parse_tree(tree_lines, arrs, T, IT)
    results = []
    k = end + 1
    for _ in queries:
        # an empty tree prints "Tree is empty; nothing to search" per search
        while not lines[k].startswith("Q "):
            k += 1
        hdr = lines[k].split()
        cnt = int(hdr[2])
        rows = lines[k + 1:k + 1 + cnt]
        results.append(rows)
        k += 1 + cnt
    return arrs, results


def ref_bench_available() -> bool:
    return os.path.exists(REF_BENCH)


def time_reference(pts: np.ndarray, kinds, min1s, max1s, min2s, threads, workdir: str) -> dict:
    """Times the reference's own StaticPSTCPUIter (oracle/_ref/ref_bench,
    f32 points with u32 IDs): construction of `pts`, then the query list
    once per entry of `threads` (IDs returned). Returns the driver's JSON:
    {"n", "nq", "build_s", "runs": [{"threads", "secs", "ids"}]}."""
    import json
    n = len(pts)
    path = os.path.join(workdir, "ref_bench_in.bin")
    nq = len(kinds)
    rec = np.zeros(nq, np.dtype([("kind", np.int32), ("rp", np.int32), ("a", np.float32), ("b", np.float32),
                                 ("c", np.float32)]))
    rec["kind"], rec["a"], rec["b"], rec["c"] = kinds, min1s, max1s, min2s
    with open(path, "wb") as f:
        f.write(b"PSTR" + struct.pack("<iiQQ", DT["f32"], DT["u32"], n, nq))
        f.write(np.ascontiguousarray(pts).tobytes())
        f.write(rec.tobytes())
    try:
        out = subprocess.run([REF_BENCH, path, *[str(t) for t in threads]], check=True, capture_output=True,
                             text=True).stdout
    finally:
        os.remove(path)
    return json.loads(out.strip().splitlines()[-1])


def _parse_tree(tree_lines, arrs, T, IT):
    """Rebuilds slot positions from StaticPSTCPUIter::print()'s drawing
    (pst/static-pst-cpu-iter.tpp:821-853): root first, then the right
    subtree, then the left; each child line carries "(R)" or "(L)" and is
    indented 7 columns per level."""
    slot_at_depth = {}
    for ln in tree_lines:
        p = ln.find("(", 0)
        marker = None
        if "─(R)─ " in ln:
            marker = "R"
            p = ln.index("─(R)─ ") + len("─(R)─ ")
        elif "─(L)─ " in ln:
            marker = "L"
            p = ln.index("─(L)─ ") + len("─(L)─ ")
        if marker is None:
            depth, slot = 0, 0
        else:
            prefix = ln[:ln.index("─(" + marker + ")─ ")]
            depth = (len(prefix) - 1) // 7 + 1
            parent = slot_at_depth[depth - 1]
            slot = 2 * parent + (1 if marker == "L" else 2)
            arrs["bits"][parent] |= 2 if marker == "L" else 1
        slot_at_depth[depth] = slot
        body = ln[p:].strip()
        assert body[0] == "(" and body[-1] == ")", ln
        parts = body[1:-1].replace(";", ",").split(",")
        parts = [x.strip() for x in parts]
        arrs["dim1"][slot] = _parse_val(parts[0], T)
        arrs["dim2"][slot] = _parse_val(parts[1], T)
        arrs["med"][slot] = _parse_val(parts[2], T)
        if IT is not None:
            arrs["ids"][slot] = _parse_val(parts[3], IT)
