"""Loads the golden fixtures written by tests/golden/make_golden.py."""
from __future__ import annotations

import glob
import os

import numpy as np

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names(prefix=""):
    fs = sorted(glob.glob(os.path.join(GOLDEN, prefix + "*.npz")))
    return [os.path.splitext(os.path.basename(f))[0] for f in fs]


class Fixture:
    def __init__(self, name):
        d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
        self.name = name
        self.t = str(d["t"])
        self.i = str(d["i"]) or None
        self.points = d["points"]
        self.n = len(self.points)
        self.T = O.NP[self.t]
        self.I = O.NP[self.i] if self.i else None
        self.q_kind = d["q_kind"]
        self.q_ret_point = d["q_ret_point"]
        self.q_min1, self.q_max1, self.q_min2 = d["q_min1"], d["q_max1"], d["q_min2"]
        self.tree = {k[5:]: d[k] for k in d.files if k.startswith("tree_")}
        self.forest = {k[7:]: d[k] for k in d.files if k.startswith("forest_")}
        counts = d["res_counts"]
        rows = d["res_rows"]
        self.results = []
        k = 0
        for c in counts:
            self.results.append(list(rows[k:k + c]))
            k += c

    def queries(self):
        for q in range(len(self.q_kind)):
            yield (q, int(self.q_kind[q]), bool(self.q_ret_point[q]), self.q_min1[q],
                   self.q_max1[q], self.q_min2[q])

    def expected_sorted(self, q):
        """Reference results of query q as a canonical sorted list of bit
        patterns: tuples (dim1, dim2[, id]) for points, ints for IDs."""
        rows = self.results[q]
        if self.q_ret_point[q]:
            out = []
            for r in rows:
                v = r.split()
                tup = (bits(O._parse_val(v[0], self.T)), bits(O._parse_val(v[1], self.T)))
                if self.i:
                    tup += (bits(O._parse_val(v[2], self.I)),)
                out.append(tup)
            return sorted(out)
        return sorted(bits(O._parse_val(r, self.I)) for r in rows)


_UINT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def bits(v):
    a = np.asarray(v)
    return int(a.reshape(1).view(_UINT[a.dtype.itemsize])[0])


def bits_arr(a):
    a = np.ascontiguousarray(a)
    return a.view(_UINT[a.dtype.itemsize]).astype(np.uint64)


def canon(res, ret_point, has_id):
    """Canonical sorted form (bit patterns) of a result array."""
    if ret_point:
        cols = [bits_arr(res["dim1"]).tolist(), bits_arr(res["dim2"]).tolist()]
        if has_id:
            cols.append(bits_arr(res["id"]).tolist())
        return sorted(zip(*cols))
    return sorted(bits_arr(res).tolist())
