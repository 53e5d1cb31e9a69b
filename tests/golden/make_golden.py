"""Writes the golden fixtures in tests/golden/*.npz.

Every fixture holds the INPUT (AoS points, queries) and the EXPECTED OUTPUT
produced by the reference's own StaticPSTCPUIter (the reference's C++ CPU
tree, compiled from its sources by `make -C oracle ref` into
oracle/_ref/pst_ref): the five tree sub-arrays (dim1, dim2, med, ids,
bitcodes, as printed by StaticPSTCPUIter::print at round-trip precision) and
each query's result list in the reference's own reporting order.

The reference ships no test vectors of its own (no tests, fixtures or golden
files exist in it), so these fixtures are the anchor that pins the oracle
restatement and, through it, the GPU path.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def pts_of(t, i, d1, d2, ids=None):
    p = np.zeros(len(d1), O.point_dtype(t, i))
    p["dim1"] = d1
    p["dim2"] = d2
    if i is not None:
        p["id"] = ids if ids is not None else np.arange(len(d1))
    return p


def interval_f32(rng, n):
    """Metacell-like intervals: dim1 ~ U[0,1), dim2 = min(dim1 + Exp(0.02), 1)."""
    d1 = rng.random(n).astype(np.float32)
    d2 = np.minimum(d1.astype(np.float64) + rng.exponential(0.02, n), 1.0).astype(np.float32)
    return d1, d2


def queries_for(rng, p, t, has_id, nq, lo, hi):
    qs = []
    T = O.NP[t]
    for _ in range(nq):
        if np.dtype(T).kind == "f":
            a, b = np.sort(rng.uniform(lo, hi, 2))
            c = rng.uniform(lo, hi)
            q = rng.uniform(lo, hi)
        else:
            a, b = np.sort(rng.integers(lo, hi, 2))
            c = rng.integers(lo, hi)
            q = rng.integers(lo, hi)
        a, b, c, q = T(a), T(b), T(c), T(q)
        for rp in ([0, 1] if has_id else [1]):
            qs.append((O.THREE, rp, a, b, c))
            qs.append((O.LEFT, rp, a, b, c))
            qs.append((O.RIGHT, rp, a, b, c))
            qs.append((O.LEFT, rp, a, q, q))  # stabbing query (q, q)
    # inverted 3-sided bounds (empty by predicate) and report-all / none
    if len(p):
        qs.append((O.THREE, 1 if not has_id else 0, T(hi), T(lo), T(lo)))
        d1, d2 = p["dim1"], p["dim2"]
        qs.append((O.LEFT, 1 if not has_id else 0, T(0), d1.max(), d2.min()))
        qs.append((O.RIGHT, 1 if not has_id else 0, d1.min(), T(0), d2.max()))
    return qs


def cases():
    rng = np.random.default_rng(20261202)
    C = []
    # shape edge cases
    for n in (0, 1, 2, 3, 4, 7, 8, 15, 16, 17):
        d1, d2 = interval_f32(rng, n)
        C.append((f"f32_u32_n{n}", "f32", "u32", pts_of("f32", "u32", d1, d2)))
    # metacell-like float intervals
    for n in (100, 1000, 4099):
        d1, d2 = interval_f32(rng, n)
        C.append((f"f32_u32_interval_n{n}", "f32", "u32", pts_of("f32", "u32", d1, d2)))
    # heavy duplicates: few distinct values, duplicate IDs (pure tie-break paths)
    n = 500
    d1 = rng.integers(0, 4, n).astype(np.float32)
    d2 = d1 + rng.integers(0, 3, n).astype(np.float32)
    C.append(("f32_u32_dups", "f32", "u32", pts_of("f32", "u32", d1, d2, rng.integers(0, 5, n))))
    # all points identical: order decided by id, then by index
    n = 257
    C.append(("f32_u32_allequal", "f32", "u32",
              pts_of("f32", "u32", np.full(n, 0.5, np.float32), np.full(n, 0.5, np.float32),
                     rng.integers(0, 3, n))))
    C.append(("f32_none_allequal", "f32", None,
              pts_of("f32", None, np.full(n, 0.5, np.float32), np.full(n, 0.5, np.float32))))
    # signed zeros (IEEE: -0.0 == +0.0, the tie-break decides)
    n = 64
    z = np.where(rng.random(n) < 0.5, np.float32(-0.0), np.float32(0.0))
    C.append(("f32_u32_signed_zero", "f32", "u32",
              pts_of("f32", "u32", z, np.where(rng.random(n) < 0.5, np.float32(-0.0), np.float32(1.0)))))
    # integer workload shape: dim2 = dim1 + geometric width
    n = 3000
    d1 = rng.integers(0, 1 << 30, n).astype(np.int32)
    d2 = (d1.astype(np.int64) + rng.geometric(2.0 ** -12, n) - 1).clip(max=2**31 - 1).astype(np.int32)
    C.append(("i32_u32_geom", "i32", "u32", pts_of("i32", "u32", d1, d2)))
    # negative ints, duplicates
    n = 700
    d1 = rng.integers(-50, 50, n).astype(np.int32)
    C.append(("i32_i32_neg", "i32", "i32", pts_of("i32", "i32", d1, d1 + rng.integers(0, 20, n),
                                                  rng.integers(-3, 3, n))))
    # wider types and ID kinds
    n = 1200
    d1 = rng.random(n)
    C.append(("f64_none", "f64", None, pts_of("f64", None, d1, d1 + rng.random(n) * 0.1)))
    d1 = rng.integers(-(1 << 40), 1 << 40, n)
    C.append(("i64_i64", "i64", "i64", pts_of("i64", "i64", d1, d1 + rng.integers(0, 1 << 35, n),
                                              rng.integers(-(1 << 40), 1 << 40, n))))
    d1, d2 = interval_f32(rng, n)
    C.append(("f32_u64", "f32", "u64", pts_of("f32", "u64", d1, d2,
                                               rng.integers(0, 1 << 62, n, dtype=np.uint64))))
    d1 = rng.random(n)
    C.append(("f64_f32ids", "f64", "f32", pts_of("f64", "f32", d1, d1 + rng.random(n),
                                                  rng.integers(0, 50, n).astype(np.float32))))
    C.append(("i32_u64", "i32", "u64", pts_of("i32", "u64", rng.integers(0, 1000, n),
                                               rng.integers(1000, 2000, n))))
    C.append(("f32_i64", "f32", "i64", pts_of("f32", "i64", *interval_f32(rng, n))))
    C.append(("i64_none", "i64", None, pts_of("i64", None, rng.integers(0, 100, n),
                                               rng.integers(50, 200, n))))
    C.append(("f64_u32", "f64", "u32", pts_of("f64", "u32", rng.random(n), rng.random(n) + 0.5)))
    return C


def chunk_cases():
    """Forest fixtures: one reference CPU tree per consecutive chunk of
    127 points (the per-tree size for 64 threads per block)."""
    rng = np.random.default_rng(7)
    n = 127 * 5 + 40
    d1, d2 = interval_f32(rng, n)
    return [("forest_f32_u32_tpb64", "f32", "u32", pts_of("f32", "u32", d1, d2), 127)]


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/pst_ref is missing: run `make -C oracle ref` first")
    wd = tempfile.mkdtemp()
    rng = np.random.default_rng(99)
    for name, t, i, p in cases():
        lo, hi = (float(p["dim1"].min()) - 0.1, float(p["dim2"].max()) + 0.1) if len(p) else (0, 1)
        if np.dtype(O.NP[t]).kind != "f":
            lo, hi = (int(p["dim1"].min()) - 2, int(p["dim2"].max()) + 2) if len(p) else (0, 10)
        qs = queries_for(rng, p, t, i is not None, 6, lo, hi)
        arrs, res = O.run_reference(p, t, i, qs, wd)
        check_against_brute_force(p, t, i, qs, res, name)
        save(name, t, i, p, qs, arrs, res)
    for name, t, i, p, chunk in chunk_cases():
        trees = []
        all_res = []
        qs = queries_for(rng, p, t, True, 4, 0.0, 1.0)
        for c0 in range(0, len(p), chunk):
            sub = p[c0:c0 + chunk]
            arrs, res = O.run_reference(sub, t, i, qs, wd)
            trees.append(arrs)
            all_res.append(res)
        # union over trees, per query (order: tree by tree, each in ref order)
        merged = [sum((r[q] for r in all_res), []) for q in range(len(qs))]
        save(name, t, i, p, qs, None, merged, forest=(chunk, trees))
    print("wrote fixtures to", OUT)


def check_against_brute_force(p, t, i, qs, res, name):
    """A vector is only kept if the reference's answer equals the predicate
    (guards against undefined behaviour in the reference leaking into a
    fixture, e.g. its under-sized blob when sizeof(IDType) > sizeof(T))."""
    for (kind, rp, a, b, c), rows in zip(qs, res):
        exp = p[O.brute_force(p, kind, a, b, c)]
        if len(rows) != len(exp):
            sys.exit(f"{name}: reference result disagrees with brute force; not saved")


def save(name, t, i, p, qs, arrs, res, forest=None):
    T = O.NP[t]
    d = {"points": p, "t": np.array(t), "i": np.array(i if i else ""),
         "q_kind": np.array([q[0] for q in qs], np.int32),
         "q_ret_point": np.array([q[1] for q in qs], np.int32),
         "q_min1": np.array([q[2] for q in qs], T),
         "q_max1": np.array([q[3] for q in qs], T),
         "q_min2": np.array([q[4] for q in qs], T)}
    if arrs is not None:
        for k, v in arrs.items():
            if v is not None:
                d["tree_" + k] = v
    if forest is not None:
        chunk, trees = forest
        d["forest_chunk"] = np.array(chunk)
        for k in ("dim1", "dim2", "med", "bits", "ids"):
            d["forest_" + k] = np.concatenate([tr[k] for tr in trees])
        d["forest_sizes"] = np.array([len(tr["dim1"]) for tr in trees], np.int64)
    # results: flatten as text rows -> typed arrays
    counts = np.array([len(r) for r in res], np.int64)
    d["res_counts"] = counts
    flat = [row for r in res for row in r]
    d["res_rows"] = np.array(flat if flat else [""], dtype=object).astype(str)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **d)


if __name__ == "__main__":
    main()
