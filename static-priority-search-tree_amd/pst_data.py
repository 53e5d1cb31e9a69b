"""Seeded synthetic workloads (points and queries) for tests and bench.

One project-owned counter-based generator: SplitMix64 over (seed, stream,
index), so any slice of any stream is reproducible on any machine and
independent of numpy's RNG implementation. Streams used:
  1 dim1, 2 width, 3 query value, 4 query width, 5 query c, 6 query kind,
  7 query side, 8 tree choice, 9 arrival gaps, 10.. volume field.
Float conversions: u in [0,1) from the top 24 bits (f32) or 53 bits (f64).
"""
from __future__ import annotations

import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


# Vertex grid dimensions (x, y, z) of the reference's isosurface datasets
# (testers/pst-dataset-tester-driver.cu:48-52); the raw files hold float32
# values, x fastest, and are read with readInVertices / read_vertices_to_device.
DATASET_DIMS = {
    "Isabel": (500, 500, 100),
    "Radiation": (600, 248, 248),
    "TeraShake": (750, 375, 100),
    "Vortex": (128, 128, 128),
}


def splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def bits(seed: int, stream: int, start: int, count: int) -> np.ndarray:
    """count 64-bit words of (seed, stream) starting at `start`."""
    with np.errstate(over="ignore"):
        base = splitmix64(np.array([seed * 0x100000001B3 + stream], dtype=np.uint64))[0]
        idx = np.arange(start, start + count, dtype=np.uint64)
        return splitmix64(idx * np.uint64(0xD1B54A32D192ED03) + base)


def u01_f32(seed, stream, start, count) -> np.ndarray:
    b = bits(seed, stream, start, count) >> np.uint64(40)
    return (b.astype(np.float64) * 2.0 ** -24).astype(np.float32)


def u01_f64(seed, stream, start, count) -> np.ndarray:
    b = bits(seed, stream, start, count) >> np.uint64(11)
    return b.astype(np.float64) * 2.0 ** -53


def interval_points(n: int, seed: int = 0, mean_width: float = 0.02, chunk: int = 1 << 22):
    """Metacell-like float32 intervals with uint32 IDs = index:
    dim1 ~ U[0,1) (f32), dim2 = min(dim1 + Exp(mean_width), 1.0) computed in
    double and rounded to f32, id = index."""
    dt = np.dtype([("dim1", np.float32), ("dim2", np.float32), ("id", np.uint32)], align=True)
    out = np.empty(n, dtype=dt)
    for s in range(0, n, chunk):
        c = min(chunk, n - s)
        d1 = u01_f32(seed, 1, s, c)
        u = u01_f64(seed, 2, s, c)
        w = -mean_width * np.log1p(-u)
        out["dim1"][s:s + c] = d1
        out["dim2"][s:s + c] = np.minimum(d1.astype(np.float64) + w, 1.0).astype(np.float32)
        out["id"][s:s + c] = np.arange(s, s + c, dtype=np.uint32)
    return out


def int_points(n: int, seed: int = 3, p: float = 2.0 ** -12, chunk: int = 1 << 22):
    """int32 dims: dim1 ~ U{0..2^30-1}, dim2 = dim1 + G with G geometric
    on {0, 1, 2, ...} (failures before the first success, mean (1-p)/p),
    uint32 IDs = index."""
    dt = np.dtype([("dim1", np.int32), ("dim2", np.int32), ("id", np.uint32)], align=True)
    out = np.empty(n, dtype=dt)
    for s in range(0, n, chunk):
        c = min(chunk, n - s)
        d1 = (bits(seed, 1, s, c) >> np.uint64(34)).astype(np.int64)
        u = u01_f64(seed, 2, s, c)
        g = np.floor(np.log1p(-u) / np.log1p(-p)).astype(np.int64)
        out["dim1"][s:s + c] = d1
        out["dim2"][s:s + c] = np.minimum(d1 + g, 2 ** 31 - 1)
        out["id"][s:s + c] = np.arange(s, s + c, dtype=np.uint32)
    return out


def stabbing_queries(nq: int, seed: int = 0, dtype=np.float32):
    """q ~ U[0,1): the reference's twoSidedLeftSearch(q, q), i.e.
    dim1 <= q && dim2 >= q."""
    q = u01_f32(seed, 3, 0, nq).astype(dtype)
    return q


def float_query_mix(nq3: int, nq2: int, seed: int = 0):
    """Query records (pst_query_f32 layout) for the float 2^26 workload:
    nq2 two-sided stabbing queries (q, q), half left and half right, and
    nq3 three-sided queries a ~ U[0,1), b = a + w with
    w = exp(U[ln 1e-4, ln 0.5]), c ~ U[0,1)."""
    qd = np.dtype([("min_dim1", np.float32), ("max_dim1", np.float32), ("min_dim2", np.float32),
                   ("kind", np.uint32)], align=True)
    out = np.zeros(nq2 + nq3, dtype=qd)
    q = u01_f32(seed, 3, 0, nq2)
    side = (bits(seed, 7, 0, nq2) & np.uint64(1)).astype(np.uint32)
    out["min_dim1"][:nq2] = q
    out["max_dim1"][:nq2] = q
    out["min_dim2"][:nq2] = q
    out["kind"][:nq2] = np.where(side == 0, 1, 2)  # 1 left, 2 right
    a = u01_f32(seed, 4, 0, nq3).astype(np.float64)
    w = np.exp(np.log(1e-4) + u01_f64(seed, 5, 0, nq3) * (np.log(0.5) - np.log(1e-4)))
    c = u01_f32(seed, 6, 0, nq3)
    out["min_dim1"][nq2:] = a.astype(np.float32)
    out["max_dim1"][nq2:] = (a + w).astype(np.float32)
    out["min_dim2"][nq2:] = c
    out["kind"][nq2:] = 0
    return out


def int_query_mix(nq: int, seed: int = 3, dense_frac: float = 0.5):
    """pst_query_i32 records for the int32 workload: 50% two-sided, 50%
    three-sided; widths w = exp(U[0, ln 2^23]) so output counts span 0 to
    ~1e6 (about w/8 points per query at 2^27 points). Three-sided:
    [a, a+w], c <= a for a dense half (everything in [a, a+w] is
    reported), c ~ U{0..2^30} for the rest. Two-sided, by a fair coin:
    left  (max_dim1 = x, min_dim2 = x - w), x ~ U{0..2^30};
    right (min_dim1 = min_dim2 = 2^30 - 1 - w), the top w of the range."""
    qd = np.dtype([("min_dim1", np.int32), ("max_dim1", np.int32), ("min_dim2", np.int32),
                   ("kind", np.uint32)], align=True)
    out = np.zeros(nq, dtype=qd)
    n2 = nq // 2
    x = (bits(seed, 3, 0, nq) >> np.uint64(34)).astype(np.int64)
    w = np.floor(np.exp(u01_f64(seed, 4, 0, nq) * np.log(2.0 ** 23))).astype(np.int64)
    side = (bits(seed, 7, 0, nq) & np.uint64(1)).astype(np.int64)
    left = side[:n2] == 0
    top = (2 ** 30 - 1) - w[:n2]
    out["min_dim1"][:n2] = np.where(left, x[:n2], top)
    out["max_dim1"][:n2] = np.where(left, x[:n2], 2 ** 31 - 1)
    out["min_dim2"][:n2] = np.where(left, np.maximum(x[:n2] - w[:n2], 0), top)
    out["kind"][:n2] = np.where(left, 1, 2)
    a = x[n2:]
    b = np.minimum(a + w[n2:], 2 ** 31 - 1)
    dense = u01_f64(seed, 5, 0, nq - n2) < dense_frac
    c_sparse = (bits(seed, 6, 0, nq - n2) >> np.uint64(34)).astype(np.int64)
    out["min_dim1"][n2:] = a
    out["max_dim1"][n2:] = b
    out["min_dim2"][n2:] = np.where(dense, np.maximum(a - 1, 0), c_sparse)
    out["kind"][n2:] = 0
    return out
