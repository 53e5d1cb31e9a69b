"""Marching cubes case table, restated in Python from the rule the library
states (include/pst/pst_capi.h, pst_marching_cubes_*): corner k of a voxel
is at (k & 1, k >> 1 & 1, k >> 2) and is below when its value is < q; on
each cube face, walked counter-clockwise as seen from outside, every maximal
run of below corners is cut off by a segment from the edge where the walk
leaves the run to the edge where it entered it; the segments close into
loops, and each loop (e0, e1, ..., ek) becomes the triangles (e0, e(i+1), ei).
Edges: x edges 0-3 (corner pairs 0-1, 2-3, 4-5, 6-7), y edges 4-7 (0-2, 1-3,
4-6, 5-7), z edges 8-11 (0-4, 1-5, 2-6, 3-7). Test infrastructure."""
import numpy as np

EDGES = [(0, 1), (2, 3), (4, 5), (6, 7), (0, 2), (1, 3), (4, 6), (5, 7), (0, 4), (1, 5), (2, 6), (3, 7)]
AXES = {1: np.array([1, 0, 0]), 2: np.array([0, 1, 0]), 4: np.array([0, 0, 1])}


def corner_pos(k):
    return np.array([k & 1, (k >> 1) & 1, k >> 2], np.float64)


def edge_index(a, b):
    return EDGES.index((min(a, b), max(a, b)))


def faces():
    """(4 corners counter-clockwise as seen from outside) for the 6 faces,
    derived from the outward normal rather than listed."""
    out = []
    for ax in (1, 2, 4):
        others = [b for b in (1, 2, 4) if b != ax]
        for side in (0, 1):
            normal = AXES[ax] * (1 if side else -1)
            u, v = others
            if np.dot(np.cross(AXES[u], AXES[v]), normal) < 0:
                u, v = v, u
            base = ax if side else 0
            out.append([base, base | u, base | u | v, base | v])
    return out


FACES = faces()


def case(cfg):
    """Edge triples of configuration cfg (bit k: corner k below)."""
    nxt = {}
    for c in FACES:
        inside = [(cfg >> x) & 1 for x in c]
        if all(inside):
            continue
        for i in range(4):
            if inside[i] and not inside[(i + 1) % 4]:
                j = i
                while inside[(j - 1) % 4]:
                    j = (j - 1) % 4
                nxt[edge_index(c[i], c[(i + 1) % 4])] = edge_index(c[(j - 1) % 4], c[j])
    tris, seen = [], set()
    for e0 in sorted(nxt):
        if e0 in seen:
            continue
        cyc, e = [], e0
        while e not in seen:
            seen.add(e)
            cyc.append(e)
            e = nxt[e]
        for k in range(1, len(cyc) - 1):
            tris.append((cyc[0], cyc[k + 1], cyc[k]))
    return tris


TABLE = [case(c) for c in range(256)]


def voxel_triangles(f, q, x, y, z, acc):
    """Triangles (float32) of the voxel with corner values f[0..7] at base
    (x, y, z), vertices interpolated from the edge endpoint of lower index."""
    cfg = sum(1 << k for k in range(8) if f[k] < q)
    out = []
    for tri in TABLE[cfg]:
        pts = []
        for e in tri:
            u, w = EDGES[e]
            t = (acc(q) - acc(f[u])) / (acc(f[w]) - acc(f[u]))
            pu = np.array([x + (u & 1), y + ((u >> 1) & 1), z + (u >> 2)], np.float32)
            pw = np.array([x + (w & 1), y + ((w >> 1) & 1), z + (w >> 2)], np.float32)
            pts.append(pu + np.float32(t) * (pw - pu))
        out.append(np.array(pts, np.float32))
    return out
