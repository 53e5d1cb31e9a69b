"""Whole-tree checks of a PST in the reference's layout, vectorised with
numpy so that they run on trees of 2^26+ points.

Together they fix the tree uniquely (the construction rule of
pst/static-pst-cpu-iter.tpp:512-545, 579-597): the slot shape is the
closed form of the point count, every input point sits in exactly one
slot, each node holds the maximum key2 = (dim2, id) of its subtree (heap
order on parent/child pairs), the left subtree's key1 = (dim1, id) values
all lie below the right subtree's, and med is dim1 of the left subtree's
largest key1 (the node's own dim1 when it has no left child). Inputs must
have IDs equal to their indices (the generators' workloads), so (value, id)
orders are the reference's (value, id, index) orders.
"""
import numpy as np

HAS_LEFT, HAS_RIGHT = 0x2, 0x1


_REV8 = np.array([int(f"{b:08b}"[::-1], 2) for b in range(256)], dtype=np.uint32)


def _brev32(x):
    """Bit reversal of 32-bit unsigned values (byte table)."""
    x = x.astype(np.uint32)
    return ((_REV8[x & 255] << 24) | (_REV8[(x >> 8) & 255] << 16) | (_REV8[(x >> 16) & 255] << 8)
            | _REV8[x >> 24])


def _bits(a):
    """The raw bit patterns of an array (floats compared bit for bit)."""
    return a.view(f"u{a.itemsize}") if a.dtype.kind == "f" else a


def shape_sizes(n: int) -> np.ndarray:
    """Points held by every slot: with M = n + 1, the slot at depth t and
    position j holds floor(M / 2^t) - 1 + [brev_t(j) < M mod 2^t]
    (0 where nothing is held)."""
    H = n.bit_length()
    S = (1 << H) - 1 if n else 0
    m = np.zeros(S, np.int64)
    M = n + 1
    for t in range(H):
        j = np.arange(1 << t, dtype=np.uint32)
        rev = (_brev32(j) >> np.uint32(32 - t)) if t else np.zeros(1, np.uint32)
        r = M & ((1 << t) - 1)
        m[(1 << t) - 1:(1 << (t + 1)) - 1] = (M >> t) - 1 + (rev < r)
    return np.maximum(m, 0)


def _ord(a):
    """Order-preserving unsigned 64-bit keys of 32-bit values (IEEE order
    for floats with -0.0 == +0.0, signed order for ints)."""
    u = a.view(np.uint32)
    if a.dtype.kind == "f":
        u = u * ((u & np.uint32(0x7FFFFFFF)) != 0)                 # -0.0 -> +0.0
        neg = (u >> np.uint32(31)).astype(bool)
        k = np.where(neg, ~u, u | np.uint32(0x80000000))
    else:
        k = u ^ np.uint32(0x80000000)
    return k.astype(np.uint64)


def check(ex: dict, pts: np.ndarray) -> None:
    """Asserts every invariant on the exported arrays `ex` (dim1, dim2,
    med, ids, bits) of the tree built from `pts` (32-bit dims, IDs =
    indices). Per level, children of level-t position j are level t+1
    positions 2j and 2j+1, so every check is a strided slice."""
    n = len(pts)
    m = shape_sizes(n)
    S = len(m)
    assert len(ex["dim1"]) == S
    valid = m > 0
    assert int(valid.sum()) == n
    d1, d2, md, ids, bits = ex["dim1"], ex["dim2"], ex["med"], ex["ids"], ex["bits"]
    # unused slots are zero
    inv = ~valid
    for arr in (d1, d2, md, ids, bits):
        assert not np.any(_bits(arr[inv])), "an unused slot is not zero"
    # every point exactly once, with its own dims
    vid = ids[valid].astype(np.int64)
    assert np.array_equal(np.sort(vid), np.arange(n)), "IDs are not a permutation of the input"
    assert np.array_equal(_bits(d1[valid]), _bits(pts["dim1"][vid])), "a slot's dim1 is not its point's"
    assert np.array_equal(_bits(d2[valid]), _bits(pts["dim2"][vid])), "a slot's dim2 is not its point's"
    del vid
    idk = ids.astype(np.uint64)
    key1 = (_ord(d1) << np.uint64(32)) | idk   # (dim1, id)
    key2 = (_ord(d2) << np.uint64(32)) | idk   # (dim2, id)
    medk = _ord(md)
    H = n.bit_length()
    lo = [(1 << t) - 1 for t in range(H + 1)]
    kmax = np.zeros(S, np.uint64)   # largest key1 of each subtree
    kmin = np.full(S, np.uint64(0xFFFFFFFFFFFFFFFF))  # smallest
    for t in range(H - 1, -1, -1):
        a, b = lo[t], lo[t + 1]
        v = valid[a:b]
        mx = np.where(v, key1[a:b], np.uint64(0))
        mn = np.where(v, key1[a:b], np.uint64(0xFFFFFFFFFFFFFFFF))
        if t + 1 < H:
            c, e = lo[t + 1], lo[t + 2]
            vl, vr = valid[c:e][0::2], valid[c:e][1::2]
            # shape and bitcodes
            want_bits = np.where(v, vl * HAS_LEFT | vr * HAS_RIGHT, 0).astype(np.uint8)
            assert np.array_equal(bits[a:b], want_bits), f"bitcodes differ from the shape at depth {t}"
            assert np.all(~v | (m[c:e][0::2] + m[c:e][1::2] == m[a:b] - 1)), "shape differs"
            # heap order on key2
            k2 = key2[a:b]
            assert np.all(~vl | (k2 > key2[c:e][0::2])), f"heap order broken at depth {t}"
            assert np.all(~vr | (k2 > key2[c:e][1::2])), f"heap order broken at depth {t}"
            lmax, rmin = kmax[c:e][0::2], kmin[c:e][1::2]
            # BST split on key1, and med = dim1 of the left subtree's largest key1
            assert np.all(~(vl & vr) | (rmin > lmax)), f"BST split broken at depth {t}"
            want_med = np.where(vl, lmax >> np.uint64(32), key1[a:b] >> np.uint64(32))
            assert np.all(~v | (medk[a:b] == want_med)), f"med differs at depth {t}"
            mx = np.maximum(mx, np.where(vl, lmax, 0))
            mx = np.maximum(mx, np.where(vr, kmax[c:e][1::2], 0))
            mn = np.minimum(mn, np.where(vl, kmin[c:e][0::2], np.uint64(0xFFFFFFFFFFFFFFFF)))
            mn = np.minimum(mn, np.where(vr, rmin, np.uint64(0xFFFFFFFFFFFFFFFF)))
        else:
            assert np.all(bits[a:b] == 0), "a last-level node has a child bit"
            assert np.all(~v | (medk[a:b] == (key1[a:b] >> np.uint64(32)))), "leaf med differs"
        kmax[a:b] = mx
        kmin[a:b] = mn
