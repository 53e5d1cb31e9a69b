"""CPU restatement of the level index scheme (csrc/pst_levels.hip), checked
against brute force on the oracle's trees. It pins the design the kernels
implement:

* the closed forms: the level-0 input order of occupied slots, the run
  starts and sizes, and the same-way counts before a run (`side_before`);
* construction: a stable sort by ~ord_key(dim2), then per level the stable
  partition of each run into its children's runs, each point placed at its
  child run's start plus its rank among same-way points;
* the invariants the searches rely on: every run is sorted, and a
  subtree's root is its run's first element;
* the search: the open paths (at most two per level) with the walk's own
  node step, one range per reported path node and one binary-searched
  prefix per REPORT_ABOVE subtree (at most 124 ranges), whose IDs equal
  brute force for every query kind, ties, signed zeros, infinities and NaN
  bounds.
"""
import numpy as np
import pytest

import oracle as O

THREE, LEFT, RIGHT = 0, 1, 2


def brev(x, t):
    r = 0
    for b in range(t):
        if (x >> b) & 1:
            r |= 1 << (t - 1 - b)
    return r


def level_size(M, t, j):
    return (M >> t) - 1 + (1 if brev(j, t) < (M & ((1 << t) - 1)) else 0)


def level_start(M, t, j):
    return sum(level_size(M, t, x) for x in range(j))


def brev_below(t, j, r):
    """#{j' < j : brev_t(j') < r}, grouped by the highest bit where j' and j
    differ (the kernels' closed form)."""
    cnt = 0
    for b in range(t):
        if (j >> b) & 1:
            c = brev((j >> b) ^ 1, t - b)
            if r > c:
                step = 1 << (t - b)
                cnt += min(-(-(r - c) // step), 1 << b)
    return cnt


def side_before(M, t, j, side):
    base = (M >> (t + 1)) - 1
    r1 = M & ((2 << t) - 1)
    off = side << t
    return j * base + brev_below(t, j, r1 - off if r1 > off else 0)


def depth(s):
    return (s + 1).bit_length() - 1


def input_slot(e, H):
    F = (1 << (H - 1)) - 1
    return e if e < F else F + brev(e - F, H - 1)


def ord_key(v, T):
    if T in (np.float32, np.float64):
        bits = 32 if T == np.float32 else 64
        u = int(np.array([v], T).view(np.uint32 if bits == 32 else np.uint64)[0])
        sign = 1 << (bits - 1)
        if u & (sign - 1) == 0:
            u = 0
        return ((~u) & ((1 << bits) - 1)) if u & sign else (u | sign)
    bits = 32 if T == np.int32 else 64
    return (int(v) + (1 << (bits - 1))) & ((1 << bits) - 1)


def node_step(d1, d2, md, m, lo, hi, min2, l, h):
    """The walk's per-query transition (pst_common.hpp: node_step, G = 1)."""
    act = min2 <= d2
    d1lo, d1hi, mdhi, mdlo = lo <= d1, d1 <= hi, md <= hi, md >= lo
    himd = (not mdhi) and hi == hi
    mdltlo = (not mdlo) and lo == lo
    ra, lf = act and not l and not h, act and h and not l
    rt, th = act and l and not h, act and l and h
    split = th and mdlo and mdhi
    tl = th and not split and himd
    tr = th and not split and not himd and mdltlo
    rep = act and (not l or d1lo) and (not h or d1hi)
    left = (ra or lf or (rt and mdlo) or split or tl, (rt and mdlo) or split or tl, (lf and not mdhi) or tl)
    right = (ra or (lf and mdhi) or rt or split or tr, (rt and not mdlo) or tr, (lf and mdhi) or split or tr)
    if m < 2:
        left = (False, False, False)
    if m < 3:
        right = (False, False, False)
    return rep, left, right


def build_index(tr, n, T):
    M, H = n + 1, n.bit_length()
    mask = (1 << (8 * np.dtype(T).itemsize)) - 1
    inp = [input_slot(e, H) for e in range(n)]
    keys = [(~ord_key(tr.dim2[s], T)) & mask for s in inp]
    order = sorted(range(n), key=lambda e: keys[e])  # stable
    lv_key, lv_id, slots = [[keys[e] for e in order]], [[int(tr.ids[inp[e]]) for e in order]], [inp[e] for e in order]
    for t in range(H - 1):
        nk = [None] * (len(slots) - (1 << t))
        ni, ns = list(nk), list(nk)
        seen = [0, 0]  # left- and right-going points so far (the global prefix count)
        for p, s in enumerate(slots):
            d = depth(s)
            if d == t:
                continue
            side = ((s + 1) >> (d - t - 1)) & 1
            j = ((s + 1) >> (d - t)) - (1 << t)
            start = side_before(M, t, j, 0) + side_before(M, t, j, 1)
            if side:
                start += level_size(M, t + 1, 2 * j)
            dst = start + seen[side] - side_before(M, t, j, side)
            assert ns[dst] is None
            nk[dst], ni[dst], ns[dst] = lv_key[t][p], lv_id[t][p], s
            seen[side] += 1
        assert None not in ns
        lv_key.append(nk)
        lv_id.append(ni)
        slots = ns
    return lv_key, lv_id


def search(tr, n, T, lv_key, lv_id, q):
    M, H = n + 1, n.bit_length()
    mask = (1 << (8 * np.dtype(T).itemsize)) - 1
    lo, hi, min2, kind = T(q["min_dim1"]), T(q["max_dim1"]), T(q["min_dim2"]), int(q["kind"])
    paths = [(0, kind != LEFT, kind in (THREE, LEFT))]
    ranges = []
    for t in range(H):
        if not paths:
            break
        nxt = []
        for slot, l, h in paths:
            j = slot + 1 - (1 << t)
            rep, L, R = node_step(tr.dim1[slot], tr.dim2[slot], tr.med[slot], level_size(M, t, j), lo, hi, min2, l, h)
            if rep:
                ranges.append((t, level_start(M, t, j), 1, False))
            for side, (a, cl, ch) in ((0, L), (1, R)):
                if a and not cl and not ch:
                    c = 2 * j + side
                    ranges.append((t + 1, level_start(M, t + 1, c), level_size(M, t + 1, c), True))
                elif a:
                    nxt.append((2 * slot + 1 + side, cl, ch))
        assert len(nxt) <= 2
        paths = nxt
    assert len(ranges) <= 124
    out = []
    for t, st, sz, srch in ranges:
        c = sz
        if srch:
            qk = (~ord_key(min2, T)) & mask
            c = int(np.searchsorted(np.array(lv_key[t][st:st + sz], dtype=object), qk, side="right"))
        out += lv_id[t][st:st + c]
    return sorted(out)


def test_closed_forms():
    for n in list(range(1, 80)) + [255, 256, 257, 1000, 4097]:
        M, H = n + 1, n.bit_length()
        occupied = [s for s in range((1 << H) - 1) if level_size(M, depth(s), s + 1 - (1 << depth(s))) >= 1]
        assert sorted(input_slot(e, H) for e in range(n)) == occupied
        for t in range(H - 1):
            for j in range(min(1 << t, 64)):
                assert level_size(M, t, j) >= 1  # every run above the last level holds its root
                for side in (0, 1):
                    assert side_before(M, t, j, side) == sum(level_size(M, t + 1, 2 * x + side) for x in range(j))


def _points(rng, t, i, n, dup):
    p = np.zeros(n, O.point_dtype(t, i))
    if dup:
        p["dim1"] = rng.integers(-4, 4, n)
        p["dim2"] = p["dim1"] + rng.integers(0, 4, n)
        if t in ("f32", "f64"):
            z = rng.random(n) < 0.3
            p["dim1"][z] = -0.0
            p["dim2"][rng.random(n) < 0.2] = -0.0
    elif t in ("f32", "f64"):
        p["dim1"] = rng.random(n)
        p["dim2"] = p["dim1"] + rng.exponential(0.1, n)
    else:
        p["dim1"] = rng.integers(-1000, 1000, n)
        p["dim2"] = p["dim1"] + rng.integers(0, 300, n)
    p["id"] = rng.permutation(n)
    return p


@pytest.mark.parametrize("t,i", [("f32", "u32"), ("i32", "u32"), ("f64", "u64"), ("i64", "i32")])
@pytest.mark.parametrize("n", [1, 2, 3, 6, 7, 8, 33, 130])
def test_scheme_equals_brute_force(t, i, n):
    rng = np.random.default_rng(n * 13 + len(t) + len(i))
    T = O.NP[t]
    for dup in (True, False):
        p = _points(rng, t, i, n, dup)
        tr = O.build(p, t, i)
        lv_key, lv_id = build_index(tr, n, T)
        M, H = n + 1, n.bit_length()
        for lvl in range(H):
            for j in range(1 << lvl):
                sz = level_size(M, lvl, j)
                if sz <= 0:
                    continue
                st = level_start(M, lvl, j)
                run = lv_key[lvl][st:st + sz]
                assert run == sorted(run)
                assert lv_id[lvl][st] == int(tr.ids[(1 << lvl) - 1 + j])  # the root leads its run
        lo, hi = float(p["dim1"].min()), float(p["dim2"].max())
        qs = np.zeros(24, [("min_dim1", T), ("max_dim1", T), ("min_dim2", T), ("kind", np.uint32)])
        for k in range(len(qs)):
            a, b = np.sort(rng.uniform(lo - 1, hi + 1, 2))
            qs[k] = (a, b, rng.uniform(lo - 1, hi + 1), k % 3)
        if t in ("f32", "f64"):
            qs[0] = (-0.0, 0.0, -0.0, THREE)
            qs[1] = (np.nan, 1.0, 0.0, THREE)
            qs[2] = (-np.inf, np.inf, np.nan, RIGHT)
            qs[3] = (0.0, np.nan, -np.inf, LEFT)
        for q in qs:
            want = sorted(p["id"][O.brute_force(p, int(q["kind"]), q["min_dim1"], q["max_dim1"],
                                                 q["min_dim2"])].tolist())
            assert search(tr, n, T, lv_key, lv_id, q) == want, (t, i, n, dup, q)
