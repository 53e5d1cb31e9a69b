"""GPU parity of the forest of shallow trees (StaticPSTGPUArr, through the C
ABI): each tree equals the oracle's tree of its input chunk (and the
reference's own forest fixture), and searches report exactly the points
that satisfy the predicate, including the right and three-sided searches
that the reference leaves as stubs."""
from collections import Counter

import numpy as np
import pytest

import golden_util as G
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.fixture(scope="module")
def P():
    import pst_amd
    pst_amd.lib()
    return pst_amd


def _forest(P, pts, t, i, tpb):
    import torch
    d = P.to_device(pts) if len(pts) else None
    f = P.StaticPSTGPUArr(d.data_ptr() if d is not None else 0, len(pts), tpb, t, i)
    torch.cuda.synchronize()
    return f, d


def _check_trees(P, f, pts, t, i):
    """Export of every tree equals the oracle's tree of its chunk."""
    if len(pts) == 0:
        return
    ex = f.export_tree()
    st = f.slots_per_tree
    off = 0
    for c, c0 in enumerate(range(0, len(pts), st)):
        tr = O.build(pts[c0:c0 + st], t, i)
        for k, v in tr.arrays().items():
            got = ex[k][off:off + tr.S]
            assert np.array_equal(G.bits_arr(got), G.bits_arr(v)), (c, k)
        off += tr.S
    assert off == len(ex["dim1"])


def test_forest_fixture(torch, P):
    fx = G.Fixture("forest_f32_u32_tpb64")
    chunk = int(fx.forest["chunk"])
    f, keep = _forest(P, fx.points, fx.t, fx.i, 64)
    assert f.slots_per_tree == chunk
    ex = f.export_tree()
    assert f.forest_sizes() == [int(x) for x in fx.forest["sizes"]]
    for k in ("dim1", "dim2", "med", "bits", "ids"):
        assert np.array_equal(G.bits_arr(ex[k]), G.bits_arr(fx.forest[k])), k
    for q, kind, rp, a, b, c in fx.queries():
        res = f.search(kind, a, b, c, ret=P.RET_POINT if rp else P.RET_ID)
        got = G.canon(res, True, True) if rp else G.canon(res, False, True)
        assert got == fx.expected_sorted(q), (q, kind)
    f.close()


def _rand(rng, t, i, n):
    dt = P_point_dtype(t, i)
    p = np.zeros(n, dt)
    if t in ("f32", "f64"):
        a = rng.integers(0, 40, n).astype(np.float64) / 8
        w = rng.integers(0, 16, n).astype(np.float64) / 8
    else:
        a = rng.integers(-50, 50, n)
        w = rng.integers(0, 20, n)
    p["dim1"] = a
    p["dim2"] = a + w
    if i is not None:
        p["id"] = rng.integers(0, max(2, n // 3), n)
    return p


def P_point_dtype(t, i):
    import pst_amd
    return pst_amd.point_dtype(t, i)


FOREST_CASES = [("f32", "u32", 64, 5000), ("f32", "u32", 1, 300), ("f32", "u32", 31, 3001),
                ("i32", "u32", 100, 20000), ("f64", None, 1024, 9000), ("i64", "i64", 128, 4099),
                ("f32", "u64", 255, 7777), ("f64", "f32", 64, 2000), ("f32", "u32", 64, 127),
                ("f32", "u32", 64, 128), ("f32", "u32", 64, 1), ("f32", "u32", 64, 0)]


@pytest.mark.parametrize("t,i,tpb,n", FOREST_CASES)
def test_forest_random(torch, P, t, i, tpb, n):
    rng = np.random.default_rng(n + tpb)
    print(f"forest {t}/{i} tpb={tpb} n={n}", flush=True)
    p = _rand(rng, t, i, n)
    f, keep = _forest(P, p, t, i, tpb)
    _check_trees(P, f, p, t, i)
    lo, hi = (-1, 8) if t in ("f32", "f64") else (-60, 80)
    for it in range(6):
        a, b = np.sort(rng.uniform(lo, hi, 2))
        c = rng.uniform(lo, hi)
        if t in ("i32", "i64"):
            a, b, c = int(a), int(b), int(c)
        for kind in (0, 1, 2):
            exp = p[O.brute_force(p, kind, a, b, c)]
            rp = i is None or (it % 2 == 1)
            if n == 0:
                continue
            res = f.search(kind, a, b, c, ret=P.RET_POINT if rp else P.RET_ID)
            if rp:
                assert G.canon(res, True, i is not None) == G.canon(exp, True, i is not None), (it, kind)
            else:
                assert G.canon(res, False, True) == G.canon(exp["id"], False, True), (it, kind)
    f.close()


@pytest.mark.parametrize("group", [8, 1])
def test_forest_batch_isovalues(torch, P, group):
    """Batched stabbing (isovalue) and mixed queries on a forest of 2^18
    metacell-like intervals, count + scan + write in one call."""
    import pst_data as D
    n = (1 << 18) + 77
    pts = D.interval_points(n, seed=8)
    f, keep = _forest(P, pts, "f32", "u32", 64)
    f.set_grouping(group)
    qs = D.float_query_mix(40, 160, seed=9)
    iso = np.linspace(0.0, 1.0, 41).astype(np.float32)
    st = np.zeros(len(iso), qs.dtype)
    st["min_dim1"] = st["max_dim1"] = st["min_dim2"] = iso
    st["kind"] = 1
    qs = np.concatenate([qs, st])
    nq = len(qs)
    print(f"forest batch nq={nq} group={group}", flush=True)
    qd = P.to_device(qs)
    exp = [pts["id"][O.brute_force(pts, int(q["kind"]), q["min_dim1"], q["max_dim1"], q["min_dim2"])]
           for q in qs]
    total = sum(len(e) for e in exp)
    counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
    offs = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
    out = torch.zeros(total + 8, dtype=torch.int32, device="cuda")
    f.batch_search(qd, nq, counts, offs, out, total + 8)
    torch.cuda.synchronize()
    cnt, off = counts.cpu().numpy(), offs.cpu().numpy()
    res = out.cpu().numpy().view(np.uint32)
    assert off[nq] == total
    for k in range(nq):
        assert cnt[k] == len(exp[k]), k
        assert np.array_equal(np.sort(res[off[k]:off[k] + cnt[k]]), np.sort(exp[k])), k
    f.close()


def test_forest_calc_global_mem_needed(P):
    n, tpb = 1 << 20, 64
    m = P.StaticPSTGPUArr.calcGlobalMemNeeded(n, tpb, "f32", "u32")
    trees = (n + 126) // 127
    assert m >= trees * 128 * 16 + n * 12


def test_forest_calc_global_mem_needed_covers_measured(torch, P):
    """pst_gpu_arr_calc_global_mem_needed (pst/static-pst-gpu-arr.tu:662-686)
    against what the device reports: the forest, a 512-query stationary
    batch's scratch (stabbing and general) and a blocking search's buffers
    fit in it. The caller's own query and output buffers are allocated
    before the first reading."""
    import pst_data as D
    n, tpb = (1 << 22) + 5, 64
    pts = D.interval_points(n, seed=62)
    d = P.to_device(pts)
    iso = np.linspace(0.0, 1.0, 512).astype(np.float32)
    st = np.zeros(512, P.query_dtype("f32"))
    st["max_dim1"] = st["min_dim2"] = iso
    st["kind"] = 1
    gen = D.float_query_mix(256, 256, seed=63)
    dst, dgen = P.to_device(st), P.to_device(gen)
    counts = torch.zeros(512, dtype=torch.int64, device="cuda")
    offs = torch.zeros(513, dtype=torch.int64, device="cuda")
    out = torch.empty(1 << 20, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    f = P.StaticPSTGPUArr(d.data_ptr(), n, tpb, "f32", "u32")
    f.batch_search(dst, 512, counts, offs, out, 1 << 20)
    f.batch_search(dgen, 512, counts, offs, out, 1 << 20)
    f.search(P.LEFT, 0, 0.5, 0.5)
    torch.cuda.synchronize()
    used = free0 - torch.cuda.mem_get_info()[0]
    calc = P.StaticPSTGPUArr.calcGlobalMemNeeded(n, tpb, "f32", "u32")
    trees = (n + 126) // 127
    print(f"forest resident after batches and a search {used}, calcGlobalMemNeeded {calc}", flush=True)
    assert trees * 128 * (16 + 4 + 4) <= used <= calc
    f.close()


STATIONARY_CASES = [("f32", "u32", 64, 5000), ("f32", "u32", 1, 300), ("f32", "u32", 31, 3001),
                    ("i32", "u32", 100, 20000), ("f64", None, 255, 9000), ("i64", "i64", 128, 4099),
                    ("f32", "u64", 256, 7777), ("f64", "f32", 64, 2000), ("f32", "u32", 64, 127),
                    ("f32", "u32", 64, 1), ("f32", "u32", 1024, 6000)]


def _query_batch(rng, t, nq, lo, hi):
    import pst_amd as P
    qs = np.zeros(nq, P.query_dtype(t))
    a = rng.uniform(lo, hi, nq)
    b = a + rng.exponential((hi - lo) / 4, nq)
    c = rng.uniform(lo, hi, nq)
    if t in ("i32", "i64"):
        a, b, c = a.astype(np.int64), b.astype(np.int64), c.astype(np.int64)
    qs["min_dim1"], qs["max_dim1"], qs["min_dim2"] = a, b, c
    qs["kind"] = rng.integers(0, 4, nq)  # kind 3 searches as a right search
    if t in ("f32", "f64") and nq > 8:
        qs["min_dim1"][0] = np.nan
        qs["max_dim1"][1] = np.nan
        qs["min_dim2"][2] = np.nan
        qs["min_dim2"][3] = -np.inf
        qs["max_dim1"][3] = np.inf
        qs["kind"][3] = 1
    return qs


def _expect_batch(p, qs):
    out = []
    for q in qs:
        kind = int(q["kind"])
        out.append(O.brute_force(p, kind if kind < 3 else 2, q["min_dim1"], q["max_dim1"], q["min_dim2"]))
    return out


@pytest.mark.parametrize("t,i,tpb,n", STATIONARY_CASES)
def test_forest_stationary_batches(torch, P, t, i, tpb, n):
    """Forest batches of up to 512 queries (two passes over the forest with
    the trees held stationary; 1024 threads per block and 513 queries take
    the walks): counts, one-call writes, truncating capacities, separate
    count and write calls with per-query caps, IDs and whole points; every
    query's results equal brute force, including NaN and infinite bounds."""
    rng = np.random.default_rng(tpb * 7 + n)
    p = _rand(rng, t, i, n)
    f, keep = _forest(P, p, t, i, tpb)
    lo, hi = (-1, 8) if t in ("f32", "f64") else (-60, 80)
    rec = np.dtype(O.NP[i]) if i is not None else P.point_dtype(t, i)
    for nq in (7, 512, 513):
        print(f"stationary forest {t}/{i} tpb={tpb} n={n} nq={nq}", flush=True)
        qs = _query_batch(rng, t, nq, lo, hi)
        exp = _expect_batch(p, qs)
        want = np.array([len(e) for e in exp], np.int64)
        total = int(want.sum())
        dq = P.to_device(qs)
        rp = P.RET_POINT if i is None else P.RET_ID
        rb = rec.itemsize
        counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
        offs = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
        f.batch_count(dq, nq, counts)
        torch.cuda.synchronize()
        assert np.array_equal(counts.cpu().numpy(), want)
        out = torch.full((total * rb + 64,), 0x5A, dtype=torch.uint8, device="cuda")
        f.batch_search(dq, nq, counts, offs, out, total, ret=rp)
        torch.cuda.synchronize()
        o = offs.cpu().numpy()
        raw = out.cpu().numpy()
        res = raw[:total * rb].view(rec)
        for k in range(nq):
            seg = res[o[k]:o[k] + want[k]]
            if rp == P.RET_POINT:
                assert G.canon(seg, True, i is not None) == G.canon(p[exp[k]], True, i is not None), k
            else:
                assert np.array_equal(np.sort(seg), np.sort(p["id"][exp[k]])), k
        assert np.all(raw[total * rb:] == 0x5A)
        # a truncating capacity writes exactly the full call's first records
        cap = total // 3
        out2 = torch.full_like(out, 0x5A)
        f.batch_search(dq, nq, counts, offs, out2, cap, ret=rp)
        torch.cuda.synchronize()
        raw2 = out2.cpu().numpy()
        if nq <= 512 and f.slots_per_tree <= 511:  # stationary passes: fixed output positions
            assert np.array_equal(raw2[:cap * rb], raw[:cap * rb])
        assert np.all(raw2[cap * rb:] == 0x5A)
        # separate count and write calls, each query capped at half its results
        caps = torch.from_numpy(want // 2).to("cuda")
        out3 = torch.full_like(out, 0x5A)
        f.batch_write(dq, nq, offs, out3, ret=rp, caps_d=caps)
        torch.cuda.synchronize()
        res3 = out3.cpu().numpy()[:total * rb].view(rec)
        for k in range(0, nq, 5):
            seg = res3[o[k]:o[k] + want[k] // 2]
            if rp == P.RET_ID:  # a sub-multiset of the query's IDs (random IDs repeat)
                assert not (Counter(seg.tolist()) - Counter(p["id"][exp[k]].tolist())), k
            else:
                full = Counter(G.canon(p[exp[k]], True, i is not None))
                got = Counter(G.canon(seg, True, i is not None))
                assert not (got - full), k
            gap = out3.cpu().numpy()[(o[k] + want[k] // 2) * rb:(o[k] + want[k]) * rb]
            assert np.all(gap == 0x5A), k
    f.close()


def _stab_batch(rng, t, nq, lo, hi):
    """Stabbing (isovalue) queries: LEFT searches with max_dim1 = min_dim2 =
    v, on the points' grid (ties), with NaN, infinite, repeated and
    out-of-range values; min_dim1 (unused by LEFT) is arbitrary."""
    import pst_amd as P
    qs = np.zeros(nq, P.query_dtype(t))
    if t in ("f32", "f64"):
        v = rng.integers(int(lo * 8) - 8, int(hi * 8) + 8, nq).astype(np.float64) / 8
    else:
        v = rng.integers(lo - 5, hi + 5, nq)
    if nq > 8:
        v[4] = v[5]
        if t in ("f32", "f64"):
            v[0], v[1], v[2] = np.nan, np.inf, -np.inf
        else:
            info = np.iinfo(np.int32 if t == "i32" else np.int64)
            v[0], v[1] = info.min, info.max
    qs["max_dim1"] = qs["min_dim2"] = v
    qs["min_dim1"] = rng.uniform(lo, hi, nq).astype(qs["min_dim1"].dtype)
    qs["kind"] = 1
    return qs


# plus whole 8-byte points (records staged like IDs)
STAB_CASES = STATIONARY_CASES + [("f32", None, 64, 5000), ("i32", None, 31, 3001)]


@pytest.mark.parametrize("t,i,tpb,n", STAB_CASES)
def test_forest_stabbing_batches(torch, P, t, i, tpb, n):
    """Batches of stabbing (isovalue) queries, which stationary forest
    passes take through sorted values and per-point runs of them: counts,
    one-call writes, truncating capacities and separate count and write
    calls equal brute force; and the records land at the same positions,
    byte for byte, as through the general passes (the same batch plus a
    final query with no hits, which is not a stabbing query)."""
    rng = np.random.default_rng(tpb * 11 + n)
    p = _rand(rng, t, i, n)
    f, keep = _forest(P, p, t, i, tpb)
    lo, hi = (-1, 8) if t in ("f32", "f64") else (-60, 80)
    rec = np.dtype(O.NP[i]) if i is not None else P.point_dtype(t, i)
    rp = P.RET_POINT if i is None else P.RET_ID
    rb = rec.itemsize
    for nq in (9, 200, 511):
        print(f"stabbing forest {t}/{i} tpb={tpb} n={n} nq={nq}", flush=True)
        qs = _stab_batch(rng, t, nq, lo, hi)
        exp = _expect_batch(p, qs)
        want = np.array([len(e) for e in exp], np.int64)
        total = int(want.sum())
        dq = P.to_device(qs)
        counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
        offs = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
        f.batch_count(dq, nq, counts)
        torch.cuda.synchronize()
        assert np.array_equal(counts.cpu().numpy(), want)
        out = torch.full((total * rb + 64,), 0x5A, dtype=torch.uint8, device="cuda")
        f.batch_search(dq, nq, counts, offs, out, total, ret=rp)
        torch.cuda.synchronize()
        o = offs.cpu().numpy()
        raw = out.cpu().numpy()
        res = raw[:total * rb].view(rec)
        for k in range(nq):
            seg = res[o[k]:o[k] + want[k]]
            if rp == P.RET_POINT:
                assert G.canon(seg, True, i is not None) == G.canon(p[exp[k]], True, i is not None), k
            else:
                assert np.array_equal(np.sort(seg), np.sort(p["id"][exp[k]])), k
        assert np.all(raw[total * rb:] == 0x5A)
        # the general passes put every record at the same position
        q2 = np.concatenate([qs, np.zeros(1, qs.dtype)])
        q2["kind"][-1] = 0
        q2["min_dim1"][-1], q2["max_dim1"][-1] = hi + 100, lo - 100  # an empty dim1 range
        d2 = P.to_device(q2)
        c2 = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
        o2 = torch.zeros(nq + 2, dtype=torch.int64, device="cuda")
        out_g = torch.full_like(out, 0x5A)
        f.batch_search(d2, nq + 1, c2, o2, out_g, total, ret=rp)
        torch.cuda.synchronize()
        assert int(c2[-1]) == 0
        if f.slots_per_tree <= 511:  # stationary passes: fixed output positions
            assert np.array_equal(out_g.cpu().numpy(), raw)
        # a truncating capacity writes exactly the full call's first records
        cap = total // 3
        out2 = torch.full_like(out, 0x5A)
        f.batch_search(dq, nq, counts, offs, out2, cap, ret=rp)
        torch.cuda.synchronize()
        raw2 = out2.cpu().numpy()
        if f.slots_per_tree <= 511:  # stationary passes: fixed output positions
            assert np.array_equal(raw2[:cap * rb], raw[:cap * rb])
        assert np.all(raw2[cap * rb:] == 0x5A)
        # separate count and write calls, each query capped at half its results
        caps = torch.from_numpy(want // 2).to("cuda")
        out3 = torch.full_like(out, 0x5A)
        f.batch_write(dq, nq, offs, out3, ret=rp, caps_d=caps)
        torch.cuda.synchronize()
        raw3 = out3.cpu().numpy()
        for k in range(nq):
            h = int(want[k] // 2)
            if f.slots_per_tree <= 511:
                assert np.array_equal(raw3[o[k] * rb:(o[k] + h) * rb], raw[o[k] * rb:(o[k] + h) * rb]), k
            assert np.all(raw3[(o[k] + h) * rb:(o[k] + want[k]) * rb] == 0x5A), k
    f.close()
