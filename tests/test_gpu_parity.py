"""GPU parity: the HIP library against the oracle, the reference's fixtures
and brute force. All calls go through the C ABI (libpst_hip.so)."""
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
    pst_amd.lib()  # fails loudly when the HIP library is not built
    return pst_amd


def _tree(P, pts, t, i):
    d = P.to_device(pts) if len(pts) else None
    import torch
    ptr = d.data_ptr() if d is not None else 0
    tr = P.StaticPSTGPU(ptr, len(pts), t, i)
    torch.cuda.synchronize()
    return tr, d


def _canon_gpu(res, rp, has_id):
    if rp:
        return G.canon(res, True, has_id)
    return G.canon(res, False, True)


@pytest.mark.parametrize("name", [n for n in G.names() if not n.startswith("forest_")])
def test_tree_and_results_match_reference_fixture(torch, P, name):
    fx = G.Fixture(name)
    print(f"fixture {name} n={fx.n} {fx.t}/{fx.i}", flush=True)
    tr, keep = _tree(P, fx.points, fx.t, fx.i)
    if fx.n:
        ex = tr.export_tree()
        for k, v in fx.tree.items():
            assert np.array_equal(G.bits_arr(ex[k]), G.bits_arr(v)), k
    for q, kind, rp, a, b, c in fx.queries():
        res = tr.search(kind, a, b, c, ret=P.RET_POINT if (rp or fx.i is None) else P.RET_ID)
        got = _canon_gpu(res, rp or fx.i is None, fx.i is not None)
        assert got == fx.expected_sorted(q), (q, kind)
    tr.close()


TYPES = [("f32", "u32"), ("i32", "u32"), ("f64", None), ("i64", "i64"), ("f32", "u64"),
         ("f64", "f32"), ("f32", None), ("i32", "i32"), ("i64", "u32"), ("f64", "u64")]


def _rand_points(rng, t, i, n, dup):
    p = np.zeros(n, O.point_dtype(t, i))
    hi = 8 if dup else 10 ** 6
    if t in ("f32", "f64") and not dup:
        p["dim1"] = rng.random(n)
        p["dim2"] = p["dim1"] + rng.exponential(0.05, n)
    else:
        p["dim1"] = rng.integers(-hi, hi, n)
        p["dim2"] = p["dim1"] + rng.integers(0, hi, n)
    if i is not None:
        p["id"] = rng.integers(0, 4, n) if dup else rng.permutation(n)
    return p


@pytest.mark.parametrize("t,i", TYPES)
@pytest.mark.parametrize("n", [5, 100, 2047, 2048, 2049, 5000, 40000])
def test_random_trees_bit_identical_to_oracle(torch, P, t, i, n):
    """Covers the LDS-only build (n <= 2048) and the level passes (n > 2048),
    duplicate-heavy and distinct inputs."""
    rng = np.random.default_rng(n * 31 + len(t) + (0 if i is None else len(i)))
    for dup in (True, False):
        print(f"case {t}/{i} n={n} dup={dup}", flush=True)
        p = _rand_points(rng, t, i, n, dup)
        tr, keep = _tree(P, p, t, i)
        ora = O.build(p, t, i)
        ex = tr.export_tree()
        for k, v in ora.arrays().items():
            assert np.array_equal(G.bits_arr(ex[k]), G.bits_arr(v)), (k, dup)
        lo, hi = float(p["dim1"].min()), float(p["dim2"].max())
        for _ in range(8):
            a, b = np.sort(rng.uniform(lo - 1, hi + 1, 2))
            c = rng.uniform(lo - 1, hi + 1)
            if t in ("i32", "i64"):
                a, b, c = int(a), int(b), int(c)
            for kind in (O.THREE, O.LEFT, O.RIGHT):
                for rp in ((False, True) if i is not None else (True,)):
                    res = tr.search(kind, a, b, c, ret=P.RET_POINT if rp else P.RET_ID)
                    exp = p[O.brute_force(p, kind, a, b, c)]
                    want = G.canon(exp, True, i is not None) if rp else G.canon(exp["id"], False, True)
                    assert _canon_gpu(res, rp, i is not None) == want, (kind, rp, dup)
        tr.close()


@pytest.mark.parametrize("n", [1 << 20, (1 << 20) + 12345])
def test_interval_workload_tree_and_stabbing(torch, P, n):
    """Metacell-shaped float intervals (dim2 = dim1 + Exp(0.02)): tree
    bit-identical to the oracle; stabbing queries (q, q) equal brute force."""
    import pst_data
    p = pst_data.interval_points(n, seed=0)
    print(f"interval n={n}", flush=True)
    tr, keep = _tree(P, p, "f32", "u32")
    ora = O.build(p, "f32", "u32")
    ex = tr.export_tree()
    for k, v in ora.arrays().items():
        assert np.array_equal(G.bits_arr(ex[k]), G.bits_arr(v)), k
    qs = pst_data.stabbing_queries(32, seed=0)
    for q in qs:
        got = np.sort(tr.twoSidedLeftSearch(q, q))
        exp = np.sort(p["id"][O.brute_force(p, O.LEFT, 0, q, q)])
        assert np.array_equal(got, exp)
    tr.close()


def test_search_into_capacity_and_preallocated(torch, P):
    import pst_data
    n = 50000
    p = pst_data.interval_points(n, seed=4)
    tr, keep = _tree(P, p, "f32", "u32")
    q = np.float32(0.5)
    exp = np.sort(p["id"][O.brute_force(p, O.LEFT, 0, q, q)])
    out = torch.zeros(len(exp) + 10, dtype=torch.int32, device="cuda")
    k = tr.search_into(O.LEFT, 0, q, q, out, len(exp) + 10)
    assert k == len(exp)
    got = np.sort(out[:k].cpu().numpy().view(np.uint32))
    assert np.array_equal(got, exp)
    # capacity smaller than the result: the count is still exact
    small = torch.zeros(7, dtype=torch.int32, device="cuda")
    k2 = tr.search_into(O.LEFT, 0, q, q, small, 7)
    assert k2 == len(exp)
    assert set(small.cpu().numpy().view(np.uint32).tolist()) <= set(exp.tolist())
    tr.close()


def test_empty_tree(torch, P, capfd):
    tr = P.StaticPSTGPU(0, 0, "f32", "u32")
    r = tr.twoSidedLeftSearch(0.5, 0.5)
    assert len(r) == 0
    assert "Tree is empty; nothing to search" in capfd.readouterr().out
    tr.close()


def test_nan_input_rejected(torch, P):
    p = np.zeros(10, O.point_dtype("f32", "u32"))
    p["dim1"] = np.arange(10)
    p["dim2"] = p["dim1"] + 1
    p["dim2"][3] = np.nan
    d = P.to_device(p)
    with pytest.raises(P.PSTError) as e:
        P.StaticPSTGPU(d.data_ptr(), 10, "f32", "u32")
    assert e.value.code == 3


@pytest.mark.parametrize("group,copy", [(8, None), (1, None), (8, (16, 65535, 7))])
@pytest.mark.parametrize("t,i", [("f32", "u32"), ("i32", "u32"), ("f64", None), ("f32", "u64"),
                                 ("i64", "u32")])
def test_batched_count_write_equal_single(torch, P, t, i, group, copy):
    """Count pass + exclusive scan + write pass over a mixed batch equals
    brute force, with grouped (8 queries per walk) and ungrouped walks, and
    with whole-subtree copies from 16 points on through a 7-entry task list
    (most copies then run inside the walk)."""
    rng = np.random.default_rng(9)
    n = 30000
    p = _rand_points(rng, t, i, n, dup=False)
    tr, keep = _tree(P, p, t, i)
    tr.set_grouping(group)
    if copy is not None:
        tr.set_subtree_copy(*copy)
    print(f"batched {t}/{i} group={group} copy={copy}", flush=True)
    nq = 300
    qd = P.query_dtype(t)
    qs = np.zeros(nq, qd)
    lo, hi = float(p["dim1"].min()), float(p["dim2"].max())
    for k in range(nq):
        a, b = np.sort(rng.uniform(lo, hi, 2))
        c = rng.uniform(lo, hi)
        qs[k] = (a, b, c, k % 3)
    qd_dev = P.to_device(qs)
    counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
    tr.batch_count(qd_dev, nq, counts)
    torch.cuda.synchronize()
    cnt = counts.cpu().numpy()
    offs = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)
    offs_d = torch.from_numpy(offs).cuda()
    rp = i is None
    rec = P.point_dtype(t, i) if rp else np.dtype(O.NP[i])
    out = torch.zeros(int(cnt.sum()) * rec.itemsize + 8, dtype=torch.uint8, device="cuda")
    tr.batch_write(qd_dev, nq, offs_d, out, ret=P.RET_POINT if rp else P.RET_ID, caps_d=counts)
    torch.cuda.synchronize()
    allres = out.cpu().numpy()[:int(cnt.sum()) * rec.itemsize].view(rec)
    for k in range(nq):
        kind = int(qs["kind"][k])
        a, b, c = qs["min_dim1"][k], qs["max_dim1"][k], qs["min_dim2"][k]
        exp = p[O.brute_force(p, kind, a, b, c)]
        assert cnt[k] == len(exp), k
        seg = allres[offs[k]:offs[k] + cnt[k]]
        if rp:
            assert G.canon(seg, True, False) == G.canon(exp, True, False)
        else:
            assert G.canon(seg, False, True) == G.canon(exp["id"], False, True)
    tr.close()


@pytest.mark.parametrize("t,i", [("f32", "u32"), ("i32", "u32"), ("f64", "u64"), ("f32", None)])
def test_interval_parallel_search_equals_brute_force(torch, P, t, i):
    rng = np.random.default_rng(3)
    n = 200003
    p = _rand_points(rng, t, i, n, dup=False)
    d = P.to_device(p)
    lo, hi = float(p["dim1"].min()), float(p["dim2"].max())
    for q in rng.uniform(lo, hi, 5):
        q = O.NP[t](q)
        for rp in ((False, True) if i is not None else (True,)):
            res = P.intervalParallelSearch(d, n, q, t, i, ret=P.RET_POINT if rp else P.RET_ID)
            exp = p[O.brute_force_ips(p, q)]
            want = G.canon(exp, True, i is not None) if rp else G.canon(exp["id"], False, True)
            got = G.canon(res, True, i is not None) if rp else G.canon(res, False, True)
            assert got == want


@pytest.mark.parametrize("n", [1, 1000, 1024, 70000, 1 << 20])
def test_interval_parallel_search_dense_sparse_and_capacity(torch, P, n):
    """Hit densities that take every path of the filter: chunks where every
    record hits (written directly), sparse chunks (staged), both mixed in
    one wave's chunks, bounds equal to stored values; a capacity below the
    hit count writes only that many results (no duplicates, nothing past
    it) and still counts every hit."""
    rng = np.random.default_rng(17)
    p = np.zeros(n, O.point_dtype("f32", "u32"))
    a = rng.random(n).astype(np.float32)
    w = np.where((np.arange(n) // 3000) % 2 == 0, 1.0, 0.001).astype(np.float32)  # dense and sparse runs
    p["dim1"], p["dim2"], p["id"] = a, a + w, np.arange(n, dtype=np.uint32)
    d = P.to_device(p)
    for q in (np.float32(0.5), p["dim1"][n // 2], p["dim2"][n // 3], np.float32(2.0)):
        exp = np.sort(p["id"][O.brute_force_ips(p, q)])
        got = np.sort(P.intervalParallelSearch(d, n, q, "f32", "u32"))
        assert np.array_equal(got, exp), (n, q)
    q = np.float32(0.5)
    exp = np.sort(p["id"][O.brute_force_ips(p, q)])
    cap = len(exp) // 2
    out = torch.full((cap + 64,), -1, dtype=torch.int32, device="cuda")
    ctr = torch.zeros(1, dtype=torch.int64, device="cuda")
    P.ips_into(d, n, q, out, cap, ctr)
    torch.cuda.synchronize()
    assert int(ctr.item()) == len(exp)
    o = out.cpu().numpy().view(np.uint32)
    head = np.sort(o[:cap])
    assert len(np.unique(head)) == cap and np.all(np.isin(head, exp))
    assert np.all(out[cap:].cpu().numpy() == -1)


@pytest.mark.parametrize("group,copy", [(8, None), (1, None), (8, (0, 0)), (8, (16, 8191)), (1, (64, 65535)),
                                        (8, (16, 65535, 5)), (8, (256, 2047, 300))])
def test_batch_search_interval_workload(torch, P, group, copy):
    """One-call count + scan + write (pst_gpu_batch_search) on a 2^20-point
    interval tree with a batch of 515 queries: stabbing queries (both
    sides), three-sided queries of log-uniform width, a duplicate, two
    queries that report every point, and a batch size that leaves the last
    group partial. Every query's results equal the oracle's."""
    import pst_data as D
    n = 1 << 20
    pts = D.interval_points(n, seed=4)
    tr, keep = _tree(P, pts, "f32", "u32")
    tr.set_grouping(group)
    if copy is not None:
        tr.set_subtree_copy(*copy)
    qs = D.float_query_mix(128, 384, seed=5)
    # a duplicate query, and two that report every point (deep walk stacks
    # that move to the overflow stack): 515 in all
    full = np.zeros(2, qs.dtype)
    full[0] = (0.0, 1.0, 0.0, 0)
    full[1] = (0.0, 0.0, 0.0, 2)
    qs = np.concatenate([qs, qs[:1], full])
    nq = len(qs)
    print(f"batch_search n=2^20 nq={nq} group={group}", flush=True)
    qd = P.to_device(qs)
    counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
    offs = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
    ora = O.build(pts, "f32", "u32")
    exp = [O.search(ora, int(q["kind"]), q["min_dim1"], q["max_dim1"], q["min_dim2"]) for q in qs]
    total = sum(len(e) for e in exp)
    out = torch.zeros(total + 16, dtype=torch.int32, device="cuda")
    tr.batch_search(qd, nq, counts, offs, out, total + 16)
    torch.cuda.synchronize()
    cnt = counts.cpu().numpy()
    off = offs.cpu().numpy()
    res = out.cpu().numpy().view(np.uint32)
    assert off[nq] == total
    for k in range(nq):
        assert cnt[k] == len(exp[k]), k
        got = np.sort(res[off[k]:off[k] + cnt[k]])
        assert np.array_equal(got, np.sort(exp[k])), k
    tr.close()
