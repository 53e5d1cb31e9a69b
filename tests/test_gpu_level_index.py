"""Batched searches through the level index (per depth, every subtree's
points sorted by dim2 descending): each query's results are the oracle's
set, for every data type pairing with IDs, tree sizes from one point up
(every last-level fill), duplicate-heavy inputs (ties in dim2 decide which
point leads a subtree's run), signed zeros, NaN and infinite bounds, empty
results, truncating capacities, the one-call count + scan + write, graph
replays, and switching the index on and off. All calls go through the C
ABI."""
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


def _points(rng, t, i, n, dup):
    p = np.zeros(n, O.point_dtype(t, i))
    if t in ("f32", "f64") and not dup:
        p["dim1"] = rng.random(n)
        p["dim2"] = p["dim1"] + rng.exponential(0.05, n)
    else:
        hi = 6 if dup else 10 ** 6
        p["dim1"] = rng.integers(-hi, hi, n)
        p["dim2"] = p["dim1"] + rng.integers(0, hi, n)
    if t in ("f32", "f64") and dup and n > 4:
        # signed zeros in both dims (equal under IEEE compares)
        k = rng.choice(n, n // 4, replace=False)
        p["dim1"][k[: len(k) // 2]] = -0.0
        p["dim2"][k[len(k) // 2:]] = -0.0
        p["dim1"][k[len(k) // 2:]] = np.minimum(p["dim1"][k[len(k) // 2:]], -0.0)
    p["id"] = rng.permutation(n)
    return p


def _queries(rng, P, t, p, nq):
    qs = np.zeros(nq, P.query_dtype(t))
    lo, hi = float(p["dim1"].min()), float(p["dim2"].max())
    span = hi - lo + 1
    for k in range(nq):
        a, b = np.sort(rng.uniform(lo - 0.1 * span, hi + 0.1 * span, 2))
        c = rng.uniform(lo - 0.1 * span, hi + 0.1 * span)
        qs[k] = (a, b, c, k % 3)
    if t in ("f32", "f64"):
        qs[0] = (-0.0, 0.0, -0.0, 0)
        qs[1] = (np.nan, 0.5, 0.0, 0)
        qs[2] = (0.0, np.nan, -np.inf, 1)
        qs[3] = (-np.inf, np.inf, np.nan, 2)
        qs[4] = (-np.inf, np.inf, -np.inf, 0)
        qs[5] = (np.nan, np.nan, -np.inf, 2)
    else:
        qs[0] = (lo, hi, lo, 0)   # every point
        qs[1] = (hi, lo, lo, 0)   # empty dim1 range
        qs[2] = (0, 0, 0, 1)
    return qs


def _expect(p, qs):
    return [np.sort(p["id"][O.brute_force(p, int(q["kind"]), q["min_dim1"], q["max_dim1"], q["min_dim2"])])
            for q in qs]


TYPES = [("f32", "u32"), ("i32", "u32"), ("f64", "u64"), ("f32", "u64"), ("i64", "u32"), ("i32", "i64")]


@pytest.mark.parametrize("t,i", TYPES)
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 9, 100, 4095, 4096, 4097, 30000])
def test_count_and_write_equal_brute_force(torch, P, t, i, n):
    rng = np.random.default_rng(n * 7 + len(t) + len(i))
    for dup in (True, False):
        print(f"level index {t}/{i} n={n} dup={dup}", flush=True)
        p = _points(rng, t, i, n, dup)
        d = P.to_device(p)
        tr = P.StaticPSTGPU(d.data_ptr(), n, t, i)
        tr.set_level_index(True)
        qs = _queries(rng, P, t, p, 96)
        nq = len(qs)
        qd = P.to_device(qs)
        counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
        tr.batch_count(qd, nq, counts)
        torch.cuda.synchronize()
        cnt = counts.cpu().numpy()
        exp = _expect(p, qs)
        assert [len(e) for e in exp] == cnt.tolist()
        offs = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)
        rec = np.dtype(O.NP[i])
        total = int(cnt.sum())
        out = torch.full(((total + 8) * rec.itemsize,), 0xAB, dtype=torch.uint8, device="cuda")
        tr.batch_write(qd, nq, torch.from_numpy(offs).cuda(), out, ret=P.RET_ID, caps_d=counts)
        torch.cuda.synchronize()
        raw = out.cpu().numpy()
        res = raw[:total * rec.itemsize].view(rec)
        assert np.all(raw[total * rec.itemsize:] == 0xAB)  # nothing past the results
        for k in range(nq):
            assert np.array_equal(np.sort(res[offs[k]:offs[k] + cnt[k]]), exp[k]), (k, qs[k])
        tr.close()


def test_caps_truncate_to_a_subset(torch, P):
    """A capacity below a query's count writes that many of its IDs (a
    subset of its results) and nothing beyond."""
    rng = np.random.default_rng(21)
    n = 50000
    p = _points(rng, "f32", "u32", n, False)
    tr = P.StaticPSTGPU(P.to_device(p).data_ptr(), n, "f32", "u32")
    tr.set_level_index(True)
    qs = _queries(rng, P, "f32", p, 64)
    nq = len(qs)
    exp = _expect(p, qs)
    caps = np.array([len(e) // 3 if k % 2 else len(e) for k, e in enumerate(exp)], np.int64)
    offs = np.concatenate([[0], np.cumsum(caps)[:-1]]).astype(np.int64)
    total = int(caps.sum())
    out = torch.full((total + 64,), -7, dtype=torch.int32, device="cuda")
    tr.batch_write(P.to_device(qs), nq, torch.from_numpy(offs).cuda(), out, ret=P.RET_ID,
                   caps_d=torch.from_numpy(caps).cuda())
    torch.cuda.synchronize()
    r = out.cpu().numpy().view(np.uint32)
    assert np.all(out.cpu().numpy()[total:] == -7)
    for k in range(nq):
        seg = r[offs[k]:offs[k] + caps[k]]
        assert len(np.unique(seg)) == caps[k] and np.all(np.isin(seg, exp[k])), k
    tr.close()


def test_batch_search_interval_workload_and_toggle(torch, P):
    """One-call count + scan + write on a 2^20-point interval tree (stabbing
    and three-sided queries, two that report every point), with the index
    on, then off, then on again: the same sets as the oracle each time."""
    import pst_data as D
    n = 1 << 20
    pts = D.interval_points(n, seed=4)
    tr = P.StaticPSTGPU(P.to_device(pts).data_ptr(), n, "f32", "u32")
    qs = D.float_query_mix(128, 384, seed=5)
    full = np.zeros(2, qs.dtype)
    full[0] = (0.0, 1.0, 0.0, 0)
    full[1] = (0.0, 0.0, 0.0, 2)
    qs = np.concatenate([qs, qs[:1], full])
    nq = len(qs)
    qd = P.to_device(qs)
    exp = _expect(pts, qs)
    total = sum(len(e) for e in exp)
    assert P.StaticPSTGPU.level_index_bytes(n, "f32", "u32") == 21 * n * 8
    assert P.StaticPSTGPU.level_index_bytes(1000, "f64", "u32") == 10 * 1024 * 12
    for on in (True, False, True):
        print(f"batch_search level index on={on}", flush=True)
        tr.set_level_index(on)
        counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
        offs = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
        out = torch.zeros(total + 16, dtype=torch.int32, device="cuda")
        tr.batch_search(qd, nq, counts, offs, out, total + 16)
        torch.cuda.synchronize()
        c, o, r = counts.cpu().numpy(), offs.cpu().numpy(), out.cpu().numpy().view(np.uint32)
        assert o[nq] == total
        for k in range(nq):
            assert np.array_equal(np.sort(r[o[k]:o[k] + c[k]]), exp[k]), k
    tr.close()


def test_graph_replays_with_level_index(torch, P):
    import pst_data as D
    n = 1 << 18
    pts = D.interval_points(n, seed=12)
    tr = P.StaticPSTGPU(P.to_device(pts).data_ptr(), n, "f32", "u32")
    tr.set_level_index(True)
    qs = D.float_query_mix(300, 700, seed=13)
    nq = len(qs)
    dq = P.to_device(qs)
    s = torch.cuda.Stream()
    counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
    offs = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
    tr.batch_count(dq, nq, counts)
    torch.cuda.synchronize()
    cap = int(counts.sum()) + 64
    out = torch.zeros(cap, dtype=torch.int32, device="cuda")
    exp = _expect(pts, qs)
    g = tr.batch_graph(dq, nq, counts, offs, out, cap, stream=s)
    for rep in range(3):
        with torch.cuda.stream(s):
            counts.zero_(), offs.zero_(), out.fill_(-1)
            tmp = torch.randn(1 << 20, device="cuda").sum()
        g.launch(s)
        s.synchronize()
        print(f"replay {rep}", flush=True)
        c, o, r = counts.cpu().numpy(), offs.cpu().numpy(), out.cpu().numpy().view(np.uint32)
        for k in range(nq):
            assert np.array_equal(np.sort(r[o[k]:o[k] + c[k]]), exp[k]), (rep, k)
        del tmp
    g.close()
    tr.close()


def test_rejected_without_ids_and_for_forests(torch, P):
    rng = np.random.default_rng(2)
    p = _points(rng, "f32", "u32", 1000, False)
    q = np.zeros(1000, O.point_dtype("f32", None))
    q["dim1"], q["dim2"] = p["dim1"], p["dim2"]
    t1 = P.StaticPSTGPU(P.to_device(q).data_ptr(), 1000, "f32", None)
    with pytest.raises(P.PSTError) as e:
        t1.set_level_index(True)
    assert e.value.code == 1 and "IDs" in str(e.value)
    t1.close()
    f = P.StaticPSTGPUArr(P.to_device(p).data_ptr(), 1000, 64, "f32", "u32")
    with pytest.raises(P.PSTError) as e:
        f.set_level_index(True)
    assert e.value.code == 1 and "forest" in str(e.value)
    f.close()
    assert P.StaticPSTGPU.level_index_bytes(1000, "f32", None) == 0


@pytest.mark.parametrize("t,i", [("f32", "u32"), ("i64", "u64"), ("f64", "i32")])
def test_single_searches_through_the_index(torch, P, t, i):
    """threeSidedSearch / twoSidedLeftSearch / twoSidedRightSearch and
    search_into (with a capacity that is too small: the count is still
    returned, at most capacity IDs written) through the index equal brute
    force; points are still returned by the walk."""
    rng = np.random.default_rng(len(t) * 5 + len(i))
    n = 20011
    p = _points(rng, t, i, n, t != "f32")
    tr = P.StaticPSTGPU(P.to_device(p).data_ptr(), n, t, i)
    tr.set_level_index(True)
    qs = _queries(rng, P, t, p, 24)
    out = torch.full((n + 16,), -1, dtype=torch.int64, device="cuda")
    for q in qs:
        kind, a, b, c = int(q["kind"]), q["min_dim1"], q["max_dim1"], q["min_dim2"]
        exp = p[O.brute_force(p, kind, a, b, c)]
        got = tr.search(kind, a, b, c)
        assert np.array_equal(np.sort(got), np.sort(exp["id"]))
        pts = tr.search(kind, a, b, c, ret=P.RET_POINT)
        assert G.canon(pts, True, True) == G.canon(exp, True, True)
        cap = max(0, len(exp) // 2)
        out.fill_(-1)
        num = tr.search_into(kind, a, b, c, out, cap)
        assert num == len(exp)
        r = out.cpu().numpy().view(O.NP[i])
        isz = np.dtype(O.NP[i]).itemsize
        w = r[:cap]
        assert len(np.unique(w)) == cap and np.all(np.isin(w, exp["id"]))
        assert np.all(out.cpu().numpy().view(np.uint8)[cap * isz:] == 0xFF)
    tr.close()


@pytest.mark.parametrize("index", [True, False])
def test_gapped_offsets_and_short_output(torch, P, index):
    """Writes land at the caller's offsets (gaps between queries untouched),
    and a one-call batch whose output holds fewer IDs than the batch
    reports writes nothing at or past the capacity, while the queries that
    fit are complete; the same with the index and with the walk."""
    rng = np.random.default_rng(31)
    n = 40000
    p = _points(rng, "f32", "u32", n, False)
    tr = P.StaticPSTGPU(P.to_device(p).data_ptr(), n, "f32", "u32")
    tr.set_level_index(index)
    qs = _queries(rng, P, "f32", p, 48)
    nq = len(qs)
    qd = P.to_device(qs)
    exp = _expect(p, qs)
    cnt = np.array([len(e) for e in exp], np.int64)
    offs = (np.concatenate([[0], np.cumsum(cnt)[:-1]]) + 7 * np.arange(nq)).astype(np.int64)
    total = int(offs[-1] + cnt[-1] + 7)
    out = torch.full((total,), -5, dtype=torch.int32, device="cuda")
    tr.batch_write(qd, nq, torch.from_numpy(offs).cuda(), out, ret=P.RET_ID)
    torch.cuda.synchronize()
    r = out.cpu().numpy()
    mask = np.ones(total, bool)
    for k in range(nq):
        assert np.array_equal(np.sort(r[offs[k]:offs[k] + cnt[k]].view(np.uint32)), exp[k]), k
        mask[offs[k]:offs[k] + cnt[k]] = False
    assert np.all(r[mask] == -5)
    # one call with a short output: the last queries are cut at the capacity
    full = int(cnt.sum())
    cap = full - int(cnt[-1]) // 2 - 3
    counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
    offd = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
    out2 = torch.full((full + 64,), -5, dtype=torch.int32, device="cuda")
    tr.batch_search(qd, nq, counts, offd, out2, cap)
    torch.cuda.synchronize()
    c, o, r2 = counts.cpu().numpy(), offd.cpu().numpy(), out2.cpu().numpy()
    assert np.array_equal(c, cnt) and o[nq] == full
    assert np.all(r2[cap:] == -5)
    for k in range(nq):
        if o[k] + c[k] <= cap:
            assert np.array_equal(np.sort(r2[o[k]:o[k] + c[k]].view(np.uint32)), exp[k]), k
        else:
            seg = r2[o[k]:cap].view(np.uint32)
            assert len(np.unique(seg)) == len(seg) and np.all(np.isin(seg, exp[k])), k
    tr.close()
