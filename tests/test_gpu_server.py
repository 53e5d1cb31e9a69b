"""The query server (continuous batching over resident trees): every
completed result equals the same query searched alone and brute force, with
queries arriving while earlier batches run, under an output budget small
enough to make admission wait, and with a query too large for the budget
failing on its own."""
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


def _pts(rng, t, i, n):
    p = np.zeros(n, __import__("pst_amd").point_dtype(t, i))
    if t in ("f32", "f64"):
        a = rng.random(n)
        p["dim1"] = a
        p["dim2"] = a + rng.exponential(0.05, n)
    else:
        a = rng.integers(0, 1 << 20, n)
        p["dim1"] = a
        p["dim2"] = a + rng.geometric(1 / 2000, n)
    if i is not None:
        p["id"] = np.arange(n)
    return p


def _query(P, rng, t, p):
    qd = P.query_dtype(t)
    q = np.zeros(1, qd)
    lo, hi = float(p["dim1"].min()), float(p["dim2"].max())
    a, b = np.sort(rng.uniform(lo, hi, 2))
    c = rng.uniform(lo, hi)
    kind = int(rng.integers(0, 3))
    if kind == 1:
        a = b = c = rng.uniform(lo, hi)   # stabbing
    if t in ("i32", "i64"):
        a, b, c = int(a), int(b), int(c)
    q[0] = (a, b, c, kind)
    return q


def _read(P, ptr, n, rec):
    return P._d2h(ptr, n * rec.itemsize)[:n * rec.itemsize].view(rec) if n else np.zeros(0, rec)


@pytest.mark.parametrize("index", [False, True])
@pytest.mark.parametrize("budget", [1 << 30, 3 << 20])
def test_server_results_equal_alone(torch, P, budget, index):
    rng = np.random.default_rng(11)
    specs = [("f32", "u32", 1 << 15, False), ("i32", "u32", 50000, False), ("f64", "u64", 20000, True)]
    trees, pts, keep = [], [], []
    for t, i, n, forest in specs:
        p = _pts(rng, t, i, n)
        d = P.to_device(p)
        tr = P.StaticPSTGPUArr(d, n, 64, t, i) if forest else P.StaticPSTGPU(d, n, t, i)
        if index and not forest:
            tr.set_level_index(True)  # before the server, which sizes the index's buffers
        trees.append(tr)
        pts.append(p)
        keep.append(d)
    torch.cuda.synchronize()
    srv = P.Server(trees, budget, max_batch=64)
    want = {}
    done = {}
    nq = 400
    print(f"server budget={budget} nq={nq} level index={index}", flush=True)
    for k in range(nq):
        ti = int(rng.integers(0, len(trees)))
        t, i = specs[ti][0], specs[ti][1]
        q = _query(P, rng, t, pts[ti])
        tk = srv.submit(ti, q[0])
        want[tk] = (ti, q[0])
        if k % 7 == 0:
            srv.step()
            for d in srv.poll():
                r = srv.result(d)
                done[d] = (r, _read(P, *r, np.dtype(O.NP[specs[want[d][0]][1]])).copy())
                srv.release(d)
    guard = 0
    while len(done) < nq:
        srv.step()
        for d in srv.poll():
            ti = want[d][0]
            r = srv.result(d)
            done[d] = (r, _read(P, *r, np.dtype(O.NP[specs[ti][1]])).copy())
            srv.release(d)
        guard += 1
        assert guard < 2_000_000, "server made no progress"
    assert srv.outstanding() == 0
    for tk, (ti, q) in want.items():
        kind = int(q["kind"])
        alone = trees[ti].search(kind, q["min_dim1"], q["max_dim1"], q["min_dim2"], ret=P.RET_ID)
        got = done[tk][1]
        assert G.canon(got, False, True) == G.canon(alone, False, True), (tk, ti, kind)
        exp = pts[ti]["id"][O.brute_force(pts[ti], kind, q["min_dim1"], q["max_dim1"], q["min_dim2"])]
        assert G.canon(got, False, True) == G.canon(exp, False, True)
    srv.close()
    for tr in trees:
        tr.close()


def test_server_query_larger_than_budget_fails_alone(torch, P):
    rng = np.random.default_rng(5)
    p = _pts(rng, "f32", "u32", 100000)
    d = P.to_device(p)
    tr = P.StaticPSTGPU(d, len(p), "f32", "u32")
    srv = P.Server([tr], 64 << 10, max_batch=16)   # 64 KiB: 16384 IDs
    qd = P.query_dtype("f32")
    big = np.zeros(1, qd)
    big[0] = (0.0, 2.0, 0.0, 0)                       # every point
    small = np.zeros(1, qd)
    small[0] = (0.5, 0.5001, 0.0, 0)
    t_big = srv.submit(0, big[0])
    t_small = srv.submit(0, small[0])
    seen = {}
    for _ in range(100000):
        srv.step()
        for x in srv.poll():
            seen[x] = True
        if len(seen) == 2:
            break
    assert set(seen) == {t_big, t_small}
    with pytest.raises(P.PSTError) as e:
        srv.result(t_big)
    assert e.value.code == 4
    ptr, n = srv.result(t_small)
    exp = p["id"][O.brute_force(p, 0, np.float32(0.5), np.float32(0.5001), np.float32(0.0))]
    assert G.canon(_read(P, ptr, n, np.dtype(np.uint32)), False, True) == G.canon(exp, False, True)
    srv.release(t_big)
    srv.release(t_small)
    srv.close()
    tr.close()


def test_server_level_index_trees_to_2e22(torch, P):
    """The measured serving setup: deep trees of 2^20..2^22 points with the
    level index on every tree, float and int32 trees, queries arriving
    while earlier batches run, an output budget that makes admission wait.
    Every result equals brute force."""
    import pst_data as D
    trees, pts, mixes = [], [], []
    for k, lg in enumerate((20, 21, 22)):
        n = 1 << lg
        if k % 2 == 0:
            p, t = D.interval_points(n, seed=300 + lg), "f32"
            qs = D.float_query_mix(64, 64, seed=400 + lg)
        else:
            p, t = D.int_points(n, seed=300 + lg), "i32"
            qs = D.int_query_mix(128, seed=400 + lg)
        tr = P.StaticPSTGPU(P.to_device(p), n, t, "u32")
        tr.set_level_index(True)
        trees.append(tr)
        pts.append(p)
        mixes.append(qs)
    torch.cuda.synchronize()
    rng = np.random.default_rng(9)
    # 96 MB of output: the larger results of the float trees make the
    # admission wait for earlier tickets' release
    srv = P.Server(trees, 96 << 20, max_batch=128)
    want, done = {}, {}
    nq = 240
    print(f"server level index 2^20..2^22 nq={nq}", flush=True)
    for k in range(nq):
        ti = int(rng.integers(0, 3))
        q = mixes[ti][int(rng.integers(0, len(mixes[ti])))]
        want[srv.submit(ti, q)] = (ti, q)
        if k % 5 == 0:
            srv.step()
            for d in srv.poll():
                r = srv.result(d)
                done[d] = _read(P, *r, np.dtype(np.uint32)).copy()
                srv.release(d)
    guard = 0
    while len(done) < nq:
        srv.step()
        for d in srv.poll():
            r = srv.result(d)
            done[d] = _read(P, *r, np.dtype(np.uint32)).copy()
            srv.release(d)
        guard += 1
        assert guard < 5_000_000, "server made no progress"
    srv.close()
    for tk, (ti, q) in want.items():
        p = pts[ti]
        kind = int(q["kind"])
        m = p["dim2"] >= q["min_dim2"]
        if kind != 2:
            m &= p["dim1"] <= q["max_dim1"]
        if kind != 1:
            m &= p["dim1"] >= q["min_dim1"]
        assert np.array_equal(np.sort(done[tk]), np.flatnonzero(m).astype(np.uint32)), (tk, ti, kind)
    for tr in trees:
        tr.close()
