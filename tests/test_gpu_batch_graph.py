"""Batched searches captured as a HIP graph and replayed, and the autotuner:
every replay (with other device work in between, and after the queries are
rewritten in place) equals an eager run; results after autotuning equal
brute force."""
import numpy as np
import pytest

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


def _segments(counts, offs, out):
    c, o, r = counts.cpu().numpy(), offs.cpu().numpy(), out.cpu().numpy()
    return c, o, [np.sort(r[o[k]:o[k] + c[k]]) for k in range(len(c))]


def test_graph_replays_equal_eager(torch, P):
    import pst_data as D
    n = 1 << 18
    pts = D.interval_points(n, seed=12)
    d = P.to_device(pts)
    tr = P.StaticPSTGPU(d, n, "f32", "u32")
    qs = D.float_query_mix(300, 700, seed=13)
    nq = len(qs)
    s = torch.cuda.Stream()
    dq = P.to_device(qs)
    counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
    offs = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
    tr.batch_count(dq, nq, counts)
    torch.cuda.synchronize()
    cap = int(counts.sum()) * 2 + 64
    out = torch.zeros(cap, dtype=torch.int32, device="cuda")
    tr.batch_search(dq, nq, counts, offs, out, cap, stream=s)
    s.synchronize()
    ref = _segments(counts, offs, out)
    c2 = torch.zeros_like(counts)
    o2 = torch.zeros_like(offs)
    out2 = torch.zeros_like(out)
    g = tr.batch_graph(dq, nq, c2, o2, out2, cap, stream=s)
    for rep in range(3):
        with torch.cuda.stream(s):
            c2.zero_(), o2.zero_(), out2.fill_(-1)
            tmp = torch.randn(1 << 20, device="cuda").sum()  # unrelated work between replays
        g.launch(s)
        s.synchronize()
        got = _segments(c2, o2, out2)
        print(f"replay {rep}", flush=True)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        assert all(np.array_equal(a, b) for a, b in zip(got[2], ref[2]))
        del tmp
    # the graph reads the queries' current contents
    qs2 = D.float_query_mix(300, 700, seed=14)
    dq.copy_(P.to_device(qs2))
    tr.batch_search(dq, nq, counts, offs, out, cap, stream=s)
    s.synchronize()
    ref2 = _segments(counts, offs, out)
    g.launch(s)
    s.synchronize()
    got2 = _segments(c2, o2, out2)
    assert np.array_equal(got2[0], ref2[0])
    assert all(np.array_equal(a, b) for a, b in zip(got2[2], ref2[2]))
    for k in (0, 500, nq - 1):
        q = qs2[k]
        exp = np.sort(pts["id"][O.brute_force(pts, int(q["kind"]), q["min_dim1"], q["max_dim1"], q["min_dim2"])])
        assert np.array_equal(got2[2][k], exp)
    g.close()
    tr.close()


def test_autotune_keeps_results(torch, P):
    import pst_data as D
    n = (1 << 17) + 3
    pts = D.int_points(n, seed=21)
    d = P.to_device(pts)
    tr = P.StaticPSTGPU(d, n, "i32", "u32")
    qs = D.int_query_mix(512, seed=22)
    nq = len(qs)
    dq = P.to_device(qs)
    cfg = tr.autotune(dq, nq)
    print("autotune", cfg, flush=True)
    assert cfg["waves_per_cu"] in (16, 24, 32, 40) and cfg["queries_per_walk"] in (1, 8)
    counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
    offs = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
    tr.batch_count(dq, nq, counts)
    torch.cuda.synchronize()
    total = int(counts.sum())
    out = torch.zeros(total + 8, dtype=torch.int32, device="cuda")
    tr.batch_search(dq, nq, counts, offs, out, total + 8)
    torch.cuda.synchronize()
    c, o, segs = _segments(counts, offs, out)
    for k in range(0, nq, 37):
        q = qs[k]
        exp = np.sort(pts["id"][O.brute_force(pts, int(q["kind"]), q["min_dim1"], q["max_dim1"], q["min_dim2"])])
        assert np.array_equal(segs[k].view(np.uint32), exp), k
    tr.close()
