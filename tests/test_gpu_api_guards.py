"""Guards of the batched entry points: a zero output capacity is a sizing
call that writes nothing, and a batch graph pins its tree's buffers (calls
that would free or reallocate them fail and change nothing while it lives)."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

GUARD = 0x5A5A5A5A


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


def _tree(P, n, seed, index):
    import pst_data as D
    pts = D.interval_points(n, seed=seed)
    tr = P.StaticPSTGPU(P.to_device(pts), n, "f32", "u32")
    if index:
        tr.set_level_index(True)
    return pts, tr


def _expected_counts(pts, qs):
    out = []
    for q in qs:
        out.append(len(O.brute_force(pts, int(q["kind"]), q["min_dim1"], q["max_dim1"], q["min_dim2"])))
    return np.array(out, np.int64)


@pytest.mark.parametrize("index", [False, True])
def test_zero_capacity_is_a_sizing_call(torch, P, index):
    import pst_data as D
    n = 40000
    pts, tr = _tree(P, n, 31, index)
    qs = D.float_query_mix(40, 40, seed=32)
    nq = len(qs)
    dq = P.to_device(qs)
    want = _expected_counts(pts, qs)
    total = int(want.sum())
    print(f"capacity 0 index={index} nq={nq} total={total}", flush=True)
    counts = torch.full((nq,), -1, dtype=torch.int64, device="cuda")
    offs = torch.full((nq + 1,), -1, dtype=torch.int64, device="cuda")
    out = torch.full((total + 64,), GUARD, dtype=torch.int32, device="cuda")
    tr.batch_search(dq, nq, counts, offs, out, 0)
    torch.cuda.synchronize()
    assert np.array_equal(counts.cpu().numpy(), want)
    o = offs.cpu().numpy()
    assert o[0] == 0 and np.array_equal(np.diff(o), want) and o[nq] == total
    assert bool((out == GUARD).all().item()), "a zero-capacity call wrote into the output"
    # the same call with room for everything then writes every result
    tr.batch_search(dq, nq, counts, offs, out, total)
    torch.cuda.synchronize()
    r = out.cpu().numpy().view(np.uint32)
    for k in range(0, nq, 7):
        q = qs[k]
        exp = np.sort(pts["id"][O.brute_force(pts, int(q["kind"]), q["min_dim1"], q["max_dim1"], q["min_dim2"])])
        assert np.array_equal(np.sort(r[o[k]:o[k] + want[k]]), exp), k
    assert np.all(r[total:] == GUARD)
    tr.close()


def test_graph_pins_the_tree_buffers(torch, P):
    import pst_data as D
    n = 1 << 16
    pts, tr = _tree(P, n, 41, index=True)
    small = D.float_query_mix(30, 34, seed=42)
    big = D.float_query_mix(3000, 3000, seed=43)   # more queries than the graph's buffers hold
    s = torch.cuda.Stream()
    dq = P.to_device(small)
    nq = len(small)
    counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
    offs = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
    want = _expected_counts(pts, small)
    cap = int(want.sum()) + 16
    out = torch.zeros(cap, dtype=torch.int32, device="cuda")
    g = tr.batch_graph(dq, nq, counts, offs, out, cap, stream=s)
    print("graph created; refused calls follow", flush=True)
    dbig = P.to_device(big)
    nb = len(big)
    cb = torch.zeros(nb, dtype=torch.int64, device="cuda")
    ob = torch.zeros(nb + 1, dtype=torch.int64, device="cuda")
    outb = torch.zeros(1 << 22, dtype=torch.int32, device="cuda")
    for call in (lambda: tr.batch_search(dbig, nb, cb, ob, outb, outb.numel()),
                 lambda: tr.set_level_index(False),
                 lambda: tr.autotune(dq, nq)):
        with pytest.raises(P.PSTError) as e:
            call()
        assert e.value.code == 1 and "graph" in str(e.value)
    rc = P.lib().pst_gpu_destroy(tr.h)
    assert rc == 1 and b"graph" in P.lib().pst_last_error()
    # the graph still runs on intact buffers
    for rep in range(2):
        with torch.cuda.stream(s):
            counts.zero_(), offs.zero_(), out.fill_(-1)
        g.launch(s)
        s.synchronize()
        assert np.array_equal(counts.cpu().numpy(), want), rep
    g.close()
    # released: the refused calls now work
    tr.batch_search(dbig, nb, cb, ob, outb, outb.numel())
    torch.cuda.synchronize()
    assert np.array_equal(cb.cpu().numpy(), _expected_counts(pts, big))
    tr.set_level_index(False)
    tr.close()


def test_calc_global_mem_needed_covers_measured(torch, P):
    """calcGlobalMemNeeded (pst/static-pst-gpu.tu:646-671) against what the
    device reports: the tree, the handle's search buffers and a blocking
    search's result array fit in it; the level index takes its documented
    bytes on top."""
    import ctypes
    import pst_data as D
    n = (1 << 22) + 5
    pts = D.interval_points(n, seed=61)
    d = P.to_device(pts)
    free0 = torch.cuda.mem_get_info()[0]
    tr = P.StaticPSTGPU(d, n, "f32", "u32")
    num, ptr = ctypes.c_size_t(), ctypes.c_void_p()
    q = np.array([0.5], np.float32)
    P._check(P.lib().pst_gpu_two_sided_left_search(tr.h, P.RET_ID, ctypes.byref(num), ctypes.byref(ptr),
                                                   q.ctypes.data, q.ctypes.data))
    torch.cuda.synchronize()
    used = free0 - torch.cuda.mem_get_info()[0]
    P._check(P.lib().pst_free(ptr))
    calc = P.StaticPSTGPU.calcGlobalMemNeeded(n, "f32", "u32")
    S = (1 << n.bit_length()) - 1
    print(f"resident after a blocking search {used}, calcGlobalMemNeeded {calc}, tree {S * 24}", flush=True)
    assert S * 24 <= used <= calc
    free1 = torch.cuda.mem_get_info()[0]
    tr.set_level_index(True)
    torch.cuda.synchronize()
    grown = free1 - torch.cuda.mem_get_info()[0]
    want = P.StaticPSTGPU.level_index_bytes(n, "f32", "u32")
    # two allocations, rounded to the allocator's pages, plus at most the
    # library's static device data if its code object loads here
    assert want <= grown <= want + (48 << 20), (grown, want)
    tr.close()
